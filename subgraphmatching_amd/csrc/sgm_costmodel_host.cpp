// Host driver of the learned-order cost model (kernels in sgm_costmodel.hip): features,
// forward / backward orchestration, Adam training over resident samples, prediction.
#include <cmath>
#include <random>

#include "sgm_costmodel.h"
#include "sgm_objects.h"

namespace {

// fp32 parameter layout (one buffer, so Adam and zeroing run once over it)
constexpr uint32_t OW0 = 0;                       // CM_F x CM_H
constexpr uint32_t OB0 = OW0 + CM_F * CM_H;       // CM_H
constexpr uint32_t LAYER = CM_X * CM_H + CM_H;    // W_l then b_l
constexpr uint32_t OL1 = OB0 + CM_H;
constexpr uint32_t OWO = OL1 + 3 * LAYER;         // CM_H
constexpr uint32_t OBO = OWO + CM_H;              // 1
constexpr uint32_t NPARAM = OBO + 1;
constexpr uint32_t NFEAT = 8;                     // used feature columns
inline uint32_t OW(int l) { return OL1 + (uint32_t)(l - 1) * LAYER; }
inline uint32_t OB(int l) { return OW(l) + CM_X * CM_H; }

uint16_t host_f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Features of one sample in order position: row i is query vertex order[i].
//   0 deg_q(u) / 8            1 log2(1 + |C(u)|) / 16     2 log2(1 + mean candidate degree) / 16
//   3 i / 32                  4 backward neighbours / deg_q(u)
//   5 backward neighbours / 8 6 [i == 0]                  7 n / 32
// f: n rows of NFEAT values (row-major, stride NFEAT).
int features(const sgm_cm_sample& s, float* f, uint8_t* valid, uint32_t* back, uint32_t* fwd)
{
    const uint32_t n = s.n;
    if (n < 1 || n > CM_ROWS || !s.qoff || !s.order || !s.cand_count || !s.cand_degsum) return SGM_EINVAL;
    uint32_t pos[CM_ROWS], seen = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t u = s.order[i];
        if (u >= n || ((seen >> u) & 1u)) return SGM_EORDER;
        seen |= 1u << u;
        pos[u] = i;
    }
    for (uint32_t i = 0; i < CM_ROWS; i++) { valid[i] = i < n; back[i] = 0; fwd[i] = 0; }
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t u = s.order[i];
        for (uint32_t p = s.qoff[u]; p < s.qoff[u + 1]; p++) {
            const uint32_t j = pos[s.qadj[p]];
            if (j < i) back[i] |= 1u << j;
            else fwd[i] |= 1u << j;
        }
        const float deg = (float)(s.qoff[u + 1] - s.qoff[u]);
        const double cnt = (double)s.cand_count[u];
        const double mean = cnt > 0 ? (double)s.cand_degsum[u] / cnt : 0.0;
        const float nb = (float)__builtin_popcount(back[i]);
        float* r = f + (size_t)i * NFEAT;
        r[0] = deg / 8.f;
        r[1] = (float)(std::log2(1.0 + cnt) / 16.0);
        r[2] = (float)(std::log2(1.0 + mean) / 16.0);
        r[3] = (float)i / 32.f;
        r[4] = deg > 0 ? nb / deg : 0.f;
        r[5] = nb / 8.f;
        r[6] = i == 0 ? 1.f : 0.f;
        r[7] = (float)n / 32.f;
    }
    return SGM_OK;
}

struct Bufs {
    uint32_t cap = 0;  // samples
    uint16_t* F = nullptr;
    uint8_t* valid = nullptr;
    uint32_t* back = nullptr;
    uint32_t* fwd = nullptr;
    float* target = nullptr;
    uint16_t* X[3] = {nullptr, nullptr, nullptr};
    float* Z[4] = {nullptr, nullptr, nullptr, nullptr};
    float* H3 = nullptr;
    float* y = nullptr;
    float* dH = nullptr;
    uint16_t* dZ = nullptr;
    float* dX = nullptr;
    uint32_t* idx = nullptr;
};

void free_bufs(Bufs& b)
{
    dfree(b.F); dfree(b.valid); dfree(b.back); dfree(b.fwd); dfree(b.target);
    for (auto& x : b.X) dfree(x);
    for (auto& z : b.Z) dfree(z);
    dfree(b.H3); dfree(b.y); dfree(b.dH); dfree(b.dZ); dfree(b.dX); dfree(b.idx);
    b = Bufs();
}

}  // namespace

struct sgm_costmodel {
    sgm_ctx* c = nullptr;
    float* p = nullptr;   // parameters
    float* g = nullptr;   // gradients
    float* m = nullptr;   // Adam moments
    float* v = nullptr;
    float* loss = nullptr;
    uint32_t step = 0;
    uint16_t* w0b = nullptr;
    uint16_t* wb[3] = {nullptr, nullptr, nullptr};
    uint16_t* wt[3] = {nullptr, nullptr, nullptr};
    Bufs work;    // minibatch / prediction chunk
    Bufs train;   // resident training set (F, valid, back, fwd, target only)
    hipEvent_t ev[2] = {nullptr, nullptr};
    double train_ms = 0, train_flops = 0, infer_ms = 0, infer_flops = 0;
};

namespace {

int alloc_bufs(sgm_ctx* c, Bufs& b, uint32_t S, bool full)
{
    if (b.cap >= S) return SGM_OK;
    free_bufs(b);
    const size_t R = (size_t)S * CM_ROWS;
    HIPCHK(c, dalloc(&b.F, R * CM_F));
    HIPCHK(c, dalloc(&b.valid, R));
    HIPCHK(c, dalloc(&b.back, R));
    HIPCHK(c, dalloc(&b.fwd, R));
    HIPCHK(c, dalloc(&b.target, S));
    if (full) {
        for (auto& x : b.X) HIPCHK(c, dalloc(&x, R * CM_X));
        for (auto& z : b.Z) HIPCHK(c, dalloc(&z, R * CM_H));
        HIPCHK(c, dalloc(&b.H3, R * CM_H));
        HIPCHK(c, dalloc(&b.y, S));
        HIPCHK(c, dalloc(&b.dH, R * CM_H));
        HIPCHK(c, dalloc(&b.dZ, R * CM_H));
        HIPCHK(c, dalloc(&b.dX, R * CM_X));
        HIPCHK(c, dalloc(&b.idx, S));
    }
    b.cap = S;
    return SGM_OK;
}

// Features of samples [0, count) into device buffers b (F as bf16).
int upload(sgm_costmodel* m, Bufs& b, const sgm_cm_sample* s, uint32_t count, const float* target)
{
    sgm_ctx* c = m->c;
    const size_t R = (size_t)count * CM_ROWS;
    std::vector<uint16_t> F(R * CM_F, 0);  // columns NFEAT.. and padding rows stay zero
    std::vector<uint8_t> valid(R);
    std::vector<uint32_t> back(R), fwd(R);
    float f[CM_ROWS * NFEAT];
    for (uint32_t k = 0; k < count; k++) {
        const int rc = features(s[k], f, valid.data() + (size_t)k * CM_ROWS, back.data() + (size_t)k * CM_ROWS,
                                fwd.data() + (size_t)k * CM_ROWS);
        if (rc) return set_err(c, rc, "cost model: bad sample %u", k);
        uint16_t* dst = F.data() + (size_t)k * CM_ROWS * CM_F;
        for (uint32_t i = 0; i < s[k].n; i++)
            for (uint32_t j = 0; j < NFEAT; j++) dst[(size_t)i * CM_F + j] = host_f2bf(f[i * NFEAT + j]);
    }
    HIPCHK(c, hipMemcpyAsync(b.F, F.data(), F.size() * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.valid, valid.data(), R, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.back, back.data(), R * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.fwd, fwd.data(), R * 4, hipMemcpyHostToDevice, c->stream));
    if (target) HIPCHK(c, hipMemcpyAsync(b.target, target, sizeof(float) * count, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SGM_OK;
}

void cast_weights(sgm_costmodel* m)
{
    hipStream_t s = m->c->stream;
    sgm::cm_cast(s, m->p + OW0, CM_F, CM_H, m->w0b, nullptr);
    for (int l = 1; l <= 3; l++) sgm::cm_cast(s, m->p + OW(l), CM_X, CM_H, m->wb[l - 1], m->wt[l - 1]);
}

// Forward over S samples in b on stream s (the context's unless a caller's inputs are on
// another); with `loss`, also the readout gradient into b.dH and the output-layer gradients
// into m->g.
void forward(sgm_costmodel* m, Bufs& b, uint32_t S, bool loss, hipStream_t s = nullptr)
{
    if (!s) s = m->c->stream;
    const uint32_t R = S * CM_ROWS;
    sgm::cm_gemm(s, b.F, CM_F, m->w0b, CM_H, m->p + OB0, b.valid, 1, b.Z[0], CM_H, b.X[0], CM_X, nullptr, 0, R, CM_H, CM_F);
    for (int l = 1; l <= 3; l++) {
        sgm::cm_aggregate(s, b.X[l - 1], b.back, b.fwd, R);
        uint16_t* next = l < 3 ? b.X[l] : nullptr;
        float* hf = l == 3 ? b.H3 : nullptr;
        sgm::cm_gemm(s, b.X[l - 1], CM_X, m->wb[l - 1], CM_H, m->p + OB(l), b.valid, 1, b.Z[l], CM_H, next, CM_X, hf,
                     CM_H, R, CM_H, CM_X);
    }
    sgm::cm_readout(s, b.H3, b.valid, m->p + OWO, m->p + OBO, b.y, loss ? b.target : nullptr, b.dH, m->g + OWO,
                    m->g + OBO, m->loss, S);
}

// Backward from b.dH (gradient of the loss w.r.t. H3) into m->g (accumulated).
void backward(sgm_costmodel* m, Bufs& b, uint32_t S)
{
    hipStream_t s = m->c->stream;
    const uint32_t R = S * CM_ROWS;
    for (int l = 3; l >= 1; l--) {
        sgm::cm_act_bwd(s, b.dH, b.Z[l], b.valid, b.dZ, m->g + OB(l), R);
        sgm::cm_gemm_tn(s, b.X[l - 1], CM_X, b.dZ, CM_H, m->g + OW(l), R, CM_X, CM_H);
        // dX = dZ W_l^T, then back through the neighbour means (the transpose of "j is a
        // backward neighbour of i" is "i is a forward neighbour of j")
        sgm::cm_gemm(s, b.dZ, CM_H, m->wt[l - 1], CM_X, nullptr, nullptr, 0, nullptr, 0, nullptr, 0, b.dX, CM_X, R, CM_X,
                     CM_H);
        sgm::cm_aggregate_bwd(s, b.dX, b.back, b.fwd, b.fwd, b.back, b.dH, R);
    }
    sgm::cm_act_bwd(s, b.dH, b.Z[0], b.valid, b.dZ, m->g + OB0, R);
    sgm::cm_gemm_tn(s, b.F, CM_F, b.dZ, CM_H, m->g + OW0, R, CM_F, CM_H);
}

double fwd_flops(uint32_t S) { return 2.0 * S * CM_ROWS * ((double)CM_F * CM_H + 3.0 * CM_X * CM_H); }
double bwd_flops(uint32_t S) { return 2.0 * S * CM_ROWS * ((double)CM_F * CM_H + 6.0 * CM_X * CM_H); }

}  // namespace

// Device-built inputs (the batch runner's ranking pass): the model's work buffers for S
// samples, then a forward pass over them in place.
int costmodel_work_inputs(sgm_costmodel* m, uint32_t S, uint16_t** F, uint8_t** valid, uint32_t** back, uint32_t** fwd)
{
    const int rc = alloc_bufs(m->c, m->work, std::max(S, 256u), true);
    if (rc) return rc;
    *F = m->work.F;
    *valid = m->work.valid;
    *back = m->work.back;
    *fwd = m->work.fwd;
    return SGM_OK;
}

int costmodel_forward_work(sgm_costmodel* m, uint32_t S, const float** y, hipStream_t s)
{
    sgm_ctx* c = m->c;
    HIPCHK(c, hipEventRecord(m->ev[0], s));
    forward(m, m->work, S, false, s);
    HIPCHK(c, hipEventRecord(m->ev[1], s));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(m->ev[1]));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
    m->infer_ms += ms;
    m->infer_flops += fwd_flops(S);
    *y = m->work.y;
    return SGM_OK;
}

extern "C" {

uint32_t sgm_costmodel_param_count(void) { return NPARAM; }

void sgm_costmodel_destroy(sgm_costmodel* m)
{
    if (!m) return;
    forget(m->c->models, m);
    (void)hipSetDevice(m->c->device);
    (void)hipStreamSynchronize(m->c->stream);
    dfree(m->p); dfree(m->g); dfree(m->m); dfree(m->v); dfree(m->loss);
    dfree(m->w0b);
    for (auto& x : m->wb) dfree(x);
    for (auto& x : m->wt) dfree(x);
    free_bufs(m->work);
    free_bufs(m->train);
    for (auto& e : m->ev)
        if (e) (void)hipEventDestroy(e);
    delete m;
}

int sgm_costmodel_create(sgm_ctx* c, uint64_t seed, sgm_costmodel** out)
{
    if (!c || !out) return SGM_EINVAL;
    *out = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    sgm_costmodel* m = new sgm_costmodel();
    m->c = c;
    hipError_t e = dalloc(&m->p, NPARAM);
    if (e == hipSuccess) e = dalloc(&m->g, NPARAM);
    if (e == hipSuccess) e = dalloc(&m->m, NPARAM);
    if (e == hipSuccess) e = dalloc(&m->v, NPARAM);
    if (e == hipSuccess) e = dalloc(&m->loss, 1);
    if (e == hipSuccess) e = dalloc(&m->w0b, CM_F * CM_H);
    for (int l = 0; l < 3 && e == hipSuccess; l++) {
        e = dalloc(&m->wb[l], CM_X * CM_H);
        if (e == hipSuccess) e = dalloc(&m->wt[l], CM_X * CM_H);
    }
    for (auto& ev : m->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) {
        sgm_costmodel_destroy(m);
        return set_err(c, SGM_ENOMEM, "cost model alloc: %s", hipGetErrorString(e));
    }
    // He-uniform initialisation from the seed; biases zero
    std::vector<float> p(NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](uint32_t off, uint32_t fan_in, uint32_t count) {
        std::uniform_real_distribution<float> d(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
        for (uint32_t i = 0; i < count; i++) p[off + i] = d(rng);
    };
    fill(OW0, NFEAT, CM_F * CM_H);
    for (uint32_t r = NFEAT; r < CM_F; r++)
        for (uint32_t c2 = 0; c2 < CM_H; c2++) p[OW0 + r * CM_H + c2] = 0.f;  // unused feature rows
    for (int l = 1; l <= 3; l++) fill(OW(l), CM_X, CM_X * CM_H);
    fill(OWO, CM_H, CM_H);
    HIPCHK(c, hipMemcpy(m->p, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(m->m, 0, sizeof(float) * NPARAM));
    HIPCHK(c, hipMemset(m->v, 0, sizeof(float) * NPARAM));
    cast_weights(m);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->models.push_back(m);
    *out = m;
    return SGM_OK;
}

int sgm_costmodel_params(sgm_costmodel* m, float* get, const float* set)
{
    if (!m) return SGM_EINVAL;
    sgm_ctx* c = m->c;
    HIPCHK(c, hipSetDevice(c->device));
    if (get) HIPCHK(c, hipMemcpy(get, m->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    if (set) {
        HIPCHK(c, hipMemcpy(m->p, set, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
        cast_weights(m);
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return SGM_OK;
}

int sgm_costmodel_features(const sgm_cm_sample* s, float* feats, uint32_t* back, uint32_t* fwd)
{
    if (!s || !feats || !back || !fwd) return SGM_EINVAL;
    uint8_t valid[CM_ROWS];
    float f[CM_ROWS * NFEAT];
    const int rc = features(*s, f, valid, back, fwd);
    if (rc) return rc;
    std::memset(feats, 0, sizeof(float) * CM_ROWS * CM_F);
    for (uint32_t i = 0; i < s->n; i++)
        for (uint32_t j = 0; j < NFEAT; j++) feats[i * CM_F + j] = f[i * NFEAT + j];
    return SGM_OK;
}

int sgm_costmodel_predict(sgm_costmodel* m, const sgm_cm_sample* s, uint32_t count, float* out)
{
    if (!m || (!s && count) || (!out && count)) return SGM_EINVAL;
    sgm_ctx* c = m->c;
    HIPCHK(c, hipSetDevice(c->device));
    const uint32_t chunk = 8192;
    for (uint32_t k0 = 0; k0 < count; k0 += chunk) {
        const uint32_t S = std::min(chunk, count - k0);
        int rc = alloc_bufs(c, m->work, std::max(S, 256u), true);
        if (!rc) rc = upload(m, m->work, s + k0, S, nullptr);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(m->ev[0], c->stream));
        forward(m, m->work, S, false);
        HIPCHK(c, hipEventRecord(m->ev[1], c->stream));
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(out + k0, m->work.y, sizeof(float) * S, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
        m->infer_ms += ms;
        m->infer_flops += fwd_flops(S);
    }
    return SGM_OK;
}

int sgm_costmodel_grad(sgm_costmodel* m, const sgm_cm_sample* s, const float* target, uint32_t count, float* grad,
                       float* loss)
{
    if (!m || !s || !target || count == 0) return SGM_EINVAL;
    sgm_ctx* c = m->c;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = alloc_bufs(c, m->work, count, true);
    if (!rc) rc = upload(m, m->work, s, count, target);
    if (rc) return rc;
    sgm::cm_zero(c->stream, m->g, NPARAM);
    sgm::cm_zero(c->stream, m->loss, 1);
    forward(m, m->work, count, true);
    backward(m, m->work, count);
    HIPCHK(c, hipGetLastError());
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, m->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->stream));
    if (loss) HIPCHK(c, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SGM_OK;
}

int sgm_costmodel_train(sgm_costmodel* m, const sgm_cm_sample* s, const float* target, uint32_t count,
                        uint32_t epochs, uint32_t minibatch, float lr, uint64_t seed, float* epoch_loss)
{
    if (!m || !s || !target || count == 0 || minibatch == 0) return SGM_EINVAL;
    sgm_ctx* c = m->c;
    HIPCHK(c, hipSetDevice(c->device));
    minibatch = std::min(minibatch, count);
    int rc = alloc_bufs(c, m->train, count, false);
    if (!rc) rc = upload(m, m->train, s, count, target);
    if (!rc) rc = alloc_bufs(c, m->work, minibatch, true);
    if (rc) return rc;
    if (m->step == 0) {  // start the output bias at the mean target
        double mean = 0;
        for (uint32_t k = 0; k < count; k++) mean += target[k];
        const float bo = (float)(mean / count);
        HIPCHK(c, hipMemcpy(m->p + OBO, &bo, sizeof(float), hipMemcpyHostToDevice));
    }
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> perm(count);
    for (uint32_t k = 0; k < count; k++) perm[k] = k;
    HIPCHK(c, hipEventRecord(m->ev[0], c->stream));
    double flops = 0;
    for (uint32_t ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        sgm::cm_zero(c->stream, m->loss, 1);
        uint32_t batches = 0;
        for (uint32_t k0 = 0; k0 < count; k0 += minibatch) {
            const uint32_t S = std::min(minibatch, count - k0);
            HIPCHK(c, hipMemcpyAsync(m->work.idx, perm.data() + k0, sizeof(uint32_t) * S, hipMemcpyHostToDevice,
                                     c->stream));
            sgm::cm_gather(c->stream, m->work.idx, S, m->train.F, m->train.valid, m->train.back, m->train.fwd,
                           m->train.target, m->work.F, m->work.valid, m->work.back, m->work.fwd, m->work.target);
            sgm::cm_zero(c->stream, m->g, NPARAM);
            forward(m, m->work, S, true);
            backward(m, m->work, S);
            sgm::cm_adam(c->stream, m->p, m->g, m->m, m->v, NPARAM, lr, ++m->step);
            cast_weights(m);
            flops += fwd_flops(S) + bwd_flops(S);
            batches++;
        }
        HIPCHK(c, hipGetLastError());
        float l = 0;
        HIPCHK(c, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (epoch_loss) epoch_loss[ep] = l / (float)batches;
    }
    HIPCHK(c, hipEventRecord(m->ev[1], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
    m->train_ms += ms;
    m->train_flops += flops;
    return SGM_OK;
}

int sgm_costmodel_timing(const sgm_costmodel* m, double* train_ms, double* train_flops, double* infer_ms,
                         double* infer_flops)
{
    if (!m) return SGM_EINVAL;
    if (train_ms) *train_ms = m->train_ms;
    if (train_flops) *train_flops = m->train_flops;
    if (infer_ms) *infer_ms = m->infer_ms;
    if (infer_flops) *infer_flops = m->infer_flops;
    return SGM_OK;
}

}  // extern "C"
