// Learned matching-order cost model: a 3-layer message-passing network over the query graph
// (hidden 64), BF16 operands on MFMA with FP32 accumulation, trained on the search-tree node
// counts the enumerator logs and used to rank candidate orders.
//
// A sample is one (query, order) pair with the query's candidate statistics.  Its vertices
// are laid out in order position (row i = order[i]), padded to 32 rows, so "backward" and
// "forward" neighbours of a row are the query neighbours placed before / after it.  Network:
//   H0 = act(F W0 + b0)                                    F: 32 x 32 features
//   X_l = [H_{l-1} | mean of backward rows | mean of forward rows]        (32 x 192)
//   H_l = act(X_l W_l + b_l),  l = 1..3                    W_l: 192 x 64
//   y   = mean over the sample's rows of H_3, dot w_o, + b_o   (predicted log2(1 + nodes))
// act(z) = max(z, 0) on real rows and 0 on padding rows.  Loss: mean squared error.
//
// Kernels: one MFMA GEMM (a wave per 32 x 32 output tile, v_mfma_f32_32x32x8bf16_1k) with a
// bias / activation epilogue, a split-K MFMA GEMM for weight gradients (A^T B over the rows),
// and scalar kernels for the sparse neighbour means, readout / loss and Adam.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sgm_costmodel.h"
#include "sgm_internal.h"
#include "sgm_wave.h"

namespace sgm {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// float -> bfloat16 bits, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One 32 x 32 output tile of A (M x K, element (i,k) at A[i*sa + k*ka]) times B (K x N,
// element (k,j) at B[k*kb + j*sb]), K a multiple of 8.  Operand layout of the 32x32x8 MFMA:
// lane l supplies A rows l%32, k = 4*(l/32) .. +3 and B columns l%32 with the same k; the
// accumulator element r of lane l is D[8*(r/4) + 4*(l/32) + r%4][l%32].
__device__ __forceinline__ v16f mfma_tile(const uint16_t* __restrict__ A, size_t sa, size_t ka,
                                          const uint16_t* __restrict__ B, size_t kb, size_t sb, uint32_t K,
                                          uint32_t i0, uint32_t j0, v16f acc)
{
    const uint32_t lane = lane_id();
    const uint32_t r = lane & 31u, kq = (lane >> 5) * 4u;
    const uint16_t* a = A + (size_t)(i0 + r) * sa;
    const uint16_t* b = B + (size_t)(j0 + r) * sb;
    for (uint32_t k = 0; k < K; k += 8) {
        v4s av, bv;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            av[t] = (short)a[(size_t)(k + kq + t) * ka];
            bv[t] = (short)b[(size_t)(k + kq + t) * kb];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ uint32_t acc_row(uint32_t r, uint32_t lane) { return 8u * (r >> 2) + 4u * (lane >> 5) + (r & 3u); }

// C = epilogue(A B + bias).  A: M x K bf16 row-major (lda), B: K x N bf16 row-major (ldb).
// Outputs (each optional): zf = A B + bias (f32, ldz); hb / cf = act(zf) as bf16 (ldh) /
// f32 (ldc), where act zeroes negative values when `relu` and whole rows whose valid[row]
// is 0 (when valid is given).
__global__ void __launch_bounds__(256) k_cm_gemm(const uint16_t* __restrict__ A, uint32_t lda,
                                                 const uint16_t* __restrict__ B, uint32_t ldb,
                                                 const float* __restrict__ bias, const uint8_t* __restrict__ valid,
                                                 int relu, float* __restrict__ zf, uint32_t ldz,
                                                 uint16_t* __restrict__ hb, uint32_t ldh, float* __restrict__ cf,
                                                 uint32_t ldc, uint32_t M, uint32_t N, uint32_t K)
{
    const uint32_t lane = lane_id();
    const uint32_t tiles_n = N / 32, tiles = (M / 32) * tiles_n;
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t t = blockIdx.x * waves + (threadIdx.x >> 6); t < tiles; t += gridDim.x * waves) {
        const uint32_t i0 = (t / tiles_n) * 32, j0 = (t % tiles_n) * 32;
        v16f acc = {0};
        acc = mfma_tile(A, lda, 1, B, ldb, 1, K, i0, j0, acc);
        const uint32_t col = j0 + (lane & 31u);
        const float bcol = bias ? bias[col] : 0.f;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t row = i0 + acc_row(r, lane);
            const float z = acc[r] + bcol;
            float h = (relu && z < 0.f) ? 0.f : z;
            if (valid && !valid[row]) h = 0.f;
            if (zf) zf[(size_t)row * ldz + col] = z;
            if (hb) hb[(size_t)row * ldh + col] = f2bf(h);
            if (cf) cf[(size_t)row * ldc + col] = h;
        }
    }
}

// out[Ka x N] += A^T B over R rows: A is R x Ka (lda), B is R x N (ldb), both bf16
// row-major; R a multiple of 8.  Blocks take slices of `chunk` rows, each wave one output
// tile, and add their partial tiles to out with atomics.
__global__ void __launch_bounds__(256) k_cm_gemm_tn(const uint16_t* __restrict__ A, uint32_t lda,
                                                    const uint16_t* __restrict__ B, uint32_t ldb,
                                                    float* __restrict__ out, uint32_t R, uint32_t Ka, uint32_t N,
                                                    uint32_t chunk)
{
    const uint32_t lane = lane_id();
    const uint32_t tiles_n = N / 32, tiles = (Ka / 32) * tiles_n;
    const uint32_t r0 = blockIdx.y * chunk;
    if (r0 >= R) return;
    const uint32_t rows = R - r0 < chunk ? R - r0 : chunk;
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t t = blockIdx.x * waves + (threadIdx.x >> 6); t < tiles; t += gridDim.x * waves) {
        const uint32_t i0 = (t / tiles_n) * 32, j0 = (t % tiles_n) * 32;
        v16f acc = {0};
        // A^T element (i, k) = A[(r0 + k) * lda + i]
        acc = mfma_tile(A + (size_t)r0 * lda, 1, lda, B + (size_t)r0 * ldb, ldb, 1, rows, i0, j0, acc);
        const uint32_t col = j0 + (lane & 31u);
#pragma unroll
        for (int r = 0; r < 16; r++) atomicAdd(&out[(size_t)(i0 + acc_row(r, lane)) * N + col], acc[r]);
    }
}

// Neighbour means of one layer: for every row i of every sample, X[i][64 + c] = mean of
// H[j][c] over backward neighbours j, X[i][128 + c] = the same over forward neighbours
// (0 without neighbours).  H is columns 0..63 of X itself.  One thread per (row, column).
__global__ void __launch_bounds__(256) k_cm_aggregate(uint16_t* __restrict__ X, const uint32_t* __restrict__ back,
                                                      const uint32_t* __restrict__ fwd, uint32_t R)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)R * 64; t += (size_t)gridDim.x * blockDim.x) {
        const uint32_t row = (uint32_t)(t >> 6), c = (uint32_t)(t & 63u);
        const uint32_t base = row & ~31u;
        uint32_t mb = back[row], mf = fwd[row];
        const uint32_t nb = __popc(mb), nf = __popc(mf);
        float sb = 0.f, sf = 0.f;
        while (mb) { const uint32_t j = __ffs(mb) - 1; mb &= mb - 1; sb += bf2f(X[(size_t)(base + j) * CM_X + c]); }
        while (mf) { const uint32_t j = __ffs(mf) - 1; mf &= mf - 1; sf += bf2f(X[(size_t)(base + j) * CM_X + c]); }
        X[(size_t)row * CM_X + 64 + c] = f2bf(nb ? sb / (float)nb : 0.f);
        X[(size_t)row * CM_X + 128 + c] = f2bf(nf ? sf / (float)nf : 0.f);
    }
}

// Backward of the neighbour means plus the direct path: dH[j][c] = dX[j][c] +
// sum over rows i with j backward of i of dX[i][64+c] / nb(i) + the same for forward.
// back_t / fwd_t are the transposed masks (bit i of back_t[j]: j is a backward neighbour of i).
__global__ void __launch_bounds__(256) k_cm_aggregate_bwd(const float* __restrict__ dX, const uint32_t* __restrict__ back,
                                                          const uint32_t* __restrict__ fwd,
                                                          const uint32_t* __restrict__ back_t,
                                                          const uint32_t* __restrict__ fwd_t, float* __restrict__ dH,
                                                          uint32_t R)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)R * 64; t += (size_t)gridDim.x * blockDim.x) {
        const uint32_t row = (uint32_t)(t >> 6), c = (uint32_t)(t & 63u);
        const uint32_t base = row & ~31u;
        float s = dX[(size_t)row * CM_X + c];
        uint32_t mb = back_t[row], mf = fwd_t[row];
        while (mb) {
            const uint32_t i = __ffs(mb) - 1; mb &= mb - 1;
            s += dX[(size_t)(base + i) * CM_X + 64 + c] / (float)__popc(back[base + i]);
        }
        while (mf) {
            const uint32_t i = __ffs(mf) - 1; mf &= mf - 1;
            s += dX[(size_t)(base + i) * CM_X + 128 + c] / (float)__popc(fwd[base + i]);
        }
        dH[(size_t)row * 64 + c] = s;
    }
}

// dZ = dH where the activation passed it (valid row, z > 0), as bf16 (the next GEMMs'
// operand); db[c] += column sums (one atomic per block and column).
__global__ void __launch_bounds__(256) k_cm_act_bwd(const float* __restrict__ dH, const float* __restrict__ Z,
                                                    const uint8_t* __restrict__ valid, uint16_t* __restrict__ dZ,
                                                    float* __restrict__ db, uint32_t R)
{
    __shared__ float colsum[64];
    if (threadIdx.x < 64) colsum[threadIdx.x] = 0.f;
    __syncthreads();
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)R * 64; t += (size_t)gridDim.x * blockDim.x) {
        const uint32_t row = (uint32_t)(t >> 6), c = (uint32_t)(t & 63u);
        const float g = (valid[row] && Z[t] > 0.f) ? dH[t] : 0.f;
        dZ[t] = f2bf(g);
        if (g != 0.f) atomicAdd(&colsum[c], g);
    }
    __syncthreads();
    if (threadIdx.x < 64 && colsum[threadIdx.x] != 0.f) atomicAdd(&db[threadIdx.x], colsum[threadIdx.x]);
}

// Readout and loss, one wave per sample: y = mean over real rows of H3 . w_o + b_o.  With
// targets: loss += (y - t)^2 / S, dH3 rows = 2 (y - t) / S / n * w_o, dw_o, db_o.
__global__ void __launch_bounds__(256) k_cm_readout(const float* __restrict__ H3, const uint8_t* __restrict__ valid,
                                                    const float* __restrict__ wo, const float* __restrict__ bo,
                                                    float* __restrict__ y, const float* __restrict__ target,
                                                    float* __restrict__ dH3, float* __restrict__ dwo,
                                                    float* __restrict__ dbo, float* __restrict__ loss, uint32_t S)
{
    const uint32_t lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t s = blockIdx.x * waves + (threadIdx.x >> 6); s < S; s += gridDim.x * waves) {
        const uint32_t base = s * 32;
        float n = 0.f, m = 0.f;  // lane = column
        for (uint32_t i = 0; i < 32; i++) {
            if (!valid[base + i]) continue;
            n += 1.f;
            m += H3[(size_t)(base + i) * 64 + lane];
        }
        m = n > 0.f ? m / n : 0.f;
        float p = m * wo[lane];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) p += __shfl_xor(p, o);
        const float ys = p + bo[0];
        if (lane == 0) y[s] = ys;
        if (target) {
            const float e = ys - target[s];
            const float g = 2.f * e / (float)S;
            if (lane == 0) { atomicAdd(loss, e * e / (float)S); atomicAdd(dbo, g); }
            atomicAdd(&dwo[lane], g * m);
            for (uint32_t i = 0; i < 32; i++)
                dH3[(size_t)(base + i) * 64 + lane] = valid[base + i] ? g / n * wo[lane] : 0.f;
        }
    }
}

// Adam over all parameters (fp32 master copies); t = step number from 1.
__global__ void __launch_bounds__(256) k_cm_adam(float* __restrict__ p, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v, uint32_t n, float lr,
                                                 float b1, float b2, float eps, float c1, float c2)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// bf16 copies of a rows x cols fp32 matrix: straight (ld = cols) and transposed (ld = rows).
__global__ void __launch_bounds__(256) k_cm_cast(const float* __restrict__ w, uint32_t rows, uint32_t cols,
                                                 uint16_t* __restrict__ wb, uint16_t* __restrict__ wt)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rows * cols; i += gridDim.x * blockDim.x) {
        const uint16_t h = f2bf(w[i]);
        if (wb) wb[i] = h;
        if (wt) wt[(size_t)(i % cols) * rows + i / cols] = h;
    }
}

// Minibatch gather: sample s of the batch is sample idx[s] of the resident training set.
__global__ void __launch_bounds__(256) k_cm_gather(const uint32_t* __restrict__ idx, uint32_t S,
                                                   const uint16_t* __restrict__ F, const uint8_t* __restrict__ valid,
                                                   const uint32_t* __restrict__ back, const uint32_t* __restrict__ fwd,
                                                   const float* __restrict__ target, uint16_t* __restrict__ oF,
                                                   uint8_t* __restrict__ ovalid, uint32_t* __restrict__ oback,
                                                   uint32_t* __restrict__ ofwd, float* __restrict__ otarget)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)S * CM_ROWS * CM_F;
         t += (size_t)gridDim.x * blockDim.x) {
        const uint32_t s = (uint32_t)(t / (CM_ROWS * CM_F)), e = (uint32_t)(t % (CM_ROWS * CM_F));
        const size_t src = (size_t)idx[s] * CM_ROWS * CM_F + e;
        oF[t] = F[src];
        if (e % CM_F == 0) {
            const uint32_t i = e / CM_F;
            const size_t sr = (size_t)idx[s] * CM_ROWS + i, dr = (size_t)s * CM_ROWS + i;
            ovalid[dr] = valid[sr];
            oback[dr] = back[sr];
            ofwd[dr] = fwd[sr];
            if (i == 0) otarget[s] = target[idx[s]];
        }
    }
}

__global__ void __launch_bounds__(256) k_cm_zero(float* __restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.f;
}

// ------------------------------------------------------------------ order ranking
// Model inputs of the ranking pass, built on the device.  Sample s is candidate k = s % K
// of batch query i = s / K: k = 0 the RI order and k = 1 the GQL order (given), k >= 2 a
// seeded random connected order (start at a uniform vertex, then a uniform unvisited
// neighbour of the visited set).  One wave per sample, lane = order position; the features
// are the host's (sgm_costmodel_host.cpp): deg/8, log2(1+|C|)/16, log2(1+mean degree)/16,
// i/32, backward/deg, backward/8, [i == 0], n/32.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t nth_bit(uint32_t x, uint32_t r)
{
    for (uint32_t t = 0; t < r; t++) x &= x - 1;
    return (uint32_t)__ffs(x) - 1u;
}

__global__ void __launch_bounds__(256) k_cm_rank_inputs(const sgm_qdesc* __restrict__ qs, const uint32_t* __restrict__ cnt,
                                                        const unsigned long long* __restrict__ degsum,
                                                        const uint8_t* __restrict__ given, uint32_t nq, uint32_t K,
                                                        uint64_t seed, uint8_t* __restrict__ orders,
                                                        uint16_t* __restrict__ F, uint8_t* __restrict__ valid,
                                                        uint32_t* __restrict__ back, uint32_t* __restrict__ fwd)
{
    __shared__ uint8_t s_ord[4][CM_ROWS];
    __shared__ uint8_t s_pos[4][CM_ROWS];
    const uint32_t lane = lane_id(), wv = (threadIdx.x >> 6) & 3u;
    const uint32_t smp = blockIdx.x * 4 + wv;
    if (smp >= nq * K) return;  // whole wave
    const uint32_t i = smp / K, k = smp % K;
    const sgm_qdesc& d = qs[i];
    const uint32_t n = d.n;
    if (lane == 0) {
        if (k < 2) {
            for (uint32_t j = 0; j < n; j++) s_ord[wv][j] = given[((size_t)i * 2 + k) * CM_ROWS + j];
        } else {
            uint64_t st = seed ^ ((uint64_t)i << 32) ^ (k * 0x9E3779B97F4A7C15ull);
            uint32_t visited = 0, frontier = 0;
            const uint32_t all = n == 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
            for (uint32_t j = 0; j < n; j++) {
                const uint32_t pool = frontier ? frontier : (all & ~visited);
                st += 0x9E3779B97F4A7C15ull;
                const uint32_t r = (uint32_t)(((mix64(st) >> 32) * (uint64_t)__popc(pool)) >> 32);
                const uint32_t u = nth_bit(pool, r);
                s_ord[wv][j] = (uint8_t)u;
                visited |= 1u << u;
                frontier = (frontier | d.nbr[u]) & ~visited;
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t u = lane < n ? s_ord[wv][lane] : 0u;
    if (lane < n) s_pos[wv][u] = (uint8_t)lane;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane >= CM_ROWS) return;
    const size_t row = (size_t)smp * CM_ROWS + lane;
    uint32_t bm = 0, fm = 0;
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (lane < n) {
        uint32_t nb = d.nbr[u];
        while (nb) {
            const uint32_t w = (uint32_t)__ffs(nb) - 1u;
            nb &= nb - 1u;
            const uint32_t p = s_pos[wv][w];
            if (p < lane) bm |= 1u << p;
            else fm |= 1u << p;
        }
        const float deg = (float)d.qdeg[u];
        const float c = (float)cnt[(size_t)i * SGM_MAXQ + u];
        const float mean = c > 0.f ? (float)degsum[(size_t)i * SGM_MAXQ + u] / c : 0.f;
        const float nbk = (float)__popc(bm);
        f[0] = deg / 8.f;
        f[1] = log2f(1.f + c) / 16.f;
        f[2] = log2f(1.f + mean) / 16.f;
        f[3] = (float)lane / 32.f;
        f[4] = deg > 0.f ? nbk / deg : 0.f;
        f[5] = nbk / 8.f;
        f[6] = lane == 0 ? 1.f : 0.f;
        f[7] = (float)n / 32.f;
    }
#pragma unroll
    for (int j = 0; j < CM_F; j++) F[row * CM_F + j] = j < 8 ? f2bf(f[j]) : (uint16_t)0;
    valid[row] = lane < n;
    back[row] = bm;
    fwd[row] = fm;
    orders[row] = (uint8_t)u;
}

// Per query: the best-scored candidate, kept only if it beats the RI prediction (candidate
// 0) by more than `margin`; its order goes to out[i*32 ..].
__global__ void __launch_bounds__(256) k_cm_choose(const float* __restrict__ y, const uint8_t* __restrict__ orders,
                                                   uint32_t nq, uint32_t K, float margin, uint32_t* __restrict__ out,
                                                   float* __restrict__ pbest, float* __restrict__ pri,
                                                   int32_t* __restrict__ learned)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const float* yi = y + (size_t)i * K;
    uint32_t best = 0;
    for (uint32_t k = 1; k < K; k++)
        if (yi[k] < yi[best]) best = k;
    const bool use = yi[best] < yi[0] - margin;
    const uint32_t pick = use ? best : 0u;
    for (uint32_t j = 0; j < CM_ROWS; j++) out[(size_t)i * CM_ROWS + j] = orders[((size_t)i * K + pick) * CM_ROWS + j];
    pbest[i] = yi[pick];
    pri[i] = yi[0];
    learned[i] = use ? 1 : 0;
}

// ------------------------------------------------------------------ launch wrappers
static inline uint32_t grid_for(size_t work, uint32_t per_block, uint32_t cap)
{
    size_t b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (uint32_t)b;
}

void cm_gemm(hipStream_t s, const uint16_t* A, uint32_t lda, const uint16_t* B, uint32_t ldb, const float* bias,
             const uint8_t* valid, int relu, float* zf, uint32_t ldz, uint16_t* hb, uint32_t ldh, float* cf,
             uint32_t ldc, uint32_t M, uint32_t N, uint32_t K)
{
    const uint32_t tiles = (M / 32) * (N / 32);
    hipLaunchKernelGGL(k_cm_gemm, dim3(grid_for(tiles, 4, 16384)), dim3(256), 0, s, A, lda, B, ldb, bias, valid, relu,
                       zf, ldz, hb, ldh, cf, ldc, M, N, K);
}

void cm_gemm_tn(hipStream_t s, const uint16_t* A, uint32_t lda, const uint16_t* B, uint32_t ldb, float* out,
                uint32_t R, uint32_t Ka, uint32_t N)
{
    const uint32_t chunk = 512;
    const uint32_t tiles = (Ka / 32) * (N / 32);
    dim3 grid(grid_for(tiles, 4, 64), (R + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_cm_gemm_tn, grid, dim3(256), 0, s, A, lda, B, ldb, out, R, Ka, N, chunk);
}

void cm_aggregate(hipStream_t s, uint16_t* X, const uint32_t* back, const uint32_t* fwd, uint32_t R)
{
    hipLaunchKernelGGL(k_cm_aggregate, dim3(grid_for((size_t)R * 64, 256, 8192)), dim3(256), 0, s, X, back, fwd, R);
}

void cm_aggregate_bwd(hipStream_t s, const float* dX, const uint32_t* back, const uint32_t* fwd, const uint32_t* back_t,
                      const uint32_t* fwd_t, float* dH, uint32_t R)
{
    hipLaunchKernelGGL(k_cm_aggregate_bwd, dim3(grid_for((size_t)R * 64, 256, 8192)), dim3(256), 0, s, dX, back, fwd,
                       back_t, fwd_t, dH, R);
}

void cm_act_bwd(hipStream_t s, const float* dH, const float* Z, const uint8_t* valid, uint16_t* dZ, float* db,
                uint32_t R)
{
    hipLaunchKernelGGL(k_cm_act_bwd, dim3(grid_for((size_t)R * 64, 256, 2048)), dim3(256), 0, s, dH, Z, valid, dZ, db, R);
}

void cm_readout(hipStream_t s, const float* H3, const uint8_t* valid, const float* wo, const float* bo, float* y,
                const float* target, float* dH3, float* dwo, float* dbo, float* loss, uint32_t S)
{
    hipLaunchKernelGGL(k_cm_readout, dim3(grid_for(S, 4, 4096)), dim3(256), 0, s, H3, valid, wo, bo, y, target, dH3,
                       dwo, dbo, loss, S);
}

void cm_adam(hipStream_t s, float* p, const float* g, float* m, float* v, uint32_t n, float lr, uint32_t step)
{
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f, c2 = 1.f;
    for (uint32_t i = 0; i < step; i++) { c1 *= b1; c2 *= b2; }
    hipLaunchKernelGGL(k_cm_adam, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, 1e-8f,
                       1.f - c1, 1.f - c2);
}

void cm_cast(hipStream_t s, const float* w, uint32_t rows, uint32_t cols, uint16_t* wb, uint16_t* wt)
{
    hipLaunchKernelGGL(k_cm_cast, dim3(grid_for((size_t)rows * cols, 256, 1024)), dim3(256), 0, s, w, rows, cols, wb, wt);
}

void cm_gather(hipStream_t s, const uint32_t* idx, uint32_t S, const uint16_t* F, const uint8_t* valid,
               const uint32_t* back, const uint32_t* fwd, const float* target, uint16_t* oF, uint8_t* ovalid,
               uint32_t* oback, uint32_t* ofwd, float* otarget)
{
    hipLaunchKernelGGL(k_cm_gather, dim3(grid_for((size_t)S * CM_ROWS * CM_F, 256, 8192)), dim3(256), 0, s, idx, S, F,
                       valid, back, fwd, target, oF, ovalid, oback, ofwd, otarget);
}

void cm_rank_inputs(hipStream_t s, const sgm_qdesc* qs, const uint32_t* cnt, const unsigned long long* degsum,
                    const uint8_t* given, uint32_t nq, uint32_t K, uint64_t seed, uint8_t* orders, uint16_t* F,
                    uint8_t* valid, uint32_t* back, uint32_t* fwd)
{
    const uint32_t S = nq * K;
    hipLaunchKernelGGL(k_cm_rank_inputs, dim3((S + 3) / 4), dim3(256), 0, s, qs, cnt, degsum, given, nq, K, seed,
                       orders, F, valid, back, fwd);
}

void cm_choose(hipStream_t s, const float* y, const uint8_t* orders, uint32_t nq, uint32_t K, float margin,
               uint32_t* out, float* pbest, float* pri, int32_t* learned)
{
    hipLaunchKernelGGL(k_cm_choose, dim3((nq + 255) / 256), dim3(256), 0, s, y, orders, nq, K, margin, out, pbest,
                       pri, learned);
}

void cm_zero(hipStream_t s, float* p, size_t n)
{
    hipLaunchKernelGGL(k_cm_zero, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n);
}

}  // namespace sgm
