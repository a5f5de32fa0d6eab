// Launch wrappers of the cost-model kernels (sgm_costmodel.hip); host driver in
// sgm_costmodel_host.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "sgm_internal.h"

#define CM_ROWS 32   // rows per sample (query vertices, padded)
#define CM_F 32      // input feature columns (8 used, zero padded to one MFMA tile)
#define CM_H 64      // hidden width
#define CM_X 192     // layer input: [H | backward mean | forward mean]

namespace sgm {
void cm_gemm(hipStream_t s, const uint16_t* A, uint32_t lda, const uint16_t* B, uint32_t ldb, const float* bias,
             const uint8_t* valid, int relu, float* zf, uint32_t ldz, uint16_t* hb, uint32_t ldh, float* cf,
             uint32_t ldc, uint32_t M, uint32_t N, uint32_t K);
void cm_gemm_tn(hipStream_t s, const uint16_t* A, uint32_t lda, const uint16_t* B, uint32_t ldb, float* out,
                uint32_t R, uint32_t Ka, uint32_t N);
void cm_aggregate(hipStream_t s, uint16_t* X, const uint32_t* back, const uint32_t* fwd, uint32_t R);
void cm_aggregate_bwd(hipStream_t s, const float* dX, const uint32_t* back, const uint32_t* fwd,
                      const uint32_t* back_t, const uint32_t* fwd_t, float* dH, uint32_t R);
void cm_act_bwd(hipStream_t s, const float* dH, const float* Z, const uint8_t* valid, uint16_t* dZ, float* db,
                uint32_t R);
void cm_readout(hipStream_t s, const float* H3, const uint8_t* valid, const float* wo, const float* bo, float* y,
                const float* target, float* dH3, float* dwo, float* dbo, float* loss, uint32_t S);
void cm_adam(hipStream_t s, float* p, const float* g, float* m, float* v, uint32_t n, float lr, uint32_t step);
void cm_cast(hipStream_t s, const float* w, uint32_t rows, uint32_t cols, uint16_t* wb, uint16_t* wt);
void cm_zero(hipStream_t s, float* p, size_t n);
// ranking: model inputs for K candidates of each batch query (RI, GQL given; sampled), then
// the per-query choice (RI unless another candidate is predicted lower by > margin)
void cm_rank_inputs(hipStream_t s, const sgm_qdesc* qs, const uint32_t* cnt, const unsigned long long* degsum,
                    const uint8_t* given, uint32_t nq, uint32_t K, uint64_t seed, uint8_t* orders, uint16_t* F,
                    uint8_t* valid, uint32_t* back, uint32_t* fwd);
void cm_choose(hipStream_t s, const float* y, const uint8_t* orders, uint32_t nq, uint32_t K, float margin,
               uint32_t* out, float* pbest, float* pri, int32_t* learned);
void cm_gather(hipStream_t s, const uint32_t* idx, uint32_t S, const uint16_t* F, const uint8_t* valid,
               const uint32_t* back, const uint32_t* fwd, const float* target, uint16_t* oF, uint8_t* ovalid,
               uint32_t* oback, uint32_t* ofwd, float* otarget);
}  // namespace sgm
