/*
 * BF16 MLP primitives of libpyscfad_amd_cc.so for the learned (T) screen
 * (no counterpart in the reference).  Device pointers; bf16 arrays are uint16_t
 * bit patterns; FP32 accumulation throughout.  Every call returns a PAMD_*
 * status (include/pyscfad_amd_ccsdt.h).
 */
#ifndef PYSCFAD_AMD_MLP_H
#define PYSCFAD_AMD_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C[M][N] (+)= sum_k A[m][k] B[n][k] (+ bias[n]); A, B bf16 with k contiguous
 * (K, lda, ldb multiples of 8, 16-byte aligned rows), C fp32 with row stride ldc. */
int pamd_bf16_gemm_tn(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                      const float *bias, float *C, int ldc, int accumulate, void *stream);

/* Y = act(Z) (act 0: identity, 1: SiLU) as bf16, row-major Y[M][N] and/or
 * transposed Yt[N][M] (either may be NULL). */
int pamd_bf16_act_fwd(int M, int N, const float *Z, int act, uint16_t *Y, uint16_t *Yt,
                      void *stream);

/* dZ = dY * act'(Z) as bf16 (row-major and transposed), db[n] = sum_m dZ[m][n]. */
int pamd_bf16_act_bwd(int M, int N, const float *Z, const float *dY, int act, uint16_t *dZ,
                      uint16_t *dZt, float *db, void *stream);

/* Adam step (decoupled weight decay wd) on FP32 masters w[R][C]; refreshes the
 * bf16 copies wb[R][C] and wt[C][R] when given.  step counts from 1. */
int pamd_adam_step(int R, int C, float *w, const float *g, float *m1, float *m2, uint16_t *wb,
                   uint16_t *wt, float lr, float beta1, float beta2, float eps, float wd, int step,
                   void *stream);

/* loss = mean (pred - y)^2 over n values; dpred = 2 (pred - y) / n (either may be NULL). */
int pamd_mse(int n, const float *pred, const float *y, float *dpred, float *loss, void *stream);

/* y = bf16(x) for n = M*N values, optionally also transposed into yt[N][M]. */
int pamd_f32_to_bf16(size_t n, const float *x, uint16_t *y, int M, int N, uint16_t *yt,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif
