// C ABI of the BF16 MLP primitives (include/pyscfad_amd_mlp.h).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/pyscfad_amd_ccsdt.h"
#include "../../include/pyscfad_amd_mlp.h"

namespace pamd {
hipError_t launch_gemm_bf16_tn(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B,
                               int ldb, const float *bias, float *C, int ldc, int accumulate,
                               hipStream_t s);
hipError_t launch_act_out(int M, int N, const float *Z, int act, uint16_t *Y, uint16_t *Yt,
                          hipStream_t s);
hipError_t launch_act_back(int M, int N, const float *Z, const float *dY, int act, uint16_t *dZ,
                           uint16_t *dZt, float *db, hipStream_t s);
hipError_t launch_adam(int R, int Ccols, float *w, const float *g, float *m1, float *m2, uint16_t *wb,
                       uint16_t *wt, float lr, float b1, float b2, float eps, float wd, int step,
                       hipStream_t s);
hipError_t launch_mse(int n, const float *pred, const float *y, float *dpred, float *loss,
                      hipStream_t s);
hipError_t launch_f32_to_bf16(size_t n, const float *x, uint16_t *y, int M, int N, uint16_t *yt,
                              hipStream_t s);
}  // namespace pamd

using namespace pamd;

static int status(hipError_t e) { return e == hipSuccess ? PAMD_OK : PAMD_ERR_HIP; }
static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

extern "C" {

int pamd_bf16_gemm_tn(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                      const float *bias, float *C, int ldc, int accumulate, void *stream) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || !C) return PAMD_ERR_ARG;
  if (K % 8 || lda % 8 || ldb % 8 || lda < K || ldb < K || ldc < N) return PAMD_ERR_ARG;
  if (!aligned16(A) || !aligned16(B)) return PAMD_ERR_ARG;
  return status(launch_gemm_bf16_tn(M, N, K, A, lda, B, ldb, bias, C, ldc, accumulate,
                                    (hipStream_t)stream));
}

int pamd_bf16_act_fwd(int M, int N, const float *Z, int act, uint16_t *Y, uint16_t *Yt,
                      void *stream) {
  if (M < 1 || N < 1 || !Z || (act != 0 && act != 1)) return PAMD_ERR_ARG;
  return status(launch_act_out(M, N, Z, act, Y, Yt, (hipStream_t)stream));
}

int pamd_bf16_act_bwd(int M, int N, const float *Z, const float *dY, int act, uint16_t *dZ,
                      uint16_t *dZt, float *db, void *stream) {
  if (M < 1 || N < 1 || !Z || !dY || !dZ || !dZt || !db || (act != 0 && act != 1))
    return PAMD_ERR_ARG;
  return status(launch_act_back(M, N, Z, dY, act, dZ, dZt, db, (hipStream_t)stream));
}

int pamd_adam_step(int R, int C, float *w, const float *g, float *m1, float *m2, uint16_t *wb,
                   uint16_t *wt, float lr, float beta1, float beta2, float eps, float wd, int step,
                   void *stream) {
  if (R < 1 || C < 1 || !w || !g || !m1 || !m2 || step < 1) return PAMD_ERR_ARG;
  return status(launch_adam(R, C, w, g, m1, m2, wb, wt, lr, beta1, beta2, eps, wd, step,
                            (hipStream_t)stream));
}

int pamd_mse(int n, const float *pred, const float *y, float *dpred, float *loss, void *stream) {
  if (n < 1 || !pred || !y) return PAMD_ERR_ARG;
  return status(launch_mse(n, pred, y, dpred, loss, (hipStream_t)stream));
}

int pamd_f32_to_bf16(size_t n, const float *x, uint16_t *y, int M, int N, uint16_t *yt,
                     void *stream) {
  if (!x || !y || n != (size_t)M * N) return PAMD_ERR_ARG;
  return status(launch_f32_to_bf16(n, x, y, M, N, yt, (hipStream_t)stream));
}

}  // extern "C"
