// Kernels of the MP2 pair energies and of the DF integral producer that are
// not contractions: the MP2 amplitude/pair-energy pass, the packing of the
// virtual-virtual DF factors and the scatter of their cotangent.
#include <hip/hip_runtime.h>
#include "eris_common.h"

namespace pamd {

// ---------------------------------------------------------------------------
// MP2 amplitudes and pair energies (pyscfad/cc/ccsd.py:364-381).
// One workgroup per unordered pair i <= j streams the two (a, b) planes
// ovov[i,:,j,:] and ovov[j,:,i,:] once, forms both amplitude planes
// t2[i,j] and t2[j,i], and reduces
//   e_ij = sum_ab (2 t2[i,j,a,b] - t2[j,i,a,b]) ovov[i,a,j,b]
//   e_ji = sum_ab (2 t2[j,i,a,b] - t2[i,j,a,b]) ovov[j,a,i,b]
// in a fixed order (per-thread strided sums, then a tree over the workgroup).
// ---------------------------------------------------------------------------
constexpr int kMp2Threads = 512;

__global__ void __launch_bounds__(kMp2Threads) k_mp2_pairs(int o, int v, const double *mo,
                                                           const double *ovov, double *t2,
                                                           double *e_pair) {
  __shared__ double red[2][kMp2Threads / 64];
  // pair index q -> (i, j), j >= i, q = j (j + 1) / 2 + i
  const int q = blockIdx.x;
  int j = 0;
  while ((j + 1) * (j + 2) / 2 <= q) ++j;
  const int i = q - j * (j + 1) / 2;
  const size_t ov = (size_t)o * v, vv = (size_t)v * v;
  const double ei = mo[i], ej = mo[j];
  double sij = 0.0, sji = 0.0;
  for (size_t x = threadIdx.x; x < vv; x += kMp2Threads) {
    const int a = (int)(x / v), b = (int)(x - (size_t)a * v);
    const double ea = mo[o + a], eb = mo[o + b];
    // eia[x,a] + eia[y,b], grouped as the reference forms the denominator of
    // t2[x,y,a,b] (the two planes differ in rounding, not in value)
    const double dij = (ei - ea) + (ej - eb);
    const double dji = (ej - ea) + (ei - eb);
    const double gij = ovov[((size_t)i * v + a) * ov + (size_t)j * v + b];
    const double gji = ovov[((size_t)j * v + a) * ov + (size_t)i * v + b];
    const double tij = gij / dij;
    const double tji = gji / dji;
    sij += (2.0 * tij - tji) * gij;
    sji += (2.0 * tji - tij) * gji;
    if (t2) {
      t2[((size_t)i * o + j) * vv + x] = tij;
      t2[((size_t)j * o + i) * vv + x] = tji;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sij += __shfl_xor(sij, off, 64);
    sji += __shfl_xor(sji, off, 64);
  }
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = sij;
    red[1][w] = sji;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int u = 0; u < kMp2Threads / 64; ++u) {
      a += red[0][u];
      b += red[1][u];
    }
    e_pair[(size_t)i * o + j] = a;
    if (i != j) e_pair[(size_t)j * o + i] = b;
  }
}

hipError_t launch_mp2_pairs(int o, int v, const double *mo, const double *ovov, double *t2,
                            double *e_pair, hipStream_t s) {
  const unsigned pairs = (unsigned)((long long)o * (o + 1) / 2);
  hipLaunchKernelGGL(k_mp2_pairs, dim3(pairs), dim3(kMp2Threads), 0, s, o, v, mo, ovov, t2,
                     e_pair);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Lvv[P, b(b+1)/2 + c] = Lpq[P, o+b, o+c]  (b >= c): lib.pack_tril over the
// virtual block, one workgroup per (b, P) row.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_pack_lvv(const double *Lpq, int o, int n, int v,
                                                 double *Lvv) {
  const int b = blockIdx.x, P = blockIdx.y;
  const size_t np = (size_t)v * (v + 1) / 2;
  const double *src = Lpq + ((size_t)P * n + o + b) * n + o;
  double *dst = Lvv + (size_t)P * np + (size_t)b * (b + 1) / 2;
  for (int c = threadIdx.x; c <= b; c += 256) dst[c] = src[c];
}

hipError_t launch_pack_lvv(const double *Lpq, int o, int n, int naux, double *Lvv,
                           hipStream_t s) {
  const int v = n - o;
  hipLaunchKernelGGL(k_pack_lvv, dim3(v, naux), dim3(256), 0, s, Lpq, o, n, v, Lvv);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Adjoint of the packing: Lpq_bar[P, o+b, o+c] = Lvv_bar[P, b(b+1)/2 + c] for
// b >= c and 0 for b < c (the reference's pack_tril adjoint), and the vo
// block Lpq_bar[P, o+b, :o] = 0.  Writes the whole row o+b of every P.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_unpack_lvv_bar(const double *Lvv_bar, int o, int n,
                                                       int v, double *Lpq_bar) {
  const int b = blockIdx.x, P = blockIdx.y;
  const size_t np = (size_t)v * (v + 1) / 2;
  const double *src = Lvv_bar + (size_t)P * np + (size_t)b * (b + 1) / 2;
  double *dst = Lpq_bar + ((size_t)P * n + o + b) * n;
  for (int x = threadIdx.x; x < n; x += 256) {
    const int c = x - o;
    dst[x] = (c >= 0 && c <= b) ? src[c] : 0.0;
  }
}

hipError_t launch_unpack_lvv_bar(const double *Lvv_bar, int o, int n, int naux, double *Lpq_bar,
                                 hipStream_t s) {
  const int v = n - o;
  hipLaunchKernelGGL(k_unpack_lvv_bar, dim3(v, naux), dim3(256), 0, s, Lvv_bar, o, n, v, Lpq_bar);
  return hipGetLastError();
}

// Zero the occupied rows Lpq_bar[P, :o, :] of every P (the oo and ov blocks,
// which the adjoint contractions then accumulate into).
__global__ void __launch_bounds__(256) k_zero_occ_rows(int o, int n, double *Lpq_bar) {
  const int P = blockIdx.y;
  double *dst = Lpq_bar + (size_t)P * n * n;
  const size_t len = (size_t)o * n;
  for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < len; x += (size_t)gridDim.x * 256)
    dst[x] = 0.0;
}

hipError_t launch_zero_occ_rows(int o, int n, int naux, double *Lpq_bar, hipStream_t s) {
  const size_t len = (size_t)o * n;
  unsigned gx = (unsigned)((len + 255) / 256);
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(k_zero_occ_rows, dim3(gx, naux), dim3(256), 0, s, o, n, Lpq_bar);
  return hipGetLastError();
}

}  // namespace pamd
