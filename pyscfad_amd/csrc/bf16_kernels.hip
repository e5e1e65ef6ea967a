// Split-BF16x3 contraction engine for gfx950.
//
// Every FP64 operand is split once into three bf16 pieces x ~ hi + mid + lo
// (hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)), packed in one
// 8-byte element (bits 0-15 hi, 16-31 mid, 32-47 lo).  The packed arrays keep
// the FP64 arrays' element layout, so the contraction descriptors are the same
// as for the FP64 engine.  A product is evaluated as the six bf16 MFMA terms
//   hi*hi + (hi*mid + mid*hi) + (hi*lo + lo*hi + mid*mid)
// (dropped: mid*lo, lo*mid, lo*lo, all below 2^-24 relative) with
// v_mfma_f32_32x32x16_bf16 into two FP32 tile accumulators (the large and the
// small terms), which are added into FP64 accumulators after every K chunk of
// 16: FP32 accumulation per tile, FP64 across tiles and chunks.
#include "ccsdt_common.h"
#include "engine_staging.h"
#include "bf16_split.h"

namespace pamd {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ void k_split_bf16x3(const double *__restrict__ x, size_t n,
                               unsigned long long *__restrict__ out) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    out[t] = split3(x[t]);
}

hipError_t launch_split_bf16x3(const double *x, size_t n, unsigned long long *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  hipLaunchKernelGGL(k_split_bf16x3, dim3((unsigned)b), dim3(256), 0, s, x, n, out);
  return hipGetLastError();
}

// Gather the three bf16 planes of 8 packed elements (one per k) into MFMA
// operand vectors.
__device__ __forceinline__ void unpack8(const unsigned long long (&u)[8], bf16x8 &h, bf16x8 &m,
                                        bf16x8 &l) {
  u32x4 vh, vm, vl;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const unsigned int L0 = (unsigned int)u[2 * d], L1 = (unsigned int)u[2 * d + 1];
    const unsigned int H0 = (unsigned int)(u[2 * d] >> 32), H1 = (unsigned int)(u[2 * d + 1] >> 32);
    vh[d] = (L0 & 0xFFFFu) | (L1 << 16);
    vm[d] = (L0 >> 16) | (L1 & 0xFFFF0000u);
    vl[d] = (H0 & 0xFFFFu) | (H1 << 16);
  }
  h = __builtin_bit_cast(bf16x8, vh);
  m = __builtin_bit_cast(bf16x8, vm);
  l = __builtin_bit_cast(bf16x8, vl);
}

// 8 consecutive-k packed elements of one operand row from the staged panel
template <int R, int NTHR>
__device__ __forceinline__ void read8(const unsigned long long *S, int rfast, int r, int k0,
                                      unsigned long long (&u)[8]) {
  using ST = OpStage<R, NTHR>;
  if (rfast) {
#pragma unroll
    for (int e = 0; e < 8; ++e) u[e] = S[(k0 + e) * ST::LR + r];
  } else {
    const u32x4 *p = (const u32x4 *)(S + r * ST::LK + k0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const u32x4 w = p[q];
      u[2 * q] = (unsigned long long)w[0] | ((unsigned long long)w[1] << 32);
      u[2 * q + 1] = (unsigned long long)w[2] | ((unsigned long long)w[3] << 32);
    }
  }
}

// Tile MT x NT with WM x WN waves of 32 x 32 (one 32x32x16 MFMA tile each).
template <int MT, int NT, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_bf16x3(GemmDesc d, int n_fast) {
  constexpr int NTHR = 64 * WM * WN;
  static_assert(MT == 32 * WM && NT == 32 * WN, "one 32x32 MFMA tile per wave");
  using SA = OpStage<MT, NTHR>;
  using SB = OpStage<NT, NTHR>;
  typedef unsigned long long u64;
  __shared__ __align__(16) u64 As[2][SA::LDS];
  __shared__ __align__(16) u64 Bs[2][SB::LDS];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int mt = (d.M + MT - 1) / MT, nt = (d.N + NT - 1) / NT;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int split = logical / (mt * nt);
  const int tile = logical - split * (mt * nt);
  const int tm = n_fast ? tile / nt : tile % mt;
  const int tn = n_fast ? tile % nt : tile / mt;
  const int m0 = tm * MT, n0 = tn * NT;
  const int total = d.chunk_start[d.nseg];
  const int cps = (total + d.splits - 1) / d.splits;
  const int g0 = split * cps;
  const int g1 = min(total, g0 + cps);

  double acc[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0;

  int fs = 0;
  while (fs + 1 < d.nseg && g0 >= d.chunk_start[fs + 1]) ++fs;
  OpDesc FA = d.seg[fs].A, FB = d.seg[fs].B;
  int FL = d.seg[fs].L, Fend = d.chunk_start[fs + 1];
  double Falpha = d.seg[fs].alpha;
  int kappa = (g0 - d.chunk_start[fs]) * GKC;
  KPos ka{kappa / FA.kdiv, kappa - (kappa / FA.kdiv) * FA.kdiv};
  KPos kb{kappa / FB.kdiv, kappa - (kappa / FB.kdiv) * FB.kdiv};
  int ra_off[SA::E], rb_off[SB::E];
  bool ra_ok[SA::E], rb_ok[SB::E];
  u64 va[SA::E], vb[SB::E];
  bool oka[SA::E], okb[SB::E];
  op_rows<MT, NTHR>(FA, d.M, m0, tid, ra_off, ra_ok);
  op_rows<NT, NTHR>(FB, d.N, n0, tid, rb_off, rb_ok);
  if (g0 < g1) {
    op_fetch<MT, NTHR>(FA, FL, kappa, ka, tid, ra_off, ra_ok, va, oka);
    op_fetch<NT, NTHR>(FB, FL, kappa, kb, tid, rb_off, rb_ok, vb, okb);
  }
  int la = FA.rfast, lb = FB.rfast;
  double alpha = Falpha;
  int buf = 0;
  const int ar = wm * 32 + (lane & 31), br = wn * 32 + (lane & 31), k0 = 8 * (lane >> 5);
  for (int g = g0; g < g1; ++g) {
    op_stash<MT, NTHR>(la, alpha, tid, va, oka, As[buf]);
    op_stash<NT, NTHR>(lb, 1.0, tid, vb, okb, Bs[buf]);
    __syncthreads();
    const int cla = la, clb = lb;
    if (g + 1 < g1) {
      if (g + 1 >= Fend) {
        ++fs;
        FA = d.seg[fs].A;
        FB = d.seg[fs].B;
        FL = d.seg[fs].L;
        Fend = d.chunk_start[fs + 1];
        Falpha = d.seg[fs].alpha;
        kappa = 0;
        ka = KPos{0, 0};
        kb = KPos{0, 0};
        op_rows<MT, NTHR>(FA, d.M, m0, tid, ra_off, ra_ok);
        op_rows<NT, NTHR>(FB, d.N, n0, tid, rb_off, rb_ok);
      } else {
        kappa += GKC;
        ka.r += GKC;
        while (ka.r >= FA.kdiv) {
          ka.r -= FA.kdiv;
          ++ka.q;
        }
        kb.r += GKC;
        while (kb.r >= FB.kdiv) {
          kb.r -= FB.kdiv;
          ++kb.q;
        }
      }
      op_fetch<MT, NTHR>(FA, FL, kappa, ka, tid, ra_off, ra_ok, va, oka);
      op_fetch<NT, NTHR>(FB, FL, kappa, kb, tid, rb_off, rb_ok, vb, okb);
      la = FA.rfast;
      lb = FB.rfast;
      alpha = Falpha;
    }
    u64 ua[8], ub[8];
    read8<MT, NTHR>(As[buf], cla, ar, k0, ua);
    read8<NT, NTHR>(Bs[buf], clb, br, k0, ub);
    bf16x8 ah, am, al, bh, bm, bl;
    unpack8(ua, ah, am, al);
    unpack8(ub, bh, bm, bl);
    f32x16 big, small;
#pragma unroll
    for (int r = 0; r < 16; ++r) big[r] = small[r] = 0.f;
    // two independent chains, smallest terms first within each
    small = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm, small, 0, 0, 0);
    big = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm, big, 0, 0, 0);
    small = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, small, 0, 0, 0);
    big = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh, big, 0, 0, 0);
    small = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, small, 0, 0, 0);
    big = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, big, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] += (double)big[r] + (double)small[r];
    buf ^= 1;
  }

  // D layout of v_mfma_f32_32x32x16_bf16 on gfx950: lane l, element r holds
  // row 8*(r/4) + 4*(l/32) + r%4, column l%32.
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int n = n0 + wn * 32 + (lane & 31);
    if (m >= d.M || n >= d.N) continue;
    const double val = acc[r];
    if (d.mode == kPartial) {
      d.partial[((size_t)split * d.M + m) * d.N + n] = val;
    } else {
      const int qm = m / d.C.mdiv, rm = m - qm * d.C.mdiv;
      const int qn = n / d.C.ndiv, rn = n - qn * d.C.ndiv;
      double *c = d.C.p + (long long)qm * d.C.mx + (long long)rm * d.C.my +
                  (long long)qn * d.C.nx + (long long)rn * d.C.ny;
      if (d.mode == kAccumulate)
        *c += val;
      else
        *c = val;
    }
  }
}

hipError_t launch_gemm_bf16x3(const GemmDesc &d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0) return hipSuccess;
  if (d.M >= d.N) {
    const long long tiles = (long long)((d.M + 127) / 128) * ((d.N + 63) / 64) * d.splits;
    hipLaunchKernelGGL((k_gemm_bf16x3<128, 64, 4, 2>), dim3((unsigned)tiles), dim3(512), 0, s, d,
                       1);
  } else {
    const long long tiles = (long long)((d.M + 63) / 64) * ((d.N + 127) / 128) * d.splits;
    hipLaunchKernelGGL((k_gemm_bf16x3<64, 128, 2, 4>), dim3((unsigned)tiles), dim3(512), 0, s, d,
                       0);
  }
  return hipGetLastError();
}

}  // namespace pamd
