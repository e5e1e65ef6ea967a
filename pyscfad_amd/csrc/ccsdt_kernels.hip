// HIP kernels for the closed-shell (T) energy and its adjoint on gfx950.
//
// * k_gemm_f64   -- the FP64 MFMA contraction engine (v_mfma_f64_16x16x4f64):
//                   128x128 (or 128x64) output tile per 512-thread workgroup, K
//                   staged through double-buffered LDS in chunks of 16, 4x2 (2x2)
//                   MFMA tiles per wave, XCD-aware tile order.
// * k_orbit      -- fused epilogue for one occupied triple: V = t1*ovov + t2*fvo
//                   terms, Z = r3(W + V), denominators, the energy reduction and,
//                   for the adjoint, Wbar / Vbar / dE/d(mo_energy) seeds.  Virtual
//                   indices are tiled 8^3 and a workgroup owns one orbit of tiles
//                   {(A,B,C) and its permutations}, so the six permuted reads that
//                   r3 needs stay in LDS.
// * k_vr_a/k_vr_b -- contractions of Vbar with t1/fvo/ovov/t2 (the V-term adjoint).
// * k_combine    -- per-triple gather of the adjoint pieces into the cotangents.
// * prepare/fold -- layout changes at entry/exit.
#include <type_traits>
#include <cstdlib>
#include <algorithm>
#include "ccsdt_common.h"
#include "engine_staging.h"
#include "bf16_split.h"

namespace pamd {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// generic helpers
// ---------------------------------------------------------------------------
__global__ void k_fill(double *p, size_t n, double val) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += st) p[t] = val;
}

hipError_t launch_fill(double *p, size_t n, double val, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, val);
  return hipGetLastError();
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// deterministic block sum for 256 threads; scratch >= 4 doubles
__device__ double block_sum256(double x, double *scratch) {
  x = wave_sum(x);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[w] = x;
  __syncthreads();
  double r = scratch[0] + scratch[1] + scratch[2] + scratch[3];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// layout preparation
//   Vp[x][p][q][f]  = ovvv[x,p,f,q]   (f contiguous: the contraction index)
//   T2m[x][p][q][m] = t2[x,m,p,q]
//   OOt[x][y][p][m] = ovoo[x,p,m,y]
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_prep_vp(const double *__restrict__ ovvv, int packed, int o, int v, T Vp) {
  const size_t total = (size_t)o * v * v * v;
  const size_t vv = (size_t)v * v;
  const size_t npair = (size_t)v * (v + 1) / 2;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    size_t f = t % v;
    size_t q = (t / v) % v;
    size_t xp = t / vv;  // x * v + p
    double val;
    if (packed) {
      size_t hi = f > q ? f : q, lo = f > q ? q : f;
      val = ovvv[xp * npair + hi * (hi + 1) / 2 + lo];
    } else {
      val = ovvv[xp * vv + f * v + q];
    }
    put(Vp, t, val);
  }
}

template <class T>
__global__ void k_prep_t2m(const double *__restrict__ t2, int o, int v, T T2m) {
  const size_t total = (size_t)o * o * v * v;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    size_t m = t % o;
    size_t q = (t / o) % v;
    size_t p = (t / ((size_t)o * v)) % v;
    size_t x = t / ((size_t)o * v * v);
    put(T2m, t, t2[((x * o + m) * v + p) * v + q]);
  }
}

template <class T>
__global__ void k_prep_oot(const double *__restrict__ ovoo, int o, int v, T OOt) {
  const size_t total = (size_t)o * o * v * o;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    size_t m = t % o;
    size_t p = (t / o) % v;
    size_t y = (t / ((size_t)o * v)) % o;
    size_t x = t / ((size_t)o * v * o);
    put(OOt, t, ovoo[((x * v + p) * o + m) * o + y]);
  }
}

static unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  if (b == 0) b = 1;
  return (unsigned)b;
}

// T: the store sink (a pointer, or Planar3)
template <class T>
static hipError_t prepare_as(const double *ovvv, int packed, const double *t2, const double *ovoo,
                             int o, int v, T Vp, T T2m, T OOt, hipStream_t s) {
  hipLaunchKernelGGL(k_prep_vp<T>, dim3(grid_for((size_t)o * v * v * v)), dim3(256), 0, s, ovvv,
                     packed, o, v, Vp);
  hipLaunchKernelGGL(k_prep_t2m<T>, dim3(grid_for((size_t)o * o * v * v)), dim3(256), 0, s, t2, o,
                     v, T2m);
  hipLaunchKernelGGL(k_prep_oot<T>, dim3(grid_for((size_t)o * o * v * o)), dim3(256), 0, s, ovoo, o,
                     v, OOt);
  return hipGetLastError();
}

// split == kSplitBf16x3: three bf16 planes (hi, mid, lo) per array, each plane
// as long as the array (in the first 6 of the 8 bytes per element allocated)
hipError_t launch_prepare(const double *ovvv, int packed, const double *t2, const double *ovoo,
                          int o, int v, double *Vp, double *T2m, double *OOt, int split,
                          hipStream_t s) {
  if (split == kSplitBf16x3) {
    const long long n_vp = (long long)o * v * v * v, n_t2m = (long long)o * o * v * v,
                    n_oot = (long long)o * o * v * o;
    return prepare_as<Planar3>(ovvv, packed, t2, ovoo, o, v,
                               Planar3{(unsigned short *)Vp, n_vp},
                               Planar3{(unsigned short *)T2m, n_t2m},
                               Planar3{(unsigned short *)OOt, n_oot}, s);
  }
  if (split == kSplitF32x2)
    return prepare_as<F2 *>(ovvv, packed, t2, ovoo, o, v, (F2 *)Vp, (F2 *)T2m, (F2 *)OOt, s);
  return prepare_as<double *>(ovvv, packed, t2, ovoo, o, v, Vp, T2m, OOt, s);
}

__global__ void k_split_f32x2(const double *__restrict__ x, size_t n, F2 *__restrict__ out) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    out[t] = split2(x[t]);
}

hipError_t launch_split_f32x2(const double *x, size_t n, unsigned long long *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  hipLaunchKernelGGL(k_split_f32x2, dim3((unsigned)b), dim3(256), 0, s, x, n, (F2 *)out);
  return hipGetLastError();
}

// Pair-symmetrised blocks, dst[x][y][p][q] = src(x,p,y,q) + src(y,q,x,p), for
//   ovov (kind 0): src(x,p,y,q) = ovov[x,p,y,q]  ->  OVS
//   t2   (kind 1): src(x,p,y,q) = t2[x,y,p,q]    ->  T2S
// so that the V terms of the epilogue and of its adjoint read them contiguous
// in the fastest virtual index.  One thread per output element; the second
// read is strided (a one-time pass).
__global__ void k_prep_sym(const double *__restrict__ src, int o, int v, int kind,
                           double *__restrict__ dst) {
  const size_t n = (size_t)o * o * v * v;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    const int q = (int)(t % v);
    const size_t r = t / v;
    const int p = (int)(r % v);
    const size_t xy = r / v;
    const int y = (int)(xy % o), x = (int)(xy / o);
    double a, b;
    if (kind == 0) {
      a = src[(((size_t)x * v + p) * o + y) * v + q];
      b = src[(((size_t)y * v + q) * o + x) * v + p];
    } else {
      a = src[(((size_t)x * o + y) * v + p) * v + q];
      b = src[(((size_t)y * o + x) * v + q) * v + p];
    }
    dst[t] = a + b;
  }
}

hipError_t launch_prep_sym(const double *src, int o, int v, int kind, double *dst, hipStream_t s) {
  const size_t n = (size_t)o * o * v * v;
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  hipLaunchKernelGGL(k_prep_sym, dim3((unsigned)b), dim3(256), 0, s, src, o, v, kind, dst);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 MFMA contraction engine
//
// Workgroup tile MT x NT: 256 threads = 4 waves in a 2 x 2 grid, each wave
// computing (MT/2) x (NT/2) as (MT/32) x (NT/32) tiles of v_mfma_f64_16x16x4f64.
// K is staged through double-buffered LDS in chunks of 16.  Each operand of
// each segment is staged in the LDS layout that keeps its global reads
// coalesced and its LDS traffic free of bank conflicts:
//   row-contiguous operand ("rfast") -> [k][r], row stride R + 16 doubles
//   K-contiguous operand   ("kfast") -> [r][k], row stride 17 doubles
// Offsets are 32-bit, relative to the segment's base pointer (the host folds
// the occupied-index slab into the pointer and checks the range); row offsets
// are computed once per segment and K offsets advance without divisions.
// ---------------------------------------------------------------------------
// Tile MT x NT with WM x WN waves; each wave (MT/WM) x (NT/WN).  The grid is
// 1-D over (split, tile) with the N tiles innermost when n_fast (the A panel
// is then read once and reused from L2), else the M tiles innermost.
// FX2: the FP32 hi/lo comparison form.  Elements are F2 pairs (hi, lo); each
// MFMA step runs three v_mfma_f32_16x16x4f32 (hi*hi, hi*lo, lo*hi) into FP32
// tile sums, added into the FP64 accumulators every kFlush2 chunks (128 k).
// The FP32 MFMA's D layout differs from the FP64 one (rows 4*(l/16) + r).
template <int MT, int NT, int WM, int WN, bool FX2>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_f64(GemmDesc d, GemmBatch bt, int n_fast) {
  constexpr int NTHR = 64 * WM * WN;
  using E = typename std::conditional<FX2, F2, double>::type;
  constexpr int kFlush2 = 8;
  using SA = OpStage<MT, NTHR>;
  using SB = OpStage<NT, NTHR>;
  constexpr int WTM = MT / WM, WTN = NT / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  __shared__ __align__(16) double As[2][SA::LDS];
  __shared__ __align__(16) double Bs[2][SB::LDS];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int mt = (d.M + MT - 1) / MT, nt = (d.N + NT - 1) / NT;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int per = mt * nt * d.splits;  // workgroups of one product of the batch
  const int batch = logical / per;
  const int rem = logical - batch * per;
  const int split = rem / (mt * nt);
  const int tile = rem - split * (mt * nt);
  const int tm = n_fast ? tile / nt : tile % mt;
  const int tn = n_fast ? tile % nt : tile / mt;
  const int m0 = tm * MT, n0 = tn * NT;
  const int total = d.chunk_start[d.nseg];
  const int cps = (total + d.splits - 1) / d.splits;
  const int g0 = split * cps;
  const int g1 = min(total, g0 + cps);

  d4 acc[TM][TN];
  f4 facc[FX2 ? TM : 1][FX2 ? TN : 1];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  if constexpr (FX2) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) facc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  }
  int nfl = 0;

  // fetch state: the segment of the next chunk to load, held in registers
  int fs = 0;
  while (fs + 1 < d.nseg && g0 >= d.chunk_start[fs + 1]) ++fs;
  OpDesc FA = d.seg[fs].A, FB = d.seg[fs].B;
  FA.p += batch * bt.a[fs];
  FB.p += batch * bt.bo[fs];
  int FL = d.seg[fs].L, Fend = d.chunk_start[fs + 1];
  double Falpha = d.seg[fs].alpha;
  int kappa = (g0 - d.chunk_start[fs]) * GKC;
  KPos ka{kappa / FA.kdiv, kappa - (kappa / FA.kdiv) * FA.kdiv};
  KPos kb{kappa / FB.kdiv, kappa - (kappa / FB.kdiv) * FB.kdiv};
  int ra_off[SA::E], rb_off[SB::E];
  bool ra_ok[SA::E], rb_ok[SB::E];
  E va[SA::E], vb[SB::E];
  bool oka[SA::E], okb[SB::E];
  op_rows<MT, NTHR>(FA, d.M, m0, tid, ra_off, ra_ok);
  op_rows<NT, NTHR>(FB, d.N, n0, tid, rb_off, rb_ok);
  if (g0 < g1) {
    op_fetch<MT, NTHR>(FA, FL, kappa, ka, tid, ra_off, ra_ok, va, oka);
    op_fetch<NT, NTHR>(FB, FL, kappa, kb, tid, rb_off, rb_ok, vb, okb);
  }
  int la = FA.rfast, lb = FB.rfast;  // LDS layouts of the chunk held in va/vb
  double alpha = Falpha;             // and its segment's scale
  int buf = 0;
  for (int g = g0; g < g1; ++g) {
    op_stash<MT, NTHR>(la, alpha, tid, va, oka, reinterpret_cast<E *>(As[buf]));
    op_stash<NT, NTHR>(lb, 1.0, tid, vb, okb, reinterpret_cast<E *>(Bs[buf]));
    __syncthreads();
    const int cla = la, clb = lb;
    if (g + 1 < g1) {
      if (g + 1 >= Fend) {  // next segment: reload its descriptors once
        ++fs;
        FA = d.seg[fs].A;
        FB = d.seg[fs].B;
        FA.p += batch * bt.a[fs];
        FB.p += batch * bt.bo[fs];
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
    const int ska = cla ? SA::LR : 1, sra = cla ? 1 : SA::LK;
    const int skb = clb ? SB::LR : 1, srb = clb ? 1 : SB::LK;
    // Within a chunk the K index is relabelled so that lane group g = lane/16
    // uses LDS k positions 4g .. 4g+3 for its four MFMA steps (A and B agree,
    // so the contraction is unchanged): a K-contiguous ([r][k]) panel then
    // gives a lane its four steps with two 16-byte reads.
    const int g4 = (lane >> 4) * 4;
    const double *Ab = As[buf] + (wm * WTM + (lane & 15)) * sra + g4 * ska;
    const double *Bb = Bs[buf] + (wn * WTN + (lane & 15)) * srb + g4 * skb;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double fa[2][TM], fb[2][TN];
      if (cla) {
#pragma unroll
        for (int t = 0; t < TM; ++t) {
          fa[0][t] = Ab[(2 * h) * ska + t * 16];
          fa[1][t] = Ab[(2 * h + 1) * ska + t * 16];
        }
      } else {
#pragma unroll
        for (int t = 0; t < TM; ++t) {
          const d2 x = *(const d2 *)(Ab + t * 16 * SA::LK + 2 * h);
          fa[0][t] = x[0];
          fa[1][t] = x[1];
        }
      }
      if (clb) {
#pragma unroll
        for (int t = 0; t < TN; ++t) {
          fb[0][t] = Bb[(2 * h) * skb + t * 16];
          fb[1][t] = Bb[(2 * h + 1) * skb + t * 16];
        }
      } else {
#pragma unroll
        for (int t = 0; t < TN; ++t) {
          const d2 x = *(const d2 *)(Bb + t * 16 * SB::LK + 2 * h);
          fb[0][t] = x[0];
          fb[1][t] = x[1];
        }
      }
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
          for (int y = 0; y < TN; ++y) {
            if constexpr (FX2) {
              const unsigned long long ua = __double_as_longlong(fa[ks][x]);
              const unsigned long long ub = __double_as_longlong(fb[ks][y]);
              const float ah = __uint_as_float((unsigned)ua), al = __uint_as_float((unsigned)(ua >> 32));
              const float bh = __uint_as_float((unsigned)ub), bl = __uint_as_float((unsigned)(ub >> 32));
              facc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(al, bh, facc[x][y], 0, 0, 0);
              facc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bl, facc[x][y], 0, 0, 0);
              facc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bh, facc[x][y], 0, 0, 0);
            } else {
              acc[x][y] =
                  __builtin_amdgcn_mfma_f64_16x16x4f64(fa[ks][x], fb[ks][y], acc[x][y], 0, 0, 0);
            }
          }
    }
    if constexpr (FX2) {
      if (++nfl == kFlush2 || g + 1 == g1) {
        nfl = 0;
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
          for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              acc[x][y][r] += (double)facc[x][y][r];
              facc[x][y][r] = 0.f;
            }
      }
    }
    buf ^= 1;
  }

  // D layout of v_mfma_f64_16x16x4f64 on gfx950: lane l, element r holds
  // row 4*r + l/16, column l%16 of the 16x16 tile.
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ml = wm * WTM + x * 16 + (FX2 ? 4 * (lane >> 4) + r : 4 * r + (lane >> 4));
        const int nl = wn * WTN + y * 16 + (lane & 15);
        const bool part = d.mode == kPartial;
        const OutPos pm = out_pos<MT>(tm, ml, part ? (1 << 30) : d.C.mdiv);
        const OutPos pn = out_pos<NT>(tn, nl, part ? (1 << 30) : d.C.ndiv);
        const int m = pm.idx, n = pn.idx;
        if (m >= d.M || n >= d.N) continue;
        const double val = acc[x][y][r];
        if (part) {
          d.partial[((size_t)split * d.M + m) * d.N + n] = val;
        } else {
          double *c = d.C.p + batch * bt.c + (long long)pm.q * d.C.mx + (long long)pm.r * d.C.my +
                      (long long)pn.q * d.C.nx + (long long)pn.r * d.C.ny;
          if (d.mode == kAccumulate)
            *c += val;
          else
            *c = val;
        }
      }
}

bool gemm_direct_enabled() {
  static const bool on = [] {
    const char *e = getenv("PYSCFAD_AMD_GEMM_DIRECT");
    return !(e && e[0] == '0');
  }();
  return on;
}

hipError_t launch_gemm_f64(const GemmDesc &d, hipStream_t s, const GemmBatch &bt) {
  if (d.M <= 0 || d.N <= 0) return hipSuccess;
  if (bt.nb < 1 || (bt.nb > 1 && d.mode == kPartial)) return hipErrorInvalidValue;
  if (gemm_direct_enabled() && gemm_direct_ok(d, bt)) return launch_gemm_f64_direct(d, s, false, bt);
  // 8 waves, two workgroups per CU; the wide side of the tile goes on the
  // larger output dimension, whose operand panel is then read once (n_fast).
  if (d.M >= d.N) {
    const long long tiles = (long long)((d.M + 127) / 128) * ((d.N + 63) / 64) * d.splits * bt.nb;
    hipLaunchKernelGGL((k_gemm_f64<128, 64, 4, 2, false>), dim3((unsigned)tiles), dim3(512), 0, s,
                       d, bt, 1);
  } else {
    const long long tiles = (long long)((d.M + 63) / 64) * ((d.N + 127) / 128) * d.splits * bt.nb;
    hipLaunchKernelGGL((k_gemm_f64<64, 128, 2, 4, false>), dim3((unsigned)tiles), dim3(512), 0, s,
                       d, bt, 0);
  }
  return hipGetLastError();
}

hipError_t launch_gemm_f32x2(const GemmDesc &d, hipStream_t s, const GemmBatch &bt) {
  if (d.M <= 0 || d.N <= 0) return hipSuccess;
  if (bt.nb < 1 || (bt.nb > 1 && d.mode == kPartial)) return hipErrorInvalidValue;
  if (gemm_direct_enabled() && gemm_direct_ok(d, bt)) return launch_gemm_f64_direct(d, s, true, bt);
  if (d.M >= d.N) {
    const long long tiles = (long long)((d.M + 127) / 128) * ((d.N + 63) / 64) * d.splits * bt.nb;
    hipLaunchKernelGGL((k_gemm_f64<128, 64, 4, 2, true>), dim3((unsigned)tiles), dim3(512), 0, s,
                       d, bt, 1);
  } else {
    const long long tiles = (long long)((d.M + 63) / 64) * ((d.N + 127) / 128) * d.splits * bt.nb;
    hipLaunchKernelGGL((k_gemm_f64<64, 128, 2, 4, true>), dim3((unsigned)tiles), dim3(512), 0, s,
                       d, bt, 0);
  }
  return hipGetLastError();
}

// (the direct engine's tiles when it is on: it takes every descriptor of even sizes)
int gemm_tiles(int M, int N, bool fx2) {
  if (gemm_direct_enabled()) return gemm_direct_tiles(M, N, fx2);
  return M >= N ? ((M + 127) / 128) * ((N + 63) / 64) : ((M + 63) / 64) * ((N + 127) / 128);
}

// ---------------------------------------------------------------------------
// Orbit epilogue
// ---------------------------------------------------------------------------
constexpr int OT = 8, OT3 = OT * OT * OT;

__device__ __forceinline__ void decode_orbit(long long q, int &A, int &B, int &C) {
  int a = 0;
  while ((long long)(a + 1) * (a + 2) * (a + 3) / 6 <= q) ++a;
  long long rem = q - (long long)a * (a + 1) * (a + 2) / 6;
  int b = 0;
  while ((long long)(b + 1) * (b + 2) / 2 <= rem) ++b;
  A = a;
  B = b;
  C = (int)(rem - (long long)b * (b + 1) / 2);
}


// Distinct tiles of an orbit and where the five non-identity permutations of a
// position land, for the four equality patterns of A >= B >= C: 0 all distinct,
// 1 A == B > C, 2 A > B == C, 3 A == B == C.  sel[p] picks the tile indices of
// the p-th distinct tile from (A, B, C) (candidates in the order abc, acb, bac,
// bca, cab, cba, first occurrence kept); pmap[s][p] is the tile that holds the
// bca, cab, cba, acb, bac permutations (s = 0..4) of a position in tile p.
// (int fields: gfx950 has no scalar byte loads, and a vector load here would
// make every wait on it also wait for the kernel's outstanding W loads and
// Wbar/Vbar stores, which share the vector-memory counter)
struct OrbitPattern {
  int np;
  int sel[6][3];
  int pmap[5][6];
};
__constant__ OrbitPattern kOrbitPat[4] = {
    {6, {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}},
     {{3, 5, 1, 4, 0, 2}, {4, 2, 5, 0, 3, 1}, {5, 3, 4, 1, 2, 0}, {1, 0, 3, 2, 5, 4}, {2, 4, 0, 5, 1, 3}}},
    {3, {{0, 0, 2}, {0, 2, 0}, {2, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
     {{1, 2, 0, 0, 0, 0}, {2, 0, 1, 0, 0, 0}, {2, 1, 0, 0, 0, 0}, {1, 0, 2, 0, 0, 0}, {0, 2, 1, 0, 0, 0}}},
    {3, {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
     {{2, 0, 1, 0, 0, 0}, {1, 2, 0, 0, 0, 0}, {2, 1, 0, 0, 0, 0}, {0, 2, 1, 0, 0, 0}, {1, 0, 2, 0, 0, 0}}},
    {1, {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
     {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}}},
};

typedef const __attribute__((address_space(4))) unsigned long long *OrbitEntryPtr;

// Orbit table: entry q = (A, B, C, pattern) of the q-th tile triple A >= B >= C;
// pattern = (A == B) + 2 (B == C) selects kOrbitPat.
__global__ void k_orbit_table(long long norb, unsigned short *tab) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= norb) return;
  int a, b, c;
  decode_orbit(q, a, b, c);
  unsigned short *e = tab + 4 * q;
  e[0] = (unsigned short)a;
  e[1] = (unsigned short)b;
  e[2] = (unsigned short)c;
  e[3] = (unsigned short)((a == b ? 1 : 0) + (b == c ? 2 : 0));
}

hipError_t launch_orbit_table(long long norb, unsigned short *tab, hipStream_t s) {
  hipLaunchKernelGGL(k_orbit_table, dim3((unsigned)((norb + 255) / 256)), dim3(256), 0, s, norb, tab);
  return hipGetLastError();
}

// The orbit epilogue.  A workgroup takes one orbit of 8x8x8 virtual tiles at a
// time: the distinct tiles (P1, P2, P3) among the permutations of (A, B, C).
// Phase 1 loads W of those tiles into LDS and forms X = W + V there; phase 2
// evaluates, for every element of every tile, r3 over the six permuted
// positions (read from LDS), the energy term, Wbar, Vbar and the energy
// denominators' adjoint.  Thread t handles position t of every tile, so wave w
// holds the slab a = 8 P1 + w: the dE/d(e_virtual) sums over two of the three
// tile axes reduce within the wave; the per-wave sums are then added over the
// waves in a fixed order (no atomics), so every output is bitwise reproducible.
// Modes: kOrbitPlain -- energy, and for the adjoint the Wbar/Vbar/dE/de seeds;
// kOrbitProjAcc -- accA[f] += u_f W, accB[f] += u_f Z/D (projected walk);
// kOrbitProjVjp -- seeds of the projected energy from the finished accA, accB
// of the pair: Q = g sum_f u_f accA[f] (dE/dY), P = g sum_f u_f accB[f]
// (dE/dW direct); Wbar = P + r3(Q)/D, Vbar = r3(Q)/D, Dbar = -Q Z/D^2, and
// per-workgroup sums of <W, accB[f]> + <accA[f], Z/D> for dE/du_f.
template <int MODE>
__global__ __launch_bounds__(512) void k_orbit(OrbitArgs A) {
  __shared__ double Ws[6][OT3];
  __shared__ double Xs[6][OT3];
  __shared__ double Qs[MODE == kOrbitProjVjp ? 6 : 1][MODE == kOrbitProjVjp ? OT3 : 1];
  __shared__ double red[8];
  __shared__ double evs[3 * OT];  // virtual orbital energies of the orbit's tile ranges
  __shared__ double ured[MODE == kOrbitProjVjp ? 8 * kMaxFrag : 1];
  extern __shared__ double vir[];  // v doubles: dE/de_virtual accumulated by this workgroup
  constexpr bool kVjp = MODE != kOrbitProjAcc;  // seeds (plain: when A.vjp)

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int o = A.o, v = A.v, n = o + v;
  const int i = A.i, j = A.j, k = A.k;
  const double eo = A.mo[i] + A.mo[j] + A.mo[k];
  const double *ev = A.mo + o;
  const double g = A.ebar * A.omega;
  for (int x = tid; x < v; x += 512) vir[x] = 0.0;
  __syncthreads();

  double eacc = 0.0, occ = 0.0;
  double udot[MODE == kOrbitProjVjp ? kMaxFrag : 1];
#pragma unroll
  for (int f = 0; f < (MODE == kOrbitProjVjp ? kMaxFrag : 1); ++f) udot[f] = 0.0;
  const size_t v3 = (size_t)v * v * v;
  // Element (a, b, c) = 8 (Pa, Pb, Pc) + (al, be, ga): every address splits into a
  // part that is uniform per tile and a part that is fixed per thread.
  const int al = tid >> 6, be = (tid >> 3) & 7, ga = tid & 7;
  const long long thr_w = ((long long)al * v + be) * v + ga;
  const int t_ab = al * v + be, t_ac = al * v + ga, t_bc = be * v + ga;
  const size_t vv = (size_t)v * v;
  const double *ovs_ij = A.ovs + ((size_t)i * o + j) * vv, *ovs_ik = A.ovs + ((size_t)i * o + k) * vv,
               *ovs_jk = A.ovs + ((size_t)j * o + k) * vv;
  const double *t2s_ij = A.t2s + ((size_t)i * o + j) * vv, *t2s_ik = A.t2s + ((size_t)i * o + k) * vv,
               *t2s_jk = A.t2s + ((size_t)j * o + k) * vv;
  const double *t1i = A.t1 + (size_t)i * v, *t1j = A.t1 + (size_t)j * v, *t1k = A.t1 + (size_t)k * v;
  const double *fvo = A.fvo;
  // LDS slot of tile position (a, b, c) = 64 a + 8 b + c, with the low five
  // bits XOR-swizzled by a function of a: then every one of the six permuted
  // read patterns below (a wave spans two of the three tile axes) spreads its
  // 64 lanes evenly over the banks, as the identity pattern does.
  auto slot = [](int pos) { const int a = pos >> 6; return pos ^ (a | ((a & 3) << 3)); };
  const int l_id = slot(tid);
  const int l_bca = slot((be << 6) | (ga << 3) | al);
  const int l_cab = slot((ga << 6) | (al << 3) | be);
  const int l_cba = slot((ga << 6) | (be << 3) | al);
  const int l_acb = slot((al << 6) | (ga << 3) | be);
  const int l_bac = slot((be << 6) | (al << 3) | ga);

  // consecutive orbits (neighbouring tiles along c, which share cache lines of
  // W) go to workgroups on the same XCD, so the halves of a line that two
  // workgroups need meet in one L2
  const int first = xcd_remap(blockIdx.x, gridDim.x);
  // W of the orbit's (up to) six tiles at this thread's position: slots past
  // the orbit's distinct tiles repeat tile 0, indices past v are clamped (their
  // values are zeroed where used).  The loads of the next orbit are issued
  // during the current orbit's phase 2, which issues no loads of its own.
  auto load_w = [&](long long qq, double (&w)[6]) {
    const unsigned long long e = ((const OrbitEntryPtr)A.tab)[qq];
    const int vq[3] = {(int)(e & 0xffff), (int)((e >> 16) & 0xffff), (int)((e >> 32) & 0xffff)};
    const OrbitPattern &pq = kOrbitPat[(int)(e >> 48)];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const int ps = p < pq.np ? p : 0;
      const int ar = min(vq[pq.sel[ps][0]] * OT + al, v - 1), br = min(vq[pq.sel[ps][1]] * OT + be, v - 1),
                cr = min(vq[pq.sel[ps][2]] * OT + ga, v - 1);
      w[p] = A.W[((long long)ar * v + br) * v + cr];
    }
  };
  double wnext[6];
  if (first < A.norb) load_w(first, wnext);
  for (long long q = first; q < A.norb; q += gridDim.x) {
    // The entry is read through the constant address space (the table is not
    // written while the kernel runs), so it and the pattern it selects are
    // scalar loads: they do not count against the vector-memory counter, and
    // so no wait for them also waits for the previous tile loads.
    const unsigned long long ent = ((const OrbitEntryPtr)A.tab)[q];
    const int val[3] = {(int)(ent & 0xffff), (int)((ent >> 16) & 0xffff), (int)((ent >> 32) & 0xffff)};
    const OrbitPattern &pt = kOrbitPat[(int)(ent >> 48)];
    const int np = pt.np;
    // phase 1: W and X = W + V of the orbit's tiles, and the orbital energies
    // of the orbit's three virtual tile ranges.  The W loads of every tile are
    // issued before anything waits on them (slots past np repeat tile 0 and are
    // not stored; indices past v are clamped and their values zeroed); the V
    // terms then follow tile by tile.
    if (tid < 3 * OT) evs[tid] = ev[min(val[tid >> 3] * OT + (tid & 7), v - 1)];
    double wreg[6], xreg[6];
    // clamped element coordinates of tile slot p (recomputed where used: they
    // are a few integer operations, cheaper than keeping them in registers)
    auto coords = [&](int p, int &ar, int &br, int &cr) {
      const int ps = p < np ? p : 0;
      const int a = val[pt.sel[ps][0]] * OT + al, b = val[pt.sel[ps][1]] * OT + be,
                c = val[pt.sel[ps][2]] * OT + ga;
      ar = min(a, v - 1);
      br = min(b, v - 1);
      cr = min(c, v - 1);
      return a < v && b < v && c < v;
    };
#pragma unroll
    for (int p = 0; p < 6; ++p) wreg[p] = wnext[p];
    auto v_terms = [&](auto with_fvo) {
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        int ar, br, cr;
        const bool in = coords(p, ar, br, cr);
        const int ab = ar * v + br, ac = ar * v + cr, bc = br * v + cr;
        double vv_ = ovs_ij[ab] * t1k[cr] + ovs_ik[ac] * t1j[br] + ovs_jk[bc] * t1i[ar];
        if constexpr (decltype(with_fvo)::value)
          vv_ += t2s_ij[ab] * fvo[cr * o + k] + t2s_ik[ac] * fvo[br * o + j] + t2s_jk[bc] * fvo[ar * o + i];
        xreg[p] = in ? wreg[p] + 0.5 * vv_ : 0.0;
        wreg[p] = in ? wreg[p] : 0.0;
      }
    };
    if (A.has_fvo)
      v_terms(std::true_type{});
    else
      v_terms(std::false_type{});
    if constexpr (MODE == kOrbitProjVjp) {
      for (int p = 0; p < np; ++p) {
        const int A0 = val[pt.sel[p][0]] * OT, B0 = val[pt.sel[p][1]] * OT, C0 = val[pt.sel[p][2]] * OT;
        double qv = 0.0;
        if (A0 + al < v && B0 + be < v && C0 + ga < v) {
          const long long tile_w = ((long long)A0 * v + B0) * v + C0;
          for (int f = 0; f < A.nfrag; ++f) qv += A.u[f] * A.accA[f * v3 + tile_w + thr_w];
        }
        Qs[p][l_id] = A.ebar * qv;
      }
    }
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      if (p < np) {
        Ws[p][l_id] = wreg[p];
        Xs[p][l_id] = xreg[p];
      }
    }
    __syncthreads();
    if (q + gridDim.x < A.norb) load_w(q + gridDim.x, wnext);  // next orbit, in flight during phase 2
    // phase 2
    // dE/de_virtual of this orbit, per lane and per tile range r of the orbit
    // (val[r]): lane l < 8 holds the c-axis sum for offset l, 8 <= l < 16 the
    // b-axis sum for offset l - 8, lane 16 the a-axis sum for offset = wave;
    // added over the orbit's tiles in program order, then summed over the waves
    // in a fixed order below (no atomics: every output is reproducible)
    double vacc[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < np; ++p) {
      const int A0 = val[pt.sel[p][0]] * OT, B0 = val[pt.sel[p][1]] * OT, C0 = val[pt.sel[p][2]] * OT;
      const int a = A0 + al, b = B0 + be, c = C0 + ga;
      double dbar = 0.0;
      if (a < v && b < v && c < v) {
        const int p0 = pt.pmap[0][p], p1 = pt.pmap[1][p], p2 = pt.pmap[2][p], p3 = pt.pmap[3][p],
                  p4 = pt.pmap[4][p];
        const double w0 = Ws[p][l_id];
        const double w1 = Ws[p0][l_bca], w2 = Ws[p1][l_cab], w3 = Ws[p2][l_cba],
                     w4 = Ws[p3][l_acb], w5 = Ws[p4][l_bac];
        const double x0 = Xs[p][l_id];
        const double x1 = Xs[p0][l_bca], x2 = Xs[p1][l_cab], x3 = Xs[p2][l_cba],
                     x4 = Xs[p3][l_acb], x5 = Xs[p4][l_bac];
        const double Z = 4.0 * x0 + x1 + x2 - 2.0 * (x3 + x4 + x5);
        const double D = eo - evs[pt.sel[p][0] * OT + al] - evs[pt.sel[p][1] * OT + be] -
                         evs[pt.sel[p][2] * OT + ga];
        // 1/D: hardware reciprocal estimate refined by two Newton-Raphson steps
        // (D <= -(HOMO-LUMO gap) < 0 is never near 0 or a denormal)
        double rD = __builtin_amdgcn_rcp(D);
        rD = fma(rD, fma(-D, rD, 1.0), rD);
        rD = fma(rD, fma(-D, rD, 1.0), rD);
        eacc += w0 * Z * rD;
        const long long idx = ((long long)A0 * v + B0) * v + C0 + thr_w;
        if constexpr (MODE == kOrbitProjAcc) {
          const double y = Z * rD;
          for (int f = 0; f < A.nfrag; ++f) {
            A.accA[f * v3 + idx] += A.u[f] * w0;
            A.accB[f * v3 + idx] += A.u[f] * y;
          }
        } else if constexpr (MODE == kOrbitProjVjp) {
          const double q0 = Qs[p][l_id];
          const double r3Q = 4.0 * q0 + Qs[p0][l_bca] + Qs[p1][l_cab] -
                             2.0 * (Qs[p2][l_cba] + Qs[p3][l_acb] + Qs[p4][l_bac]);
          const double y = Z * rD;
          double pw = 0.0;
#pragma unroll
          for (int f = 0; f < kMaxFrag; ++f) {
            if (f < A.nfrag) {
              const double bf = A.accB[f * v3 + idx];
              pw += A.u[f] * bf;
              udot[f] += w0 * bf + A.accA[f * v3 + idx] * y;
            }
          }
          const double vbar = r3Q * rD;
          const double wbar = A.ebar * pw + vbar;
          if (A.Wb3)
            put3(A.Wb3, v3, idx, wbar);
          else if (A.Wb8)
            A.Wb8[idx] = split2(wbar).u;
          else
            A.W[idx] = wbar;
          A.Vb[idx] = vbar;
          dbar = -q0 * Z * rD * rD;
          occ += dbar;
        } else {
          if (A.vjp) {
            const double r3W = 4.0 * w0 + w1 + w2 - 2.0 * (w3 + w4 + w5);
            const double vbar = g * r3W * rD;
            const double wbar = g * Z * rD + vbar;
            if (A.Wb3)
              put3(A.Wb3, v3, idx, wbar);
            else if (A.Wb8)
              A.Wb8[idx] = split2(wbar).u;
            else
              A.W[idx] = wbar;
            A.Vb[idx] = vbar;
            dbar = -g * w0 * Z * rD * rD;
            occ += dbar;
          }
        }
      }
      if (kVjp && (MODE == kOrbitProjVjp || A.vjp)) {
        // axis c: sum over be (lane bits 3-5); axis b: sum over ga (bits 0-2); axis a: all
        double sc = dbar;
        sc += __shfl_xor(sc, 8, 64);
        sc += __shfl_xor(sc, 16, 64);
        sc += __shfl_xor(sc, 32, 64);
        double sb = dbar;
        sb += __shfl_xor(sb, 1, 64);
        sb += __shfl_xor(sb, 2, 64);
        sb += __shfl_xor(sb, 4, 64);
        double sa = sb;
        sa += __shfl_xor(sa, 8, 64);
        sa += __shfl_xor(sa, 16, 64);
        sa += __shfl_xor(sa, 32, 64);
        // lane l in 8..15 takes the b-axis sum of offset l - 8 (held by lane 8 (l - 8))
        const double sbl = __shfl(sb, 8 * ((lane - 8) & 7), 64);
        const double t = lane < 8 ? sc : (lane < 16 ? sbl : sa);
        const int axis = lane < 8 ? 2 : (lane < 16 ? 1 : 0);
        const int rr = pt.sel[p][axis];
#pragma unroll
        for (int r = 0; r < 3; ++r) vacc[r] += rr == r ? -t : 0.0;
      }
    }
    __syncthreads();
    if (kVjp && (MODE == kOrbitProjVjp || A.vjp)) {
      // stage the per-wave sums in the (now idle) W tiles, then one thread per
      // (range, offset) of the orbit adds them over the waves, and over the
      // ranges that name the same tile, in a fixed order
      double *stg = &Ws[0][0];  // [wave][range][17]
      if (lane < 17) {
#pragma unroll
        for (int r = 0; r < 3; ++r) stg[(wave * 3 + r) * 17 + lane] = vacc[r];
      }
      __syncthreads();
      if (tid < 3 * OT) {
        const int r = tid >> 3, off = tid & 7;
        bool owner = true;
        for (int r2 = 0; r2 < r; ++r2) owner = owner && val[r2] != val[r];
        const int x = val[r] * OT + off;
        if (owner && x < v) {
          double sum = 0.0;
          for (int r2 = r; r2 < 3; ++r2) {
            if (val[r2] != val[r]) continue;
            for (int w = 0; w < 8; ++w) {
              const double *q = stg + (w * 3 + r2) * 17;
              sum += q[off] + q[8 + off] + (off == w ? q[16] : 0.0);
            }
          }
          vir[x] += sum;
        }
      }
      __syncthreads();
    }
  }
  // block sums (512 threads, deterministic order)
  eacc = wave_sum(eacc);
  occ = wave_sum(occ);
  if (lane == 0) red[wave] = eacc;
  __syncthreads();
  double etot = 0.0;
  for (int w = 0; w < 8; ++w) etot += red[w];
  __syncthreads();
  if (lane == 0) red[wave] = occ;
  __syncthreads();
  double otot = 0.0;
  for (int w = 0; w < 8; ++w) otot += red[w];
  if (tid == 0) A.part_t[blockIdx.x] = etot;
  if constexpr (MODE == kOrbitProjVjp) {
    // per-workgroup sums of <W, accB[f]> + <accA[f], Z/D>, wave order fixed
#pragma unroll
    for (int f = 0; f < kMaxFrag; ++f) {
      if (f < A.nfrag) {
        const double s_ = wave_sum(udot[f]);
        if (lane == 0) ured[f * 8 + wave] = s_;
      }
    }
    __syncthreads();
    if (tid < A.nfrag) {
      double s_ = 0.0;
      for (int w = 0; w < 8; ++w) s_ += ured[tid * 8 + w];
      A.part_u[(size_t)blockIdx.x * kMaxFrag + tid] = s_;
    }
  }
  if (MODE == kOrbitProjVjp || (MODE == kOrbitPlain && A.vjp)) {
    double *pm = A.part_mo + (size_t)blockIdx.x * n;
    for (int x = tid; x < v; x += 512) pm[o + x] += vir[x];
    if (tid == 0) {
      pm[i] += otot;
      pm[j] += otot;
      pm[k] += otot;
    }
  }
}

int orbit_grid(int v) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  }
  // one round of resident workgroups (each loops over its share of the orbits):
  // a second, partial round would leave most CUs idle at the end
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_orbit<kOrbitPlain>, 512,
                                                   sizeof(double) * (size_t)v) != hipSuccess ||
      per_cu < 1)
    per_cu = 2;
  long long nt = (v + OT - 1) / OT;
  long long norb = nt * (nt + 1) * (nt + 2) / 6;
  long long g = (long long)per_cu * cus;
  if (g > norb) g = norb;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_orbit(const OrbitArgs &a, int grid, hipStream_t s, int mode) {
  if (mode == kOrbitProjAcc)
    hipLaunchKernelGGL(k_orbit<kOrbitProjAcc>, dim3(grid), dim3(512), sizeof(double) * a.v, s, a);
  else if (mode == kOrbitProjVjp)
    hipLaunchKernelGGL(k_orbit<kOrbitProjVjp>, dim3(grid), dim3(512), sizeof(double) * a.v, s, a);
  else
    hipLaunchKernelGGL(k_orbit<kOrbitPlain>, dim3(grid), dim3(512), sizeof(double) * a.v, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// V-bar contractions for one triple.  Scratch R (v x v blocks):
//   0 R1[a,b] = sum_c Vb t1[k,c]        1 R2[a,b] = sum_c Vb fvo[c,k]
//   2 R3[a,c] = sum_b Vb t1[j,b]        3 R4[a,c] = sum_b Vb fvo[b,j]
//   4 R5[b,c] = sum_a Vb t1[i,a]        5 R6[b,c] = sum_a Vb fvo[a,i]
//   6 Pk[a,c] = sum_b Vb (ovov[i,a,j,b] + ovov[j,b,i,a])
//   7 Fk[a,c] = sum_b Vb (t2[j,i,b,a] + t2[i,j,a,b])
//   8 Pj[a,b] = sum_c Vb (ovov[i,a,k,c] + ovov[k,c,i,a])
//   9 Fj[a,b] = sum_c Vb (t2[k,i,c,a] + t2[i,k,a,c])
//  10*v^2 + a : t1i[a] = sum_bc Vb (ovov[j,b,k,c] + ovov[k,c,j,b])
//  10*v^2 + v + a : fvoi[a] = sum_bc Vb (t2[k,j,c,b] + t2[j,k,b,c])
// Blocks 2-7 and the two vectors are sums over b (k_vr_a) or a (k_vr_b); the
// kVRS parts of those sums are stored separately (parts 1.. after the slots
// above: vr_scratch_doubles) and added in part order by the consumers.  The
// virtual index c is processed in column blocks of kVRC (grid z), so blocks 0,
// 1, 8, 9 (sums over c) and the two vectors also have one part per column
// block; every part is written by exactly one workgroup.
// ---------------------------------------------------------------------------
constexpr int kMaxVQ = kVRC / 64;  // per-lane columns of one column block

constexpr int kVRW = 8;  // waves per V-bar reduction workgroup

// sum over the kVRW waves of per-lane arrays indexed by the block column
// c = lane + 64 qq (c < vl, the width of this column block), written to
// dst[c] in a fixed order
#define PAMD_WAVE_COMBINE(arr, dst)                                          \
  do {                                                                       \
    _Pragma("unroll") for (int qq = 0; qq < kMaxVQ; ++qq) {                  \
      const int c_ = lane + 64 * qq;                                         \
      if (c_ < vl) red[wave * kVRC + c_] = (arr)[qq];                        \
    }                                                                        \
    __syncthreads();                                                         \
    for (int c_ = tid; c_ < vl; c_ += 64 * kVRW) {                           \
      double s_ = 0.0;                                                       \
      for (int w_ = 0; w_ < kVRW; ++w_) s_ += red[w_ * kVRC + c_];           \
      (dst)[c_] = s_;                                                        \
    }                                                                        \
    __syncthreads();                                                         \
  } while (0)

// Sums of four per-lane values over the wave: on return lane 16 q holds the
// total of value q (q = 0..3), in 7 cross-lane steps instead of 4 x 6.
__device__ __forceinline__ double wave_sum4(double v0, double v1, double v2, double v3) {
  const int lane = threadIdx.x & 63;
  const bool hi32 = (lane & 32) != 0, hi16 = (lane & 16) != 0;
  // lanes below 32 collect values 0 and 1, the others 2 and 3
  double x = hi32 ? v2 : v0, y = hi32 ? v3 : v1;
  const double sx = hi32 ? v0 : v2, sy = hi32 ? v1 : v3;
  x += __shfl_xor(sx, 32, 64);
  y += __shfl_xor(sy, 32, 64);
  // then bit 4 of the lane picks one of the two
  double z = hi16 ? y : x;
  const double sz = hi16 ? x : y;
  z += __shfl_xor(sz, 16, 64);
  z += __shfl_xor(z, 8, 64);
  z += __shfl_xor(z, 4, 64);
  z += __shfl_xor(z, 2, 64);
  z += __shfl_xor(z, 1, 64);
  return z;
}

// deterministic block sum over kVRW waves; scratch >= kVRW doubles
__device__ double block_sum_vr(double x, double *scratch) {
  x = wave_sum(x);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[w] = x;
  __syncthreads();
  double r = 0.0;
  for (int q = 0; q < kVRW; ++q) r += scratch[q];
  __syncthreads();
  return r;
}

// Part s of kVRS of the b-sums (k_vr_a) and a-sums (k_vr_b): arrays that are
// sums over the split index are written per part; part 0 uses the slots of R
// listed above, parts 1.. an extension after them (vr_part, vr_part_vec).  The
// vectors have kVRS x ncb parts (index cb * kVRS + s); blocks 0, 1, 8, 9 have
// ncb parts (vr_cpart), column block cb >= 1 in a second extension.
__device__ __forceinline__ double *vr_part(double *R, int v, int s, int q) {
  const size_t vv = (size_t)v * v;
  if (s == 0) return R + (size_t)q * vv;
  return R + 10 * vv + 2 * (size_t)v + ((size_t)(s - 1) * 6 + (q - 2)) * vv;
}
__device__ __forceinline__ double *vr_part_vec(double *R, int v, int s, int which) {
  const size_t vv = (size_t)v * v;
  if (s == 0) return R + 10 * vv + (size_t)which * v;
  return R + 10 * vv + 2 * (size_t)v + (size_t)(kVRS - 1) * 6 * vv + ((size_t)(s - 1) * 2 + which) * v;
}
__device__ __forceinline__ double *vr_cpart(double *R, int v, int cb, int q) {
  const size_t vv = (size_t)v * v;
  if (cb == 0) return R + (size_t)q * vv;
  const int slot = q < 2 ? q : q - 6;  // 0, 1, 8, 9 -> 0, 1, 2, 3
  return R + 10 * vv + 2 * (size_t)v + (size_t)(kVRS - 1) * 6 * vv +
         (size_t)(kVRS * vr_col_blocks(v) - 1) * 2 * v + ((size_t)(cb - 1) * 4 + slot) * vv;
}

// sums over the parts, in part order
__device__ __forceinline__ double vr_sum(const double *R, int v, int q, size_t idx) {
  double x = 0.0;
#pragma unroll
  for (int s = 0; s < kVRS; ++s) x += vr_part(const_cast<double *>(R), v, s, q)[idx];
  return x;
}
__device__ __forceinline__ double vr_sum_vec(const double *R, int v, int which, int idx) {
  double x = 0.0;
  const int np = kVRS * vr_col_blocks(v);
  for (int s = 0; s < np; ++s) x += vr_part_vec(const_cast<double *>(R), v, s, which)[idx];
  return x;
}
__device__ __forceinline__ double vr_csum(const double *R, int v, int q, size_t idx) {
  double x = 0.0;
  const int ncb = vr_col_blocks(v);
  for (int cb = 0; cb < ncb; ++cb) x += vr_cpart(const_cast<double *>(R), v, cb, q)[idx];
  return x;
}

// LDS of the V-bar reductions (one launch runs both kinds of workgroup)
struct VRLds {
  double red[kVRW * kVRC];
  double cst[4][kVRC];  // per-c constants of the b-sums: t1[k,c], fvo[c,k], S2[a,c], T2[a,c]
  double sred[kVRW];
};

template <bool FVO>
__device__ __forceinline__ void vr_a_block(const VRArgs &A, VRLds &L, int a, int part, int cb) {
  double *red = L.red;
  double (*cst)[kVRC] = L.cst;
  double *sred = L.sred;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int o = A.o, v = A.v, i = A.i, j = A.j, k = A.k;
  const size_t vv = (size_t)v * v;
  const double *ovs = A.ovs, *t2s = A.t2s, *t1 = A.t1, *fvo = A.fvo;
#define SYM(arr, x, y, p, q) arr[(((size_t)(x) * o + (y)) * v + (p)) * v + (q)]
  // this workgroup's columns: c = c0 + cl, cl < vl
  const int c0 = cb * kVRC, vl = min(kVRC, v - c0);
  for (int cl = tid; cl < vl; cl += 64 * kVRW) {
    const int c = c0 + cl;
    cst[0][cl] = t1[k * v + c];
    cst[2][cl] = SYM(ovs, i, k, a, c);
    cst[1][cl] = FVO ? fvo[c * o + k] : 0.0;
    cst[3][cl] = SYM(t2s, i, k, a, c);
  }
  __syncthreads();
  double r3[kMaxVQ], r4[kMaxVQ], pk[kMaxVQ], fk[kMaxVQ];
#pragma unroll
  for (int qq = 0; qq < kMaxVQ; ++qq) r3[qq] = r4[qq] = pk[qq] = fk[qq] = 0.0;
  double t1i = 0.0, fvoi = 0.0;
  double *R = A.R;
  for (int b = wave + kVRW * part; b < v; b += kVRW * kVRS) {
    const double t1j = t1[j * v + b];
    const double S1 = SYM(ovs, i, j, a, b);
    const double fvj = FVO ? fvo[b * o + j] : 0.0;
    const double T1 = SYM(t2s, i, j, a, b);
    const double *sjk = ovs + (((size_t)j * o + k) * v + b) * v + c0;
    const double *tjk = t2s + (((size_t)j * o + k) * v + b) * v + c0;
    const double *row = A.Vb + (size_t)a * vv + (size_t)b * v + c0;
    double sr1 = 0.0, sr2 = 0.0, spj = 0.0, sfj = 0.0;
#pragma unroll
    for (int qq = 0; qq < kMaxVQ; ++qq) {
      const int c = lane + 64 * qq;
      if (c < vl) {
        const double x = row[c];
        sr1 += x * cst[0][c];
        spj += x * cst[2][c];
        r3[qq] += x * t1j;
        pk[qq] += x * S1;
        t1i += x * sjk[c];
        // fvo_bar pieces depend on t2 only; the fvo-valued pieces only matter when fvo != 0
        sfj += x * cst[3][c];
        fk[qq] += x * T1;
        fvoi += x * tjk[c];
        if (FVO) {
          sr2 += x * cst[1][c];
          r4[qq] += x * fvj;
        }
      }
    }
    // R1, R2, S_pj, S_fj of row (a, b): lane 16 q holds sum q
    const double s4 = wave_sum4(sr1, sr2, spj, sfj);
    if ((lane & 15) == 0) {
      const int q = lane >> 4;
      if (FVO || q != 1) vr_cpart(R, v, cb, q < 2 ? q : q + 6)[(size_t)a * v + b] = s4;
    }
  }
  __syncthreads();
  PAMD_WAVE_COMBINE(r3, vr_part(R, v, part, 2) + (size_t)a * v + c0);
  if (FVO) PAMD_WAVE_COMBINE(r4, vr_part(R, v, part, 3) + (size_t)a * v + c0);
  PAMD_WAVE_COMBINE(pk, vr_part(R, v, part, 6) + (size_t)a * v + c0);
  PAMD_WAVE_COMBINE(fk, vr_part(R, v, part, 7) + (size_t)a * v + c0);
  const double t1i_s = block_sum_vr(t1i, sred);
  const double fvoi_s = block_sum_vr(fvoi, sred);
  if (tid == 0) {
    vr_part_vec(R, v, cb * kVRS + part, 0)[a] = t1i_s;
    vr_part_vec(R, v, cb * kVRS + part, 1)[a] = fvoi_s;
  }
#undef SYM
}

template <bool FVO>
__device__ __forceinline__ void vr_b_block(const VRArgs &A, VRLds &L, int b, int part, int cb) {
  double *red = L.red;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int o = A.o, v = A.v, i = A.i;
  const size_t vv = (size_t)v * v;
  const int c0 = cb * kVRC, vl = min(kVRC, v - c0);
  double r5[kMaxVQ], r6[kMaxVQ];
#pragma unroll
  for (int qq = 0; qq < kMaxVQ; ++qq) r5[qq] = r6[qq] = 0.0;
  for (int a = wave + kVRW * part; a < v; a += kVRW * kVRS) {
    const double t1a = A.t1[i * v + a];
    const double fa = FVO ? A.fvo[a * o + i] : 0.0;
    const double *row = A.Vb + (size_t)a * vv + (size_t)b * v + c0;
#pragma unroll
    for (int qq = 0; qq < kMaxVQ; ++qq) {
      const int c = lane + 64 * qq;
      if (c < vl) {
        const double x = row[c];
        r5[qq] += x * t1a;
        if (FVO) r6[qq] += x * fa;
      }
    }
  }
  PAMD_WAVE_COMBINE(r5, vr_part(A.R, v, part, 4) + (size_t)b * v + c0);
  if (FVO) PAMD_WAVE_COMBINE(r6, vr_part(A.R, v, part, 5) + (size_t)b * v + c0);
}
#undef PAMD_WAVE_COMBINE

// both V-bar reductions in one launch: blockIdx.y < kVRS runs part y of the
// b-sums, the rest part y - kVRS of the a-sums
template <bool FVO>
__global__ __launch_bounds__(64 * kVRW) void k_vr(VRArgs A) {
  __shared__ VRLds L;
  const int y = blockIdx.y;
  if (y < kVRS)
    vr_a_block<FVO>(A, L, blockIdx.x, y, blockIdx.z);
  else
    vr_b_block<FVO>(A, L, blockIdx.x, y - kVRS, blockIdx.z);
}

hipError_t launch_vr(const VRArgs &a, hipStream_t s) {
  const dim3 grid(a.v, 2 * kVRS, vr_col_blocks(a.v)), block(64 * kVRW);
  if (a.has_fvo)
    hipLaunchKernelGGL(k_vr<true>, grid, block, 0, s, a);
  else
    hipLaunchKernelGGL(k_vr<false>, grid, block, 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-triple gather into the cotangents.  One thread owns one (p,q) position
// of every slab it touches, so coinciding occupied indices (i == j etc.) are
// summed sequentially in the same thread: no atomics, deterministic.
// ---------------------------------------------------------------------------
// Split-K partial sums of the six particle and six hole problems of a triple,
// in place: part[pr][0][t] = sum_s part[pr][s][t], s = 0..S-1, in a fixed
// order (four interleaved partial sums); blockIdx.y < 6 picks a particle
// problem, the rest a hole problem.
__global__ void k_splitk_reduce(double *part_p, int S_p, size_t T_p, double *part_h, int S_h,
                                size_t T_h) {
  const int y = blockIdx.y;
  const bool hole = y >= 6;
  const size_t T = hole ? T_h : T_p;
  const int S = hole ? S_h : S_p;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  double *p = (hole ? part_h + (size_t)(y - 6) * S * T : part_p + (size_t)y * S * T) + t;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int s = 0;
  for (; s + 4 <= S; s += 4) {
    a0 += p[(size_t)s * T];
    a1 += p[(size_t)(s + 1) * T];
    a2 += p[(size_t)(s + 2) * T];
    a3 += p[(size_t)(s + 3) * T];
  }
  for (; s < S; ++s) a0 += p[(size_t)s * T];
  p[0] = (a0 + a1) + (a2 + a3);
}

__device__ __forceinline__ void combine_block(const CombineArgs &A, size_t t) {
  const int o = A.o, v = A.v, i = A.i, j = A.j, k = A.k;
  const size_t vv = (size_t)v * v;
  const double *R = A.R;
  if (t < vv) {
    const int p = (int)(t / v), q = (int)(t % v);
    const size_t pq = (size_t)p * v + q, qp = (size_t)q * v + p;
    // ovov_bar[x,p,y,q]
    auto ob = [&](int x, int y) -> double & {
      return A.ovov_bar[(((size_t)x * v + p) * o + y) * v + q];
    };
    ob(i, j) += 0.5 * vr_csum(R, v, 0, pq);
    ob(j, i) += 0.5 * vr_csum(R, v, 0, qp);
    ob(i, k) += 0.5 * vr_sum(R, v, 2, pq);
    ob(k, i) += 0.5 * vr_sum(R, v, 2, qp);
    ob(j, k) += 0.5 * vr_sum(R, v, 4, pq);
    ob(k, j) += 0.5 * vr_sum(R, v, 4, qp);
    // t2_bar[x,y,p,q]: particle split-K partials (problem order P1..P6) + V terms
    auto tb = [&](int x, int y) -> double & {
      return A.t2_bar[(((size_t)x * o + y) * v + p) * v + q];
    };
    const int xs[6] = {k, j, k, i, j, i};
    const int ys[6] = {j, k, i, k, i, j};
    for (int pr = 0; pr < 6; ++pr) {
      double s = A.part_p[(size_t)pr * A.S_p * vv + pq];  // split-K sum (k_splitk_reduce)
      if (A.has_fvo) {
        switch (pr) {
          case 0: s += 0.5 * vr_sum(R, v, 5, qp); break;  // t2_bar[k,j,c,b] += R6[b,c]/2
          case 1: s += 0.5 * vr_sum(R, v, 5, pq); break;  // t2_bar[j,k,b,c] += R6[b,c]/2
          case 2: s += 0.5 * vr_sum(R, v, 3, qp); break;  // t2_bar[k,i,c,a] += R4[a,c]/2
          case 3: s += 0.5 * vr_sum(R, v, 3, pq); break;  // t2_bar[i,k,a,c] += R4[a,c]/2
          case 4: s += 0.5 * vr_csum(R, v, 1, qp); break;  // t2_bar[j,i,b,a] += R2[a,b]/2
          case 5: s += 0.5 * vr_csum(R, v, 1, pq); break;  // t2_bar[i,j,a,b] += R2[a,b]/2
        }
      }
      tb(xs[pr], ys[pr]) += s;
    }
  }
  if (t < (size_t)v * o) {
    // ovoo_bar[x,p,m,y] from the hole split-K partials H1..H6
    const int p = (int)(t / o), m = (int)(t % o);
    const int xs[6] = {i, i, j, j, k, k};
    const int ys[6] = {j, k, i, k, i, j};
    const size_t vo = (size_t)v * o;
    for (int pr = 0; pr < 6; ++pr) {
      const double s = A.part_h[(size_t)pr * A.S_h * vo + t];  // split-K sum (k_splitk_reduce)
      A.ovoo_bar[(((size_t)xs[pr] * v + p) * o + m) * o + ys[pr]] += s;
    }
  }
}

// t1_bar / fvo_bar pieces of one triple: column sums over a of four v x v
// slices of R, one block per 64 columns (the 4 waves split the rows).
__device__ __forceinline__ void colsums_block(const CombineArgs &A, int blk, double (*red)[4][64]) {
  const int o = A.o, v = A.v, i = A.i, j = A.j, k = A.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blk * 64 + lane;
  const double *R = A.R;
  double sk = 0.0, sj = 0.0, fk = 0.0, fj = 0.0;
  if (c < v) {
    for (int a = wave; a < v; a += 4) {
      const size_t ac = (size_t)a * v + c;
      sk += vr_sum(R, v, 6, ac);
      sj += vr_csum(R, v, 8, ac);
      fk += vr_sum(R, v, 7, ac);
      fj += vr_csum(R, v, 9, ac);
    }
  }
  red[0][wave][lane] = sk;
  red[1][wave][lane] = sj;
  red[2][wave][lane] = fk;
  red[3][wave][lane] = fj;
  __syncthreads();
  if (wave == 0 && c < v) {
    double s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = red[q][0][lane] + red[q][1][lane] + red[q][2][lane] + red[q][3][lane];
    A.t1_bar[k * v + c] += 0.5 * s[0];
    A.t1_bar[j * v + c] += 0.5 * s[1];
    A.t1_bar[i * v + c] += 0.5 * vr_sum_vec(R, v, 0, c);
    A.fvo_bar[(size_t)c * o + k] += 0.5 * s[2];
    A.fvo_bar[(size_t)c * o + j] += 0.5 * s[3];
    A.fvo_bar[(size_t)c * o + i] += 0.5 * vr_sum_vec(R, v, 1, c);
  }
}

// The gather of one triple in one launch (after k_splitk_reduce): blocks
// [0, nc) gather the split-K sums and the V terms into t2_bar / ovov_bar /
// ovoo_bar (one thread per
// position), the next ceil(v / 64) the t1_bar / fvo_bar column sums, and,
// when A.part_t is set, one more the triple's energy from the orbit
// epilogue's per-workgroup sums.
__global__ __launch_bounds__(256) void k_gather(CombineArgs A, int nc) {
  __shared__ double red[4][4][64];
  const int blk = blockIdx.x;
  if (blk < nc) {
    combine_block(A, (size_t)blk * 256 + threadIdx.x);
    return;
  }
  const int ncs = (A.v + 63) / 64;
  if (blk < nc + ncs) {
    colsums_block(A, blk - nc, red);
    return;
  }
  double s = 0.0;
  for (int r = threadIdx.x; r < A.G; r += 256) s += A.part_t[r];
  s = block_sum256(s, &red[0][0][0]);
  if (threadIdx.x == 0) *A.tri_dst = s;
}

hipError_t launch_combine(const CombineArgs &a, hipStream_t s) {
  {
    const size_t Tp = (size_t)a.v * a.v, Th = (size_t)a.v * a.o;
    hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)((std::max(Tp, Th) + 255) / 256), 12),
                       dim3(256), 0, s, const_cast<double *>(a.part_p), a.S_p, Tp,
                       const_cast<double *>(a.part_h), a.S_h, Th);
  }
  size_t n = (size_t)a.v * a.v;
  if ((size_t)a.v * a.o > n) n = (size_t)a.v * a.o;
  const int nc = (int)((n + 255) / 256), ncs = (a.v + 63) / 64;
  const int blocks = nc + ncs + (a.part_t ? 1 : 0);
  hipLaunchKernelGGL(k_gather, dim3((unsigned)blocks), dim3(256), 0, s, a, nc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// exit folds and small reductions
// ---------------------------------------------------------------------------
// ovvv_bar from Vpb[x][p][q][f] (= dE/d ovvv[x,p,f,q]):
//   full:   out[x,p,f,q] = Vpb[x][p][q][f]
//   packed: out[x,p,f(f+1)/2+q] = g[f,q] + g[q,f] (q < f), g[f,f] (q == f)
__global__ void k_fold_ovvv(const double *__restrict__ Vpb, int o, int v, int packed,
                            double *__restrict__ out) {
  const size_t vv = (size_t)v * v;
  const size_t npair = (size_t)v * (v + 1) / 2;
  const size_t total = (size_t)o * v * (packed ? npair : vv);
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    if (packed) {
      const size_t xp = t / npair, r = t % npair;
      size_t f = (size_t)((sqrt(8.0 * (double)r + 1.0) - 1.0) * 0.5);
      while (f * (f + 1) / 2 > r) --f;
      while ((f + 1) * (f + 2) / 2 <= r) ++f;
      const size_t q = r - f * (f + 1) / 2;
      const double *base = Vpb + xp * vv;
      double val = base[q * v + f];
      if (q != f) val += base[f * v + q];
      out[t] = val;
    } else {
      const size_t xp = t / vv, r = t % vv;
      const size_t f = r / v, q = r % v;
      out[t] = Vpb[xp * vv + q * v + f];
    }
  }
}

hipError_t launch_fold_ovvv(const double *Vpb, int o, int v, int packed, double *out, hipStream_t s) {
  size_t n = (size_t)o * v * (packed ? (size_t)v * (v + 1) / 2 : (size_t)v * v);
  hipLaunchKernelGGL(k_fold_ovvv, dim3(grid_for(n)), dim3(256), 0, s, Vpb, o, v, packed, out);
  return hipGetLastError();
}

// t2_bar[x,m,p,q] += T2mb[x][p][q][m]
__global__ void k_fold_t2m(const double *__restrict__ T2mb, int o, int v, double *__restrict__ t2_bar) {
  const size_t total = (size_t)o * o * v * v;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t q = t % v, p = (t / v) % v, m = (t / ((size_t)v * v)) % o,
                 x = t / ((size_t)v * v * o);
    t2_bar[t] += T2mb[((x * v + p) * v + q) * o + m];
  }
}

hipError_t launch_fold_t2m(const double *T2mb, int o, int v, double *t2_bar, hipStream_t s) {
  hipLaunchKernelGGL(k_fold_t2m, dim3(grid_for((size_t)o * o * v * v)), dim3(256), 0, s, T2mb, o,
                     v, t2_bar);
  return hipGetLastError();
}

// out[c] (+)= sum_r part[r][c], fixed order
__global__ void k_reduce_rows(const double *part, int rows, int cols, double *out, int accumulate) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double s = 0.0;
  for (int r = 0; r < rows; ++r) s += part[(size_t)r * cols + c];
  out[c] = accumulate ? out[c] + s : s;
}

hipError_t launch_reduce_rows(const double *part, int rows, int cols, double *out, int accumulate,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_rows, dim3((cols + 255) / 256), dim3(256), 0, s, part, rows, cols,
                     out, accumulate);
  return hipGetLastError();
}

__global__ void k_triple_energy(const double *part_t, int G, double *dst) {
  __shared__ double red[4];
  double s = 0.0;
  for (int r = threadIdx.x; r < G; r += 256) s += part_t[r];
  s = block_sum256(s, red);
  if (threadIdx.x == 0) *dst = s;
}

hipError_t launch_triple_energy(const double *part_t, int G, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_triple_energy, dim3(1), dim3(256), 0, s, part_t, G, dst);
  return hipGetLastError();
}

// tri_e[i][j][k] holds sum_abc W Z / D for sorted i <= j <= k (zero elsewhere).
// E = 1/3 sum over ordered triples = sum_{i<=j<=k} 2 s_ijk tri_e.
__global__ void k_total_energy(const double *tri_e, int o, double *e_out) {
  __shared__ double red[4];
  double s = 0.0;
  const int n3 = o * o * o;
  for (int t = threadIdx.x; t < n3; t += 256) {
    const int i = t / (o * o), j = (t / o) % o, k = t % o;
    if (i <= j && j <= k) {
      const double w = (i == j && j == k) ? (1.0 / 3.0) : ((i == j || j == k) ? 1.0 : 2.0);
      s += w * tri_e[t];
    }
  }
  s = block_sum256(s, red);
  if (threadIdx.x == 0) *e_out = s;
}

hipError_t launch_total_energy(const double *tri_e, int o, double *e_out, hipStream_t s) {
  hipLaunchKernelGGL(k_total_energy, dim3(1), dim3(256), 0, s, tri_e, o, e_out);
  return hipGetLastError();
}

// Dot product of two length-n arrays: per-workgroup strided partial sums, then
// one workgroup sums the G partials (fixed order for a fixed G).
__global__ void k_dot_partial(const double *a, const double *b, size_t n, double *part) {
  __shared__ double red[4];
  double s = 0.0;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256)
    s += a[t] * b[t];
  s = block_sum256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

hipError_t launch_dot(const double *a, const double *b, size_t n, double *part, int G, double *dst,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partial, dim3(G), dim3(256), 0, s, a, b, n, part);
  hipLaunchKernelGGL(k_triple_energy, dim3(1), dim3(256), 0, s, part, G, dst);
  return hipGetLastError();
}

// dst[f * stride] += scale * sum_g part_u[g][f] for f < nfrag (fixed order)
__global__ void k_reduce_ulo(const double *part_u, int G, int nfrag, double scale, double *dst,
                             int stride) {
  const int f = threadIdx.x;
  if (f >= nfrag) return;
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += part_u[(size_t)g * kMaxFrag + f];
  dst[(size_t)f * stride] += scale * s;
}

hipError_t launch_reduce_ulo(const double *part_u, int G, int nfrag, double scale, double *dst,
                             int stride, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_ulo, dim3(1), dim3(64), 0, s, part_u, G, nfrag, scale, dst, stride);
  return hipGetLastError();
}

// E = 1/3 sum over ordered occupied pairs (j, k) of pair_e[j, k], from the
// unordered entries j <= k (the (k, j) term equals the (j, k) one).
__global__ void k_projected_total(const double *pair_e, int o, double *e_out) {
  __shared__ double red[4];
  double s = 0.0;
  for (int t = threadIdx.x; t < o * o; t += 256) {
    const int j = t / o, k = t % o;
    if (j <= k) s += (j == k ? 1.0 : 2.0) * pair_e[t];
  }
  s = block_sum256(s, red);
  if (threadIdx.x == 0) *e_out = s / 3.0;
}

hipError_t launch_projected_total(const double *pair_e, int o, double *e_out, hipStream_t s) {
  hipLaunchKernelGGL(k_projected_total, dim3(1), dim3(256), 0, s, pair_e, o, e_out);
  return hipGetLastError();
}

// Weighted total for a diagonal occupied projector diag(m):
//   E_m = sum_t w_t tri_e[t] (m_i + m_j + m_k) / 3 over i <= j <= k (w_t as k_total_energy),
//   m_bar[x] = ebar * E_x, E_x = 1/3 sum_{y,z} tri_e[sort(x,y,z)] (row sums of e_pair).
// One workgroup, fixed summation order.
__global__ void k_weighted_energy(const double *tri_e, int o, const double *m, double ebar,
                                  double *e_out, double *m_bar) {
  __shared__ double red[4];
  double s = 0.0;
  const int n3 = o * o * o;
  for (int t = threadIdx.x; t < n3; t += 256) {
    const int i = t / (o * o), j = (t / o) % o, k = t % o;
    if (i <= j && j <= k) {
      const double w = (i == j && j == k) ? (1.0 / 3.0) : ((i == j || j == k) ? 1.0 : 2.0);
      s += w * tri_e[t] * ((m[i] + m[j] + m[k]) / 3.0);
    }
  }
  s = block_sum256(s, red);
  if (threadIdx.x == 0) *e_out = s;
  if (m_bar) {
    for (int x = threadIdx.x; x < o; x += 256) {
      double r = 0.0;
      for (int y = 0; y < o; ++y)
        for (int z = 0; z < o; ++z) {
          int a = x, b = y, c = z, tmp;
          if (a > b) { tmp = a; a = b; b = tmp; }
          if (b > c) { tmp = b; b = c; c = tmp; }
          if (a > b) { tmp = a; a = b; b = tmp; }
          r += tri_e[((size_t)a * o + b) * o + c];
        }
      m_bar[x] = ebar * (r / 3.0);
    }
  }
}

hipError_t launch_weighted_energy(const double *tri_e, int o, const double *m, double ebar,
                                  double *e_out, double *m_bar, hipStream_t s) {
  hipLaunchKernelGGL(k_weighted_energy, dim3(1), dim3(256), 0, s, tri_e, o, m, ebar, e_out, m_bar);
  return hipGetLastError();
}

// e_pair[x][y] = 1/3 sum_z tri_e[sort(x,y,z)]
__global__ void k_pair_energy(const double *tri_e, int o, double *e_pair) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= o * o) return;
  const int x = t / o, y = t % o;
  double s = 0.0;
  for (int z = 0; z < o; ++z) {
    int a = x, b = y, c = z, tmp;
    if (a > b) { tmp = a; a = b; b = tmp; }
    if (b > c) { tmp = b; b = c; c = tmp; }
    if (a > b) { tmp = a; a = b; b = tmp; }
    s += tri_e[((size_t)a * o + b) * o + c];
  }
  e_pair[t] = s / 3.0;
}

hipError_t launch_pair_energy(const double *tri_e, int o, double *e_pair, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_energy, dim3((o * o + 255) / 256), dim3(256), 0, s, tri_e, o, e_pair);
  return hipGetLastError();
}

}  // namespace pamd
