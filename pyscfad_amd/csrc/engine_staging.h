// Operand staging shared by the contraction engines (FP64 and split-BF16x3).
//
// An operand panel of R rows x GKC contraction indices is fetched from global
// memory through an OpDesc (see ccsdt_common.h) and written to LDS in the
// layout that keeps its global reads coalesced and its LDS traffic free of bank
// conflicts:
//   row-contiguous operand ("rfast") -> [k][r], row stride R + 16 elements
//   K-contiguous operand   ("kfast") -> [r][k], row stride 18 elements
// Elements are 8 bytes: a double, or a packed bf16 (hi, mid, lo) triplet.
#pragma once
#include "bf16_split.h"
#include "ccsdt_common.h"

namespace pamd {

constexpr int GKC = 16;

template <int R, int NTHR>
struct OpStage {
  static constexpr int E = R * GKC / NTHR;  // staged elements per thread
  static constexpr int RSTEP = NTHR / R;    // rfast: k step between a thread's elements
  static constexpr int KSTEP = NTHR / 16;   // kfast: row step between a thread's elements
  static constexpr int LR = R + 16;         // [k][r] row stride (doubles)
  static constexpr int LK = GKC + 2;        // [r][k] row stride (doubles; 16-B aligned rows)
  static constexpr int LDS = (GKC * LR > R * LK) ? GKC * LR : R * LK;
  // every (row, k) of the R x 16 panel is staged exactly once by both mappings
  static_assert((R * GKC) % NTHR == 0 && NTHR % R == 0 && NTHR % 16 == 0, "staging map");
  static_assert(KSTEP * E == R && RSTEP * E == GKC, "staging map covers the panel");
};

// Output coordinate of position w of tile t along one dimension: the index,
// and its split (q, r) = divmod(index, div) for the output descriptor (no
// division when div is the "no split" value 2^30).
struct OutPos {
  int idx, q, r;
};

template <int R>
__device__ __forceinline__ OutPos out_pos(int t, int w, int div) {
  OutPos p;
  p.idx = t * R + w;
  if (div >= (1 << 30)) {
    p.q = 0;
    p.r = p.idx;
  } else {
    p.q = p.idx / div;
    p.r = p.idx - p.q * div;
  }
  return p;
}

template <int R, int NTHR>
__device__ __forceinline__ void op_rows(const OpDesc &d, int R_total, int r0, int tid,
                                        int (&roff)[OpStage<R, NTHR>::E],
                                        bool (&rok)[OpStage<R, NTHR>::E]) {
  using S = OpStage<R, NTHR>;
  if (d.rfast) {
    const int r = r0 + tid % R;
    rok[0] = r < R_total;
    const int q = r / d.rdiv;
    roff[0] = rok[0] ? (int)(q * d.rx + (r - q * d.rdiv) * d.ry) : 0;
#pragma unroll
    for (int e = 1; e < S::E; ++e) {
      rok[e] = false;
      roff[e] = 0;
    }
  } else {
#pragma unroll
    for (int e = 0; e < S::E; ++e) {
      const int r = r0 + (tid >> 4) + S::KSTEP * e;
      rok[e] = r < R_total;
      const int q = r / d.rdiv;
      roff[e] = rok[e] ? (int)(q * d.rx + (r - q * d.rdiv) * d.ry) : 0;
    }
  }
}

// K position of a chunk as (q, r) = divmod(kappa0, kdiv), advanced without division
struct KPos {
  int q, r;
};

__device__ __forceinline__ int k_offset(const OpDesc &d, KPos kp, int add) {
  int r = kp.r + add, q = kp.q;
  if (d.kdiv >= GKC) {
    if (r >= d.kdiv) {
      r -= d.kdiv;
      ++q;
    }
  } else {
    while (r >= d.kdiv) {
      r -= d.kdiv;
      ++q;
    }
  }
  return (int)(q * d.kx + r * d.ky);
}

// Loads are issued unconditionally from a clamped (in-range) address; the raw
// values stay in registers and the mask and alpha are applied when they are
// written to LDS, so nothing waits on the loads before the MFMA phase.
template <int R, int NTHR, class T>
__device__ __forceinline__ void op_fetch(const OpDesc &d, int L, int kappa0, KPos kp, int tid,
                                         const int (&roff)[OpStage<R, NTHR>::E],
                                         const bool (&rok)[OpStage<R, NTHR>::E],
                                         T (&v)[OpStage<R, NTHR>::E],
                                         bool (&ok)[OpStage<R, NTHR>::E]) {
  const T *p = reinterpret_cast<const T *>(d.p);
  using S = OpStage<R, NTHR>;
  if (d.rfast) {
#pragma unroll
    for (int e = 0; e < S::E; ++e) {
      const int kl = tid / R + S::RSTEP * e;
      ok[e] = rok[0] && (kappa0 + kl) < L;
      const int off = ok[e] ? roff[0] + k_offset(d, kp, kl) : 0;
      v[e] = p[off];
    }
  } else {
    const int kl = tid & 15;
    const bool kok = (kappa0 + kl) < L;
    const int ko = kok ? k_offset(d, kp, kl) : 0;
#pragma unroll
    for (int e = 0; e < S::E; ++e) {
      ok[e] = kok && rok[e];
      v[e] = p[ok[e] ? roff[e] + ko : 0];
    }
  }
}

// masked, scaled element as stored in LDS: FP64 values are multiplied by
// alpha; packed bf16 triplets (alpha = +-1) get their three sign bits flipped
__device__ __forceinline__ double staged(double v, bool ok, double alpha) {
  return ok ? v * alpha : 0.0;
}
__device__ __forceinline__ unsigned long long staged(unsigned long long v, bool ok, double alpha) {
  return ok ? (alpha < 0.0 ? v ^ 0x0000800080008000ull : v) : 0ull;
}

// FP32 hi/lo pairs (alpha = +-1): the two sign bits are flipped
__device__ __forceinline__ F2 staged(F2 v, bool ok, double alpha) {
  F2 r;
  r.u = ok ? (alpha < 0.0 ? v.u ^ 0x8000000080000000ull : v.u) : 0ull;
  return r;
}

template <int R, int NTHR, class T>
__device__ __forceinline__ void op_stash(int rfast, double alpha, int tid,
                                         const T (&v)[OpStage<R, NTHR>::E],
                                         const bool (&ok)[OpStage<R, NTHR>::E], T *Sm) {
  using S = OpStage<R, NTHR>;
  if (rfast) {
    const int r = tid % R;
#pragma unroll
    for (int e = 0; e < S::E; ++e)
      Sm[(tid / R + S::RSTEP * e) * S::LR + r] = staged(v[e], ok[e], alpha);
  } else {
    const int k = tid & 15;
#pragma unroll
    for (int e = 0; e < S::E; ++e)
      Sm[((tid >> 4) + S::KSTEP * e) * S::LK + k] = staged(v[e], ok[e], alpha);
  }
}



// Workgroup w runs on XCD w % 8; give each XCD a contiguous range of logical
// tiles so that tiles sharing an operand panel share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int w, int n) {
  const int q = n >> 3, r = n & 7, x = w & 7, k = w >> 3;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

}  // namespace pamd
