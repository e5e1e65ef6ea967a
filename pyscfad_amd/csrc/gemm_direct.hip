// FP64 contraction engine with direct-to-LDS operand loads (gfx950).
//
// Same problem and descriptors as k_gemm_f64 (ccsdt_kernels.hip): a segmented
// strided GEMM C (=, +=, split-K partial) sum_s alpha_s A_s B_s over K chunks
// of 16.  The difference is how operands reach LDS: every lane moves 16 bytes
// (two consecutive elements) per `global_load_lds_dwordx4`, which writes LDS
// at (wave base + 16 * lane) without passing through registers.  The staging
// layout is chosen by picking WHICH global element pair each lane fetches:
//
//   K-contiguous operand ("kfast"), R rows x 16 k: [r][k], 128 bytes per row,
//     the eight 16-byte chunks of row r stored at chunk c ^ f(r), with
//     f(r) = (r ^ (r >> 3)) & 7, so the 16 rows a fragment read touches hit 16
//     distinct 16-byte slots of every 256-byte bank window;
//   row-contiguous operand ("rfast"): [k][r], R doubles per k, the R/2 chunks
//     of k-row k stored at chunk p ^ (((k >> 2) & 3) << 3), so the four lane
//     groups of a fragment read (k = 4g + s) sit 128 bytes apart (R = 80:
//     128-byte row blocks placed across k-rows, rfast_at).
//
// Three LDS stages for the 64-wide tiles (two for the 80-wide ones): while
// chunk g is multiplied, the loads of chunks g+1 and g+2 are in flight.  Each
// iteration waits for its own chunk's loads
// (s_waitcnt vmcnt), then one s_barrier (without the memory fence that
// __syncthreads adds, which would also wait for the chunks still in flight)
// makes every wave's loads visible and retires the slot read one iteration
// earlier, which the next loads then overwrite.
//
// The kernel is compiled per launch for the operand layouts its segments use
// (each operand all K-contiguous, all row-contiguous, or mixed with a runtime
// choice), so the fragment reads and MFMAs of a chunk form one straight-line
// block, and the fetch tracker keeps each operand's chunk offset by additions
// (no 64-bit multiplies per chunk): every wave runs that scalar work in the
// same phase after the barrier, while the MFMA pipes wait.  On W-build-shaped
// descriptors 42-43 -> 52 TFLOP/s.
//
// A pair of elements must be contiguous and 16-byte aligned, so an operand is
// eligible when its fast stride is 1, every other stride and index divisor is
// even, its base is 16-byte aligned, the rows of a row-contiguous operand are
// even in number, and every segment length is even (gemm_direct_ok).  Rows past
// M or N and k past L are clamped to valid elements (finite inputs); k past L
// is then zeroed in the A fragments, rows past M/N are never stored.
#include <cstdlib>
#include "ccsdt_common.h"
#include "engine_staging.h"

namespace pamd {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kDStages = 3;
constexpr int kDK = 16;  // K chunk

typedef __attribute__((address_space(3))) void *lds_ptr_t;

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

__device__ __forceinline__ int kfast_swz(int r) { return (r ^ (r >> 3)) & 7; }
// row-contiguous tile of R rows: 16-byte chunk p (rows 2p, 2p+1) of k-row k is
// stored at p ^ rfast_swz<R>(k), so the four k of a fragment read (k = 4g + s,
// g = 0..3) sit at different chunk offsets; the XOR stays below R / 2
template <int R>
__device__ __forceinline__ int rfast_swz(int k) {
  return (R % 64 == 0) ? (((k >> 2) & 3) << 3) : (((k >> 2) & 3) << 1);
}

// LDS offset (doubles) of row r, k-row k in a row-contiguous stage of R rows,
// and its inverse for the loads (16-byte chunk id -> k, row pair p).  R = 80:
// a k-row is 640 bytes, so k = 4g + s for the four lane groups g of a fragment
// read falls on one bank window whatever the XOR within the row; instead the
// 128-byte blocks (16 rows) of lane groups 0 and 2 take the even block slots
// and those of groups 1 and 3 the odd ones (block slot 2 (5 kr + r / 16) +
// g % 2, kr = s + 4 (g / 2)), so the two groups of each half wave read
// different halves of every 256-byte bank window.  Other R: the XOR above.
template <int R>
__device__ __forceinline__ int rfast_at(int k, int r) {
  if constexpr (R == 80) {
    const int kr = (k & 3) + 4 * (k >> 3);
    const int bslot = 2 * (5 * kr + (r >> 4)) + ((k >> 2) & 1);
    return (bslot * 8 + ((r >> 1) & 7)) * 2 + (r & 1);
  } else {
    return k * R + 2 * ((r >> 1) ^ rfast_swz<R>(k)) + (r & 1);
  }
}

template <int R>
__device__ __forceinline__ void rfast_chunk(int id, int &k, int &p) {
  if constexpr (R == 80) {
    const int bs = id >> 3, idx = bs >> 1, kr = idx / 5;
    k = (kr & 3) + 8 * (kr >> 2) + 4 * (bs & 1);
    p = 8 * (idx - 5 * kr) + (id & 7);
  } else {
    k = id / (R / 2);
    p = (id % (R / 2)) ^ rfast_swz<R>(k);
  }
}

// 16-byte chunks of an R-row x 16-k tile, and loads per thread of 512 (the
// last load of a thread past the tile's 8 R chunks fills padding)
template <int R>
struct DTile {
  static constexpr int CH = 8 * R;
  static constexpr int NI = (CH + 511) / 512;
  static constexpr int SLOTS = NI * 512;
};

// per-thread load plan of one operand (R rows) for one segment
// When no K chunk of the segment straddles a divisor boundary of the K index
// (kdiv a multiple of 16, or K not split), every load of a full chunk sits at
// a per-thread pointer (row offset + its chunk-local k times ky, fixed for the
// segment) plus one wave-uniform offset of the chunk start: ptr[q].
template <int R>
struct DPlan {
  static constexpr int NI = DTile<R>::NI;  // 16-byte loads per thread per chunk
  const double *ptr[NI];             // p + row offset + kl * ky (used when `fast`)
  bool fast;
};

// row offset (elements, clamped row) and chunk-local k of this thread's load q
template <int R>
__device__ __forceinline__ void dslot(const OpDesc &d, int Rtot, int r0, int tid, int q, int &roff,
                                      int &kl) {
  int r;
  int id = q * 512 + tid;
  if (id >= DTile<R>::CH) id = 0;  // padding: a valid chunk, never read
  if (d.rfast) {
    int k, p;
    rfast_chunk<R>(id, k, p);
    r = r0 + 2 * p;
    if (r > Rtot - 2) r = Rtot - 2;  // Rtot is even for a row-contiguous operand
    kl = k;
  } else {
    const int rl = id >> 3, pc = id & 7;
    kl = 2 * (pc ^ kfast_swz(rl));
    r = r0 + rl;
    if (r > Rtot - 1) r = Rtot - 1;
  }
  const int qd = r / d.rdiv;
  roff = (int)(qd * d.rx + (r - qd * d.rdiv) * d.ry);
}

template <int R>
__device__ __forceinline__ void dplan(const OpDesc &d, int Rtot, int r0, int tid, DPlan<R> &pl) {
#pragma unroll
  for (int q = 0; q < DPlan<R>::NI; ++q) {
    int roff, kl;
    dslot<R>(d, Rtot, r0, tid, q, roff, kl);
    pl.ptr[q] = d.p + roff + (long long)kl * d.ky;
  }
  pl.fast = d.kdiv >= (1 << 30) || (d.kdiv % kDK) == 0;
}

// K offset of chunk-local k = kl at chunk start kappa (kp = divmod(kappa, kdiv)),
// clamped to the last valid element (pair) when past L
__device__ __forceinline__ int dkoff(const OpDesc &d, int L, int kappa, KPos kp, int kl) {
  const int lim = L - (d.rfast ? 1 : 2) - kappa;
  if (kl > lim) kl = lim;
  int r = kp.r + kl, q = kp.q;
  while (r >= d.kdiv) {
    r -= d.kdiv;
    ++q;
  }
  return (int)(q * d.kx + r * d.ky);
}

// The loads of one chunk.  Full chunks of a `fast` operand (kappa <= flim)
// use the stored pointers plus kb, the chunk start's K offset (kept by the
// tracker); otherwise (divisor boundary inside a chunk, or the last chunk of
// a segment whose length is not a multiple of 16) each load's row and k are
// recomputed and k is clamped to the segment.
template <int R>
__device__ __forceinline__ void dissue(const OpDesc &d, int Rtot, int r0, int tid, int L,
                                       int kappa, KPos kp, long long kb, int flim,
                                       const DPlan<R> &pl, double *lds, int wave) {
  if (kappa <= flim) {
#pragma unroll
    for (int q = 0; q < DPlan<R>::NI; ++q)
      __builtin_amdgcn_global_load_lds((const void *)(pl.ptr[q] + kb),
                                       (lds_ptr_t)(lds + (q * 512 + wave * 64) * 2), 16, 0, 0);
    return;
  }
#pragma unroll
  for (int q = 0; q < DPlan<R>::NI; ++q) {
    int roff, kl;
    dslot<R>(d, Rtot, r0, tid, q, roff, kl);
    const double *src = d.p + roff + dkoff(d, L, kappa, kp, kl);
    __builtin_amdgcn_global_load_lds((const void *)src, (lds_ptr_t)(lds + (q * 512 + wave * 64) * 2),
                                     16, 0, 0);
  }
}

// K tracker of one operand: position (q, r) of the next chunk start in the
// operand's K divisor and its element offset q kx + r ky, advanced by adding
// 16 ky per chunk and kx - kdiv ky per wrap of r (no multiplies per chunk)
struct DTrack {
  KPos kp;
  long long off;
  int flim;  // last chunk start that the stored pointers cover (-1: none)
};

template <int R>
__device__ __forceinline__ void dtrack_init(DTrack &t, const OpDesc &d, int kappa, int L,
                                            const DPlan<R> &pl) {
  t.kp = KPos{kappa / d.kdiv, kappa % d.kdiv};
  t.off = (long long)t.kp.q * d.kx + (long long)t.kp.r * d.ky;
  t.flim = pl.fast ? L - kDK : -1;
}

__device__ __forceinline__ void dtrack_next(DTrack &t, const OpDesc &d) {
  t.kp.r += kDK;
  t.off += d.ky << 4;
  while (t.kp.r >= d.kdiv) {
    t.kp.r -= d.kdiv;
    ++t.kp.q;
    t.off += d.kx - (long long)d.kdiv * d.ky;
  }
}

// One K chunk (16 k) of a wave's TM x TN MFMA tiles, in KH parts of SPH = 4 / KH
// MFMA steps (s = SPH h + u; two parts keep half the fragments live at a
// time).  LAM / LBM: the A / B stages of every segment of the launch are
// K-contiguous (0) or row-contiguous (1), or differ between segments (2: the
// runtime flags la / lb decide).  With the layouts fixed the reads and MFMAs
// of a part form one block, so the reads overlap the MFMAs.  k past the
// segment (kv < 16) is zeroed and alpha applied on the A side.
template <int MT, int NT, int WTM, int WTN, bool FX2, int KH, int LAM, int LBM, class AccD,
          class AccF>
__device__ __forceinline__ void f64d_chunk(const double *As, const double *Bs, int wm, int wn,
                                           int lane, int la, int lb, int kv, double alpha,
                                           AccD &acc, AccF &facc) {
  constexpr int TM = WTM / 16, TN = WTN / 16, SPH = 4 / KH;
  const bool ra = LAM == 2 ? la != 0 : LAM == 1;
  const bool rb = LBM == 2 ? lb != 0 : LBM == 1;
  const int g4 = (lane >> 4) * 4;
#pragma unroll
  for (int h = 0; h < KH; ++h) {
    double fa[SPH][TM], fb[SPH][TN];
#pragma unroll
    for (int t = 0; t < TM; ++t) {
      const int r = wm * WTM + t * 16 + (lane & 15);
      if (ra) {
#pragma unroll
        for (int u = 0; u < SPH; ++u) {
          fa[u][t] = As[rfast_at<MT>(g4 + SPH * h + u, r)];
        }
      } else {
#pragma unroll
        for (int w = 0; w < SPH / 2; ++w) {
          const d2v x = *(const d2v *)(As + r * kDK + 2 * (((g4 >> 1) + SPH / 2 * h + w) ^ kfast_swz(r)));
          fa[2 * w][t] = x[0];
          fa[2 * w + 1][t] = x[1];
        }
      }
    }
#pragma unroll
    for (int t = 0; t < TN; ++t) {
      const int r = wn * WTN + t * 16 + (lane & 15);
      if (rb) {
#pragma unroll
        for (int u = 0; u < SPH; ++u) {
          fb[u][t] = Bs[rfast_at<NT>(g4 + SPH * h + u, r)];
        }
      } else {
#pragma unroll
        for (int w = 0; w < SPH / 2; ++w) {
          const d2v x = *(const d2v *)(Bs + r * kDK + 2 * (((g4 >> 1) + SPH / 2 * h + w) ^ kfast_swz(r)));
          fb[2 * w][t] = x[0];
          fb[2 * w + 1][t] = x[1];
        }
      }
    }
    if (kv < kDK) {  // k past the segment: zero the A side (B holds clamped, finite values)
#pragma unroll
      for (int u = 0; u < SPH; ++u)
        if (g4 + SPH * h + u >= kv) {
#pragma unroll
          for (int t = 0; t < TM; ++t) fa[u][t] = 0.0;
        }
    }
    if constexpr (FX2) {
      if (alpha < 0.0) {  // FP32 hi/lo pairs: flip both signs
#pragma unroll
        for (int u = 0; u < SPH; ++u)
#pragma unroll
          for (int t = 0; t < TM; ++t)
            fa[u][t] = __longlong_as_double(__double_as_longlong(fa[u][t]) ^ 0x8000000080000000ll);
      }
    } else if (alpha != 1.0) {
#pragma unroll
      for (int u = 0; u < SPH; ++u)
#pragma unroll
        for (int t = 0; t < TM; ++t) fa[u][t] *= alpha;
    }
#pragma unroll
    for (int u = 0; u < SPH; ++u)
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) {
          if constexpr (FX2) {
            const unsigned long long ua = __double_as_longlong(fa[u][x]);
            const unsigned long long ub = __double_as_longlong(fb[u][y]);
            const float ah = __uint_as_float((unsigned)ua), al = __uint_as_float((unsigned)(ua >> 32));
            const float bh = __uint_as_float((unsigned)ub), bl = __uint_as_float((unsigned)(ub >> 32));
            facc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(al, bh, facc[x][y], 0, 0, 0);
            facc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bl, facc[x][y], 0, 0, 0);
            facc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bh, facc[x][y], 0, 0, 0);
          } else {
            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[u][x], fb[u][y], acc[x][y], 0, 0, 0);
          }
        }
  }
}

}  // namespace

template <int MT, int NT, int WM, int WN, bool FX2, int NST, int LAM, int LBM, int KH = 2>
__global__ __launch_bounds__(512, 4) void k_gemm_f64d(GemmDesc d, GemmBatch bt, int n_fast) {
  static_assert(WM * WN == 8, "8 waves");
  constexpr int WTM = MT / WM, WTN = NT / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  static_assert(KH == 1 || KH == 2, "one or two parts per chunk");
  constexpr int SPH = 4 / KH;  // MFMA k-steps per part
  constexpr int NIA = DTile<MT>::NI, NIB = DTile<NT>::NI, NI = NIA + NIB;
  constexpr int SA = DTile<MT>::SLOTS * 2, SB = DTile<NT>::SLOTS * 2, SS = SA + SB;  // doubles per stage
  static_assert(2 * NST * SS * 8 <= 160 * 1024, "LDS of two workgroups per CU");
  constexpr int kFlush2 = 8;
  __shared__ __align__(16) double S[NST * SS];
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

  d4v acc[TM][TN];
  f4v facc[FX2 ? TM : 1][FX2 ? TN : 1];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = d4v{0.0, 0.0, 0.0, 0.0};
  if constexpr (FX2) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) facc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
  }
  int nfl = 0;

  // fetch tracker: segment, chunk start and K position of the next chunk to load
  int fs = 0;
  while (fs + 1 < d.nseg && g0 >= d.chunk_start[fs + 1]) ++fs;
  int fkappa = (g0 - d.chunk_start[fs]) * kDK;
  // fetch-side segment fields (operand pointers moved to this product of the
  // batch), cached while the segment lasts
  OpDesc FA = d.seg[fs].A, FB = d.seg[fs].B;
  FA.p += batch * bt.a[fs];
  FB.p += batch * bt.bo[fs];
  int FL = d.seg[fs].L, fend = d.chunk_start[fs + 1];
  DPlan<MT> pa;
  DPlan<NT> pb;
  dplan<MT>(FA, d.M, m0, tid, pa);
  dplan<NT>(FB, d.N, n0, tid, pb);
  DTrack ta, tb;
  dtrack_init<MT>(ta, FA, fkappa, FL, pa);
  dtrack_init<NT>(tb, FB, fkappa, FL, pb);
  int fst = 0, cst = 0;  // LDS stages of the next chunk to load and to compute
  // compute tracker (segment fields cached while the segment lasts)
  int cs = fs, ckappa = fkappa, cend = d.chunk_start[fs + 1];
  int la = d.seg[fs].A.rfast, lb = d.seg[fs].B.rfast, cL = d.seg[fs].L;
  double alpha = d.seg[fs].alpha;

  auto issue = [&](int g) {  // loads of chunk g (the fetch tracker's chunk) into its stage
    double *st = S + fst * SS;
    fst = fst + 1 == NST ? 0 : fst + 1;
    dissue<MT>(FA, d.M, m0, tid, FL, fkappa, ta.kp, ta.off, ta.flim, pa, st, wave);
    dissue<NT>(FB, d.N, n0, tid, FL, fkappa, tb.kp, tb.off, tb.flim, pb, st + SA, wave);
    // advance the tracker to chunk g + 1
    if (g + 1 < g1) {
      if (g + 1 >= fend) {
        while (g + 1 >= d.chunk_start[fs + 1]) ++fs;  // (skips segments of length 0)
        FA = d.seg[fs].A;
        FB = d.seg[fs].B;
        FA.p += batch * bt.a[fs];
        FB.p += batch * bt.bo[fs];
        FL = d.seg[fs].L;
        fend = d.chunk_start[fs + 1];
        fkappa = 0;
        dplan<MT>(FA, d.M, m0, tid, pa);
        dplan<NT>(FB, d.N, n0, tid, pb);
        dtrack_init<MT>(ta, FA, 0, FL, pa);
        dtrack_init<NT>(tb, FB, 0, FL, pb);
      } else {
        fkappa += kDK;
        dtrack_next(ta, FA);
        dtrack_next(tb, FB);
      }
    }
  };

  static_assert(NST == 2 || NST == 3, "two or three LDS stages");
  if (g0 < g1) issue(g0);
  if (NST == 3 && g0 + 1 < g1) issue(g0 + 1);

  // a wave whose rows or columns of the tile lie wholly past M or N only loads
  // and synchronises (its MFMA issue slots go to the other waves of its SIMD)
  const bool live = m0 + wm * WTM < d.M && n0 + wn * WTN < d.N;
  for (int g = g0; g < g1; ++g) {
    if (NST == 3 && g + 1 < g1)
      wait_vmcnt<NI>();
    else
      wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (g + NST - 1 < g1) issue(g + NST - 1);

    // compute chunk g
    if (g >= cend) {
      while (g >= d.chunk_start[cs + 1]) ++cs;
      ckappa = 0;
      cend = d.chunk_start[cs + 1];
      la = d.seg[cs].A.rfast;
      lb = d.seg[cs].B.rfast;
      alpha = d.seg[cs].alpha;
      cL = d.seg[cs].L;
    }
    const int kv = cL - ckappa;  // valid k in this chunk (>= 16: all)
    ckappa += kDK;
    const double *As = S + cst * SS;
    cst = cst + 1 == NST ? 0 : cst + 1;
    if (!live) continue;
    const double *Bs = As + SA;
    f64d_chunk<MT, NT, WTM, WTN, FX2, KH, LAM, LBM>(As, Bs, wm, wn, lane, la, lb, kv, alpha, acc, facc);
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
  }

  // D layout of v_mfma_f64_16x16x4f64: lane l, element r holds row 4 r + l / 16,
  // column l % 16 (FP32 16x16x4: row 4 (l / 16) + r)
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

// An operand can be staged with 16-byte direct loads: two consecutive elements
// along its fast index are adjacent in memory and 16-byte aligned, for every
// row and k the kernel may address (rows clamped to [0, rows), k to [0, L)).
static bool op_direct_ok(const OpDesc &o, int rows, int L) {
  auto even = [](long long x) { return (x & 1) == 0; };
  const bool rsplit = o.rdiv < (1 << 30), ksplit = o.kdiv < (1 << 30);
  if (((uintptr_t)o.p & 15) != 0) return false;
  if (o.rfast) {
    if (o.ry != 1 || (rows & 1)) return false;
    if (rsplit && ((o.rdiv & 1) || !even(o.rx))) return false;
    if (!even(o.ky) || (ksplit && !even(o.kx))) return false;
  } else {
    if (o.ky != 1 || (L & 1)) return false;
    if (ksplit && ((o.kdiv & 1) || !even(o.kx))) return false;
    if (!even(o.ry) || (rsplit && !even(o.rx))) return false;
  }
  return true;
}

bool gemm_direct_ok(const GemmDesc &d, const GemmBatch &bt) {
  if (d.M <= 0 || d.N <= 0 || d.nseg <= 0) return false;
  for (int q = 0; q < d.nseg; ++q) {
    const SegDesc &s = d.seg[q];
    if (s.L & 1) return false;
    if (!op_direct_ok(s.A, d.M, s.L) || !op_direct_ok(s.B, d.N, s.L)) return false;
    // every product of a batch keeps the operands 16-byte aligned
    if (bt.nb > 1 && ((bt.a[q] | bt.bo[q]) & 1)) return false;
  }
  return true;
}

// Tile shapes {MT, NT}: 128 x 64 / 64 x 128 (4 x 2 / 2 x 4 waves, three LDS
// stages) and 128 x 80 / 80 x 128 (8 x 1 / 1 x 8 waves, two LDS stages: the
// 80-wide tiles cover nvir = 400 or 160 without padding).  The pick maximises
// the padded tiling's efficiency times MT NT / (MT + NT).
static const int kF64dShape[4][2] = {{128, 64}, {64, 128}, {128, 80}, {80, 128}};

static double f64d_eff(int M, int N, int MT, int NT) {
  const double pm = (double)((M + MT - 1) / MT) * MT, pn = (double)((N + NT - 1) / NT) * NT;
  return (double)M * N / (pm * pn);
}

// fx2: the FP32 hi/lo form keeps FP32 and FP64 accumulators, which do not fit
// the 80-wide tiles' registers; it takes the 64-wide shapes only
int gemm_direct_pick(int M, int N, bool fx2) {
  int best = 0;
  double bs = -1.0;
  for (int c = 0; c < (fx2 ? 2 : 4); ++c) {
    const int MT = kF64dShape[c][0], NT = kF64dShape[c][1];
    const double sc = f64d_eff(M, N, MT, NT) * (double)MT * NT / (MT + NT);
    if (sc > bs + 1e-9) {
      bs = sc;
      best = c;
    }
  }
  return best;
}

int gemm_direct_tiles(int M, int N, bool fx2) {
  const int c = gemm_direct_pick(M, N, fx2);
  return ((M + kF64dShape[c][0] - 1) / kF64dShape[c][0]) * ((N + kF64dShape[c][1] - 1) / kF64dShape[c][1]);
}

// layout mode of one operand over the descriptor's segments: 0 all
// K-contiguous, 1 all row-contiguous, 2 mixed
static int layout_mode(const GemmDesc &d, bool b) {
  int any0 = 0, any1 = 0;
  for (int q = 0; q < d.nseg; ++q) {
    if ((b ? d.seg[q].B.rfast : d.seg[q].A.rfast) != 0)
      any1 = 1;
    else
      any0 = 1;
  }
  return any0 && any1 ? 2 : any1;
}

template <int MT, int NT, int WM, int WN, bool FX2, int NST, int LAM>
static void launch_f64d_b(int lbm, dim3 g, hipStream_t s, const GemmDesc &d, const GemmBatch &bt,
                          int n_fast) {
  switch (lbm) {
    case 0: hipLaunchKernelGGL((k_gemm_f64d<MT, NT, WM, WN, FX2, NST, LAM, 0>), g, dim3(512), 0, s, d, bt, n_fast); break;
    case 1: hipLaunchKernelGGL((k_gemm_f64d<MT, NT, WM, WN, FX2, NST, LAM, 1>), g, dim3(512), 0, s, d, bt, n_fast); break;
    default: hipLaunchKernelGGL((k_gemm_f64d<MT, NT, WM, WN, FX2, NST, LAM, 2>), g, dim3(512), 0, s, d, bt, n_fast); break;
  }
}

template <int MT, int NT, int WM, int WN, bool FX2, int NST>
static void launch_f64d_ab(dim3 g, hipStream_t s, const GemmDesc &d, const GemmBatch &bt, int n_fast) {
  const int lbm = layout_mode(d, true);
  switch (layout_mode(d, false)) {
    case 0: launch_f64d_b<MT, NT, WM, WN, FX2, NST, 0>(lbm, g, s, d, bt, n_fast); break;
    case 1: launch_f64d_b<MT, NT, WM, WN, FX2, NST, 1>(lbm, g, s, d, bt, n_fast); break;
    default: launch_f64d_b<MT, NT, WM, WN, FX2, NST, 2>(lbm, g, s, d, bt, n_fast); break;
  }
}

template <bool FX2>
static void launch_f64d(const GemmDesc &d, const GemmBatch &bt, hipStream_t s) {
  const int c = gemm_direct_pick(d.M, d.N, FX2);
  const int MT = kF64dShape[c][0], NT = kF64dShape[c][1];
  const long long tiles = (long long)((d.M + MT - 1) / MT) * ((d.N + NT - 1) / NT) * d.splits * bt.nb;
  const int n_fast = d.M >= d.N ? 1 : 0;
  const dim3 g((unsigned)tiles);
  switch (c) {
    case 0: launch_f64d_ab<128, 64, 4, 2, FX2, kDStages>(g, s, d, bt, n_fast); break;
    case 1: launch_f64d_ab<64, 128, 2, 4, FX2, kDStages>(g, s, d, bt, n_fast); break;
    default:
      if constexpr (!FX2) {
        if (c == 2)
          launch_f64d_ab<128, 80, 8, 1, false, 2>(g, s, d, bt, n_fast);
        else
          launch_f64d_ab<80, 128, 1, 8, false, 2>(g, s, d, bt, n_fast);
      }
      break;
  }
}

hipError_t launch_gemm_f64_direct(const GemmDesc &d, hipStream_t s, bool fx2,
                                  const GemmBatch &bt) {
  if (d.M <= 0 || d.N <= 0) return hipSuccess;
  if (bt.nb < 1 || (bt.nb > 1 && d.mode == kPartial)) return hipErrorInvalidValue;
  if (fx2)
    launch_f64d<true>(d, bt, s);
  else
    launch_f64d<false>(d, bt, s);
  return hipGetLastError();
}

}  // namespace pamd
