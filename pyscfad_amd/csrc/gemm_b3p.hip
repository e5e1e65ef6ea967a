// Split-BF16x3 contraction engine with direct-to-LDS operand loads (gfx950).
//
// Same problem and descriptors as the register-staged split-BF16x3 engine
// (bf16_kernels.hip): C (=, +=, split-K partial) sum_s alpha_s A_s B_s, each
// product as the six bf16 MFMA terms hi*hi + hi*mid + mid*hi + hi*lo + lo*hi +
// mid*mid summed in FP32 per workgroup, alpha_s = +-1.  Operands are planar
// (OpDesc::ps): the hi, mid and lo pieces of an element sit in three bf16
// planes of the FP64 array's layout.  Every lane moves 16 bytes (8 consecutive
// bf16 of one plane) per `global_load_lds_dwordx4`, which writes LDS at (wave
// base + 16 * lane) without passing through registers; the LDS layout is set
// by which global chunk each lane fetches.  Per plane of an R-row x 32-k stage:
//
//   K-contiguous operand: [r][32 k], 64 bytes per row; the chunk holding k
//     8c..8c+7 of row r is stored at chunk c ^ ksw(r).  The fragment read
//     (16 rows, one chunk, 16 bytes per lane) then touches 16 distinct 16-byte
//     slots of every 256-byte window, and 8 distinct ones of every 128-byte
//     window for the first 8 lanes;
//   row-contiguous operand: [32 k][R], 2R bytes per k-row; the chunk holding
//     rows 8c..8c+7 of k-row kk is stored at chunk c ^ rsw(kk).  A 16x32 MFMA
//     operand (row = lane % 16, k = 8 (lane / 16) + 0..7) is read with two
//     ds_read_b64_tr_b16: within each 16-lane group that instruction hands
//     lane j element j % 4 of the 8 bytes read by lane 4e + j / 4, as element
//     e.  Lane 4e + q of group G therefore reads rows 4q..4q+3 of k-row
//     8G + e (then 8G + 4 + e), and lane j receives row j at k 8G + e.
//
// Two LDS stages: while stage g is multiplied, the loads of stage g + 1 are
// in flight.  Each iteration waits for its own loads (s_waitcnt vmcnt(0)),
// then one s_barrier (no memory fence) makes every wave's loads visible and
// retires the stage read one iteration earlier, which the new loads overwrite.
//
// Eligibility (gemm_b3p_ok): 8 consecutive elements along the fast index are
// adjacent and 16-byte aligned for every row and k the kernel addresses: fast
// stride 1, every other stride, divisor, plane stride and batch offset a
// multiple of 8, the base 16-byte aligned; K-contiguous operands need
// segment lengths that are multiples of 8, row-contiguous ones a row count
// that is.  Rows past M or N and k past L are clamped to valid elements
// (finite inputs); k past L is zeroed in the fragments of one side, rows past
// M or N are never stored.
#include <cstdlib>
#include <cstring>
#include "ccsdt_common.h"
#include "engine_staging.h"

namespace pamd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) void *lds_ptr_t;
typedef __attribute__((address_space(3))) v4i16 lds_v4i16;
typedef unsigned short bf16_t;

constexpr int kPK = 32;      // k per stage: one 16x16x32 MFMA step
constexpr int kPStages = 2;

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

__device__ __forceinline__ int ksw(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
template <int R>
__device__ __forceinline__ int rsw(int kk) {
  if constexpr (R % 128 == 0) return 2 * (kk & 3) + 8 * ((kk >> 3) & 1);
  if constexpr (R % 64 == 0) return 2 * (kk & 3);
  return 0;  // 2R = 32 x odd bytes per k-row: consecutive k-rows already fall in distinct banks
}

// per-thread load plan of one operand (R rows, 3 planes) for one segment, in
// a workgroup of T threads
template <int R, int T>
struct PPlan {
  static constexpr int CH = 12 * R;             // chunks per stage
  static constexpr int NI = (CH + T - 1) / T;   // loads per thread
  static constexpr int SLOTS = NI * T;          // LDS chunks (with padding)
  const bf16_t *ptr[NI];  // source of load q at the segment's k = 0, used by full stages
};

// slot of load q: plane and, within the plane's tile, the row (K-contiguous)
// or first row of 8 (row-contiguous) rl, and the chunk-local k kl.  Loads past
// the 12 R chunks fill padding with a valid chunk.
template <int R, int T>
__device__ __forceinline__ void pslot(int q, int tid, int rc, int &plane, int &rl, int &kl) {
  const int id = q * T + tid;
  if (id >= 12 * R) {
    plane = 0;
    rl = 0;
    kl = 0;
    return;
  }
  plane = id / (4 * R);
  const int rem = id - plane * 4 * R;
  if (!rc) {
    rl = rem >> 2;
    kl = 8 * ((rem & 3) ^ ksw(rl));
  } else {
    constexpr int CPR = R / 8;
    const int kk = rem / CPR, pc = rem - kk * CPR;
    rl = 8 * (pc ^ rsw<R>(kk));
    kl = kk;
  }
}

// element offset of row r (clamped: K-contiguous to Rtot - 1, row-contiguous
// chunk starts to Rtot - 8)
__device__ __forceinline__ long long prow(const OpDesc &d, int rc, int Rtot, int r) {
  const int lim = rc ? Rtot - 8 : Rtot - 1;
  if (r > lim) r = lim;
  const int q = r / d.rdiv;
  return (long long)q * d.rx + (long long)(r - q * d.rdiv) * d.ry;
}

// element offset of contraction index k (clamped: K-contiguous chunk starts
// to L - 8, row-contiguous k-rows to L - 1)
__device__ __forceinline__ long long pkoff(const OpDesc &d, int rc, int L, int k) {
  const int lim = rc ? L - 1 : L - 8;
  if (k > lim) k = lim;
  const int q = k / d.kdiv;
  return (long long)q * d.kx + (long long)(k - q * d.kdiv) * d.ky;
}

template <int R, int T>
__device__ __forceinline__ void pplan(const OpDesc &d, int rc, int Rtot, int r0, int tid,
                                      PPlan<R, T> &pl) {
  const bf16_t *p = reinterpret_cast<const bf16_t *>(d.p);
#pragma unroll
  for (int q = 0; q < PPlan<R, T>::NI; ++q) {
    int plane, rl, kl;
    pslot<R, T>(q, tid, rc, plane, rl, kl);
    pl.ptr[q] = p + plane * d.ps + prow(d, rc, Rtot, r0 + rl) + (long long)kl * d.ky;
  }
}

// Fetch-side state of one operand for the current segment: its K strides and
// divisor (all that full stages need), whether it is row-contiguous, the next
// stage start's remainder r in the K divisor and its element offset
// kb = q kx + r ky, advanced by adding 32 ky per stage and kx - kdiv ky per
// wrap of r (no multiplies per stage).  The full descriptor is read from the
// kernel argument only for partial stages and at segment switches.
struct POp {
  long long kx, ky, kb;
  int kdiv, rc, r;
};

__device__ __forceinline__ void pop_init(POp &f, const OpDesc &d, int kappa) {
  f.kx = d.kx;
  f.ky = d.ky;
  f.kdiv = d.kdiv;
  f.rc = d.rfast;
  const int q = kappa / d.kdiv;
  f.r = kappa - q * d.kdiv;
  f.kb = (long long)q * d.kx + (long long)f.r * d.ky;
}

__device__ __forceinline__ void pop_next(POp &f) {
  f.r += kPK;
  f.kb += f.ky << 5;
  while (f.r >= f.kdiv) {
    f.r -= f.kdiv;
    f.kb += f.kx - (long long)f.kdiv * f.ky;
  }
}

// stage k = kappa .. kappa + 31 lies in the segment and in one quotient of the
// K divisor: the plan's pointers plus kb address it
__device__ __forceinline__ bool pop_full(const POp &f, int kappa, int L) {
  return kappa + kPK <= L && (f.kdiv >= (1 << 30) || f.r + kPK <= f.kdiv);
}

template <int R, int T>
__device__ __forceinline__ void pissue_full(const PPlan<R, T> &pl, long long kb, bf16_t *lds,
                                            int wave) {
#pragma unroll
  for (int q = 0; q < PPlan<R, T>::NI; ++q)
    __builtin_amdgcn_global_load_lds((const void *)(pl.ptr[q] + kb),
                                     (lds_ptr_t)(lds + (q * T + wave * 64) * 8), 16, 0, 0);
}

// the loads of a partial stage: each load's row and k recomputed, k clamped
template <int R, int T>
__device__ __forceinline__ void pissue_part(const OpDesc &d, int Rtot, int r0, int tid, int L,
                                            int kappa, bf16_t *lds, int wave) {
  const int rc = d.rfast;
  const bf16_t *p = reinterpret_cast<const bf16_t *>(d.p);
#pragma unroll
  for (int q = 0; q < PPlan<R, T>::NI; ++q) {
    int plane, rl, kl;
    pslot<R, T>(q, tid, rc, plane, rl, kl);
    const bf16_t *src = p + plane * d.ps + prow(d, rc, Rtot, r0 + rl) + pkoff(d, rc, L, kappa + kl);
    __builtin_amdgcn_global_load_lds((const void *)src,
                                     (lds_ptr_t)(lds + (q * T + wave * 64) * 8), 16, 0, 0);
  }
}

// MFMA operand (rows r16 .. r16 + 15, k 8 (lane / 16) .. + 7) of one plane
template <int R>
__device__ __forceinline__ u32x4 frag_k(const bf16_t *P, int r16, int lane) {
  const int r = r16 + (lane & 15), G = lane >> 4;
  return *(const u32x4 *)(P + r * kPK + 8 * (G ^ ksw(r)));
}

template <int R>
__device__ __forceinline__ u32x4 frag_r(const bf16_t *P, int r16, int lane) {
  const int G = lane >> 4, i = lane & 15, e = i >> 2, q = i & 3;
  const int rr = r16 + 4 * q;
  const int c = rr >> 3, half = rr & 7;
  const int k0 = 8 * G + e, k1 = k0 + 4;
  const v4i16 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
      (lds_v4i16 *)(P + k0 * R + 8 * (c ^ rsw<R>(k0)) + half));
  const v4i16 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
      (lds_v4i16 *)(P + k1 * R + 8 * (c ^ rsw<R>(k1)) + half));
  const u32x2 l2 = __builtin_bit_cast(u32x2, lo), h2 = __builtin_bit_cast(u32x2, hi);
  return u32x4{l2[0], l2[1], h2[0], h2[1]};
}

template <int R>
__device__ __forceinline__ void frags(const bf16_t *T, int rc, int r16, int lane, u32x4 &h,
                                      u32x4 &m, u32x4 &l) {
  constexpr int PL = R * kPK;  // bf16 per plane
  if (rc) {
    h = frag_r<R>(T, r16, lane);
    m = frag_r<R>(T + PL, r16, lane);
    l = frag_r<R>(T + 2 * PL, r16, lane);
  } else {
    h = frag_k<R>(T, r16, lane);
    m = frag_k<R>(T + PL, r16, lane);
    l = frag_k<R>(T + 2 * PL, r16, lane);
  }
}

// WM x WN waves; two workgroups per CU (80 KB of LDS each at most)
// LDS of one workgroup (bf16 elements): two stages of the A and B tiles
template <int MT, int NT, int WM, int WN>
struct B3pLds {
  static constexpr int T = 64 * WM * WN;
  static constexpr int SA = PPlan<MT, T>::SLOTS * 8, SB = PPlan<NT, T>::SLOTS * 8, SS = SA + SB;
  static constexpr int N = kPStages * SS;
  static_assert(2 * N * 2 <= 160 * 1024, "LDS of two workgroups per CU");
};

// One workgroup's tile of one product: workgroup number `logical` of the
// product's (or the batch's) grid.  D is GemmDesc or GemmDescS (same fields).
template <int MT, int NT, int WM, int WN, class D>
__device__ __forceinline__ void b3p_tile(const D &d, const GemmBatch &bt, int n_fast, int logical,
                                         bf16_t *S) {
  constexpr int T = 64 * WM * WN;
  constexpr int WTM = MT / WM, WTN = NT / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  static_assert(TM * 16 == WTM && TN * 16 == WTN, "16x16 MFMA tiles");
  constexpr int SA = B3pLds<MT, NT, WM, WN>::SA, SS = B3pLds<MT, NT, WM, WN>::SS;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int mt = (d.M + MT - 1) / MT, nt = (d.N + NT - 1) / NT;
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

  f32x4 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  // fetch tracker: segment, stage start and K position of the next stage to load
  int fs = 0;
  while (fs + 1 < d.nseg && g0 >= d.chunk_start[fs + 1]) ++fs;
  int fkappa = (g0 - d.chunk_start[fs]) * kPK;
  // operand b (0: A, 1: B) of the fetch segment, moved to this product of the batch
  auto opdesc = [&](int b) {
    OpDesc o = b ? d.seg[fs].B : d.seg[fs].A;
    o.p = reinterpret_cast<const double *>(reinterpret_cast<const bf16_t *>(o.p) +
                                           batch * (b ? bt.bo[fs] : bt.a[fs]));
    return o;
  };
  POp fa, fb;
  PPlan<MT, T> pa;
  PPlan<NT, T> pb;
  auto fetch_segment = [&](int kappa) {
    const OpDesc A = opdesc(0), B = opdesc(1);
    pop_init(fa, A, kappa);
    pop_init(fb, B, kappa);
    pplan<MT, T>(A, A.rfast, d.M, m0, tid, pa);
    pplan<NT, T>(B, B.rfast, d.N, n0, tid, pb);
  };
  fetch_segment(fkappa);
  int FL = d.seg[fs].L, fend = d.chunk_start[fs + 1];
  // compute tracker
  int cs = fs, ckappa = fkappa, cend = d.chunk_start[fs + 1];
  int la = d.seg[fs].A.rfast, lb = d.seg[fs].B.rfast, cL = d.seg[fs].L;
  bool neg = d.seg[fs].alpha < 0.0;

  auto issue = [&](int g) {  // loads of stage g (the fetch tracker's stage)
    bf16_t *st = S + ((g - g0) & 1) * SS;
    if (pop_full(fa, fkappa, FL))
      pissue_full<MT, T>(pa, fa.kb, st, wave);
    else
      pissue_part<MT, T>(opdesc(0), d.M, m0, tid, FL, fkappa, st, wave);
    if (pop_full(fb, fkappa, FL))
      pissue_full<NT, T>(pb, fb.kb, st + SA, wave);
    else
      pissue_part<NT, T>(opdesc(1), d.N, n0, tid, FL, fkappa, st + SA, wave);
    if (g + 1 < g1) {
      if (g + 1 >= fend) {
        while (g + 1 >= d.chunk_start[fs + 1]) ++fs;  // (skips segments of length 0)
        FL = d.seg[fs].L;
        fend = d.chunk_start[fs + 1];
        fkappa = 0;
        fetch_segment(0);
      } else {
        fkappa += kPK;
        pop_next(fa);
        pop_next(fb);
      }
    }
  };

  if (g0 < g1) issue(g0);
  const int G = lane >> 4;
  // a wave whose rows or columns of the tile lie wholly past M or N (the last
  // tile of a side that is not a multiple of the tile, e.g. 400 = 3 x 128 + 16)
  // only loads and synchronises: its MFMA issue slots go to the other waves of
  // its SIMD
  const bool live = m0 + wm * WTM < d.M && n0 + wn * WTN < d.N;
  for (int g = g0; g < g1; ++g) {
    wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (g + 1 < g1) issue(g + 1);

    if (g >= cend) {
      while (g >= d.chunk_start[cs + 1]) ++cs;
      ckappa = 0;
      cend = d.chunk_start[cs + 1];
      la = d.seg[cs].A.rfast;
      lb = d.seg[cs].B.rfast;
      neg = d.seg[cs].alpha < 0.0;
      cL = d.seg[cs].L;
    }
    const int kv = cL - ckappa;  // valid k in this stage (>= 32: all)
    ckappa += kPK;
    if (!live) continue;
    const bf16_t *As = S + ((g - g0) & 1) * SS;
    const bf16_t *Bs = As + SA;
    // k past the segment is zeroed and alpha = -1 flips the sign bits on the
    // side whose fragments stay in registers (both sides hold finite values
    // there: clamped elements), and only on stages that need it
    const bool fix = neg || kv < kPK;
    u32x4 keep = {~0u, ~0u, ~0u, ~0u};
    if (kv < kPK) {
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int k = 8 * G + 2 * w;
        keep[w] = (k < kv ? 0xffffu : 0u) | (k + 1 < kv ? 0xffff0000u : 0u);
      }
    }
    const unsigned sg = neg ? 0x80008000u : 0u;
    const u32x4 sgn = {sg, sg, sg, sg};
    // the fragments of the narrower side stay in registers; the other side is
    // read two tiles at a time, and the six products of those two tiles
    // alternate between their accumulators
    if constexpr (TM <= TN) {
      u32x4 ah[TM], am[TM], al[TM];
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        frags<MT>(As, la, wm * WTM + 16 * t, lane, ah[t], am[t], al[t]);
      }
      if (fix) {
#pragma unroll
        for (int t = 0; t < TM; ++t) {
          ah[t] = (ah[t] ^ sgn) & keep;
          am[t] = (am[t] ^ sgn) & keep;
          al[t] = (al[t] ^ sgn) & keep;
        }
      }
#pragma unroll
      for (int b0 = 0; b0 < TN; b0 += 2) {
        constexpr int W2 = 2;
        u32x4 bh[W2], bm[W2], bl[W2];
#pragma unroll
        for (int u = 0; u < W2; ++u)
          if (b0 + u < TN) frags<NT>(Bs, lb, wn * WTN + 16 * (b0 + u), lane, bh[u], bm[u], bl[u]);
#pragma unroll
        for (int a = 0; a < TM; ++a) {
#define B3P_PRODUCT(AX, BX)                                                                   \
  _Pragma("unroll") for (int u = 0; u < W2; ++u) if (b0 + u < TN) acc[a][b0 + u] =            \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, AX),                 \
                                              __builtin_bit_cast(bf16x8, BX[u]), acc[a][b0 + u], 0, 0, 0);
          // smallest terms first
          B3P_PRODUCT(am[a], bm)
          B3P_PRODUCT(ah[a], bl)
          B3P_PRODUCT(al[a], bh)
          B3P_PRODUCT(ah[a], bm)
          B3P_PRODUCT(am[a], bh)
          B3P_PRODUCT(ah[a], bh)
#undef B3P_PRODUCT
        }
      }
    } else {
      u32x4 bh[TN], bm[TN], bl[TN];
#pragma unroll
      for (int t = 0; t < TN; ++t) frags<NT>(Bs, lb, wn * WTN + 16 * t, lane, bh[t], bm[t], bl[t]);
      if (fix) {
#pragma unroll
        for (int t = 0; t < TN; ++t) {
          bh[t] = (bh[t] ^ sgn) & keep;
          bm[t] = (bm[t] ^ sgn) & keep;
          bl[t] = (bl[t] ^ sgn) & keep;
        }
      }
#pragma unroll
      for (int a0 = 0; a0 < TM; a0 += 2) {
        constexpr int W2 = 2;
        u32x4 ah[W2], am[W2], al[W2];
#pragma unroll
        for (int u = 0; u < W2; ++u)
          if (a0 + u < TM) frags<MT>(As, la, wm * WTM + 16 * (a0 + u), lane, ah[u], am[u], al[u]);
#pragma unroll
        for (int b = 0; b < TN; ++b) {
#define B3P_PRODUCT(AX, BX)                                                                   \
  _Pragma("unroll") for (int u = 0; u < W2; ++u) if (a0 + u < TM) acc[a0 + u][b] =            \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, AX[u]),              \
                                              __builtin_bit_cast(bf16x8, BX), acc[a0 + u][b], 0, 0, 0);
          B3P_PRODUCT(am, bm[b])
          B3P_PRODUCT(ah, bl[b])
          B3P_PRODUCT(al, bh[b])
          B3P_PRODUCT(ah, bm[b])
          B3P_PRODUCT(am, bh[b])
          B3P_PRODUCT(ah, bh[b])
#undef B3P_PRODUCT
        }
      }
    }
  }

  // D layout of v_mfma_f32_16x16x32_bf16: lane l, element r -> row 4 (l / 16) + r, column l % 16
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ml = wm * WTM + 16 * a + 4 * (lane >> 4) + r;
        const int nl = wn * WTN + 16 * b + (lane & 15);
        const bool part = d.mode == kPartial;
        const OutPos pm = out_pos<MT>(tm, ml, part ? (1 << 30) : d.C.mdiv);
        const OutPos pn = out_pos<NT>(tn, nl, part ? (1 << 30) : d.C.ndiv);
        const int m = pm.idx, n = pn.idx;
        if (m >= d.M || n >= d.N) continue;
        const double val = (double)acc[a][b][r];
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

template <int MT, int NT, int WM, int WN, int MINW>
__global__ __launch_bounds__(64 * WM * WN, MINW) void k_gemm_b3p(GemmDesc d, GemmBatch bt, int n_fast) {
  __shared__ __align__(16) bf16_t S[B3pLds<MT, NT, WM, WN>::N];
  b3p_tile<MT, NT, WM, WN>(d, bt, n_fast, xcd_remap(blockIdx.x, gridDim.x), S);
}

// A group of up to kMaxGroup independent products with equal M, N, split-K
// factor and mode, each of at most two segments, in one launch: the grid is
// the products' grids one after the other.
template <int MT, int NT, int WM, int WN, int MINW>
__global__ __launch_bounds__(64 * WM * WN, MINW) void k_gemm_b3p_group(GemmGroup g, int n_fast) {
  __shared__ __align__(16) bf16_t S[B3pLds<MT, NT, WM, WN>::N];
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int per = (int)(gridDim.x / (unsigned)g.n);
  const int p = logical / per;
  GemmBatch bt;
  for (int q = 0; q < kMaxSeg; ++q) bt.a[q] = bt.bo[q] = 0;
  bt.c = 0;
  bt.nb = 1;
  b3p_tile<MT, NT, WM, WN>(g.d[p], bt, n_fast, logical - p * per, S);
}

// An operand can be staged with 16-byte direct loads of 8 bf16 along its fast
// index, for every row and k the kernel may address (rows clamped to
// [0, rows), k to [0, L)).
bool op_b3p_ok(const OpDesc &o, int rows, int L) {
  auto m8 = [](long long x) { return (x & 7) == 0; };
  const bool rsplit = o.rdiv < (1 << 30), ksplit = o.kdiv < (1 << 30);
  if (((uintptr_t)o.p & 15) != 0 || !m8(o.ps) || o.ps <= 0) return false;
  if (o.rfast) {
    if (o.ry != 1 || rows < 8 || (rows & 7)) return false;
    if (rsplit && ((o.rdiv & 7) || !m8(o.rx))) return false;
    if (!m8(o.ky) || (ksplit && !m8(o.kx))) return false;
  } else {
    if (o.ky != 1 || L < 8 || (L & 7)) return false;
    if (ksplit && ((o.kdiv & 7) || !m8(o.kx))) return false;
    if (!m8(o.ry) || (rsplit && !m8(o.rx))) return false;
  }
  return true;
}

}  // namespace

template <class D>
static bool b3p_ok(const D &d, const GemmBatch &bt) {
  if (d.M <= 0 || d.N <= 0 || d.nseg <= 0) return false;
  for (int q = 0; q < d.nseg; ++q) {
    const SegDesc &s = d.seg[q];
    if (s.alpha != 1.0 && s.alpha != -1.0) return false;
    if (!op_b3p_ok(s.A, d.M, s.L) || !op_b3p_ok(s.B, d.N, s.L)) return false;
    if (bt.nb > 1 && ((bt.a[q] | bt.bo[q]) & 7)) return false;
  }
  return true;
}

bool gemm_b3p_ok(const GemmDesc &d, const GemmBatch &bt) { return b3p_ok(d, bt); }

bool gemm_b3p_enabled() {
  static const bool on = [] {
    const char *e = getenv("PYSCFAD_AMD_B3_DIRECT");
    return !(e && e[0] == '0');
  }();
  return on;
}

// fraction of a padded MT x NT tiling that is real output
static double b3p_eff(int M, int N, int MT, int NT) {
  const double pm = (double)((M + MT - 1) / MT) * MT, pn = (double)((N + NT - 1) / NT) * NT;
  return (double)M * N / (pm * pn);
}

// Tile shapes {MT, NT}: 8-wave 128 x 80 / 128 x 48 (and transposes), and
// 5-wave 80 x 48 (and transpose) for outputs with one side a multiple of 80
// and the other short (400 x 40 at nvir = 400, nocc = 40: 83% of the padded
// tiling is real output, against 65% with 128 x 48).  The pick maximises the
// padded tiling's efficiency times the tile's flops per staged byte,
// MT NT / (MT + NT).  (A 5-wave 80 x 80 tile for 400 x 400 was measured at
// half the 128 x 80 rate: ten waves per CU instead of sixteen.)
static const int kB3pShape[][2] = {{128, 80}, {80, 128}, {128, 48}, {48, 128}, {80, 48}, {48, 80}};

int gemm_b3p_pick(int M, int N) {
  int best = 0;
  double bs = -1.0;
  for (int c = 0; c < (int)(sizeof(kB3pShape) / sizeof(kB3pShape[0])); ++c) {
    const int MT = kB3pShape[c][0], NT = kB3pShape[c][1];
    const double sc = b3p_eff(M, N, MT, NT) * (double)MT * NT / (MT + NT);
    if (sc > bs + 1e-9) {
      bs = sc;
      best = c;
    }
  }
  return best;
}

int gemm_b3p_tiles(int M, int N) {
  const int c = gemm_b3p_pick(M, N);
  return ((M + kB3pShape[c][0] - 1) / kB3pShape[c][0]) * ((N + kB3pShape[c][1] - 1) / kB3pShape[c][1]);
}

// The products form a group: equal sizes, split-K factor and mode, at most two
// segments each, at most kMaxGroup of them, each eligible for direct loads.
bool gemm_b3p_group_ok(const GemmDesc *ds, int n) {
  if (n < 1 || n > kMaxGroup) return false;
  for (int p = 0; p < n; ++p) {
    const GemmDesc &d = ds[p];
    if (d.nseg > 2 || d.M != ds[0].M || d.N != ds[0].N || d.splits != ds[0].splits ||
        d.mode != ds[0].mode || !gemm_b3p_ok(d))
      return false;
  }
  const GemmDesc &d = ds[0];
  const int c = gemm_b3p_pick(d.M, d.N);
  const long long per = (long long)((d.M + kB3pShape[c][0] - 1) / kB3pShape[c][0]) *
                        ((d.N + kB3pShape[c][1] - 1) / kB3pShape[c][1]) * d.splits;
  return per * n < (1LL << 31);
}

// Group launch of n products (chunk_start of each in stages of 32) that
// satisfy gemm_b3p_group_ok.
hipError_t launch_gemm_b3p_group(const GemmDesc *ds, int n, hipStream_t s) {
  if (!gemm_b3p_group_ok(ds, n)) return hipErrorInvalidValue;
  GemmGroup g;
  memset(&g, 0, sizeof(g));
  g.n = n;
  for (int p = 0; p < n; ++p) {
    const GemmDesc &d = ds[p];
    GemmDescS &e = g.d[p];
    e.M = d.M;
    e.N = d.N;
    e.nseg = d.nseg;
    e.mode = d.mode;
    for (int q = 0; q <= d.nseg; ++q) e.chunk_start[q] = d.chunk_start[q];
    for (int q = d.nseg + 1; q <= 2; ++q) e.chunk_start[q] = d.chunk_start[d.nseg];
    e.splits = d.splits;
    for (int q = 0; q < d.nseg; ++q) e.seg[q] = d.seg[q];
    e.C = d.C;
    e.partial = d.partial;
  }
  const GemmDesc &d = ds[0];
  const int c = gemm_b3p_pick(d.M, d.N);
  const int MT = kB3pShape[c][0], NT = kB3pShape[c][1];
  const long long per = (long long)((d.M + MT - 1) / MT) * ((d.N + NT - 1) / NT) * d.splits;
  const unsigned grid = (unsigned)(per * n);
  const int n_fast = d.M >= d.N ? 1 : 0;
  switch (c) {
    case 0: hipLaunchKernelGGL((k_gemm_b3p_group<128, 80, 8, 1, 4>), dim3(grid), dim3(512), 0, s, g, n_fast); break;
    case 1: hipLaunchKernelGGL((k_gemm_b3p_group<80, 128, 1, 8, 4>), dim3(grid), dim3(512), 0, s, g, n_fast); break;
    case 2: hipLaunchKernelGGL((k_gemm_b3p_group<128, 48, 8, 1, 4>), dim3(grid), dim3(512), 0, s, g, n_fast); break;
    case 3: hipLaunchKernelGGL((k_gemm_b3p_group<48, 128, 1, 8, 4>), dim3(grid), dim3(512), 0, s, g, n_fast); break;
    case 4: hipLaunchKernelGGL((k_gemm_b3p_group<80, 48, 5, 1, 3>), dim3(grid), dim3(320), 0, s, g, n_fast); break;
    default: hipLaunchKernelGGL((k_gemm_b3p_group<48, 80, 3, 1, 2>), dim3(grid), dim3(192), 0, s, g, n_fast); break;
  }
  return hipGetLastError();
}

// d: chunk_start in stages of 32
hipError_t launch_gemm_b3p(const GemmDesc &d, hipStream_t s, const GemmBatch &bt) {
  const int c = gemm_b3p_pick(d.M, d.N);
  const int MT = kB3pShape[c][0], NT = kB3pShape[c][1];
  const unsigned grid =
      (unsigned)((long long)((d.M + MT - 1) / MT) * ((d.N + NT - 1) / NT) * d.splits * bt.nb);
  const int n_fast = d.M >= d.N ? 1 : 0;
  switch (c) {
    case 0: hipLaunchKernelGGL((k_gemm_b3p<128, 80, 8, 1, 4>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case 1: hipLaunchKernelGGL((k_gemm_b3p<80, 128, 1, 8, 4>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case 2: hipLaunchKernelGGL((k_gemm_b3p<128, 48, 8, 1, 4>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case 3: hipLaunchKernelGGL((k_gemm_b3p<48, 128, 1, 8, 4>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case 4: hipLaunchKernelGGL((k_gemm_b3p<80, 48, 5, 1, 3>), dim3(grid), dim3(320), 0, s, d, bt, n_fast); break;
    default: hipLaunchKernelGGL((k_gemm_b3p<48, 80, 3, 1, 2>), dim3(grid), dim3(192), 0, s, d, bt, n_fast); break;
  }
  return hipGetLastError();
}

}  // namespace pamd
