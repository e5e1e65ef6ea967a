// Split-BF16x3 contraction engine for gfx950 (register-staged form).
//
// Every FP64 operand is split once into three bf16 pieces x ~ hi + mid + lo
// (hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)), stored as three
// bf16 planes with the FP64 array's element layout (OpDesc::ps apart), so the
// contraction descriptors are the same as for the FP64 engine.  A product is
// evaluated as the six bf16 MFMA terms
//   hi*hi + (hi*mid + mid*hi) + (hi*lo + lo*hi + mid*mid)
// (dropped: mid*lo, lo*mid, lo*lo, all below 2^-24 relative) with FP32
// accumulation per tile.  Descriptors whose operands allow 16-byte direct
// loads run on the direct-to-LDS form (gemm_b3p.hip); this one takes the rest
// (odd sizes and strides).
#include "ccsdt_common.h"
#include "engine_staging.h"
#include "bf16_split.h"

namespace pamd {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// x (n doubles) -> three bf16 planes of n at out, out + n, out + 2 n
__global__ void k_split_bf16x3(const double *__restrict__ x, size_t n,
                               unsigned short *__restrict__ out) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    put3(out, (long long)n, t, x[t]);
}

hipError_t launch_split_bf16x3(const double *x, size_t n, unsigned short *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  hipLaunchKernelGGL(k_split_bf16x3, dim3((unsigned)b), dim3(256), 0, s, x, n, out);
  return hipGetLastError();
}

// VpT[x][f][p][q] (mode 0) or VpU[x][f][q][p] (mode 1) from Vp[x][p][q][f]; all
// three are planar split-BF16x3 arrays of o v^3 elements (plane stride o v^3).
// One 64 x 64 tile transpose through LDS per block: y is the index that keeps
// its place (p for mode 0, q for mode 1), a the one that moves innermost;
// reads run along f and writes along a, 64 consecutive bf16 per wave.
__global__ __launch_bounds__(256) void k_vp_transpose(const unsigned short *__restrict__ src,
                                                      int o, int v, int mode,
                                                      unsigned short *__restrict__ dst) {
  __shared__ unsigned short t[3][64][65];
  const long long v2 = (long long)v * v, v3 = v2 * v, n = (long long)o * v3;
  const int nt = (v + 63) / 64;
  long long b = blockIdx.x;
  const int tf = (int)(b % nt);
  b /= nt;
  const int ta = (int)(b % nt);
  b /= nt;
  const int y = (int)(b % v);
  const long long x = b / v;
  const int a0 = ta * 64, f0 = tf * 64;
  const unsigned short *S = src + x * v3;
  unsigned short *D = dst + x * v3;
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int al = e >> 6, fl = e & 63;
    const int a = a0 + al, f = f0 + fl;
    if (a < v && f < v) {
      const long long so = (mode == 0 ? y * v2 + a * (long long)v : a * v2 + y * (long long)v) + f;
#pragma unroll
      for (int q = 0; q < 3; ++q) t[q][al][fl] = S[so + q * n];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int fl = e >> 6, al = e & 63;
    const int a = a0 + al, f = f0 + fl;
    if (a < v && f < v) {
      const long long dof = f * v2 + y * (long long)v + a;
#pragma unroll
      for (int q = 0; q < 3; ++q) D[dof + q * n] = t[q][al][fl];
    }
  }
}

hipError_t launch_vp_transposes(const unsigned short *Vp, int o, int v, unsigned short *VpT,
                                unsigned short *VpU, hipStream_t s) {
  const long long nt = (v + 63) / 64;
  const long long blocks = (long long)o * v * nt * nt;
  if (blocks <= 0) return hipSuccess;
  if (blocks >= (1LL << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_vp_transpose, dim3((unsigned)blocks), dim3(256), 0, s, Vp, o, v, 0, VpT);
  hipLaunchKernelGGL(k_vp_transpose, dim3((unsigned)blocks), dim3(256), 0, s, Vp, o, v, 1, VpU);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split-BF16x3 engine, K stages of 32 with planar LDS.
//
// Each stage stages R rows x 32 k of an operand and writes them to LDS as
// three planes (hi, mid, lo) of bf16, 64 bytes per row and plane, so that
// an MFMA operand (8 consecutive k of one row) is one 16-byte LDS read per
// plane with no unpacking.  The four 16-byte chunks of row r are stored
// XOR-swizzled by ((r / 4) ^ (r / 16)) % 4, which keeps the 16-lane groups of
// those reads (16 consecutive rows, one chunk) conflict-free.
// A thread stages "quads" (4 consecutive k of one row, in each of the three
// planes), R / 64 of them per stage, and writes each as one 8-byte LDS store
// per plane:
//   K-contiguous operand: row = tid / 8 + 64 j, quad = tid % 8 (8 lanes cover
//     a row's 32 k; one 8-byte load per plane and quad);
//   row-contiguous operand: row = tid % 64 + 64 j, quad = tid / 64 (64 lanes
//     read 64 consecutive rows at one k per load).
// Accumulation: FP32 MFMA tile sums over four stages (128 k), then added into
// FP64 accumulators.  Two stages of global loads are in flight at a time.
// ---------------------------------------------------------------------------
constexpr int BKC = 32;

template <int R>
struct B3Stage {
  static constexpr int NQ = (R + 63) / 64;  // quads per thread per stage
  static constexpr bool FULL = R % 64 == 0; // otherwise rows row0 + 64 j >= R are skipped
  static constexpr int PLANE = R * 64;
  static constexpr int BYTES = 3 * PLANE;
  static_assert(R % 16 == 0 && R <= 512, "B3 staging map");
};

__device__ __forceinline__ int b3_lds_off(int r, int chunk) {
  return r * 64 + 16 * (chunk ^ (((r >> 2) ^ (r >> 4)) & 3));
}

// this thread's staging map for one operand of one segment
struct B3Rows {
  int rfast, row0, slot;  // rows row0 + 64 j, quad index slot
};

__device__ __forceinline__ B3Rows b3_map(int rfast, int tid) {
  B3Rows m;
  m.rfast = rfast;
  m.row0 = rfast ? (tid & 63) : (tid >> 3);
  m.slot = rfast ? (tid >> 6) : (tid & 7);
  return m;
}

template <int R>
__device__ __forceinline__ void b3_rows(const OpDesc &d, int R_total, int r0, const B3Rows &m,
                                        int (&roff)[B3Stage<R>::NQ],
                                        bool (&rok)[B3Stage<R>::NQ]) {
#pragma unroll
  for (int j = 0; j < B3Stage<R>::NQ; ++j) {
    const int r = r0 + m.row0 + 64 * j;
    rok[j] = r < R_total && (B3Stage<R>::FULL || m.row0 + 64 * j < R);
    const int q = r / d.rdiv;
    roff[j] = rok[j] ? (int)(q * d.rx + (r - q * d.rdiv) * d.ry) : 0;
  }
}

// K position of this thread's quad in the current stage: the element offset
// of its first k, the step between its k (ky) and which of its k lie inside
// the segment (bit u).  With kdiv % 4 == 0 a quad never straddles a kdiv
// boundary (kappa0 % 32 == 0); otherwise `wrap` is set and the offsets are
// formed one by one.
struct B3K {
  int off, step;
  unsigned kmask;
  bool wrap;
};

__device__ __forceinline__ int b3_koff(const OpDesc &d, KPos kp, int add) {
  int r = kp.r + add, q = kp.q;
  while (r >= d.kdiv) {
    r -= d.kdiv;
    ++q;
  }
  return q * (int)d.kx + r * (int)d.ky;
}

__device__ __forceinline__ B3K b3_kpos(const OpDesc &d, int L, int kappa0, KPos kp,
                                       const B3Rows &m) {
  B3K k;
  const int add = 4 * m.slot;
  k.off = b3_koff(d, kp, add);
  k.step = (int)d.ky;
  const int rem = L - (kappa0 + add);
  k.kmask = rem >= 4 ? 15u : (rem <= 0 ? 0u : (1u << rem) - 1u);
  k.wrap = (d.kdiv & 3) != 0;
  return k;
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// the quads of one stage: v[j] = (hi, mid, lo) planes, two dwords (4 bf16) each
template <int R>
__device__ __forceinline__ void b3_fetch(const OpDesc &d, const B3Rows &m,
                                         const int (&roff)[B3Stage<R>::NQ],
                                         const bool (&rok)[B3Stage<R>::NQ], const B3K &k,
                                         KPos kp, bool vec, unsigned (&v)[B3Stage<R>::NQ][6],
                                         unsigned &rowmask, unsigned &kmask) {
  const unsigned short *p = reinterpret_cast<const unsigned short *>(d.p);
  const long long ps = d.ps;
  unsigned rm = 0;
#pragma unroll
  for (int j = 0; j < B3Stage<R>::NQ; ++j) rm |= rok[j] ? (1u << j) : 0u;
  rowmask = rm;
  kmask = k.kmask;
#pragma unroll
  for (int j = 0; j < B3Stage<R>::NQ; ++j) {
    const unsigned msk = rok[j] ? k.kmask : 0u;
    const int base = roff[j] + k.off;
    if (msk == 15u && vec) {
      // four consecutive elements per plane, 8-byte aligned (checked per segment)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const u32x2 x = *reinterpret_cast<const u32x2 *>(p + base + q * ps);
        v[j][2 * q] = x[0];
        v[j][2 * q + 1] = x[1];
      }
    } else {
      unsigned e[3][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        int o = 0;
        if ((msk >> u) & 1u) o = k.wrap ? roff[j] + b3_koff(d, kp, 4 * m.slot + u) : base + u * k.step;
#pragma unroll
        for (int q = 0; q < 3; ++q) e[q][u] = p[o + q * ps];
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        v[j][2 * q] = e[q][0] | (e[q][1] << 16);
        v[j][2 * q + 1] = e[q][2] | (e[q][3] << 16);
      }
    }
  }
}

// a segment's operand qualifies for 8-byte loads of whole quads
__device__ __forceinline__ bool b3_vec_ok(const OpDesc &d) {
  return !d.rfast && d.ky == 1 && ((d.rx | d.ry | d.kx | d.ps) & 3) == 0 && (d.kdiv & 3) == 0 &&
         ((unsigned long long)d.p & 7ull) == 0;
}

template <int R>
__device__ __forceinline__ void b3_stash(const B3Rows &m, bool neg,
                                         const unsigned (&v)[B3Stage<R>::NQ][6],
                                         unsigned rowmask, unsigned kmask, char *S) {
  const unsigned flip = neg ? 0x80008000u : 0u;
#pragma unroll
  for (int j = 0; j < B3Stage<R>::NQ; ++j) {
    if (!B3Stage<R>::FULL && m.row0 + 64 * j >= R) continue;  // beyond the panel
    const unsigned okj = ((rowmask >> j) & 1u) ? kmask : 0u;
    const unsigned keep0 = ((okj & 1u) ? 0xffffu : 0u) | ((okj & 2u) ? 0xffff0000u : 0u);
    const unsigned keep1 = ((okj & 4u) ? 0xffffu : 0u) | ((okj & 8u) ? 0xffff0000u : 0u);
    const u32x2 hi = {(v[j][0] ^ flip) & keep0, (v[j][1] ^ flip) & keep1};
    const u32x2 mid = {(v[j][2] ^ flip) & keep0, (v[j][3] ^ flip) & keep1};
    const u32x2 lo = {(v[j][4] ^ flip) & keep0, (v[j][5] ^ flip) & keep1};
    const int r = m.row0 + 64 * j;
    const int off = b3_lds_off(r, m.slot >> 1) + 8 * (m.slot & 1);
    *(u32x2 *)(S + off) = hi;
    *(u32x2 *)(S + B3Stage<R>::PLANE + off) = mid;
    *(u32x2 *)(S + 2 * B3Stage<R>::PLANE + off) = lo;
  }
}

__device__ __forceinline__ bf16x8 b3_read(const char *S, int plane_bytes, int plane, int r, int chunk) {
  const u32x4 w = *(const u32x4 *)(S + plane * plane_bytes + b3_lds_off(r, chunk));
  return __builtin_bit_cast(bf16x8, w);
}

// Operand data of one stage held in registers between its global loads and
// its LDS stash, with the staging map and sign it was fetched under.
template <int MT, int NT>
struct B3Set {
  unsigned va[B3Stage<MT>::NQ][6], vb[B3Stage<NT>::NQ][6];
  unsigned rma, kma, rmb, kmb;  // valid rows (bit per group) and valid k of the group
  B3Rows mA, mB;
  bool neg;
};

// The stage producer: walks the K stages of this workgroup's split in order,
// across segment boundaries, and fetches one stage per call.
template <int MT, int NT>
struct B3Producer {
  const GemmDesc *d;
  const GemmBatch *bt;
  long long batch;
  int tid, m0, n0, gp, g1;
  int fs, FL, Fend, kappa;
  OpDesc FA, FB;
  bool Fneg, vA, vB;
  KPos ka, kb;
  B3Rows mA, mB;
  int ra[B3Stage<MT>::NQ], rb[B3Stage<NT>::NQ];
  bool raok[B3Stage<MT>::NQ], rbok[B3Stage<NT>::NQ];

  __device__ __forceinline__ void segment(int f) {
    fs = f;
    FA = d->seg[fs].A;
    FB = d->seg[fs].B;
    FA.p = reinterpret_cast<const double *>(reinterpret_cast<const unsigned short *>(FA.p) +
                                            batch * bt->a[fs]);
    FB.p = reinterpret_cast<const double *>(reinterpret_cast<const unsigned short *>(FB.p) +
                                            batch * bt->bo[fs]);
    FL = d->seg[fs].L;
    Fend = d->chunk_start[fs + 1];
    Fneg = d->seg[fs].alpha < 0.0;
    mA = b3_map(FA.rfast, tid);
    mB = b3_map(FB.rfast, tid);
    b3_rows<MT>(FA, d->M, m0, mA, ra, raok);
    b3_rows<NT>(FB, d->N, n0, mB, rb, rbok);
    vA = b3_vec_ok(FA);
    vB = b3_vec_ok(FB);
  }

  __device__ __forceinline__ void start(const GemmDesc &dd, const GemmBatch &bb, int b, int t,
                                        int mm0, int nn0, int g0, int gg1) {
    d = &dd;
    bt = &bb;
    batch = b;
    tid = t;
    m0 = mm0;
    n0 = nn0;
    gp = g0;
    g1 = gg1;
    int f = 0;
    while (f + 1 < dd.nseg && g0 >= dd.chunk_start[f + 1]) ++f;
    segment(f);
    kappa = (g0 - dd.chunk_start[f]) * BKC;
    ka = KPos{kappa / FA.kdiv, kappa - (kappa / FA.kdiv) * FA.kdiv};
    kb = KPos{kappa / FB.kdiv, kappa - (kappa / FB.kdiv) * FB.kdiv};
  }

  __device__ __forceinline__ void produce(B3Set<MT, NT> &S) {
    if (gp >= g1) return;
    const B3K kA = b3_kpos(FA, FL, kappa, ka, mA), kB = b3_kpos(FB, FL, kappa, kb, mB);
    b3_fetch<MT>(FA, mA, ra, raok, kA, ka, vA, S.va, S.rma, S.kma);
    b3_fetch<NT>(FB, mB, rb, rbok, kB, kb, vB, S.vb, S.rmb, S.kmb);
    S.mA = mA;
    S.mB = mB;
    S.neg = Fneg;
    ++gp;
    if (gp >= g1) return;
    if (gp >= Fend) {
      segment(fs + 1);
      kappa = 0;
      ka = KPos{0, 0};
      kb = KPos{0, 0};
    } else {
      kappa += BKC;
      ka.r += BKC;
      while (ka.r >= FA.kdiv) {
        ka.r -= FA.kdiv;
        ++ka.q;
      }
      kb.r += BKC;
      while (kb.r >= FB.kdiv) {
        kb.r -= FB.kdiv;
        ++kb.q;
      }
    }
  }
};

template <int MT, int NT, int WM, int WN>
__global__ __launch_bounds__(512) void k_gemm_b3(GemmDesc d, GemmBatch bt, int n_fast) {
  static_assert(WM * WN == 8, "8 waves");
  constexpr int WTM = MT / WM, WTN = NT / WN;
  constexpr int TM = WTM / 32, TN = WTN / 32;
  static_assert(TM * 32 == WTM && TN * 32 == WTN, "32x32 MFMA tiles");
  using SA = B3Stage<MT>;
  using SB = B3Stage<NT>;
  __shared__ __align__(16) char As[2][SA::BYTES];
  __shared__ __align__(16) char Bs[2][SB::BYTES];
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

  double acc[TM][TN][16];
  f32x16 fa[TM][TN];  // FP32 tile sums (all six products, smallest first)
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc[a][b][r] = 0.0;
        fa[a][b][r] = 0.f;
      }

  // two stages of loads in flight: stage g is stashed while g + 1 is landing
  // and g + 2 is issued
  B3Producer<MT, NT> P;
  P.start(d, bt, batch, tid, m0, n0, g0, g1);
  B3Set<MT, NT> S0, S1;
  P.produce(S0);
  P.produce(S1);
  int buf = 0, nfl = 0;
  constexpr int kFlush = 4;

  auto consume = [&](B3Set<MT, NT> &S, int g) {
    b3_stash<MT>(S.mA, S.neg, S.va, S.rma, S.kma, As[buf]);
    b3_stash<NT>(S.mB, false, S.vb, S.rmb, S.kmb, Bs[buf]);
    __syncthreads();
    P.produce(S);
    const char *SAb = As[buf];
    const char *SBb = Bs[buf];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int chunk = 2 * s + (lane >> 5);
      bf16x8 bh[TN], bm[TN], bl[TN];
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        const int r = wn * WTN + 32 * b + (lane & 31);
        bh[b] = b3_read(SBb, SB::PLANE, 0, r, chunk);
        bm[b] = b3_read(SBb, SB::PLANE, 1, r, chunk);
        bl[b] = b3_read(SBb, SB::PLANE, 2, r, chunk);
      }
#pragma unroll
      for (int a = 0; a < TM; ++a) {
        const int r = wm * WTM + 32 * a + (lane & 31);
        const bf16x8 ah = b3_read(SAb, SA::PLANE, 0, r, chunk);
        const bf16x8 am = b3_read(SAb, SA::PLANE, 1, r, chunk);
        const bf16x8 al = b3_read(SAb, SA::PLANE, 2, r, chunk);
#pragma unroll
        for (int b = 0; b < TN; ++b) {
          fa[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm[b], fa[a][b], 0, 0, 0);
          fa[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[b], fa[a][b], 0, 0, 0);
          fa[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[b], fa[a][b], 0, 0, 0);
          fa[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm[b], fa[a][b], 0, 0, 0);
          fa[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh[b], fa[a][b], 0, 0, 0);
          fa[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[b], fa[a][b], 0, 0, 0);
        }
      }
    }
    // FP32 tile sums go into the FP64 accumulators every kFlush stages (128 k)
    if (++nfl == kFlush || g + 1 == g1) {
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            acc[a][b][r] += (double)fa[a][b][r];
            fa[a][b][r] = 0.f;
          }
      nfl = 0;
    }
    buf ^= 1;
  };
  for (int g = g0; g < g1; g += 2) {
    consume(S0, g);
    if (g + 1 < g1) consume(S1, g + 1);
  }

  // D layout of v_mfma_f32_32x32x16_bf16: lane l, element r -> row 8 (r/4) + 4 (l/32) + r%4,
  // column l%32
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ml = wm * WTM + 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int nl = wn * WTN + 32 * b + (lane & 31);
        const bool part = d.mode == kPartial;
        const OutPos pm = out_pos<MT>(tm, ml, part ? (1 << 30) : d.C.mdiv);
        const OutPos pn = out_pos<NT>(tn, nl, part ? (1 << 30) : d.C.ndiv);
        const int m = pm.idx, n = pn.idx;
        if (m >= d.M || n >= d.N) continue;
        const double val = acc[a][b][r];
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

typedef float f32x4 __attribute__((ext_vector_type(4)));

// The same engine on v_mfma_f32_16x16x32_bf16 (16 x 16 output tiles, one MFMA
// per product per stage), for tile widths that are multiples of 16 but not of
// 32: 80 (5 x 80 = 400 = v at the benchmark size, so the A panel is read five
// times instead of seven and nothing is padded) and 48.  One stage of loads in
// flight (the wider wave tiles need the registers).
// the six split products of one 16x16 tile, smallest first, into one FP32 sum
__device__ __forceinline__ void b3n_products(bf16x8 ah, bf16x8 am, bf16x8 al, bf16x8 bh, bf16x8 bm,
                                             bf16x8 bl, f32x4 &acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

// MINW = 4 (with F32ACC): two workgroups per CU.  The FP32 tile sums then
// cover the workgroup's whole K range (no FP64 accumulators, which would not
// fit 128 VGPRs); the cross-workgroup and split-K sums stay FP64.
template <int MT, int NT, int WM, int WN, int MINW = 2, bool F32ACC = false>
__global__ __launch_bounds__(512, MINW) void k_gemm_b3n(GemmDesc d, GemmBatch bt, int n_fast) {
  static_assert(WM * WN == 8, "8 waves");
  constexpr int WTM = MT / WM, WTN = NT / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  static_assert(TM * 16 == WTM && TN * 16 == WTN, "16x16 MFMA tiles");
  using SA = B3Stage<MT>;
  using SB = B3Stage<NT>;
  __shared__ __align__(16) char As[2][SA::BYTES];
  __shared__ __align__(16) char Bs[2][SB::BYTES];
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

  double acc[F32ACC ? 1 : TM][F32ACC ? 1 : TN][4];
  f32x4 fa[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if constexpr (!F32ACC) acc[a][b][r] = 0.0;
        fa[a][b][r] = 0.f;
      }

  B3Producer<MT, NT> P;
  P.start(d, bt, batch, tid, m0, n0, g0, g1);
  B3Set<MT, NT> S;
  P.produce(S);
  int buf = 0, nfl = 0;
  constexpr int kFlush = 4;
  const int chunk = lane >> 4;
  for (int g = g0; g < g1; ++g) {
    b3_stash<MT>(S.mA, S.neg, S.va, S.rma, S.kma, As[buf]);
    b3_stash<NT>(S.mB, false, S.vb, S.rmb, S.kmb, Bs[buf]);
    __syncthreads();
    P.produce(S);
    const char *SAb = As[buf];
    const char *SBb = Bs[buf];
    // keep the fragments of the narrower side in registers, stream the other
    if constexpr (TM <= TN) {
      bf16x8 ah[TM], am[TM], al[TM];
#pragma unroll
      for (int a = 0; a < TM; ++a) {
        const int r = wm * WTM + 16 * a + (lane & 15);
        ah[a] = b3_read(SAb, SA::PLANE, 0, r, chunk);
        am[a] = b3_read(SAb, SA::PLANE, 1, r, chunk);
        al[a] = b3_read(SAb, SA::PLANE, 2, r, chunk);
      }
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        const int r = wn * WTN + 16 * b + (lane & 15);
        const bf16x8 bh = b3_read(SBb, SB::PLANE, 0, r, chunk);
        const bf16x8 bm = b3_read(SBb, SB::PLANE, 1, r, chunk);
        const bf16x8 bl = b3_read(SBb, SB::PLANE, 2, r, chunk);
#pragma unroll
        for (int a = 0; a < TM; ++a) b3n_products(ah[a], am[a], al[a], bh, bm, bl, fa[a][b]);
      }
    } else {
      bf16x8 bh[TN], bm[TN], bl[TN];
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        const int r = wn * WTN + 16 * b + (lane & 15);
        bh[b] = b3_read(SBb, SB::PLANE, 0, r, chunk);
        bm[b] = b3_read(SBb, SB::PLANE, 1, r, chunk);
        bl[b] = b3_read(SBb, SB::PLANE, 2, r, chunk);
      }
#pragma unroll
      for (int a = 0; a < TM; ++a) {
        const int r = wm * WTM + 16 * a + (lane & 15);
        const bf16x8 ah = b3_read(SAb, SA::PLANE, 0, r, chunk);
        const bf16x8 am = b3_read(SAb, SA::PLANE, 1, r, chunk);
        const bf16x8 al = b3_read(SAb, SA::PLANE, 2, r, chunk);
#pragma unroll
        for (int b = 0; b < TN; ++b) b3n_products(ah, am, al, bh[b], bm[b], bl[b], fa[a][b]);
      }
    }
    if (!F32ACC && (++nfl == kFlush || g + 1 == g1)) {
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if constexpr (!F32ACC) {
              acc[a][b][r] += (double)fa[a][b][r];
              fa[a][b][r] = 0.f;
            }
          }
      nfl = 0;
    }
    buf ^= 1;
  }

  // D layout of v_mfma_f32_16x16x32_bf16: lane l, element r -> row 4 (l/16) + r, column l%16
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
        double val;
        if constexpr (F32ACC)
          val = (double)fa[a][b][r];
        else
          val = acc[a][b][r];
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

// fraction of a padded MT x NT tiling that is real output
static double tile_eff(int M, int N, int MT, int NT) {
  const double pm = (double)((M + MT - 1) / MT) * MT, pn = (double)((N + NT - 1) / NT) * NT;
  return (double)M * N / (pm * pn);
}

// Tile shapes of the split-BF16x3 engine: {MT, NT, workgroups per CU}.  The
// 16x16-MFMA shapes run two workgroups per CU (FP32 tile sums over the whole
// K range of a workgroup), which hides one workgroup's operand staging and
// epilogue behind the other's MFMA work; they are used whenever one of them
// covers the output at least half-efficiently.  Among the candidates the
// score is the padded tiling's efficiency times the tile's flops per staged
// byte, MT NT / (MT + NT).
enum B3Cfg { kB3_128x80 = 0, kB3_80x128, kB3_128x48, kB3_48x128, kB3_256x64, kB3_64x256, kB3Num };
static const int kB3Shape[kB3Num][3] = {{128, 80, 2}, {80, 128, 2}, {128, 48, 2},
                                        {48, 128, 2}, {256, 64, 1}, {64, 256, 1}};
constexpr int kB3Wide = 4;  // configurations 0 .. 3 are the 16x16-MFMA ones

static int b3_pick(int M, int N) {
  double best_eff = 0.0;
  for (int c = 0; c < kB3Wide; ++c)
    best_eff = std::max(best_eff, tile_eff(M, N, kB3Shape[c][0], kB3Shape[c][1]));
  const int ncand = best_eff >= 0.5 ? kB3Wide : kB3Num;
  int best = 0;
  double bs = -1.0;
  for (int c = 0; c < ncand; ++c) {
    const int MT = kB3Shape[c][0], NT = kB3Shape[c][1];
    const double sc = tile_eff(M, N, MT, NT) * (double)MT * NT / (MT + NT);
    if (sc > bs + 1e-9) {
      bs = sc;
      best = c;
    }
  }
  return best;
}

// (sides that are multiples of 8 normally run on the direct engine, with its own tiles)
int gemm_tiles_b3(int M, int N) {
  if (gemm_b3p_enabled() && M % 8 == 0 && N % 8 == 0) return gemm_b3p_tiles(M, N);
  const int c = b3_pick(M, N);
  const int MT = kB3Shape[c][0], NT = kB3Shape[c][1];
  return ((M + MT - 1) / MT) * ((N + NT - 1) / NT);
}

int gemm_wgs_per_cu_b3(int M, int N) {
  if (gemm_b3p_enabled() && M % 8 == 0 && N % 8 == 0) return 2;
  return kB3Shape[b3_pick(M, N)][2];
}

hipError_t launch_gemm_b3(const GemmDesc &d0, hipStream_t s, const GemmBatch &bt) {
  if (d0.M <= 0 || d0.N <= 0) return hipSuccess;
  if (bt.nb < 1 || (bt.nb > 1 && d0.mode == kPartial)) return hipErrorInvalidValue;
  GemmDesc d = d0;  // chunks of BKC
  d.chunk_start[0] = 0;
  for (int q = 0; q < d.nseg; ++q) d.chunk_start[q + 1] = d.chunk_start[q] + (d.seg[q].L + BKC - 1) / BKC;
  for (int q = d.nseg + 1; q <= kMaxSeg; ++q) d.chunk_start[q] = d.chunk_start[d.nseg];
  if (gemm_b3p_enabled() && gemm_b3p_ok(d, bt)) return launch_gemm_b3p(d, s, bt);
  const int c = b3_pick(d.M, d.N);
  const int MT = kB3Shape[c][0], NT = kB3Shape[c][1];
  const unsigned grid =
      (unsigned)((long long)((d.M + MT - 1) / MT) * ((d.N + NT - 1) / NT) * d.splits * bt.nb);
  const int n_fast = d.M >= d.N ? 1 : 0;
  switch (c) {
    case kB3_256x64: hipLaunchKernelGGL((k_gemm_b3<256, 64, 4, 2>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case kB3_64x256: hipLaunchKernelGGL((k_gemm_b3<64, 256, 2, 4>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case kB3_128x80: hipLaunchKernelGGL((k_gemm_b3n<128, 80, 8, 1, 4, true>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case kB3_80x128: hipLaunchKernelGGL((k_gemm_b3n<80, 128, 1, 8, 4, true>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    case kB3_128x48: hipLaunchKernelGGL((k_gemm_b3n<128, 48, 8, 1, 4, true>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
    default: hipLaunchKernelGGL((k_gemm_b3n<48, 128, 1, 8, 4, true>), dim3(grid), dim3(512), 0, s, d, bt, n_fast); break;
  }
  return hipGetLastError();
}

hipError_t launch_gemm_bf16x3(const GemmDesc &d, hipStream_t s, const GemmBatch &bt) {
  return launch_gemm_b3(d, s, bt);
}

// chunk_start of a descriptor in stages of BKC
static GemmDesc to_stages(const GemmDesc &d0) {
  GemmDesc d = d0;
  d.chunk_start[0] = 0;
  for (int q = 0; q < d.nseg; ++q) d.chunk_start[q + 1] = d.chunk_start[q] + (d.seg[q].L + BKC - 1) / BKC;
  for (int q = d.nseg + 1; q <= kMaxSeg; ++q) d.chunk_start[q] = d.chunk_start[d.nseg];
  return d;
}

hipError_t launch_gemm_bf16x3_group(const GemmDesc *ds, int n, hipStream_t s) {
  if (n >= 2 && n <= kMaxGroup && gemm_b3p_enabled()) {
    GemmDesc g[kMaxGroup];
    for (int p = 0; p < n; ++p) g[p] = to_stages(ds[p]);
    if (gemm_b3p_group_ok(g, n)) return launch_gemm_b3p_group(g, n, s);
  }
  for (int p = 0; p < n; ++p) {
    const hipError_t e = launch_gemm_b3(ds[p], s, no_batch());
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace pamd
