// Host driver and C ABI for the (T) energy and its adjoint.
//
// Walk: occupied triples i <= j <= k (i == j == k contributes nothing: X is then
// fully symmetric and r3(X) = 0).  Per triple:
//   1. W build -- three segmented MFMA GEMMs write W_ijk[a,b,c] (v^3):
//        pass C  (rows (a,b), cols c):  P1 + P3 + H5 + H6
//        pass B  (rows b, cols (a,c)):  P2 + P5 + H3 + H4
//        pass A  (rows a, cols (b,c)):  P4 + P6 + H1 + H2
//      (P1..P6 / H1..H6: the six joint permutations of the particle / hole
//      term of w_abc[ijk], pyscfadlib/pyscfadlib/vjp/cc/ccsd_t.c:121-142 with
//      the permutation tables of ccsd_t.c:74-96).
//   2. orbit epilogue: V, Z = r3(W + V), 1/D, energy; for the adjoint also
//      Wbar, Vbar and dE/d(mo_energy) (ccsd_t_vjp.c:42-117).
//   3. adjoint (ccsd_t_vjp.c:139-269 restated as large contractions):
//        ovvv_bar  -- 3 GEMMs into the Vp-layout accumulator (rows = pairs)
//        t2_bar    -- 6 split-K GEMMs (K = a virtual pair) + V-term pieces
//        ovoo_bar  -- 6 split-K GEMMs (hole terms)
//        t2_bar    -- 3 GEMMs for the hole-term t2 reads (T2m layout)
//        ovov_bar, t1_bar, fvo_bar -- Vbar contractions (k_vr_*) + gather.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>
#include <atomic>
#include <chrono>
#include <mutex>
#include "ccsdt_common.h"
#include "gemm_desc.h"
#include "../../include/pyscfad_amd_ccsdt.h"
#include "../../include/pyscfad_amd_debug.h"

namespace pamd {

static const int kGmax = 1024;

// When set, GEMM launches are recorded instead of run and the other per-triple
// kernels are skipped (host-only descriptor inspection, no GPU work).  Per host
// thread, like every other piece of per-call state here.
static thread_local std::vector<GemmDesc> *g_record = nullptr;

// Sampled launch timing (pamd_profile_*): for one triple in g_prof_every, every
// launch is bracketed by hipEvents on its own stream; the events are read after
// the call has been queued completely.
// GEMM kinds: W build; adjoint to ovvv (Vpb), to t2 via split-K partials
// (particle), to ovoo (hole split-K partials), to t2 through T2mb (hole)
enum ProfKind {
  kProfWBuild = 0,
  kProfAdjOvvv,
  kProfAdjT2P,
  kProfAdjOvoo,
  kProfAdjT2H,
  kProfOrbit,
  kProfVR,
  kProfCombine,
  kProfNum
};
static const char *kProfNames[kProfNum] = {"gemm_wbuild", "gemm_adj_ovvv", "gemm_adj_t2p",
                                           "gemm_adj_ovoo", "gemm_adj_t2h", "orbit_epilogue",
                                           "vbar_reduce", "combine"};
struct ProfRec {
  int kind;
  hipEvent_t a, b;
  double flops, bytes;
};
struct ProfStat {
  long long launches;
  double ms, flops, bytes;
};
// The switches and the accumulated statistics are process-wide (read and reset
// through pamd_profile_*, the statistics under g_prof_mutex); whether the current
// triple is sampled, and its pending event records, belong to the calling thread,
// so concurrent calls from different host threads never see each other's state.
static std::atomic<int> g_prof_every{0};
static std::atomic<int> g_progress_every{0};  // report (to stderr) every N triples, after a stream sync
static thread_local bool g_prof_on = false;  // sampling the current triple
// set for the duration of a call that must not allocate: no event records then
static thread_local bool g_prof_blocked = false;
static thread_local std::vector<ProfRec> g_prof_recs;
static ProfStat g_prof_stat[kProfNum];
static std::mutex g_prof_mutex;

static hipError_t prof_begin(hipStream_t s, ProfRec &r, int kind, double flops, double bytes) {
  r.kind = kind;
  r.flops = flops;
  r.bytes = bytes;
  hipError_t e = hipEventCreate(&r.a);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&r.b);
  if (e != hipSuccess) return e;
  return hipEventRecord(r.a, s);
}

static hipError_t prof_end(hipStream_t s, ProfRec &r) {
  hipError_t e = hipEventRecord(r.b, s);
  g_prof_recs.push_back(r);
  return e;
}

static hipError_t prof_collect() {
  hipError_t first = hipSuccess;
  for (ProfRec &r : g_prof_recs) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(r.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.a, r.b);
    if (e == hipSuccess) {
      std::lock_guard<std::mutex> lock(g_prof_mutex);
      ProfStat &st = g_prof_stat[r.kind];
      st.launches += 1;
      st.ms += ms;
      st.flops += r.flops;
      st.bytes += r.bytes;
    } else if (first == hipSuccess) {
      first = e;
    }
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  g_prof_recs.clear();
  return first;
}

static double gemm_flops(const GemmDesc &d) {
  double k = 0;
  for (int q = 0; q < d.nseg; ++q) k += d.seg[q].L;
  return 2.0 * d.M * (double)d.N * k;
}

// timed launch wrapper for any kernel
template <class F>
static hipError_t timed(hipStream_t s, int kind, double flops, double bytes, F &&launch) {
  if (!g_prof_on || g_prof_blocked) return launch();
  ProfRec r;
  hipError_t e = prof_begin(s, r, kind, flops, bytes);
  if (e != hipSuccess) return e;
  e = launch();
  if (e != hipSuccess) return e;
  return prof_end(s, r);
}

struct Layout {
  size_t off_Vp, off_T2m, off_OOt, off_OVS, off_T2S, off_orb, off_W, off_tri, off_part_t, off_T2g, off_W8;
  int wbatch;  // W blocks of the energy walk's batched W build (1 with the adjoint)
  size_t off_Vb, off_Vpb, off_T2mb, off_part_mo, off_R, off_part_p, off_part_h, off_wt;
  // transposed planar copies of Vp (split-BF16x3 W build; see vpt_fits)
  int vpt;
  size_t off_VpT, off_VpU;
  // pipelined adjoint walk: second copies of the per-triple buffers
  int pipe;
  size_t off_W2, off_W8b, off_R2, off_part_p2, off_part_h2;
  size_t total;
  int S_p, S_h;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static int device_cus() {
  static int cus = [] {
    int dev = 0, n = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      n = prop.multiProcessorCount;
    return n > 0 ? n : 256;
  }();
  return cus;
}

// Split-K factor for an M x N output with `chunks` K chunks of 16: one full
// round of resident workgroups (two per CU for the FP64 engine and for the
// split-BF16x3 16x16-MFMA tiles, one for its 32x32 tiles), rounded down so
// that no workgroup waits for a second round.  (One round measured faster than two or three: fewer split-K
// partials to write and reduce, same engine time.)
// The split-BF16x3 two-workgroup tiles sum each output tile in FP32 over the
// workgroup's whole K range, so the split count is also raised until no part
// spans more than kB3MaxK terms: the FP32 rounding of those sums then stays
// bounded as v grows (K = v^2 for the particle-side t2 adjoint).  At v = 400 one
// round of workgroups already gives 25 parts of 6400 terms, so this bound only
// acts at larger v.
constexpr int kB3MaxK = 8192;
static int splits_for(int M, int N, int chunks, int prec) {
  const bool b3 = prec == PAMD_PREC_BF16X3;
  const int tiles = std::max(1, b3 ? gemm_tiles_b3(M, N) : gemm_tiles(M, N, prec == PAMD_PREC_FP32X2));
  const int slots = (b3 ? gemm_wgs_per_cu_b3(M, N) : 2) * device_cus();
  int s = slots / tiles;
  if (b3) s = std::max(s, (int)(((long long)chunks * KC + kB3MaxK - 1) / kB3MaxK));
  s = std::max(1, std::min(s, 512));
  return std::min(s, std::max(1, b3 ? chunks / 2 : chunks));
}

// Consecutive k of one occupied pair (i, j) share the W build of the energy walk
// when one triple's W-build GEMM has too few workgroup tiles to fill the GPU:
// about four rounds of the resident workgroups per launch, at most 8 triples,
// their W blocks within 2 GB.  (At nvir = 400 one triple already has ~6000
// tiles: no batching.)
static int energy_batch(int o, int v, int prec) {
  const long long tiles = prec == PAMD_PREC_BF16X3 ? gemm_tiles_b3(v * v, v)
                                                    : gemm_tiles(v * v, v, prec == PAMD_PREC_FP32X2);
  const long long slots = 2LL * device_cus();
  long long b = (4 * slots + tiles - 1) / std::max(1LL, tiles);
  b = std::max(1LL, std::min(8LL, b));
  const long long v3 = (long long)v * v * v;
  while (b > 1 && b * v3 * (long long)sizeof(double) > (2LL << 30)) --b;
  return (int)std::min<long long>(b, o);
}

// Vp, T2m and OOt hold FP64 copies, or their split forms: three bf16 planes
// (split-BF16x3, 6 of the 8 bytes per element) or packed FP32 hi/lo pairs; the
// split precisions add split copies of t2 and of Wbar.
// Transposed planar copies of Vp make every particle-term ovvv operand of the
// split-BF16x3 W build row-contiguous, the faster of the direct engine's two
// staging forms (the big operand of all three passes).  They take two o v^3
// arrays; held when the direct engine runs the W build (nocc and nvir
// multiples of 8) and the copies stay within 48 GB.
static bool vpt_fits(int o, int v, int prec) {
  const double bytes = 2.0 * 8.0 * o * (double)v * v * v;
  return prec == PAMD_PREC_BF16X3 && o % 8 == 0 && v % 8 == 0 && bytes <= 48e9 &&
         gemm_b3p_enabled();
}

// pipe: the adjoint walk runs two triples in flight (run_vjp_pipelined) and
// needs second copies of W, Wbar, the V-bar scratch and the split-K partials
static Layout make_layout(int o, int v, int vjp, int prec, bool pipe = false) {
  Layout L;
  memset(&L, 0, sizeof(L));
  const size_t d = sizeof(double);
  const size_t v2 = (size_t)v * v, v3 = v2 * v, n = (size_t)o + v;
  size_t off = 0;
  auto take = [&](size_t count) {
    size_t at = off;
    off = align256(off + count * d);
    return at;
  };
  L.off_Vp = take((size_t)o * v3);
  L.off_T2m = take((size_t)o * o * v2);
  L.off_OOt = take((size_t)o * o * o * v);
  L.off_OVS = take((size_t)o * o * v2);
  L.off_T2S = take((size_t)o * o * v2);
  {
    const size_t nt = (v + 7) / 8;
    L.off_orb = take(nt * (nt + 1) * (nt + 2) / 6);  // 4 x uint16 per orbit = one double
  }
  L.wbatch = vjp ? 1 : energy_batch(o, v, prec);
  L.off_W = take(v3 * L.wbatch);
  L.off_tri = take((size_t)o * o * o);
  L.off_part_t = take(kGmax);
  L.off_wt = take((size_t)o);
  if (prec == PAMD_PREC_BF16X3 || prec == PAMD_PREC_FP32X2) {
    L.off_T2g = take((size_t)o * o * v2);
    if (vjp) L.off_W8 = take(v3);
  }
  const int chunks = (int)((v2 + KC - 1) / KC);
  L.S_p = splits_for(v, v, chunks, prec);
  L.S_h = splits_for(v, o, chunks, prec);
  if (vjp) {
    L.off_Vb = take(v3);
    L.off_Vpb = take((size_t)o * v3);
    L.off_T2mb = take((size_t)o * o * v2);
    L.off_part_mo = take((size_t)kGmax * n);
    L.off_R = take(vr_scratch_doubles(v));
    L.off_part_p = take((size_t)6 * L.S_p * v2);
    L.off_part_h = take((size_t)6 * L.S_h * v * o);
  }
  L.vpt = vpt_fits(o, v, prec) ? 1 : 0;
  if (L.vpt) {
    L.off_VpT = take((size_t)o * v3);
    L.off_VpU = take((size_t)o * v3);
  }
  L.pipe = (vjp && pipe) ? 1 : 0;
  if (L.pipe) {
    L.off_W2 = take(v3);
    if (prec == PAMD_PREC_BF16X3 || prec == PAMD_PREC_FP32X2) L.off_W8b = take(v3);
    L.off_R2 = take(vr_scratch_doubles(v));
    L.off_part_p2 = take((size_t)6 * L.S_p * v2);
    L.off_part_h2 = take((size_t)6 * L.S_h * v * o);
  }
  L.total = off;
  return L;
}

// Planar split-BF16x3 operands: the arrays the walk's descriptors point into
// (Vp, T2m, OOt, t2 copy, Wbar copy), allocated as FP64 (8 bytes per element)
// and holding three bf16 planes of their element count in the first 6 bytes.
// A descriptor built with FP64 element arithmetic is moved onto them: element
// e of an array becomes bf16 element e of its hi plane, with the array's
// element count as the plane stride.
struct PlanarMap {
  const char *base[8];
  long long n[8];
  int count;
};

struct Ctx {
  int o, v, packed, has_fvo, vjp, prec;
  const double *gT2;  // t2 as a GEMM operand (FP64, or its split copy)
  unsigned short *orb;  // orbit table
  double *gW;         // Wbar as a GEMM operand (= W, or its split copy)
  hipStream_t s;
  const double *mo, *t1, *t2, *ovvv, *ovoo, *ovov, *fvo;
  double *Vp, *T2m, *OOt, *OVS, *T2S, *W, *tri, *part_t;
  double *Vb, *Vpb, *T2mb, *part_mo, *R, *part_p, *part_h;
  int S_p, S_h, G;
  int wbatch;  // W blocks available to the energy walk's batched W build
  double ebar;
  int shard, nshards;     // this call covers walk triples with index % nshards == shard
  const double *wt_host;  // occupied weights m (host) of a diagonal projector, or null
  double *wt;             // device copy of m
  PlanarMap planes;       // split-BF16x3 operand arrays
  // transposed planar copies of Vp for the W build (split-BF16x3, when the
  // layout holds them): VpT[x][f][p][q] = VpU[x][f][q][p] = Vp[x][p][q][f]
  double *VpT, *VpU;
  // pipelined adjoint walk (pipe = 1): the second set of per-triple buffers
  int pipe;
  double *W2, *gW2, *R2, *part_p2, *part_h2;
  // outputs
  double *t1_bar, *t2_bar, *ovoo_bar, *ovov_bar, *fvo_bar;
};

static bool planarize(const PlanarMap &pm, OpDesc &o) {
  const char *p = (const char *)o.p;
  for (int a = 0; a < pm.count; ++a) {
    if (p < pm.base[a] || p >= pm.base[a] + pm.n[a] * (long long)sizeof(double)) continue;
    const long long e = (p - pm.base[a]) / (long long)sizeof(double);
    o.p = (const double *)(pm.base[a] + 2 * e);
    o.ps = pm.n[a];
    return true;
  }
  return false;
}

// one contraction in the precision of the call that issues it
static hipError_t run_gemm(const Ctx &c, const GemmDesc &d0, int kind,
                           const GemmBatch &bt = no_batch()) {
  if (g_record) {
    g_record->push_back(d0);
    return hipSuccess;
  }
  const double fl = gemm_flops(d0) * bt.nb;
  const hipStream_t s = c.s;
  if (c.prec == PAMD_PREC_BF16X3) {
    GemmDesc d = d0;
    for (int q = 0; q < d.nseg; ++q)
      if (!planarize(c.planes, d.seg[q].A) || !planarize(c.planes, d.seg[q].B))
        return hipErrorInvalidValue;
    return timed(s, kind, fl, 0.0, [&] { return launch_gemm_bf16x3(d, s, bt); });
  }
  if (c.prec == PAMD_PREC_FP32X2)
    return timed(s, kind, fl, 0.0, [&] { return launch_gemm_f32x2(d0, s, bt); });
  return timed(s, kind, fl, 0.0, [&] { return launch_gemm_f64(d0, s, bt); });
}

// Group launches of the adjoint families: env PYSCFAD_AMD_GROUP = 0 never,
// 1 always; unset: when the three bf16 planes of Wbar (6 v^3 bytes, read by
// every product of a family) fit in half of the GPU's 256 MB last-level
// cache.  Measured E+VJP: o16/v160 0.634 -> 0.602 s and o24/v240 5.61 ->
// 5.44 s with groups, o40/v400 134.7 -> 138.6 s (the products of a group then
// stream different ovvv slabs and Wbar orders through the cache at once).
static int group_mode() {
  static const int m = [] {
    const char *e = getenv("PYSCFAD_AMD_GROUP");
    return e ? (e[0] == '0' ? 0 : 1) : -1;
  }();
  return m;
}
constexpr double kGroupWbarBytes = 128.0 * (1 << 20);

// n independent contractions of one kind: in the split-BF16x3 form one group
// launch when their outputs are disjoint (out_disjoint), they fit the direct
// engine's group and grouping is on for their size, else one launch each
static hipError_t run_gemm_group(const Ctx &c, const GemmDesc *ds, int n, int kind,
                                 bool out_disjoint) {
  const int gm = group_mode();
  const bool group = gm == 1 || (gm < 0 && 6.0 * c.v * (double)c.v * c.v <= kGroupWbarBytes);
  if (g_record || c.prec != PAMD_PREC_BF16X3 || !out_disjoint || n > kMaxGroup || !group) {
    for (int q = 0; q < n; ++q) {
      const hipError_t e = run_gemm(c, ds[q], kind);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  GemmDesc d[kMaxGroup];
  double fl = 0.0;
  for (int q = 0; q < n; ++q) {
    d[q] = ds[q];
    for (int p = 0; p < d[q].nseg; ++p)
      if (!planarize(c.planes, d[q].seg[p].A) || !planarize(c.planes, d[q].seg[p].B))
        return hipErrorInvalidValue;
    fl += gemm_flops(d[q]);
  }
  return timed(c.s, kind, fl, 0.0, [&] { return launch_gemm_bf16x3_group(d, n, c.s); });
}

#define PAMD_CHECK(expr)                                                      \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "pyscfad_amd: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return PAMD_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

// ---------------------------------------------------------------------------
// W build for one triple
// ---------------------------------------------------------------------------
// nb > 1: the W blocks of triples (i, j, k), (i, j, k + 1), ..., (i, j, k + nb - 1)
// in one launch per pass, written to W + b v^3 (every operand slab that names
// k moves by a fixed stride per step of k).
static int build_w(const Ctx &c, int i, int j, int k, int nb = 1) {
  const long long o = c.o, v = c.v;
  const long long v2 = v * v, v3 = v2 * v, vo = v * o;
  // element strides per step of k of: Vp(k), T2(k, x), T2(x, k), OOt(k, x), OOt(x, k), T2m(k)
  const long long sVp = v3, sT2k_ = o * v2, sT2_k = v2, sOOk_ = o * vo, sOO_k = vo, sT2m = v2 * o;
  auto batch = [&](std::initializer_list<long long> a, std::initializer_list<long long> bo) {
    GemmBatch bt = no_batch();
    int q = 0;
    for (long long x : a) bt.a[q++] = x;
    q = 0;
    for (long long x : bo) bt.bo[q++] = x;
    bt.c = v3;
    bt.nb = nb;
    return bt;
  };
  auto Vp = [&](int x) { return c.Vp + x * v3; };
  auto T2 = [&](int x, int y) { return c.gT2 + (x * o + y) * v2; };
  auto T2m = [&](int x) { return c.T2m + x * v2 * o; };
  auto OOt = [&](int x, int y) { return c.OOt + (x * o + y) * vo; };
  const int V = c.v, O = c.o;
  // the ovvv slab of a particle term with its row pair (q, s) read as
  // Vp(x)[q][s][f] ("qs") or Vp(x)[s][q][f] ("sq"), K = f: from Vp itself
  // (K-contiguous), or, when the call holds the transposed copies, from
  // VpT(x)[f][q][s] / VpU(x)[f][q][s] (row-contiguous)
  auto Pqs = [&](int x) {
    return c.VpT ? opR(c.VpT + x * v3, v, 1, V, v2, 1) : opR(Vp(x), v2, v, V, 1, 0);
  };
  auto Psq = [&](int x) {
    return c.VpU ? opR(c.VpU + x * v3, v, 1, V, v2, 1) : opR(Vp(x), v, v2, V, 1, 0);
  };
  // pass C: rows (a,b), cols c
  {
    GemmDesc d = gemm(V * V, V, kStore,
                      {seg(Pqs(i), op1(T2(k, j), v, 1, 0), V, 1.0),      // P1
                       seg(Psq(j), op1(T2(k, i), v, 1, 0), V, 1.0),      // P3
                       seg(opR(T2m(j), o, vo, V, 1, 0), op1(OOt(k, i), o, 1, 0), O, -1.0),   // H5
                       seg(opR(T2m(i), vo, o, V, 1, 0), op1(OOt(k, j), o, 1, 0), O, -1.0)},  // H6
                      out(c.W, v2, v, V, 0, 1, kBig));
    // P1: T2(k,j), P3: T2(k,i), H5: OOt(k,i), H6: OOt(k,j)
    PAMD_CHECK(run_gemm(c, d, kProfWBuild,
                        batch({0, 0, 0, 0}, {sT2k_, sT2k_, sOOk_, sOOk_})));
  }
  // pass B: rows b, cols (a,c)
  {
    GemmDesc d = gemm(V, V * V, kAccumulate,
                      {seg(op1(T2(j, k), v, 1, 0), Pqs(i), V, 1.0),      // P2
                       seg(op1(T2(j, i), v, 1, 0), Psq(k), V, 1.0),      // P5
                       seg(op1(OOt(j, i), o, 1, 0), opR(T2m(k), o, vo, V, 1, 0), O, -1.0),   // H3
                       seg(op1(OOt(j, k), o, 1, 0), opR(T2m(i), vo, o, V, 1, 0), O, -1.0)},  // H4
                      out(c.W, 0, v, kBig, v2, 1, V));
    // P2: T2(j,k) | Vp(i), P5: T2(j,i) | Vp(k), H3: OOt(j,i) | T2m(k), H4: OOt(j,k) | T2m(i)
    PAMD_CHECK(run_gemm(c, d, kProfWBuild,
                        batch({sT2_k, 0, 0, sOO_k}, {0, sVp, sT2m, 0})));
  }
  // pass A: rows a, cols (b,c)
  {
    GemmDesc d = gemm(V, V * V, kAccumulate,
                      {seg(op1(T2(i, k), v, 1, 0), Pqs(j), V, 1.0),      // P4
                       seg(op1(T2(i, j), v, 1, 0), Psq(k), V, 1.0),      // P6
                       seg(op1(OOt(i, j), o, 1, 0), opR(T2m(k), o, vo, V, 1, 0), O, -1.0),   // H1
                       seg(op1(OOt(i, k), o, 1, 0), opR(T2m(j), vo, o, V, 1, 0), O, -1.0)},  // H2
                      out(c.W, 0, v2, kBig, v, 1, V));
    // P4: T2(i,k) | Vp(j), P6: T2(i,j) | Vp(k), H1: OOt(i,j) | T2m(k), H2: OOt(i,k) | T2m(j)
    PAMD_CHECK(run_gemm(c, d, kProfWBuild,
                        batch({sT2_k, 0, 0, sOO_k}, {0, sVp, sT2m, 0})));
  }
  return PAMD_OK;
}

// ---------------------------------------------------------------------------
// adjoint of one triple (c.gW holds Wbar, c.Vb holds Vbar): the V-bar
// reductions into c.R, the adjoint contractions (stream c.s) into Vpb, T2mb
// and the split-K partials c.part_p, c.part_h, and the gather of R and the
// partials into the cotangents
// ---------------------------------------------------------------------------
static int adjoint_vr(const Ctx &c, int i, int j, int k, hipStream_t st) {
  const double v3 = (double)c.v * c.v * c.v;
  VRArgs vr;
  vr.Vb = c.Vb;
  vr.t1 = c.t1;
  vr.fvo = c.fvo;
  vr.ovs = c.OVS;
  vr.t2s = c.T2S;
  vr.o = c.o;
  vr.v = c.v;
  vr.i = i;
  vr.j = j;
  vr.k = k;
  vr.has_fvo = c.has_fvo;
  vr.R = c.R;
  if (!g_record) {
    const double vb = 2.0 * 8.0 * v3;  // Vbar read twice
    PAMD_CHECK(timed(st, kProfVR, 0.0, vb, [&] { return launch_vr(vr, st); }));
  }
  return PAMD_OK;
}

// pair: 0 all three ovvv adjoints (and hole-term t2 adjoints) of the triple;
// 1 only those into Vpb[k] and T2mb[k] (those into index i and j wait for the
// next triple (i, j, k'), which shares them); 2 those into Vpb[i], Vpb[j],
// T2mb[i], T2mb[j] as four-segment launches that also take the terms of the
// previous triple (i, j, kp) from its own Wbar buffer (prev->gW), plus the
// ones into Vpb[k] and T2mb[k]
static int adjoint_gemms(const Ctx &c, int i, int j, int k, int pair = 0, const Ctx *prev = nullptr,
                         int kp = 0) {
  const long long o = c.o, v = c.v;
  const long long v2 = v * v, v3 = v2 * v, vo = v * o;
  const int V = c.v, O = c.o;
  const double *Wb = c.gW;
  auto Vp = [&](int x) { return c.Vp + x * v3; };
  auto Vpb = [&](int x) { return c.Vpb + x * v3; };
  auto T2 = [&](int x, int y) { return c.gT2 + (x * o + y) * v2; };
  auto T2m = [&](int x) { return c.T2m + x * v2 * o; };
  auto T2mb = [&](int x) { return c.T2mb + x * v2 * o; };
  auto OOt = [&](int x, int y) { return c.OOt + (x * o + y) * vo; };

  // the three adjoints into Vpb[i], Vpb[j], Vpb[k] and into T2mb[i], T2mb[j],
  // T2mb[k] write disjoint slabs unless two occupied indices coincide
  const bool distinct = i != j && j != k;
  // ovvv_bar, rows = virtual pairs, cols = f
  if (pair != 0) {
    const double *Wq = prev ? prev->gW : nullptr;
    GemmDesc ds[3];
    int n = 0;
    if (pair == 2) {
      ds[n++] = gemm(V * V, V, kAccumulate,
                     {seg(opR(Wq, v2, v, V, 1, 0), op1(T2(kp, j), 1, v, 1), V, 1.0),
                      seg(opR(Wq, v2, 1, V, v, 1), op1(T2(j, kp), 1, v, 1), V, 1.0),
                      seg(opR(Wb, v2, v, V, 1, 0), op1(T2(k, j), 1, v, 1), V, 1.0),
                      seg(opR(Wb, v2, 1, V, v, 1), op1(T2(j, k), 1, v, 1), V, 1.0)},
                     out(Vpb(i), v2, v, V, 0, 1, kBig));
      ds[n++] = gemm(V * V, V, kAccumulate,
                     {seg(opR(Wq, v, v2, V, 1, 0), op1(T2(kp, i), 1, v, 1), V, 1.0),
                      seg(opR(Wq, v, 1, V, v2, 1), op1(T2(i, kp), 1, v, 1), V, 1.0),
                      seg(opR(Wb, v, v2, V, 1, 0), op1(T2(k, i), 1, v, 1), V, 1.0),
                      seg(opR(Wb, v, 1, V, v2, 1), op1(T2(i, k), 1, v, 1), V, 1.0)},
                     out(Vpb(j), v2, v, V, 0, 1, kBig));
    }
    ds[n++] = gemm(V * V, V, kAccumulate,
                   {seg(opR(Wb, v2, 1, V, v, 1), op1(T2(j, i), 1, v, 1), V, 1.0),
                    seg(opR(Wb, v, 1, V, v2, 1), op1(T2(i, j), 1, v, 1), V, 1.0)},
                   out(Vpb(k), v, v2, V, 0, 1, kBig));
    for (int q = 0; q < n; ++q) PAMD_CHECK(run_gemm(c, ds[q], kProfAdjOvvv));
  } else {
    const GemmDesc ds[3] = {
        // Vpb[i][a][b|c][f]: P1 (K=c), P2 (K=b)
        gemm(V * V, V, kAccumulate,
             {seg(opR(Wb, v2, v, V, 1, 0), op1(T2(k, j), 1, v, 1), V, 1.0),
              seg(opR(Wb, v2, 1, V, v, 1), op1(T2(j, k), 1, v, 1), V, 1.0)},
             out(Vpb(i), v2, v, V, 0, 1, kBig)),
        // Vpb[j][b][a|c][f]: P3 (K=c), P4 (K=a)
        gemm(V * V, V, kAccumulate,
             {seg(opR(Wb, v, v2, V, 1, 0), op1(T2(k, i), 1, v, 1), V, 1.0),
              seg(opR(Wb, v, 1, V, v2, 1), op1(T2(i, k), 1, v, 1), V, 1.0)},
             out(Vpb(j), v2, v, V, 0, 1, kBig)),
        // Vpb[k][c][a|b][f]: rows m = y*v + x with x = c: P5 (K=b), P6 (K=a)
        gemm(V * V, V, kAccumulate,
             {seg(opR(Wb, v2, 1, V, v, 1), op1(T2(j, i), 1, v, 1), V, 1.0),
              seg(opR(Wb, v, 1, V, v2, 1), op1(T2(i, j), 1, v, 1), V, 1.0)},
             out(Vpb(k), v, v2, V, 0, 1, kBig)),
    };
    PAMD_CHECK(run_gemm_group(c, ds, 3, kProfAdjOvvv, distinct));
  }

  // particle t2_bar pieces: split-K over the virtual pair, partials per problem
  const size_t pp = (size_t)c.S_p * v2;
  {
    const GemmDesc ds[6] = {
        // P1: t2_bar[k,j][c][f] = sum_(a,b) Wb[a,b,c] Vp(i)[a][b][f]
        gemm(V, V, kPartial, {seg(opK(Wb, 1, v2, v, V, 1), opK(Vp(i), 1, v2, v, V, 1), V * V, 1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_p, c.part_p + 0 * pp),
        // P2: t2_bar[j,k][b][f] = sum_(a,c) Wb[a,b,c] Vp(i)[a][c][f]
        gemm(V, V, kPartial, {seg(opK(Wb, v, v2, 1, V, 0), opK(Vp(i), 1, v2, v, V, 1), V * V, 1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_p, c.part_p + 1 * pp),
        // P3: t2_bar[k,i][c][f] = sum_(a,b) Wb[a,b,c] Vp(j)[b][a][f]
        gemm(V, V, kPartial, {seg(opK(Wb, 1, v2, v, V, 1), opK(Vp(j), 1, v, v2, V, 1), V * V, 1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_p, c.part_p + 2 * pp),
        // P4: t2_bar[i,k][a][f] = sum_(b,c) Wb[a,b,c] Vp(j)[b][c][f]
        gemm(V, V, kPartial, {seg(opK(Wb, v2, v, 1, V, 0), opK(Vp(j), 1, v2, v, V, 1), V * V, 1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_p, c.part_p + 3 * pp),
        // P5: t2_bar[j,i][b][f] = sum_(a,c) Wb[a,b,c] Vp(k)[c][a][f]
        gemm(V, V, kPartial, {seg(opK(Wb, v, v2, 1, V, 0), opK(Vp(k), 1, v, v2, V, 1), V * V, 1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_p, c.part_p + 4 * pp),
        // P6: t2_bar[i,j][a][f] = sum_(b,c) Wb[a,b,c] Vp(k)[c][b][f]
        gemm(V, V, kPartial, {seg(opK(Wb, v2, v, 1, V, 0), opK(Vp(k), 1, v, v2, V, 1), V * V, 1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_p, c.part_p + 5 * pp),
    };
    PAMD_CHECK(run_gemm_group(c, ds, 6, kProfAdjT2P, true));
  }
  // hole ovoo_bar pieces: split-K over the virtual pair
  const size_t ph = (size_t)c.S_h * v * o;
  {
    const GemmDesc ds[6] = {
        // H1: ovoo_bar[i,a,m,j] = -sum_(b,c) Wb[a,b,c] T2m(k)[c][b][m]
        gemm(V, O, kPartial, {seg(opK(Wb, v2, v, 1, V, 0), opK(T2m(k), 1, o, vo, V, 1), V * V, -1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_h, c.part_h + 0 * ph),
        // H2: ovoo_bar[i,a,m,k] = -sum_(b,c) Wb[a,b,c] T2m(j)[b][c][m]
        gemm(V, O, kPartial, {seg(opK(Wb, v2, v, 1, V, 0), opK(T2m(j), 1, vo, o, V, 1), V * V, -1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_h, c.part_h + 1 * ph),
        // H3: ovoo_bar[j,b,m,i] = -sum_(a,c) Wb[a,b,c] T2m(k)[c][a][m]
        gemm(V, O, kPartial, {seg(opK(Wb, v, v2, 1, V, 0), opK(T2m(k), 1, o, vo, V, 1), V * V, -1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_h, c.part_h + 2 * ph),
        // H4: ovoo_bar[j,b,m,k] = -sum_(a,c) Wb[a,b,c] T2m(i)[a][c][m]
        gemm(V, O, kPartial, {seg(opK(Wb, v, v2, 1, V, 0), opK(T2m(i), 1, vo, o, V, 1), V * V, -1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_h, c.part_h + 3 * ph),
        // H5: ovoo_bar[k,c,m,i] = -sum_(a,b) Wb[a,b,c] T2m(j)[b][a][m]
        gemm(V, O, kPartial, {seg(opK(Wb, 1, v2, v, V, 1), opK(T2m(j), 1, o, vo, V, 1), V * V, -1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_h, c.part_h + 4 * ph),
        // H6: ovoo_bar[k,c,m,j] = -sum_(a,b) Wb[a,b,c] T2m(i)[a][b][m]
        gemm(V, O, kPartial, {seg(opK(Wb, 1, v2, v, V, 1), opK(T2m(i), 1, vo, o, V, 1), V * V, -1.0)},
             out(nullptr, 0, 0, 1, 0, 0, 1), c.S_h, c.part_h + 5 * ph),
    };
    PAMD_CHECK(run_gemm_group(c, ds, 6, kProfAdjOvoo, true));
  }
  // hole-term t2 reads: T2mb[x][p][q][m] (= dE/d t2[x,m,p,q])
  if (pair != 0) {
    const double *Wq = prev ? prev->gW : nullptr;
    GemmDesc ds[3];
    int n = 0;
    if (pair == 2) {
      ds[n++] = gemm(V * V, O, kAccumulate,
                     {seg(opR(Wq, v2, 1, V, v, 1), op1(OOt(j, kp), 1, o, 1), V, -1.0),
                      seg(opR(Wq, v2, v, V, 1, 0), op1(OOt(kp, j), 1, o, 1), V, -1.0),
                      seg(opR(Wb, v2, 1, V, v, 1), op1(OOt(j, k), 1, o, 1), V, -1.0),
                      seg(opR(Wb, v2, v, V, 1, 0), op1(OOt(k, j), 1, o, 1), V, -1.0)},
                     out(T2mb(i), vo, o, V, 0, 1, kBig));
      ds[n++] = gemm(V * V, O, kAccumulate,
                     {seg(opR(Wq, v, 1, V, v2, 1), op1(OOt(i, kp), 1, o, 1), V, -1.0),
                      seg(opR(Wq, v, v2, V, 1, 0), op1(OOt(kp, i), 1, o, 1), V, -1.0),
                      seg(opR(Wb, v, 1, V, v2, 1), op1(OOt(i, k), 1, o, 1), V, -1.0),
                      seg(opR(Wb, v, v2, V, 1, 0), op1(OOt(k, i), 1, o, 1), V, -1.0)},
                     out(T2mb(j), vo, o, V, 0, 1, kBig));
    }
    ds[n++] = gemm(V * V, O, kAccumulate,
                   {seg(opR(Wb, v, 1, V, v2, 1), op1(OOt(i, j), 1, o, 1), V, -1.0),
                    seg(opR(Wb, v2, 1, V, v, 1), op1(OOt(j, i), 1, o, 1), V, -1.0)},
                   out(T2mb(k), o, vo, V, 0, 1, kBig));
    for (int q = 0; q < n; ++q) PAMD_CHECK(run_gemm(c, ds[q], kProfAdjT2H));
  } else {
    const GemmDesc ds[3] = {
        // T2mb[i]: H4 (p=a,q=c,K=b), H6 (p=a,q=b,K=c)
        gemm(V * V, O, kAccumulate,
             {seg(opR(Wb, v2, 1, V, v, 1), op1(OOt(j, k), 1, o, 1), V, -1.0),
              seg(opR(Wb, v2, v, V, 1, 0), op1(OOt(k, j), 1, o, 1), V, -1.0)},
             out(T2mb(i), vo, o, V, 0, 1, kBig)),
        // T2mb[j]: H2 (p=b,q=c,K=a), H5 (p=b,q=a,K=c)
        gemm(V * V, O, kAccumulate,
             {seg(opR(Wb, v, 1, V, v2, 1), op1(OOt(i, k), 1, o, 1), V, -1.0),
              seg(opR(Wb, v, v2, V, 1, 0), op1(OOt(k, i), 1, o, 1), V, -1.0)},
             out(T2mb(j), vo, o, V, 0, 1, kBig)),
        // T2mb[k]: rows m = q*v + p with p = c: H1 (q=b,K=a), H3 (q=a,K=b)
        gemm(V * V, O, kAccumulate,
             {seg(opR(Wb, v, 1, V, v2, 1), op1(OOt(i, j), 1, o, 1), V, -1.0),
              seg(opR(Wb, v2, 1, V, v, 1), op1(OOt(j, i), 1, o, 1), V, -1.0)},
             out(T2mb(k), o, vo, V, 0, 1, kBig)),
    };
    PAMD_CHECK(run_gemm_group(c, ds, 3, kProfAdjT2H, distinct));
  }
  return PAMD_OK;
}

// with_energy: the same launch also sums the orbit epilogue's per-workgroup
// energies of this triple into tri (the walk then skips launch_triple_energy)
static int adjoint_combine(const Ctx &c, int i, int j, int k, hipStream_t st,
                           bool with_energy = false) {
  const double v = c.v, o = c.o, v2 = v * v;
  CombineArgs ca;
  ca.R = c.R;
  ca.part_p = c.part_p;
  ca.part_h = c.part_h;
  ca.S_p = c.S_p;
  ca.S_h = c.S_h;
  ca.o = c.o;
  ca.v = c.v;
  ca.i = i;
  ca.j = j;
  ca.k = k;
  ca.has_fvo = c.has_fvo;
  ca.t2_bar = c.t2_bar;
  ca.ovov_bar = c.ovov_bar;
  ca.ovoo_bar = c.ovoo_bar;
  ca.t1_bar = c.t1_bar;
  ca.fvo_bar = c.fvo_bar;
  ca.part_t = with_energy ? c.part_t : nullptr;
  ca.G = c.G;
  ca.tri_dst = with_energy ? c.tri + ((size_t)i * c.o + j) * c.o + k : nullptr;
  if (!g_record) {
    const double cb = 8.0 * (6.0 * c.S_p * v2 + 6.0 * c.S_h * v * o + 12.0 * v2);
    PAMD_CHECK(timed(st, kProfCombine, 0.0, cb, [&] { return launch_combine(ca, st); }));
  }
  return PAMD_OK;
}

static int adjoint(const Ctx &c, int i, int j, int k, bool with_energy = false) {
  int st = adjoint_vr(c, i, j, k, c.s);
  if (st == PAMD_OK) st = adjoint_gemms(c, i, j, k);
  if (st == PAMD_OK) st = adjoint_combine(c, i, j, k, c.s, with_energy);
  return st;
}

// operand format of a precision
static int split_format(int prec) {
  return prec == PAMD_PREC_BF16X3 ? kSplitBf16x3
                                  : (prec == PAMD_PREC_FP32X2 ? kSplitF32x2 : kSplitNone);
}

// where the orbit epilogue writes Wbar: the GEMM operand copy in the engine's
// format (split forms), or over W (FP64)
static void wbar_out(const Ctx &c, int fmt, OrbitArgs &oa) {
  oa.Wb3 = fmt == kSplitBf16x3 ? (unsigned short *)c.gW : nullptr;
  oa.Wb8 = fmt == kSplitF32x2 ? (unsigned long long *)c.gW : nullptr;
}

// Zero the adjoint accumulators and the cotangent outputs of one call.
static int zero_vjp(const Ctx &c) {
  const int o = c.o, v = c.v;
  const size_t v2 = (size_t)v * v;
  PAMD_CHECK(launch_fill(c.Vpb, (size_t)o * v2 * v, 0.0, c.s));
  PAMD_CHECK(launch_fill(c.T2mb, (size_t)o * o * v2, 0.0, c.s));
  PAMD_CHECK(launch_fill(c.part_mo, (size_t)kGmax * (o + v), 0.0, c.s));
  PAMD_CHECK(launch_fill(c.t1_bar, (size_t)o * v, 0.0, c.s));
  PAMD_CHECK(launch_fill(c.t2_bar, (size_t)o * o * v2, 0.0, c.s));
  PAMD_CHECK(launch_fill(c.ovoo_bar, (size_t)o * v * o * o, 0.0, c.s));
  PAMD_CHECK(launch_fill(c.ovov_bar, (size_t)o * v * o * v, 0.0, c.s));
  PAMD_CHECK(launch_fill(c.fvo_bar, (size_t)v * o, 0.0, c.s));
  return PAMD_OK;
}

// Layout changes at entry, shared by every walk; returns the orbit count.
static int setup(Ctx &c, long long &norb) {
  const int o = c.o, v = c.v;
  const int fmt = split_format(c.prec);
  PAMD_CHECK(launch_prepare(c.ovvv, c.packed, c.t2, c.ovoo, o, v, c.Vp, c.T2m, c.OOt, fmt, c.s));
  PAMD_CHECK(launch_prep_sym(c.ovov, o, v, 0, c.OVS, c.s));
  PAMD_CHECK(launch_prep_sym(c.t2, o, v, 1, c.T2S, c.s));
  if (fmt == kSplitBf16x3)
    PAMD_CHECK(launch_split_bf16x3(c.t2, (size_t)o * o * v * v, (unsigned short *)c.gT2, c.s));
  if (c.VpT)
    PAMD_CHECK(launch_vp_transposes((const unsigned short *)c.Vp, o, v, (unsigned short *)c.VpT,
                                    (unsigned short *)c.VpU, c.s));
  if (fmt == kSplitF32x2)
    PAMD_CHECK(launch_split_f32x2(c.t2, (size_t)o * o * v * v, (unsigned long long *)c.gT2, c.s));
  PAMD_CHECK(launch_fill(c.tri, (size_t)o * o * o, 0.0, c.s));
  const int nt = (v + 7) / 8;
  norb = (long long)nt * (nt + 1) * (nt + 2) / 6;
  PAMD_CHECK(launch_orbit_table(norb, c.orb, c.s));
  c.G = std::min(orbit_grid(v), kGmax);
  return PAMD_OK;
}

static OrbitArgs orbit_args(const Ctx &c, int i, int j, int k, long long norb) {
  OrbitArgs oa;
  memset(&oa, 0, sizeof(oa));
  oa.W = c.W;
  oa.tab = c.orb;
  oa.Vb = c.Vb;
  oa.t1 = c.t1;
  oa.ovs = c.OVS;
  oa.t2s = c.T2S;
  oa.fvo = c.fvo;
  oa.mo = c.mo;
  oa.o = c.o;
  oa.v = c.v;
  oa.i = i;
  oa.j = j;
  oa.k = k;
  oa.has_fvo = c.has_fvo;
  oa.norb = norb;
  oa.part_t = c.part_t;
  oa.part_mo = c.part_mo;
  return oa;
}

// Projected walk: for every unordered occupied pair (j, k), all o triples
// (l, j, k) in order of l.  Each W_ljk is built, and the orbit epilogue adds
// ulo[f, l] W and ulo[f, l] Z / D into accA[f], accB[f]; then
// pair_e[j, k] = sum_f <accA[f], accB[f]> = sum_{l,i} mat[l,i] <W_ljk, Z_ijk / D_ijk>.
// With c.vjp, a second pass over l rebuilds W_ljk and runs the projected
// adjoint epilogue (seeds from the finished accA, accB, cotangent
// g = ebar w_jk / 3 with w_jk = 2 off the diagonal) and the plain adjoint
// contractions of that triple; dE/dulo[f, l] gathers g (<W, accB[f]> +
// <accA[f], Z/D>) into ulo_bar.
static int run_projected(Ctx &c, const double *ulo, int nfrag, double *accA, double *accB,
                         double *pair_e, double *dot_part, double *part_u, double *ulo_bar) {
  long long norb = 0;
  int st = setup(c, norb);
  if (st != PAMD_OK) return st;
  const int o = c.o;
  const size_t v3 = (size_t)c.v * c.v * c.v;
  if (c.vjp) {
    st = zero_vjp(c);
    if (st != PAMD_OK) return st;
    PAMD_CHECK(launch_fill(ulo_bar, (size_t)nfrag * o, 0.0, c.s));
  }
  long long tcount = 0;
  const long long nbuild = (long long)o * o * (o + 1) / 2 * (c.vjp ? 2 : 1);
  const auto t_start = std::chrono::steady_clock::now();
  auto progress = [&]() -> int {
    if (g_progress_every > 0 && tcount % g_progress_every == 0) {
      PAMD_CHECK(hipStreamSynchronize(c.s));
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                       t_start).count();
      fprintf(stderr, "pyscfad_amd: projected (T)%s triples %lld/%lld done, %.1f s\n",
              c.vjp ? "+VJP" : "", tcount, nbuild, el);
      fflush(stderr);
    }
    return PAMD_OK;
  };
  const int fmt = split_format(c.prec);
  for (int j = 0; j < o; ++j)
    for (int k = j; k < o; ++k) {
      PAMD_CHECK(launch_fill(accA, (size_t)nfrag * v3, 0.0, c.s));
      PAMD_CHECK(launch_fill(accB, (size_t)nfrag * v3, 0.0, c.s));
      for (int l = 0; l < o; ++l) {
        g_prof_on = g_prof_every > 0 && (tcount % g_prof_every) == 0;
        ++tcount;
        st = build_w(c, l, j, k);
        if (st != PAMD_OK) return st;
        OrbitArgs oa = orbit_args(c, l, j, k, norb);
        oa.accA = accA;
        oa.accB = accB;
        oa.nfrag = nfrag;
        for (int f = 0; f < nfrag; ++f) oa.u[f] = ulo[(size_t)f * o + l];
        const double ob = 8.0 * (double)v3 * (1.0 + 4.0 * nfrag);  // W read, 2 nfrag RMW
        PAMD_CHECK(timed(c.s, kProfOrbit, 0.0, ob,
                         [&] { return launch_orbit(oa, c.G, c.s, kOrbitProjAcc); }));
        st = progress();
        if (st != PAMD_OK) return st;
      }
      PAMD_CHECK(launch_dot(accA, accB, (size_t)nfrag * v3, dot_part, kGmax, pair_e + j * o + k,
                            c.s));
      if (!c.vjp) continue;
      const double g = c.ebar * (j == k ? 1.0 : 2.0) / 3.0;
      for (int l = 0; l < o; ++l) {
        g_prof_on = g_prof_every > 0 && (tcount % g_prof_every) == 0;
        ++tcount;
        st = build_w(c, l, j, k);
        if (st != PAMD_OK) return st;
        OrbitArgs oa = orbit_args(c, l, j, k, norb);
        wbar_out(c, fmt, oa);
        oa.vjp = 1;
        oa.ebar = g;
        oa.accA = accA;
        oa.accB = accB;
        oa.nfrag = nfrag;
        oa.part_u = part_u;
        for (int f = 0; f < nfrag; ++f) oa.u[f] = ulo[(size_t)f * o + l];
        const double ob = 8.0 * (double)v3 * (3.0 + 2.0 * nfrag);  // W, accA, accB read; Wbar, Vbar
        PAMD_CHECK(timed(c.s, kProfOrbit, 0.0, ob,
                         [&] { return launch_orbit(oa, c.G, c.s, kOrbitProjVjp); }));
        PAMD_CHECK(launch_reduce_ulo(part_u, c.G, nfrag, g, ulo_bar + l, o, c.s));
        st = adjoint(c, l, j, k);
        if (st != PAMD_OK) return st;
        st = progress();
        if (st != PAMD_OK) return st;
      }
    }
  g_prof_on = false;
  if (!g_prof_recs.empty()) PAMD_CHECK(prof_collect());
  return PAMD_OK;
}

// Side stream and events of the pipelined adjoint walk, per host thread and
// device: created once (default priority, non-blocking) and kept for the life
// of the process.
struct PipeRes {
  hipStream_t e;
  hipEvent_t ev[10];  // start, end, W built [2], V-bar reduced [2], adjoint done [2], gathered [2]
  bool ok;
};

static PipeRes *pipe_resources() {
  constexpr int kMaxDev = 64;
  static thread_local PipeRes res[kMaxDev];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
  PipeRes &r = res[dev];
  if (!r.ok) {
    if (hipStreamCreateWithFlags(&r.e, hipStreamNonBlocking) != hipSuccess) return nullptr;
    for (hipEvent_t &e : r.ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    r.ok = true;
  }
  return &r;
}

// The adjoint walk with two triples in flight.  The contractions (W builds
// and adjoint GEMMs) run in order on the call's stream G; the memory-bound
// epilogue of each triple (orbit epilogue, triple energy, V-bar reductions,
// gather into the cotangents) runs on the side stream E while G builds the
// next triple's W:
//   G: WB(0) WB(1) ADJ(0) WB(2) ADJ(1) WB(3) ADJ(2) ...
//   E:   ORB(0) VR(0) | COMB(0) ORB(1) VR(1) | COMB(1) ...
// Triple t uses buffer set t % 2 (W, Wbar, V-bar scratch R, split-K
// partials).  Waits: ORB(t) after WB(t); ADJ(t) after VR(t) and after COMB(t-2)
// (which read the partials ADJ(t) overwrites); COMB(t) after ADJ(t).  The
// orders within G and E cover the rest: WB(t+2) follows ADJ(t), which followed
// ORB(t) (the last reader of W[t % 2]); ORB(t+2) follows COMB(t+1) and so
// ADJ(t) (the last reader of Wbar[t % 2]); VR(t+2) follows COMB(t) (the last
// reader of R[t % 2]).  The accumulators shared by all triples (Vpb and T2mb
// on G; the cotangents, triple energies and dE/de partials on E) are each
// written by one stream, in walk order, so the sums are the sequential walk's.
static int run_vjp_pipelined(Ctx &c, long long norb, PipeRes &pr) {
  const int o = c.o, fmt = split_format(c.prec);
  std::vector<int> tr;  // triples of this call in walk order, i | j << 10 | k << 20
  {
    long long windex = -1;
    for (int i = 0; i < o; ++i)
      for (int j = i; j < o; ++j)
        for (int k = j; k < o; ++k) {
          if (i == j && j == k) continue;
          ++windex;
          if (c.nshards > 1 && windex % c.nshards != c.shard) continue;
          tr.push_back(i | (j << 10) | (k << 20));
        }
  }
  Ctx cx[2] = {c, c};
  cx[1].W = c.W2;
  cx[1].gW = c.gW2;
  cx[1].R = c.R2;
  cx[1].part_p = c.part_p2;
  cx[1].part_h = c.part_h2;
  for (int a = 0; a < cx[1].planes.count; ++a)
    if (cx[1].planes.base[a] == (const char *)c.gW) cx[1].planes.base[a] = (const char *)c.gW2;
  // Paired ovvv adjoints (env PYSCFAD_AMD_PAIR_OVVV=0 turns them off): consecutive triples
  // (i, j, k), (i, j, k') add into the same Vpb[i] and Vpb[j], so the second
  // one's launches for them also take the first one's terms, from the other
  // Wbar buffer.  That buffer is next written by the orbit epilogue two
  // triples on, which the side stream runs after the gather of the second
  // triple, which waits for its contractions.  Only where Wbar has its own
  // buffer (split forms; FP64 writes Wbar over W, which the W build of the
  // triple after next overwrites before the second triple's contractions).
  // The hole-term t2 adjoints into T2mb[i], T2mb[j] pair the same way.
  // Measured (full walks, split-BF16x3, ovvv adjoints paired): o16/v160 E+VJP
  // 0.541 -> 0.532 s, o24/v240 5.39 -> 5.30 s, o40/v400 127.8 -> 125.2 s.
  // Only ovvv_bar and t2_bar change (the same terms grouped in other FP32
  // sums, ~1e-8 relative).
  static const bool pair_env = [] {
    const char *e = getenv("PYSCFAD_AMD_PAIR_OVVV");
    return !(e && e[0] == '0');
  }();
  const bool pair_on = pair_env && c.gW != c.W && c.gW2 != nullptr;
  if (pair_on && c.prec == PAMD_PREC_BF16X3) {
    const long long v3 = (long long)c.v * c.v * c.v;
    for (int p = 0; p < 2; ++p) {
      PlanarMap &pm = cx[p].planes;
      if (pm.count >= 8) return PAMD_ERR_ARG;
      pm.base[pm.count] = (const char *)cx[1 - p].gW;
      pm.n[pm.count] = v3;
      ++pm.count;
    }
  }
  const hipStream_t G = c.s, E = pr.e;
  hipEvent_t *wb = pr.ev + 2, *vr = pr.ev + 4, *adj = pr.ev + 6, *cmb = pr.ev + 8;
  PAMD_CHECK(hipEventRecord(pr.ev[0], G));
  PAMD_CHECK(hipStreamWaitEvent(E, pr.ev[0], 0));
  const long long n = (long long)tr.size();
  const long long ntrip = (long long)o * (o + 1) * (o + 2) / 6 - o;
  const auto t_start = std::chrono::steady_clock::now();
  auto walk = [&]() -> int {
    if (n == 0) return PAMD_OK;
    auto ijk = [&](long long t, int &i, int &j, int &k) {
      i = tr[t] & 1023;
      j = (tr[t] >> 10) & 1023;
      k = tr[t] >> 20;
    };
    int i, j, k, st;
    bool pend = false;  // the previous triple left its Vpb[i], Vpb[j] adjoints to this one
    int kpend = 0;
    g_prof_on = g_prof_every > 0;
    ijk(0, i, j, k);
    st = build_w(cx[0], i, j, k);
    if (st != PAMD_OK) return st;
    PAMD_CHECK(hipEventRecord(wb[0], G));
    for (long long t = 0; t < n; ++t) {
      const int p = (int)(t & 1);
      const Ctx &C = cx[p];
      g_prof_on = g_prof_every > 0 && (t % g_prof_every) == 0;
      ijk(t, i, j, k);
      // E: epilogue and V-bar reductions of t
      PAMD_CHECK(hipStreamWaitEvent(E, wb[p], 0));
      OrbitArgs oa = orbit_args(C, i, j, k, norb);
      wbar_out(C, fmt, oa);
      oa.vjp = 1;
      oa.omega = (i == j || j == k) ? 1.0 : 2.0;
      oa.ebar = c.wt_host ? c.ebar * ((c.wt_host[i] + c.wt_host[j] + c.wt_host[k]) / 3.0) : c.ebar;
      {
        const double ob = 8.0 * (double)c.v * c.v * c.v * 3.0;  // W read, Wbar and Vbar written
        PAMD_CHECK(timed(E, kProfOrbit, 0.0, ob, [&] { return launch_orbit(oa, C.G, E); }));
      }
      st = adjoint_vr(C, i, j, k, E);
      if (st != PAMD_OK) return st;
      PAMD_CHECK(hipEventRecord(vr[p], E));
      // G: W build of t + 1
      if (t + 1 < n) {
        int i1, j1, k1;
        ijk(t + 1, i1, j1, k1);
        st = build_w(cx[1 - p], i1, j1, k1);
        if (st != PAMD_OK) return st;
        PAMD_CHECK(hipEventRecord(wb[1 - p], G));
      }
      // G: adjoint contractions of t
      PAMD_CHECK(hipStreamWaitEvent(G, vr[p], 0));
      if (t >= 2) PAMD_CHECK(hipStreamWaitEvent(G, cmb[p], 0));
      if (pend) {
        st = adjoint_gemms(C, i, j, k, 2, &cx[1 - p], kpend);
        pend = false;
      } else if (pair_on && t + 1 < n && (tr[t + 1] & 0xfffff) == (tr[t] & 0xfffff)) {
        st = adjoint_gemms(C, i, j, k, 1);
        pend = true;
        kpend = k;
      } else {
        st = adjoint_gemms(C, i, j, k);
      }
      if (st != PAMD_OK) return st;
      PAMD_CHECK(hipEventRecord(adj[p], G));
      // E: gather of t
      PAMD_CHECK(hipStreamWaitEvent(E, adj[p], 0));
      st = adjoint_combine(C, i, j, k, E, true);  // with the triple's energy
      if (st != PAMD_OK) return st;
      PAMD_CHECK(hipEventRecord(cmb[p], E));
      if (g_progress_every > 0 && (t + 1) % g_progress_every == 0) {
        PAMD_CHECK(hipStreamSynchronize(E));
        PAMD_CHECK(hipStreamSynchronize(G));
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                         t_start).count();
        fprintf(stderr, "pyscfad_amd: (T)+VJP triples %lld/%lld done, %.1f s\n", t + 1,
                c.nshards > 1 ? n : ntrip, el);
        fflush(stderr);
      }
    }
    return PAMD_OK;
  };
  int st = walk();
  // the call's stream waits for the side stream, also after an error
  hipError_t e = hipEventRecord(pr.ev[1], E);
  if (e == hipSuccess) e = hipStreamWaitEvent(G, pr.ev[1], 0);
  if (e != hipSuccess && st == PAMD_OK) {
    fprintf(stderr, "pyscfad_amd: joining the side stream failed: %s\n", hipGetErrorString(e));
    st = PAMD_ERR_HIP;
  }
  g_prof_on = false;
  if (st == PAMD_OK && !g_prof_recs.empty()) PAMD_CHECK(prof_collect());
  return st;
}

static int run(Ctx &c) {
  const int o = c.o, v = c.v;
  const int fmt = split_format(c.prec);
  long long norb = 0;
  {
    const int st0 = setup(c, norb);
    if (st0 != PAMD_OK) return st0;
  }
  if (c.vjp) {
    const int st0 = zero_vjp(c);
    if (st0 != PAMD_OK) return st0;
  }
  static const bool pipe_on = [] {
    const char *e = getenv("PYSCFAD_AMD_PIPELINE");
    return !(e && e[0] == '0');
  }();
  if (c.vjp && c.pipe && pipe_on && !g_record) {
    PipeRes *pr = pipe_resources();
    if (pr) return run_vjp_pipelined(c, norb, *pr);
  }
  long long tcount = 0, windex = -1;
  const long long ntrip = (long long)o * (o + 1) * (o + 2) / 6 - o;
  const auto t_start = std::chrono::steady_clock::now();
  if (!c.vjp && c.nshards == 1 && c.wbatch > 1) {
    // energy only, small nvir: the W blocks of up to wbatch consecutive k of one
    // pair (i, j) are built by one launch per pass; each triple's epilogue then
    // reads its own block.  Same triples, same per-triple arithmetic and the
    // same fixed-order sum over tri as the walk below.
    const size_t v3 = (size_t)v * v * v;
    for (int i = 0; i < o; ++i)
      for (int j = i; j < o; ++j)
        for (int kb = (i == j ? j + 1 : j); kb < o; kb += c.wbatch) {
          const int nb = std::min(c.wbatch, o - kb);
          g_prof_on = g_prof_every > 0 && (tcount % g_prof_every) < nb;
          int st = build_w(c, i, j, kb, nb);
          if (st != PAMD_OK) return st;
          for (int b = 0; b < nb; ++b) {
            const int k = kb + b;
            ++tcount;
            OrbitArgs oa = orbit_args(c, i, j, k, norb);
            oa.W = c.W + b * v3;
            oa.omega = (i == j || j == k) ? 1.0 : 2.0;
            oa.ebar = c.ebar;
            const double ob = 8.0 * (double)v3;  // W read
            PAMD_CHECK(timed(c.s, kProfOrbit, 0.0, ob, [&] { return launch_orbit(oa, c.G, c.s); }));
            PAMD_CHECK(launch_triple_energy(c.part_t, c.G, c.tri + ((size_t)i * o + j) * o + k, c.s));
            if (g_progress_every > 0 && tcount % g_progress_every == 0) {
              PAMD_CHECK(hipStreamSynchronize(c.s));
              const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                               t_start).count();
              fprintf(stderr, "pyscfad_amd: (T) triples %lld/%lld done, %.1f s\n", tcount, ntrip,
                      el);
              fflush(stderr);
            }
          }
        }
    g_prof_on = false;
    if (!g_prof_recs.empty()) PAMD_CHECK(prof_collect());
    return PAMD_OK;
  }
  for (int i = 0; i < o; ++i)
    for (int j = i; j < o; ++j)
      for (int k = j; k < o; ++k) {
        if (i == j && j == k) continue;
        // round-robin shard of the walk: triple number windex goes to shard windex % nshards
        ++windex;
        if (c.nshards > 1 && windex % c.nshards != c.shard) continue;
        g_prof_on = g_prof_every > 0 && (tcount % g_prof_every) == 0;
        ++tcount;
        int st = build_w(c, i, j, k);
        if (st != PAMD_OK) return st;
        OrbitArgs oa = orbit_args(c, i, j, k, norb);
        if (c.vjp) wbar_out(c, fmt, oa);
        oa.vjp = c.vjp;
        oa.omega = (i == j || j == k) ? 1.0 : 2.0;
        // a diagonal occupied projector diag(m) weighs the whole triple by
        // (m_i + m_j + m_k) / 3 (the triple energy is symmetric in i, j, k)
        oa.ebar = c.wt_host ? c.ebar * ((c.wt_host[i] + c.wt_host[j] + c.wt_host[k]) / 3.0) : c.ebar;
        {
          const double v3 = (double)v * v * v;
          const double ob = 8.0 * v3 * (c.vjp ? 3.0 : 1.0);  // W read (+ Wbar, Vbar written)
          PAMD_CHECK(timed(c.s, kProfOrbit, 0.0, ob, [&] { return launch_orbit(oa, c.G, c.s); }));
        }
        if (!c.vjp)
          PAMD_CHECK(launch_triple_energy(c.part_t, c.G, c.tri + ((size_t)i * o + j) * o + k, c.s));
        if (c.vjp) {
          st = adjoint(c, i, j, k, true);  // the gather also sums the triple's energy
          if (st != PAMD_OK) return st;
        }
        if (g_progress_every > 0 && tcount % g_progress_every == 0) {
          PAMD_CHECK(hipStreamSynchronize(c.s));
          const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                           t_start).count();
          fprintf(stderr, "pyscfad_amd: (T)%s triples %lld/%lld done, %.1f s\n",
                  c.vjp ? "+VJP" : "", tcount, ntrip, el);
          fflush(stderr);
        }
      }
  g_prof_on = false;
  if (!g_prof_recs.empty()) PAMD_CHECK(prof_collect());
  return PAMD_OK;
}

static int check_common(const pamd_ccsdt_config *cfg, const double *mo, const double *t1,
                        const double *t2, const double *ovvv, const double *ovoo,
                        const double *ovov, const double *fvo) {
  if (!cfg || cfg->nocc < 1 || cfg->nvir < 1) return PAMD_ERR_ARG;
  if (!mo || !t1 || !t2 || !ovvv || !ovoo || !ovov) return PAMD_ERR_ARG;
  if (cfg->has_fvo && !fvo) return PAMD_ERR_ARG;
  if (cfg->precision != PAMD_PREC_FP64 && cfg->precision != PAMD_PREC_BF16X3 &&
      cfg->precision != PAMD_PREC_FP32X2)
    return PAMD_ERR_UNSUPPORTED;
  if (cfg->nocc > 1024) return PAMD_ERR_ARG;
  // the engines address an operand within one occupied slab with 32-bit
  // element offsets, and the largest slab is v^3 (W, Wbar, Vp[x])
  if ((long long)cfg->nvir * cfg->nvir * cfg->nvir >= (1LL << 31)) return PAMD_ERR_UNSUPPORTED;
  return PAMD_OK;
}

static void bind_workspace(Ctx &c, const Layout &L, void *ws) {
  char *b = (char *)ws;
  c.Vp = (double *)(b + L.off_Vp);
  c.T2m = (double *)(b + L.off_T2m);
  c.OOt = (double *)(b + L.off_OOt);
  c.OVS = (double *)(b + L.off_OVS);
  c.T2S = (double *)(b + L.off_T2S);
  c.orb = (unsigned short *)(b + L.off_orb);
  c.W = (double *)(b + L.off_W);
  c.wbatch = L.wbatch;
  c.tri = (double *)(b + L.off_tri);
  c.part_t = (double *)(b + L.off_part_t);
  c.wt = (double *)(b + L.off_wt);
  const bool packed_ops = c.prec == PAMD_PREC_BF16X3 || c.prec == PAMD_PREC_FP32X2;
  c.gT2 = packed_ops ? (const double *)(b + L.off_T2g) : c.t2;
  c.gW = (packed_ops && c.vjp) ? (double *)(b + L.off_W8) : c.W;
  c.VpT = L.vpt ? (double *)(b + L.off_VpT) : nullptr;
  c.VpU = L.vpt ? (double *)(b + L.off_VpU) : nullptr;
  {
    const long long o = c.o, v = c.v, v3 = v * v * v;
    PlanarMap &pm = c.planes;
    pm.count = 0;
    auto add = [&](const void *p, long long n) {
      pm.base[pm.count] = (const char *)p;
      pm.n[pm.count] = n;
      ++pm.count;
    };
    add(c.Vp, o * v3);
    add(c.T2m, o * o * v * v);
    add(c.OOt, o * o * o * v);
    add(c.gT2, o * o * v * v);
    if (c.vjp) add(c.gW, v3);
    if (c.VpT) {
      add(c.VpT, o * v3);
      add(c.VpU, o * v3);
    }
  }
  if (c.vjp) {
    c.Vb = (double *)(b + L.off_Vb);
    c.Vpb = (double *)(b + L.off_Vpb);
    c.T2mb = (double *)(b + L.off_T2mb);
    c.part_mo = (double *)(b + L.off_part_mo);
    c.R = (double *)(b + L.off_R);
    c.part_p = (double *)(b + L.off_part_p);
    c.part_h = (double *)(b + L.off_part_h);
    c.pipe = L.pipe;
    if (L.pipe) {
      c.W2 = (double *)(b + L.off_W2);
      c.gW2 = packed_ops ? (double *)(b + L.off_W8b) : c.W2;
      c.R2 = (double *)(b + L.off_R2);
      c.part_p2 = (double *)(b + L.off_part_p2);
      c.part_h2 = (double *)(b + L.off_part_h2);
    }
  } else {
    c.Vb = c.Vpb = c.T2mb = c.part_mo = c.R = c.part_p = c.part_h = nullptr;
  }
  c.S_p = L.S_p;
  c.S_h = L.S_h;
}

// ---------------------------------------------------------------------------
// Library-owned scratch for pamd_ccsdt_energy_vjp_noalloc: the workspace of
// pamd_ccsdt_energy_vjp in a file-scope device array, so that call allocates
// nothing.  Each GPU holds its own copy for the life of the process; its
// address on a device is looked up once (hipGetSymbolAddress, with that device
// current) and kept per device.  The host-side name `scratch` is never passed
// to a kernel or a copy: only the looked-up address is.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[kMaxDevices];

// Address of this device's scratch array (the device current on the calling
// thread), or a PAMD_* error.
static int scratch_on_current_device(void **ptr) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    fprintf(stderr, "pyscfad_amd: hipGetDevice failed: %s\n", hipGetErrorString(e));
    return PAMD_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return PAMD_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "pyscfad_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
              hipGetErrorString(e));
      return PAMD_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return PAMD_OK;
}

}  // namespace pamd

using namespace pamd;

extern "C" {

const char *pamd_version(void) { return "pyscfad_amd 0.1.0 (gfx950)"; }

int pamd_ccsdt_workspace_size(const pamd_ccsdt_config *cfg, int with_vjp, size_t *bytes) {
  if (!cfg || !bytes || cfg->nocc < 1 || cfg->nvir < 1) return PAMD_ERR_ARG;
  *bytes = make_layout(cfg->nocc, cfg->nvir, with_vjp ? 1 : 0, cfg->precision, true).total;
  return PAMD_OK;
}

static int energy_impl(const pamd_ccsdt_config *cfg, const double *mo, const double *t1,
                       const double *t2, const double *ovvv, const double *ovoo,
                       const double *ovov, const double *fvo, void *ws, size_t ws_bytes,
                       hipStream_t s, Ctx &c, int shard = 0, int nshards = 1) {
  int st = check_common(cfg, mo, t1, t2, ovvv, ovoo, ovov, fvo);
  if (st != PAMD_OK) return st;
  Layout L = make_layout(cfg->nocc, cfg->nvir, 0, cfg->precision);
  if (!ws || ws_bytes < L.total) return PAMD_ERR_WORKSPACE;
  memset(&c, 0, sizeof(c));
  c.o = cfg->nocc;
  c.v = cfg->nvir;
  c.packed = cfg->ovvv_packed ? 1 : 0;
  c.has_fvo = cfg->has_fvo ? 1 : 0;
  c.vjp = 0;
  c.prec = cfg->precision;
  c.s = s;
  c.mo = mo;
  c.t1 = t1;
  c.t2 = t2;
  c.ovvv = ovvv;
  c.ovoo = ovoo;
  c.ovov = ovov;
  c.fvo = fvo;
  c.shard = shard;
  c.nshards = nshards;
  bind_workspace(c, L, ws);
  return run(c);
}

int pamd_ccsdt_energy(const pamd_ccsdt_config *cfg, const double *mo_energy, const double *t1,
                      const double *t2, const double *ovvv, const double *ovoo, const double *ovov,
                      const double *fvo, void *workspace, size_t workspace_bytes, double *e_out,
                      void *stream) {
  if (!e_out) return PAMD_ERR_ARG;
  Ctx c;
  hipStream_t s = (hipStream_t)stream;
  int st = energy_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, workspace, workspace_bytes,
                       s, c);
  if (st != PAMD_OK) return st;
  PAMD_CHECK(launch_total_energy(c.tri, c.o, e_out, s));
  return PAMD_OK;
}

int pamd_ccsdt_num_triples(int nocc, long long *n) {
  if (nocc < 1 || !n) return PAMD_ERR_ARG;
  *n = (long long)nocc * (nocc + 1) * (nocc + 2) / 6 - nocc;
  return PAMD_OK;
}

int pamd_ccsdt_energy_shard(const pamd_ccsdt_config *cfg, const double *mo_energy,
                            const double *t1, const double *t2, const double *ovvv,
                            const double *ovoo, const double *ovov, const double *fvo, int shard,
                            int nshards, void *workspace, size_t workspace_bytes, double *e_out,
                            void *stream) {
  if (!e_out || nshards < 1 || shard < 0 || shard >= nshards) return PAMD_ERR_ARG;
  Ctx c;
  hipStream_t s = (hipStream_t)stream;
  int st = energy_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, workspace, workspace_bytes,
                       s, c, shard, nshards);
  if (st != PAMD_OK) return st;
  PAMD_CHECK(launch_total_energy(c.tri, c.o, e_out, s));
  return PAMD_OK;
}

int pamd_ccsdt_pair_energy(const pamd_ccsdt_config *cfg, const double *mo_energy, const double *t1,
                           const double *t2, const double *ovvv, const double *ovoo,
                           const double *ovov, const double *fvo, void *workspace,
                           size_t workspace_bytes, double *e_pair, void *stream) {
  if (!e_pair) return PAMD_ERR_ARG;
  Ctx c;
  hipStream_t s = (hipStream_t)stream;
  int st = energy_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, workspace, workspace_bytes,
                       s, c);
  if (st != PAMD_OK) return st;
  PAMD_CHECK(launch_pair_energy(c.tri, c.o, e_pair, s));
  return PAMD_OK;
}

static int vjp_impl(const pamd_ccsdt_config *cfg, const double *mo_energy, const double *t1,
                    const double *t2, const double *ovvv, const double *ovoo, const double *ovov,
                    const double *fvo, const double *wt_host, double et_bar, void *workspace,
                    size_t workspace_bytes, double *e_out, double *mo_energy_bar, double *t1_bar,
                    double *t2_bar, double *ovvv_bar, double *ovoo_bar, double *ovov_bar,
                    double *fvo_bar, double *wt_bar, void *stream, int shard = 0,
                    int nshards = 1, bool pipe = true) {
  int st = check_common(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo);
  if (st != PAMD_OK) return st;
  if (!e_out || !mo_energy_bar || !t1_bar || !t2_bar || !ovvv_bar || !ovoo_bar || !ovov_bar ||
      !fvo_bar)
    return PAMD_ERR_ARG;
  Layout L = make_layout(cfg->nocc, cfg->nvir, 1, cfg->precision, pipe);
  if (!workspace || workspace_bytes < L.total) return PAMD_ERR_WORKSPACE;
  Ctx c;
  memset(&c, 0, sizeof(c));
  c.o = cfg->nocc;
  c.v = cfg->nvir;
  c.packed = cfg->ovvv_packed ? 1 : 0;
  c.has_fvo = cfg->has_fvo ? 1 : 0;
  c.vjp = 1;
  c.prec = cfg->precision;
  c.s = (hipStream_t)stream;
  c.mo = mo_energy;
  c.t1 = t1;
  c.t2 = t2;
  c.ovvv = ovvv;
  c.ovoo = ovoo;
  c.ovov = ovov;
  c.fvo = fvo;
  c.ebar = et_bar;
  c.wt_host = wt_host;
  c.t1_bar = t1_bar;
  c.t2_bar = t2_bar;
  c.ovoo_bar = ovoo_bar;
  c.ovov_bar = ovov_bar;
  c.fvo_bar = fvo_bar;
  c.shard = shard;
  c.nshards = nshards;
  bind_workspace(c, L, workspace);
  st = run(c);
  if (st != PAMD_OK) return st;
  const int o = c.o, v = c.v;
  if (wt_host) {
    PAMD_CHECK(hipMemcpyAsync(c.wt, wt_host, sizeof(double) * (size_t)o, hipMemcpyHostToDevice,
                              c.s));
    PAMD_CHECK(launch_weighted_energy(c.tri, o, c.wt, et_bar, e_out, wt_bar, c.s));
  } else {
    PAMD_CHECK(launch_total_energy(c.tri, o, e_out, c.s));
  }
  PAMD_CHECK(launch_reduce_rows(c.part_mo, c.G, o + v, mo_energy_bar, 0, c.s));
  PAMD_CHECK(launch_fold_ovvv(c.Vpb, o, v, c.packed, ovvv_bar, c.s));
  PAMD_CHECK(launch_fold_t2m(c.T2mb, o, v, t2_bar, c.s));
  return PAMD_OK;
}

int pamd_ccsdt_energy_vjp(const pamd_ccsdt_config *cfg, const double *mo_energy, const double *t1,
                          const double *t2, const double *ovvv, const double *ovoo,
                          const double *ovov, const double *fvo, double et_bar, void *workspace,
                          size_t workspace_bytes, double *e_out, double *mo_energy_bar,
                          double *t1_bar, double *t2_bar, double *ovvv_bar, double *ovoo_bar,
                          double *ovov_bar, double *fvo_bar, void *stream) {
  return vjp_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, nullptr, et_bar, workspace,
                  workspace_bytes, e_out, mo_energy_bar, t1_bar, t2_bar, ovvv_bar, ovoo_bar,
                  ovov_bar, fvo_bar, nullptr, stream);
}

int pamd_ccsdt_energy_vjp_shard(const pamd_ccsdt_config *cfg, const double *mo_energy,
                                const double *t1, const double *t2, const double *ovvv,
                                const double *ovoo, const double *ovov, const double *fvo,
                                double et_bar, int shard, int nshards, void *workspace,
                                size_t workspace_bytes, double *e_out, double *mo_energy_bar,
                                double *t1_bar, double *t2_bar, double *ovvv_bar,
                                double *ovoo_bar, double *ovov_bar, double *fvo_bar,
                                void *stream) {
  if (nshards < 1 || shard < 0 || shard >= nshards) return PAMD_ERR_ARG;
  return vjp_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, nullptr, et_bar, workspace,
                  workspace_bytes, e_out, mo_energy_bar, t1_bar, t2_bar, ovvv_bar, ovoo_bar,
                  ovov_bar, fvo_bar, nullptr, stream, shard, nshards);
}

size_t pamd_ccsdt_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int pamd_ccsdt_energy_vjp_noalloc(const pamd_ccsdt_config *cfg, const double *mo_energy,
                                  const double *t1, const double *t2, const double *ovvv,
                                  const double *ovoo, const double *ovov, const double *fvo,
                                  double et_bar, double *e_out, double *mo_energy_bar,
                                  double *t1_bar, double *t2_bar, double *ovvv_bar,
                                  double *ovoo_bar, double *ovov_bar, double *fvo_bar,
                                  void *stream) {
  int st = check_common(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo);
  if (st != PAMD_OK) return st;
  const Layout L = make_layout(cfg->nocc, cfg->nvir, 1, cfg->precision);
  if (L.total > kNoAllocScratchBytes) {
    fprintf(stderr,
            "pyscfad_amd: pamd_ccsdt_energy_vjp_noalloc: nocc=%d nvir=%d needs %zu bytes of "
            "scratch, more than the library's %zu; use pamd_ccsdt_energy_vjp with a workspace\n",
            cfg->nocc, cfg->nvir, L.total, kNoAllocScratchBytes);
    return PAMD_ERR_WORKSPACE;
  }
  void *ws = nullptr;
  st = scratch_on_current_device(&ws);
  if (st != PAMD_OK) return st;
  g_prof_blocked = true;
  // one triple in flight: no second stream, events or buffer copies
  st = vjp_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, nullptr, et_bar, ws,
                kNoAllocScratchBytes, e_out, mo_energy_bar, t1_bar, t2_bar, ovvv_bar, ovoo_bar,
                ovov_bar, fvo_bar, nullptr, stream, 0, 1, false);
  g_prof_blocked = false;
  return st;
}

int pamd_ccsdt_weighted_energy_vjp(const pamd_ccsdt_config *cfg, const double *mo_energy,
                                   const double *t1, const double *t2, const double *ovvv,
                                   const double *ovoo, const double *ovov, const double *fvo,
                                   const double *occ_weight, double et_bar, void *workspace,
                                   size_t workspace_bytes, double *e_out, double *mo_energy_bar,
                                   double *t1_bar, double *t2_bar, double *ovvv_bar,
                                   double *ovoo_bar, double *ovov_bar, double *fvo_bar,
                                   double *occ_weight_bar, void *stream) {
  if (!occ_weight || !occ_weight_bar) return PAMD_ERR_ARG;
  return vjp_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, occ_weight, et_bar, workspace,
                  workspace_bytes, e_out, mo_energy_bar, t1_bar, t2_bar, ovvv_bar, ovoo_bar,
                  ovov_bar, fvo_bar, occ_weight_bar, stream);
}

// workspace of the projected walk: the plain layout (forward or adjoint), then
// accA, accB (nfrag v^3 each), the pair energies (o^2), the dot partials
// (kGmax) and the per-workgroup dE/dulo sums (kGmax x kMaxFrag)
struct ProjLayout {
  Layout L;
  size_t off_a, off_b, off_p, off_d, off_u, total;
};

static ProjLayout projected_layout(int o, int v, int nfrag, int vjp, int prec) {
  ProjLayout P;
  P.L = make_layout(o, v, vjp, prec);
  const size_t blk = align256(sizeof(double) * (size_t)nfrag * v * v * v);
  P.off_a = P.L.total;
  P.off_b = P.off_a + blk;
  P.off_p = P.off_b + blk;
  P.off_d = P.off_p + align256(sizeof(double) * (size_t)o * o);
  P.off_u = P.off_d + align256(sizeof(double) * kGmax);
  P.total = P.off_u + align256(sizeof(double) * kGmax * kMaxFrag);
  return P;
}

int pamd_ccsdt_projected_workspace_size(const pamd_ccsdt_config *cfg, int nfrag, int with_vjp,
                                        size_t *bytes) {
  if (!cfg || !bytes || cfg->nocc < 1 || cfg->nvir < 1 || nfrag < 1 || nfrag > kMaxFrag)
    return PAMD_ERR_ARG;
  *bytes = projected_layout(cfg->nocc, cfg->nvir, nfrag, with_vjp ? 1 : 0, cfg->precision).total;
  return PAMD_OK;
}

static int projected_impl(const pamd_ccsdt_config *cfg, const double *mo_energy, const double *t1,
                          const double *t2, const double *ovvv, const double *ovoo,
                          const double *ovov, const double *fvo, const double *ulo, int nfrag,
                          int vjp, double et_bar, void *workspace, size_t workspace_bytes,
                          double *e_out, double *mo_energy_bar, double *t1_bar, double *t2_bar,
                          double *ovvv_bar, double *ovoo_bar, double *ovov_bar, double *fvo_bar,
                          double *ulo_bar, void *stream) {
  int st = check_common(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo);
  if (st != PAMD_OK) return st;
  if (!ulo || !e_out || nfrag < 1 || nfrag > kMaxFrag) return PAMD_ERR_ARG;
  if (vjp && (!mo_energy_bar || !t1_bar || !t2_bar || !ovvv_bar || !ovoo_bar || !ovov_bar ||
              !fvo_bar || !ulo_bar))
    return PAMD_ERR_ARG;
  const ProjLayout P = projected_layout(cfg->nocc, cfg->nvir, nfrag, vjp, cfg->precision);
  if (!workspace || workspace_bytes < P.total) return PAMD_ERR_WORKSPACE;
  Ctx c;
  memset(&c, 0, sizeof(c));
  c.o = cfg->nocc;
  c.v = cfg->nvir;
  c.packed = cfg->ovvv_packed ? 1 : 0;
  c.has_fvo = cfg->has_fvo ? 1 : 0;
  c.vjp = vjp;
  c.prec = cfg->precision;
  c.s = (hipStream_t)stream;
  c.mo = mo_energy;
  c.t1 = t1;
  c.t2 = t2;
  c.ovvv = ovvv;
  c.ovoo = ovoo;
  c.ovov = ovov;
  c.fvo = fvo;
  c.ebar = et_bar;
  c.t1_bar = t1_bar;
  c.t2_bar = t2_bar;
  c.ovoo_bar = ovoo_bar;
  c.ovov_bar = ovov_bar;
  c.fvo_bar = fvo_bar;
  bind_workspace(c, P.L, workspace);
  char *b = (char *)workspace;
  double *pair_e = (double *)(b + P.off_p);
  st = run_projected(c, ulo, nfrag, (double *)(b + P.off_a), (double *)(b + P.off_b), pair_e,
                     (double *)(b + P.off_d), (double *)(b + P.off_u), ulo_bar);
  if (st != PAMD_OK) return st;
  PAMD_CHECK(launch_projected_total(pair_e, c.o, e_out, c.s));
  if (vjp) {
    const int o = c.o, v = c.v;
    PAMD_CHECK(launch_reduce_rows(c.part_mo, c.G, o + v, mo_energy_bar, 0, c.s));
    PAMD_CHECK(launch_fold_ovvv(c.Vpb, o, v, c.packed, ovvv_bar, c.s));
    PAMD_CHECK(launch_fold_t2m(c.T2mb, o, v, t2_bar, c.s));
  }
  return PAMD_OK;
}

int pamd_ccsdt_projected_energy(const pamd_ccsdt_config *cfg, const double *mo_energy,
                                const double *t1, const double *t2, const double *ovvv,
                                const double *ovoo, const double *ovov, const double *fvo,
                                const double *ulo, int nfrag, void *workspace,
                                size_t workspace_bytes, double *e_out, void *stream) {
  return projected_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, ulo, nfrag, 0, 0.0,
                        workspace, workspace_bytes, e_out, nullptr, nullptr, nullptr, nullptr,
                        nullptr, nullptr, nullptr, nullptr, stream);
}

int pamd_ccsdt_projected_energy_vjp(const pamd_ccsdt_config *cfg, const double *mo_energy,
                                    const double *t1, const double *t2, const double *ovvv,
                                    const double *ovoo, const double *ovov, const double *fvo,
                                    const double *ulo, int nfrag, double et_bar, void *workspace,
                                    size_t workspace_bytes, double *e_out, double *mo_energy_bar,
                                    double *t1_bar, double *t2_bar, double *ovvv_bar,
                                    double *ovoo_bar, double *ovov_bar, double *fvo_bar,
                                    double *ulo_bar, void *stream) {
  return projected_impl(cfg, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, ulo, nfrag, 1, et_bar,
                        workspace, workspace_bytes, e_out, mo_energy_bar, t1_bar, t2_bar,
                        ovvv_bar, ovoo_bar, ovov_bar, fvo_bar, ulo_bar, stream);
}

int pamd_debug_gemm_direct_ok(const void *desc) {
  if (!desc) return -PAMD_ERR_ARG;
  GemmDesc d;
  memcpy(&d, desc, sizeof(d));
  return gemm_direct_ok(d) ? 1 : 0;
}

int pamd_debug_gemm_b3p_ok(const void *desc, long long plane_stride) {
  if (!desc) return -PAMD_ERR_ARG;
  GemmDesc d;
  memcpy(&d, desc, sizeof(d));
  for (int q = 0; q < d.nseg; ++q)
    for (OpDesc *op : {&d.seg[q].A, &d.seg[q].B}) {
      // FP64 element address e * 8 -> bf16 element address e * 2
      op->p = (const double *)((uintptr_t)op->p / 4);
      op->ps = plane_stride;
    }
  return gemm_b3p_ok(d) ? 1 : 0;
}

int pamd_debug_gemm_descriptors(int nocc, int nvir, int i, int j, int k, int with_vjp,
                                const unsigned long long *bases, void *out, int max_desc,
                                int *n_desc, int *desc_bytes) {
  if (nocc < 1 || nvir < 1 || !bases || !n_desc || !desc_bytes) return PAMD_ERR_ARG;
  if (i < 0 || j < i || k < j || k >= nocc) return PAMD_ERR_ARG;
  Layout L = make_layout(nocc, nvir, with_vjp ? 1 : 0, PAMD_PREC_FP64);
  Ctx c;
  memset(&c, 0, sizeof(c));
  c.o = nocc;
  c.v = nvir;
  c.vjp = with_vjp ? 1 : 0;
  c.Vp = (double *)bases[0];
  c.T2m = (double *)bases[1];
  c.OOt = (double *)bases[2];
  c.W = (double *)bases[3];
  c.gW = c.W;
  c.t2 = (const double *)bases[4];
  c.gT2 = c.t2;
  c.Vb = (double *)bases[5];
  c.Vpb = (double *)bases[6];
  c.T2mb = (double *)bases[7];
  c.part_p = (double *)bases[8];
  c.part_h = (double *)bases[9];
  c.S_p = L.S_p;
  c.S_h = L.S_h;
  std::vector<GemmDesc> rec;
  g_record = &rec;
  int st = build_w(c, i, j, k);
  if (st == PAMD_OK && with_vjp) st = adjoint(c, i, j, k);
  g_record = nullptr;
  if (st != PAMD_OK) return st;
  *desc_bytes = (int)sizeof(GemmDesc);
  *n_desc = (int)rec.size();
  if (out) {
    const int n = std::min((int)rec.size(), max_desc);
    memcpy(out, rec.data(), sizeof(GemmDesc) * (size_t)n);
  }
  return PAMD_OK;
}

int pamd_debug_paired_adjoint_descriptors(int nocc, int nvir, int i, int j, int kp, int k,
                                          const unsigned long long *bases, void *out,
                                          int max_desc, int *n_desc, int *desc_bytes) {
  if (nocc < 1 || nvir < 1 || !bases || !n_desc || !desc_bytes) return PAMD_ERR_ARG;
  if (i < 0 || j < i || kp < j || k <= kp || k >= nocc) return PAMD_ERR_ARG;
  Layout L = make_layout(nocc, nvir, 1, PAMD_PREC_FP64);
  Ctx c1;
  memset(&c1, 0, sizeof(c1));
  c1.o = nocc;
  c1.v = nvir;
  c1.vjp = 1;
  c1.Vp = (double *)bases[0];
  c1.T2m = (double *)bases[1];
  c1.OOt = (double *)bases[2];
  c1.W = (double *)bases[3];
  c1.gW = c1.W;
  c1.t2 = (const double *)bases[4];
  c1.gT2 = c1.t2;
  c1.Vb = (double *)bases[5];
  c1.Vpb = (double *)bases[6];
  c1.T2mb = (double *)bases[7];
  c1.part_p = (double *)bases[8];
  c1.part_h = (double *)bases[9];
  c1.S_p = L.S_p;
  c1.S_h = L.S_h;
  Ctx c2 = c1;
  c2.W = (double *)bases[10];
  c2.gW = c2.W;
  std::vector<GemmDesc> rec;
  g_record = &rec;
  int st = adjoint_gemms(c1, i, j, kp, 1);
  if (st == PAMD_OK) st = adjoint_gemms(c2, i, j, k, 2, &c1, kp);
  g_record = nullptr;
  if (st != PAMD_OK) return st;
  *desc_bytes = (int)sizeof(GemmDesc);
  *n_desc = (int)rec.size();
  if (out) {
    const int n = std::min((int)rec.size(), max_desc);
    memcpy(out, rec.data(), sizeof(GemmDesc) * (size_t)n);
  }
  return PAMD_OK;
}

int pamd_profile_enable(int every) {
  if (every < 0) return PAMD_ERR_ARG;
  g_prof_every = every;
  return PAMD_OK;
}

int pamd_profile_progress(int every) {
  if (every < 0) return PAMD_ERR_ARG;
  g_progress_every = every;
  return PAMD_OK;
}

int pamd_profile_reset(void) {
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  memset(g_prof_stat, 0, sizeof(g_prof_stat));
  return PAMD_OK;
}

int pamd_profile_num_kinds(void) { return kProfNum; }

const char *pamd_profile_kind_name(int kind) {
  return (kind >= 0 && kind < kProfNum) ? kProfNames[kind] : "";
}

int pamd_profile_read(int kind, long long *launches, double *ms, double *flops, double *bytes) {
  if (kind < 0 || kind >= kProfNum || !launches || !ms || !flops || !bytes) return PAMD_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  *launches = g_prof_stat[kind].launches;
  *ms = g_prof_stat[kind].ms;
  *flops = g_prof_stat[kind].flops;
  *bytes = g_prof_stat[kind].bytes;
  return PAMD_OK;
}

int pamd_debug_split_factors(int nocc, int nvir, int *S_p, int *S_h) {
  return pamd_debug_split_factors_prec(nocc, nvir, PAMD_PREC_FP64, S_p, S_h);
}

int pamd_debug_split_factors_prec(int nocc, int nvir, int precision, int *S_p, int *S_h) {
  if (nocc < 1 || nvir < 1 || !S_p || !S_h) return PAMD_ERR_ARG;
  if (precision != PAMD_PREC_FP64 && precision != PAMD_PREC_BF16X3 && precision != PAMD_PREC_FP32X2)
    return PAMD_ERR_ARG;
  Layout L = make_layout(nocc, nvir, 1, precision);
  *S_p = L.S_p;
  *S_h = L.S_h;
  return PAMD_OK;
}

}  // extern "C"
