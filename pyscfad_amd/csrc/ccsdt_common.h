// Shared declarations of the (T) kernels and their host driver.
//
// The triples are walked over occupied index triples i <= j <= k.  For one
// triple the connected triples amplitude W_ijk[a,b,c] (v^3) is the sum of six
// particle and six hole contractions (the joint (ijk,abc) permutations of
// w_abc[ijk] = sum_f ovvv[i,a,f,b] t2[k,j,c,f] - sum_m ovoo[i,a,m,j] t2[k,m,c,b],
// pyscfadlib/pyscfadlib/vjp/cc/ccsd_t.c:121-142).  Every contraction -- the
// W build and all adjoint contractions of the backward -- is one instance of a
// single strided "segmented GEMM" problem run by the MFMA engine:
//
//   C[m, n]  (=, +=, or split-K partial)  sum_s alpha_s sum_k A_s(m, k) B_s(k, n)
//
// where every operand element is addressed as
//   A_s(m, k) = A.p[(m / A.rdiv) * A.rx + (m % A.rdiv) * A.ry
//                   + (k / A.kdiv) * A.kx + (k % A.kdiv) * A.ky]
// (likewise B_s(k, n) with n in the role of m), so that pairs of virtual indices
// can sit on the M, N or K side of a contraction without repacking tensors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pamd {

constexpr int kMaxSeg = 4;

struct OpDesc {
  const double *p;
  long long rx, ry, kx, ky;  // strides: row index (m or n) split (r / rdiv, r % rdiv), K likewise
  int rdiv, kdiv;
  int rfast;                 // 1: contiguous along the row index -> stage with threads along rows
  int pad_;
  // split-BF16x3 operands are planar: p addresses bf16 elements, and element e
  // of plane q (0 hi, 1 mid, 2 lo) is ((const uint16_t *)p)[e + q * ps].  The
  // FP64 and FP32 hi/lo engines (8-byte elements) ignore ps.
  long long ps;
};

struct SegDesc {
  OpDesc A, B;
  int L;          // contraction length of this segment
  int pad_;
  double alpha;
};

struct OutDesc {
  double *p;
  long long mx, my, nx, ny;
  int mdiv, ndiv;
};

enum GemmMode { kStore = 0, kAccumulate = 1, kPartial = 2 };

struct GemmDesc {
  int M, N, nseg, mode;
  int chunk_start[kMaxSeg + 1];  // prefix sums of ceil(L_s / KC) in K chunks
  int splits;                    // split-K factor (grid z); partial mode only when > 1
  int pad_;
  SegDesc seg[kMaxSeg];
  OutDesc C;
  double *partial;               // kPartial: partial[(split * M + m) * N + n]
};

// A descriptor of at most two segments (same fields as GemmDesc), for group
// launches whose kernel argument holds several of them.
struct GemmDescS {
  int M, N, nseg, mode;
  int chunk_start[3];
  int splits;
  SegDesc seg[2];
  OutDesc C;
  double *partial;
};
constexpr int kMaxGroup = 6;
struct GemmGroup {
  GemmDescS d[kMaxGroup];
  int n;
  int pad_;
};

// A batch of nb products that share one descriptor up to pointer offsets:
// product b reads seg[s].A.p + b * a[s] and seg[s].B.p + b * bo[s] and writes
// C.p + b * c (element offsets; not for the split-K partial mode).  One launch
// then runs nb products, which fills the GPU when one product has too few
// output tiles (small nvir).
struct GemmBatch {
  long long a[kMaxSeg], bo[kMaxSeg], c;
  int nb;
  int pad_;
};
inline GemmBatch no_batch() {
  GemmBatch b;
  for (int q = 0; q < kMaxSeg; ++q) b.a[q] = b.bo[q] = 0;
  b.c = 0;
  b.nb = 1;
  b.pad_ = 0;
  return b;
}

constexpr int kMaxFrag = 8;  // rows of the local-orbital projector ulo

// operand formats of the GEMM engines: FP64; split-BF16x3 as three planes of
// bf16 (hi, mid, lo; OpDesc::ps); FP32 hi/lo pairs packed in 8-byte elements
enum SplitFormat { kSplitNone = 0, kSplitBf16x3 = 1, kSplitF32x2 = 2 };

// Epilogue over orbits of virtual tiles for one occupied triple.
struct OrbitArgs {
  double *W;            // v^3; W in (and Wbar out when Wb8 and Wb3 are null, vjp)
  unsigned long long *Wb8;  // v^3 packed FP32 hi/lo pairs of Wbar (FP32 hi/lo vjp) or null
  unsigned short *Wb3;      // 3 planes of v^3 bf16 (hi, mid, lo) of Wbar (split-BF16x3 vjp) or null
  double *Vb;           // v^3; Vbar out (vjp)
  // OVS[x][y][p][q] = ovov[x,p,y,q] + ovov[y,q,x,p], T2S[x][y][p][q] = t2[x,y,p,q] + t2[y,x,q,p]
  const double *t1, *ovs, *t2s, *fvo, *mo;
  const unsigned short *tab;  // [norb][4] orbit table (launch_orbit_table)
  int o, v, i, j, k, has_fvo, vjp;
  long long norb;
  double omega;         // 2 s_ijk: 2, 1 (two indices equal)
  double ebar;          // cotangent of E
  double *part_t;       // [G] per-workgroup unweighted triple energy (overwritten)
  double *part_mo;      // [G][n] per-workgroup mo_energy_bar accumulators (vjp)
  // projector accumulation (pamd_ccsdt_projected_energy): when accA is set,
  // accA[f] += u[f] W and accB[f] += u[f] Z / D over the v^3 block, f < nfrag
  double *accA, *accB;
  int nfrag;
  double u[kMaxFrag];
  double *part_u;       // [G][kMaxFrag] per-workgroup dE/du_f sums (projected adjoint)
};

enum OrbitMode { kOrbitPlain = 0, kOrbitProjAcc = 1, kOrbitProjVjp = 2 };

// V-bar reductions: the sums over one virtual index are split into kVRS parts
// (more workgroups in flight), and the last virtual index is processed in
// column blocks of kVRC; the scratch R holds every part.
constexpr int kVRS = 4;
constexpr int kVRC = 512;
__host__ __device__ inline int vr_col_blocks(int v) { return (v + kVRC - 1) / kVRC; }
inline size_t vr_scratch_doubles(int v) {
  const size_t vv = (size_t)v * v, ncb = (size_t)vr_col_blocks(v);
  return 10 * vv + 2 * (size_t)v + (size_t)(kVRS - 1) * 6 * vv +
         (kVRS * ncb - 1) * 2 * (size_t)v + (ncb - 1) * 4 * vv;
}

// V-bar reductions for one triple.
struct VRArgs {
  const double *Vb, *t1, *fvo, *ovs, *t2s;  // ovs, t2s: pair-symmetrised blocks (OrbitArgs)
  int o, v, i, j, k, has_fvo;
  double *R;            // scratch: vr_scratch_doubles(v) (layout in kernels)
};

// Per-triple gather of the adjoint pieces into the output cotangents.
struct CombineArgs {
  const double *R;                // VR scratch
  double *part_p;                 // [6][S][v][v] particle t2_bar partials (reduced in place)
  double *part_h;                 // [6][S][v][o] hole ovoo_bar partials (reduced in place)
  int S_p, S_h;
  int o, v, i, j, k, has_fvo;
  double *t2_bar, *ovov_bar, *ovoo_bar, *t1_bar, *fvo_bar;
  // when part_t is set: *tri_dst = sum of the orbit epilogue's G per-workgroup
  // energy sums (fixed order), in the same launch
  const double *part_t;
  int G;
  double *tri_dst;
};

// host-side launchers (ccsdt_kernels.hip)
hipError_t launch_fill(double *p, size_t n, double val, hipStream_t s);
hipError_t launch_orbit_table(long long norb, unsigned short *tab, hipStream_t s);
hipError_t launch_prep_sym(const double *src, int o, int v, int kind, double *dst, hipStream_t s);
hipError_t launch_prepare(const double *ovvv, int packed, const double *t2, const double *ovoo,
                          int o, int v, double *Vp, double *T2m, double *OOt, int split,
                          hipStream_t s);
hipError_t launch_split_f32x2(const double *x, size_t n, unsigned long long *out, hipStream_t s);
hipError_t launch_gemm_f32x2(const GemmDesc &d, hipStream_t s, const GemmBatch &bt = no_batch());
hipError_t launch_split_bf16x3(const double *x, size_t n, unsigned short *out, hipStream_t s);
// planar split-BF16x3 VpT[x][f][p][q] and VpU[x][f][q][p] from Vp[x][p][q][f]
hipError_t launch_vp_transposes(const unsigned short *Vp, int o, int v, unsigned short *VpT,
                                unsigned short *VpU, hipStream_t s);
hipError_t launch_gemm_bf16x3(const GemmDesc &d, hipStream_t s, const GemmBatch &bt = no_batch());
// split-BF16x3 engine with direct-to-LDS operand loads (gemm_b3p.hip), on
// descriptors whose chunk_start counts stages of 32; launch_gemm_bf16x3 picks
// it when gemm_b3p_ok(d)
bool gemm_b3p_ok(const GemmDesc &d, const GemmBatch &bt = no_batch());
bool gemm_b3p_enabled();  // env PYSCFAD_AMD_B3_DIRECT=0 turns it off
int gemm_b3p_tiles(int M, int N);  // workgroup tiles of one split
hipError_t launch_gemm_b3p(const GemmDesc &d, hipStream_t s, const GemmBatch &bt = no_batch());
// n independent products in one launch of the direct engine (chunk_start in
// stages of 32), when gemm_b3p_group_ok (see gemm_b3p.hip)
bool gemm_b3p_group_ok(const GemmDesc *ds, int n);
hipError_t launch_gemm_b3p_group(const GemmDesc *ds, int n, hipStream_t s);
// split-BF16x3 group launch (chunk_start in chunks of KC): one launch when the
// products form a group on the direct engine, else one launch each
hipError_t launch_gemm_bf16x3_group(const GemmDesc *ds, int n, hipStream_t s);
hipError_t launch_gemm_f64(const GemmDesc &d, hipStream_t s, const GemmBatch &bt = no_batch());
// FP64 / FP32 hi/lo engine with direct-to-LDS operand loads (gemm_direct.hip):
// usable when gemm_direct_ok(d); launch_gemm_f64 / launch_gemm_f32x2 pick it then
bool gemm_direct_ok(const GemmDesc &d, const GemmBatch &bt = no_batch());
hipError_t launch_gemm_f64_direct(const GemmDesc &d, hipStream_t s, bool fx2,
                                  const GemmBatch &bt = no_batch());
bool gemm_direct_enabled();  // env PYSCFAD_AMD_GEMM_DIRECT=0 turns the direct engine off
int gemm_direct_tiles(int M, int N, bool fx2);  // direct FP64 / FP32 hi/lo engine: tiles of one split
int gemm_tiles(int M, int N, bool fx2 = false);  // FP64 / FP32 hi/lo engine: tiles of one split
int gemm_tiles_b3(int M, int N);  // split-BF16x3 engine: workgroup tiles of one split
int gemm_wgs_per_cu_b3(int M, int N);  // split-BF16x3 engine: resident workgroups per CU
hipError_t launch_orbit(const OrbitArgs &a, int grid, hipStream_t s, int mode = kOrbitPlain);
hipError_t launch_vr(const VRArgs &a, hipStream_t s);
hipError_t launch_combine(const CombineArgs &a, hipStream_t s);
hipError_t launch_fold_ovvv(const double *Vpb, int o, int v, int packed, double *out, hipStream_t s);
hipError_t launch_fold_t2m(const double *T2mb, int o, int v, double *t2_bar, hipStream_t s);
hipError_t launch_reduce_rows(const double *part, int rows, int cols, double *out, int accumulate,
                              hipStream_t s);
hipError_t launch_triple_energy(const double *part_t, int G, double *dst, hipStream_t s);
int orbit_grid(int v);
hipError_t launch_pair_energy(const double *tri_e, int o, double *e_pair, hipStream_t s);
hipError_t launch_total_energy(const double *tri_e, int o, double *e_out, hipStream_t s);
hipError_t launch_dot(const double *a, const double *b, size_t n, double *part, int G, double *dst,
                      hipStream_t s);
hipError_t launch_reduce_ulo(const double *part_u, int G, int nfrag, double scale, double *dst,
                             int stride, hipStream_t s);
hipError_t launch_projected_total(const double *pair_e, int o, double *e_out, hipStream_t s);
hipError_t launch_weighted_energy(const double *tri_e, int o, const double *m, double ebar,
                                  double *e_out, double *m_bar, hipStream_t s);

}  // namespace pamd
