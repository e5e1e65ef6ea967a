// Host driver and C ABI of the MP2 pair energies and of the DF integral
// producer (include/pyscfad_amd_eris.h).
//
// The DF products are segmented GEMMs on the FP64 MFMA contraction engine
// (GemmDesc, ccsdt_common.h); the Lvv packing and the scatter of its
// cotangent are streaming kernels (eris_kernels.hip).  With n = nmo,
// v = n - o, np = v (v + 1) / 2 and ia = i v + a:
//   forward   ovvv[ia, bc] = sum_P Lov[P, ia] Lvv[P, bc]          (M o v, N np, K naux)
//             ovoo[ia, jk] = sum_P Lov[P, ia] Loo[P, jk]          (M o v, N o^2)
//             ovov[ia, jb] = sum_P Lov[P, ia] Lov[P, jb]          (M o v, N o v)
//   adjoint   Loo_bar[P, jk] += sum_ia Lov[P, ia] ovoo_bar[ia, jk]
//             Lov_bar[P, ia] += sum_bc Lvv[P, bc] ovvv_bar[ia, bc] + sum_jk Loo[P, jk] ovoo_bar[ia, jk]
//                             + sum_jb Lov[P, jb] (ovov_bar[ia, jb] + ovov_bar[jb, ia])
//             Lvv_bar[P, bc]  = sum_ia Lov[P, ia] ovvv_bar[ia, bc]
// The engine addresses operands with 32-bit element offsets from each
// descriptor's base, so products that read ovvv_bar are blocked over the
// occupied index i to keep every offset below 2^31.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <vector>
#include "ccsdt_common.h"
#include "gemm_desc.h"
#include "eris_common.h"
#include "../../include/pyscfad_amd_ccsdt.h"
#include "../../include/pyscfad_amd_eris.h"

namespace pamd {
namespace {

#define ERIS_CHECK(expr)                                                      \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "pyscfad_amd: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return PAMD_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

const long long kMaxOffset = (1LL << 31) - 1;

struct DfShape {
  int o, n, v, naux;
  long long np, n2, ov;
};

int df_shape(int nocc, int nmo, int naux, DfShape &s) {
  if (nocc < 1 || nmo <= nocc || naux < 1 || naux > 65535 || nmo > 65535) return PAMD_ERR_ARG;
  s.o = nocc;
  s.n = nmo;
  s.v = nmo - nocc;
  s.naux = naux;
  s.np = (long long)s.v * (s.v + 1) / 2;
  s.n2 = (long long)nmo * nmo;
  s.ov = (long long)s.o * s.v;
  // every operand array other than ovvv(_bar) must be addressable in one descriptor
  if ((long long)naux * s.n2 > kMaxOffset || (long long)naux * s.np > kMaxOffset ||
      s.ov * s.ov > kMaxOffset || s.ov * s.o * s.o > kMaxOffset || s.v * s.np > kMaxOffset ||
      s.ov > kMaxOffset / 8 || s.np > kMaxOffset / 8)
    return PAMD_ERR_UNSUPPORTED;
  return PAMD_OK;
}

// occupied block of the ovvv(_bar) products: i0 .. i0 + ib - 1
int ovvv_block(const DfShape &s, int forced) {
  if (forced > 0) return std::min(forced, s.o);
  const long long per_i = (long long)s.v * s.np;
  return (int)std::max(1LL, std::min<long long>(s.o, kMaxOffset / per_i));
}

size_t lvv_bytes(const DfShape &s) { return ((size_t)s.naux * s.np * 8 + 255) & ~(size_t)255; }

struct DfPtrs {
  const double *Lpq;
  double *Lvv, *Lvv_bar;
  double *ovvv, *ovoo, *ovov;  // outputs (forward) or cotangents (adjoint; read only)
  double *Lpq_bar;
};

void forward_descs(const DfShape &s, const DfPtrs &p, std::vector<GemmDesc> &out_d) {
  const int o = s.o, v = s.v, n = s.n;
  const int OV = (int)s.ov, NP = (int)s.np;
  // Lov^T as the row operand: A(ia, P) = Lpq[P, i, o + a]
  const OpDesc lovA = opR(p.Lpq + o, n, 1, v, s.n2, 1);
  if (p.ovvv)
    out_d.push_back(gemm(OV, NP, kStore, {seg(lovA, op1(p.Lvv, 1, s.np, 1), s.naux, 1.0)},
                         out(p.ovvv, 0, s.np, kBig, 0, 1, kBig)));
  if (p.ovoo)
    out_d.push_back(gemm(OV, o * o, kStore,
                         {seg(lovA, opR(p.Lpq, n, 1, o, s.n2, 1), s.naux, 1.0)},
                         out(p.ovoo, 0, (long long)o * o, kBig, 0, 1, kBig)));
  if (p.ovov)
    out_d.push_back(gemm(OV, OV, kStore, {seg(lovA, opR(p.Lpq + o, n, 1, v, s.n2, 1), s.naux, 1.0)},
                         out(p.ovov, 0, s.ov, kBig, 0, 1, kBig)));
}

void adjoint_descs(const DfShape &s, const DfPtrs &p, std::vector<GemmDesc> &out_d,
                   int forced_block = 0) {
  const int o = s.o, v = s.v, n = s.n, naux = s.naux;
  const int OV = (int)s.ov, NP = (int)s.np, OO = o * o;
  // Lov and Loo as row operands over P with the contraction index on K
  const OpDesc lovK = opK(p.Lpq + o, s.n2, n, 1, v, 0);    // A(P, ia) = Lpq[P, i, o + a]
  const OpDesc looK = opK(p.Lpq, s.n2, n, 1, o, 0);        // A(P, jk) = Lpq[P, j, k]
  // Loo_bar[P, jk] += sum_ia Lov[P, ia] ovoo_bar[ia, jk]
  if (p.ovoo)
    out_d.push_back(gemm(naux, OO, kAccumulate, {seg(lovK, op1(p.ovoo, 1, OO, 1), OV, 1.0)},
                         out(p.Lpq_bar, 0, s.n2, kBig, n, 1, o)));
  const int ib = ovvv_block(s, forced_block);
  for (int i0 = 0; i0 < o; i0 += ib) {
    const int ni = std::min(ib, o - i0);
    const int NB = ni * v;  // columns ia of this block
    // Lov_bar[P, ia] for i in [i0, i0 + ni)
    std::vector<SegDesc> sg;
    if (p.ovvv)
      sg.push_back(seg(op1(p.Lvv, s.np, 1, 0), op1(p.ovvv + (long long)i0 * v * s.np, s.np, 1, 0),
                       NP, 1.0));
    if (p.ovoo)
      sg.push_back(seg(looK, op1(p.ovoo + (long long)i0 * v * OO, OO, 1, 0), OO, 1.0));
    if (p.ovov) {
      sg.push_back(seg(lovK, op1(p.ovov + (long long)i0 * v * s.ov, s.ov, 1, 0), OV, 1.0));
      sg.push_back(seg(lovK, op1(p.ovov + (long long)i0 * v, 1, s.ov, 1), OV, 1.0));
    }
    if (!sg.empty())
      out_d.push_back(gemm(naux, NB, kAccumulate, sg.data(), sg.data() + sg.size(),
                           out(p.Lpq_bar + o + (long long)i0 * n, 0, s.n2, kBig, n, 1, v)));
    // Lvv_bar[P, bc] (+)= sum_{ia in block} Lov[P, ia] ovvv_bar[ia, bc]
    if (p.ovvv)
      out_d.push_back(gemm(naux, NP, i0 == 0 ? kStore : kAccumulate,
                           {seg(opK(p.Lpq + o + (long long)i0 * n, s.n2, n, 1, v, 0),
                                op1(p.ovvv + (long long)i0 * v * s.np, 1, s.np, 1), NB, 1.0)},
                           out(p.Lvv_bar, 0, s.np, kBig, 0, 1, kBig)));
  }
}

int run_descs(const std::vector<GemmDesc> &ds, hipStream_t st) {
  for (const GemmDesc &d : ds) ERIS_CHECK(launch_gemm_f64(d, st));
  return PAMD_OK;
}

}  // namespace
}  // namespace pamd

using namespace pamd;

extern "C" {

int pamd_mp2_init_amps(int nocc, int nvir, const double *mo_energy, const double *ovov,
                       double *t2_out, double *e_pair, void *stream) {
  if (nocc < 1 || nvir < 1 || nocc > 4096 || !mo_energy || !ovov || !e_pair) return PAMD_ERR_ARG;
  ERIS_CHECK(launch_mp2_pairs(nocc, nvir, mo_energy, ovov, t2_out, e_pair, (hipStream_t)stream));
  return PAMD_OK;
}

int pamd_df_eris_workspace_size(int nocc, int nmo, int naux, int with_vjp, size_t *bytes) {
  if (!bytes) return PAMD_ERR_ARG;
  DfShape s;
  const int st = df_shape(nocc, nmo, naux, s);
  if (st != PAMD_OK) return st;
  *bytes = lvv_bytes(s) * (with_vjp ? 2 : 1);
  return PAMD_OK;
}

int pamd_df_eris(int nocc, int nmo, int naux, const double *Lpq, double *ovvv, double *ovoo,
                 double *ovov, void *workspace, size_t workspace_bytes, void *stream) {
  DfShape s;
  int st = df_shape(nocc, nmo, naux, s);
  if (st != PAMD_OK) return st;
  if (!Lpq) return PAMD_ERR_ARG;
  if (ovvv && (!workspace || workspace_bytes < lvv_bytes(s))) return PAMD_ERR_WORKSPACE;
  hipStream_t hs = (hipStream_t)stream;
  DfPtrs p{Lpq, (double *)workspace, nullptr, ovvv, ovoo, ovov, nullptr};
  if (ovvv) ERIS_CHECK(launch_pack_lvv(Lpq, nocc, nmo, naux, p.Lvv, hs));
  std::vector<GemmDesc> ds;
  forward_descs(s, p, ds);
  return run_descs(ds, hs);
}

int pamd_df_eris_vjp(int nocc, int nmo, int naux, const double *Lpq, const double *ovvv_bar,
                     const double *ovoo_bar, const double *ovov_bar, double *Lpq_bar,
                     void *workspace, size_t workspace_bytes, void *stream) {
  DfShape s;
  int st = df_shape(nocc, nmo, naux, s);
  if (st != PAMD_OK) return st;
  if (!Lpq || !Lpq_bar) return PAMD_ERR_ARG;
  if (!workspace || workspace_bytes < 2 * lvv_bytes(s)) return PAMD_ERR_WORKSPACE;
  hipStream_t hs = (hipStream_t)stream;
  DfPtrs p{Lpq, (double *)workspace, (double *)((char *)workspace + lvv_bytes(s)),
           const_cast<double *>(ovvv_bar), const_cast<double *>(ovoo_bar),
           const_cast<double *>(ovov_bar), Lpq_bar};
  ERIS_CHECK(launch_zero_occ_rows(nocc, nmo, naux, Lpq_bar, hs));
  if (ovvv_bar) ERIS_CHECK(launch_pack_lvv(Lpq, nocc, nmo, naux, p.Lvv, hs));
  else ERIS_CHECK(launch_fill(p.Lvv_bar, (size_t)naux * s.np, 0.0, hs));
  std::vector<GemmDesc> ds;
  adjoint_descs(s, p, ds);
  st = run_descs(ds, hs);
  if (st != PAMD_OK) return st;
  ERIS_CHECK(launch_unpack_lvv_bar(p.Lvv_bar, nocc, nmo, naux, Lpq_bar, hs));
  return PAMD_OK;
}

int pamd_debug_df_descriptors(int nocc, int nmo, int naux, int with_vjp, int block_i,
                              const unsigned long long *bases, void *out, int max_desc,
                              int *n_desc, int *desc_bytes) {
  if (!bases || !n_desc || !desc_bytes) return PAMD_ERR_ARG;
  DfShape s;
  const int st = df_shape(nocc, nmo, naux, s);
  if (st != PAMD_OK) return st;
  DfPtrs p{(const double *)bases[0], (double *)bases[1], (double *)bases[2], (double *)bases[3],
           (double *)bases[4],       (double *)bases[5], (double *)bases[6]};
  std::vector<GemmDesc> ds;
  if (with_vjp) adjoint_descs(s, p, ds, block_i);
  else forward_descs(s, p, ds);
  *desc_bytes = (int)sizeof(GemmDesc);
  *n_desc = (int)ds.size();
  if (out) memcpy(out, ds.data(), sizeof(GemmDesc) * (size_t)std::min((int)ds.size(), max_desc));
  return PAMD_OK;
}

}  // extern "C"
