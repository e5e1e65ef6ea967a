/*
 * pyscfad_amd: closed-shell CCSD(T) triples energy E(T) and its adjoint on one
 * AMD MI355X (gfx950).  C ABI of libpyscfad_amd_cc.so.
 *
 * All array arguments are DEVICE pointers to C-contiguous float64 arrays in the
 * reference's own layouts (pyscfad/cc/ccsd_t.py:36-42, pyscfad/cc/ccsd.py:30-31):
 *   mo_energy (n)            n = nocc + nvir
 *   t1        (o, v)
 *   t2        (o, o, v, v)
 *   ovvv      (o, v, v, v) full, or (o, v, v(v+1)/2) lower-triangle packed
 *             (the DF layout, pyscfad/cc/dfccsd.py:88) when cfg->ovvv_packed
 *   ovoo      (o, v, o, o)   chemists' (ia|jk)
 *   ovov      (o, v, o, v)   chemists' (ia|jb)
 *   fvo       (v, o)         fock[o:, :o]; may be NULL when cfg->has_fvo == 0
 * The stream argument is a hipStream_t (0 = the null stream).  Calls on one
 * workspace must not overlap.  Every entry point returns a PAMD_* status.
 * Entry points keep no per-call state outside their arguments, so calls on
 * different workspaces may run from different host threads; only the sampled
 * timing switches of pyscfad_amd_profile.h are process-global.
 */
#ifndef PYSCFAD_AMD_CCSDT_H
#define PYSCFAD_AMD_CCSDT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PAMD_OK 0
#define PAMD_ERR_ARG 1          /* bad sizes, NULL pointer, unknown option     */
#define PAMD_ERR_HIP 2          /* a HIP runtime call failed                  */
#define PAMD_ERR_WORKSPACE 3    /* workspace smaller than pamd_ccsdt_workspace_size */
#define PAMD_ERR_UNSUPPORTED 4  /* option not built into this library         */

#define PAMD_PREC_FP64 0        /* plain form: FP64 MFMA                       */
#define PAMD_PREC_BF16X3 1      /* split-BF16x3 MFMA, FP32 tile accumulation   */
#define PAMD_PREC_FP32X2 2      /* FP32 hi/lo split on FP32 MFMA (comparison form:
                                   three products per multiply-add, FP32 tile sums
                                   added into FP64 every 128 k)                 */

typedef struct {
    int nocc;
    int nvir;
    int ovvv_packed;  /* 1: ovvv (and ovvv_bar) are (o, v, v(v+1)/2) packed   */
    int has_fvo;      /* 0: fvo is zero (the t2*fvo terms are skipped)          */
    int precision;    /* PAMD_PREC_*                                            */
    int reserved;
} pamd_ccsdt_config;

/* Device workspace needed by pamd_ccsdt_energy (with_vjp = 0) or
 * pamd_ccsdt_energy_vjp (with_vjp = 1).  Host-only call; no GPU work. */
int pamd_ccsdt_workspace_size(const pamd_ccsdt_config *cfg, int with_vjp, size_t *bytes);

/* E(T) -> *e_out (device, 1 double, overwritten).
 * Replaces the forward of the reference's (T) call: pyscf ccsd_t.kernel ->
 * libcc CCsd_t_contract at pyscfad/cc/ccsd_t.py:41, and its in-tree C sibling
 * lnoccsdt_contract (pyscfadlib/pyscfadlib/vjp/cc/lno_ccsd_t.c:96-100) with
 * mat = I, et = 2/3 e_tot (pyscfad/lno/ccsd_t.py:109-112). */
int pamd_ccsdt_energy(const pamd_ccsdt_config *cfg,
                      const double *mo_energy, const double *t1, const double *t2,
                      const double *ovvv, const double *ovoo, const double *ovov,
                      const double *fvo,
                      void *workspace, size_t workspace_bytes,
                      double *e_out, void *stream);

/* E(T) and its VJP for the cotangent et_bar, in one pass over the triples
 * (W/V are recomputed per triple block, nothing o^3 v^3 is stored).
 * Replaces ccsd_t_energy_vjp (pyscfadlib/pyscfadlib/vjp/cc/ccsd_t_vjp.c:332-346)
 * together with its host driver _ccsd_t_energy_vjp (pyscfad/cc/ccsd_t.py:82-192):
 * outputs are OVERWRITTEN with the cotangents in the reference's output layout
 * (ccsd_t.py:179-191): mo_energy_bar (n), t1_bar (o,v), t2_bar (o,o,v,v),
 * ovvv_bar (packed fold g+g^T, diagonal halved, when cfg->ovvv_packed; else the
 * unsymmetrised full (o,v,v,v) block), ovoo_bar (o,v,o,o), ovov_bar (o,v,o,v),
 * fvo_bar (v,o) (the vo block of fock_bar; required even when fvo == 0, since
 * dE/dfvo = 1/2 sum Vbar t2 does not vanish with fvo).  e_out receives E(T). */
int pamd_ccsdt_energy_vjp(const pamd_ccsdt_config *cfg,
                          const double *mo_energy, const double *t1, const double *t2,
                          const double *ovvv, const double *ovoo, const double *ovov,
                          const double *fvo, double et_bar,
                          void *workspace, size_t workspace_bytes,
                          double *e_out, double *mo_energy_bar, double *t1_bar,
                          double *t2_bar, double *ovvv_bar, double *ovoo_bar,
                          double *ovov_bar, double *fvo_bar, void *stream);

/* Number of occupied triples i <= j <= k (i == j == k excluded) of the walk. */
int pamd_ccsdt_num_triples(int nocc, long long *n);

/* One shard of E(T): the triples whose walk index t (i outer, then j, then k,
 * i <= j <= k, i == j == k skipped) has t % nshards == shard.  The shards'
 * results add up to pamd_ccsdt_energy's; this is the per-GPU share of a
 * sharded run (round-robin keeps the shares balanced).  Same workspace as
 * pamd_ccsdt_energy. */
int pamd_ccsdt_energy_shard(const pamd_ccsdt_config *cfg,
                            const double *mo_energy, const double *t1, const double *t2,
                            const double *ovvv, const double *ovoo, const double *ovov,
                            const double *fvo, int shard, int nshards,
                            void *workspace, size_t workspace_bytes,
                            double *e_out, void *stream);

/* One shard of E(T) and of its VJP: the triples whose walk index t has
 * t % nshards == shard (the walk and rounding order of
 * pamd_ccsdt_energy_shard), with the outputs of pamd_ccsdt_energy_vjp for that
 * share of the walk (overwritten; e_out is the shard's energy).  The shards'
 * cotangents add up to pamd_ccsdt_energy_vjp's.  Plays the role of the
 * reference's block-callable VJP, ccsd_t_energy_vjp(..., a0, a1, b0, b1, ...)
 * (pyscfadlib/pyscfadlib/vjp/cc/ccsd_t_vjp.c:332-346), which its host driver
 * calls once per virtual block (pyscfad/cc/ccsd_t.py:145-177); here the unit of
 * work is the occupied triple.  Same workspace as pamd_ccsdt_energy_vjp. */
int pamd_ccsdt_energy_vjp_shard(const pamd_ccsdt_config *cfg,
                                const double *mo_energy, const double *t1, const double *t2,
                                const double *ovvv, const double *ovoo, const double *ovov,
                                const double *fvo, double et_bar, int shard, int nshards,
                                void *workspace, size_t workspace_bytes,
                                double *e_out, double *mo_energy_bar, double *t1_bar,
                                double *t2_bar, double *ovvv_bar, double *ovoo_bar,
                                double *ovov_bar, double *fvo_bar, void *stream);

/* pamd_ccsdt_energy_vjp without a workspace argument and without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a library-owned device
 * array of pamd_ccsdt_noalloc_scratch_bytes() bytes (128 MiB), one per GPU,
 * held from the first call on that GPU until the process ends; problems whose
 * pamd_ccsdt_workspace_size(cfg, 1, ...) exceeds it are rejected with
 * PAMD_ERR_WORKSPACE (e.g. nocc=10/nvir=64 fits, nocc=16/nvir=64 does not).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (several streams or host threads) race and must be serialised by the caller.
 * The call overwrites every part of the array it reads, so nothing carries
 * over from an earlier call.  Arguments, outputs and results are those of
 * pamd_ccsdt_energy_vjp; sampled profiling is not recorded for these calls. */
int pamd_ccsdt_energy_vjp_noalloc(const pamd_ccsdt_config *cfg,
                                  const double *mo_energy, const double *t1, const double *t2,
                                  const double *ovvv, const double *ovoo, const double *ovov,
                                  const double *fvo, double et_bar,
                                  double *e_out, double *mo_energy_bar, double *t1_bar,
                                  double *t2_bar, double *ovvv_bar, double *ovoo_bar,
                                  double *ovov_bar, double *fvo_bar, void *stream);
size_t pamd_ccsdt_noalloc_scratch_bytes(void);

/* Occupied-pair resolved E(T): e_pair[i*o + j] = 1/3 sum_{k,abc} W Z / D,
 * sum(e_pair) = E(T).  Row sums are the occupied-weighted (T) of
 * lnoccsdt_contract with mat = e_i e_i^T (lno_ccsd_t.c:16-26); the labels of the
 * learned screen.  e_pair: device, o*o doubles, overwritten. */
int pamd_ccsdt_pair_energy(const pamd_ccsdt_config *cfg,
                           const double *mo_energy, const double *t1, const double *t2,
                           const double *ovvv, const double *ovoo, const double *ovov,
                           const double *fvo,
                           void *workspace, size_t workspace_bytes,
                           double *e_pair, void *stream);

/* Occupied-weighted E(T) for a diagonal occupied projector mat = diag(m) and
 * its adjoint: the weighted (T) of lnoccsdt_contract (lno_ccsd_t.c:6-29,
 * 96-138, with et *= 2/3 of pyscfad/lno/ccsd_t.py:104) and of its VJP
 * (lno_ccsd_t_vjp.c:400-566) when mat is diagonal.
 *   e_out = E_m = sum_i m_i E_i,   E_i = sum_j e_pair[i,j] (pamd_ccsdt_pair_energy)
 * occ_weight: HOST array of nocc doubles (m).  occ_weight_bar: device, nocc
 * doubles, overwritten with et_bar * E_i -- the derivative along the diagonal
 * of mat (the off-diagonal part of the reference's mat_bar is not formed).
 * The other cotangents are those of pamd_ccsdt_energy_vjp for E_m; with m = 1
 * they equal pamd_ccsdt_energy_vjp's bit for bit. */
int pamd_ccsdt_weighted_energy_vjp(const pamd_ccsdt_config *cfg,
                                   const double *mo_energy, const double *t1, const double *t2,
                                   const double *ovvv, const double *ovoo, const double *ovov,
                                   const double *fvo, const double *occ_weight, double et_bar,
                                   void *workspace, size_t workspace_bytes,
                                   double *e_out, double *mo_energy_bar, double *t1_bar,
                                   double *t2_bar, double *ovvv_bar, double *ovoo_bar,
                                   double *ovov_bar, double *fvo_bar, double *occ_weight_bar,
                                   void *stream);

/* Occupied-weighted E(T) of a local-orbital projector and its adjoint:
 * pyscfad/lno/ccsd_t.py:35-60 (mat = ulo^T ulo) with lnoccsdt_contract
 * (lno_ccsd_t.c:6-29, 96-138) and et *= 2/3 (lno/ccsd_t.py:104), and the VJP
 * of _ccsd_t_energy (lno/ccsd_t.py:124-237, lno_ccsd_t_vjp.c:400-566) chained
 * through mat = ulo^T ulo:
 *   E = 1/3 sum_{l,i,j,k,abc} mat[l,i] W_ljk[abc] Z_ijk[abc] / D_ijk[abc].
 * ulo: HOST array (nfrag, nocc), 1 <= nfrag <= 8 per call (E is linear in
 * mat = sum_f u_f u_f^T, so a larger fragment is the sum of calls on row groups
 * of ulo; pyscfad_amd.cc._ops does that).  e_out: device scalar.
 * The walk takes every unordered occupied pair (j, k) and all o triples
 * (l, j, k): about three times the W builds of pamd_ccsdt_energy, and the
 * adjoint builds each W a second time.  Cotangents as pamd_ccsdt_energy_vjp,
 * plus ulo_bar (device, nfrag x nocc, overwritten) = ulo (mat_bar + mat_bar^T).
 * Workspace: pamd_ccsdt_projected_workspace_size(cfg, nfrag, with_vjp, ...). */
int pamd_ccsdt_projected_workspace_size(const pamd_ccsdt_config *cfg, int nfrag, int with_vjp,
                                        size_t *bytes);
int pamd_ccsdt_projected_energy(const pamd_ccsdt_config *cfg,
                                const double *mo_energy, const double *t1, const double *t2,
                                const double *ovvv, const double *ovoo, const double *ovov,
                                const double *fvo, const double *ulo, int nfrag,
                                void *workspace, size_t workspace_bytes, double *e_out,
                                void *stream);
int pamd_ccsdt_projected_energy_vjp(const pamd_ccsdt_config *cfg,
                                    const double *mo_energy, const double *t1, const double *t2,
                                    const double *ovvv, const double *ovoo, const double *ovov,
                                    const double *fvo, const double *ulo, int nfrag,
                                    double et_bar, void *workspace, size_t workspace_bytes,
                                    double *e_out, double *mo_energy_bar, double *t1_bar,
                                    double *t2_bar, double *ovvv_bar, double *ovoo_bar,
                                    double *ovov_bar, double *fvo_bar, double *ulo_bar,
                                    void *stream);

/* Library version string (static storage). */
const char *pamd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PYSCFAD_AMD_CCSDT_H */
