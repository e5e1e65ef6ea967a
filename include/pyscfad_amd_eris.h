/*
 * C ABI of the MP2 amplitude/pair-energy kernel and the density-fitted (DF)
 * integral producer for the (T) energy (libpyscfad_amd_cc.so).
 *
 * All arrays are device pointers to C-contiguous float64 arrays; `stream` is a
 * hipStream_t (NULL = default stream).  Calls queue their work and return
 * without synchronising.  Status codes are those of pyscfad_amd_ccsdt.h.
 */
#ifndef PYSCFAD_AMD_ERIS_H
#define PYSCFAD_AMD_ERIS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MP2 amplitudes and occupied-pair MP2 energies, as CCSD.init_amps forms them
 * (pyscfad/cc/ccsd.py:364-381):
 *   t2[i,j,a,b] = ovov[i,a,j,b] / (e_i + e_j - e_a - e_b)      (e = mo_energy)
 *   e_pair[i*o+j] = 2 sum_ab t2[i,j,a,b] ovov[i,a,j,b]
 *                   - sum_ab t2[j,i,a,b] ovov[i,a,j,b]
 * so that sum(e_pair) is the reference's emp2 (ccsd.py:380-381).
 * mo_energy (o+v), ovov (o,v,o,v); t2_out (o,o,v,v) is written when not NULL;
 * e_pair (o*o) is overwritten.  Every e_pair element is reduced in a fixed
 * order (bitwise reproducible). */
int pamd_mp2_init_amps(int nocc, int nvir, const double *mo_energy, const double *ovov,
                       double *t2_out, double *e_pair, void *stream);

/* Device bytes of the workspace of pamd_df_eris (with_vjp = 0) or of
 * pamd_df_eris_vjp (with_vjp = 1): the packed virtual-virtual factors Lvv, and
 * for the adjoint also their cotangent. */
int pamd_df_eris_workspace_size(int nocc, int nmo, int naux, int with_vjp, size_t *bytes);

/* The (T) blocks of the in-core DF integrals, _make_df_eris_incore
 * (pyscfad/cc/dfccsd.py:62-89), from the MO-basis factors Lpq (naux, nmo, nmo):
 *   Loo = Lpq[:, :o, :o], Lov = Lpq[:, :o, o:], Lvv = pack_tril(Lpq[:, o:, o:])
 *   ovvv = Lov^T Lvv   -> (o, v, v(v+1)/2)   packed lower triangle, b >= c
 *   ovoo = Lov^T Loo   -> (o, v, o, o)
 *   ovov = Lov^T Lov   -> (o, v, o, v)
 * Any output pointer may be NULL (that block is skipped).  The products run
 * on the FP64 MFMA contraction engine. */
int pamd_df_eris(int nocc, int nmo, int naux, const double *Lpq, double *ovvv, double *ovoo,
                 double *ovov, void *workspace, size_t workspace_bytes, void *stream);

/* Reverse mode of pamd_df_eris: the cotangent of Lpq from those of ovvv
 * (packed, as pamd_ccsdt_energy_vjp writes it for packed input), ovoo and ovov.
 * Lpq_bar (naux, nmo, nmo) is overwritten.  Its layout follows the reference's
 * adjoints: np.dot for the three products, and pack_tril's adjoint for Lvv,
 * which puts the packed cotangent on the lower triangle b >= c and zeros above
 * it (pyscfad/lib/_numpy_helper_opt.py:70-95, NPdpack_tril_vjp in
 * pyscfadlib/pyscfadlib/vjp/np_helper/pack_tril_vjp.c:62-70).  The vo block
 * Lpq_bar[:, o:, :o] is zero.  A NULL cotangent counts as zero. */
int pamd_df_eris_vjp(int nocc, int nmo, int naux, const double *Lpq, const double *ovvv_bar,
                     const double *ovoo_bar, const double *ovov_bar, double *Lpq_bar,
                     void *workspace, size_t workspace_bytes, void *stream);

/* Host-only: record (do not launch) the contraction descriptors of
 * pamd_df_eris (with_vjp = 0) or pamd_df_eris_vjp (with_vjp = 1).  block_i > 0
 * forces the occupied block of the products that read the ovvv cotangent
 * (0 = the library's choice, which keeps every operand offset below 2^31).
 * bases[7]: addresses to put in them for Lpq, Lvv, Lvv_bar, ovvv (or its
 * cotangent), ovoo (or cotangent), ovov (or cotangent), Lpq_bar.
 * out receives min(n, max_desc) descriptors of *desc_bytes bytes each.  The
 * packing of Lvv and the scatter of its cotangent are not GEMMs and are not
 * recorded.  No GPU work. */
int pamd_debug_df_descriptors(int nocc, int nmo, int naux, int with_vjp, int block_i,
                              const unsigned long long *bases, void *out, int max_desc,
                              int *n_desc, int *desc_bytes);

#ifdef __cplusplus
}
#endif
#endif
