/*
 * Host-only inspection entry points of libpyscfad_amd_cc.so.  They run no GPU
 * work and have no counterpart in the reference: they let tests check, on the
 * CPU, every address the contraction engine will touch before a kernel runs.
 */
#ifndef PYSCFAD_AMD_DEBUG_H
#define PYSCFAD_AMD_DEBUG_H

#ifdef __cplusplus
extern "C" {
#endif

/* Record (do not launch) the contraction descriptors of one occupied triple
 * i <= j <= k: the three W-build GEMMs, plus the adjoint GEMMs when with_vjp.
 * bases[10]: base addresses to put in the descriptors for
 *   Vp, T2m, OOt, W, t2, Vb, Vpb, T2mb, part_p, part_h.
 * out receives min(n, max_desc) descriptors of *desc_bytes bytes each. */
int pamd_debug_gemm_descriptors(int nocc, int nvir, int i, int j, int k, int with_vjp,
                                const unsigned long long *bases, void *out, int max_desc,
                                int *n_desc, int *desc_bytes);

/* Record the adjoint contractions of two consecutive triples (i, j, kp) and
 * (i, j, k), kp < k, as the two-stream adjoint walk pairs them in the split
 * forms: the first triple's launches without its ovvv and hole-term t2
 * adjoints into index i and j, then the second triple's, those adjoints as
 * four-segment launches that also take the first triple's terms.
 * bases[11]: those of pamd_debug_gemm_descriptors, then the second triple's
 * Wbar (the first triple's Wbar is at bases[3]). */
int pamd_debug_paired_adjoint_descriptors(int nocc, int nvir, int i, int j, int kp, int k,
                                          const unsigned long long *bases, void *out,
                                          int max_desc, int *n_desc, int *desc_bytes);

/* 1 when a recorded descriptor (one of pamd_debug_gemm_descriptors' outputs)
 * runs on the FP64 engine's direct-to-LDS path (16-byte element pairs: fast
 * strides 1, other strides and index divisors even, 16-byte aligned bases),
 * 0 when it runs on the register-staged path. */
int pamd_debug_gemm_direct_ok(const void *desc);

/* 1 when a recorded descriptor, moved onto planar split-BF16x3 operands (each
 * FP64 element address 8 e becomes the bf16 address 2 e, bases being multiples
 * of 64 bytes, with plane stride plane_stride), runs on the split-BF16x3
 * engine's direct-to-LDS path (16-byte runs of 8 bf16: fast strides 1, other
 * strides, divisors, segment lengths or row counts multiples of 8), 0 when it
 * runs on the register-staged path. */
int pamd_debug_gemm_b3p_ok(const void *desc, long long plane_stride);

/* Split-K factors the adjoint uses for (particle, hole) partial buffers. */
int pamd_debug_split_factors(int nocc, int nvir, int *S_p, int *S_h);

/* The same for a given precision (PAMD_PREC_*): the split-BF16x3 engine sizes
 * its split-K for its own tile shapes and bounds the K range of one part. */
int pamd_debug_split_factors_prec(int nocc, int nvir, int precision, int *S_p, int *S_h);

#ifdef __cplusplus
}
#endif
#endif
