/*
 * Sampled kernel timing of libpyscfad_amd_cc.so (no counterpart in the
 * reference, whose only timing is the host logger: pyscfad/lib/logger.py:63-79).
 * When enabled, every launch of one occupied triple in `every` is bracketed by
 * hipEvents on the stream it runs on; per-kernel totals accumulate until reset.
 * The switches and the totals are process-wide (the totals are updated and read
 * under a lock); which triple is sampled, and its pending events, are per host
 * thread, so concurrent calls from several threads can be profiled together.
 */
#ifndef PYSCFAD_AMD_PROFILE_H
#define PYSCFAD_AMD_PROFILE_H

#ifdef __cplusplus
extern "C" {
#endif

int pamd_profile_enable(int every);   /* 0 disables */
/* Progress line on stderr every `every` occupied triples (0 disables); each
 * report first synchronises the stream, so it reflects finished GPU work. */
int pamd_profile_progress(int every);
int pamd_profile_reset(void);
int pamd_profile_num_kinds(void);
const char *pamd_profile_kind_name(int kind);
/* launches timed, their summed duration (ms), algorithmic flops and bytes */
int pamd_profile_read(int kind, long long *launches, double *ms, double *flops, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
