// Launchers of eris_kernels.hip (MP2 pair energies, DF integral producer).
#pragma once
#include <hip/hip_runtime.h>

namespace pamd {

hipError_t launch_mp2_pairs(int o, int v, const double *mo, const double *ovov, double *t2,
                            double *e_pair, hipStream_t s);
hipError_t launch_pack_lvv(const double *Lpq, int o, int n, int naux, double *Lvv,
                           hipStream_t s);
hipError_t launch_unpack_lvv_bar(const double *Lvv_bar, int o, int n, int naux, double *Lpq_bar,
                                 hipStream_t s);
hipError_t launch_zero_occ_rows(int o, int n, int naux, double *Lpq_bar, hipStream_t s);

}  // namespace pamd
