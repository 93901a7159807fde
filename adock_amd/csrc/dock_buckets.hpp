// Entry points of the chain kernels per torsion bucket KM (defined by ADOCK_BUCKET_DEFINE in dock_km<KM>.hip);
// dock.hip dispatches on KM.
#pragma once
#include "kernels.hpp"

// every compiled bucket: KM = 8 .. 56 torsions in steps of 8 (state lanes 7 + KM of one wave)
#define ADOCK_KM_LIST(X) X(8) X(16) X(24) X(32) X(40) X(48) X(56)

namespace adock {
#define ADOCK_BUCKET_DECL(KM)                                                                                          \
    int launch_mc_km##KM(const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*, int,         \
                         const McArgs&, hipStream_t);                                                                  \
    int launch_eg_km##KM(const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*, int,         \
                         const float*, float*, float*, hipStream_t);                                                   \
    int launch_mc_persistent_km##KM(const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*,   \
                                    int, const McArgs&, const QueueArgs&, int, hipStream_t);                           \
    int launch_mc_step_km##KM(const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, const int*, int,        \
                              const McArgs&, const StepArgs&, hipStream_t);                                            \
    int launch_mc_learned_km##KM(const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, const int*, int,     \
                                 const McArgs&, const StepArgs&, const NetArgs&, const QueueArgs&, hipStream_t);       \
    int mc_resident_waves_km##KM(int P_MAX, int persistent, int APL);
ADOCK_KM_LIST(ADOCK_BUCKET_DECL)
#undef ADOCK_BUCKET_DECL
}  // namespace adock
