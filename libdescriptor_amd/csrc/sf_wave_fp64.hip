// The FP64 forward and adjoint instances of the wave kernels (sf_wave.hip), in
// a compilation unit of their own: they keep the compiler's default scheduler
// (build.py).  sf_wave.hip defines launch_fwd_wave_fp64 and
// launch_adj_wave_fp64 when LDSF_WAVE_FP64_UNIT is 1, and nothing else with
// external linkage.
#define LDSF_WAVE_FP64_UNIT 1
#include "sf_wave.hip"
