// The FP64 instances of the fused forward + adjoint wave kernel (sf_wave.hip),
// in a compilation unit of their own.  Like the FP32-bulk kernels they take
// the iterative ILP scheduler (build.py), which overlaps more of their long
// dependent chains at 2 waves per SIMD; the separate FP64 forward and adjoint
// (sf_wave_fp64.hip) keep the default scheduler, under which they are faster.
// sf_wave.hip defines launch_fused_wave_fp64 when LDSF_WAVE_FP64_UNIT is 2,
// and nothing else with external linkage.
#define LDSF_WAVE_FP64_UNIT 2
#include "sf_wave.hip"
