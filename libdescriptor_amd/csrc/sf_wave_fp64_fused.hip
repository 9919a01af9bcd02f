// The FP64 instances of the fused forward + adjoint wave kernel (sf_wave.hip),
// in a compilation unit of their own so that they take the iterative ILP
// scheduler (build.py), which overlaps more of their long dependent chains of
// FP64 library sequences at 2 waves per SIMD.  (The other instances keep the
// default scheduler: under it the FP32-bulk kernels need fewer registers and
// no scratch, and the separate FP64 kernels are as fast or faster.)
// sf_wave.hip defines launch_fused_wave_fp64 when LDSF_WAVE_FP64_UNIT is 2,
// and nothing else with external linkage.
#define LDSF_WAVE_FP64_UNIT 2
#include "sf_wave.hip"
