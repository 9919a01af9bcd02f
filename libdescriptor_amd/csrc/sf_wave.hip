// Wave-per-centre kernels (placeholder until the LDS-staged kernels land).
#include <hip/hip_runtime.h>

#include "sf_kernels.h"

namespace ldsf {

bool wave_supported(const WaveShape &, int) { return false; }

hipError_t launch_fwd_wave(const SFTab &, const WaveShape &, int, int, const int *, const int *,
                           const int *, const int *, const double *, double *, int, hipStream_t)
{
    return hipErrorNotSupported;
}

hipError_t launch_adj_wave(const SFTab &, const WaveShape &, int, int, const int *, const int *,
                           const int *, const int *, const double *, const double *, double *,
                           double *, int, hipStream_t)
{
    return hipErrorNotSupported;
}

}  // namespace ldsf
