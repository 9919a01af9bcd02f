// Launchers of the HIP kernels (host-callable).
#ifndef LDSF_SF_KERNELS_H_
#define LDSF_SF_KERNELS_H_

#include <hip/hip_runtime.h>

#include "ldsf_internal.h"

namespace ldsf {

// sf_general.hip: one thread per centre, any parameter set
hipError_t launch_fwd_general(const SFTab &tab, int precision, int n_centers, const int *centers,
                              const int *offs, const int *nl, const int *species, const double *x,
                              double *desc, hipStream_t s);
hipError_t launch_adj_general(const SFTab &tab, int precision, int n_centers, const int *centers,
                              const int *offs, const int *nl, const int *species, const double *x,
                              const double *dE, double *edge, double *self, hipStream_t s);

// sf_wave.hip: one wavefront per centre, neighbours staged in LDS
struct WaveShape {
    int n2;      // two-body rows, all g2
    int n3;      // three-body rows, all g4
    int nsb;     // species blocks of the two-body part (1 or n_species)
    int npb;     // pair blocks of the three-body part (1 or n_pairs)
};
// true if the wave kernels cover this parameter set
bool wave_supported(const WaveShape &shape, int n_species);
hipError_t launch_fwd_wave(const SFTab &tab, const WaveShape &shape, int precision, int n_centers,
                           const int *centers, const int *offs, const int *nl, const int *species,
                           const double *x, double *desc, int max_nn, hipStream_t s);
hipError_t launch_adj_wave(const SFTab &tab, const WaveShape &shape, int precision, int n_centers,
                           const int *centers, const int *offs, const int *nl, const int *species,
                           const double *x, const double *dE, double *edge, double *self,
                           int max_nn, hipStream_t s);

// sf_plan.hip: neighbour-list plan building and the adjoint gather
hipError_t launch_plan_validate(int n_centers, const int *centers, int n_total, const int *numneigh,
                                const int *species, int n_species, int *flags, hipStream_t s);
hipError_t launch_edge_targets_count(long n_edges, int n_centers, const int *nl, const int *centers,
                                     int n_total, int *count, int *flags, hipStream_t s);
hipError_t launch_edge_targets_fill(long n_edges, int n_centers, const int *nl, const int *centers,
                                    const int *in_start, int *cursor, int *in_edges, hipStream_t s);
hipError_t launch_sort_buckets(int n_total, const int *in_start, int *in_edges, hipStream_t s);
hipError_t launch_gather(int n_total, const int *in_start, const int *in_edges, long n_edges,
                         const double *edge, const double *self, double *dcoords, hipStream_t s);
hipError_t launch_max_int(int n, const int *v, int *out, hipStream_t s);
hipError_t launch_scan_int(int n, const int *in, int *out_exclusive_plus_total, void *tmp,
                           size_t *tmp_bytes, hipStream_t s);

}  // namespace ldsf

#endif  // LDSF_SF_KERNELS_H_
