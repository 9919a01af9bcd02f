// Neighbour-list plans and the atomic-free adjoint gather.
//
// The adjoint kernels write, for every edge e of the CSR neighbour list,
// the 3-vector dE/d r_{nl[e]} contributed by that edge's centre, and for
// every centre t the vector dE/d r_{centre t}.  Summing these per atom needs
// the transposed list: for each coordinate row a, the ids of the edges that
// point at it.  It is built once per plan with an integer count, a prefix
// sum and a scatter whose buckets are then sorted, so the gather visits
// every row's edges in ascending id order and the result is bitwise the same
// on every run and in every batch.  Centre self-terms are edges too (ids
// n_edges + t), so one gather writes d_coordinates, padding rows included
// (the layout the reference's gradient returns, python_bindings.cpp:235,252).
// The adjoint kernels store each vector at its position in that transposed
// order (inv_pos), so the gather reads every row's records contiguously.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "sf_bucket.h"
#include "sf_device.h"
#include "sf_kernels.h"

namespace ldsf {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// flags: 1 neighbour row out of range, 2 negative neighbour count,
//        4 species code out of range, 8 centre row out of range
__global__ void k_validate(int n_centers, const int *__restrict__ centers, int n_total,
                           const int *__restrict__ numneigh, const int *__restrict__ species,
                           int n_species, int *flags)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_centers) {
        if (numneigh[t] < 0) atomicOr(flags, 2);
        int c = centers ? centers[t] : (int)t;
        if (c < 0 || c >= n_total) atomicOr(flags, 8);
    }
    if (t < n_total) {
        int s = species[t];
        if (s < 0 || s >= n_species) atomicOr(flags, 4);
    }
}

__device__ __forceinline__ int edge_target(long e, long n_edges, const int *nl, const int *centers)
{
    if (e < n_edges) return nl[e];
    int t = (int)(e - n_edges);
    return centers ? centers[t] : t;
}

// ---- the transposed list by buckets of rows (sf_bucket.h) ----
// A tile is kPlanTile consecutive entries (edges, then the centres'
// self-terms); each counts its entries per bucket in LDS, so that a plan's
// random targets cost one global atomic per tile and bucket, not per entry.
constexpr int kPlanPer = 64;                   // entries per thread and tile
constexpr long kPlanTile = (long)kBlock * kPlanPer;

// the tile's valid entries per bucket in hist[0..nb) (zeroed first); an
// entry whose target is outside [0, n_total) sets flags bit 1 and is skipped
__device__ __forceinline__ void plan_tile_hist(int *hist, int nb, long base, long n_ent,
                                               long n_edges, const int *nl, const int *centers,
                                               int n_total, int *flags)
{
    for (int j = threadIdx.x; j < nb; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    int bad = 0;
    for (int r = 0; r < kPlanPer; r += 8) {
        int tg[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const long i = base + (long)(r + k) * kBlock + threadIdx.x;
            tg[k] = i < n_ent ? edge_target(i, n_edges, nl, centers) : -1;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const long i = base + (long)(r + k) * kBlock + threadIdx.x;
            if (i >= n_ent) continue;
            if (tg[k] < 0 || tg[k] >= n_total) bad = 1;
            else atomicAdd(&hist[tg[k] >> kBucketShift], 1);
        }
    }
    if (bad) atomicOr(flags, 1);
    __syncthreads();
}

__global__ void k_bucket_count(long n_ent, long n_edges, const int *__restrict__ nl,
                               const int *__restrict__ centers, int n_total, int nb, int *bcount,
                               int *flags)
{
    __shared__ int hist[kHistCap];
    const long ntiles = (n_ent + kPlanTile - 1) / kPlanTile;
    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        plan_tile_hist(hist, nb, t * kPlanTile, n_ent, n_edges, nl, centers, n_total, flags);
        for (int j = threadIdx.x; j < nb; j += blockDim.x)
            if (hist[j]) atomicAdd(bcount + j, hist[j]);
        __syncthreads();  // hist is the next tile's
    }
}

// every valid entry's (id, row) into its bucket's range of stage: a run per
// tile and bucket, claimed with one atomic, its slots taken in any order
__global__ void k_bucket_scatter(long n_ent, long n_edges, const int *__restrict__ nl,
                                 const int *__restrict__ centers, int n_total, int nb,
                                 const int *__restrict__ bbase, int *bcursor, int2 *stage,
                                 int *flags)
{
    __shared__ int hist[kHistCap];
    const long ntiles = (n_ent + kPlanTile - 1) / kPlanTile;
    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const long base = t * kPlanTile;
        plan_tile_hist(hist, nb, base, n_ent, n_edges, nl, centers, n_total, flags);
        for (int j = threadIdx.x; j < nb; j += blockDim.x)
            if (hist[j]) hist[j] = bbase[j] + atomicAdd(bcursor + j, hist[j]);
        __syncthreads();
        for (int r = 0; r < kPlanPer; r += 8) {
            int tg[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const long i = base + (long)(r + k) * kBlock + threadIdx.x;
                tg[k] = i < n_ent ? edge_target(i, n_edges, nl, centers) : -1;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const long i = base + (long)(r + k) * kBlock + threadIdx.x;
                if (i >= n_ent || tg[k] < 0 || tg[k] >= n_total) continue;
                const int slot = atomicAdd(&hist[tg[k] >> kBucketShift], 1);
                stage[slot] = make_int2((int)i, tg[k]);
            }
        }
        __syncthreads();  // hist is the next tile's
    }
}

__global__ __launch_bounds__(kBucketThreads) void k_bucket_rows(int nb, int n_total,
                                                               const int2 *__restrict__ stage,
                                                               const int *__restrict__ bbase,
                                                               int *in_start, int *in_edges,
                                                               int *inv_pos)
{
    __shared__ BucketLds sh;
    for (int b = blockIdx.x; b < nb; b += gridDim.x)
        bucket_pass(b, nb, n_total, stage, bbase, in_start, in_edges, inv_pos, sh);
}

__global__ void k_count(long n_edges, int n_centers, const int *__restrict__ nl,
                        const int *__restrict__ centers, int n_total, int *count, int *flags)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_edges + n_centers) return;
    int a = edge_target(e, n_edges, nl, centers);
    if (a < 0 || a >= n_total) {
        atomicOr(flags, 1);
        return;
    }
    atomicAdd(count + a, 1);
}

__global__ void k_fill(long n_edges, int n_centers, const int *__restrict__ nl,
                       const int *__restrict__ centers, int n_total,
                       const int *__restrict__ in_start, int *cursor, int *in_edges)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_edges + n_centers) return;
    int a = edge_target(e, n_edges, nl, centers);
    if (a < 0 || a >= n_total) return;  // flagged by k_count; the plan is refused
    int pos = atomicAdd(cursor + a, 1);
    in_edges[in_start[a] + pos] = (int)e;
}

// each row's bucket (a few tens of ids) sorted by one wavefront
// every row's bucket sorted, and inv_pos[edge id] = its position
__global__ void k_sort_inverse(int n_total, const int *__restrict__ in_start, int *in_edges,
                               int *inv_pos)
{
    const long nw = ((long)gridDim.x * blockDim.x) >> 6;
    sort_rows_inverse(n_total, in_start, in_edges, inv_pos,
                      ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw, threadIdx.x & 63);
}

// records of row a are contiguous: [in_start[a], in_start[a+1]), in ascending
// edge id, written there by the adjoint kernels; nothing is added when *skip
// is non-zero (skip may be NULL)
__global__ void k_gather(int n_total, const int *__restrict__ in_start,
                         const double *__restrict__ edge, double *__restrict__ dcoords,
                         const int *__restrict__ skip, int assign)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_total || (skip && *skip)) return;
    int b = in_start[a], e = in_start[a + 1];
    if (b == e && !assign) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    row_sum3(edge, b, e, s0, s1, s2);
    double *o = dcoords + 3 * (long)a;
    if (assign) {  // = the row sum, bitwise the += into zeros
        o[0] = s0;
        o[1] = s1;
        o[2] = s2;
    } else {
        o[0] += s0;
        o[1] += s1;
        o[2] += s2;
    }
}

__global__ void k_max(int n, const int *__restrict__ v, int *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int m = 0;
    for (; i < n; i += gridDim.x * blockDim.x) m = max(m, v[i]);
    // wave reduction, then the block's waves through LDS: one atomic per block
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    __shared__ int wm[kBlock / 64];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = max(m, wm[k]);
        atomicMax(out, m);
    }
}

}  // namespace

hipError_t launch_plan_validate(int n_centers, const int *centers, int n_total, const int *numneigh,
                                const int *species, int n_species, int *flags, hipStream_t s)
{
    long n = n_centers > n_total ? n_centers : n_total;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_validate, grid_for(n), dim3(kBlock), 0, s, n_centers, centers, n_total,
                       numneigh, species, n_species, flags);
    return hipGetLastError();
}

hipError_t launch_edge_targets_count(long n_edges, int n_centers, const int *nl, const int *centers,
                                     int n_total, int *count, int *flags, hipStream_t s)
{
    long n = n_edges + n_centers;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_count, grid_for(n), dim3(kBlock), 0, s, n_edges, n_centers, nl, centers,
                       n_total, count, flags);
    return hipGetLastError();
}

hipError_t launch_edge_targets_fill(long n_edges, int n_centers, const int *nl, const int *centers,
                                    int n_total, const int *in_start, int *cursor, int *in_edges,
                                    hipStream_t s)
{
    long n = n_edges + n_centers;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, s, n_edges, n_centers, nl, centers,
                       n_total, in_start, cursor, in_edges);
    return hipGetLastError();
}

hipError_t launch_sort_inverse(int n_total, const int *in_start, int *in_edges, int *inv_pos,
                               hipStream_t s)
{
    if (n_total <= 0) return hipSuccess;
    // a wave per two rows, at most 8192 blocks of 4 waves
    const long blocks = std::min<long>(((long)n_total * 32 + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_sort_inverse, dim3((unsigned)blocks), dim3(kBlock), 0, s, n_total,
                       in_start, in_edges, inv_pos);
    return hipGetLastError();
}

// out[e] = nl[e] | species[nl[e]] << kNlSpeciesShift: the wave kernels read a
// neighbour's species from its list entry instead of gathering it (indices
// validated before, in [0, n_total) with n_total <= kNlIndexMask + 1)
__global__ void k_pack_species(long n_edges, const int *__restrict__ nl,
                               const int *__restrict__ species, int *__restrict__ out)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n_edges;
         e += (long)gridDim.x * blockDim.x) {
        const int j = nl[e];
        out[e] = j | (species[j] << kNlSpeciesShift);
    }
}

hipError_t launch_pack_species(long n_edges, const int *nl, const int *species, int *out,
                               hipStream_t s)
{
    if (n_edges <= 0) return hipSuccess;
    const long blocks = std::min<long>((n_edges + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_pack_species, dim3((unsigned)blocks), dim3(kBlock), 0, s, n_edges, nl,
                       species, out);
    return hipGetLastError();
}

__global__ void k_zero_f64(long n, double *__restrict__ v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v[i] = 0.0;
}

hipError_t launch_zero_f64(long n, double *v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const long blocks = std::min<long>((n + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)blocks), dim3(kBlock), 0, s, n, v);
    return hipGetLastError();
}

hipError_t launch_gather(int n_total, const int *in_start, const double *edge, double *dcoords,
                         hipStream_t s, const int *skip, bool assign)
{
    if (n_total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, grid_for(n_total), dim3(kBlock), 0, s, n_total, in_start, edge,
                       dcoords, skip, assign ? 1 : 0);
    return hipGetLastError();
}

int bucket_count_of(int n_total) { return (n_total + kBucketRows - 1) >> kBucketShift; }

bool bucketed_fits(int n_total) { return bucket_count_of(n_total) <= kHistCap; }

hipError_t launch_transpose_bucketed(long n_edges, int n_centers, const int *nl,
                                     const int *centers, int n_total, int *bcount, int *bbase,
                                     int *bcursor, int2 *stage, int *in_start, int *in_edges,
                                     int *inv_pos, int *flags, void *tmp, size_t tmp_bytes,
                                     hipStream_t s)
{
    const int nb = bucket_count_of(n_total);
    if (nb > kHistCap) return hipErrorInvalidValue;
    const long n_ent = n_edges + n_centers;
    hipError_t e;
    if (n_total <= 0) return hipSuccess;
    if ((e = hipMemsetAsync(bcount, 0, sizeof(int) * (nb + 1), s))) return e;
    if ((e = hipMemsetAsync(bcursor, 0, sizeof(int) * nb, s))) return e;
    const long ntiles = (n_ent + kPlanTile - 1) / kPlanTile;
    const unsigned tg = (unsigned)std::max<long>(1, std::min<long>(ntiles, 2048));
    if (n_ent > 0)
        hipLaunchKernelGGL(k_bucket_count, dim3(tg), dim3(kBlock), 0, s, n_ent, n_edges, nl,
                           centers, n_total, nb, bcount, flags);
    if ((e = hipGetLastError())) return e;
    size_t tb = tmp_bytes;
    if ((e = launch_scan_int(nb, bcount, bbase, tmp, &tb, s))) return e;
    if (n_ent > 0)
        hipLaunchKernelGGL(k_bucket_scatter, dim3(tg), dim3(kBlock), 0, s, n_ent, n_edges, nl,
                           centers, n_total, nb, bbase, bcursor, stage, flags);
    hipLaunchKernelGGL(k_bucket_rows, dim3((unsigned)std::min(nb, 4096)), dim3(kBucketThreads), 0,
                       s, nb, n_total, stage, bbase, in_start, in_edges, inv_pos);
    return hipGetLastError();
}

hipError_t launch_max_int(int n, const int *v, int *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_max, dim3((unsigned)blocks), dim3(kBlock), 0, s, n, v, out);
    return hipGetLastError();
}

// out[0] = 0, out[k] = in[0] + ... + in[k-1] for k = 1..n.  With tmp == NULL
// only reports the scratch size in *tmp_bytes.
hipError_t launch_scan_int(int n, const int *in, int *out, void *tmp, size_t *tmp_bytes,
                           hipStream_t s)
{
    if (!tmp) {
        *tmp_bytes = 0;
        if (n <= 0) return hipSuccess;
        return hipcub::DeviceScan::InclusiveSum(nullptr, *tmp_bytes, in, out + 1, n, s);
    }
    hipError_t err = hipMemsetAsync(out, 0, sizeof(int), s);
    if (err != hipSuccess || n <= 0) return err;
    return hipcub::DeviceScan::InclusiveSum(tmp, *tmp_bytes, in, out + 1, n, s);
}

}  // namespace ldsf
