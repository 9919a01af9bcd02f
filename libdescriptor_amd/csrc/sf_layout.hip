// Device-resident input-layout producer of the SF path: periodic padding
// atoms, the full neighbour list of the contributing atoms in KLIFF order and
// the flat int32 numneigh/neighlist layout, all built on the GPU from device
// coordinates; and the fold of padding-row gradients back onto their
// contributing atoms.
//
// Reference behaviour reproduced (same values, same order):
//   nbl_create_paddings  python_bindings/neighbor/neighbor_list.cpp:283-418
//   nbl_build            neighbor_list.cpp:78-243 (cell grid from a min/max box
//                        whose initial max is coord0 + 1, helper.hpp:96-108)
//   flattening           python_package/libdescriptor/neighbor.py:108-154
//   assemble_forces      neighbor.py:227-263 (both of its summation orders)
//
// Padding: the images (i, j, k) are enumerated on the host in the reference's
// order (x shift outermost); one flag per (image, atom) and an exclusive scan
// give every padding atom its output row, so the rows come out image-major,
// atoms ascending, as the reference's push_back loop writes them.
// Neighbour list: atoms are radix-sorted by cell (stable, so ascending index
// within a cell); one thread per centre walks the 3x3x3 block x-outermost and
// each cell's atoms in order, counting in a first pass and writing at its
// scanned offset in a second.  Threads take centres in cell order, so the
// threads of a wavefront read the same few cells.  Every distance and
// coordinate is the reference's FP64 expression without FMA contraction, so
// membership and padding coordinates are bitwise those of the host producers.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ldsf.h"
#include "ldsf_internal.h"
#include "sf_device.h"
#include "sf_kernels.h"

namespace {

constexpr double kTol = 1.0e-10;  // neighbor_list.cpp:14, helper.hpp:10
constexpr int kBlock = 256;

int hip_fail(hipError_t e, const char *where)
{
    return ldsf::fail(e == hipErrorOutOfMemory ? LDSF_ERR_NOMEM : LDSF_ERR_DEVICE,
                      std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, where)                                \
    do {                                                    \
        hipError_t e__ = (expr);                            \
        if (e__ != hipSuccess) return hip_fail(e__, where); \
    } while (0)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    hipError_t alloc(size_t count)
    {
        release();
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, sizeof(T) * count);
    }
    T *take()
    {
        T *q = p;
        p = nullptr;
        return q;
    }
};

inline unsigned grid_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---- host 3x3 helpers, in the reference's operation order (helper.hpp:19-93) ----

inline double hdot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

inline void hcross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

inline double hnorm(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

inline double det2(double a11, double a12, double a21, double a22) { return (a11 * a22) - (a12 * a21); }

bool hinvert(const double *m, double *inv)
{
    inv[0] = det2(m[4], m[5], m[7], m[8]);
    inv[1] = det2(m[2], m[1], m[8], m[7]);
    inv[2] = det2(m[1], m[2], m[4], m[5]);
    inv[3] = det2(m[5], m[3], m[8], m[6]);
    inv[4] = det2(m[0], m[2], m[6], m[8]);
    inv[5] = det2(m[2], m[0], m[5], m[3]);
    inv[6] = det2(m[3], m[4], m[6], m[7]);
    inv[7] = det2(m[1], m[0], m[7], m[6]);
    inv[8] = det2(m[0], m[1], m[3], m[4]);
    double dd = m[0] * m[4] * m[8] - m[0] * m[5] * m[7] - m[1] * m[3] * m[8] + m[1] * m[5] * m[6]
                + m[2] * m[3] * m[7] - m[2] * m[4] * m[6];
    if (std::abs(dd) < kTol) return false;
    for (int i = 0; i < 9; ++i) inv[i] /= dd;
    return true;
}

// ---- device helpers ----

struct Mat3 {
    double m[9];
};

__device__ __forceinline__ double ddot(const double *a, double b0, double b1, double b2)
{
#pragma clang fp contract(off)
    return a[0] * b0 + a[1] * b1 + a[2] * b2;
}

__device__ __forceinline__ double drsq(double dx, double dy, double dz)
{
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// min / max of three columns: per block, then over the blocks (exact, so the
// order of the reduction does not matter)
struct Box {
    double lo[3], hi[3];
};

__device__ __forceinline__ void box_reduce_block(Box &b)
{
    __shared__ double sh[6][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int off = 32; off > 0; off >>= 1) {
            b.lo[d] = fmin(b.lo[d], __shfl_xor(b.lo[d], off));
            b.hi[d] = fmax(b.hi[d], __shfl_xor(b.hi[d], off));
        }
    }
    if (lane == 0)
        for (int d = 0; d < 3; ++d) {
            sh[d][w] = b.lo[d];
            sh[3 + d][w] = b.hi[d];
        }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < kBlock / 64; ++k)
            for (int d = 0; d < 3; ++d) {
                b.lo[d] = fmin(b.lo[d], sh[d][k]);
                b.hi[d] = fmax(b.hi[d], sh[3 + d][k]);
            }
}

// fractional coordinates frac = fcell x (neighbor_list.cpp:308-326) and their
// per-block extent
__global__ void k_frac(int n, Mat3 fcell, const double *__restrict__ x, double *__restrict__ frac,
                       Box *partial)
{
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = 1e10;
        b.hi[d] = -1e10;
    }
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
        for (int d = 0; d < 3; ++d) {
            const double f = ddot(fcell.m + 3 * d, x0, x1, x2);
            frac[3 * i + d] = f;
            b.lo[d] = fmin(b.lo[d], f);
            b.hi[d] = fmax(b.hi[d], f);
        }
    }
    box_reduce_block(b);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// extent of the coordinates themselves, for the cell grid (neighbor_list.cpp:87-109)
__global__ void k_extent(int n, const double *__restrict__ x, Box *partial)
{
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = x[d];
        b.hi[d] = x[d];
    }
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        for (int d = 0; d < 3; ++d) {
            const double v = x[3 * i + d];
            b.lo[d] = fmin(b.lo[d], v);
            b.hi[d] = fmax(b.hi[d], v);
        }
    box_reduce_block(b);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// one block: the threads take the partial boxes k = t, t + kBlock, ..., then
// the block reduction (min and max are exact, so the order does not matter)
__global__ void k_box_final(int nb, const Box *partial, Box *out)
{
    Box b = partial[0];
    for (int k = threadIdx.x; k < nb; k += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            b.lo[d] = fmin(b.lo[d], partial[k].lo[d]);
            b.hi[d] = fmax(b.hi[d], partial[k].hi[d]);
        }
    box_reduce_block(b);
    if (threadIdx.x == 0) *out = b;
}

struct PadParams {
    Mat3 tcell;
    double lo[3], hi[3];  // extent of the fractional coordinates, widened by kTol
    double slack[3];      // size - ratio (neighbor_list.cpp:359-361)
    int size[3];
};

// keep flag of padding candidate (image g, atom at) (neighbor_list.cpp:388-396)
__global__ void k_pad_flag(long total, int n, const int3 *__restrict__ img, PadParams pp,
                           const double *__restrict__ frac, int *__restrict__ flag)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int g = (int)(t / n), at = (int)(t - (long)g * n);
    const int3 s = img[g];
    const double x = frac[3 * (long)at], y = frac[3 * (long)at + 1], z = frac[3 * (long)at + 2];
    bool keep = true;
    if (s.x == -pp.size[0] && x - pp.lo[0] < pp.slack[0]) keep = false;
    if (s.x == pp.size[0] && pp.hi[0] - x < pp.slack[0]) keep = false;
    if (s.y == -pp.size[1] && y - pp.lo[1] < pp.slack[1]) keep = false;
    if (s.y == pp.size[1] && pp.hi[1] - y < pp.slack[1]) keep = false;
    if (s.z == -pp.size[2] && z - pp.lo[2] < pp.slack[2]) keep = false;
    if (s.z == pp.size[2] && pp.hi[2] - z < pp.slack[2]) keep = false;
    flag[t] = keep ? 1 : 0;
}

// padding atom rows n + pos[t] (neighbor_list.cpp:398-410)
__global__ void k_pad_emit(long total, int n, const int3 *__restrict__ img, Mat3 tcell,
                           const double *__restrict__ frac, const int *__restrict__ flag,
                           const int *__restrict__ pos, const int *__restrict__ species_in,
                           double *__restrict__ coords, int *__restrict__ species,
                           int *__restrict__ image)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total || !flag[t]) return;
    const int g = (int)(t / n), at = (int)(t - (long)g * n);
    const int3 s = img[g];
    double a0, a1, a2;
    {
#pragma clang fp contract(off)
        a0 = (double)s.x + frac[3 * (long)at];
        a1 = (double)s.y + frac[3 * (long)at + 1];
        a2 = (double)s.z + frac[3 * (long)at + 2];
    }
    const long r = (long)n + pos[t];
    coords[3 * r] = ddot(tcell.m, a0, a1, a2);
    coords[3 * r + 1] = ddot(tcell.m + 3, a0, a1, a2);
    coords[3 * r + 2] = ddot(tcell.m + 6, a0, a1, a2);
    species[r] = species_in[at];
    image[r] = at;
}

__global__ void k_iota(int n, int *__restrict__ v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

struct Grid {
    double lo[3], hi[3];
    int size[3];
};

// coords_to_index (helper.hpp:96-108)
__device__ __forceinline__ void cell_index(const Grid &g, double x0, double x1, double x2, int *id)
{
    const double x[3] = {x0, x1, x2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int v;
        {
#pragma clang fp contract(off)
            v = (int)(((x[d] - g.lo[d]) / (g.hi[d] - g.lo[d])) * g.size[d]);
        }
        // x >= lo, so v >= 0 for finite input; the clamp keeps a NaN row in bounds
        id[d] = max(0, min(v, g.size[d] - 1));
    }
}

__global__ void k_cell_of(int n, Grid g, const double *__restrict__ x, int *__restrict__ cell,
                          int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int id[3];
    cell_index(g, x[3 * (long)i], x[3 * (long)i + 1], x[3 * (long)i + 2], id);
    const int c = id[0] + id[1] * g.size[0] + id[2] * g.size[0] * g.size[1];
    cell[i] = c;
    atomicAdd(count + c, 1);
}

// One centre's walk over the 3x3x3 block around its cell, in the reference's
// order (neighbor_list.cpp:162-231): FILL = false counts the neighbours
// within the cutoff and reports collisions, FILL = true writes them.
template <bool FILL>
__global__ void k_neigh(int n_sorted, int n_query, Grid g, double cutsq,
                        const int *__restrict__ sorted, const int *__restrict__ cell_start,
                        const double *__restrict__ x, int *__restrict__ count,
                        const int *__restrict__ begin, int *__restrict__ nl,
                        unsigned long long *collision)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_sorted) return;
    const int i = sorted[p];
    if (i >= n_query) return;
    const double xi = x[3 * (long)i], yi = x[3 * (long)i + 1], zi = x[3 * (long)i + 2];
    int id[3];
    cell_index(g, xi, yi, zi, id);
    int c = 0;
    long out = FILL ? begin[i] : 0;
    int seen = 0;  // candidates visited, for the collision report's order
    for (int ii = max(0, id[0] - 1); ii <= min(id[0] + 1, g.size[0] - 1); ++ii)
        for (int jj = max(0, id[1] - 1); jj <= min(id[1] + 1, g.size[1] - 1); ++jj)
            for (int kk = max(0, id[2] - 1); kk <= min(id[2] + 1, g.size[2] - 1); ++kk) {
                const int cc = ii + jj * g.size[0] + kk * g.size[0] * g.size[1];
                for (int m = cell_start[cc]; m < cell_start[cc + 1]; ++m) {
                    const int nb = sorted[m];
                    if (nb == i) continue;
                    const double rsq = drsq(x[3 * (long)nb] - xi, x[3 * (long)nb + 1] - yi,
                                            x[3 * (long)nb + 2] - zi);
                    if (!FILL && rsq < kTol)
                        atomicMin(collision, ((unsigned long long)i << 32) | (unsigned)seen);
                    ++seen;
                    if (rsq < cutsq) {
                        if (FILL) nl[out++] = nb;
                        else ++c;
                    }
                }
            }
    if (!FILL) count[i] = c;
}

// the collision partner: candidate number `seen` of centre i's walk
__global__ void k_collision_partner(int i, int seen, Grid g, const int *__restrict__ sorted,
                                    const int *__restrict__ cell_start, const double *__restrict__ x,
                                    int *partner)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int id[3];
    cell_index(g, x[3 * (long)i], x[3 * (long)i + 1], x[3 * (long)i + 2], id);
    int k = 0;
    for (int ii = max(0, id[0] - 1); ii <= min(id[0] + 1, g.size[0] - 1); ++ii)
        for (int jj = max(0, id[1] - 1); jj <= min(id[1] + 1, g.size[1] - 1); ++jj)
            for (int kk = max(0, id[2] - 1); kk <= min(id[2] + 1, g.size[2] - 1); ++kk) {
                const int cc = ii + jj * g.size[0] + kk * g.size[0] * g.size[1];
                for (int m = cell_start[cc]; m < cell_start[cc + 1]; ++m) {
                    const int nb = sorted[m];
                    if (nb == i) continue;
                    if (k++ == seen) {
                        *partner = nb;
                        return;
                    }
                }
            }
}

__global__ void k_max_int(int n, const int *__restrict__ v, int *out)
{
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = max(m, v[i]);
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

// ---- fold of padding rows onto contributing atoms (neighbor.py:227-263) ----

// bucket k of the images: padding rows fold_rows[fold_start[k] .. fold_start[k+1]),
// ascending.  pairwise = 1 is the reference's branch for n < n_padding:
// total = f_k + (0 + p_0 + p_1 + ...) (np.sum over the rows, then one add);
// pairwise = 0 is its other branch: ((f_k + p_0) + p_1) + ...
__global__ void k_fold(int n, int pairwise, const int *__restrict__ fold_start,
                       const int *__restrict__ fold_rows, const double *__restrict__ f,
                       double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int b = fold_start[k], e = fold_start[k + 1];
    double t0 = f[3 * (long)k], t1 = f[3 * (long)k + 1], t2 = f[3 * (long)k + 2];
    if (b < e || pairwise) {
        if (pairwise) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int q = b; q < e; ++q) {
                const long r = (long)n + fold_rows[q];
                s0 += f[3 * r];
                s1 += f[3 * r + 1];
                s2 += f[3 * r + 2];
            }
            t0 += s0;
            t1 += s1;
            t2 += s2;
        } else {
            for (int q = b; q < e; ++q) {
                const long r = (long)n + fold_rows[q];
                t0 += f[3 * r];
                t1 += f[3 * r + 1];
                t2 += f[3 * r + 2];
            }
        }
    }
    out[3 * (long)k] = t0;
    out[3 * (long)k + 1] = t1;
    out[3 * (long)k + 2] = t2;
}

__global__ void k_fold_count(int n_pad, int n, const int *__restrict__ image, int *count, int *bad)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_pad) return;
    const int k = image[q];
    if (k < 0 || k >= n) {
        atomicOr(bad, 1);
        return;
    }
    atomicAdd(count + k, 1);
}

// radix-sort helpers
hipError_t sort_pairs(void *tmp, size_t &bytes, const int *kin, int *kout, const int *vin,
                      int *vout, int n, int bits, hipStream_t s)
{
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, bits, s);
}

hipError_t exclusive_scan(void *tmp, size_t &bytes, const int *in, int *out, int n, hipStream_t s)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
}

int bits_for(long v)
{
    int b = 1;
    while ((1L << b) <= v) ++b;
    return b;
}

// A stable counting sort of padding rows by image: rows (0-based among the
// padding rows) grouped by contributing atom, ascending within a group.
struct FoldTable {
    int *start = nullptr;  // n + 1
    int *rows = nullptr;   // n_pad
};

int build_fold(int n, int n_pad, const int *d_image, FoldTable &ft, hipStream_t s)
{
    DevBuf<int> start, rows, count, keys, iota, bad;
    HIP_TRY(start.alloc(n + 1), "hipMalloc(fold start)");
    HIP_TRY(count.alloc(n + 1), "hipMalloc(fold count)");
    HIP_TRY(bad.alloc(1), "hipMalloc(flag)");
    HIP_TRY(hipMemsetAsync(count.p, 0, sizeof(int) * (n + 1), s), "hipMemsetAsync");
    HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), s), "hipMemsetAsync");
    if (n_pad > 0) {
        hipLaunchKernelGGL(k_fold_count, dim3(grid_for(n_pad)), dim3(kBlock), 0, s, n_pad, n,
                           d_image, count.p, bad.p);
        HIP_TRY(hipGetLastError(), "fold count");
    }
    size_t b1 = 0, b2 = 0;
    HIP_TRY(exclusive_scan(nullptr, b1, count.p, start.p, n + 1, s), "scan size");
    if (n_pad > 0) {
        HIP_TRY(rows.alloc(n_pad), "hipMalloc(fold rows)");
        HIP_TRY(keys.alloc(n_pad), "hipMalloc(fold keys)");
        HIP_TRY(iota.alloc(n_pad), "hipMalloc(fold iota)");
        HIP_TRY(sort_pairs(nullptr, b2, d_image, keys.p, iota.p, rows.p, n_pad, bits_for(n), s),
                "sort size");
    }
    DevBuf<char> tmp;
    HIP_TRY(tmp.alloc(std::max(b1, b2) + 16), "hipMalloc(scratch)");
    size_t bb = std::max(b1, b2) + 16;
    HIP_TRY(exclusive_scan(tmp.p, bb, count.p, start.p, n + 1, s), "scan");
    int hbad = 0;
    HIP_TRY(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy flag");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (hbad) return ldsf::fail(LDSF_ERR_ARG, "assemble_forces: padding image out of range [0, n)");
    if (n_pad > 0) {
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n_pad)), dim3(kBlock), 0, s, n_pad, iota.p);
        HIP_TRY(hipGetLastError(), "iota");
        bb = std::max(b1, b2) + 16;
        HIP_TRY(sort_pairs(tmp.p, bb, d_image, keys.p, iota.p, rows.p, n_pad, bits_for(n), s),
                "sort");
        HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    ft.start = start.take();
    ft.rows = rows.take();
    return LDSF_OK;
}

// The gather and the fold in one pass: atom k's dE/dr is the sum of its own
// row's adjoint records (ascending edge order, as k_gather sums it into a
// zeroed row) folded with its padding rows' sums in the reference's order,
// as k_fold folds them -- bitwise the two launches' result, without writing
// and re-reading the padding rows.
__global__ void k_gather_fold(int n, int pairwise, const int *__restrict__ fold_start,
                              const int *__restrict__ fold_rows, const int *__restrict__ in_start,
                              const double *__restrict__ edge, double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    ldsf::row_sum3(edge, in_start[k], in_start[k + 1], t0, t1, t2);
    const int b = fold_start[k], e = fold_start[k + 1];
    if (pairwise) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int q = b; q < e; ++q) {
            const long r = (long)n + fold_rows[q];
            double f0 = 0.0, f1 = 0.0, f2 = 0.0;
            ldsf::row_sum3(edge, in_start[r], in_start[r + 1], f0, f1, f2);
            s0 += f0;
            s1 += f1;
            s2 += f2;
        }
        t0 += s0;
        t1 += s1;
        t2 += s2;
    } else {
        for (int q = b; q < e; ++q) {
            const long r = (long)n + fold_rows[q];
            double f0 = 0.0, f1 = 0.0, f2 = 0.0;
            ldsf::row_sum3(edge, in_start[r], in_start[r + 1], f0, f1, f2);
            t0 += f0;
            t1 += f1;
            t2 += f2;
        }
    }
    out[3 * (long)k] = t0;
    out[3 * (long)k + 1] = t1;
    out[3 * (long)k + 2] = t2;
}

int run_fold(int n, int n_pad, const FoldTable &ft, const double *d_forces, double *d_out,
             hipStream_t s)
{
    if (n <= 0) return LDSF_OK;
    // the reference's branch: per-atom np.sum when n < n_padding
    const int pairwise = (n_pad > 0 && n < n_pad) ? 1 : 0;
    hipLaunchKernelGGL(k_fold, dim3(grid_for(n)), dim3(kBlock), 0, s, n, pairwise, ft.start,
                       ft.rows, d_forces, d_out);
    HIP_TRY(hipGetLastError(), "fold");
    return LDSF_OK;
}

}  // namespace

struct ldsf_layout {
    int device = 0;
    int n_contrib = 0, n_pad = 0, max_nn = 0;
    long n_neigh = 0;
    double *coords = nullptr;
    int *species = nullptr, *image = nullptr, *numneigh = nullptr, *neighlist = nullptr;
    FoldTable fold;

    ~ldsf_layout()
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(device);
        for (void *p : {(void *)coords, (void *)species, (void *)image, (void *)numneigh,
                        (void *)neighlist, (void *)fold.start, (void *)fold.rows})
            if (p) (void)hipFree(p);
        (void)hipSetDevice(cur);
    }
};

namespace {

// the extent over the blocks' partial boxes, read back to the host
hipError_t read_box(const Box *partial, int nb, Box &out, Box *dev_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(kBlock), 0, s, nb, partial, dev_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&out, dev_out, sizeof(Box), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

// padding rows appended to the contributing rows (nbl_create_paddings)
int make_padding(ldsf_layout *L, int n, double cutoff, const double *cell, const int *pbc,
                 const double *d_coords, const int *d_species, hipStream_t s)
{
    double tcell[9], fcell[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) tcell[3 * r + c] = cell[3 * c + r];
    if (!hinvert(tcell, fcell))
        return ldsf::fail(LDSF_ERR_SINGULAR_CELL, "In inverting the cell matrix, the determinant is 0!");
    PadParams pp;
    std::memcpy(pp.tcell.m, tcell, sizeof tcell);
    Mat3 fm;
    std::memcpy(fm.m, fcell, sizeof fcell);
    const int nb = (int)std::min<long>(std::max<long>(1, grid_for(n)), 1024);
    DevBuf<double> frac;
    DevBuf<Box> partial, boxd;
    HIP_TRY(frac.alloc(3 * (size_t)std::max(n, 1)), "hipMalloc(frac)");
    HIP_TRY(partial.alloc(nb), "hipMalloc(box)");
    HIP_TRY(boxd.alloc(1), "hipMalloc(box)");
    Box box;
    if (n > 0) {
        hipLaunchKernelGGL(k_frac, dim3(nb), dim3(kBlock), 0, s, n, fm, d_coords, frac.p, partial.p);
        HIP_TRY(hipGetLastError(), "frac");
        HIP_TRY(read_box(partial.p, nb, box, boxd.p, s), "extent");
    } else {
        for (int d = 0; d < 3; ++d) {
            box.lo[d] = 1e10;
            box.hi[d] = -1e10;
        }
    }
    for (int d = 0; d < 3; ++d) {
        pp.lo[d] = box.lo[d] - kTol;
        pp.hi[d] = box.hi[d] + kTol;
    }
    double xp[3], dist[3];
    hcross(cell + 3, cell + 6, xp);
    const double volume = std::abs(hdot(cell, xp));
    hcross(cell + 3, cell + 6, xp);
    dist[0] = volume / hnorm(xp);
    hcross(cell + 6, cell + 0, xp);
    dist[1] = volume / hnorm(xp);
    hcross(cell + 0, cell + 3, xp);
    dist[2] = volume / hnorm(xp);
    for (int d = 0; d < 3; ++d) {
        const double ratio = cutoff / dist[d];
        if (!(ratio < 1e6)) return ldsf::fail(LDSF_ERR_NEIGHBOR, "create_paddings: cutoff / cell width too large");
        pp.size[d] = (int)std::ceil(ratio);
        pp.slack[d] = (double)pp.size[d] - ratio;
    }
    std::vector<int3> imgs;
    for (int i = -pp.size[0]; i <= pp.size[0]; ++i)
        for (int j = -pp.size[1]; j <= pp.size[1]; ++j)
            for (int k = -pp.size[2]; k <= pp.size[2]; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                if ((!pbc[0] && i) || (!pbc[1] && j) || (!pbc[2] && k)) continue;
                imgs.push_back(make_int3(i, j, k));
            }
    const long total = (long)imgs.size() * n;
    if (total > 2000000000L)
        return ldsf::fail(LDSF_ERR_UNSUPPORTED, "create_paddings: more than 2e9 padding candidates");
    int n_pad = 0;
    DevBuf<int3> dimg;
    DevBuf<int> flag, pos;
    if (total > 0) {
        HIP_TRY(dimg.alloc(imgs.size()), "hipMalloc(images)");
        HIP_TRY(hipMemcpyAsync(dimg.p, imgs.data(), sizeof(int3) * imgs.size(), hipMemcpyHostToDevice, s),
                "copy images");
        HIP_TRY(flag.alloc(total), "hipMalloc(flags)");
        HIP_TRY(pos.alloc(total), "hipMalloc(positions)");
        hipLaunchKernelGGL(k_pad_flag, dim3(grid_for(total)), dim3(kBlock), 0, s, total, n, dimg.p,
                           pp, frac.p, flag.p);
        HIP_TRY(hipGetLastError(), "padding flags");
        size_t bytes = 0;
        HIP_TRY(exclusive_scan(nullptr, bytes, flag.p, pos.p, (int)total, s), "scan size");
        DevBuf<char> tmp;
        HIP_TRY(tmp.alloc(bytes + 16), "hipMalloc(scan scratch)");
        HIP_TRY(exclusive_scan(tmp.p, bytes, flag.p, pos.p, (int)total, s), "scan");
        int last[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(last, pos.p + total - 1, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
        HIP_TRY(hipMemcpyAsync(last + 1, flag.p + total - 1, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
        HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
        n_pad = last[0] + last[1];
    }
    const long n_total = (long)n + n_pad;
    if (n_total > 0x7fffffffL) return ldsf::fail(LDSF_ERR_UNSUPPORTED, "more than 2^31-1 rows");
    DevBuf<double> coords;
    DevBuf<int> species, image;
    HIP_TRY(coords.alloc(3 * n_total), "hipMalloc(coords)");
    HIP_TRY(species.alloc(n_total), "hipMalloc(species)");
    HIP_TRY(image.alloc(n_total), "hipMalloc(image)");
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(coords.p, d_coords, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s), "copy coords");
        HIP_TRY(hipMemcpyAsync(species.p, d_species, sizeof(int) * n, hipMemcpyDeviceToDevice, s), "copy species");
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, s, n, image.p);
        HIP_TRY(hipGetLastError(), "iota");
    }
    if (n_pad > 0) {
        hipLaunchKernelGGL(k_pad_emit, dim3(grid_for(total)), dim3(kBlock), 0, s, total, n, dimg.p,
                           pp.tcell, frac.p, flag.p, pos.p, d_species, coords.p, species.p, image.p);
        HIP_TRY(hipGetLastError(), "padding rows");
    }
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    L->n_contrib = n;
    L->n_pad = n_pad;
    L->coords = coords.take();
    L->species = species.take();
    L->image = image.take();
    return LDSF_OK;
}

// the neighbour lists of rows 0..n_query-1 among rows 0..n_total-1 (nbl_build
// with needNeighbors = 1 for the contributing rows, then the flattening)
int make_neighbors(ldsf_layout *L, int n_total, int n_query, double infl, hipStream_t s)
{
    const double *x = L->coords;
    const int nb = (int)std::min<long>(grid_for(n_total), 1024);
    DevBuf<Box> partial, boxd;
    HIP_TRY(partial.alloc(nb), "hipMalloc(box)");
    HIP_TRY(boxd.alloc(1), "hipMalloc(box)");
    hipLaunchKernelGGL(k_extent, dim3(nb), dim3(kBlock), 0, s, n_total, x, partial.p);
    HIP_TRY(hipGetLastError(), "extent");
    Box box;
    HIP_TRY(read_box(partial.p, nb, box, boxd.p, s), "extent");
    double x0[3];
    HIP_TRY(hipMemcpy(x0, x, sizeof x0, hipMemcpyDeviceToHost), "copy coords");
    Grid g;
    long cells = 1;
    for (int d = 0; d < 3; ++d) {
        // min starts at coord0 (an element of the set); max at coord0 + 1
        g.lo[d] = box.lo[d];
        g.hi[d] = std::max(x0[d] + 1.0, box.hi[d]);
        const double q = (g.hi[d] - g.lo[d]) / infl;
        if (!(q < 2e9))
            return ldsf::fail(LDSF_ERR_NEIGHBOR,
                              "Cell size too large! (partilces fly away) or\nCollision of atoms happened!");
        g.size[d] = static_cast<int>(q);
        if (g.size[d] <= 0) g.size[d] = 1;
        cells *= g.size[d];
    }
    if (cells > 1000000000L)
        return ldsf::fail(LDSF_ERR_NEIGHBOR,
                          "Cell size too large! (partilces fly away) or\nCollision of atoms happened!");
    DevBuf<int> cell, cell_sorted, iota, sorted, cnt, start;
    HIP_TRY(cell.alloc(n_total), "hipMalloc(cells)");
    HIP_TRY(cell_sorted.alloc(n_total), "hipMalloc(cells)");
    HIP_TRY(iota.alloc(n_total), "hipMalloc(iota)");
    HIP_TRY(sorted.alloc(n_total), "hipMalloc(sorted)");
    HIP_TRY(cnt.alloc(cells + 1), "hipMalloc(cell counts)");
    HIP_TRY(start.alloc(cells + 1), "hipMalloc(cell starts)");
    HIP_TRY(hipMemsetAsync(cnt.p, 0, sizeof(int) * (cells + 1), s), "hipMemsetAsync");
    hipLaunchKernelGGL(k_cell_of, dim3(grid_for(n_total)), dim3(kBlock), 0, s, n_total, g, x, cell.p, cnt.p);
    HIP_TRY(hipGetLastError(), "cells");
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n_total)), dim3(kBlock), 0, s, n_total, iota.p);
    HIP_TRY(hipGetLastError(), "iota");
    size_t b1 = 0, b2 = 0, b3 = 0;
    const int bits = bits_for(cells);
    HIP_TRY(sort_pairs(nullptr, b1, cell.p, cell_sorted.p, iota.p, sorted.p, n_total, bits, s), "sort size");
    HIP_TRY(exclusive_scan(nullptr, b2, cnt.p, start.p, (int)cells + 1, s), "scan size");
    HIP_TRY(exclusive_scan(nullptr, b3, cnt.p, start.p, n_query + 1, s), "scan size");
    const size_t tb = std::max(b1, std::max(b2, b3)) + 16;
    DevBuf<char> tmp;
    HIP_TRY(tmp.alloc(tb), "hipMalloc(scratch)");
    size_t bb = tb;
    HIP_TRY(sort_pairs(tmp.p, bb, cell.p, cell_sorted.p, iota.p, sorted.p, n_total, bits, s), "sort");
    bb = tb;
    HIP_TRY(exclusive_scan(tmp.p, bb, cnt.p, start.p, (int)cells + 1, s), "scan");

    DevBuf<int> count, begin, mx;
    DevBuf<unsigned long long> coll;
    HIP_TRY(count.alloc(n_query + 1), "hipMalloc(counts)");
    HIP_TRY(begin.alloc(n_query + 1), "hipMalloc(offsets)");
    HIP_TRY(mx.alloc(2), "hipMalloc(max)");
    HIP_TRY(coll.alloc(1), "hipMalloc(collision)");
    HIP_TRY(hipMemsetAsync(count.p, 0, sizeof(int) * (n_query + 1), s), "hipMemsetAsync");
    HIP_TRY(hipMemsetAsync(mx.p, 0, sizeof(int) * 2, s), "hipMemsetAsync");
    HIP_TRY(hipMemsetAsync(coll.p, 0xff, sizeof(unsigned long long), s), "hipMemsetAsync");
    const double cutsq = infl * infl;
    hipLaunchKernelGGL(k_neigh<false>, dim3(grid_for(n_total)), dim3(kBlock), 0, s, n_total, n_query,
                       g, cutsq, sorted.p, start.p, x, count.p, (const int *)nullptr, (int *)nullptr,
                       coll.p);
    HIP_TRY(hipGetLastError(), "neighbour count");
    bb = tb;
    HIP_TRY(exclusive_scan(tmp.p, bb, count.p, begin.p, n_query + 1, s), "scan");
    if (n_query > 0) {
        hipLaunchKernelGGL(k_max_int, dim3(std::min<unsigned>(grid_for(n_query), 1024)), dim3(kBlock),
                           0, s, n_query, count.p, mx.p);
        HIP_TRY(hipGetLastError(), "max");
    }
    unsigned long long hcoll = 0;
    int total = 0, hmax = 0;
    HIP_TRY(hipMemcpyAsync(&hcoll, coll.p, sizeof hcoll, hipMemcpyDeviceToHost, s), "copy");
    HIP_TRY(hipMemcpyAsync(&total, begin.p + n_query, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
    HIP_TRY(hipMemcpyAsync(&hmax, mx.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (hcoll != ~0ULL) {
        const int i = (int)(hcoll >> 32), seen = (int)(hcoll & 0xffffffffu);
        hipLaunchKernelGGL(k_collision_partner, dim3(1), dim3(64), 0, s, i, seen, g, sorted.p,
                           start.p, x, mx.p + 1);
        int partner = -1;
        HIP_TRY(hipMemcpyAsync(&partner, mx.p + 1, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
        HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
        char msg[160];
        std::snprintf(msg, sizeof msg, "Collision of atoms %d and %d.", i + 1, partner + 1);
        return ldsf::fail(LDSF_ERR_NEIGHBOR, msg);
    }
    if (total < 0) return ldsf::fail(LDSF_ERR_UNSUPPORTED, "more than 2^31-1 neighbour-list entries");
    DevBuf<int> nl;
    HIP_TRY(nl.alloc(std::max(total, 1)), "hipMalloc(neighbour list)");
    hipLaunchKernelGGL(k_neigh<true>, dim3(grid_for(n_total)), dim3(kBlock), 0, s, n_total, n_query, g,
                       cutsq, sorted.p, start.p, x, (int *)nullptr, begin.p, nl.p, coll.p);
    HIP_TRY(hipGetLastError(), "neighbour fill");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    L->n_neigh = total;
    L->max_nn = hmax;
    L->numneigh = count.take();
    L->neighlist = nl.take();
    return LDSF_OK;
}

}  // namespace

namespace ldsf {
int layout_gather_fold(const ldsf_layout *L, const int *in_start, const double *edge,
                       double *d_out, hipStream_t s)
{
    const int n = L->n_contrib;
    if (n <= 0) return LDSF_OK;
    const int pairwise = (L->n_pad > 0 && n < L->n_pad) ? 1 : 0;
    hipLaunchKernelGGL(k_gather_fold, dim3(grid_for(n)), dim3(kBlock), 0, s, n, pairwise,
                       L->fold.start, L->fold.rows, in_start, edge, d_out);
    HIP_TRY(hipGetLastError(), "gather fold");
    return LDSF_OK;
}

int layout_rows(const ldsf_layout *L) { return L->n_contrib + L->n_pad; }
}  // namespace ldsf

extern "C" {

LDSF_API int ldsf_layout_create(ldsf_layout **out, int n_contrib, double cutoff, const double *cell,
                                const int *pbc, const double *d_coords, const int *d_species,
                                void *stream)
{
    if (!out) return ldsf::fail(LDSF_ERR_ARG, "layout_create: null output");
    *out = nullptr;
    if (n_contrib <= 0 || !(cutoff > 0.0) || !cell || !pbc || !d_coords || !d_species)
        return ldsf::fail(LDSF_ERR_ARG, "layout_create: bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto L = std::make_unique<ldsf_layout>();
    HIP_TRY(hipGetDevice(&L->device), "hipGetDevice");
    if (int rc = make_padding(L.get(), n_contrib, cutoff, cell, pbc, d_coords, d_species, s)) return rc;
    if (int rc = make_neighbors(L.get(), n_contrib + L->n_pad, n_contrib, cutoff, s)) return rc;
    if (int rc = build_fold(n_contrib, L->n_pad, L->image + n_contrib, L->fold, s)) return rc;
    *out = L.release();
    return LDSF_OK;
}

LDSF_API void ldsf_layout_destroy(ldsf_layout *layout) { delete layout; }

LDSF_API int ldsf_layout_info(const ldsf_layout *layout, int *n_contrib, int *n_padding,
                              long *n_neigh, int *max_neighbors)
{
    if (!layout) return ldsf::fail(LDSF_ERR_ARG, "layout_info: null layout");
    if (n_contrib) *n_contrib = layout->n_contrib;
    if (n_padding) *n_padding = layout->n_pad;
    if (n_neigh) *n_neigh = layout->n_neigh;
    if (max_neighbors) *max_neighbors = layout->max_nn;
    return LDSF_OK;
}

LDSF_API int ldsf_layout_arrays(const ldsf_layout *layout, const double **coords,
                                const int **species, const int **image, const int **numneigh,
                                const int **neighlist)
{
    if (!layout) return ldsf::fail(LDSF_ERR_ARG, "layout_arrays: null layout");
    if (coords) *coords = layout->coords;
    if (species) *species = layout->species;
    if (image) *image = layout->image;
    if (numneigh) *numneigh = layout->numneigh;
    if (neighlist) *neighlist = layout->neighlist;
    return LDSF_OK;
}

LDSF_API int ldsf_layout_copy(const ldsf_layout *layout, double *d_coords, int *d_species,
                              int *d_image, int *d_numneigh, int *d_neighlist, void *stream)
{
    if (!layout) return ldsf::fail(LDSF_ERR_ARG, "layout_copy: null layout");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t rows = (size_t)layout->n_contrib + layout->n_pad;
    auto cp = [&](void *dst, const void *src, size_t bytes) {
        return (dst && bytes) ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s)
                              : hipSuccess;
    };
    HIP_TRY(cp(d_coords, layout->coords, sizeof(double) * 3 * rows), "copy coords");
    HIP_TRY(cp(d_species, layout->species, sizeof(int) * rows), "copy species");
    HIP_TRY(cp(d_image, layout->image, sizeof(int) * rows), "copy image");
    HIP_TRY(cp(d_numneigh, layout->numneigh, sizeof(int) * layout->n_contrib), "copy numneigh");
    HIP_TRY(cp(d_neighlist, layout->neighlist, sizeof(int) * layout->n_neigh), "copy neighlist");
    return LDSF_OK;
}


LDSF_API int ldsf_layout_assemble_forces(const ldsf_layout *layout, const double *d_forces,
                                         double *d_out, void *stream)
{
    if (!layout || !d_forces || !d_out)
        return ldsf::fail(LDSF_ERR_ARG, "layout_assemble_forces: bad arguments");
    return run_fold(layout->n_contrib, layout->n_pad, layout->fold, d_forces, d_out,
                    static_cast<hipStream_t>(stream));
}

LDSF_API int ldsf_assemble_forces(int n_contrib, int n_padding, const double *d_forces,
                                  const int *d_padding_image, double *d_out, void *stream)
{
    if (n_contrib < 0 || n_padding < 0 || (n_contrib > 0 && (!d_forces || !d_out)) ||
        (n_padding > 0 && !d_padding_image))
        return ldsf::fail(LDSF_ERR_ARG, "assemble_forces: bad arguments");
    if (n_contrib == 0) {
        if (n_padding > 0) return ldsf::fail(LDSF_ERR_ARG, "assemble_forces: padding image out of range [0, n)");
        return LDSF_OK;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    FoldTable ft;
    if (int rc = build_fold(n_contrib, n_padding, d_padding_image, ft, s)) return rc;
    DevBuf<int> g1, g2;
    g1.p = ft.start;
    g2.p = ft.rows;
    if (int rc = run_fold(n_contrib, n_padding, ft, d_forces, d_out, s)) return rc;
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    return LDSF_OK;
}

}  // extern "C"
