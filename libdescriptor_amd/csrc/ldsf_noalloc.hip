// ldsf_gradient_noalloc: the gradient entry point for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It has ldsf_gradient's semantics
// (Descriptor::gradient, src/Descriptors.cpp:165-190: d_coordinates += J^T
// dE/dzeta, desc += the descriptors) on device pointers, and keeps everything
// a plan would allocate -- the parameter tables, CSR offsets, the transposed
// neighbour list, the adjoint's edge records, flags and scan partials -- in
// one static device array, `scratch`, of LDSF_NOALLOC_SCRATCH_BYTES per GPU.
//
// Host code never uses the array's name as an address: its device address is
// looked up once per device with hipGetSymbolAddress (with that device
// current) and kept, and every launch and copy takes that pointer.  Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU
// race; everything a call needs at zero is cleared by a kernel at its start.
// The inputs are validated on the device before any kernel reads through
// them: a flagged call skips every centre and the gather, so desc and
// d_coordinates are left as they were, and the call returns the error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/ldsf.h"
#include "ldsf_internal.h"
#include "sf_kernels.h"
#include "sf_scan.h"

namespace {

constexpr size_t kScratchBytes = LDSF_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

int hip_fail(hipError_t e, const char *where)
{
    return ldsf::fail(LDSF_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

#define NA_TRY(expr, where)                                   \
    do {                                                      \
        hipError_t e__ = (expr);                              \
        if (e__ != hipSuccess) return hip_fail(e__, where);   \
    } while (0)

// the scratch array of the current device (looked up on its first use there)
int scratch_of(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices)
        return ldsf::fail(LDSF_ERR_DEVICE, "gradient_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return ldsf::fail(LDSF_ERR_DEVICE, "hipGetSymbolAddress(scratch): null address");
        g_scratch[device] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch[device]);
    return LDSF_OK;
}

// flags[0] bits: 1 neighbour index out of [0, n_total) (set by the edge
// count), 2 negative neighbour count, 4 species code out of range,
// 16 neighbour count above max_neighbors, 32 counts do not sum to n_neigh
enum { F_INDEX = 1, F_NEGATIVE = 2, F_SPECIES = 4, F_MAXNN = 16, F_SUM = 32 };

// Byte offsets of one call's regions in the scratch array, each 256-aligned.
struct Carve {
    size_t dd, di, wpar, flags, zeros, offs, count, cursor, in_start, in_edges, inv_pos, part,
        edge, end;
};

inline size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

Carve carve(const ldsf_sf *sf, long n, long R, long E)
{
    const ldsf::PackedTable &T = sf->packed;
    Carve c;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = up256(o + bytes);
        return at;
    };
    c.dd = take(sizeof(double) * (T.d.size() + 1));
    c.di = take(sizeof(int) * (T.i.size() + 1));
    c.wpar = take(sizeof(ldsf::WaveParams));
    c.flags = take(sizeof(int) * 64);
    c.zeros = take(sizeof(int) * n);  // request of every centre: 0
    c.offs = take(sizeof(int) * (n + 1));
    c.count = take(sizeof(int) * (R + 1));
    c.cursor = take(sizeof(int) * R);
    c.in_start = take(sizeof(int) * (R + 1));
    c.in_edges = take(sizeof(int) * (E + n));
    c.inv_pos = take(sizeof(int) * (E + n));
    c.part = take(sizeof(int) * (ldsf::kScanBlocks + 1));
    c.edge = take(sizeof(double) * 3 * (E + n));
    c.end = o;
    return c;
}

constexpr int kBlock = 256;

inline dim3 blocks_for(long n) { return dim3((unsigned)std::max<long>((n + kBlock - 1) / kBlock, 1)); }

// the counters the call starts from (edge counts and fill cursors per row,
// the flags) and the centres' request index, all to zero
__global__ void k_na_zero(int R, int n, int *count, int *cursor, int *zeros, int *flags)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= R) count[t] = 0;
    if (t < R) cursor[t] = 0;
    if (t < n) zeros[t] = 0;
    if (t < 64) flags[t] = 0;
}

__global__ void k_na_validate(int n, int R, int max_nn, const int *__restrict__ numneigh,
                              const int *__restrict__ species, int n_species, int *flags)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int f = 0;
    if (t < n) {
        const int nn = numneigh[t];
        if (nn < 0) f |= F_NEGATIVE;
        if (nn > max_nn) f |= F_MAXNN;
    }
    if (t < R) {
        const int s = species[t];
        if (s < 0 || s >= n_species) f |= F_SPECIES;
    }
    if (f) atomicOr(flags, f);
}

__global__ void k_na_sum(const int *__restrict__ offs, int n, int E, int *flags)
{
    if (threadIdx.x == 0 && offs[n] != E) atomicOr(flags, F_SUM);
}

}  // namespace

extern "C" LDSF_API int ldsf_gradient_noalloc(ldsf_sf *sf, int n_atoms, int n_total, int n_neigh,
                                              int max_neighbors, const int *d_species,
                                              const int *d_neighbor_list,
                                              const int *d_number_of_neighs,
                                              const double *d_coordinates,
                                              double *d_d_coordinates, double *d_desc,
                                              const double *d_dE_dzeta, int precision,
                                              void *stream)
{
    using namespace ldsf;
    if (!sf) return fail(LDSF_ERR_ARG, "gradient_noalloc: null descriptor");
    if (n_atoms < 0 || n_total < n_atoms || n_neigh < 0 || max_neighbors < 0)
        return fail(LDSF_ERR_ARG, "gradient_noalloc: bad sizes");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return fail(LDSF_ERR_ARG, "gradient_noalloc: unknown precision");
    if ((long)max_neighbors * n_atoms < n_neigh)
        return fail(LDSF_ERR_ARG, "gradient_noalloc: max_neighbors below n_neigh / n_atoms");
    if (n_atoms > 0 && (!d_species || !d_number_of_neighs || !d_coordinates ||
                        !d_d_coordinates || !d_dE_dzeta))
        return fail(LDSF_ERR_ARG, "gradient_noalloc: null input or output");
    if (n_neigh > 0 && !d_neighbor_list)
        return fail(LDSF_ERR_ARG, "gradient_noalloc: null neighbour list");
    if ((long)n_neigh + n_atoms > 0x7fffffffL)
        return fail(LDSF_ERR_UNSUPPORTED, "gradient_noalloc: more than 2^31-1 records");
    const Carve cv = carve(sf, n_atoms, n_total, n_neigh);
    if (cv.end > kScratchBytes)
        return fail(LDSF_ERR_UNSUPPORTED,
                    "gradient_noalloc: the input needs " + std::to_string(cv.end) +
                        " bytes of scratch, more than LDSF_NOALLOC_SCRATCH_BYTES (" +
                        std::to_string(kScratchBytes) + ")");
    if (n_atoms == 0) return LDSF_OK;

    hipStream_t s = (hipStream_t)stream;
    int device = 0;
    NA_TRY(hipGetDevice(&device), "hipGetDevice");
    unsigned char *base = nullptr;
    if (int rc = scratch_of(device, &base)) return rc;
    double *dd = reinterpret_cast<double *>(base + cv.dd);
    int *di = reinterpret_cast<int *>(base + cv.di);
    auto *wpar = reinterpret_cast<WaveParams *>(base + cv.wpar);
    int *flags = reinterpret_cast<int *>(base + cv.flags);
    int *zeros = reinterpret_cast<int *>(base + cv.zeros);
    int *offs = reinterpret_cast<int *>(base + cv.offs);
    int *count = reinterpret_cast<int *>(base + cv.count);
    int *cursor = reinterpret_cast<int *>(base + cv.cursor);
    int *in_start = reinterpret_cast<int *>(base + cv.in_start);
    int *in_edges = reinterpret_cast<int *>(base + cv.in_edges);
    int *inv_pos = reinterpret_cast<int *>(base + cv.inv_pos);
    int *part = reinterpret_cast<int *>(base + cv.part);
    double *edge = reinterpret_cast<double *>(base + cv.edge);

    // the parameter tables: the set's packed image and, if the wave kernels
    // cover it, their parameter block (both copied from host memory)
    const PackedTable &T = sf->packed;
    if (!T.d.empty())
        NA_TRY(hipMemcpyAsync(dd, T.d.data(), sizeof(double) * T.d.size(), hipMemcpyHostToDevice, s),
               "upload parameter table");
    if (!T.i.empty())
        NA_TRY(hipMemcpyAsync(di, T.i.data(), sizeof(int) * T.i.size(), hipMemcpyHostToDevice, s),
               "upload parameter table");
    const SFTab tab = table_view(sf, dd, di);
    WaveTab wt;
    WaveParams wp;
    const bool wave = wave_build(sf, &wt, &wp) && wave_runs(wt, precision) &&
                      wave_group_size(max_neighbors, wt.n_species) > 0;
    if (wave) {
        NA_TRY(hipMemcpyAsync(wpar, &wp, sizeof(wp), hipMemcpyHostToDevice, s),
               "upload wave parameters");
        wt.rc = tab.rc;
        wt.par = wpar;
    }

    // validation and the neighbour list's CSR offsets and transpose
    const int n = n_atoms, R = n_total, E = n_neigh;
    hipLaunchKernelGGL(k_na_zero, blocks_for(std::max(R + 1, std::max(n, 64))), dim3(kBlock), 0, s,
                       R, n, count, cursor, zeros, flags);
    hipLaunchKernelGGL(k_na_validate, blocks_for(std::max(n, R)), dim3(kBlock), 0, s, n, R,
                       max_neighbors, d_number_of_neighs, d_species, tab.n_species, flags);
    enqueue_scan(Count{nullptr, n}, d_number_of_neighs, part, offs, s);
    hipLaunchKernelGGL(k_na_sum, dim3(1), dim3(64), 0, s, offs, n, E, flags);
    NA_TRY(hipGetLastError(), "validation launch");
    NA_TRY(launch_edge_targets_count(E, n, d_neighbor_list, nullptr, R, count, flags, s),
           "edge count");
    enqueue_scan(Count{nullptr, R}, count, part, in_start, s);
    NA_TRY(launch_edge_targets_fill(E, n, d_neighbor_list, nullptr, R, in_start, cursor, in_edges,
                                    s),
           "edge fill");
    NA_TRY(launch_sort_inverse(R, in_start, in_edges, inv_pos, s), "edge sort");

    // the kernels skip every centre of a flagged call (its one "request")
    DevCounts dc;
    dc.req_of = zeros;
    dc.req_flag = flags;
    hipError_t e;
    if (wave) {
        e = launch_adj_wave(wt, precision, n, nullptr, offs, d_neighbor_list, d_species,
                            d_coordinates, d_dE_dzeta, inv_pos, edge, E, tab.width, max_neighbors,
                            wave_group_size(max_neighbors, wt.n_species), s, dc, d_desc);
    } else {
        e = hipSuccess;
        if (d_desc)
            e = launch_fwd_general(tab, precision, n, nullptr, offs, d_neighbor_list, d_species,
                                   d_coordinates, d_desc, s, dc);
        if (e == hipSuccess)
            e = launch_adj_general(tab, precision, n, nullptr, offs, d_neighbor_list, d_species,
                                   d_coordinates, d_dE_dzeta, inv_pos, edge, E, s, dc);
    }
    NA_TRY(e, "kernel launch");
    NA_TRY(launch_gather(R, in_start, edge, d_d_coordinates, s, flags), "gather launch");

    int h_flags = 0;
    NA_TRY(hipMemcpyAsync(&h_flags, flags, sizeof(int), hipMemcpyDeviceToHost, s), "copy flags");
    NA_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (h_flags & F_NEGATIVE) return fail(LDSF_ERR_ARG, "gradient_noalloc: negative number_of_neighs entry");
    if (h_flags & F_MAXNN)
        return fail(LDSF_ERR_ARG, "gradient_noalloc: a number_of_neighs entry exceeds max_neighbors");
    if (h_flags & F_SUM)
        return fail(LDSF_ERR_ARG, "gradient_noalloc: number_of_neighs does not sum to n_neigh");
    if (h_flags & F_SPECIES) return fail(LDSF_ERR_ARG, "gradient_noalloc: species code out of range");
    if (h_flags & F_INDEX)
        return fail(LDSF_ERR_ARG, "gradient_noalloc: neighbour index out of range [0, n_total)");
    return LDSF_OK;
}
