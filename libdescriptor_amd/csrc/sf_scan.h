// Exclusive prefix sums of int32 counts without library scratch or memset
// nodes, in three fixed-grid launches: per-block partial sums, one block
// scanning the partials, then block-local scans offset by them; two
// independent scans can share the three launches.  The element
// count is read on the device (a pointer into a graph-replayed launch's
// header) or passed by value, so the same kernels serve the serving loop's
// replayed graphs and the allocation-free entry point.
#ifndef LDSF_SF_SCAN_H_
#define LDSF_SF_SCAN_H_

#include <hip/hip_runtime.h>

namespace ldsf {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanBlocks = 256;  // partials: kScanBlocks + 1 ints of scratch

// an element count: *p when p is set, else v
struct Count {
    const int *p;
    int v;
    __device__ __forceinline__ long get() const { return p ? *p : v; }
};

__device__ inline int scan_block_sum(int v)
{
    __shared__ int ws[kScanThreads / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    int s = 0;
    for (int k = 0; k < kScanThreads / 64; ++k) s += ws[k];
    __syncthreads();
    return s;
}

// inclusive scan over the block; *total gets the block's sum
__device__ inline int scan_block_inclusive(int v, int *total)
{
    __shared__ int ws[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) ws[w] = v;
    __syncthreads();
    int before = 0, all = 0;
    for (int k = 0; k < kScanThreads / 64; ++k) {
        if (k < w) before += ws[k];
        all += ws[k];
    }
    __syncthreads();
    *total = all;
    return v + before;
}

// up to two independent scans in the same three launches (blockIdx.y picks
// one): out = exclusive scan of in[0..n), out[n] = the total; part holds
// kScanBlocks + 1 ints of each scan's partials
struct ScanJobs {
    Count cnt[2];
    const int *in[2];
    int *part[2];
    int *out[2];
};

// part[b] = sum of block b's contiguous chunk of in[0..n)
__global__ void k_scan_partial(ScanJobs j)
{
    const int y = blockIdx.y;
    const long n = j.cnt[y].get();
    const int *__restrict__ in = j.in[y];
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    int s = 0;
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) s += in[i];
    s = scan_block_sum(s);
    if (threadIdx.x == 0) j.part[y][blockIdx.x] = s;
}

// one block per scan: exclusive scan of part[0..nb), total into part[nb]
__global__ void k_scan_top(ScanJobs j, int nb)
{
    int *part = j.part[blockIdx.y];
    int run = 0;
    for (int base = 0; base < nb; base += blockDim.x) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? part[i] : 0;
        int tot;
        const int inc = scan_block_inclusive(v, &tot);
        if (i < nb) part[i] = run + inc - v;
        run += tot;
    }
    if (threadIdx.x == 0) part[nb] = run;
}

// out[i] = in[0] + ... + in[i-1] for i < n, out[n] = the total
__global__ void k_scan_final(ScanJobs j)
{
    const int y = blockIdx.y;
    const long n = j.cnt[y].get();
    const int *__restrict__ in = j.in[y];
    const int *__restrict__ part = j.part[y];
    int *out = j.out[y];
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    int run = part[blockIdx.x];
    for (long base = lo; base < hi; base += blockDim.x) {
        const long i = base + threadIdx.x;
        const int v = i < hi ? in[i] : 0;
        int tot;
        const int inc = scan_block_inclusive(v, &tot);
        if (i < hi) out[i] = run + inc - v;
        run += tot;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = part[gridDim.x];
}

// the three launches of `jobs` scans (1 or 2) on stream s
inline void enqueue_scans(const ScanJobs &j, int jobs, hipStream_t s)
{
    const dim3 tb(kScanThreads), gb(kScanBlocks, jobs);
    hipLaunchKernelGGL(k_scan_partial, gb, tb, 0, s, j);
    hipLaunchKernelGGL(k_scan_top, dim3(1, jobs), tb, 0, s, j, kScanBlocks);
    hipLaunchKernelGGL(k_scan_final, gb, tb, 0, s, j);
}

// one scan; part holds kScanBlocks + 1 ints
inline void enqueue_scan(Count cnt, const int *in, int *part, int *out, hipStream_t s)
{
    ScanJobs j{{cnt, cnt}, {in, in}, {part, part}, {out, out}};
    enqueue_scans(j, 1, s);
}

}  // namespace
}  // namespace ldsf

#endif  // LDSF_SF_SCAN_H_
