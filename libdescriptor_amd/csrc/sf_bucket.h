// The transposed neighbour list (the edges and centre self-terms pointing at
// each row, each row's ids ascending, and the inverse positions), built
// bucket by bucket: a bucket is kBucketRows consecutive rows.  The entries
// are first staged as (id, row) pairs in their bucket's range
// [bbase[b], bbase[b + 1]) of a staging array, in any order; then one block
// per bucket counts its rows in LDS, writes their in_start, places the ids
// per row (in LDS when the bucket holds at most kBucketCap entries, else in
// place in in_edges), sorts each row and writes the inverse.  Used by the
// serving loop's batches and by plans.
#ifndef LDSF_SF_BUCKET_H_
#define LDSF_SF_BUCKET_H_

#include <hip/hip_runtime.h>

#include "sf_device.h"

namespace ldsf {
namespace {

constexpr int kBucketShift = 8;
constexpr int kBucketRows = 1 << kBucketShift;
constexpr int kHistCap = 4096;     // buckets a tile counts in LDS
constexpr int kBucketThreads = 512;
constexpr int kBucketCap = 12288;  // ids of a bucket sorted in LDS (48 KiB); more: in place
constexpr int kBucketPer = 8;      // staged entries per thread and load round

struct BucketLds {
    int cnt[kBucketRows], start[kBucketRows + 1], wsum[kBucketRows / 64];
    int buf[kBucketCap];
};

// Bucket b of nb over R rows, by the whole block (kBucketThreads threads):
// its entries counted per row in LDS, in_start of its rows (and in_start[R]
// by the last bucket), the ids placed per row, each row sorted ascending with
// inv_pos[id] = its position, and the sorted ids written out.
__device__ inline void bucket_pass(int b, int nb, int R, const int2 *__restrict__ stage,
                                   const int *__restrict__ bbase, int *in_start, int *in_edges,
                                   int *inv_pos, BucketLds &sh)
{
    int *cnt = sh.cnt, *start = sh.start, *wsum = sh.wsum, *buf_l = sh.buf;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lo = bbase[b], hi = bbase[b + 1], m = hi - lo;
    const int r0 = b << kBucketShift, nr = min(kBucketRows, R - r0);
    if (threadIdx.x < kBucketRows) cnt[threadIdx.x] = 0;
    __syncthreads();
    // kBucketPer loads in flight per thread, then their counts
    for (int base = lo; base < hi; base += kBucketThreads * kBucketPer) {
        int row[kBucketPer];
#pragma unroll
        for (int k = 0; k < kBucketPer; ++k) {
            const int i = base + threadIdx.x + k * kBucketThreads;
            row[k] = i < hi ? stage[i].y - r0 : -1;
        }
#pragma unroll
        for (int k = 0; k < kBucketPer; ++k)
            if (row[k] >= 0) atomicAdd(&cnt[row[k]], 1);
    }
    __syncthreads();
    // exclusive scan of the row counts (the first kBucketRows threads)
    int incl = 0;
    if (threadIdx.x < kBucketRows) {
        incl = cnt[threadIdx.x];
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(incl, off);
            if (lane >= off) incl += u;
        }
        if (lane == 63) wsum[wv] = incl;
    }
    __syncthreads();
    if (threadIdx.x < kBucketRows) {
        int add = 0;
        for (int k = 0; k < wv; ++k) add += wsum[k];
        const int ex = add + incl - cnt[threadIdx.x];
        start[threadIdx.x] = ex;
        if (threadIdx.x == kBucketRows - 1) start[kBucketRows] = add + incl;
        if (threadIdx.x < nr) in_start[r0 + threadIdx.x] = lo + ex;
        cnt[threadIdx.x] = 0;  // the rows' cursors from here on
    }
    if (b == nb - 1 && threadIdx.x == 0) in_start[R] = hi;
    __syncthreads();
    const bool in_lds = m <= kBucketCap;
    int *buf = in_lds ? buf_l : in_edges + lo;
    for (int base = lo; base < hi; base += kBucketThreads * kBucketPer) {
        int2 q[kBucketPer];
#pragma unroll
        for (int k = 0; k < kBucketPer; ++k) {
            const int i = base + threadIdx.x + k * kBucketThreads;
            q[k] = i < hi ? stage[i] : make_int2(-1, -1);
        }
#pragma unroll
        for (int k = 0; k < kBucketPer; ++k)
            if (q[k].x >= 0) {
                const int r = q[k].y - r0;
                buf[start[r] + atomicAdd(&cnt[r], 1)] = q[k].x;
            }
    }
    __threadfence_block();
    __syncthreads();
    sort_rows_inverse(nr, start, buf, inv_pos, wv, kBucketThreads / 64, lane, lo);
    __threadfence_block();
    __syncthreads();
    if (in_lds)
        for (int i = threadIdx.x; i < m; i += kBucketThreads) in_edges[lo + i] = buf_l[i];
    __syncthreads();  // cnt, start and buf_l serve the block's next bucket
}

}  // namespace
}  // namespace ldsf

#endif  // LDSF_SF_BUCKET_H_
