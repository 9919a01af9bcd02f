// Device-resident serving loop for descriptor + dE/dr requests.
//
// A request names device buffers: its species, neighbour counts, neighbour
// list, coordinates and dE/dzeta, and the outputs desc[n][width] and
// d_coordinates[n_total][3] (both overwritten).  The server holds a few
// *lanes*; a lane is a set of device buffers sized for one batch (centres,
// coordinate rows, neighbour-list entries, requests), a HIP stream and one
// instantiated hipGraph per neighbour-count bucket.  A worker thread admits
// queued requests in arrival order into the next free lane while they fit its
// capacities (the lanes together are sized from the HBM budget), writes the
// batch's table into pinned memory and replays the lane's graph; requests that
// arrive meanwhile wait only for a lane to free up, and lanes run concurrently
// on their own streams.  One completion thread per lane waits for the lane's
// event, reads the per-request flags and releases the lane.
//
// The graphs of a lane (one per bucket, all captured when the server is
// created), replayed for every batch (sizes are read on the device
// from the table's header, grids are fixed by the lane's capacity):
//   table H2D -> prepare (one block per tile of a request's rows, centres or
//   neighbour entries: copy and validate them into the lane, offset to batch
//   rows; per centre the addresses of its descriptor and dE/dzeta rows in the
//   request's own buffers; the transposed list's entry counts per bucket of
//   256 rows) -> scans of the neighbour counts and the bucket counts ->
//   CSR offsets per request, neighbour sums check -> transposed CSR (entries
//   scattered into their buckets, then one block per bucket: in_start, each
//   row's ids sorted, the inverse) -> the fused forward +
//   adjoint kernel (descriptors straight into each request's desc; the
//   general kernels of the largest bucket run as forward, then adjoint) ->
//   gather straight into each request's d_coordinates -> flags D2H.
// No memset nodes (zeroing is done by kernels).  A request with an index,
// species code or neighbour count out of range is flagged on the device; the
// kernels skip its centres, so nothing reads out of bounds, and its wait
// returns LDSF_ERR_NEIGHBOR.  Each request's results are bitwise those of the
// same request run alone: every centre is computed on its own and every row's
// dE/dr records are summed in ascending edge order, whatever else is in the
// batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ldsf.h"
#include "ldsf_internal.h"
#include "sf_device.h"
#include "sf_kernels.h"
#include "sf_bucket.h"
#include "sf_scan.h"

namespace ldsf {

namespace {

// one request of a batch, as the device sees it
struct Seg {
    const int *species;
    const int *numneigh;
    const int *nl;
    const double *coords;
    const double *dE;
    double *desc;
    double *dcoords;
    int n_atoms, n_total, n_neigh, max_nn;
    int centre_base, row_base, edge_base, pad;
};

// a block's share of k_prepare: [begin, end) of one request's rows, centres
// or neighbour entries
struct Tile {
    int r, kind, begin, end;
};
enum { kTileRows = 0, kTileCentres = 1, kTileEdges = 2 };

// header of the table: batch sizes
enum { H_CENTERS = 0, H_ROWS = 1, H_EDGES = 2, H_REQ = 3, H_TILES = 4, H_BUCKETS = 5,
       H_INTS = 16 };

constexpr int kThreads = 256;
constexpr int kPerThread = 8;
constexpr int kTile = kThreads * kPerThread;
// The transposed list is built bucket by bucket (sf_bucket.h): a bucket is
// kBucketRows consecutive batch rows.  A tile counts its entries per bucket
// in LDS (kHistCap buckets: requests of up to 1M rows; past that, per entry
// in global memory), so the global atomics are one per tile and bucket.
constexpr int kGridStride = 2048;
constexpr int kBuckets = 9;        // wave kernels for max_nn <= 8, 16, ..., 64; general kernels
                                   // (max_nn > 64, or a set the wave kernels do not cover)
constexpr int kGeneralBucket = 8;
// the wave kernels' grid: 32 one-wave blocks per CU, each a run of a few
// centres of the batch (a batch of ~60,000 centres: runs of ~7), so that the
// last runs of a batch leave little of the GPU idle (measured against 4,096
// and 12,288 blocks: 2 % more offline throughput, a 6 % lower median latency)
constexpr int kWaveBlocks = 8192;

__device__ __forceinline__ long gtid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ long gstride() { return (long)gridDim.x * blockDim.x; }

// the LDS histogram of a tile's entries over its request's buckets
// [b0, b0 + nbk): zeroed, counted (valid[k]: entry k is real), ready for the
// caller after the barrier
__device__ __forceinline__ void tile_hist(int *hist, int b0, int nbk, const int *tg, const bool *valid)
{
    for (int j = threadIdx.x; j < nbk; j += kThreads) hist[j] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPerThread; ++k)
        if (valid[k]) atomicAdd(&hist[(tg[k] >> kBucketShift) - b0], 1);
    __syncthreads();
}

// flags: 1 index/species/count out of range, 2 neighbour counts do not sum
// to n_neigh
// One tile per block: up to kTile rows, centres or neighbour entries of one
// request, so that a large request spreads over many blocks and every block
// has enough independent loads in flight.  The centre and entry tiles add
// their entries' counts per bucket of target rows into bcount (the entries
// are the transposed list's: an edge points at its neighbour's row, a
// centre's self-term at the centre's own; bcount starts at zero).
__global__ void k_prepare(const int *__restrict__ hdr, const Seg *__restrict__ seg,
                          const Tile *__restrict__ tiles, int width, int n_species, int *species,
                          int *numneigh, int *nl, int *centres, int *req_of, int *row_req,
                          double *x, double **out_row, const double **in_row,
                          const double *zero_row, int *req_sum, int *flags, int *bcount)
{
    __shared__ int hist[kHistCap];
    const int nt = hdr[H_TILES];
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile T = tiles[t];
        const int r = T.r;
        const Seg &S = seg[r];
        const int rb = S.row_base;
        int bad = 0, sum = 0;
        int tg[kPerThread];
        bool valid[kPerThread];
        if (T.kind == kTileRows) {
            const int *sp_in = S.species;
            const double *xin = S.coords;
#pragma unroll
            for (int k = 0; k < kPerThread; ++k) {
                const long i = T.begin + threadIdx.x + k * kThreads;
                if (i >= T.end) break;
                int sp = sp_in[i];
                if (sp < 0 || sp >= n_species) {
                    bad = 1;
                    sp = 0;
                }
                const long a = rb + i;
                species[a] = sp;
                row_req[a] = r;
                x[3 * a] = xin[3 * i];
                x[3 * a + 1] = xin[3 * i + 1];
                x[3 * a + 2] = xin[3 * i + 2];
            }
        } else if (T.kind == kTileCentres) {
            const int cb = S.centre_base, mx = S.max_nn;
            double *dsc = S.desc;
            const double *dE = S.dE;
#pragma unroll
            for (int k = 0; k < kPerThread; ++k) {
                const long i = T.begin + threadIdx.x + k * kThreads;
                valid[k] = i < T.end;
                tg[k] = rb + (int)i;
                if (!valid[k]) continue;
                int nn = S.numneigh[i];
                if (nn < 0 || nn > mx) {
                    bad = 1;
                    nn = 0;
                }
                const long c = cb + i;
                numneigh[c] = nn;
                centres[c] = rb + (int)i;
                req_of[c] = r;
                out_row[c] = dsc + i * width;
                in_row[c] = dE ? dE + i * width : zero_row;
                sum += nn;
            }
        } else {
            const int *nl_in = S.nl;
            const int eb = S.edge_base, nt_r = S.n_total;
            int v[kPerThread];
#pragma unroll
            for (int k = 0; k < kPerThread; ++k) {
                const long e = T.begin + threadIdx.x + k * kThreads;
                valid[k] = e < T.end;
                v[k] = valid[k] ? nl_in[e] : 0;
            }
#pragma unroll
            for (int k = 0; k < kPerThread; ++k) {
                if (valid[k] && (v[k] < 0 || v[k] >= nt_r)) {
                    bad = 1;
                    v[k] = 0;
                }
                tg[k] = rb + v[k];
                const long e = T.begin + threadIdx.x + k * kThreads;
                if (valid[k]) nl[eb + e] = tg[k];
            }
        }
        if (T.kind != kTileRows) {
            // the tile's counts per bucket of target rows
            const int b0 = rb >> kBucketShift;
            const int nbk = ((rb + S.n_total - 1) >> kBucketShift) - b0 + 1;
            if (nbk <= kHistCap) {
                tile_hist(hist, b0, nbk, tg, valid);
                for (int j = threadIdx.x; j < nbk; j += kThreads)
                    if (hist[j]) atomicAdd(bcount + b0 + j, hist[j]);
                __syncthreads();  // hist is the next tile's
            } else {
#pragma unroll
                for (int k = 0; k < kPerThread; ++k)
                    if (valid[k]) atomicAdd(bcount + (tg[k] >> kBucketShift), 1);
            }
        }
        // one atomic per wave for the neighbour-count sum
        if (T.kind == kTileCentres) {
            for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
            if ((threadIdx.x & 63) == 0 && sum) atomicAdd(req_sum + r, sum);
        }
        if (bad) atomicOr(flags + r, 1);
    }
}

// CSR offsets of every request from its own edge base; block 0 also flags
// the requests whose neighbour counts do not sum to n_neigh
__global__ void k_offsets(const int *__restrict__ hdr, const Seg *__restrict__ seg,
                          const int *__restrict__ req_of, const int *__restrict__ scan, int *offs,
                          const int *__restrict__ req_sum, int *flags)
{
    if (blockIdx.x == 0)
        for (int r = threadIdx.x; r < hdr[H_REQ]; r += blockDim.x)
            if (req_sum[r] != seg[r].n_neigh) atomicOr(flags + r, 2);
    const int n = hdr[H_CENTERS];
    for (long c = gtid(); c < n; c += gstride()) {
        const Seg &S = seg[req_of[c]];
        offs[c] = S.edge_base + scan[c] - scan[S.centre_base];
    }
    if (gtid() == 0) offs[n] = hdr[H_EDGES];
}

// ---- transposed CSR: the edges (and centre self-terms) pointing at each row ----

// bcount[0..n_buckets] and bcursor[0..n_buckets) to zero, with the
// per-request sums and flags (Q entries each): every counter a batch starts
// from zero, in one launch
__global__ void k_zero_rows(const int *__restrict__ hdr, int *bcount, int *bcursor, int *req_sum,
                            int *flags, int Q)
{
    const long nb = hdr[H_BUCKETS], n = max(nb + 1, (long)Q);
    for (long i = gtid(); i < n; i += gstride()) {
        if (i <= nb) bcount[i] = 0;
        if (i < nb) bcursor[i] = 0;
        if (i < Q) {
            req_sum[i] = 0;
            flags[i] = 0;
        }
    }
}

// Every entry of the transposed list (id: edge e, or n_edges + centre for a
// self-term; target row) into its bucket's range [bbase[b], bbase[b + 1]) of
// stage: each tile claims a run per bucket with one atomic, and its entries
// take the run's slots in an arbitrary order (the per-bucket pass sorts).
__global__ void k_scatter_t(const int *__restrict__ hdr, const Seg *__restrict__ seg,
                            const Tile *__restrict__ tiles, const int *__restrict__ nl,
                            const int *__restrict__ bbase, int *bcursor, int2 *stage)
{
    __shared__ int hist[kHistCap];
    const int nt = hdr[H_TILES];
    const long ne = hdr[H_EDGES];
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile T = tiles[t];
        if (T.kind == kTileRows) continue;
        const Seg &S = seg[T.r];
        const int rb = S.row_base;
        int id[kPerThread], tg[kPerThread];
        bool valid[kPerThread];
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const long i = T.begin + threadIdx.x + k * kThreads;
            valid[k] = i < T.end;
            if (T.kind == kTileCentres) {
                id[k] = (int)(ne + S.centre_base + i);
                tg[k] = rb + (int)i;
            } else {
                id[k] = (int)(S.edge_base + i);
                tg[k] = valid[k] ? nl[S.edge_base + i] : rb;
            }
        }
        const int b0 = rb >> kBucketShift;
        const int nbk = ((rb + S.n_total - 1) >> kBucketShift) - b0 + 1;
        if (nbk <= kHistCap) {
            tile_hist(hist, b0, nbk, tg, valid);
            for (int j = threadIdx.x; j < nbk; j += kThreads)
                if (hist[j]) hist[j] = bbase[b0 + j] + atomicAdd(bcursor + b0 + j, hist[j]);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kPerThread; ++k)
                if (valid[k]) {
                    const int slot = atomicAdd(&hist[(tg[k] >> kBucketShift) - b0], 1);
                    stage[slot] = make_int2(id[k], tg[k]);
                }
            __syncthreads();  // hist is the next tile's
        } else {
#pragma unroll
            for (int k = 0; k < kPerThread; ++k)
                if (valid[k]) {
                    const int b = tg[k] >> kBucketShift;
                    stage[bbase[b] + atomicAdd(bcursor + b, 1)] = make_int2(id[k], tg[k]);
                }
        }
    }
}

// One block per bucket (the pass of sf_bucket.h), sizes from the table
__global__ __launch_bounds__(kBucketThreads) void k_bucket_t(const int *__restrict__ hdr,
                                                            const int2 *__restrict__ stage,
                                                            const int *__restrict__ bbase,
                                                            int *in_start, int *in_edges,
                                                            int *inv_pos)
{
    __shared__ BucketLds sh;
    const int R = hdr[H_ROWS], nb = hdr[H_BUCKETS];
    for (int b = blockIdx.x; b < nb; b += gridDim.x)
        bucket_pass(b, nb, R, stage, bbase, in_start, in_edges, inv_pos, sh);
}

// per-row sums of the records, written (=) into the owning request's output
__global__ void k_gather_out(const int *__restrict__ hdr, const Seg *__restrict__ seg,
                             const int *__restrict__ row_req, const int *__restrict__ in_start,
                             const double *__restrict__ edge)
{
    const int n = hdr[H_ROWS];
    for (long a = gtid(); a < n; a += gstride()) {
        const Seg &S = seg[row_req[a]];
        if (!S.dcoords) continue;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        row_sum3(edge, in_start[a], in_start[a + 1], s0, s1, s2);
        double *o = S.dcoords + 3 * (a - S.row_base);
        o[0] = s0;
        o[1] = s1;
        o[2] = s2;
    }
}

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

}  // namespace ldsf

struct ldsf_ticket_rec {
    long long id = 0;
    ldsf_request req{};
    double t_submit = 0, t_done = 0;
    int status = LDSF_OK;
    bool done = false;
};

struct ldsf_lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // device
    char *d_table = nullptr;  // header + Seg[max_req]
    int *species = nullptr, *numneigh = nullptr, *nl = nullptr, *centres = nullptr;
    int *req_of = nullptr, *row_req = nullptr, *req_sum = nullptr, *flags = nullptr;
    int *scan = nullptr, *offs = nullptr, *in_start = nullptr;
    int *bcount = nullptr, *bbase = nullptr, *bcursor = nullptr;  // per bucket of rows
    int *in_edges = nullptr, *inv_pos = nullptr, *part = nullptr;
    double *x = nullptr, *edge = nullptr, *zero_row = nullptr;
    double **out_row = nullptr;
    const double **in_row = nullptr;
    // pinned host
    char *h_table = nullptr;
    int *h_flags = nullptr;
    hipGraphExec_t exec[ldsf::kBuckets] = {};
    std::vector<ldsf_ticket_rec *> batch;
    bool busy = false, launched = false;
};

struct ldsf_server {
    const ldsf_sf *sf = nullptr;
    ldsf_server_config cfg{};
    int device = 0;
    int width = 0, n_species = 0;
    ldsf::SFTab tab{};
    ldsf::WaveTab wave{};
    ldsf::WaveParams *wave_par = nullptr;
    bool wave_ok = false;
    size_t table_bytes = 0;
    long max_tiles = 0;  // k_prepare's tiles a batch can need
    std::vector<ldsf_lane> lanes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_lane;
    std::deque<ldsf_ticket_rec *> queue;
    std::unordered_map<long long, ldsf_ticket_rec *> tickets;
    long long next_id = 1;
    long long queued_centers = 0;
    long long n_batches = 0, n_requests = 0, n_atoms = 0, n_graphs = 0;
    bool stop = false;
    std::thread worker;
    std::vector<std::thread> completers;
    int worker_error = LDSF_OK;
};

namespace {

using ldsf::Seg;

int hip_fail(hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return ldsf::fail(LDSF_ERR_DEVICE, m.c_str());
}

#define SRV_TRY(expr, what)                                  \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(e_, what);     \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, long n)
{
    return hipMalloc((void **)p, sizeof(T) * (size_t)std::max<long>(n, 1));
}

// sum over a batch's requests of ceil(n / kTile) for rows, centres and
// entries: at most the lane's totals over kTile plus one partial tile per
// request and kind
long max_tiles(const ldsf_server_config &c)
{
    return ((long)c.lane_rows + c.lane_centers + c.lane_edges) / ldsf::kTile + 3L * c.max_requests +
           3;
}

// buckets of kBucketRows rows a lane's rows span
long n_buckets(const ldsf_server_config &c)
{
    return ((long)c.lane_rows + ldsf::kBucketRows - 1) >> ldsf::kBucketShift;
}

size_t lane_bytes(const ldsf_server_config &c, int width)
{
    const size_t C = c.lane_centers, R = c.lane_rows, E = c.lane_edges, Q = c.max_requests;
    size_t b = 0;
    b += sizeof(int) * (R /*species*/ + C /*numneigh*/ + E /*nl*/ + C /*centres*/ + C /*req_of*/ +
                        R /*row_req*/ + 2 * Q /*sums, flags*/ + (C + 1) * 2 /*scan, offs*/ +
                        (R + 1) /*in_start*/ + 3 * (n_buckets(c) + 1) /*bcount, bbase, bcursor*/ +
                        2 * (E + C) + 2 * (ldsf::kScanBlocks + 1));
    b += sizeof(double) * (3 * R + 2 * C + (size_t)width + 3 * (E + C));
    b += ldsf::H_INTS * sizeof(int) + Q * sizeof(Seg) + sizeof(ldsf::Tile) * (size_t)max_tiles(c);
    return b;
}

int lane_alloc(ldsf_server *S, ldsf_lane &L)
{
    const ldsf_server_config &c = S->cfg;
    const long C = c.lane_centers, R = c.lane_rows, E = c.lane_edges, Q = c.max_requests;
    SRV_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking), "hipStreamCreate");
    SRV_TRY(hipEventCreateWithFlags(&L.done, hipEventDisableTiming), "hipEventCreate");
    SRV_TRY(hipMalloc((void **)&L.d_table, S->table_bytes), "hipMalloc(table)");
    SRV_TRY(dalloc(&L.species, R), "hipMalloc");
    SRV_TRY(dalloc(&L.numneigh, C), "hipMalloc");
    SRV_TRY(dalloc(&L.nl, E), "hipMalloc");
    SRV_TRY(dalloc(&L.centres, C), "hipMalloc");
    SRV_TRY(dalloc(&L.req_of, C), "hipMalloc");
    SRV_TRY(dalloc(&L.row_req, R), "hipMalloc");
    SRV_TRY(dalloc(&L.req_sum, Q), "hipMalloc");
    SRV_TRY(dalloc(&L.flags, Q), "hipMalloc");
    SRV_TRY(dalloc(&L.scan, C + 1), "hipMalloc");
    SRV_TRY(dalloc(&L.offs, C + 1), "hipMalloc");
    SRV_TRY(dalloc(&L.bcount, n_buckets(c) + 1), "hipMalloc");
    SRV_TRY(dalloc(&L.bbase, n_buckets(c) + 1), "hipMalloc");
    SRV_TRY(dalloc(&L.bcursor, n_buckets(c) + 1), "hipMalloc");
    SRV_TRY(dalloc(&L.in_start, R + 1), "hipMalloc");
    SRV_TRY(dalloc(&L.in_edges, E + C), "hipMalloc");
    SRV_TRY(dalloc(&L.inv_pos, E + C), "hipMalloc");
    SRV_TRY(dalloc(&L.part, 2 * (ldsf::kScanBlocks + 1)), "hipMalloc");
    SRV_TRY(dalloc(&L.x, 3 * R), "hipMalloc");
    SRV_TRY(dalloc(&L.out_row, C), "hipMalloc");
    SRV_TRY(dalloc(&L.in_row, C), "hipMalloc");
    SRV_TRY(dalloc(&L.zero_row, S->width), "hipMalloc");
    SRV_TRY(hipMemset(L.zero_row, 0, sizeof(double) * std::max(S->width, 1)), "hipMemset");
    SRV_TRY(dalloc(&L.edge, 3 * (E + C)), "hipMalloc");
    SRV_TRY(hipHostMalloc((void **)&L.h_table, S->table_bytes, hipHostMallocDefault),
            "hipHostMalloc(table)");
    SRV_TRY(hipHostMalloc((void **)&L.h_flags, sizeof(int) * Q, hipHostMallocDefault),
            "hipHostMalloc(flags)");
    return LDSF_OK;
}

void lane_free(ldsf_lane &L)
{
    for (auto &g : L.exec)
        if (g) (void)hipGraphExecDestroy(g);
    void *dev[] = {L.d_table, L.species, L.numneigh, L.nl, L.centres, L.req_of, L.row_req,
                   L.req_sum, L.flags, L.scan, L.offs, L.bcount, L.bbase, L.bcursor, L.in_start,
                   L.in_edges, L.inv_pos, L.part, L.x, L.edge, L.zero_row, L.out_row,
                   (void *)L.in_row};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (L.h_table) (void)hipHostFree(L.h_table);
    if (L.h_flags) (void)hipHostFree(L.h_flags);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
}

// the batch sequence of a lane for one neighbour-count bucket, on its stream
hipError_t lane_enqueue(ldsf_server *S, ldsf_lane &L, int bucket)
{
    using namespace ldsf;
    const ldsf_server_config &c = S->cfg;
    hipStream_t s = L.stream;
    const int *hdr = reinterpret_cast<const int *>(L.d_table);
    const Seg *seg = reinterpret_cast<const Seg *>(L.d_table + H_INTS * sizeof(int));
    const int Q = c.max_requests;
    const dim3 tb(kThreads), gs(kGridStride);
    const Tile *tiles = reinterpret_cast<const Tile *>(L.d_table + H_INTS * sizeof(int) +
                                                       sizeof(Seg) * (size_t)Q);
    hipError_t e;
    if ((e = hipMemcpyAsync(L.d_table, L.h_table, S->table_bytes, hipMemcpyHostToDevice, s)))
        return e;
    // the bucket counters start at zero (k_prepare counts as it copies), and
    // so do the per-request sums and flags
    hipLaunchKernelGGL(k_zero_rows, gs, tb, 0, s, hdr, L.bcount, L.bcursor, L.req_sum, L.flags, Q);
    hipLaunchKernelGGL(k_prepare, gs, tb, 0, s, hdr, seg, tiles, S->width, S->n_species, L.species,
                       L.numneigh, L.nl, L.centres, L.req_of, L.row_req, L.x, L.out_row,
                       L.in_row, L.zero_row, L.req_sum, L.flags, L.bcount);
    // two scans in one set of launches: the neighbour counts (CSR offsets,
    // rebased per request by k_offsets) and the buckets' entry counts
    ScanJobs sj{{Count{hdr + H_CENTERS, 0}, Count{hdr + H_BUCKETS, 0}},
                {L.numneigh, L.bcount},
                {L.part, L.part + kScanBlocks + 1},
                {L.scan, L.bbase}};
    enqueue_scans(sj, 2, s);
    hipLaunchKernelGGL(k_offsets, gs, tb, 0, s, hdr, seg, L.req_of, L.scan, L.offs, L.req_sum,
                       L.flags);
    // transposed CSR: entries into their buckets (staged in the record
    // buffer, which the kernels below only write later), then per bucket:
    // in_start, the rows' ids sorted, inv_pos
    int2 *stage = reinterpret_cast<int2 *>(L.edge);
    hipLaunchKernelGGL(k_scatter_t, gs, tb, 0, s, hdr, seg, tiles, L.nl, L.bbase, L.bcursor, stage);
    hipLaunchKernelGGL(k_bucket_t, dim3(kGridStride), dim3(kBucketThreads), 0, s, hdr, stage,
                       L.bbase, L.in_start, L.in_edges, L.inv_pos);
    if ((e = hipGetLastError())) return e;
    DevCounts dc;
    dc.hdr = hdr;
    dc.req_of = L.req_of;
    dc.req_flag = L.flags;
    dc.max_blocks = kWaveBlocks;
    dc.out_row = L.out_row;
    dc.in_row = L.in_row;
    const int prec = c.precision;
    if (bucket == kGeneralBucket) {
        if ((e = launch_fwd_general(S->tab, prec, c.lane_centers, L.centres, L.offs, L.nl,
                                    L.species, L.x, nullptr, s, dc)))
            return e;
        if ((e = launch_adj_general(S->tab, prec, c.lane_centers, L.centres, L.offs, L.nl,
                                    L.species, L.x, nullptr, L.inv_pos, L.edge, c.lane_edges, s,
                                    dc)))
            return e;
    } else {
        const int maxn = 8 * (bucket + 1);
        // the fused kernel: descriptors (assigned through dc.out_row) and the
        // adjoint's records from one pass over the pairs and triplets
        if ((e = launch_adj_wave(S->wave, prec, c.lane_centers, L.centres, L.offs, L.nl,
                                 L.species, L.x, nullptr, L.inv_pos, L.edge, c.lane_edges, S->width,
                                 maxn, 64, s, dc)))
            return e;
    }
    hipLaunchKernelGGL(k_gather_out, gs, tb, 0, s, hdr, seg, L.row_req, L.in_start, L.edge);
    if ((e = hipGetLastError())) return e;
    return hipMemcpyAsync(L.h_flags, L.flags, sizeof(int) * Q, hipMemcpyDeviceToHost, s);
}

int lane_graph(ldsf_server *S, ldsf_lane &L, int bucket)
{
    if (L.exec[bucket]) return LDSF_OK;
    hipGraph_t g = nullptr;
    SRV_TRY(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal), "begin capture");
    hipError_t e = lane_enqueue(S, L, bucket);
    hipError_t e2 = hipStreamEndCapture(L.stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e != hipSuccess ? e : e2, "graph capture");
    }
    e = hipGraphInstantiate(&L.exec[bucket], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "graph instantiate");
    ++S->n_graphs;
    return LDSF_OK;
}

int bucket_for(const ldsf_server *S, int max_nn)
{
    if (!S->wave_ok || max_nn > 64) return ldsf::kGeneralBucket;
    return std::max(1, (max_nn + 7) / 8) - 1;  // 0..7: max_nn <= 8, 16, ..., 64
}

bool fits(const ldsf_server_config &c, long centres, long rows, long edges, int nreq)
{
    return centres <= c.lane_centers && rows <= c.lane_rows && edges <= c.lane_edges &&
           nreq <= c.max_requests;
}

void finish_batch(ldsf_server *S, ldsf_lane &L, int status)
{
    const double t = ldsf::now_s();
    for (size_t k = 0; k < L.batch.size(); ++k) {
        ldsf_ticket_rec *T = L.batch[k];
        T->t_done = t;
        if (status != LDSF_OK) T->status = status;
        else if (L.h_flags[k]) T->status = LDSF_ERR_NEIGHBOR;
        else T->status = LDSF_OK;
        T->done = true;
    }
    L.batch.clear();
    L.busy = false;
    L.launched = false;
}

void worker_main(ldsf_server *S)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        ldsf_lane *L = nullptr;
        // A free lane takes the queue when (a) the queue holds
        // dispatch_min_centers (> 0) centres, (b) its oldest request has
        // waited batch_window_us (> 0), (c) no window is set and either no
        // threshold is set or the device is idle, or (d) the server is
        // stopping -- so that batches grow with the load.
        const double window = 1e-6 * S->cfg.batch_window_us;
        for (;;) {
            if (S->queue.empty() && S->stop) break;
            bool free_lane = false;
            int running = 0;
            for (auto &l : S->lanes) {
                if (l.busy) ++running;
                else free_lane = true;
            }
            if (!S->queue.empty() && free_lane) {
                const double age = ldsf::now_s() - S->queue.front()->t_submit;
                const int min_c = S->cfg.dispatch_min_centers;
                if (S->stop || (min_c > 0 && S->queued_centers >= (long long)min_c) ||
                    (window > 0 && age >= window) ||
                    (window <= 0 && (running == 0 || min_c == 0)))
                    break;
                if (window > 0) {
                    S->cv_work.wait_for(lk, std::chrono::duration<double>(window - age));
                    continue;
                }
            }
            S->cv_work.wait(lk);
        }
        if (S->queue.empty() && S->stop) break;
        for (auto &l : S->lanes)
            if (!l.busy) {
                L = &l;
                break;
            }
        // admit in arrival order while the batch fits the lane.  A batch runs
        // one kernel family (the general kernels for neighbourhoods above 64,
        // else the wave kernels), and the two round differently: a request
        // of the other family closes the batch, so that every request is
        // computed by the kernels that would compute it alone.
        long C = 0, R = 0, E = 0;
        int maxnn = 0;
        bool general = false;
        while (!S->queue.empty()) {
            const ldsf_request &q = S->queue.front()->req;
            const bool qg = bucket_for(S, q.max_neighbors) == ldsf::kGeneralBucket;
            if (!L->batch.empty() &&
                (qg != general ||
                 !fits(S->cfg, C + q.n_atoms, R + q.n_total, E + q.n_neigh,
                       (int)L->batch.size() + 1)))
                break;
            general = qg;
            ldsf_ticket_rec *T = S->queue.front();
            S->queue.pop_front();
            S->queued_centers -= q.n_atoms;
            C += q.n_atoms;
            R += q.n_total;
            E += q.n_neigh;
            maxnn = std::max(maxnn, q.max_neighbors);
            L->batch.push_back(T);
        }
        L->busy = true;
        lk.unlock();
        // table: header + one Seg per request
        int *h = reinterpret_cast<int *>(L->h_table);
        std::memset(h, 0, ldsf::H_INTS * sizeof(int));
        Seg *sg = reinterpret_cast<Seg *>(L->h_table + ldsf::H_INTS * sizeof(int));
        int cb = 0, rb = 0, eb = 0;
        for (size_t k = 0; k < L->batch.size(); ++k) {
            const ldsf_request &q = L->batch[k]->req;
            Seg &g = sg[k];
            g.species = q.species;
            g.numneigh = q.number_of_neighs;
            g.nl = q.neighbor_list;
            g.coords = q.coordinates;
            g.dE = q.dE_dzeta;
            g.desc = q.desc;
            g.dcoords = q.d_coordinates;
            g.n_atoms = q.n_atoms;
            g.n_total = q.n_total;
            g.n_neigh = q.n_neigh;
            g.max_nn = q.max_neighbors;
            g.centre_base = cb;
            g.row_base = rb;
            g.edge_base = eb;
            g.pad = 0;
            cb += q.n_atoms;
            rb += q.n_total;
            eb += q.n_neigh;
        }
        // k_prepare's tiles: each request's rows, centres and entries in
        // chunks of kTile
        ldsf::Tile *tl = reinterpret_cast<ldsf::Tile *>(
            L->h_table + ldsf::H_INTS * sizeof(int) + sizeof(Seg) * (size_t)S->cfg.max_requests);
        int ntiles = 0;
        for (size_t k = 0; k < L->batch.size(); ++k) {
            const ldsf_request &q = L->batch[k]->req;
            const int n[3] = {q.n_total, q.n_atoms, q.n_neigh};
            for (int kind = 0; kind < 3; ++kind)
                for (int b = 0; b < n[kind]; b += ldsf::kTile)
                    tl[ntiles++] = {(int)k, kind, b, std::min(n[kind], b + ldsf::kTile)};
        }
        h[ldsf::H_CENTERS] = cb;
        h[ldsf::H_ROWS] = rb;
        h[ldsf::H_EDGES] = eb;
        h[ldsf::H_REQ] = (int)L->batch.size();
        h[ldsf::H_TILES] = ntiles;
        h[ldsf::H_BUCKETS] = (rb + ldsf::kBucketRows - 1) >> ldsf::kBucketShift;
        const int bucket = bucket_for(S, maxnn);
        int rc = lane_graph(S, *L, bucket);
        hipError_t e = hipSuccess;
        if (rc == LDSF_OK) {
            e = hipGraphLaunch(L->exec[bucket], L->stream);
            if (e == hipSuccess) e = hipEventRecord(L->done, L->stream);
            if (e != hipSuccess) rc = hip_fail(e, "graph launch");
        }
        lk.lock();
        S->n_batches += 1;
        S->n_requests += (long long)L->batch.size();
        S->n_atoms += cb;
        if (rc != LDSF_OK) {
            S->worker_error = rc;
            finish_batch(S, *L, rc);
            S->cv_done.notify_all();
            continue;
        }
        L->launched = true;
        S->cv_lane.notify_all();
    }
}

void completer_main(ldsf_server *S, size_t li)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    ldsf_lane &L = S->lanes[li];
    for (;;) {
        S->cv_lane.wait(lk, [&] { return L.launched || (S->stop && !L.busy && S->queue.empty()); });
        if (!L.launched) break;
        lk.unlock();
        const hipError_t e = hipEventSynchronize(L.done);
        lk.lock();
        finish_batch(S, L, e == hipSuccess ? LDSF_OK : hip_fail(e, "batch"));
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

LDSF_API int ldsf_server_create(ldsf_server **out, const ldsf_sf *sf,
                                const ldsf_server_config *config)
{
    if (!out || !sf || !config) return ldsf::fail(LDSF_ERR_ARG, "server_create: null argument");
    *out = nullptr;
    const ldsf_server_config &c = *config;
    if ((c.precision != LDSF_FP64 && c.precision != LDSF_FP32_BULK) || c.lane_centers < 1 ||
        c.lane_rows < c.lane_centers || c.lane_edges < 0 || c.max_requests < 1 ||
        c.max_requests > 65535 || c.max_lanes < 1 || c.dispatch_min_centers < 0 ||
        c.batch_window_us < 0 || (long)c.lane_edges + c.lane_centers > 0x7fffffffL ||
        (long)c.lane_rows * 3 > 0x7fffffffL)
        return ldsf::fail(LDSF_ERR_ARG, "server_create: bad configuration");
    auto *S = new ldsf_server();
    S->sf = sf;
    S->cfg = c;
    S->width = ldsf_sf_width(sf);
    S->n_species = ldsf_sf_num_species(sf);
    if (hipGetDevice(&S->device) != hipSuccess) {
        delete S;
        return ldsf::fail(LDSF_ERR_DEVICE, "server_create: no device");
    }
    int rc = ldsf::device_table(sf, &S->tab);
    if (rc) {
        delete S;
        return rc;
    }
    ldsf::WaveParams wp;
    // (FP32-bulk sets whose exponents leave the FP32 range run the general kernels)
    if (ldsf::wave_build(sf, &S->wave, &wp) && ldsf::wave_runs(S->wave, c.precision)) {
        if (hipMalloc((void **)&S->wave_par, sizeof(wp)) != hipSuccess ||
            hipMemcpy(S->wave_par, &wp, sizeof(wp), hipMemcpyHostToDevice) != hipSuccess) {
            if (S->wave_par) (void)hipFree(S->wave_par);
            delete S;
            return ldsf::fail(LDSF_ERR_DEVICE, "server_create: wave parameters");
        }
        S->wave.rc = S->tab.rc;
        S->wave.par = S->wave_par;
        S->wave_ok = true;
    }
    S->max_tiles = max_tiles(c);
    S->table_bytes = ldsf::H_INTS * sizeof(int) + sizeof(Seg) * (size_t)c.max_requests +
                     sizeof(ldsf::Tile) * (size_t)S->max_tiles;
    const size_t per = lane_bytes(c, S->width);
    const long n_lanes = std::min<long>(c.max_lanes, (long)(c.hbm_budget_bytes / per));
    if (n_lanes < 1) {
        if (S->wave_par) (void)hipFree(S->wave_par);
        delete S;
        return ldsf::fail(LDSF_ERR_NOMEM, "server_create: HBM budget below one lane");
    }
    S->lanes.resize(n_lanes);
    // every lane's graphs are captured here, before any request: a capture in
    // progress would make device-wide synchronisation in other threads fail
    for (auto &L : S->lanes) {
        rc = lane_alloc(S, L);
        for (int b = S->wave_ok ? 0 : ldsf::kGeneralBucket; !rc && b < ldsf::kBuckets; ++b)
            rc = lane_graph(S, L, b);
        if (rc) break;
    }
    if (rc) {
        for (auto &L : S->lanes) lane_free(L);
        if (S->wave_par) (void)hipFree(S->wave_par);
        delete S;
        return rc;
    }
    S->worker = std::thread(worker_main, S);
    for (size_t k = 0; k < S->lanes.size(); ++k) S->completers.emplace_back(completer_main, S, k);
    *out = S;
    return LDSF_OK;
}

LDSF_API int ldsf_server_num_lanes(const ldsf_server *S)
{
    return S ? (int)S->lanes.size() : 0;
}

LDSF_API int ldsf_server_submit(ldsf_server *S, const ldsf_request *req, long long *ticket)
{
    if (!S || !req || !ticket) return ldsf::fail(LDSF_ERR_ARG, "server_submit: null argument");
    const ldsf_request &q = *req;
    if (q.n_atoms < 0 || q.n_total < q.n_atoms || q.n_neigh < 0 || q.max_neighbors < 0 ||
        (q.n_total > 0 && (!q.species || !q.coordinates)) ||
        (q.n_atoms > 0 && (!q.number_of_neighs || !q.desc)) || (q.n_neigh > 0 && !q.neighbor_list))
        return ldsf::fail(LDSF_ERR_ARG, "server_submit: bad request");
    if (!fits(S->cfg, q.n_atoms, q.n_total, q.n_neigh, 1))
        return ldsf::fail(LDSF_ERR_ARG, "server_submit: request exceeds the lane capacity");
    if ((long)q.max_neighbors * q.n_atoms < q.n_neigh)
        return ldsf::fail(LDSF_ERR_ARG, "server_submit: max_neighbors below n_neigh / n_atoms");
    auto *T = new ldsf_ticket_rec();
    T->req = q;
    T->t_submit = ldsf::now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete T;
            return ldsf::fail(LDSF_ERR_ARG, "server_submit: server is stopping");
        }
        T->id = S->next_id++;
        S->tickets[T->id] = T;
        S->queue.push_back(T);
        S->queued_centers += q.n_atoms;
        *ticket = T->id;
    }
    S->cv_work.notify_one();
    return LDSF_OK;
}

LDSF_API int ldsf_server_wait(ldsf_server *S, long long ticket, double timeout_s,
                              double *latency_s)
{
    if (!S) return ldsf::fail(LDSF_ERR_ARG, "server_wait: null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return ldsf::fail(LDSF_ERR_ARG, "server_wait: unknown ticket");
    ldsf_ticket_rec *T = it->second;
    auto ready = [&] { return T->done; };
    if (timeout_s < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return ldsf::fail(LDSF_ERR_TIMEOUT, "server_wait: not finished");
    const int st = T->status;
    if (latency_s) *latency_s = T->t_done - T->t_submit;
    S->tickets.erase(it);
    delete T;
    if (st == LDSF_ERR_NEIGHBOR)
        return ldsf::fail(st, "server_wait: request rejected on the device (index, species "
                              "code or neighbour count out of range)");
    if (st != LDSF_OK) return st;
    return LDSF_OK;
}

LDSF_API int ldsf_server_stats(const ldsf_server *S, long long *batches, long long *requests,
                               long long *atoms, long long *graphs)
{
    if (!S) return ldsf::fail(LDSF_ERR_ARG, "server_stats: null server");
    auto *M = const_cast<ldsf_server *>(S);
    std::lock_guard<std::mutex> lk(M->mu);
    if (batches) *batches = S->n_batches;
    if (requests) *requests = S->n_requests;
    if (atoms) *atoms = S->n_atoms;
    if (graphs) *graphs = S->n_graphs;
    return LDSF_OK;
}

LDSF_API void ldsf_server_destroy(ldsf_server *S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_lane.notify_all();
    if (S->worker.joinable()) S->worker.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
    }
    S->cv_lane.notify_all();
    for (auto &t : S->completers)
        if (t.joinable()) t.join();
    for (auto &L : S->lanes) lane_free(L);
    if (S->wave_par) (void)hipFree(S->wave_par);
    for (auto &kv : S->tickets) delete kv.second;
    delete S;
}
