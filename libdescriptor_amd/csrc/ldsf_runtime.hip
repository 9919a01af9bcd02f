// Host runtime of the C-ABI: per-device parameter tables, neighbour-list
// plans, and the reference-signature host-pointer entry points, which stage
// their arguments in device memory and run the same plans.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ldsf.h"
#include "ldsf_internal.h"
#include "sf_kernels.h"
#include "sf_stage.h"

namespace {

int hip_fail(hipError_t e, const char *where)
{
    return ldsf::fail(e == hipErrorOutOfMemory ? LDSF_ERR_NOMEM : LDSF_ERR_DEVICE,
                      std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, where)                            \
    do {                                                \
        hipError_t e__ = (expr);                        \
        if (e__ != hipSuccess) return hip_fail(e__, where); \
    } while (0)

// device buffer that frees itself
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, sizeof(T) * count);
    }
    T *release_ownership()
    {
        T *q = p;
        p = nullptr;
        n = 0;
        return q;
    }
};

}  // namespace

// frees a parameter table uploaded to `device` (ldsf_sf::Dev::release)
static void release_table(int device, void *mem)
{
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    (void)hipFree(mem);
    (void)hipSetDevice(cur);
}

namespace ldsf {

SFTab table_view(const ldsf_sf *sf, const double *dd, const int *di)
{
    const PackedTable &T = sf->packed;
    SFTab t{};
    t.n_species = (int)sf->species.size();
    t.width = sf->width;
    t.n2 = (int)sf->k2.size();
    t.n3 = (int)sf->k3.size();
    t.species_resolved = sf->species_resolved;
    t.rc = dd + T.o_rc;
    t.a2 = dd + T.o_a2;
    t.b2 = dd + T.o_b2;
    t.zeta3 = dd + T.o_z3;
    t.lam3 = dd + T.o_l3;
    t.eta3 = dd + T.o_e3;
    t.coef3 = dd + T.o_c3;
    t.k2 = di + T.o_k2;
    t.slot2 = di + T.o_s2;
    t.str2 = di + T.o_t2;
    t.k3 = di + T.o_k3;
    t.izeta3 = di + T.o_iz;
    t.slot3 = di + T.o_s3;
    t.str3 = di + T.o_t3;
    t.pidx2 = di + T.o_p2;
    t.pidx3 = di + T.o_p3;
    t.n_params = sf->n_params;
    return t;
}

int device_table(const ldsf_sf *csf, SFTab *out)
{
    auto *sf = const_cast<ldsf_sf *>(csf);
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    std::lock_guard<std::mutex> lk(sf->dev_mu);
    for (auto &d : sf->dev)
        if (d.device == device && d.kind == 0) {
            *out = d.tab;
            return LDSF_OK;
        }
    // one allocation: the packed doubles, then the packed ints
    const PackedTable &T = sf->packed;
    void *mem = nullptr;
    HIP_TRY(hipMalloc(&mem, T.bytes()), "hipMalloc(parameter table)");
    double *dd = (double *)mem;
    int *di = (int *)(dd + T.d.size() + 1);
    hipError_t e = hipSuccess;
    if (!T.d.empty()) e = hipMemcpy(dd, T.d.data(), sizeof(double) * T.d.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !T.i.empty())
        e = hipMemcpy(di, T.i.data(), sizeof(int) * T.i.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(mem);
        return hip_fail(e, "upload parameter table");
    }
    ldsf_sf::Dev d;
    d.device = device;
    d.mem = mem;
    d.tab = table_view(sf, dd, di);
    d.release = release_table;
    sf->dev.push_back(d);
    *out = d.tab;
    return LDSF_OK;
}

}  // namespace ldsf

struct ldsf_plan {
    const ldsf_sf *sf = nullptr;
    int device = 0;
    int n_centers = 0;
    int n_total = 0;
    long n_edges = 0;
    int max_nn = 0;
    int family = 0;
    bool wave_ok = false;
    ldsf::WaveTab wave{};
    int wave_group = 0;
    bool edges_ready = false;  // plan->edge holds the last adjoint's records
    const int *species = nullptr;  // borrowed
    const int *nl = nullptr;       // borrowed
    int *centers = nullptr;        // owned (NULL = identity)
    int *offs = nullptr;
    int *in_start = nullptr;
    int *in_edges = nullptr;
    int *inv_pos = nullptr;
    int *nlp = nullptr;            // wave kernels: the list with each neighbour's species code
    double *edge = nullptr;
    bool edge_owned = true;        // false: a caller's buffer (ldsf_plan_set_records)
    ldsf::WaveParams *wave_par = nullptr;
    bool borrowed = false;         // true: the arrays above live in a PlanArena

    ~ldsf_plan()
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(device);
        if (!borrowed)
            for (void *p : {(void *)centers, (void *)offs, (void *)in_start, (void *)in_edges,
                            (void *)inv_pos, (void *)nlp, (void *)wave_par})
                if (p) (void)hipFree(p);
        if (edge_owned && edge) (void)hipFree(edge);
        (void)hipSetDevice(cur);
    }
};

// Device arrays a plan can be built in instead of its own allocations: the
// host-pointer calls build one plan per call, in one arena per descriptor
// and device, which keeps the largest sizes asked of it.
struct PlanArena {
    ldsf::Grow<int> centers, offs, flags, count, in_start, in_edges, inv_pos, stage, nlp;
    ldsf::Grow<char> tmp;
    ldsf::Grow<double> edge;
    ldsf::Grow<ldsf::WaveParams> wpar;
};

namespace {

// an array of the plan: from the arena when there is one, else its own
template <typename T>
hipError_t take(DevBuf<T> &own, ldsf::Grow<T> *arena, size_t n, T **out)
{
    hipError_t e = arena ? arena->reserve(n) : own.alloc(n);
    *out = arena ? arena->p : own.p;
    return e;
}

// With an arena, centers (NULL = identity) lives in it and the plan's arrays
// are taken from it; without, the plan owns centers and allocates its own.
int plan_build(ldsf_plan **out, const ldsf_sf *sf, int n_centers, int *owned_centers, int n_total,
               const int *d_species, const int *d_nl, const int *d_numneigh, hipStream_t s,
               PlanArena *arena = nullptr)
{
    DevBuf<int> centers_guard;
    if (!arena) centers_guard.p = owned_centers;
    if (!sf) return ldsf::fail(LDSF_ERR_ARG, "plan_create: null descriptor");
    if (n_centers < 0 || n_total < 0 || (n_total > 0 && !d_species) ||
        (n_centers > 0 && !d_numneigh))
        return ldsf::fail(LDSF_ERR_ARG, "plan_create: bad arguments");
    ldsf::SFTab tab;
    if (int rc = ldsf::device_table(sf, &tab)) return rc;
    auto plan = new ldsf_plan;
    std::unique_ptr<ldsf_plan> guard(plan);
    HIP_TRY(hipGetDevice(&plan->device), "hipGetDevice");
    plan->borrowed = arena != nullptr;
    plan->sf = sf;
    plan->n_centers = n_centers;
    plan->n_total = n_total;
    plan->species = d_species;
    plan->nl = d_nl;
    plan->centers = arena ? owned_centers : centers_guard.release_ownership();

    DevBuf<int> offs, scratch_i, count;
    int *offs_p = nullptr, *flags_p = nullptr, *count_p = nullptr;
    HIP_TRY(take(offs, arena ? &arena->offs : nullptr, n_centers + 1, &offs_p), "hipMalloc(offsets)");
    HIP_TRY(take(scratch_i, arena ? &arena->flags : nullptr, 4, &flags_p), "hipMalloc(flags)");
    HIP_TRY(hipMemsetAsync(flags_p, 0, sizeof(int) * 4, s), "hipMemsetAsync");
    HIP_TRY(ldsf::launch_plan_validate(n_centers, plan->centers, n_total, d_numneigh, d_species,
                                       tab.n_species, flags_p, s), "validate");
    size_t tmp_bytes = 0, tmp2 = 0;
    HIP_TRY(ldsf::launch_scan_int(n_centers, d_numneigh, offs_p, nullptr, &tmp_bytes, s), "scan size");
    HIP_TRY(ldsf::launch_scan_int(n_total, d_numneigh, offs_p, nullptr, &tmp2, s), "scan size");
    tmp_bytes = std::max(tmp_bytes, tmp2) + 256;
    DevBuf<char> tmp;
    char *tmp_p = nullptr;
    HIP_TRY(take(tmp, arena ? &arena->tmp : nullptr, tmp_bytes, &tmp_p), "hipMalloc(scan scratch)");
    // a negative count would make the scan meaningless; it is caught below
    HIP_TRY(ldsf::launch_scan_int(n_centers, d_numneigh, offs_p, tmp_p, &tmp_bytes, s), "scan");
    HIP_TRY(ldsf::launch_max_int(n_centers, d_numneigh, flags_p + 1, s), "max");
    int host_flags[4] = {0, 0, 0, 0};
    int total = 0;
    HIP_TRY(hipMemcpyAsync(host_flags, flags_p, sizeof(int) * 4, hipMemcpyDeviceToHost, s), "copy flags");
    HIP_TRY(hipMemcpyAsync(&total, offs_p + n_centers, sizeof(int), hipMemcpyDeviceToHost, s), "copy total");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (host_flags[0] & 2) return ldsf::fail(LDSF_ERR_ARG, "negative number_of_neighs entry");
    if (host_flags[0] & 8) return ldsf::fail(LDSF_ERR_ARG, "centre row out of range");
    if (host_flags[0] & 4) return ldsf::fail(LDSF_ERR_ARG, "species code out of range");
    if (total < 0) return ldsf::fail(LDSF_ERR_UNSUPPORTED, "more than 2^31-1 neighbour-list entries");
    if (total > 0 && !d_nl) return ldsf::fail(LDSF_ERR_ARG, "plan_create: null neighbour list");
    plan->n_edges = total;
    plan->max_nn = host_flags[1];

    // transposed list; out-of-range targets are left out and flagged, and the
    // flag is read once the list is built
    HIP_TRY(hipMemsetAsync(flags_p, 0, sizeof(int), s), "hipMemsetAsync");
    DevBuf<int> in_start, in_edges, inv_pos, stage;
    int *in_start_p = nullptr, *in_edges_p = nullptr, *inv_pos_p = nullptr, *stage_p = nullptr;
    const long n_ent = plan->n_edges + n_centers;
    HIP_TRY(take(in_start, arena ? &arena->in_start : nullptr, n_total + 1, &in_start_p),
            "hipMalloc(in_start)");
    HIP_TRY(take(in_edges, arena ? &arena->in_edges : nullptr, n_ent, &in_edges_p),
            "hipMalloc(in_edges)");
    HIP_TRY(take(inv_pos, arena ? &arena->inv_pos : nullptr, n_ent, &inv_pos_p),
            "hipMalloc(inv_pos)");
    if (ldsf::bucketed_fits(n_total)) {
        // by buckets of rows: one global atomic per tile and bucket
        const size_t nb1 = (size_t)ldsf::bucket_count_of(n_total) + 1;
        HIP_TRY(take(count, arena ? &arena->count : nullptr, 3 * nb1, &count_p),
                "hipMalloc(bucket counts)");
        HIP_TRY(take(stage, arena ? &arena->stage : nullptr, 2 * (size_t)n_ent, &stage_p),
                "hipMalloc(staging)");
        HIP_TRY(ldsf::launch_transpose_bucketed(plan->n_edges, n_centers, d_nl, plan->centers,
                                                n_total, count_p, count_p + nb1,
                                                count_p + 2 * nb1, reinterpret_cast<int2 *>(stage_p),
                                                in_start_p, in_edges_p, inv_pos_p, flags_p, tmp_p,
                                                tmp_bytes, s),
                "transposed list");
    } else {
        // more than 4096 buckets: count, scan, fill with per-entry atomics, sort
        HIP_TRY(take(count, arena ? &arena->count : nullptr, n_total + 1, &count_p),
                "hipMalloc(counts)");
        HIP_TRY(hipMemsetAsync(count_p, 0, sizeof(int) * (n_total + 1), s), "hipMemsetAsync");
        HIP_TRY(ldsf::launch_edge_targets_count(plan->n_edges, n_centers, d_nl, plan->centers,
                                                n_total, count_p, flags_p, s), "count");
        HIP_TRY(ldsf::launch_scan_int(n_total, count_p, in_start_p, tmp_p, &tmp_bytes, s), "scan");
        HIP_TRY(hipMemsetAsync(count_p, 0, sizeof(int) * (n_total + 1), s), "hipMemsetAsync");
        HIP_TRY(ldsf::launch_edge_targets_fill(plan->n_edges, n_centers, d_nl, plan->centers,
                                               n_total, in_start_p, count_p, in_edges_p, s), "fill");
        HIP_TRY(ldsf::launch_sort_inverse(n_total, in_start_p, in_edges_p, inv_pos_p, s),
                "sort and inverse");
    }
    HIP_TRY(hipMemcpyAsync(host_flags, flags_p, sizeof(int), hipMemcpyDeviceToHost, s), "copy flags");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (host_flags[0] & 1) return ldsf::fail(LDSF_ERR_ARG, "neighbour index out of range [0, n_total)");
    plan->offs = arena ? offs_p : offs.release_ownership();
    plan->in_start = arena ? in_start_p : in_start.release_ownership();
    plan->in_edges = arena ? in_edges_p : in_edges.release_ownership();
    plan->inv_pos = arena ? inv_pos_p : inv_pos.release_ownership();
    // wave kernels: parameter set covered and the largest neighbourhood fits in LDS
    ldsf::WaveParams wpar;
    int group = 0;
    if (ldsf::wave_build(sf, &plan->wave, &wpar) &&
        (group = ldsf::wave_group_size(plan->max_nn, plan->wave.n_species)) > 0) {
        plan->wave_group = group;
        if (arena) {
            HIP_TRY(arena->wpar.reserve(1), "hipMalloc(wave parameters)");
            plan->wave_par = arena->wpar.p;
        } else {
            HIP_TRY(hipMalloc((void **)&plan->wave_par, sizeof(wpar)), "hipMalloc(wave parameters)");
        }
        HIP_TRY(hipMemcpy(plan->wave_par, &wpar, sizeof(wpar), hipMemcpyHostToDevice),
                "hipMemcpy(wave parameters)");
        plan->wave.rc = tab.rc;
        plan->wave.par = plan->wave_par;
        plan->wave_ok = true;
        // each entry with its neighbour's species code: the wave kernels skip
        // the per-neighbour species gather (the list is validated above)
        if (n_total <= ldsf::kNlIndexMask + 1 && plan->n_edges > 0) {
            if (arena) {
                HIP_TRY(arena->nlp.reserve(plan->n_edges), "hipMalloc(packed list)");
                plan->nlp = arena->nlp.p;
            } else {
                HIP_TRY(hipMalloc((void **)&plan->nlp, sizeof(int) * plan->n_edges),
                        "hipMalloc(packed list)");
            }
            HIP_TRY(ldsf::launch_pack_species(plan->n_edges, d_nl, d_species, plan->nlp, s),
                    "packed list");
            plan->wave.nl_species = 1;
        }
    }
    plan->family = plan->wave_ok ? 1 : 0;
    *out = guard.release();
    return LDSF_OK;
}

// the neighbour list the wave kernels read: the packed one when the plan has it
const int *wave_nl(const ldsf_plan *plan) { return plan->nlp ? plan->nlp : plan->nl; }

int plan_compute(ldsf_plan *plan, const double *x, double *desc, int precision, hipStream_t s)
{
    ldsf::SFTab tab;
    if (int rc = ldsf::device_table(plan->sf, &tab)) return rc;
    hipError_t e;
    if (plan->family == 1 && plan->wave_ok && ldsf::wave_runs(plan->wave, precision))
        e = ldsf::launch_fwd_wave(plan->wave, precision, plan->n_centers, plan->centers,
                                  plan->offs, wave_nl(plan), plan->species, x, desc, tab.width,
                                  plan->max_nn, plan->wave_group, s);
    else
        e = ldsf::launch_fwd_general(tab, precision, plan->n_centers, plan->centers, plan->offs,
                                     plan->nl, plan->species, x, desc, s);
    HIP_TRY(e, "forward kernel launch");
    return LDSF_OK;
}

// adjoint kernel only: one dE/dr record per edge and per centre in plan->edge
int plan_adjoint_edges(ldsf_plan *plan, const double *x, const double *dE, int precision,
                       hipStream_t s)
{
    ldsf::SFTab tab;
    if (int rc = ldsf::device_table(plan->sf, &tab)) return rc;
    const long n_rec = plan->n_edges + plan->n_centers;
    if (!plan->edge && n_rec > 0)
        HIP_TRY(hipMalloc((void **)&plan->edge, sizeof(double) * 3 * n_rec), "hipMalloc(edge)");
    hipError_t e;
    if (plan->family == 1 && plan->wave_ok && ldsf::wave_runs(plan->wave, precision))
        e = ldsf::launch_adj_wave(plan->wave, precision, plan->n_centers, plan->centers,
                                  plan->offs, wave_nl(plan), plan->species, x, dE, plan->inv_pos,
                                  plan->edge, (int)plan->n_edges, tab.width, plan->max_nn,
                                  plan->wave_group, s);
    else
        e = ldsf::launch_adj_general(tab, precision, plan->n_centers, plan->centers, plan->offs,
                                     plan->nl, plan->species, x, dE, plan->inv_pos, plan->edge,
                                     (int)plan->n_edges, s);
    HIP_TRY(e, "adjoint kernel launch");
    plan->edges_ready = true;
    return LDSF_OK;
}

// the fused kernel: desc += descriptors and the adjoint's records, from one
// pass over the pairs and triplets (general family: the two kernels in turn)
int plan_forward_adjoint_edges(ldsf_plan *plan, const double *x, const double *dE, double *desc,
                               int precision, hipStream_t s, bool assign = false)
{
    ldsf::SFTab tab;
    if (int rc = ldsf::device_table(plan->sf, &tab)) return rc;
    const long n_rec = plan->n_edges + plan->n_centers;
    if (!plan->edge && n_rec > 0)
        HIP_TRY(hipMalloc((void **)&plan->edge, sizeof(double) * 3 * n_rec), "hipMalloc(edge)");
    hipError_t e;
    if (plan->family == 1 && plan->wave_ok && ldsf::wave_runs(plan->wave, precision)) {
        e = ldsf::launch_adj_wave(plan->wave, precision, plan->n_centers, plan->centers,
                                  plan->offs, wave_nl(plan), plan->species, x, dE, plan->inv_pos,
                                  plan->edge, (int)plan->n_edges, tab.width, plan->max_nn,
                                  plan->wave_group, s, ldsf::DevCounts(), desc, assign);
    } else {
        // the general forward adds: an assigning call starts its rows from zero
        e = assign ? ldsf::launch_zero_f64((long)plan->n_centers * tab.width, desc, s) : hipSuccess;
        if (e == hipSuccess)
            e = ldsf::launch_fwd_general(tab, precision, plan->n_centers, plan->centers, plan->offs,
                                         plan->nl, plan->species, x, desc, s);
        if (e == hipSuccess)
            e = ldsf::launch_adj_general(tab, precision, plan->n_centers, plan->centers,
                                         plan->offs, plan->nl, plan->species, x, dE,
                                         plan->inv_pos, plan->edge, (int)plan->n_edges, s);
    }
    HIP_TRY(e, "fused kernel launch");
    plan->edges_ready = true;
    return LDSF_OK;
}

int plan_gather(ldsf_plan *plan, double *dcoords, hipStream_t s, bool assign = false)
{
    if (!plan->edges_ready && plan->n_edges + plan->n_centers > 0)
        return ldsf::fail(LDSF_ERR_ARG, "plan_gather: no adjoint records (run the adjoint first)");
    HIP_TRY(ldsf::launch_gather(plan->n_total, plan->in_start, plan->edge, dcoords, s, nullptr,
                                assign),
            "gather launch");
    return LDSF_OK;
}

int plan_gradient(ldsf_plan *plan, const double *x, const double *dE, double *dcoords,
                  int precision, hipStream_t s)
{
    if (int rc = plan_adjoint_edges(plan, x, dE, precision, s)) return rc;
    return plan_gather(plan, dcoords, s);
}

int plan_forward_gradient(ldsf_plan *plan, const double *x, const double *dE, double *desc,
                          double *dcoords, int precision, hipStream_t s, bool assign = false)
{
    if (int rc = plan_forward_adjoint_edges(plan, x, dE, desc, precision, s, assign)) return rc;
    return plan_gather(plan, dcoords, s, assign);
}

// ---- host-pointer entry points: stage, run a plan, copy back ----

struct HostCall {
    const ldsf_sf *sf;
    int n_centers;
    std::vector<int> centers;  // empty = identity
    const int *species;
    const int *nl;
    const int *numneigh;
    const double *coords;
    int n_rows = 0;
    long n_edges = 0;
};

int host_prepare(HostCall &c)
{
    if (!c.sf) return ldsf::fail(LDSF_ERR_ARG, "null descriptor");
    if (c.n_centers < 0) return ldsf::fail(LDSF_ERR_ARG, "negative atom count");
    if (c.n_centers > 0 && (!c.species || !c.numneigh || !c.coords))
        return ldsf::fail(LDSF_ERR_ARG, "null input array");
    long total = 0;
    int rows = 0;
    for (int t = 0; t < c.n_centers; ++t) {
        if (c.numneigh[t] < 0) return ldsf::fail(LDSF_ERR_ARG, "negative number_of_neighs entry");
        total += c.numneigh[t];
        int ctr = c.centers.empty() ? t : c.centers[t];
        if (ctr < 0) return ldsf::fail(LDSF_ERR_ARG, "negative centre index");
        rows = std::max(rows, ctr + 1);
    }
    if (total > 0 && !c.nl) return ldsf::fail(LDSF_ERR_ARG, "null neighbour list");
    if (total > 0x7fffffffL) return ldsf::fail(LDSF_ERR_UNSUPPORTED, "more than 2^31-1 neighbour-list entries");
    for (long e = 0; e < total; ++e) {
        if (c.nl[e] < 0) return ldsf::fail(LDSF_ERR_ARG, "negative neighbour index");
        rows = std::max(rows, c.nl[e] + 1);
    }
    const long limit = ldsf::host_rows_limit();
    if (limit >= 0 && rows > limit) return ldsf::fail(LDSF_ERR_ARG, "neighbour or centre index out of range");
    c.n_rows = rows;
    c.n_edges = total;
    return LDSF_OK;
}

// the counts and centres of a call (O(n_centers) on the host): n_edges, and
// n_rows as far as the centres fix it
int host_counts(HostCall &c)
{
    if (!c.sf) return ldsf::fail(LDSF_ERR_ARG, "null descriptor");
    if (c.n_centers < 0) return ldsf::fail(LDSF_ERR_ARG, "negative atom count");
    if (c.n_centers > 0 && (!c.species || !c.numneigh || !c.coords))
        return ldsf::fail(LDSF_ERR_ARG, "null input array");
    long total = 0;
    int rows = 0;
    for (int t = 0; t < c.n_centers; ++t) {
        if (c.numneigh[t] < 0) return ldsf::fail(LDSF_ERR_ARG, "negative number_of_neighs entry");
        total += c.numneigh[t];
        const int ctr = c.centers.empty() ? t : c.centers[t];
        if (ctr < 0) return ldsf::fail(LDSF_ERR_ARG, "negative centre index");
        rows = std::max(rows, ctr + 1);
    }
    if (total > 0 && !c.nl) return ldsf::fail(LDSF_ERR_ARG, "null neighbour list");
    if (total > 0x7fffffffL) return ldsf::fail(LDSF_ERR_UNSUPPORTED, "more than 2^31-1 neighbour-list entries");
    c.n_rows = rows;
    c.n_edges = total;
    return LDSF_OK;
}

// The host-pointer calls' workspace on one device: a stream, the staging
// copies, the staged arrays and the plan arena, all kept at the largest
// size a call has needed until the descriptor is destroyed.  `mu` makes the
// calls that share it run one at a time.
struct HostWork {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::unique_ptr<ldsf::Stager> st;
    ldsf::Grow<int> species, nl, nn;
    ldsf::Grow<double> x, desc, dE, dx;
    PlanArena arena;
    ~HostWork()
    {
        st.reset();  // waits for the DMAs still reading or writing pinned chunks
        if (stream) (void)hipStreamDestroy(stream);
    }
};

void release_host_work(int device, void *mem)
{
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    delete static_cast<HostWork *>(mem);
    (void)hipSetDevice(cur);
}

int host_work(const ldsf_sf *csf, HostWork **out)
{
    auto *sf = const_cast<ldsf_sf *>(csf);
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    std::lock_guard<std::mutex> lk(sf->dev_mu);
    for (auto &d : sf->dev)
        if (d.device == device && d.kind == 1) {
            *out = static_cast<HostWork *>(d.mem);
            return LDSF_OK;
        }
    auto W = std::make_unique<HostWork>();
    W->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking), "hipStreamCreate");
    W->st = std::make_unique<ldsf::Stager>(W->stream);
    ldsf_sf::Dev d;
    d.device = device;
    d.kind = 1;
    d.mem = W.get();
    d.release = release_host_work;
    sf->dev.push_back(d);
    *out = W.release();
    return LDSF_OK;
}

// desc_io: [n_centers][width] (+=, or = with assign_desc) or NULL; dE +
// dcoords_io for the adjoint.  The neighbour list goes up first, with its
// smallest and largest index taken on the way (they fix how many species
// and coordinate rows are read); the plan is built in the workspace's
// arena; results come back through the same pinned chunks.
int host_run(HostCall &c, double *desc_io, const double *dE, double *dcoords_io, int precision,
             bool assign_desc = false)
{
    if (int rc = host_counts(c)) return rc;
    if (c.n_centers == 0) return LDSF_OK;
    const int W = c.sf->width;
    HostWork *H = nullptr;
    if (int rc = host_work(c.sf, &H)) return rc;
    std::lock_guard<std::mutex> lk(H->mu);
    ldsf::Stager &st = *H->st;
    hipStream_t s = H->stream;
    if (c.n_edges > 0) {
        HIP_TRY(H->nl.reserve(c.n_edges), "hipMalloc(neighbour list)");
        int mn = 0, mx = 0;
        HIP_TRY(st.h2d_minmax(H->nl.p, c.nl, c.n_edges, &mn, &mx), "upload neighbour list");
        if (mn < 0) return ldsf::fail(LDSF_ERR_ARG, "negative neighbour index");
        c.n_rows = std::max(c.n_rows, mx + 1);
    }
    const long limit = ldsf::host_rows_limit();
    if (limit >= 0 && c.n_rows > limit)
        return ldsf::fail(LDSF_ERR_ARG, "neighbour or centre index out of range");
    const size_t C = c.n_centers, R = c.n_rows;
    HIP_TRY(H->species.reserve(R), "hipMalloc(species)");
    HIP_TRY(H->x.reserve(3 * R), "hipMalloc(coordinates)");
    HIP_TRY(H->nn.reserve(C), "hipMalloc(neighbour counts)");
    HIP_TRY(st.h2d(H->species.p, c.species, sizeof(int) * R), "upload species");
    HIP_TRY(st.h2d(H->x.p, c.coords, sizeof(double) * 3 * R), "upload coordinates");
    HIP_TRY(st.h2d(H->nn.p, c.numneigh, sizeof(int) * C), "upload neighbour counts");
    int *d_centers = nullptr;
    if (!c.centers.empty()) {
        HIP_TRY(H->arena.centers.reserve(C), "hipMalloc(centres)");
        HIP_TRY(st.h2d(H->arena.centers.p, c.centers.data(), sizeof(int) * C), "upload centres");
        d_centers = H->arena.centers.p;
    }
    ldsf_plan *plan = nullptr;
    int rc = plan_build(&plan, c.sf, c.n_centers, d_centers, c.n_rows, H->species.p, H->nl.p,
                        H->nn.p, s, &H->arena);
    if (rc) return rc;
    std::unique_ptr<ldsf_plan> guard(plan);
    const size_t DW = C * (size_t)W;
    if (desc_io) {
        HIP_TRY(H->desc.reserve(DW), "hipMalloc(descriptors)");
        if (assign_desc)  // starts from zero on the device: no upload
            HIP_TRY(hipMemsetAsync(H->desc.p, 0, sizeof(double) * DW, s), "hipMemsetAsync");
        else
            HIP_TRY(st.h2d(H->desc.p, desc_io, sizeof(double) * DW), "upload descriptors");
    }
    if (dE) {
        HIP_TRY(H->dE.reserve(DW), "hipMalloc(dE/dzeta)");
        HIP_TRY(H->dx.reserve(3 * R), "hipMalloc(dE/dr)");
        HIP_TRY(st.h2d(H->dE.p, dE, sizeof(double) * DW), "upload dE/dzeta");
        HIP_TRY(st.h2d(H->dx.p, dcoords_io, sizeof(double) * 3 * R), "upload dE/dr");
        HIP_TRY(H->arena.edge.reserve(3 * (size_t)(plan->n_edges + plan->n_centers)),
                "hipMalloc(edge records)");
        plan->edge = H->arena.edge.p;
        plan->edge_owned = false;
    }
    if (desc_io && dE)  // the reference's gradient: the forward runs too (one fused pass)
        rc = plan_forward_gradient(plan, H->x.p, H->dE.p, H->desc.p, H->dx.p, precision, s);
    else if (desc_io)
        rc = plan_compute(plan, H->x.p, H->desc.p, precision, s);
    else if (dE)
        rc = plan_gradient(plan, H->x.p, H->dE.p, H->dx.p, precision, s);
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (desc_io) HIP_TRY(st.d2h(desc_io, H->desc.p, sizeof(double) * DW), "download descriptors");
    if (dE) HIP_TRY(st.d2h(dcoords_io, H->dx.p, sizeof(double) * 3 * R), "download dE/dr");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    return LDSF_OK;
}

std::vector<int> batch_centers(int n_conf, const int *n_atoms, const int *ptr, int *n_centers, int *rc)
{
    std::vector<int> c;
    *rc = LDSF_OK;
    long total = 0;
    for (int k = 0; k < n_conf; ++k) {
        if (n_atoms[k] < 0) { *rc = ldsf::fail(LDSF_ERR_ARG, "negative n_atoms entry"); return c; }
        total += n_atoms[k];
    }
    if (total > 0x7fffffffL) { *rc = ldsf::fail(LDSF_ERR_UNSUPPORTED, "too many centres"); return c; }
    c.reserve(total);
    for (int k = 0; k < n_conf; ++k)
        for (int j = 0; j < n_atoms[k]; ++j) c.push_back(j + ptr[k]);
    *n_centers = (int)total;
    return c;
}

}  // namespace

extern "C" {

LDSF_API int ldsf_plan_create(ldsf_plan **out, const ldsf_sf *sf, int n_centers,
                              const int *d_centers, int n_total, const int *d_species,
                              const int *d_neighbor_list, const int *d_number_of_neighs,
                              void *stream)
{
    if (!out) return ldsf::fail(LDSF_ERR_ARG, "plan_create: null output");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    int *owned = nullptr;
    if (d_centers && n_centers > 0) {
        HIP_TRY(hipMalloc((void **)&owned, sizeof(int) * n_centers), "hipMalloc(centres)");
        hipError_t e = hipMemcpyAsync(owned, d_centers, sizeof(int) * n_centers,
                                      hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) { (void)hipFree(owned); return hip_fail(e, "copy centres"); }
    }
    return plan_build(out, sf, n_centers, owned, n_total, d_species, d_neighbor_list,
                      d_number_of_neighs, s);
}

LDSF_API int ldsf_plan_create_batch(ldsf_plan **out, const ldsf_sf *sf, int n_configurations,
                                    const int *n_atoms, const int *config_ptr, int n_total,
                                    const int *d_species, const int *d_neighbor_list,
                                    const int *d_number_of_neighs, void *stream)
{
    if (!out) return ldsf::fail(LDSF_ERR_ARG, "plan_create_batch: null output");
    *out = nullptr;
    if (n_configurations < 0 || (n_configurations > 0 && (!n_atoms || !config_ptr)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_create_batch: bad arguments");
    int n_centers = 0, rc = 0;
    std::vector<int> c = batch_centers(n_configurations, n_atoms, config_ptr, &n_centers, &rc);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    int *owned = nullptr;
    if (n_centers > 0) {
        HIP_TRY(hipMalloc((void **)&owned, sizeof(int) * n_centers), "hipMalloc(centres)");
        hipError_t e = hipMemcpy(owned, c.data(), sizeof(int) * n_centers, hipMemcpyHostToDevice);
        if (e != hipSuccess) { (void)hipFree(owned); return hip_fail(e, "copy centres"); }
    }
    return plan_build(out, sf, n_centers, owned, n_total, d_species, d_neighbor_list,
                      d_number_of_neighs, s);
}

LDSF_API void ldsf_plan_destroy(ldsf_plan *plan) { delete plan; }

LDSF_API int ldsf_plan_num_centers(const ldsf_plan *plan) { return plan ? plan->n_centers : 0; }

LDSF_API long ldsf_plan_num_edges(const ldsf_plan *plan) { return plan ? plan->n_edges : 0; }

LDSF_API int ldsf_plan_max_neighbors(const ldsf_plan *plan) { return plan ? plan->max_nn : 0; }

LDSF_API int ldsf_plan_kernel_family(const ldsf_plan *plan) { return plan ? plan->family : -1; }

LDSF_API int ldsf_plan_set_kernel_family(ldsf_plan *plan, int family)
{
    if (!plan || (family != 0 && family != 1)) return ldsf::fail(LDSF_ERR_ARG, "set_kernel_family: bad arguments");
    if (family == 1 && !plan->wave_ok)
        return ldsf::fail(LDSF_ERR_UNSUPPORTED, "wave kernels do not cover this parameter set");
    plan->family = family;
    return LDSF_OK;
}

LDSF_API long ldsf_plan_num_records(const ldsf_plan *plan)
{
    return plan ? plan->n_edges + plan->n_centers : 0;
}

LDSF_API int ldsf_plan_set_records(ldsf_plan *plan, double *d_records, long n_doubles)
{
    if (!plan) return ldsf::fail(LDSF_ERR_ARG, "plan_set_records: null plan");
    const long need = 3 * (plan->n_edges + plan->n_centers);
    if (need > 0 && (!d_records || n_doubles < need))
        return ldsf::fail(LDSF_ERR_ARG, "plan_set_records: buffer smaller than 3 * ldsf_plan_num_records");
    if (plan->edge && plan->edge_owned) {
        HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
        (void)hipFree(plan->edge);
    }
    plan->edge = need > 0 ? d_records : nullptr;
    plan->edge_owned = false;
    plan->edges_ready = false;
    return LDSF_OK;
}

LDSF_API int ldsf_plan_compute(ldsf_plan *plan, const double *d_coordinates, double *d_desc,
                               int precision, void *stream)
{
    if (!plan || (plan->n_centers > 0 && (!d_coordinates || !d_desc)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_compute: bad arguments");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return ldsf::fail(LDSF_ERR_ARG, "plan_compute: unknown precision");
    return plan_compute(plan, d_coordinates, d_desc, precision, (hipStream_t)stream);
}

LDSF_API int ldsf_plan_gradient(ldsf_plan *plan, const double *d_coordinates,
                                const double *d_dE_dzeta, double *d_dcoords, int precision,
                                void *stream)
{
    if (!plan || (plan->n_centers > 0 && (!d_coordinates || !d_dE_dzeta || !d_dcoords)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_gradient: bad arguments");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return ldsf::fail(LDSF_ERR_ARG, "plan_gradient: unknown precision");
    return plan_gradient(plan, d_coordinates, d_dE_dzeta, d_dcoords, precision, (hipStream_t)stream);
}

LDSF_API int ldsf_plan_adjoint_edges(ldsf_plan *plan, const double *d_coordinates,
                                     const double *d_dE_dzeta, int precision, void *stream)
{
    if (!plan || (plan->n_centers > 0 && (!d_coordinates || !d_dE_dzeta)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_adjoint_edges: bad arguments");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return ldsf::fail(LDSF_ERR_ARG, "plan_adjoint_edges: unknown precision");
    return plan_adjoint_edges(plan, d_coordinates, d_dE_dzeta, precision, (hipStream_t)stream);
}

LDSF_API int ldsf_plan_forward_gradient(ldsf_plan *plan, const double *d_coordinates,
                                        const double *d_dE_dzeta, double *d_desc,
                                        double *d_dcoords, int precision, void *stream)
{
    if (!plan || (plan->n_centers > 0 && (!d_coordinates || !d_dE_dzeta || !d_desc || !d_dcoords)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_gradient: bad arguments");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_gradient: unknown precision");
    return plan_forward_gradient(plan, d_coordinates, d_dE_dzeta, d_desc, d_dcoords, precision,
                                 (hipStream_t)stream);
}

LDSF_API int ldsf_plan_forward_gradient_assign(ldsf_plan *plan, const double *d_coordinates,
                                               const double *d_dE_dzeta, double *d_desc,
                                               double *d_dcoords, int precision, void *stream)
{
    if (!plan || (plan->n_centers > 0 && (!d_coordinates || !d_dE_dzeta || !d_desc || !d_dcoords)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_gradient_assign: bad arguments");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_gradient_assign: unknown precision");
    return plan_forward_gradient(plan, d_coordinates, d_dE_dzeta, d_desc, d_dcoords, precision,
                                 (hipStream_t)stream, true);
}

LDSF_API int ldsf_plan_forward_adjoint_edges(ldsf_plan *plan, const double *d_coordinates,
                                             const double *d_dE_dzeta, double *d_desc,
                                             int precision, void *stream)
{
    if (!plan || (plan->n_centers > 0 && (!d_coordinates || !d_dE_dzeta || !d_desc)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_adjoint_edges: bad arguments");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_adjoint_edges: unknown precision");
    return plan_forward_adjoint_edges(plan, d_coordinates, d_dE_dzeta, d_desc, precision,
                                      (hipStream_t)stream);
}

LDSF_API int ldsf_plan_forward_adjoint_edges_assign(ldsf_plan *plan, const double *d_coordinates,
                                                    const double *d_dE_dzeta, double *d_desc,
                                                    int precision, void *stream)
{
    if (!plan || (plan->n_centers > 0 && (!d_coordinates || !d_dE_dzeta || !d_desc)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_adjoint_edges_assign: bad arguments");
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        return ldsf::fail(LDSF_ERR_ARG, "plan_forward_adjoint_edges_assign: unknown precision");
    return plan_forward_adjoint_edges(plan, d_coordinates, d_dE_dzeta, d_desc, precision,
                                      (hipStream_t)stream, true);
}

LDSF_API int ldsf_plan_gather_assign(ldsf_plan *plan, double *d_dcoords, void *stream)
{
    if (!plan || (plan->n_total > 0 && !d_dcoords))
        return ldsf::fail(LDSF_ERR_ARG, "plan_gather_assign: bad arguments");
    return plan_gather(plan, d_dcoords, (hipStream_t)stream, true);
}

LDSF_API int ldsf_plan_param_gradient(ldsf_plan *plan, const double *d_coordinates,
                                      const double *d_dE_dzeta, double *d_dparams,
                                      double *d_drcut, void *stream)
{
    if (!plan || !d_dparams || (plan->n_centers > 0 && (!d_coordinates || !d_dE_dzeta)))
        return ldsf::fail(LDSF_ERR_ARG, "plan_param_gradient: bad arguments");
    ldsf::SFTab tab;
    if (int rc = ldsf::device_table(plan->sf, &tab)) return rc;
    if (tab.n_species > 8)
        return ldsf::fail(LDSF_ERR_UNSUPPORTED, "plan_param_gradient: more than 8 species");
    HIP_TRY(ldsf::launch_param_gradient(tab, plan->n_centers, plan->centers, plan->offs, plan->nl,
                                        plan->species, d_coordinates, d_dE_dzeta, d_dparams,
                                        d_drcut, (hipStream_t)stream),
            "parameter gradient");
    return LDSF_OK;
}

LDSF_API int ldsf_param_gradient(ldsf_sf *sf, int n_atoms, const int *species,
                                 const int *neighbor_list, const int *number_of_neighs,
                                 const double *coordinates, const double *dE_dzeta,
                                 double *dparams, double *drcut)
{
    if (!sf || !dparams || (n_atoms > 0 && !dE_dzeta))
        return ldsf::fail(LDSF_ERR_ARG, "param_gradient: bad arguments");
    HostCall c{sf, n_atoms, {}, species, neighbor_list, number_of_neighs, coordinates};
    if (int rc = host_prepare(c)) return rc;
    if (n_atoms == 0) return LDSF_OK;
    const int W = sf->width, ns = (int)sf->species.size();
    DevBuf<int> d_species, d_nl, d_nn;
    DevBuf<double> d_x, d_dE, d_dp, d_drc;
    HIP_TRY(d_species.alloc(c.n_rows), "hipMalloc");
    HIP_TRY(d_x.alloc(3 * (size_t)c.n_rows), "hipMalloc");
    HIP_TRY(d_nl.alloc(c.n_edges), "hipMalloc");
    HIP_TRY(d_nn.alloc(n_atoms), "hipMalloc");
    HIP_TRY(d_dE.alloc((size_t)n_atoms * W), "hipMalloc");
    HIP_TRY(d_dp.alloc(std::max(sf->n_params, 1)), "hipMalloc");
    HIP_TRY(d_drc.alloc(ns * ns), "hipMalloc");
    HIP_TRY(hipMemcpy(d_species.p, species, sizeof(int) * c.n_rows, hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(hipMemcpy(d_x.p, coordinates, sizeof(double) * 3 * c.n_rows, hipMemcpyHostToDevice), "hipMemcpy");
    if (c.n_edges)
        HIP_TRY(hipMemcpy(d_nl.p, neighbor_list, sizeof(int) * c.n_edges, hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(hipMemcpy(d_nn.p, number_of_neighs, sizeof(int) * n_atoms, hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(hipMemcpy(d_dE.p, dE_dzeta, sizeof(double) * n_atoms * W, hipMemcpyHostToDevice), "hipMemcpy");
    if (sf->n_params)
        HIP_TRY(hipMemcpy(d_dp.p, dparams, sizeof(double) * sf->n_params, hipMemcpyHostToDevice), "hipMemcpy");
    if (drcut) HIP_TRY(hipMemcpy(d_drc.p, drcut, sizeof(double) * ns * ns, hipMemcpyHostToDevice), "hipMemcpy");
    ldsf_plan *plan = nullptr;
    int rc = plan_build(&plan, sf, n_atoms, nullptr, c.n_rows, d_species.p, d_nl.p, d_nn.p, nullptr);
    if (rc) return rc;
    std::unique_ptr<ldsf_plan> guard(plan);
    if ((rc = ldsf_plan_param_gradient(plan, d_x.p, d_dE.p, d_dp.p, drcut ? d_drc.p : nullptr, nullptr)))
        return rc;
    if (sf->n_params)
        HIP_TRY(hipMemcpy(dparams, d_dp.p, sizeof(double) * sf->n_params, hipMemcpyDeviceToHost), "hipMemcpy");
    if (drcut) HIP_TRY(hipMemcpy(drcut, d_drc.p, sizeof(double) * ns * ns, hipMemcpyDeviceToHost), "hipMemcpy");
    return LDSF_OK;
}

LDSF_API int ldsf_plan_gather_fold(ldsf_plan *plan, const ldsf_layout *layout, double *d_out,
                                   void *stream)
{
    if (!plan || !layout) return ldsf::fail(LDSF_ERR_ARG, "plan_gather_fold: null plan or layout");
    if (plan->n_total != ldsf::layout_rows(layout))
        return ldsf::fail(LDSF_ERR_ARG, "plan_gather_fold: the plan's rows are not the layout's");
    if (plan->n_centers > 0 && !d_out) return ldsf::fail(LDSF_ERR_ARG, "plan_gather_fold: null output");
    if (!plan->edges_ready && plan->n_edges + plan->n_centers > 0)
        return ldsf::fail(LDSF_ERR_ARG, "plan_gather_fold: no adjoint records (run the adjoint first)");
    return ldsf::layout_gather_fold(layout, plan->in_start, plan->edge, d_out, (hipStream_t)stream);
}

LDSF_API int ldsf_plan_gather(ldsf_plan *plan, double *d_dcoords, void *stream)
{
    if (!plan || (plan->n_total > 0 && !d_dcoords))
        return ldsf::fail(LDSF_ERR_ARG, "plan_gather: bad arguments");
    return plan_gather(plan, d_dcoords, (hipStream_t)stream);
}

LDSF_API int ldsf_compute(ldsf_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                          const int *number_of_neighs, const double *coordinates, double *desc)
{
    if (!sf || (n_atoms > 0 && !desc)) return ldsf::fail(LDSF_ERR_ARG, "compute: bad arguments");
    HostCall c{sf, n_atoms, {}, species, neighbor_list, number_of_neighs, coordinates};
    return host_run(c, desc, nullptr, nullptr, sf->precision);
}

LDSF_API int ldsf_compute_overwrite(ldsf_sf *sf, int n_atoms, const int *species,
                                    const int *neighbor_list, const int *number_of_neighs,
                                    const double *coordinates, double *desc)
{
    if (!sf || (n_atoms > 0 && !desc)) return ldsf::fail(LDSF_ERR_ARG, "compute: bad arguments");
    HostCall c{sf, n_atoms, {}, species, neighbor_list, number_of_neighs, coordinates};
    return host_run(c, desc, nullptr, nullptr, sf->precision, true);
}

LDSF_API int ldsf_compute_single_atom(ldsf_sf *sf, int index, int n_atoms, const int *species,
                                      const int *neighbor_list, int number_of_neighs,
                                      const double *coordinates, double *desc)
{
    (void)n_atoms;
    if (!sf || !desc || index < 0) return ldsf::fail(LDSF_ERR_ARG, "compute_single_atom: bad arguments");
    int nn = number_of_neighs;
    HostCall c{sf, 1, {index}, species, neighbor_list, &nn, coordinates};
    return host_run(c, desc, nullptr, nullptr, sf->precision);
}

LDSF_API int ldsf_compute_batch(ldsf_sf *sf, int n_configurations, const int *n_atoms,
                                const int *config_ptr, const int *species, const int *neighbor_list,
                                const int *number_of_neighs, const double *coordinates, double *desc)
{
    if (!sf || n_configurations < 0 || (n_configurations > 0 && (!n_atoms || !config_ptr)))
        return ldsf::fail(LDSF_ERR_ARG, "compute_batch: bad arguments");
    int n = 0, rc = 0;
    std::vector<int> ctr = batch_centers(n_configurations, n_atoms, config_ptr, &n, &rc);
    if (rc) return rc;
    if (n > 0 && !desc) return ldsf::fail(LDSF_ERR_ARG, "compute_batch: null desc");
    HostCall c{sf, n, std::move(ctr), species, neighbor_list, number_of_neighs, coordinates};
    return host_run(c, desc, nullptr, nullptr, sf->precision);
}

LDSF_API int ldsf_compute_batch_overwrite(ldsf_sf *sf, int n_configurations, const int *n_atoms,
                                          const int *config_ptr, const int *species,
                                          const int *neighbor_list, const int *number_of_neighs,
                                          const double *coordinates, double *desc)
{
    if (!sf || n_configurations < 0 || (n_configurations > 0 && (!n_atoms || !config_ptr)))
        return ldsf::fail(LDSF_ERR_ARG, "compute_batch: bad arguments");
    int n = 0, rc = 0;
    std::vector<int> ctr = batch_centers(n_configurations, n_atoms, config_ptr, &n, &rc);
    if (rc) return rc;
    if (n > 0 && !desc) return ldsf::fail(LDSF_ERR_ARG, "compute_batch: null desc");
    HostCall c{sf, n, std::move(ctr), species, neighbor_list, number_of_neighs, coordinates};
    return host_run(c, desc, nullptr, nullptr, sf->precision, true);
}

LDSF_API int ldsf_gradient(ldsf_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                           const int *number_of_neighs, const double *coordinates,
                           double *d_coordinates, double *desc, const double *dE_dzeta)
{
    if (!sf || (n_atoms > 0 && (!d_coordinates || !dE_dzeta)))
        return ldsf::fail(LDSF_ERR_ARG, "gradient: bad arguments");
    HostCall c{sf, n_atoms, {}, species, neighbor_list, number_of_neighs, coordinates};
    return host_run(c, desc, dE_dzeta, d_coordinates, sf->precision);
}

LDSF_API int ldsf_gradient_single_atom(ldsf_sf *sf, int index, int n_atoms, const int *species,
                                       const int *neighbor_list, int number_of_neighs,
                                       const double *coordinates, double *d_coordinates,
                                       double *desc, const double *dE_dzeta)
{
    (void)n_atoms;
    if (!sf || !d_coordinates || !dE_dzeta || index < 0)
        return ldsf::fail(LDSF_ERR_ARG, "gradient_single_atom: bad arguments");
    int nn = number_of_neighs;
    HostCall c{sf, 1, {index}, species, neighbor_list, &nn, coordinates};
    return host_run(c, desc, dE_dzeta, d_coordinates, sf->precision);
}

LDSF_API int ldsf_gradient_batch(ldsf_sf *sf, int n_configurations, const int *n_atoms,
                                 const int *config_ptr, const int *species,
                                 const int *neighbor_list, const int *number_of_neighs,
                                 const double *coordinates, double *d_coordinates, double *desc,
                                 const double *dE_dzeta)
{
    if (!sf || n_configurations < 0 || (n_configurations > 0 && (!n_atoms || !config_ptr)))
        return ldsf::fail(LDSF_ERR_ARG, "gradient_batch: bad arguments");
    int n = 0, rc = 0;
    std::vector<int> ctr = batch_centers(n_configurations, n_atoms, config_ptr, &n, &rc);
    if (rc) return rc;
    if (n > 0 && (!d_coordinates || !dE_dzeta)) return ldsf::fail(LDSF_ERR_ARG, "gradient_batch: null output");
    HostCall c{sf, n, std::move(ctr), species, neighbor_list, number_of_neighs, coordinates};
    return host_run(c, desc, dE_dzeta, d_coordinates, sf->precision);
}

LDSF_API int ldsf_jacobian(ldsf_sf *sf, int n_atoms, int n_total_atoms, const int *species,
                           const int *neighbor_list, const int *number_of_neighs,
                           const double *coordinates, double *J_coordinates)
{
    if (!sf || n_atoms < 0 || n_total_atoms < 0 || (n_atoms > 0 && !J_coordinates))
        return ldsf::fail(LDSF_ERR_ARG, "jacobian: bad arguments");
    HostCall c{sf, n_atoms, {}, species, neighbor_list, number_of_neighs, coordinates};
    if (int rc = host_prepare(c)) return rc;
    if (n_atoms == 0) return LDSF_OK;
    if (c.n_rows > n_total_atoms) return ldsf::fail(LDSF_ERR_ARG, "jacobian: neighbour index >= n_total_atoms");
    const int W = sf->width;
    const size_t rows = (size_t)W * n_atoms;
    const size_t cols = 3 * (size_t)n_total_atoms;
    DevBuf<int> d_species, d_nl, d_nn;
    DevBuf<double> d_x, d_g, d_J;
    HIP_TRY(d_species.alloc(c.n_rows), "hipMalloc");
    HIP_TRY(d_x.alloc(3 * (size_t)c.n_rows), "hipMalloc");
    HIP_TRY(d_nl.alloc(c.n_edges), "hipMalloc");
    HIP_TRY(d_nn.alloc(n_atoms), "hipMalloc");
    HIP_TRY(hipMemcpy(d_species.p, species, sizeof(int) * c.n_rows, hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(hipMemcpy(d_x.p, coordinates, sizeof(double) * 3 * c.n_rows, hipMemcpyHostToDevice), "hipMemcpy");
    if (c.n_edges)
        HIP_TRY(hipMemcpy(d_nl.p, neighbor_list, sizeof(int) * c.n_edges, hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(hipMemcpy(d_nn.p, number_of_neighs, sizeof(int) * n_atoms, hipMemcpyHostToDevice), "hipMemcpy");
    ldsf_plan *plan = nullptr;
    int rc = plan_build(&plan, sf, n_atoms, nullptr, c.n_rows, d_species.p, d_nl.p, d_nn.p, nullptr);
    if (rc) return rc;
    std::unique_ptr<ldsf_plan> guard(plan);
    // one-hot VJPs, one per descriptor entry (Descriptors.cpp:467-483), each
    // into its own row of J (whose first n_rows*3 columns can be non-zero)
    HIP_TRY(d_g.alloc(rows), "hipMalloc");
    HIP_TRY(d_J.alloc(3 * (size_t)c.n_rows), "hipMalloc");
    HIP_TRY(hipMemset(d_g.p, 0, sizeof(double) * rows), "hipMemset");
    std::vector<double> row(3 * (size_t)c.n_rows);
    const double one = 1.0, zero = 0.0;
    for (size_t j = 0; j < rows; ++j) {
        if (j > 0) HIP_TRY(hipMemcpy(d_g.p + j - 1, &zero, sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        HIP_TRY(hipMemcpy(d_g.p + j, &one, sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        double *Jrow = J_coordinates + j * cols;
        HIP_TRY(hipMemcpy(d_J.p, Jrow, sizeof(double) * row.size(), hipMemcpyHostToDevice), "hipMemcpy");
        if ((rc = plan_gradient(plan, d_x.p, d_g.p, d_J.p, sf->precision, nullptr))) return rc;
        HIP_TRY(hipMemcpy(Jrow, d_J.p, sizeof(double) * row.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    return LDSF_OK;
}

// numdiff::make_xph_representable (include/finite_difference.hpp:46-57)
static double make_xph_representable(double x, double h)
{
    volatile double temp = x + h;
    h = temp - x;
    if (h == 0) h = std::nextafter(x, std::numeric_limits<double>::max()) - x;
    return h;
}

LDSF_API int ldsf_num_gradient_single_atom(ldsf_sf *sf, int index, int n_atoms, const int *species,
                                           const int *neighbor_list, int number_of_neighs,
                                           const double *coordinates, double *d_coordinates,
                                           const double *dE_dzeta)
{
    if (!sf || !coordinates || !d_coordinates || !dE_dzeta || index < 0 || index >= n_atoms)
        return ldsf::fail(LDSF_ERR_ARG, "num_gradient_single_atom: bad arguments");
    const int W = sf->width;
    std::vector<double> x(coordinates, coordinates + 3 * (size_t)n_atoms);
    std::vector<double> y(8 * (size_t)W);
    static const int mult[8] = {1, -1, -2, 2, 3, -3, -4, 4};
    for (int dim = 0; dim < 3; ++dim) {
        const int pos = index * 3 + dim;
        double h = std::pow(551.25 * std::numeric_limits<double>::epsilon(), (double)1 / (double)9);
        h = make_xph_representable(coordinates[pos], h);
        for (int e = 0; e < 8; ++e) {
            std::fill(y.begin() + e * W, y.begin() + (e + 1) * W, 0.0);
            x[pos] = coordinates[pos] + mult[e] * h;
            if (int rc = ldsf_compute_single_atom(sf, index, n_atoms, species, neighbor_list,
                                                  number_of_neighs, x.data(), y.data() + e * W))
                return rc;
        }
        x[pos] = coordinates[pos];
        // stencil of finite_difference.hpp:120-128
        for (int w = 0; w < W; ++w) {
            double y1 = y[0 * W + w] - y[1 * W + w];
            double y2 = y[2 * W + w] - y[3 * W + w];
            double y3 = y[4 * W + w] - y[5 * W + w];
            double y4 = y[6 * W + w] - y[7 * W + w];
            double t1 = 3 * y4 / 8 + 4 * y3;
            double t2 = 21 * y2 + 84 * y1;
            d_coordinates[dim] += dE_dzeta[w] * ((t1 + t2) / (105 * h));
        }
    }
    return LDSF_OK;
}

}  // extern "C"
