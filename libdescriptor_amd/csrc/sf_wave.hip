// Wave-per-centre symmetry-function kernels for gfx950.
//
// One 64-lane wavefront owns one centre atom at a time (persistent grid,
// XCD-contiguous centre ranges).  Per centre:
//   1. pair phase: lanes over the centre's neighbour-list entries compute the
//      FP64 displacement and distance; the entries inside the cutoff are
//      compacted, in list order, into LDS (ballot + popcount) with their
//      cutoff value; each lane adds its g2 rows into FP64 accumulators.
//   2. triplet enumeration: lanes over all pairs (a < b) of compacted
//      neighbours, walked along diagonals b - a = 1, 2, ... so that the lanes
//      of one instruction touch distinct neighbours; pairs with r_jk < rc_jk
//      are compacted into an LDS list (tagged with the species-pair block).
//   3. triplet phase: lanes over the active list only (full lane occupancy in
//      the expensive part).  cos(theta), fc_jk, one exp per distinct g4 eta
//      and one power per distinct (zeta, lambda) are shared by all rows of a
//      triplet; integer zeta uses exponentiation by squaring.
//   4. the per-lane FP64 accumulators are summed across the wave with a
//      fixed-order butterfly (reduce-scatter) and added into the caller's
//      descriptor row (the reference's += semantics).
// The adjoint kernel runs the same phases, forms dE/d r_ij, dE/d r_ik and
// dE/d r_jk per triplet from three row sums, and adds the resulting vectors
// into per-neighbour FP64 LDS accumulators (ds_add_f64; one wave owns the
// centre, so the order is fixed); it then writes one vector per neighbour-list
// edge and the centre's own vector for the gather in sf_plan.hip.
// Nothing is atomic in global memory, and results do not depend on which
// other centres share the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "sf_device.h"
#include "sf_kernels.h"

namespace ldsf {

namespace {

constexpr int kW = 64;

__device__ __forceinline__ int prefix_count(unsigned long long m, int lane)
{
    return __popcll(m & ((1ull << lane) - 1ull));
}

// Sum K per-lane values over the wave.  Afterwards v[0] of lane l holds the
// total of value (l >> (6 - log2 K)) & (K - 1).  The butterfly order is fixed.
// Halving step with H live values, partner lane l ^ M (all compile-time).
template <int K, int H, int M>
struct Halve {
    static __device__ __forceinline__ void run(double (&v)[K], int lane)
    {
        const bool hi = (lane & M) != 0;
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const double send = hi ? v[q] : v[q + H];
            const double keep = hi ? v[q + H] : v[q];
            v[q] = keep + __shfl_xor(send, M);
        }
        Halve<K, H / 2, M / 2>::run(v, lane);
    }
};

template <int K, int M>
struct Halve<K, 0, M> {
    static __device__ __forceinline__ void run(double (&v)[K], int)
    {
        v[0] += __shfl_xor(v[0], M);
        Halve<K, 0, M / 2>::run(v, 0);
    }
};

template <int K>
struct Halve<K, 0, 0> {
    static __device__ __forceinline__ void run(double (&)[K], int) {}
};

template <int K>
__device__ __forceinline__ void reduce_scatter(double (&v)[K], int lane)
{
    Halve<K, K / 2, 32>::run(v, lane);
}

template <int K>
struct Log2;
template <> struct Log2<1> { static constexpr int v = 0; };
template <> struct Log2<2> { static constexpr int v = 1; };
template <> struct Log2<4> { static constexpr int v = 2; };
template <> struct Log2<8> { static constexpr int v = 3; };
template <> struct Log2<16> { static constexpr int v = 4; };

// reduce and add into row[slot(idx)] (slot < 0: nothing)
template <int K>
__device__ __forceinline__ void reduce_add(double (&v)[K], int lane, const int *__restrict__ slots,
                                           double *__restrict__ row)
{
    reduce_scatter<K>(v, lane);
    constexpr int sh = 6 - Log2<K>::v;
    if ((lane & ((1 << sh) - 1)) == 0) {
        const int slot = slots[(lane >> sh) & (K - 1)];
        if (slot >= 0) row[slot] += v[0];
    }
}

__device__ __forceinline__ void diag_norm(int &d, int &a, int M)
{
    while (d < M && a >= M - d) {
        a -= M - d;
        ++d;
    }
}

// centre range of this block: blocks are dealt to the 8 XCDs round-robin, so
// block b works on XCD b % 8; each XCD takes one contiguous slice of centres
struct CentreRange {
    int lo, hi, step, first;
};

__device__ __forceinline__ CentreRange centre_range(int n)
{
    const int G = gridDim.x, b = blockIdx.x;
    CentreRange r;
    if (G >= 8 && (G & 7) == 0) {
        const int xcd = b & 7, local = b >> 3, per = G >> 3;
        const int chunk = (n + 7) >> 3;
        r.lo = min(n, xcd * chunk);
        r.hi = min(n, r.lo + chunk);
        r.first = r.lo + local;
        r.step = per;
    } else {
        r.lo = 0;
        r.hi = n;
        r.first = b;
        r.step = G;
    }
    return r;
}

// shared geometry of one triplet
template <int P>
struct Trip {
    using T = typename Real<P>::T;
    double ra, rb, rjk, rjk2;  // r_ij, r_ik, r_jk, r_jk^2
    double e0, e1, e2;         // x_k - x_j
    T fa, fb, fjk;             // cutoff values
    T B[2];                    // 1 + lambda cos(theta)
    T E[2];                    // exp(-eta (r_ij^2 + r_ik^2 + r_jk^2)) per distinct eta
};

template <int P>
__device__ __forceinline__ void triplet_geometry(const WaveTab &wt, const WaveParams &wp,
                                                 const double *s_dx,
                                                 const double *s_dy, const double *s_dz,
                                                 const double *s_r, const double *s_r2, int a,
                                                 int b, double rcjk, Trip<P> &t)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    t.B[0] = t.B[1] = (T)0;
    t.E[0] = t.E[1] = (T)0;
    t.ra = s_r[a];
    t.rb = s_r[b];
    t.e0 = s_dx[b] - s_dx[a];
    t.e1 = s_dy[b] - s_dy[a];
    t.e2 = s_dz[b] - s_dz[a];
    t.rjk2 = t.e0 * t.e0 + t.e1 * t.e1 + t.e2 * t.e2;
    t.rjk = sqrt(t.rjk2);
    if (P == 0) {
        // sym_g4's own arithmetic (src/SymmetryFunctions.cpp:112-122)
        const double rijsq = t.ra * t.ra, riksq = t.rb * t.rb, rjksq = t.rjk * t.rjk;
        const double cosv = (rijsq + riksq - rjksq) / (2 * t.ra * t.rb);
        const double S = rijsq + riksq + rjksq;
#pragma unroll
        for (int l = 0; l < 2; ++l)
            if (l < wt.nl) t.B[l] = (T)(1 + wp.lam[l] * cosv);
#pragma unroll
        for (int e = 0; e < 2; ++e)
            if (e < wt.ne) t.E[e] = (T)exp(-wp.eta3[e] * S);
    } else {
        const double r2a = s_r2[a], r2b = s_r2[b];
        const float inv = rcp_refined((float)(2 * t.ra * t.rb));
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            if (l < wt.nl) {
                const int code = __builtin_amdgcn_readfirstlane(wp.lamcode[l]);
                if (code == 1)
                    t.B[l] = (float)((t.ra + t.rb) * (t.ra + t.rb) - t.rjk2) * inv;
                else if (code == -1)
                    t.B[l] = (float)(t.rjk2 - (t.ra - t.rb) * (t.ra - t.rb)) * inv;
                else
                    t.B[l] = 1.0f + (float)wp.lam[l] * ((float)(r2a + r2b - t.rjk2) * inv);
            }
        }
        const double S = r2a + r2b + t.rjk2;
#pragma unroll
        for (int e = 0; e < 2; ++e)
            if (e < wt.ne) t.E[e] = dexp(-wp.eta3[e] * S, R);
    }
    (void)R;
}

// B^(2^k), k = 0..nsq (nsq uniform): every integer zeta <= 63 of the grid is a
// product of these, so one squaring chain serves all rows of a triplet
template <typename T>
__device__ __forceinline__ void squarings(T b, int nsq, T (&sq)[6])
{
    sq[0] = b;
#pragma unroll
    for (int k = 1; k < 6; ++k) sq[k] = (k <= nsq) ? sq[k - 1] * sq[k - 1] : (T)0;
}

// product of sq[k] over the set bits of `bits` (uniform)
template <typename T>
__device__ __forceinline__ T pow_bits(const T (&sq)[6], int bits)
{
    T r = (T)1;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (bits & (1 << k)) r *= sq[k];
    return r;
}

// b^zeta_z (with zeta_z - 1 when minus_one) for b > 0; 0 for b <= 0 as the
// reference's guard (src/SymmetryFunctions.cpp:121)
template <typename T>
__device__ __forceinline__ T power_z(T b, const T (&sq)[6], int z, const WaveParams &wp,
                                     bool minus_one)
{
    const int iz = __builtin_amdgcn_readfirstlane(wp.izeta[z]);
    T v;
    if (iz >= (minus_one ? 1 : 0)) {
        v = pow_bits(sq, minus_one ? iz - 1 : iz);
    } else {
        const double zz = wp.zeta[z] - (minus_one ? 1.0 : 0.0);
        if (sizeof(T) == 8) v = (T)pow((double)b, zz);
        else v = (T)__builtin_amdgcn_exp2f((float)zz * __builtin_amdgcn_logf((float)b));
    }
    return b > (T)0 ? v : (T)0;
}

// LDS carve-up shared by both kernels (all arrays sized for maxn neighbours)
struct Lds {
    WaveParams *par;  // staged row parameters
    double *w, *wv;   // adjoint row weights of the current pair block (16 each)
    double *rc;       // cutoff matrix, n_species^2 <= kMaxWaveSpecies^2
    double *dx, *dy, *dz, *r, *r2, *fc, *dfc, *fx, *fy, *fz;
    int *sp, *orig, *l0, *l1;
};

__device__ __forceinline__ Lds carve(double *base, int maxn, int tmax, bool adjoint)
{
    Lds L;
    double *p = base;
    L.par = reinterpret_cast<WaveParams *>(p);
    p += (sizeof(WaveParams) + 7) / 8;
    L.rc = p; p += kMaxWaveSpecies * kMaxWaveSpecies;
    L.w = p; p += 16;
    L.wv = p; p += 16;
    L.dx = p; p += maxn;
    L.dy = p; p += maxn;
    L.dz = p; p += maxn;
    L.r = p; p += maxn;
    L.r2 = p; p += maxn;
    L.fc = p; p += maxn;
    if (adjoint) {
        L.dfc = p; p += maxn;
        L.fx = p; p += maxn;
        L.fy = p; p += maxn;
        L.fz = p; p += maxn;
    } else {
        L.dfc = L.fx = L.fy = L.fz = nullptr;
    }
    int *q = reinterpret_cast<int *>(p);
    L.sp = q; q += maxn;
    L.orig = q; q += maxn;
    L.l0 = q; q += tmax;
    L.l1 = q;
    return L;
}

// Phase 2: active pairs of compacted neighbours into L.l0, diagonal order.
// Returns the count (uniform).
template <int P, int NPB>
__device__ __forceinline__ int enumerate_triplets(const WaveTab &wt, const Lds &L, int M, int lane)
{
    const int ns = wt.n_species;
    const int T = M * (M - 1) / 2;
    int cnt = 0;
    int d = 1, a = lane;
    diag_norm(d, a, M);
    for (int t0 = 0; t0 < T; t0 += kW) {
        bool act = false;
        int code = 0;
        if (d < M) {
            const int b = a + d;
            const double e0 = L.dx[b] - L.dx[a], e1 = L.dy[b] - L.dy[a], e2 = L.dz[b] - L.dz[a];
            const double rjk2 = e0 * e0 + e1 * e1 + e2 * e2;
            const int sa = L.sp[a], sb = L.sp[b];
            const double rcjk = L.rc[sa * ns + sb];
            // the reference's test is sqrt(r_jk^2) < rc_jk; the squared form differs
            // only within an ulp of rc_jk, where fc_jk vanishes quadratically
            act = (P == 0) ? (sqrt(rjk2) < rcjk) : (rjk2 < rcjk * rcjk);
            code = a | (b << 10) | ((NPB > 1 ? pair_index(ns, sa, sb) : 0) << 20);
        }
        const unsigned long long m = __ballot(act);
        if (act) L.l0[cnt + prefix_count(m, lane)] = code;
        cnt += __popcll(m);
        a += kW;
        diag_norm(d, a, M);
    }
    return cnt;
}

// the active triplets of pair block p: L.l0 itself, or a compacted copy in L.l1
template <int NPB>
__device__ __forceinline__ const int *block_list(const Lds &L, int cnt, int p, int lane, int &nb)
{
    if (NPB == 1) {
        nb = cnt;
        return L.l0;
    }
    int k = 0;
    for (int t0 = 0; t0 < cnt; t0 += kW) {
        const int t = t0 + lane;
        const int code = t < cnt ? L.l0[t] : 0;
        const bool mine = t < cnt && (code >> 20) == p;
        const unsigned long long m = __ballot(mine);
        if (mine) L.l1[k + prefix_count(m, lane)] = code;
        k += __popcll(m);
    }
    __syncthreads();
    nb = k;
    return L.l1;
}

template <int P, int NSB, int NPB>
__global__ __launch_bounds__(64) void k_fwd_wave(WaveTab wt, int n_centers,
                                                 const int *__restrict__ centers,
                                                 const int *__restrict__ offs,
                                                 const int *__restrict__ nl,
                                                 const int *__restrict__ species,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ desc, int width, int maxn,
                                                 int tmax)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    extern __shared__ double lds_base[];
    const Lds L = carve(lds_base, maxn, tmax, false);
    T *s_fc = reinterpret_cast<T *>(L.fc);
    const int lane = threadIdx.x;
    const int ns = wt.n_species;
    const CentreRange cr = centre_range(n_centers);
    for (int k = lane; k < ns * ns; k += kW) L.rc[k] = wt.rc[k];
    {
        const int *src = reinterpret_cast<const int *>(wt.par);
        int *dst = reinterpret_cast<int *>(L.par);
        for (int k = lane; k < (int)(sizeof(WaveParams) / 4); k += kW) dst[k] = src[k];
    }
    __syncthreads();
    const WaveParams &wp = *L.par;
    const int wp_nsq = __builtin_amdgcn_readfirstlane(wp.nsq);
    for (int c = cr.first; c < cr.hi; c += cr.step) {
        const int i = centers ? centers[c] : c;
        const int beg = offs[c], N = offs[c + 1] - beg;
        const int si = species[i];
        const double xi0 = x[3 * (size_t)i], xi1 = x[3 * (size_t)i + 1], xi2 = x[3 * (size_t)i + 2];
        double *row = desc + (size_t)c * width;

        // ---- pairs ----
        double acc2[NSB * 8];
#pragma unroll
        for (int k = 0; k < NSB * 8; ++k) acc2[k] = 0.0;
        int M = 0;
        for (int j0 = 0; j0 < N; j0 += kW) {
            const int jj = j0 + lane;
            bool valid = false;
            int sj = 0;
            double d0 = 0, d1 = 0, d2 = 0, r2 = 0, r = 0, rcij = 0;
            if (jj < N) {
                const int j = nl[beg + jj];
                sj = species[j];
                d0 = x[3 * (size_t)j] - xi0;
                d1 = x[3 * (size_t)j + 1] - xi1;
                d2 = x[3 * (size_t)j + 2] - xi2;
                r2 = d0 * d0 + d1 * d1 + d2 * d2;
                r = sqrt(r2);
                rcij = L.rc[si * ns + sj];
                valid = r < rcij;  // r > rc is skipped; r == rc adds exactly 0
            }
            const unsigned long long m = __ballot(valid);
            if (valid) {
                const int pos = M + prefix_count(m, lane);
                L.dx[pos] = d0;
                L.dy[pos] = d1;
                L.dz[pos] = d2;
                L.r[pos] = r;
                L.r2[pos] = r2;
                L.sp[pos] = sj;
                const T fc = fcut_value(r, rcij, R);
                s_fc[pos] = fc;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (q < wt.nr2) {
                        const double dr = r - wp.rs2[q];
                        const double v = (double)(dexp(-wp.eta2[q] * dr * dr, R) * fc);
#pragma unroll
                        for (int s = 0; s < NSB; ++s)
                            if (NSB == 1 || s == sj) acc2[s * 8 + q] += v;
                    }
                }
            }
            M += __popcll(m);
        }
        reduce_add<NSB * 8>(acc2, lane, wp.slots, row);
        __syncthreads();

        // ---- triplets ----
        if (wt.has3 && M > 1) {
            const int cnt = enumerate_triplets<P, NPB>(wt, L, M, lane);
            __syncthreads();
            for (int p = 0; p < NPB; ++p) {
                int nb = 0;
                const int *lst = block_list<NPB>(L, cnt, p, lane, nb);
                double acc3[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) acc3[k] = 0.0;
                for (int t0 = 0; t0 < nb; t0 += kW) {
                    const int t = t0 + lane;
                    if (t < nb) {
                        const int code = lst[t];
                        const int a = code & 1023, b = (code >> 10) & 1023;
                        const double rcjk = L.rc[L.sp[a] * ns + L.sp[b]];
                        Trip<P> g;
                        triplet_geometry<P>(wt, wp, L.dx, L.dy, L.dz, L.r, L.r2, a, b, rcjk, g);
                        const T fjk = fcut_value(g.rjk, rcjk, R);
                        const T fc3 = s_fc[a] * s_fc[b] * fjk;
                        // row value = 2^(1-zeta) B^zeta * exp(-eta S) fc3, accumulated in
                        // FP64 as one fma of the two converted factors
                        double te[2];
#pragma unroll
                        for (int e = 0; e < 2; ++e) te[e] = (double)(g.E[e] * fc3);
#pragma unroll
                        for (int l = 0; l < 2; ++l) {
                            if (l < wt.nl) {
                                T sq[6];
                                squarings(g.B[l], wp_nsq, sq);
#pragma unroll
                                for (int z = 0; z < 4; ++z) {
                                    if (z < wt.nz) {
                                        const double pz =
                                            (double)((T)wp.coef[z] * power_z<T>(g.B[l], sq, z, wp, false));
#pragma unroll
                                        for (int e = 0; e < 2; ++e)
                                            if (e < wt.ne)
                                                acc3[(z * 2 + l) * 2 + e] = fma(pz, te[e], acc3[(z * 2 + l) * 2 + e]);
                                    }
                                }
                            }
                        }
                    }
                }
                reduce_add<16>(acc3, lane, wp.slots + 16 + p * 16, row);
                if (NPB > 1) __syncthreads();
            }
        }
        __syncthreads();
    }
}

template <int P, int NSB, int NPB>
__global__ __launch_bounds__(64) void k_adj_wave(WaveTab wt, int n_centers,
                                                 const int *__restrict__ centers,
                                                 const int *__restrict__ offs,
                                                 const int *__restrict__ nl,
                                                 const int *__restrict__ species,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ dE,
                                                 const int *__restrict__ inv_pos,
                                                 double *__restrict__ edge, int n_edges, int width,
                                                 int maxn, int tmax)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    extern __shared__ double lds_base[];
    const Lds L = carve(lds_base, maxn, tmax, true);
    T *s_fc = reinterpret_cast<T *>(L.fc);
    T *s_dfc = reinterpret_cast<T *>(L.dfc);
    const int lane = threadIdx.x;
    const int ns = wt.n_species;
    const CentreRange cr = centre_range(n_centers);
    for (int k = lane; k < ns * ns; k += kW) L.rc[k] = wt.rc[k];
    {
        const int *src = reinterpret_cast<const int *>(wt.par);
        int *dst = reinterpret_cast<int *>(L.par);
        for (int k = lane; k < (int)(sizeof(WaveParams) / 4); k += kW) dst[k] = src[k];
    }
    __syncthreads();
    const WaveParams &wp = *L.par;
    const int wp_nsq = __builtin_amdgcn_readfirstlane(wp.nsq);
    for (int c = cr.first; c < cr.hi; c += cr.step) {
        const int i = centers ? centers[c] : c;
        const int beg = offs[c], N = offs[c + 1] - beg;
        const int si = species[i];
        const double xi0 = x[3 * (size_t)i], xi1 = x[3 * (size_t)i + 1], xi2 = x[3 * (size_t)i + 2];
        const double *g = dE + (size_t)c * width;

        // ---- pairs: radial part of dE/d r_j, written as the accumulator's start ----
        int M = 0;
        for (int j0 = 0; j0 < N; j0 += kW) {
            const int jj = j0 + lane;
            bool valid = false;
            int sj = 0;
            double d0 = 0, d1 = 0, d2 = 0, r2 = 0, r = 0, rcij = 0;
            if (jj < N) {
                const int j = nl[beg + jj];
                sj = species[j];
                d0 = x[3 * (size_t)j] - xi0;
                d1 = x[3 * (size_t)j + 1] - xi1;
                d2 = x[3 * (size_t)j + 2] - xi2;
                r2 = d0 * d0 + d1 * d1 + d2 * d2;
                r = sqrt(r2);
                rcij = L.rc[si * ns + sj];
                valid = r < rcij;
                if (!valid) {
                    double *ej = edge + 3 * (size_t)inv_pos[beg + jj];
                    ej[0] = 0.0;
                    ej[1] = 0.0;
                    ej[2] = 0.0;
                }
            }
            const unsigned long long m = __ballot(valid);
            if (valid) {
                const int pos = M + prefix_count(m, lane);
                T fc, dfc;
                fcut(r, rcij, fc, dfc, R);
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (q < wt.nr2) {
                        const int slot = wp.slots[(NSB == 1 ? 0 : sj) * 8 + q];
                        const double dr = r - wp.rs2[q];
                        const T ex = dexp(-wp.eta2[q] * dr * dr, R);
                        const T dphi = ex * (T)(-2.0 * wp.eta2[q] * dr) * fc + ex * dfc;
                        s += g[slot] * (double)dphi;
                    }
                }
                const double ir = 1.0 / r;
                L.dx[pos] = d0;
                L.dy[pos] = d1;
                L.dz[pos] = d2;
                L.r[pos] = r;
                L.r2[pos] = r2;
                L.sp[pos] = sj;
                L.orig[pos] = jj;
                s_fc[pos] = fc;
                s_dfc[pos] = dfc;
                L.fx[pos] = s * d0 * ir;
                L.fy[pos] = s * d1 * ir;
                L.fz[pos] = s * d2 * ir;
            }
            M += __popcll(m);
        }
        __syncthreads();

        // ---- triplets ----
        if (wt.has3 && M > 1) {
            const int cnt = enumerate_triplets<P, NPB>(wt, L, M, lane);
            __syncthreads();
            for (int p = 0; p < NPB; ++p) {
                int nb = 0;
                const int *lst = block_list<NPB>(L, cnt, p, lane, nb);
                // row weights of this block: dE/dzeta * 2^(1-zeta), and times zeta lambda
                if (lane < 16) {
                    const int z = lane >> 2, l = (lane >> 1) & 1;
                    const int slot = wp.slots[16 + p * 16 + lane];
                    const double ww = slot >= 0 ? g[slot] * wp.coef[z] : 0.0;
                    L.w[lane] = ww;
                    L.wv[lane] = ww * wp.zeta[z] * wp.lam[l];
                }
                __syncthreads();
                for (int t0 = 0; t0 < nb; t0 += kW) {
                    const int t = t0 + lane;
                    if (t < nb) {
                        const int code = lst[t];
                        const int a = code & 1023, b = (code >> 10) & 1023;
                        const double rcjk = L.rc[L.sp[a] * ns + L.sp[b]];
                        Trip<P> q;
                        triplet_geometry<P>(wt, wp, L.dx, L.dy, L.dz, L.r, L.r2, a, b, rcjk, q);
                        T fjk, dfjk;
                        fcut(q.rjk, rcjk, fjk, dfjk, R);
                        const T fa = s_fc[a], fb = s_fc[b], dfa = s_dfc[a], dfb = s_dfc[b];
                        T U[2] = {(T)0, (T)0}, V[2] = {(T)0, (T)0};
#pragma unroll
                        for (int l = 0; l < 2; ++l) {
                            if (l < wt.nl) {
                                T sq[6];
                                squarings(q.B[l], wp_nsq, sq);
#pragma unroll
                                for (int z = 0; z < 4; ++z) {
                                    if (z < wt.nz) {
                                        const T pm1 = power_z<T>(q.B[l], sq, z, wp, true);
                                        const T pw = pm1 * q.B[l];
#pragma unroll
                                        for (int e = 0; e < 2; ++e) {
                                            U[e] += (T)L.w[(z * 2 + l) * 2 + e] * pw;
                                            V[e] += (T)L.wv[(z * 2 + l) * 2 + e] * pm1;
                                        }
                                    }
                                }
                            }
                        }
                        T SA = (T)0, SB = (T)0, SC = (T)0;
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            if (e < wt.ne) {
                                SA += V[e] * q.E[e];
                                const T ue = U[e] * q.E[e];
                                SB += ue;
                                SC += (T)wp.eta3[e] * ue;
                            }
                        }
                        // partial derivatives of cos(theta) w.r.t. r_ij, r_ik, r_jk
                        const double rijsq = q.ra * q.ra, riksq = q.rb * q.rb, rjksq = q.rjk * q.rjk;
                        T dcij, dcik, dcjk;
                        if (P == 0) {
                            dcij = (T)((rijsq - riksq + rjksq) / (2 * rijsq * q.rb));
                            dcik = (T)((riksq - rijsq + rjksq) / (2 * q.ra * riksq));
                            dcjk = (T)(-q.rjk / (q.ra * q.rb));
                        } else {
                            const float inv = rcp_refined((float)(2 * q.ra * q.rb));
                            dcij = (float)(rijsq - riksq + rjksq) * inv * (float)(1.0 / q.ra);
                            dcik = (float)(riksq - rijsq + rjksq) * inv * (float)(1.0 / q.rb);
                            dcjk = -(float)q.rjk * inv * 2.0f;
                        }
                        const T fc3 = fa * fb * fjk;
                        const T Aij = fc3 * (dcij * SA - (T)(2 * q.ra) * SC) + fb * fjk * dfa * SB;
                        const T Aik = fc3 * (dcik * SA - (T)(2 * q.rb) * SC) + fa * fjk * dfb * SB;
                        const T Ajk = fc3 * (dcjk * SA - (T)(2 * q.rjk) * SC) + fa * fb * dfjk * SB;
                        const double a1 = (double)Aij / q.ra, a2 = (double)Aik / q.rb,
                                     a3 = (double)Ajk / q.rjk;
                        atomicAdd(&L.fx[a], a1 * L.dx[a] - a3 * q.e0);
                        atomicAdd(&L.fy[a], a1 * L.dy[a] - a3 * q.e1);
                        atomicAdd(&L.fz[a], a1 * L.dz[a] - a3 * q.e2);
                        atomicAdd(&L.fx[b], a2 * L.dx[b] + a3 * q.e0);
                        atomicAdd(&L.fy[b], a2 * L.dy[b] + a3 * q.e1);
                        atomicAdd(&L.fz[b], a2 * L.dz[b] + a3 * q.e2);
                    }
                }
                if (NPB > 1) __syncthreads();
            }
            __syncthreads();
        }

        // ---- write edge vectors and the centre's own vector ----
        double sv[4] = {0.0, 0.0, 0.0, 0.0};
        for (int a0 = 0; a0 < M; a0 += kW) {
            const int a = a0 + lane;
            if (a < M) {
                const double f0 = L.fx[a], f1 = L.fy[a], f2 = L.fz[a];
                double *ej = edge + 3 * (size_t)inv_pos[beg + L.orig[a]];
                ej[0] = f0;
                ej[1] = f1;
                ej[2] = f2;
                sv[0] -= f0;
                sv[1] -= f1;
                sv[2] -= f2;
            }
        }
        reduce_scatter<4>(sv, lane);
        // lane l holds component (l >> 4) & 3
        if ((lane & 15) == 0 && (lane >> 4) < 3)
            edge[3 * (size_t)inv_pos[n_edges + c] + (lane >> 4)] = sv[0];
        __syncthreads();
    }
}

int grid_for(int n_centers)
{
    int g = std::min(n_centers, 16384);
    if (g >= 8) g = (g + 7) & ~7;
    return std::max(g, 1);
}

}  // namespace

size_t wave_lds_bytes(int max_nn, bool adjoint)
{
    const size_t maxn = std::max(max_nn, 1);
    const size_t tmax = std::max<size_t>(maxn * (maxn - 1) / 2, 1);
    return sizeof(double) * ((sizeof(WaveParams) + 7) / 8 + kMaxWaveSpecies * kMaxWaveSpecies + 32 +
                             maxn * (adjoint ? 10 : 6)) +
           sizeof(int) * (2 * maxn + 2 * tmax);
}

bool wave_build(const ldsf_sf *sf, WaveTab *wt, WaveParams *wp)
{
    std::memset(wt, 0, sizeof(*wt));
    std::memset(wp, 0, sizeof(*wp));
    int *slots = wp->slots;
    for (int k = 0; k < kWaveSlots; ++k) slots[k] = -1;
    const int ns = (int)sf->species.size();
    wt->n_species = ns;
    if (ns > kMaxWaveSpecies) return false;
    if (sf->species_resolved && ns > 2) return false;
    wt->nsb = (sf->species_resolved && ns == 2) ? 2 : 1;
    wt->npb = (sf->species_resolved && ns == 2) ? 3 : 1;
    // two-body rows: g2 only, at most 8
    const int n2 = (int)sf->k2.size();
    if (n2 > 8) return false;
    for (int q = 0; q < n2; ++q) {
        if (sf->k2[q] != G2) return false;
        wp->eta2[q] = sf->a2[q];
        wp->rs2[q] = sf->b2[q];
        for (int s = 0; s < wt->nsb; ++s) slots[s * 8 + q] = sf->slot2[q] + s * sf->str2[q];
    }
    wt->nr2 = n2;
    // three-body rows: g4 forming a zeta x lambda x eta grid
    const int n3 = (int)sf->k3.size();
    if (n3 == 0) return true;
    std::vector<double> zs, ls, es;
    auto idx = [](std::vector<double> &v, double x) {
        for (size_t k = 0; k < v.size(); ++k)
            if (v[k] == x) return (int)k;
        v.push_back(x);
        return (int)v.size() - 1;
    };
    std::map<int, int> seen;
    for (int q = 0; q < n3; ++q) {
        if (sf->k3[q] != G4) return false;
        const int z = idx(zs, sf->zeta3[q]), l = idx(ls, sf->lam3[q]), e = idx(es, sf->eta3[q]);
        if (zs.size() > 4 || ls.size() > 2 || es.size() > 2) return false;
        const int k = (z * 2 + l) * 2 + e;
        if (seen.count(k)) return false;
        seen[k] = q;
    }
    if ((int)seen.size() != (int)(zs.size() * ls.size() * es.size())) return false;
    wt->has3 = 1;
    wt->nz = (int)zs.size();
    wt->nl = (int)ls.size();
    wt->ne = (int)es.size();
    for (int z = 0; z < wt->nz; ++z) {
        wp->zeta[z] = zs[z];
        wp->coef[z] = std::pow(2, 1 - zs[z]);
        int iz = -1;
        if (zs[z] >= 0 && zs[z] <= 63 && zs[z] == std::floor(zs[z])) iz = (int)zs[z];
        wp->izeta[z] = iz;
        for (int k = 5; k >= 1; --k)
            if (iz >= (1 << k)) { wp->nsq = std::max(wp->nsq, k); break; }
    }
    for (int l = 0; l < wt->nl; ++l) {
        wp->lam[l] = ls[l];
        wp->lamcode[l] = ls[l] == 1.0 ? 1 : (ls[l] == -1.0 ? -1 : 0);
    }
    for (int e = 0; e < wt->ne; ++e) wp->eta3[e] = es[e];
    for (auto &kv : seen) {
        const int q = kv.second;
        for (int p = 0; p < wt->npb; ++p) slots[16 + p * 16 + kv.first] = sf->slot3[q] + p * sf->str3[q];
    }
    return true;
}

hipError_t launch_fwd_wave(const WaveTab &wt, int precision, int n_centers, const int *centers,
                           const int *offs, const int *nl, const int *species, const double *x,
                           double *desc, int width, int max_nn, hipStream_t s)
{
    if (n_centers <= 0) return hipSuccess;
    const int maxn = std::max(max_nn, 1);
    const int tmax = std::max(maxn * (maxn - 1) / 2, 1);
    const size_t lds = wave_lds_bytes(max_nn, false);
    const dim3 grid(grid_for(n_centers)), block(kW);
#define LDSF_FWD(P, NSB, NPB)                                                                     \
    hipLaunchKernelGGL((k_fwd_wave<P, NSB, NPB>), grid, block, lds, s, wt, n_centers, centers,    \
                       offs, nl, species, x, desc, width, maxn, tmax)
    if (wt.nsb == 1) {
        if (precision == 0) LDSF_FWD(0, 1, 1); else LDSF_FWD(1, 1, 1);
    } else {
        if (precision == 0) LDSF_FWD(0, 2, 3); else LDSF_FWD(1, 2, 3);
    }
#undef LDSF_FWD
    return hipGetLastError();
}

hipError_t launch_adj_wave(const WaveTab &wt, int precision, int n_centers, const int *centers,
                           const int *offs, const int *nl, const int *species, const double *x,
                           const double *dE, const int *inv_pos, double *edge, int n_edges,
                           int width, int max_nn, hipStream_t s)
{
    if (n_centers <= 0) return hipSuccess;
    const int maxn = std::max(max_nn, 1);
    const int tmax = std::max(maxn * (maxn - 1) / 2, 1);
    const size_t lds = wave_lds_bytes(max_nn, true);
    const dim3 grid(grid_for(n_centers)), block(kW);
#define LDSF_ADJ(P, NSB, NPB)                                                                     \
    hipLaunchKernelGGL((k_adj_wave<P, NSB, NPB>), grid, block, lds, s, wt, n_centers, centers,    \
                       offs, nl, species, x, dE, inv_pos, edge, n_edges, width, maxn, tmax)
    if (wt.nsb == 1) {
        if (precision == 0) LDSF_ADJ(0, 1, 1); else LDSF_ADJ(1, 1, 1);
    } else {
        if (precision == 0) LDSF_ADJ(0, 2, 3); else LDSF_ADJ(1, 2, 3);
    }
#undef LDSF_ADJ
    return hipGetLastError();
}

}  // namespace ldsf
