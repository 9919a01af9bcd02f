// Device-side scalar pieces of the symmetry functions, in two precision forms.
//
//  Form 0 (plain, LDSF_FP64): everything in FP64; the formulas and operation
//  order follow src/SymmetryFunctions.cpp:67-139 so results stay within a few
//  ulp of the reference (libm vs ocml is the only systematic difference).
//
//  Form 1 (LDSF_FP32_BULK): displacements, squared distances, the cutoff
//  tests and every accumulation stay FP64; the transcendental terms (exp, cos,
//  sin, powers) and the products that form one row's value are FP32 using the
//  gfx950 hardware transcendentals (v_exp_f32, v_cos_f32/v_sin_f32 take the
//  angle in revolutions, v_sqrt_f32, v_rcp_f32).  1 +/- cos(theta) is formed
//  from FP64 distances in the cancellation-free forms
//      1 + cos = ((rij + rik)^2 - rjk^2) / (2 rij rik)
//      1 - cos = (rjk^2 - (rij - rik)^2) / (2 rij rik).
#ifndef LDSF_SF_DEVICE_H_
#define LDSF_SF_DEVICE_H_

#include <hip/hip_runtime.h>

#include "ldsf_internal.h"

namespace ldsf {

// same literal as src/SymmetryFunctions.cpp:18
constexpr double kPi = 3.1415926535897932;
constexpr double kLn2 = 0.69314718055994531;  // ln 2: d 2^(1 - zeta) / d zeta = -ln 2 2^(1 - zeta)

template <int P>
struct Real;
template <>
struct Real<0> { using T = double; };
template <>
struct Real<1> { using T = float; };

// ---- the reference's FP64 expressions, evaluated without FMA contraction ----

// |d|^2 as d0*d0 + d1*d1 + d2*d2 (src/SymmetryFunctions.cpp:162,218-219)
__device__ __forceinline__ double dist2(double d0, double d1, double d2)
{
#pragma clang fp contract(off)
    return d0 * d0 + d1 * d1 + d2 * d2;
}

// argument of cut_cos: MY_PI * r / rcut (src/SymmetryFunctions.cpp:68)
__device__ __forceinline__ double ref_arg(double r, double rc)
{
#pragma clang fp contract(off)
    return kPi * r / rc;
}

struct RefTrip {
    double rijsq, riksq, rjksq, cosv, s2;
};

// squares from the magnitudes, cos_ijk by division, and the exponent's sum
// (src/SymmetryFunctions.cpp:112-122)
__device__ __forceinline__ RefTrip ref_trip(double rij, double rik, double rjk)
{
#pragma clang fp contract(off)
    RefTrip t;
    t.rijsq = rij * rij;
    t.riksq = rik * rik;
    t.rjksq = rjk * rjk;
    t.cosv = (t.rijsq + t.riksq - t.rjksq) / (2 * rij * rik);
    t.s2 = t.rijsq + t.riksq + t.rjksq;
    return t;
}

__device__ __forceinline__ double ref_base(double lam, double cosv)
{
#pragma clang fp contract(off)
    return 1 + lam * cosv;
}

// ---- the gather: a row's dE/dr records summed in ascending order ----

// s += records b .. e-1 of `edge` (3 doubles each), added one record after the
// other; four records' loads are issued before their adds, so a thread keeps
// 12 loads in flight (the order of the adds, and so the sum, is unchanged)
__device__ __forceinline__ void row_sum3(const double *__restrict__ edge, long b, long e,
                                         double &s0, double &s1, double &s2)
{
    long p = b;
    for (; p + 4 <= e; p += 4) {
        double v[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = edge[3 * p + k];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            s0 += v[3 * r];
            s1 += v[3 * r + 1];
            s2 += v[3 * r + 2];
        }
    }
    for (; p < e; ++p) {
        s0 += edge[3 * p];
        s1 += edge[3 * p + 1];
        s2 += edge[3 * p + 2];
    }
}

// ---- transposed neighbour lists ----

// Sorts ids[b .. b+m) ascending with one wavefront (all 64 lanes call it with
// the same b, m): a bitonic network across the lanes for m <= 64, an insertion
// sort on lane 0 for longer buckets.
__device__ inline void sort_bucket_wave(int *ids, long b, int m, int lane)
{
    if (m <= 1) return;
    if (m <= 64) {
        int v = lane < m ? ids[b + lane] : 0x7fffffff;
#pragma unroll
        for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const int u = __shfl_xor(v, j);
                const bool up = (lane & k) == 0, lower = (lane & j) == 0;
                v = (lower == up) ? min(v, u) : max(v, u);
            }
        }
        if (lane < m) ids[b + lane] = v;
        return;
    }
    if (lane != 0) return;
    for (long p = b + 1; p < b + m; ++p) {
        const int v = ids[p];
        long q = p - 1;
        while (q >= b && ids[q] > v) {
            ids[q + 1] = ids[q];
            --q;
        }
        ids[q + 1] = v;
    }
}

// One wavefront's share of the transposed list's last step: rows w, w + nw,
// ... of [0, n_rows) (w: this wave's index, nw: the waves sharing the rows),
// two rows at a time.  Each row's ids[in_start[a] .. in_start[a+1]) are
// sorted ascending and inv[id] = base + position written for every id
// (in_start and ids may sit in LDS: base then places them in the whole
// list).  Two rows of at most 32 ids each take a 32-lane bitonic network in
// each half-wave; longer ones take sort_bucket_wave one row after the other.
__device__ inline void sort_rows_inverse(int n_rows, const int *in_start, int *ids, int *inv,
                                         long w, long nw, int lane, long base = 0)
{
    const int half = lane >> 5, hl = lane & 31;
    for (long a0 = 2 * w; a0 < n_rows; a0 += 2 * nw) {
        const long a = a0 + half;
        const int b = a < n_rows ? in_start[a] : 0;
        const int m = a < n_rows ? in_start[a + 1] - b : 0;
        const int mmax = max(m, __shfl_xor(m, 32));  // the same in both halves
        if (mmax <= 32) {
            int v = hl < m ? ids[b + hl] : 0x7fffffff;
#pragma unroll
            for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
                for (int j = k >> 1; j > 0; j >>= 1) {
                    const int u = __shfl_xor(v, j);
                    const bool up = (hl & k) == 0, lower = (hl & j) == 0;
                    v = (lower == up) ? min(v, u) : max(v, u);
                }
            }
            if (hl < m) {
                ids[b + hl] = v;
                inv[v] = base + b + hl;
            }
        } else {
            for (int r = 0; r < 2; ++r) {
                const int br = __shfl(b, 32 * r), mr = __shfl(m, 32 * r);
                sort_bucket_wave(ids, br, mr, lane);
                // the sorted ids, re-read after the sort's own stores (m > 64
                // is sorted by lane 0 alone)
                __threadfence_block();
                for (int k = lane; k < mr; k += 64) inv[ids[br + k]] = base + br + k;
            }
        }
    }
}

// ---- cutoff function fc(r) = 0.5 (cos(pi r / rc) + 1) for r < rc, and dfc/dr ----

__device__ __forceinline__ void fcut(double r, double rc, double &fc, double &dfc, Real<0>)
{
    double a = kPi * r / rc;
    double s, c;
    sincos(a, &s, &c);
    fc = 0.5 * (c + 1.0);
    dfc = -0.5 * s * (kPi / rc);
}

// fc(r) = 0.5 (cos(pi r / rc) + 1) = sin^2(y) with y = (pi / 2) (rc - r) / rc,
// and dfc/dr = -(pi / rc) sin(y) cos(y).  y is formed in FP64; sin and cos of
// y in [0, pi/2] come from their Taylor series in FP32 (truncation below 1e-8
// there), so fc keeps ~1e-7 relative accuracy also where it vanishes at the
// cutoff -- a hardware cosine's absolute error would not.
__device__ __forceinline__ void sincos_quarter(float y, float &s, float &c)
{
    const float y2 = y * y;
    float t = 1.0f - y2 * (1.0f / 156.0f);
    t = 1.0f - y2 * (1.0f / 110.0f) * t;
    t = 1.0f - y2 * (1.0f / 72.0f) * t;
    t = 1.0f - y2 * (1.0f / 42.0f) * t;
    t = 1.0f - y2 * (1.0f / 20.0f) * t;
    t = 1.0f - y2 * (1.0f / 6.0f) * t;
    s = y * t;
    float u = 1.0f - y2 * (1.0f / 132.0f);
    u = 1.0f - y2 * (1.0f / 90.0f) * u;
    u = 1.0f - y2 * (1.0f / 56.0f) * u;
    u = 1.0f - y2 * (1.0f / 30.0f) * u;
    u = 1.0f - y2 * (1.0f / 12.0f) * u;
    c = 1.0f - y2 * 0.5f * u;
}

__device__ __forceinline__ void fcut(double r, double rc, float &fc, float &dfc, Real<1>)
{
    float s, c;
    sincos_quarter((float)(0.5 * kPi * (rc - r) / rc), s, c);
    fc = s * s;
    dfc = -(float)(kPi / rc) * s * c;
}

__device__ __forceinline__ double fcut_value(double r, double rc, Real<0>)
{
    return 0.5 * (cos(kPi * r / rc) + 1.0);
}

__device__ __forceinline__ float fcut_value(double r, double rc, Real<1>)
{
    float s, c;
    sincos_quarter((float)(0.5 * kPi * (rc - r) / rc), s, c);
    return s * s;
}

// ---- exp of an FP64 argument ----
__device__ __forceinline__ double dexp(double x, Real<0>) { return exp(x); }

// FP32-bulk exp: x log2(e) = n + f is split in FP64 (|f| <= 1/2), 2^f comes
// from v_exp_f32 and 2^n from ldexp, so the relative error stays ~1e-7 for
// any |x| (scaling the argument in FP32 would lose |x| * 6e-8).
__device__ __forceinline__ float dexp(double x, Real<1>)
{
    const double t = x * 1.4426950408889634;
    const double n = rint(t);
    return ldexpf(__builtin_amdgcn_exp2f((float)(t - n)), (int)n);
}

// ---- cos / sin for g3 ----
__device__ __forceinline__ void dsincos(double x, double &s, double &c, Real<0>) { sincos(x, &s, &c); }
__device__ __forceinline__ void dsincos(double x, float &s, float &c, Real<1>) { sincosf((float)x, &s, &c); }

// FP32 reciprocal refined by one Newton step (~0.5 ulp)
__device__ __forceinline__ float rcp_refined(float d)
{
    float r = __builtin_amdgcn_rcpf(d);
    return r * (2.0f - d * r);
}

// ---- powers: base^zeta for base > 0 ----
template <typename T>
__device__ __forceinline__ T ipow(T b, int n)
{
    T r = (T)1;
    while (n) {
        if (n & 1) r *= b;
        n >>= 1;
        if (n) b *= b;
    }
    return r;
}

__device__ __forceinline__ double dpow(double b, double zeta, int izeta, Real<0>)
{
    return izeta >= 0 ? ipow(b, izeta) : pow(b, zeta);
}

__device__ __forceinline__ float dpow(float b, double zeta, int izeta, Real<1>)
{
    return izeta >= 0 ? ipow(b, izeta) : __builtin_amdgcn_exp2f((float)zeta * __builtin_amdgcn_logf(b));
}

// base^(zeta-1) for the adjoint (zeta >= 1 on the integer path)
__device__ __forceinline__ double dpowm1(double b, double zeta, int izeta, Real<0>)
{
    return izeta >= 1 ? ipow(b, izeta - 1) : pow(b, zeta - 1.0);
}

__device__ __forceinline__ float dpowm1(float b, double zeta, int izeta, Real<1>)
{
    return izeta >= 1 ? ipow(b, izeta - 1)
                      : __builtin_amdgcn_exp2f((float)(zeta - 1.0) * __builtin_amdgcn_logf(b));
}

__device__ __forceinline__ int pair_index(int ns, int a, int b)
{
    if (a > b) { int t = a; a = b; b = t; }
    return a * ns - a * (a - 1) / 2 + (b - a);
}

}  // namespace ldsf

#endif  // LDSF_SF_DEVICE_H_
