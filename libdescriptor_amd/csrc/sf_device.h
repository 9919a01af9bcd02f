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

template <int P>
struct Real;
template <>
struct Real<0> { using T = double; };
template <>
struct Real<1> { using T = float; };

// ---- cutoff function fc(r) = 0.5 (cos(pi r / rc) + 1) for r < rc, and dfc/dr ----

__device__ __forceinline__ void fcut(double r, double rc, double &fc, double &dfc, Real<0>)
{
    double a = kPi * r / rc;
    double s, c;
    sincos(a, &s, &c);
    fc = 0.5 * (c + 1.0);
    dfc = -0.5 * s * (kPi / rc);
}

__device__ __forceinline__ void fcut(double r, double rc, float &fc, float &dfc, Real<1>)
{
    // v_cos_f32 / v_sin_f32 compute cos(2 pi x), sin(2 pi x)
    float x = (float)(0.5 * r / rc);
    float c = __builtin_amdgcn_cosf(x);
    float s = __builtin_amdgcn_sinf(x);
    fc = 0.5f * (c + 1.0f);
    dfc = -0.5f * s * (float)(kPi / rc);
}

__device__ __forceinline__ double fcut_value(double r, double rc, Real<0>)
{
    return 0.5 * (cos(kPi * r / rc) + 1.0);
}

__device__ __forceinline__ float fcut_value(double r, double rc, Real<1>)
{
    return 0.5f * (__builtin_amdgcn_cosf((float)(0.5 * r / rc)) + 1.0f);
}

// ---- exp ----
__device__ __forceinline__ double dexp(double x, Real<0>) { return exp(x); }
__device__ __forceinline__ float dexp(float x, Real<1>) { return __expf(x); }

// ---- cos / sin for g3 ----
__device__ __forceinline__ void dsincos(double x, double &s, double &c, Real<0>) { sincos(x, &s, &c); }
__device__ __forceinline__ void dsincos(float x, float &s, float &c, Real<1>)
{
    constexpr float inv2pi = 0.15915494309189535f;
    float t = x * inv2pi;
    s = __builtin_amdgcn_sinf(t);
    c = __builtin_amdgcn_cosf(t);
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
