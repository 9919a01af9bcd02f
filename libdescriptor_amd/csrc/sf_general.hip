// General symmetry-function kernels: one thread per centre atom, any mix of
// g1..g5 blocks, any neighbour count, plain or species-resolved layout, both
// precision forms.  They walk the neighbour list in the reference's order
// (src/SymmetryFunctions.cpp:141-259): pair terms of neighbour jj, then the
// triplets (jj, kk > jj).  The adjoint writes one 3-vector per neighbour-list
// edge (dE/d r_neighbour from this centre) and one per centre (dE/d r_centre);
// the atomic-free gather in sf_plan.hip sums them per atom.
//
// These kernels are the reference for every parameter set; the wave-per-centre
// kernels in sf_wave.hip take over where their preconditions hold.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sf_device.h"
#include "sf_kernels.h"

namespace ldsf {

namespace {

template <int P>
__global__ __launch_bounds__(256) void k_fwd_general(SFTab tab, int n_centers,
                                                     const int *__restrict__ centers,
                                                     const int *__restrict__ offs,
                                                     const int *__restrict__ nl,
                                                     const int *__restrict__ species,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ desc, DevCounts dc)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (dc.hdr) n_centers = dc.hdr[0];
    if (t >= n_centers) return;
    if (dc.req_flag && dc.req_flag[dc.req_of[t]]) return;
    const int i = centers ? centers[t] : t;
    const int ns = tab.n_species;
    double *out = desc + (size_t)t * tab.width;
    if (dc.out_row) {  // written (=): start from zero
        out = dc.out_row[t];
        for (int k = 0; k < tab.width; ++k) out[k] = 0.0;
    }
    const int beg = offs[t], end = offs[t + 1];
    const int si = species[i];
    const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
    for (int jj = beg; jj < end; ++jj) {
        const int j = nl[jj];
        const int sj = species[j];
        const double rcij = tab.rc[si * ns + sj];
        const double dj0 = x[3 * j] - xi0, dj1 = x[3 * j + 1] - xi1, dj2 = x[3 * j + 2] - xi2;
        const double rij = sqrt(dist2(dj0, dj1, dj2));
        if (!(rij < rcij)) continue;  // r > rc is skipped; r == rc contributes exactly 0
        const T fij = fcut_value(rij, rcij, R);
        for (int q = 0; q < tab.n2; ++q) {
            const int slot = tab.slot2[q] + sj * tab.str2[q];
            T v;
            const int kd = tab.k2[q];
            if (kd == G2) {
                const double dr = rij - tab.b2[q];
                v = dexp(-tab.a2[q] * dr * dr, R) * fij;
            } else if (kd == G3) {
                T s, c;
                dsincos(tab.a2[q] * rij, s, c, R);
                v = c * fij;
            } else {
                v = fij;
            }
            out[slot] += (double)v;
        }
        if (tab.n3 == 0) continue;
        for (int kk = jj + 1; kk < end; ++kk) {
            const int k = nl[kk];
            const int sk = species[k];
            const double rcik = tab.rc[si * ns + sk];
            const double rcjk = tab.rc[sj * ns + sk];
            const double dk0 = x[3 * k] - xi0, dk1 = x[3 * k + 1] - xi1, dk2 = x[3 * k + 2] - xi2;
            const double e0 = x[3 * k] - x[3 * j], e1 = x[3 * k + 1] - x[3 * j + 1],
                         e2 = x[3 * k + 2] - x[3 * j + 2];
            const double rik = sqrt(dist2(dk0, dk1, dk2));
            const double rjk = sqrt(dist2(e0, e1, e2));
            if (!(rik < rcik)) continue;
            const RefTrip rt = ref_trip(rij, rik, rjk);
            const double rijsq = rt.rijsq, riksq = rt.riksq, rjksq = rt.rjksq;
            const bool jk_in = rjk < rcjk;
            const T fik = fcut_value(rik, rcik, R);
            const T fjk = jk_in ? fcut_value(rjk, rcjk, R) : (T)0;
            const int pidx = pair_index(ns, sj, sk);
            // cos(theta_jik) and 1 +/- cos
            const double den = 2 * rij * rik;
            T cosv, onep, onem;
            if (P == 0) {
                cosv = (T)rt.cosv;
                onep = (T)0;
                onem = (T)0;
            } else {
                const float inv = 1.0f / (float)den;
                cosv = (T)(rijsq + riksq - rjksq) * inv;
                onep = (T)((rij + rik) * (rij + rik) - rjksq) * inv;
                onem = (T)(rjksq - (rij - rik) * (rij - rik)) * inv;
            }
            for (int q = 0; q < tab.n3; ++q) {
                const int kd = tab.k3[q];
                if (kd == G4 && !jk_in) continue;
                const double lam = tab.lam3[q];
                T base;
                if (P == 0) base = (T)ref_base(lam, (double)cosv);
                else base = lam == 1.0 ? onep : (lam == -1.0 ? onem : (T)1 + (T)lam * cosv);
                if (base <= (T)0) continue;
                const T cost = dpow(base, tab.zeta3[q], tab.izeta3[q], R);
                const double esum = kd == G4 ? rt.s2 : rijsq + riksq;
                const T eterm = dexp(-tab.eta3[q] * esum, R);
                T v = (T)tab.coef3[q] * cost * eterm * fij * fik;
                if (kd == G4) v *= fjk;
                out[tab.slot3[q] + pidx * tab.str3[q]] += (double)v;
            }
        }
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_adj_general(SFTab tab, int n_centers,
                                                     const int *__restrict__ centers,
                                                     const int *__restrict__ offs,
                                                     const int *__restrict__ nl,
                                                     const int *__restrict__ species,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ dE,
                                                     const int *__restrict__ inv_pos,
                                                     double *__restrict__ edge, int n_edges,
                                                     DevCounts dc)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (dc.hdr) {
        n_centers = dc.hdr[0];
        n_edges = dc.hdr[2];
    }
    if (t >= n_centers) return;
    if (dc.req_flag && dc.req_flag[dc.req_of[t]]) {
        double *v = edge + 3 * (size_t)inv_pos[n_edges + t];
        v[0] = 0.0;
        v[1] = 0.0;
        v[2] = 0.0;
        return;
    }
    const int i = centers ? centers[t] : t;
    const int ns = tab.n_species;
    const double *g = dc.in_row ? dc.in_row[t] : dE + (size_t)t * tab.width;
    const int beg = offs[t], end = offs[t + 1];
    for (int e = beg; e < end; ++e) {
        double *v = edge + 3 * (size_t)inv_pos[e];
        v[0] = 0.0;
        v[1] = 0.0;
        v[2] = 0.0;
    }
    const int si = species[i];
    const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int jj = beg; jj < end; ++jj) {
        const int j = nl[jj];
        const int sj = species[j];
        const double rcij = tab.rc[si * ns + sj];
        const double dj0 = x[3 * j] - xi0, dj1 = x[3 * j + 1] - xi1, dj2 = x[3 * j + 2] - xi2;
        const double rij = sqrt(dist2(dj0, dj1, dj2));
        if (!(rij < rcij)) continue;
        T fij, dfij;
        fcut(rij, rcij, fij, dfij, R);
        // d/dr of the pair rows, weighted by dE/dzeta
        double sp = 0.0;
        for (int q = 0; q < tab.n2; ++q) {
            const double w = g[tab.slot2[q] + sj * tab.str2[q]];
            const int kd = tab.k2[q];
            T d;
            if (kd == G2) {
                const double dr = rij - tab.b2[q];
                const T ex = dexp(-tab.a2[q] * dr * dr, R);
                d = ex * (T)(-2.0 * tab.a2[q] * dr) * fij + ex * dfij;
            } else if (kd == G3) {
                T s, c;
                dsincos(tab.a2[q] * rij, s, c, R);
                d = -(T)tab.a2[q] * s * fij + c * dfij;
            } else {
                d = dfij;
            }
            sp += w * (double)d;
        }
        const double ir = 1.0 / rij;
        double v0 = sp * dj0 * ir, v1 = sp * dj1 * ir, v2 = sp * dj2 * ir;
        double *ej = edge + 3 * (size_t)inv_pos[jj];
        ej[0] += v0; ej[1] += v1; ej[2] += v2;
        s0 -= v0; s1 -= v1; s2 -= v2;
        if (tab.n3 == 0) continue;
        for (int kk = jj + 1; kk < end; ++kk) {
            const int k = nl[kk];
            const int sk = species[k];
            const double rcik = tab.rc[si * ns + sk];
            const double rcjk = tab.rc[sj * ns + sk];
            const double dk0 = x[3 * k] - xi0, dk1 = x[3 * k + 1] - xi1, dk2 = x[3 * k + 2] - xi2;
            const double e0 = x[3 * k] - x[3 * j], e1 = x[3 * k + 1] - x[3 * j + 1],
                         e2 = x[3 * k + 2] - x[3 * j + 2];
            const double rik = sqrt(dist2(dk0, dk1, dk2));
            const double rjk = sqrt(dist2(e0, e1, e2));
            if (!(rik < rcik)) continue;
            const RefTrip rt = ref_trip(rij, rik, rjk);
            const double rijsq = rt.rijsq, riksq = rt.riksq, rjksq = rt.rjksq;
            const bool jk_in = rjk < rcjk;
            T fik, dfik, fjk = (T)0, dfjk = (T)0;
            fcut(rik, rcik, fik, dfik, R);
            if (jk_in) fcut(rjk, rcjk, fjk, dfjk, R);
            const int pidx = pair_index(ns, sj, sk);
            const double den = 2 * rij * rik;
            // cos(theta) and its partial derivatives w.r.t. rij, rik, rjk
            const T cosv = P == 0 ? (T)rt.cosv : (T)((rijsq + riksq - rjksq) / den);
            const T dc_ij = (T)((rijsq - riksq + rjksq) / (2 * rijsq * rik));
            const T dc_ik = (T)((riksq - rijsq + rjksq) / (2 * rij * riksq));
            const T dc_jk = (T)(-rjk / (rij * rik));
            T onep = (T)0, onem = (T)0;
            if (P == 1) {
                const float inv = 1.0f / (float)den;
                onep = (T)((rij + rik) * (rij + rik) - rjksq) * inv;
                onem = (T)(rjksq - (rij - rik) * (rij - rik)) * inv;
            }
            T aij = 0, aik = 0, ajk = 0;
            for (int q = 0; q < tab.n3; ++q) {
                const int kd = tab.k3[q];
                if (kd == G4 && !jk_in) continue;
                const double lam = tab.lam3[q];
                T base;
                if (P == 0) base = (T)ref_base(lam, (double)cosv);
                else base = lam == 1.0 ? onep : (lam == -1.0 ? onem : (T)1 + (T)lam * cosv);
                if (base <= (T)0) continue;
                const double zeta = tab.zeta3[q];
                const T cost = dpow(base, zeta, tab.izeta3[q], R);
                const T dcost = (T)zeta * dpowm1(base, zeta, tab.izeta3[q], R) * (T)lam;
                const double eta = tab.eta3[q];
                const T w = (T)g[tab.slot3[q] + pidx * tab.str3[q]] * (T)tab.coef3[q];
                if (kd == G4) {
                    const T ex = dexp(-eta * (rijsq + riksq + rjksq), R);
                    const T f3 = fij * fik * fjk;
                    const T we = w * ex;
                    aij += we * (dcost * dc_ij * f3 + cost * (T)(-2 * eta * rij) * f3 + cost * dfij * fik * fjk);
                    aik += we * (dcost * dc_ik * f3 + cost * (T)(-2 * eta * rik) * f3 + cost * fij * dfik * fjk);
                    ajk += we * (dcost * dc_jk * f3 + cost * (T)(-2 * eta * rjk) * f3 + cost * fij * fik * dfjk);
                } else {
                    const T ex = dexp(-eta * (rijsq + riksq), R);
                    const T f2 = fij * fik;
                    const T we = w * ex;
                    aij += we * (dcost * dc_ij * f2 + cost * (T)(-2 * eta * rij) * f2 + cost * dfij * fik);
                    aik += we * (dcost * dc_ik * f2 + cost * (T)(-2 * eta * rik) * f2 + cost * fij * dfik);
                    ajk += we * (dcost * dc_jk * f2);
                }
            }
            const double iik = 1.0 / rik, ijk = 1.0 / rjk;
            const double a1 = (double)aij * ir, a2 = (double)aik * iik, a3 = (double)ajk * ijk;
            const double vij0 = a1 * dj0, vij1 = a1 * dj1, vij2 = a1 * dj2;
            const double vik0 = a2 * dk0, vik1 = a2 * dk1, vik2 = a2 * dk2;
            const double vjk0 = a3 * e0, vjk1 = a3 * e1, vjk2 = a3 * e2;
            ej[0] += vij0 - vjk0; ej[1] += vij1 - vjk1; ej[2] += vij2 - vjk2;
            double *ek = edge + 3 * (size_t)inv_pos[kk];
            ek[0] += vik0 + vjk0; ek[1] += vik1 + vjk1; ek[2] += vik2 + vjk2;
            s0 -= vij0 + vik0; s1 -= vij1 + vik1; s2 -= vij2 + vik2;
        }
    }
    double *vs = edge + 3 * (size_t)inv_pos[n_edges + t];
    vs[0] = s0;
    vs[1] = s1;
    vs[2] = s2;
}


// ---- hyperparameter gradients (FP64) ----
//
// The reference's gradient duplicates the descriptor object itself for
// Enzyme (src/Descriptors.cpp:176-188, enzyme_dup desc_kind, d_desc_kind): the
// shadow receives dE/d(parameters) and dE/d(cutoff matrix), which the
// reference then deletes unread.  These kernels compute them: thread (row r,
// centre t) walks the centre's neighbours and triplets in the reference's
// order for its one parameter row and writes the row's parameter derivatives
// and its cutoff-matrix derivatives into column-major partials
// part[col][centre]; the columns are then summed over the centres in a fixed
// order (a block per (column, 4096 centres), each thread 16 consecutive
// centres then an LDS tree; the blocks in order), so the result is the same on
// every run.
constexpr int kParamMaxSpecies = 8;

// d cut_cos(r; rc) / d rc for r < rc
__device__ __forceinline__ double dcut_drc(double r, double rc)
{
    return 0.5 * sin(ref_arg(r, rc)) * kPi * r / (rc * rc);
}

__global__ __launch_bounds__(256) void k_param_rows(SFTab tab, int n_rows, int c0, int nc,
                                                    const int *__restrict__ centers,
                                                    const int *__restrict__ offs,
                                                    const int *__restrict__ nl,
                                                    const int *__restrict__ species,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ dE,
                                                    double *__restrict__ part)
{
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)n_rows * nc) return;
    const int r = (int)(id / nc), tl = (int)(id - (long)r * nc), t = c0 + tl;
    const int i = centers ? centers[t] : t;
    const int ns = tab.n_species;
    const double *g = dE + (size_t)t * tab.width;
    double dp[3] = {0.0, 0.0, 0.0};
    double drc[kParamMaxSpecies * kParamMaxSpecies];
    for (int k = 0; k < ns * ns; ++k) drc[k] = 0.0;
    const int beg = offs[t], end = offs[t + 1];
    const int si = species[i];
    const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
    int kd, ncols;
    if (r < tab.n2) {
        const int q = r;
        kd = tab.k2[q];
        ncols = kd == G2 ? 2 : (kd == G3 ? 1 : 0);
        const double a = tab.a2[q], b = tab.b2[q];
        for (int jj = beg; jj < end; ++jj) {
            const int j = nl[jj];
            const int sj = species[j];
            const double rcij = tab.rc[si * ns + sj];
            const double rij = sqrt(dist2(x[3 * j] - xi0, x[3 * j + 1] - xi1, x[3 * j + 2] - xi2));
            if (!(rij < rcij)) continue;  // r >= rc: the term and its derivatives vanish
            const double w = g[tab.slot2[q] + sj * tab.str2[q]];
            const double fc = fcut_value(rij, rcij, Real<0>());
            double val = 1.0;  // the term without its cutoff factor
            if (kd == G2) {
                const double dr = rij - b;
                val = exp(-a * dr * dr);
                dp[0] += w * (-dr * dr) * val * fc;
                dp[1] += w * (2.0 * a * dr) * val * fc;
            } else if (kd == G3) {
                val = cos(a * rij);
                dp[0] += w * (-rij * sin(a * rij)) * fc;
            }
            drc[si * ns + sj] += w * val * dcut_drc(rij, rcij);
        }
    } else {
        const int q = r - tab.n2;
        kd = tab.k3[q];
        ncols = 3;
        const double zeta = tab.zeta3[q], lam = tab.lam3[q], eta = tab.eta3[q];
        const double coef = tab.coef3[q];
        const int iz = tab.izeta3[q];
        for (int jj = beg; jj < end; ++jj) {
            const int j = nl[jj];
            const int sj = species[j];
            const double rcij = tab.rc[si * ns + sj];
            const double rij = sqrt(dist2(x[3 * j] - xi0, x[3 * j + 1] - xi1, x[3 * j + 2] - xi2));
            if (!(rij < rcij)) continue;
            const double fij = fcut_value(rij, rcij, Real<0>()), dfij = dcut_drc(rij, rcij);
            for (int kk = jj + 1; kk < end; ++kk) {
                const int k = nl[kk];
                const int sk = species[k];
                const double rcik = tab.rc[si * ns + sk];
                const double rcjk = tab.rc[sj * ns + sk];
                const double rik = sqrt(dist2(x[3 * k] - xi0, x[3 * k + 1] - xi1, x[3 * k + 2] - xi2));
                const double rjk = sqrt(dist2(x[3 * k] - x[3 * j], x[3 * k + 1] - x[3 * j + 1],
                                              x[3 * k + 2] - x[3 * j + 2]));
                if (!(rik < rcik)) continue;
                const bool jk_in = rjk < rcjk;
                if (kd == G4 && !jk_in) continue;
                const RefTrip rt = ref_trip(rij, rik, rjk);
                const double base = ref_base(lam, rt.cosv);
                if (base <= 0.0) continue;
                const double w = g[tab.slot3[q] + pair_index(ns, sj, sk) * tab.str3[q]];
                const double fik = fcut_value(rik, rcik, Real<0>()), dfik = dcut_drc(rik, rcik);
                const double fjk = kd == G4 ? fcut_value(rjk, rcjk, Real<0>()) : 1.0;
                const double dfjk = kd == G4 ? dcut_drc(rjk, rcjk) : 0.0;
                const double cost = dpow(base, zeta, iz, Real<0>());
                const double S = kd == G4 ? rt.s2 : rt.rijsq + rt.riksq;
                const double e = exp(-eta * S);
                const double F = fij * fik * fjk;
                const double term = coef * cost * e * F;
                dp[0] += w * term * (log(base) - kLn2);
                dp[1] += w * coef * zeta * dpowm1(base, zeta, iz, Real<0>()) * rt.cosv * e * F;
                dp[2] += w * (-S) * term;
                const double A = w * coef * cost * e;
                drc[si * ns + sj] += A * dfij * fik * fjk;
                drc[si * ns + sk] += A * fij * dfik * fjk;
                if (kd == G4) drc[sj * ns + sk] += A * fij * fik * dfjk;
            }
        }
    }
    const int pc = r < tab.n2 ? tab.pidx2[r] : tab.pidx3[r - tab.n2];
    for (int c = 0; c < ncols; ++c) part[(size_t)(pc + c) * nc + tl] = dp[c];
    const int rcbase = tab.n_params + r * ns * ns;
    for (int k = 0; k < ns * ns; ++k) part[(size_t)(rcbase + k) * nc + tl] = drc[k];
}

constexpr int kColChunk = 4096;  // centres per column-sum block: 256 threads x 16

__global__ __launch_bounds__(256) void k_colsum(int nc, const double *__restrict__ part,
                                                double *__restrict__ part2, int nb)
{
    __shared__ double sh[256];
    const int c = blockIdx.x, b = blockIdx.y;
    const size_t base = (size_t)c * nc;
    const int lo = b * kColChunk + threadIdx.x * 16;
    double s = 0.0;
    for (int k = 0; k < 16; ++k)
        if (lo + k < nc) s += part[base + lo + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part2[(size_t)c * nb + b] = sh[0];
}

__global__ void k_colsum_final(int ncol, int nb, const double *__restrict__ part2,
                               double *__restrict__ acc)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part2[(size_t)c * nb + b];
    acc[c] += s;
}

// dparams += the parameter columns; drcut[k] += the rows' cutoff columns, rows ascending
__global__ void k_param_finish(int n_params, int n_rows, int ns2, const double *__restrict__ acc,
                               double *__restrict__ dparams, double *__restrict__ drcut)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n_params) dparams[c] += acc[c];
    if (c < ns2 && drcut) {
        double s = 0.0;
        for (int r = 0; r < n_rows; ++r) s += acc[n_params + r * ns2 + c];
        drcut[c] += s;
    }
}

}  // namespace

hipError_t launch_fwd_general(const SFTab &tab, int precision, int n_centers, const int *centers,
                              const int *offs, const int *nl, const int *species, const double *x,
                              double *desc, hipStream_t s, DevCounts dc)
{
    if (n_centers <= 0) return hipSuccess;
    dim3 block(256), grid((n_centers + 255) / 256);
    if (precision == 0)
        hipLaunchKernelGGL(k_fwd_general<0>, grid, block, 0, s, tab, n_centers, centers, offs, nl,
                           species, x, desc, dc);
    else
        hipLaunchKernelGGL(k_fwd_general<1>, grid, block, 0, s, tab, n_centers, centers, offs, nl,
                           species, x, desc, dc);
    return hipGetLastError();
}

hipError_t launch_adj_general(const SFTab &tab, int precision, int n_centers, const int *centers,
                              const int *offs, const int *nl, const int *species, const double *x,
                              const double *dE, const int *inv_pos, double *edge, int n_edges,
                              hipStream_t s, DevCounts dc)
{
    if (n_centers <= 0) return hipSuccess;
    dim3 block(256), grid((n_centers + 255) / 256);
    if (precision == 0)
        hipLaunchKernelGGL(k_adj_general<0>, grid, block, 0, s, tab, n_centers, centers, offs, nl,
                           species, x, dE, inv_pos, edge, n_edges, dc);
    else
        hipLaunchKernelGGL(k_adj_general<1>, grid, block, 0, s, tab, n_centers, centers, offs, nl,
                           species, x, dE, inv_pos, edge, n_edges, dc);
    return hipGetLastError();
}

hipError_t launch_param_gradient(const SFTab &tab, int n_centers, const int *centers,
                                 const int *offs, const int *nl, const int *species,
                                 const double *x, const double *dE, double *dparams,
                                 double *drcut, hipStream_t s)
{
    const int n_rows = tab.n2 + tab.n3, ns2 = tab.n_species * tab.n_species;
    if (n_centers <= 0 || n_rows == 0) return hipSuccess;
    if (tab.n_species > kParamMaxSpecies) return hipErrorInvalidValue;
    const int ncol = tab.n_params + n_rows * ns2;
    // centres per chunk: partials of at most 32 M doubles (256 MB)
    const int chunk = (int)std::max<long>(1, std::min<long>(n_centers, (32L << 20) / ncol));
    const int nbmax = (chunk + kColChunk - 1) / kColChunk;
    double *part = nullptr, *part2 = nullptr, *acc = nullptr;
    hipError_t e = hipMalloc((void **)&part, sizeof(double) * (size_t)ncol * chunk);
    if (e == hipSuccess) e = hipMalloc((void **)&part2, sizeof(double) * (size_t)ncol * nbmax);
    if (e == hipSuccess) e = hipMalloc((void **)&acc, sizeof(double) * ncol);
    if (e == hipSuccess) e = hipMemsetAsync(acc, 0, sizeof(double) * ncol, s);
    for (int c0 = 0; e == hipSuccess && c0 < n_centers; c0 += chunk) {
        const int nc = std::min(chunk, n_centers - c0);
        const int nb = (nc + kColChunk - 1) / kColChunk;
        e = hipMemsetAsync(part, 0, sizeof(double) * (size_t)ncol * nc, s);
        if (e != hipSuccess) break;
        const long threads = (long)n_rows * nc;
        hipLaunchKernelGGL(k_param_rows, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                           tab, n_rows, c0, nc, centers, offs, nl, species, x, dE, part);
        hipLaunchKernelGGL(k_colsum, dim3(ncol, nb), dim3(256), 0, s, nc, part, part2, nb);
        hipLaunchKernelGGL(k_colsum_final, dim3((ncol + 255) / 256), dim3(256), 0, s, ncol, nb,
                           part2, acc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const int m = std::max(tab.n_params, ns2);
        hipLaunchKernelGGL(k_param_finish, dim3((m + 255) / 256), dim3(256), 0, s, tab.n_params,
                           n_rows, ns2, acc, dparams, drcut);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (double *p : {part, part2, acc})
        if (p) (void)hipFree(p);
    return e;
}

}  // namespace ldsf
