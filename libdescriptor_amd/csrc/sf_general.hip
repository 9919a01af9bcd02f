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

}  // namespace ldsf
