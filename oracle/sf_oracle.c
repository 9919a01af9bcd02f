/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, FP64 restatement of libdescriptor's Behler symmetry-function
 * path, written to be read side by side with the reference:
 *
 *   forward   : src/SymmetryFunctions.cpp:67-69   (cut_cos)
 *               src/SymmetryFunctions.cpp:75-139  (sym_g1 .. sym_g5)
 *               src/SymmetryFunctions.cpp:141-259 (SymmetryFunctions::compute,
 *                                                  loop order and += order kept)
 *               src/SymmetryFunctions.cpp:43-65, 429-450 (slot layout)
 *               src/Descriptors.cpp:121-163       (atom walk and batch walk)
 *   adjoint   : the reverse-mode VJP that the reference obtains from Enzyme
 *               (src/Descriptors.cpp:165-190, 263-292, 368-395), derived by
 *               hand here term by term, in the same loop order.
 *   FD check  : include/finite_difference.hpp:46-143 (order-8 stencil) and
 *               src/Descriptors.cpp:488-518 (num_gradient_single_atom).
 *
 * The species-resolved layout (per-neighbour-species G1/G2/G3 blocks and
 * per-unordered-species-pair G4/G5 blocks) is a build-side extension; with one
 * species it is the reference layout, and its blocks sum to the reference
 * descriptor.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline; the
 * product path never calls it.
 *
 * Pins (see tests/test_oracle.py): closed forms for one neighbour and for an
 * equilateral triple, mpmath brute force at 50 digits, the compiled reference
 * forward (oracle/_ref, built from the reference sources where they lie),
 * central finite differences, translation/rotation invariance of the adjoint
 * and linearity in dE/dzeta.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* same literal as src/SymmetryFunctions.cpp:18 */
#define ORC_PI 3.1415926535897932

typedef struct {
    int n_species;
    double *rcut;      /* n_species x n_species, row-major (helper.hpp:2056-2067) */
    int n_blocks;
    int *kind;         /* 1..5 for g1..g5 (SymmetryFunctions.cpp:47-51) */
    int *rows;
    int *cols;
    int *start;        /* first slot of each block */
    double **par;      /* per block, rows x cols row-major */
    int has_three_body;
    int species_resolved;
    int width;
} orc_sf;

static int is_two_body(int k) { return k == 1 || k == 2 || k == 3; }
static int is_three_body(int k) { return k == 4 || k == 5; }

/* number of unordered species pairs */
static int n_pairs(int ns) { return ns * (ns + 1) / 2; }

/* index of the unordered pair (a,b), a<=b, lexicographic (0,0),(0,1)..(1,1).. */
static int pair_index(int ns, int a, int b)
{
    if (a > b) { int t = a; a = b; b = t; }
    return a * ns - a * (a - 1) / 2 + (b - a);
}

orc_sf *orc_sf_create(int n_species, const double *rcut, int n_blocks,
                      const int *kind, const int *rows, const int *cols,
                      const double *params, int species_resolved)
{
    orc_sf *sf = (orc_sf *)calloc(1, sizeof(orc_sf));
    sf->n_species = n_species;
    sf->rcut = (double *)malloc(sizeof(double) * n_species * n_species);
    memcpy(sf->rcut, rcut, sizeof(double) * n_species * n_species);
    sf->n_blocks = n_blocks;
    sf->kind = (int *)malloc(sizeof(int) * n_blocks);
    sf->rows = (int *)malloc(sizeof(int) * n_blocks);
    sf->cols = (int *)malloc(sizeof(int) * n_blocks);
    sf->start = (int *)malloc(sizeof(int) * n_blocks);
    sf->par = (double **)malloc(sizeof(double *) * n_blocks);
    sf->species_resolved = species_resolved;
    int off = 0, w = 0;
    for (int p = 0; p < n_blocks; ++p) {
        sf->kind[p] = kind[p];
        sf->rows[p] = rows[p];
        sf->cols[p] = cols[p];
        sf->par[p] = (double *)malloc(sizeof(double) * (rows[p] * cols[p] + 1));
        memcpy(sf->par[p], params + off, sizeof(double) * rows[p] * cols[p]);
        off += rows[p] * cols[p];
        sf->start[p] = w;
        int mult = 1;
        if (species_resolved)
            mult = is_three_body(kind[p]) ? n_pairs(n_species) : n_species;
        w += rows[p] * mult;
        if (is_three_body(kind[p])) sf->has_three_body = 1;
    }
    sf->width = w;
    return sf;
}

void orc_sf_destroy(orc_sf *sf)
{
    if (!sf) return;
    for (int p = 0; p < sf->n_blocks; ++p) free(sf->par[p]);
    free(sf->par); free(sf->kind); free(sf->rows); free(sf->cols);
    free(sf->start); free(sf->rcut); free(sf);
}

int orc_sf_width(const orc_sf *sf) { return sf->width; }

/* ---- the reference's scalar functions (SymmetryFunctions.cpp:67-139) ---- */

static double cut_cos(double r, double rcut)
{
    return (r < rcut) ? 0.5 * (cos(ORC_PI * r / rcut) + 1.0) : 0.0;
}

/* d cut_cos / dr; zero outside the cutoff, where cut_cos is the constant 0 */
static double d_cut_cos(double r, double rcut)
{
    return (r < rcut) ? -0.5 * sin(ORC_PI * r / rcut) * (ORC_PI / rcut) : 0.0;
}

static double sym_g5(double zeta, double lambda, double eta, const double *r, const double *rcut)
{
    double rij = r[0], rik = r[1];
    if (rij > rcut[0] || rik > rcut[1]) return 0.0;
    double rjk = r[2];
    double rijsq = rij * rij, riksq = rik * rik, rjksq = rjk * rjk;
    double cos_ijk = (rijsq + riksq - rjksq) / (2 * rij * rik);
    double base = 1.0 + lambda * cos_ijk;
    double costerm = (base <= 0) ? 0.0 : pow(base, zeta);
    double eterm = exp(-eta * (rijsq + riksq));
    return pow(2, 1 - zeta) * costerm * eterm * cut_cos(rij, rcut[0]) * cut_cos(rik, rcut[1]);
}

static double sym_g4(double zeta, double lambda, double eta, const double *r, const double *rcut)
{
    double rij = r[0], rik = r[1], rjk = r[2];
    if (rij > rcut[0] || rik > rcut[1] || rjk > rcut[2]) return 0.0;
    double rijsq = rij * rij, riksq = rik * rik, rjksq = rjk * rjk;
    double cos_ijk = (rijsq + riksq - rjksq) / (2 * rij * rik);
    double base = 1 + lambda * cos_ijk;
    double costerm = (base <= 0) ? 0.0 : pow(base, zeta);
    double eterm = exp(-eta * (rijsq + riksq + rjksq));
    return pow(2, 1 - zeta) * costerm * eterm * cut_cos(rij, rcut[0])
           * cut_cos(rik, rcut[1]) * cut_cos(rjk, rcut[2]);
}

static double sym_g3(double kappa, double r, double rcut) { return cos(kappa * r) * cut_cos(r, rcut); }

static double sym_g2(double eta, double Rs, double r, double rcut)
{
    return exp(-eta * (r - Rs) * (r - Rs)) * cut_cos(r, rcut);
}

static double sym_g1(double r, double rcut) { return cut_cos(r, rcut); }

/* ---- derivatives of the scalar functions w.r.t. their distances ---- */

/* d phi / d r for the two-body functions */
static double d_sym_2b(int kind, const double *prm, double r, double rcut)
{
    if (kind == 1) return d_cut_cos(r, rcut);
    if (kind == 2) {
        double eta = prm[0], Rs = prm[1];
        double e = exp(-eta * (r - Rs) * (r - Rs));
        return e * (-2.0 * eta * (r - Rs)) * cut_cos(r, rcut) + e * d_cut_cos(r, rcut);
    }
    /* kind 3 */
    double kappa = prm[0];
    return -kappa * sin(kappa * r) * cut_cos(r, rcut) + cos(kappa * r) * d_cut_cos(r, rcut);
}

/* d phi / d (rij, rik, rjk) for g4 (three_body == 4) and g5 (== 5) */
static void d_sym_3b(int kind, double zeta, double lambda, double eta, const double *r,
                     const double *rcut, double *d)
{
    double rij = r[0], rik = r[1], rjk = r[2];
    d[0] = d[1] = d[2] = 0.0;
    if (rij > rcut[0] || rik > rcut[1]) return;
    if (kind == 4 && rjk > rcut[2]) return;
    double rijsq = rij * rij, riksq = rik * rik, rjksq = rjk * rjk;
    double cos_ijk = (rijsq + riksq - rjksq) / (2 * rij * rik);
    /* partial derivatives of cos(theta_ijk) = (a^2+b^2-c^2)/(2ab) */
    double dc_ij = (rijsq - riksq + rjksq) / (2 * rijsq * rik);
    double dc_ik = (riksq - rijsq + rjksq) / (2 * rij * riksq);
    double dc_jk = -rjk / (rij * rik);
    double base = 1 + lambda * cos_ijk;
    if (base <= 0) return;
    double costerm = pow(base, zeta);
    double dcost = zeta * pow(base, zeta - 1) * lambda; /* d costerm / d cos */
    double p2 = pow(2, 1 - zeta);
    double fij = cut_cos(rij, rcut[0]), fik = cut_cos(rik, rcut[1]);
    double dfij = d_cut_cos(rij, rcut[0]), dfik = d_cut_cos(rik, rcut[1]);
    if (kind == 4) {
        double fjk = cut_cos(rjk, rcut[2]), dfjk = d_cut_cos(rjk, rcut[2]);
        double e = exp(-eta * (rijsq + riksq + rjksq));
        double f3 = fij * fik * fjk;
        d[0] = p2 * (dcost * dc_ij * e * f3 + costerm * e * (-2 * eta * rij) * f3
                     + costerm * e * dfij * fik * fjk);
        d[1] = p2 * (dcost * dc_ik * e * f3 + costerm * e * (-2 * eta * rik) * f3
                     + costerm * e * fij * dfik * fjk);
        d[2] = p2 * (dcost * dc_jk * e * f3 + costerm * e * (-2 * eta * rjk) * f3
                     + costerm * e * fij * fik * dfjk);
    } else {
        double e = exp(-eta * (rijsq + riksq));
        double f2 = fij * fik;
        d[0] = p2 * (dcost * dc_ij * e * f2 + costerm * e * (-2 * eta * rij) * f2
                     + costerm * e * dfij * fik);
        d[1] = p2 * (dcost * dc_ik * e * f2 + costerm * e * (-2 * eta * rik) * f2
                     + costerm * e * fij * dfik);
        d[2] = p2 * (dcost * dc_jk * e * f2);
    }
}

static int slot2(const orc_sf *sf, int p, int js)
{
    return sf->start[p] + (sf->species_resolved ? js * sf->rows[p] : 0);
}

static int slot3(const orc_sf *sf, int p, int js, int ks)
{
    return sf->start[p]
           + (sf->species_resolved ? pair_index(sf->n_species, js, ks) * sf->rows[p] : 0);
}

/*
 * One centre's descriptor, added into desc[0:width].
 * Restates SymmetryFunctions::compute (src/SymmetryFunctions.cpp:141-259).
 */
void orc_compute_atom(const orc_sf *sf, int index, const int *species, const int *neigh_list,
                      int number_of_neigh, const double *coords, double *desc)
{
    const int ns = sf->n_species;
    const int iS = species[index];
    const double *xi = coords + 3 * index;
    for (int jj = 0; jj < number_of_neigh; ++jj) {
        int j = neigh_list[jj];
        int jS = species[j];
        double rcutij = sf->rcut[iS * ns + jS];
        double rij[3];
        for (int d = 0; d < 3; ++d) rij[d] = coords[3 * j + d] - xi[d];
        double rijsq = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
        double rijmag = sqrt(rijsq);
        if (rijmag > rcutij) continue;

        for (int p = 0; p < sf->n_blocks; ++p) {
            int kd = sf->kind[p];
            if (!is_two_body(kd)) continue;
            int idx = slot2(sf, p, jS);
            for (int q = 0; q < sf->rows[p]; ++q) {
                const double *prm = sf->par[p] + q * sf->cols[p];
                double gc = 0.0;
                if (kd == 1) gc = sym_g1(rijmag, rcutij);
                else if (kd == 2) gc = sym_g2(prm[0], prm[1], rijmag, rcutij);
                else gc = sym_g3(prm[0], rijmag, rcutij);
                desc[idx] += gc;
                ++idx;
            }
        }
        if (!sf->has_three_body) continue;

        for (int kk = jj + 1; kk < number_of_neigh; ++kk) {
            int k = neigh_list[kk];
            int kS = species[k];
            double rcutik = sf->rcut[iS * ns + kS];
            double rcutjk = sf->rcut[jS * ns + kS];
            double rik[3], rjk[3];
            for (int d = 0; d < 3; ++d) {
                rik[d] = coords[3 * k + d] - xi[d];
                rjk[d] = coords[3 * k + d] - coords[3 * j + d];
            }
            double riksq = rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2];
            double rjksq = rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2];
            double rikmag = sqrt(riksq);
            double rjkmag = sqrt(rjksq);
            if (rikmag > rcutik) continue;
            double rvec[3] = {rijmag, rikmag, rjkmag};
            double rcvec[3] = {rcutij, rcutik, rcutjk};
            for (int p = 0; p < sf->n_blocks; ++p) {
                int kd = sf->kind[p];
                if (!is_three_body(kd)) continue;
                int idx = slot3(sf, p, jS, kS);
                for (int q = 0; q < sf->rows[p]; ++q) {
                    const double *prm = sf->par[p] + q * sf->cols[p];
                    double gc = (kd == 4) ? sym_g4(prm[0], prm[1], prm[2], rvec, rcvec)
                                          : sym_g5(prm[0], prm[1], prm[2], rvec, rcvec);
                    desc[idx] += gc;
                    ++idx;
                }
            }
        }
    }
}

/*
 * One centre's VJP: d_coords[n_total][3] += (d desc_i / d coords)^T g
 * where g = dE/dzeta_i (width entries).  Same loop order as the forward.
 */
void orc_gradient_atom(const orc_sf *sf, int index, const int *species, const int *neigh_list,
                       int number_of_neigh, const double *coords, const double *g,
                       double *d_coords)
{
    const int ns = sf->n_species;
    const int iS = species[index];
    const double *xi = coords + 3 * index;
    double *gi = d_coords + 3 * index;
    for (int jj = 0; jj < number_of_neigh; ++jj) {
        int j = neigh_list[jj];
        int jS = species[j];
        double rcutij = sf->rcut[iS * ns + jS];
        double rij[3];
        for (int d = 0; d < 3; ++d) rij[d] = coords[3 * j + d] - xi[d];
        double rijsq = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
        double rijmag = sqrt(rijsq);
        if (rijmag > rcutij) continue;
        double *gj = d_coords + 3 * j;

        double s = 0.0;
        for (int p = 0; p < sf->n_blocks; ++p) {
            int kd = sf->kind[p];
            if (!is_two_body(kd)) continue;
            int idx = slot2(sf, p, jS);
            for (int q = 0; q < sf->rows[p]; ++q) {
                const double *prm = sf->par[p] + q * sf->cols[p];
                s += g[idx] * d_sym_2b(kd, prm, rijmag, rcutij);
                ++idx;
            }
        }
        for (int d = 0; d < 3; ++d) {
            double v = s * rij[d] / rijmag;
            gj[d] += v;
            gi[d] -= v;
        }
        if (!sf->has_three_body) continue;

        for (int kk = jj + 1; kk < number_of_neigh; ++kk) {
            int k = neigh_list[kk];
            int kS = species[k];
            double rcutik = sf->rcut[iS * ns + kS];
            double rcutjk = sf->rcut[jS * ns + kS];
            double rik[3], rjk[3];
            for (int d = 0; d < 3; ++d) {
                rik[d] = coords[3 * k + d] - xi[d];
                rjk[d] = coords[3 * k + d] - coords[3 * j + d];
            }
            double riksq = rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2];
            double rjksq = rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2];
            double rikmag = sqrt(riksq);
            double rjkmag = sqrt(rjksq);
            if (rikmag > rcutik) continue;
            double rvec[3] = {rijmag, rikmag, rjkmag};
            double rcvec[3] = {rcutij, rcutik, rcutjk};
            double a_ij = 0.0, a_ik = 0.0, a_jk = 0.0;
            for (int p = 0; p < sf->n_blocks; ++p) {
                int kd = sf->kind[p];
                if (!is_three_body(kd)) continue;
                int idx = slot3(sf, p, jS, kS);
                for (int q = 0; q < sf->rows[p]; ++q) {
                    const double *prm = sf->par[p] + q * sf->cols[p];
                    double dphi[3];
                    d_sym_3b(kd, prm[0], prm[1], prm[2], rvec, rcvec, dphi);
                    a_ij += g[idx] * dphi[0];
                    a_ik += g[idx] * dphi[1];
                    a_jk += g[idx] * dphi[2];
                    ++idx;
                }
            }
            double *gk = d_coords + 3 * k;
            for (int d = 0; d < 3; ++d) {
                double vij = a_ij * rij[d] / rijmag;
                double vik = a_ik * rik[d] / rikmag;
                double vjk = a_jk * rjk[d] / rjkmag;
                gj[d] += vij - vjk;
                gk[d] += vik + vjk;
                gi[d] -= vij + vik;
            }
        }
    }
}

/* Descriptor::compute (src/Descriptors.cpp:121-136): desc += over all centres */
void orc_compute(const orc_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                 const int *number_of_neighs, const double *coords, double *desc)
{
    const int *nptr = neighbor_list;
    double *dptr = desc;
    for (int i = 0; i < n_atoms; ++i) {
        orc_compute_atom(sf, i, species, nptr, number_of_neighs[i], coords, dptr);
        nptr += number_of_neighs[i];
        dptr += sf->width;
    }
}

/* Descriptor::gradient (src/Descriptors.cpp:165-190), SF arm, without the
 * re-accumulation of the forward into desc that Enzyme's augmented pass does */
void orc_gradient(const orc_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                  const int *number_of_neighs, const double *coords, const double *dE_dzeta,
                  double *d_coords)
{
    const int *nptr = neighbor_list;
    for (int i = 0; i < n_atoms; ++i) {
        orc_gradient_atom(sf, i, species, nptr, number_of_neighs[i], coords,
                          dE_dzeta + (size_t)i * sf->width, d_coords);
        nptr += number_of_neighs[i];
    }
}

/* Descriptor::compute_batch (src/Descriptors.cpp:138-163) */
void orc_compute_batch(const orc_sf *sf, int n_configurations, const int *n_atoms, const int *ptr,
                       const int *species, const int *neighbor_list, const int *number_of_neighs,
                       const double *coords, double *desc)
{
    const int *nptr = neighbor_list;
    double *dptr = desc;
    int idx = 0;
    for (int c = 0; c < n_configurations; ++c) {
        for (int j = 0; j < n_atoms[c]; ++j) {
            orc_compute_atom(sf, j + ptr[c], species, nptr, number_of_neighs[idx], coords, dptr);
            nptr += number_of_neighs[idx];
            dptr += sf->width;
            idx++;
        }
    }
}

/* Descriptor::gradient_batch (src/Descriptors.cpp:263-292), SF arm */
void orc_gradient_batch(const orc_sf *sf, int n_configurations, const int *n_atoms, const int *ptr,
                        const int *species, const int *neighbor_list, const int *number_of_neighs,
                        const double *coords, const double *dE_dzeta, double *d_coords)
{
    const int *nptr = neighbor_list;
    int idx = 0;
    for (int c = 0; c < n_configurations; ++c) {
        for (int j = 0; j < n_atoms[c]; ++j) {
            orc_gradient_atom(sf, j + ptr[c], species, nptr, number_of_neighs[idx], coords,
                              dE_dzeta + (size_t)idx * sf->width, d_coords);
            nptr += number_of_neighs[idx];
            idx++;
        }
    }
}

/*
 * Multi-threaded forward + adjoint over the first n_centers centres, for the
 * CPU baseline timing: OpenMP over centres, per-thread d_coords buffers
 * reduced at the end.  offsets[n_centers+1] is the CSR start of each centre.
 */
void orc_fwd_adj_omp(const orc_sf *sf, int n_centers, int n_total, const int *species,
                     const int *neighbor_list, const int *offsets, const double *coords,
                     const double *dE_dzeta, double *desc, double *d_coords, int n_threads)
{
#ifdef _OPENMP
    if (n_threads < 1) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    double *buf = (double *)calloc((size_t)n_threads * n_total * 3, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(n_threads)
#endif
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *mine = buf + (size_t)t * n_total * 3;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int i = 0; i < n_centers; ++i) {
            int nn = offsets[i + 1] - offsets[i];
            orc_compute_atom(sf, i, species, neighbor_list + offsets[i], nn, coords,
                             desc + (size_t)i * sf->width);
            orc_gradient_atom(sf, i, species, neighbor_list + offsets[i], nn, coords,
                              dE_dzeta + (size_t)i * sf->width, mine);
        }
    }
    for (int t = 0; t < n_threads; ++t) {
        const double *src = buf + (size_t)t * n_total * 3;
        for (size_t a = 0; a < (size_t)n_total * 3; ++a) d_coords[a] += src[a];
    }
    free(buf);
}

/* The batch walk of src/Descriptors.cpp:138-163 / 263-292 (centre j of
 * configuration c is row config_ptr[c] + j, its neighbours follow those of the
 * previous centre), OpenMP over configurations, per-thread dE/dr buffers. */
void orc_fwd_adj_batch_omp(const orc_sf *sf, int n_configs, const int *n_atoms,
                           const int *config_ptr, int n_total, const int *species,
                           const int *neighbor_list, const int *number_of_neighs,
                           const double *coords, const double *dE_dzeta, double *desc,
                           double *d_coords, int n_threads)
{
#ifdef _OPENMP
    if (n_threads < 1) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    /* first centre and first neighbour entry of each configuration */
    long *c0 = (long *)malloc(sizeof(long) * (n_configs + 1));
    long *e0 = (long *)malloc(sizeof(long) * (n_configs + 1));
    c0[0] = 0;
    e0[0] = 0;
    for (int c = 0; c < n_configs; ++c) {
        long e = 0;
        for (int j = 0; j < n_atoms[c]; ++j) e += number_of_neighs[c0[c] + j];
        c0[c + 1] = c0[c] + n_atoms[c];
        e0[c + 1] = e0[c] + e;
    }
    double *buf = (double *)calloc((size_t)n_threads * n_total * 3, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(n_threads)
#endif
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *mine = buf + (size_t)t * n_total * 3;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int c = 0; c < n_configs; ++c) {
            const int *nl = neighbor_list + e0[c];
            for (int j = 0; j < n_atoms[c]; ++j) {
                const long k = c0[c] + j;
                const int nn = number_of_neighs[k];
                const int row = config_ptr[c] + j;
                orc_compute_atom(sf, row, species, nl, nn, coords, desc + (size_t)k * sf->width);
                orc_gradient_atom(sf, row, species, nl, nn, coords,
                                  dE_dzeta + (size_t)k * sf->width, mine);
                nl += nn;
            }
        }
    }
    for (int t = 0; t < n_threads; ++t) {
        const double *src = buf + (size_t)t * n_total * 3;
        for (size_t a = 0; a < (size_t)n_total * 3; ++a) d_coords[a] += src[a];
    }
    free(buf);
    free(c0);
    free(e0);
}

/* ---- finite differences ---- */

/* numdiff::make_xph_representable (include/finite_difference.hpp:46-57) */
static double make_xph_representable(double x, double h)
{
    volatile double temp = x + h;
    h = temp - x;
    if (h == 0) h = nextafter(x, DBL_MAX) - x;
    return h;
}

/*
 * Descriptor::num_gradient_single_atom (src/Descriptors.cpp:488-518) with the
 * order-8 central stencil of numdiff::vec_finite_difference_derivative
 * (include/finite_difference.hpp:59-143):
 *   f' = (84 D1 - 21 D2 + 4 D3 - (3/8) D4) / (105 h),  Dk = f(x+kh) - f(x-kh),
 *   h = (551.25 eps)^(1/9), adjusted so that x+h is representable.
 * out[0:3] += sum_w g[w] * d desc_index[w] / d x_index.
 * n_total is the number of coordinate rows (coordinates.shape[0]).
 */
void orc_num_gradient_single_atom(const orc_sf *sf, int index, int n_total, const int *species,
                                  const int *neigh_list, int number_of_neigh,
                                  const double *coords, const double *g, double *out)
{
    int W = sf->width;
    double *x = (double *)malloc(sizeof(double) * 3 * n_total);
    double *y = (double *)malloc(sizeof(double) * W * 8);
    double *dy = (double *)malloc(sizeof(double) * W);
    const double eps = DBL_EPSILON;
    for (int dim = 0; dim < 3; ++dim) {
        int pos = index * 3 + dim;
        double h = pow(551.25 * eps, (double)1 / (double)9);
        h = make_xph_representable(coords[pos], h);
        memcpy(x, coords, sizeof(double) * 3 * n_total);
        static const int mult[8] = {1, -1, -2, 2, 3, -3, -4, 4};
        for (int e = 0; e < 8; ++e) {
            double *ye = y + e * W;
            for (int w = 0; w < W; ++w) ye[w] = 0.0;
            x[pos] = coords[pos] + mult[e] * h;
            orc_compute_atom(sf, index, species, neigh_list, number_of_neigh, x, ye);
        }
        for (int w = 0; w < W; ++w) {
            double y1 = y[0 * W + w] - y[1 * W + w];       /* f(x+h)  - f(x-h)  */
            double y2 = y[2 * W + w] - y[3 * W + w];       /* f(x-2h) - f(x+2h) */
            double y3 = y[4 * W + w] - y[5 * W + w];       /* f(x+3h) - f(x-3h) */
            double y4 = y[6 * W + w] - y[7 * W + w];       /* f(x-4h) - f(x+4h) */
            double t1 = 3 * y4 / 8 + 4 * y3;
            double t2 = 21 * y2 + 84 * y1;
            dy[w] = (t1 + t2) / (105 * h);
        }
        for (int w = 0; w < W; ++w) out[dim] += g[w] * dy[w];
    }
    free(x); free(y); free(dy);
}

/* ---- hyperparameter gradients ----
 *
 * The reference's gradient calls Enzyme with the descriptor object itself
 * duplicated (src/Descriptors.cpp:176-188, enzyme_dup desc_kind, d_desc_kind):
 * the shadow receives dE/d(parameters) and dE/d(cutoff matrix) and is then
 * deleted unread.  This restates what that shadow holds, term by term in the
 * forward's loop order (src/SymmetryFunctions.cpp:141-259):
 *   dparams[n_params]  flat in the parameter layout of initDescriptor (per
 *                      block, rows x cols row-major): g2 (eta, Rs), g3 (kappa),
 *                      g4/g5 (zeta, lambda, eta); g1 has none
 *   drcut[ns * ns]     rcut_2D_ entries, row-major (helper.hpp:2056-2067)
 * both +=, for the centres 0..n_atoms-1 weighted by g = dE/dzeta.
 */

/* d cut_cos(r; rcut) / d rcut */
static double dcut_drc(double r, double rcut)
{
    return (r < rcut) ? 0.5 * sin(ORC_PI * r / rcut) * ORC_PI * r / (rcut * rcut) : 0.0;
}

static void param_gradient_atom(const orc_sf *sf, int index, const int *species,
                                const int *neigh_list, int number_of_neigh, const double *coords,
                                const double *g, double *dparams, double *drcut)
{
    const int ns = sf->n_species;
    const int iS = species[index];
    const double *xi = coords + 3 * index;
    int *poff = (int *)malloc(sizeof(int) * (sf->n_blocks + 1));
    poff[0] = 0;
    for (int p = 0; p < sf->n_blocks; ++p) poff[p + 1] = poff[p] + sf->rows[p] * sf->cols[p];
    for (int jj = 0; jj < number_of_neigh; ++jj) {
        int j = neigh_list[jj];
        int jS = species[j];
        double rcutij = sf->rcut[iS * ns + jS];
        double rij[3];
        for (int d = 0; d < 3; ++d) rij[d] = coords[3 * j + d] - xi[d];
        double rijmag = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
        if (rijmag > rcutij) continue;
        double fij = cut_cos(rijmag, rcutij), dfij = dcut_drc(rijmag, rcutij);
        for (int p = 0; p < sf->n_blocks; ++p) {
            int kd = sf->kind[p];
            if (!is_two_body(kd)) continue;
            int idx = slot2(sf, p, jS);
            for (int q = 0; q < sf->rows[p]; ++q, ++idx) {
                const double *prm = sf->par[p] + q * sf->cols[p];
                double *dp = dparams + poff[p] + q * sf->cols[p];
                double w = g[idx];
                double val = 1.0; /* the term without its cutoff factor */
                if (kd == 2) {
                    double eta = prm[0], Rs = prm[1], dr = rijmag - Rs;
                    val = exp(-eta * dr * dr);
                    dp[0] += w * (-dr * dr) * val * fij;
                    dp[1] += w * (2.0 * eta * dr) * val * fij;
                } else if (kd == 3) {
                    double kappa = prm[0];
                    val = cos(kappa * rijmag);
                    dp[0] += w * (-rijmag * sin(kappa * rijmag)) * fij;
                }
                drcut[iS * ns + jS] += w * val * dfij;
            }
        }
        if (!sf->has_three_body) continue;
        for (int kk = jj + 1; kk < number_of_neigh; ++kk) {
            int k = neigh_list[kk];
            int kS = species[k];
            double rcutik = sf->rcut[iS * ns + kS];
            double rcutjk = sf->rcut[jS * ns + kS];
            double rik[3], rjk[3];
            for (int d = 0; d < 3; ++d) {
                rik[d] = coords[3 * k + d] - xi[d];
                rjk[d] = coords[3 * k + d] - coords[3 * j + d];
            }
            double rikmag = sqrt(rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2]);
            double rjkmag = sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
            if (rikmag > rcutik) continue;
            double rijsq = rijmag * rijmag, riksq = rikmag * rikmag, rjksq = rjkmag * rjkmag;
            double cos_ijk = (rijsq + riksq - rjksq) / (2 * rijmag * rikmag);
            double fik = cut_cos(rikmag, rcutik), dfik = dcut_drc(rikmag, rcutik);
            double fjk = cut_cos(rjkmag, rcutjk), dfjk = dcut_drc(rjkmag, rcutjk);
            for (int p = 0; p < sf->n_blocks; ++p) {
                int kd = sf->kind[p];
                if (!is_three_body(kd)) continue;
                if (kd == 4 && rjkmag > rcutjk) continue;
                int idx = slot3(sf, p, jS, kS);
                for (int q = 0; q < sf->rows[p]; ++q, ++idx) {
                    const double *prm = sf->par[p] + q * sf->cols[p];
                    double *dp = dparams + poff[p] + q * sf->cols[p];
                    double zeta = prm[0], lambda = prm[1], eta = prm[2];
                    double base = 1 + lambda * cos_ijk;
                    if (base <= 0) continue;
                    double w = g[idx];
                    double p2 = pow(2, 1 - zeta);
                    double costerm = pow(base, zeta);
                    double S = kd == 4 ? rijsq + riksq + rjksq : rijsq + riksq;
                    double e = exp(-eta * S);
                    double F = fij * fik * (kd == 4 ? fjk : 1.0);
                    double term = p2 * costerm * e * F;
                    dp[0] += w * term * (log(base) - log(2.0));
                    dp[1] += w * p2 * zeta * pow(base, zeta - 1) * cos_ijk * e * F;
                    dp[2] += w * (-S) * term;
                    double a = w * p2 * costerm * e;
                    drcut[iS * ns + jS] += a * dfij * fik * (kd == 4 ? fjk : 1.0);
                    drcut[iS * ns + kS] += a * fij * dfik * (kd == 4 ? fjk : 1.0);
                    if (kd == 4) drcut[jS * ns + kS] += a * fij * fik * dfjk;
                }
            }
        }
    }
    free(poff);
}

int orc_num_params(const orc_sf *sf)
{
    int n = 0;
    for (int p = 0; p < sf->n_blocks; ++p) n += sf->rows[p] * sf->cols[p];
    return n;
}

void orc_param_gradient(const orc_sf *sf, int n_atoms, const int *species,
                        const int *neighbor_list, const int *number_of_neighs,
                        const double *coords, const double *dE_dzeta, double *dparams,
                        double *drcut)
{
    const int *nl = neighbor_list;
    for (int i = 0; i < n_atoms; ++i) {
        param_gradient_atom(sf, i, species, nl, number_of_neighs[i], coords,
                            dE_dzeta + (size_t)i * sf->width, dparams, drcut);
        nl += number_of_neighs[i];
    }
}
