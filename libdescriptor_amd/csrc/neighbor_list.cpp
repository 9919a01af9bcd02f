// Host-side layout producer for the SF path: periodic padding atoms and the
// full neighbour list in KLIFF cell order, with the flat int32
// numneigh/neighlist layout the descriptor entry points consume.
//
// Interfaces mirrored (python_bindings/neighbor/neighbor_list.h):
//   nbl_create_paddings  neighbor_list.cpp:283-418  -> ldsf_nbl_create_paddings
//   nbl_build            neighbor_list.cpp:78-243   -> ldsf_nbl_build
//   nbl_get_neigh        neighbor_list.cpp:246-280  -> ldsf_nbl_get_neigh (KIM signature)
//   NeighborList.get_numneigh_and_neighlist_1D  neighbor.py:108-154 -> ldsf_nbl_flatten
//
// The order of neighbours is part of the drop-in contract: neighbours of an
// atom come cell by cell (x index outermost, then y, then z) over the 3x3x3
// block around the atom's cell, and in ascending atom index inside a cell.
// Cells are stored as a counting-sorted CSR instead of vectors of vectors,
// and the per-atom scans run in parallel over atoms (two passes: count, fill)
// so the result does not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldsf.h"
#include "ldsf_internal.h"

namespace {

constexpr double kTol = 1.0e-10;  // neighbor_list.cpp:14, helper.hpp:10

inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

inline double norm3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

inline double det2(double a11, double a12, double a21, double a22) { return (a11 * a22) - (a12 * a21); }

// 3x3 inverse by cofactors; the arithmetic follows helper.hpp:73-93 so that
// padding coordinates round exactly as the reference's do.
bool invert3(const double *m, double *inv)
{
    inv[0] = det2(m[4], m[5], m[7], m[8]);
    inv[1] = det2(m[2], m[1], m[8], m[7]);
    inv[2] = det2(m[1], m[2], m[4], m[5]);
    inv[3] = det2(m[5], m[3], m[8], m[6]);
    inv[4] = det2(m[0], m[2], m[6], m[8]);
    inv[5] = det2(m[2], m[0], m[5], m[3]);
    inv[6] = det2(m[3], m[4], m[6], m[7]);
    inv[7] = det2(m[1], m[0], m[7], m[6]);
    inv[8] = det2(m[0], m[1], m[3], m[4]);
    double dd = m[0] * m[4] * m[8] - m[0] * m[5] * m[7] - m[1] * m[3] * m[8] + m[1] * m[5] * m[6]
                + m[2] * m[3] * m[7] - m[2] * m[4] * m[6];
    if (std::abs(dd) < kTol) return false;
    for (int i = 0; i < 9; ++i) inv[i] /= dd;
    return true;
}

struct CellGrid {
    int size[3];
    double lo[3], hi[3];
    std::vector<int> start;  // size_total + 1
    std::vector<int> atoms;  // atoms sorted by cell, ascending index within a cell

    void index_of(const double *x, int *idx) const
    {
        for (int d = 0; d < 3; ++d) {
            int v = static_cast<int>(((x[d] - lo[d]) / (hi[d] - lo[d])) * size[d]);
            idx[d] = std::min(v, size[d] - 1);
        }
    }
};

}  // namespace

struct ldsf_neighlist {
    int n_particles = 0;
    int n_lists = 0;
    std::vector<double> cutoff;
    std::vector<std::vector<int>> numneigh;  // per list, per particle
    std::vector<std::vector<long>> begin;    // per list, per particle
    std::vector<std::vector<int>> neigh;     // per list, flat
};

struct ldsf_padding {
    std::vector<double> coords;
    std::vector<int> species;
    std::vector<int> image;
};

extern "C" {

LDSF_API int ldsf_nbl_create_paddings(int n_particles, double cutoff, const double *cell,
                                      const int *pbc, const double *coords, const int *species,
                                      ldsf_padding **out)
{
    *out = nullptr;
    if (n_particles < 0 || !cell || !pbc || (n_particles > 0 && (!coords || !species)))
        return ldsf::fail(LDSF_ERR_ARG, "create_paddings: bad arguments");
    double tcell[9], fcell[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) tcell[3 * r + c] = cell[3 * c + r];
    if (!invert3(tcell, fcell))
        return ldsf::fail(LDSF_ERR_SINGULAR_CELL,
                          "In inverting the cell matrix, the determinant is 0!");
    std::vector<double> frac(3 * (size_t)n_particles);
    double mn[3] = {1e10, 1e10, 1e10}, mx[3] = {-1e10, -1e10, -1e10};
    for (int i = 0; i < n_particles; ++i) {
        const double *x = coords + 3 * (size_t)i;
        for (int d = 0; d < 3; ++d) {
            double f = dot3(fcell + 3 * d, x);
            frac[3 * (size_t)i + d] = f;
            mn[d] = std::min(mn[d], f);
            mx[d] = std::max(mx[d], f);
        }
    }
    for (int d = 0; d < 3; ++d) { mn[d] -= kTol; mx[d] += kTol; }
    double xp[3];
    cross3(cell + 3, cell + 6, xp);
    double volume = std::abs(dot3(cell, xp));
    double dist[3];
    cross3(cell + 3, cell + 6, xp); dist[0] = volume / norm3(xp);
    cross3(cell + 6, cell + 0, xp); dist[1] = volume / norm3(xp);
    cross3(cell + 0, cell + 3, xp); dist[2] = volume / norm3(xp);
    double ratio[3] = {cutoff / dist[0], cutoff / dist[1], cutoff / dist[2]};
    int size[3];
    double slack[3];
    for (int d = 0; d < 3; ++d) {
        size[d] = static_cast<int>(std::ceil(ratio[d]));
        slack[d] = static_cast<double>(size[d]) - ratio[d];
    }
    auto *pad = new ldsf_padding;
    // images with x shift outermost, then y, then z; inside an image, atoms
    // in ascending index (neighbor_list.cpp:369-413)
    for (int i = -size[0]; i <= size[0]; ++i)
        for (int j = -size[1]; j <= size[1]; ++j)
            for (int k = -size[2]; k <= size[2]; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                if ((!pbc[0] && i) || (!pbc[1] && j) || (!pbc[2] && k)) continue;
                for (int at = 0; at < n_particles; ++at) {
                    const double *f = &frac[3 * (size_t)at];
                    if (i == -size[0] && f[0] - mn[0] < slack[0]) continue;
                    if (i == size[0] && mx[0] - f[0] < slack[0]) continue;
                    if (j == -size[1] && f[1] - mn[1] < slack[1]) continue;
                    if (j == size[1] && mx[1] - f[1] < slack[1]) continue;
                    if (k == -size[2] && f[2] - mn[2] < slack[2]) continue;
                    if (k == size[2] && mx[2] - f[2] < slack[2]) continue;
                    double shifted[3] = {i + f[0], j + f[1], k + f[2]};
                    for (int d = 0; d < 3; ++d) pad->coords.push_back(dot3(tcell + 3 * d, shifted));
                    pad->species.push_back(species[at]);
                    pad->image.push_back(at);
                }
            }
    *out = pad;
    return LDSF_OK;
}

LDSF_API int ldsf_padding_count(const ldsf_padding *p) { return p ? (int)p->image.size() : 0; }

LDSF_API int ldsf_padding_copy(const ldsf_padding *p, double *coords, int *species, int *image)
{
    if (!p) return ldsf::fail(LDSF_ERR_ARG, "padding_copy: null padding");
    size_t n = p->image.size();
    if (n) {
        if (coords) std::memcpy(coords, p->coords.data(), sizeof(double) * 3 * n);
        if (species) std::memcpy(species, p->species.data(), sizeof(int) * n);
        if (image) std::memcpy(image, p->image.data(), sizeof(int) * n);
    }
    return LDSF_OK;
}

LDSF_API void ldsf_padding_free(ldsf_padding *p) { delete p; }

LDSF_API ldsf_neighlist *ldsf_nbl_create(void) { return new ldsf_neighlist; }

LDSF_API void ldsf_nbl_free(ldsf_neighlist *nl) { delete nl; }

LDSF_API int ldsf_nbl_build(ldsf_neighlist *nl, int n_particles, const double *coords,
                            double influence_distance, int n_cutoffs, const double *cutoffs,
                            const int *need_neighbors)
{
    if (!nl || n_particles <= 0 || !coords || n_cutoffs <= 0 || !cutoffs || !need_neighbors)
        return ldsf::fail(LDSF_ERR_ARG, "nbl_build: bad arguments");
    CellGrid g;
    for (int d = 0; d < 3; ++d) { g.lo[d] = coords[d]; g.hi[d] = coords[d] + 1.0; }
    for (int i = 0; i < n_particles; ++i)
        for (int d = 0; d < 3; ++d) {
            double v = coords[3 * (size_t)i + d];
            if (g.hi[d] < v) g.hi[d] = v;
            if (g.lo[d] > v) g.lo[d] = v;
        }
    long total_cells = 1;
    for (int d = 0; d < 3; ++d) {
        g.size[d] = static_cast<int>((g.hi[d] - g.lo[d]) / influence_distance);
        if (g.size[d] <= 0) g.size[d] = 1;
        total_cells *= g.size[d];
    }
    if (total_cells > 1000000000L)
        return ldsf::fail(LDSF_ERR_NEIGHBOR,
                          "Cell size too large! (partilces fly away) or\nCollision of atoms happened!");
    // counting sort of atoms into cells; a stable scatter keeps ascending index
    std::vector<int> cell_of(n_particles);
    g.start.assign(total_cells + 1, 0);
    for (int i = 0; i < n_particles; ++i) {
        int id[3];
        g.index_of(coords + 3 * (size_t)i, id);
        int c = id[0] + id[1] * g.size[0] + id[2] * g.size[0] * g.size[1];
        cell_of[i] = c;
        g.start[c + 1]++;
    }
    for (long c = 0; c < total_cells; ++c) g.start[c + 1] += g.start[c];
    g.atoms.resize(n_particles);
    {
        std::vector<int> fill(g.start.begin(), g.start.end() - 1);
        for (int i = 0; i < n_particles; ++i) g.atoms[fill[cell_of[i]]++] = i;
    }
    std::vector<double> cutsq(n_cutoffs);
    double maxcut = 0.0;
    for (int k = 0; k < n_cutoffs; ++k) {
        cutsq[k] = cutoffs[k] * cutoffs[k];
        maxcut = std::max(maxcut, cutoffs[k]);
    }
    (void)maxcut;
    nl->n_particles = n_particles;
    nl->n_lists = n_cutoffs;
    nl->cutoff.assign(cutoffs, cutoffs + n_cutoffs);
    nl->numneigh.assign(n_cutoffs, std::vector<int>(n_particles, 0));
    nl->begin.assign(n_cutoffs, std::vector<long>(n_particles, 0));
    nl->neigh.assign(n_cutoffs, std::vector<int>());
    int collision = -1, collision_with = -1;

    // visit the 3x3x3 block around atom i in the reference's order and call
    // f(n, rsq) for every other atom n in it
    auto scan = [&](int i, auto &&f) {
        int id[3];
        const double *xi = coords + 3 * (size_t)i;
        g.index_of(xi, id);
        for (int ii = std::max(0, id[0] - 1); ii <= std::min(id[0] + 1, g.size[0] - 1); ++ii)
            for (int jj = std::max(0, id[1] - 1); jj <= std::min(id[1] + 1, g.size[1] - 1); ++jj)
                for (int kk = std::max(0, id[2] - 1); kk <= std::min(id[2] + 1, g.size[2] - 1); ++kk) {
                    long c = ii + (long)jj * g.size[0] + (long)kk * g.size[0] * g.size[1];
                    for (int m = g.start[c]; m < g.start[c + 1]; ++m) {
                        int n = g.atoms[m];
                        if (n == i) continue;
                        double dx = coords[3 * (size_t)n] - xi[0];
                        double dy = coords[3 * (size_t)n + 1] - xi[1];
                        double dz = coords[3 * (size_t)n + 2] - xi[2];
                        f(n, dx * dx + dy * dy + dz * dz);
                    }
                }
    };
    // pass 1: counts (and collision detection)
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n_particles; ++i) {
        if (!need_neighbors[i]) continue;
        scan(i, [&](int n, double rsq) {
            if (rsq < kTol) {
#pragma omp critical(ldsf_nbl_collision)
                {
                    if (collision < 0 || i < collision) { collision = i; collision_with = n; }
                }
                return;
            }
            for (int k = 0; k < n_cutoffs; ++k)
                if (rsq < cutsq[k]) nl->numneigh[k][i]++;
        });
    }
    if (collision >= 0) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "Collision of atoms %d and %d.", collision + 1, collision_with + 1);
        nl->numneigh.clear(); nl->begin.clear(); nl->neigh.clear(); nl->n_lists = 0;
        return ldsf::fail(LDSF_ERR_NEIGHBOR, msg);
    }
    for (int k = 0; k < n_cutoffs; ++k) {
        long t = 0;
        for (int i = 0; i < n_particles; ++i) { nl->begin[k][i] = t; t += nl->numneigh[k][i]; }
        nl->neigh[k].resize(t);
    }
    // pass 2: fill in order
#pragma omp parallel
    {
        std::vector<long> pos(n_cutoffs);
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n_particles; ++i) {
            if (!need_neighbors[i]) continue;
            for (int k = 0; k < n_cutoffs; ++k) pos[k] = nl->begin[k][i];
            scan(i, [&](int n, double rsq) {
                for (int k = 0; k < n_cutoffs; ++k)
                    if (rsq < cutsq[k]) nl->neigh[k][pos[k]++] = n;
            });
        }
    }
    return LDSF_OK;
}

// KIM-API GetNeighbor signature, as nbl_get_neigh (neighbor_list.cpp:246-280)
LDSF_API int ldsf_nbl_get_neigh(void const *data_object, int n_cutoffs, double const *cutoffs,
                                int neighbor_list_index, int particle_number,
                                int *number_of_neighbors, int const **neighbors_of_particle)
{
    (void)n_cutoffs;
    auto *nl = static_cast<const ldsf_neighlist *>(data_object);
    if (!nl || neighbor_list_index < 0 || neighbor_list_index >= nl->n_lists) return 1;
    if (cutoffs[neighbor_list_index] > nl->cutoff[neighbor_list_index] + kTol) return 1;
    if (particle_number >= nl->n_particles || particle_number < 0) return 1;
    *number_of_neighbors = nl->numneigh[neighbor_list_index][particle_number];
    *neighbors_of_particle = nl->neigh[neighbor_list_index].data()
                             + nl->begin[neighbor_list_index][particle_number];
    return 0;
}

LDSF_API int ldsf_nbl_num_lists(const ldsf_neighlist *nl) { return nl ? nl->n_lists : 0; }

LDSF_API double ldsf_nbl_list_cutoff(const ldsf_neighlist *nl, int k)
{
    return (nl && k >= 0 && k < nl->n_lists) ? nl->cutoff[k] : 0.0;
}

LDSF_API int ldsf_nbl_num_particles(const ldsf_neighlist *nl) { return nl ? nl->n_particles : 0; }

// Total neighbour count over atoms [0, n_query) of list k (-1 on bad input).
LDSF_API long ldsf_nbl_total(const ldsf_neighlist *nl, int k, int n_query)
{
    if (!nl || k < 0 || k >= nl->n_lists || n_query < 0 || n_query > nl->n_particles) return -1;
    long t = 0;
    for (int i = 0; i < n_query; ++i) t += nl->numneigh[k][i];
    return t;
}

// The flat layout of get_numneigh_and_neighlist_1D (neighbor.py:108-154):
// numneigh[n_query] and the concatenated neighbour indices of atoms 0..n_query-1.
LDSF_API int ldsf_nbl_flatten(const ldsf_neighlist *nl, int k, int n_query, int *numneigh,
                              int *neighlist)
{
    if (!nl || k < 0 || k >= nl->n_lists || n_query < 0 || n_query > nl->n_particles)
        return ldsf::fail(LDSF_ERR_ARG, "nbl_flatten: bad arguments");
    long t = 0;
    for (int i = 0; i < n_query; ++i) {
        int c = nl->numneigh[k][i];
        numneigh[i] = c;
        if (c) std::memcpy(neighlist + t, nl->neigh[k].data() + nl->begin[k][i], sizeof(int) * c);
        t += c;
    }
    return LDSF_OK;
}

}  // extern "C"
