// The host-only code of the path under AddressSanitizer and
// UndefinedBehaviorSanitizer: the parameter sets and file reader
// (csrc/sf_params.cpp), the neighbour list and padding producer
// (csrc/neighbor_list.cpp) and the oracle (oracle/sf_oracle.c).  The
// reference forces a Release build (CMakeLists.txt:12) and has no sanitizer
// runs; this drives each of those sources through the cases the Python tests
// give them (valid and invalid parameter sets, the golden parameter files, a
// truncated and a missing file, random cubic, triclinic, partly open and
// small periodic cells, a cluster, a collision and a singular cell, and every
// oracle entry point on a small cluster and batch), so that the sanitizers see
// every allocation and index those tests touch.  No GPU call: it links none.
//
// usage: sanitize_host <tests/golden directory>; exit 0 when every check holds
// (the sanitizers abort the run on the first error they find).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "ldsf.h"

extern "C" {
typedef struct orc_sf orc_sf;
orc_sf *orc_sf_create(int n_species, const double *rcut, int n_blocks, const int *kind,
                      const int *rows, const int *cols, const double *params,
                      int species_resolved);
void orc_sf_destroy(orc_sf *sf);
int orc_sf_width(const orc_sf *sf);
void orc_compute(const orc_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                 const int *numneigh, const double *coords, double *desc);
void orc_gradient(const orc_sf *sf, int n_atoms, const int *species, const int *neighbor_list,
                  const int *numneigh, const double *coords, const double *dE_dzeta,
                  double *d_coords);
void orc_compute_batch(const orc_sf *sf, int n_configurations, const int *n_atoms, const int *ptr,
                       const int *species, const int *neighbor_list, const int *numneigh,
                       const double *coords, double *desc);
void orc_gradient_batch(const orc_sf *sf, int n_configurations, const int *n_atoms, const int *ptr,
                        const int *species, const int *neighbor_list, const int *numneigh,
                        const double *coords, const double *dE_dzeta, double *d_coords);
void orc_fwd_adj_omp(const orc_sf *sf, int n_centers, int n_total, const int *species,
                     const int *neighbor_list, const int *offsets, const double *coords,
                     const double *dE_dzeta, double *desc, double *d_coords, int n_threads);
void orc_num_gradient_single_atom(const orc_sf *sf, int index, int n_total, const int *species,
                                  const int *neigh, int n_neigh, const double *coords,
                                  const double *dE_dzeta, double *d_coords);
int orc_num_params(const orc_sf *sf);
void orc_param_gradient(const orc_sf *sf, int n_atoms, const int *species,
                        const int *neighbor_list, const int *numneigh, const double *coords,
                        const double *dE_dzeta, double *dparams, double *drcut);
}

static int failures = 0;
#define CHECK(c)                                                                   \
    do {                                                                           \
        if (!(c)) {                                                                \
            std::fprintf(stderr, "check failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                            \
        }                                                                          \
    } while (0)

// g1..g5 in one set: sizes [rows, cols] per block, flat parameters
static const char *kTypes[] = {"g1", "g2", "g3", "g4", "g5"};
static const int kSizes[] = {1, 0, 3, 2, 2, 1, 3, 3, 1, 3};
static const double kParams[] = {0.7, 0.3, 0.05, 1.2, 1.9, 1.0, 1.0, 0.003, 2.0, -1.0,
                                 0.02, 5.0, 1.0, 0.004, 1.0, -1.0, 0.01, 3.5, 1.0, 0.002};
static const int kKinds[] = {1, 2, 3, 4, 5};
static const int kRows[] = {1, 3, 2, 3, 1};
static const int kCols[] = {0, 2, 1, 3, 3};

static ldsf_sf *make_set(int ns, double rc, int resolved, int *rc_out)
{
    std::vector<std::string> names;
    std::vector<const char *> pn;
    for (int k = 0; k < ns; ++k) names.push_back("S" + std::to_string(k));
    for (auto &s : names) pn.push_back(s.c_str());
    std::vector<double> cut(ns * ns, rc);
    ldsf_sf *h = nullptr;
    *rc_out = ldsf_sf_create(&h, ns, pn.data(), cut.data(), 5, kTypes, kSizes, kParams, resolved);
    return h;
}

static void parameter_sets(const std::string &golden)
{
    int rc = 0;
    for (int ns = 1; ns <= 3; ++ns)
        for (int res = 0; res <= 1; ++res) {
            ldsf_sf *h = make_set(ns, 4.0, res, &rc);
            CHECK(rc == LDSF_OK && h);
            if (!h) continue;
            const int npairs = ns * (ns + 1) / 2;
            CHECK(ldsf_sf_width(h) == (res ? 6 * ns + 4 * npairs : 10));
            CHECK(ldsf_sf_num_params(h) == 20);
            CHECK(ldsf_sf_num_species(h) == ns);
            CHECK(ldsf_sf_species_name(h, ns) == nullptr);
            CHECK(ldsf_sf_set_precision(h, LDSF_FP32_BULK) == LDSF_OK);
            CHECK(ldsf_sf_set_precision(h, 7) != LDSF_OK);
            ldsf_sf_destroy(h);
        }
    // invalid sets are refused
    const char *n1[] = {"Si"};
    double cut = 4.0;
    ldsf_sf *h = nullptr;
    const char *bad_type[] = {"g9"};
    const int s1[] = {1, 2};
    const double p1[] = {0.1, 0.0};
    CHECK(ldsf_sf_create(&h, 1, n1, &cut, 1, bad_type, s1, p1, 0) != LDSF_OK);
    const char *g2[] = {"g2"};
    const int s_neg[] = {-1, 2};
    CHECK(ldsf_sf_create(&h, 1, n1, &cut, 1, g2, s_neg, p1, 0) != LDSF_OK);
    const int s_cols[] = {2, 1};
    CHECK(ldsf_sf_create(&h, 1, n1, &cut, 1, g2, s_cols, p1, 0) != LDSF_OK);
    const char *g4[] = {"g4"};
    const int s4[] = {1, 2};
    CHECK(ldsf_sf_create(&h, 1, n1, &cut, 1, g4, s4, p1, 0) != LDSF_OK);
    CHECK(ldsf_last_error() != nullptr && ldsf_last_error()[0] != '\0');

    // the golden parameter files, a truncated copy and a missing file
    for (const char *f : {"sf_g2g4.params", "sf_two_species.params"}) {
        const std::string path = golden + "/" + f;
        ldsf_sf *hf = nullptr;
        CHECK(ldsf_sf_create_from_file(&hf, path.c_str()) == LDSF_OK && hf);
        if (!hf) continue;
        CHECK(ldsf_sf_width(hf) > 0);
        for (int k = 0; k < ldsf_sf_num_species(hf); ++k) CHECK(ldsf_sf_species_name(hf, k));
        ldsf_sf_destroy(hf);
        std::ifstream in(path);
        std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        for (size_t cut_at : {text.size() / 3, text.size() / 2, text.size() - 20}) {
            const std::string tpath = "/tmp/ldsf_sanitize_truncated.params";
            std::ofstream(tpath) << text.substr(0, cut_at);
            ldsf_sf *ht = nullptr;
            CHECK(ldsf_sf_create_from_file(&ht, tpath.c_str()) == LDSF_ERR_FILE);
            std::remove(tpath.c_str());
        }
    }
    CHECK(ldsf_sf_create_from_file(&h, "/nonexistent/ldsf.params") == LDSF_ERR_FILE);
}

// atoms uniform in a cell, rows of `cell` are the cell vectors
static std::vector<double> random_atoms(int n, const double *cell, std::mt19937_64 &rng)
{
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> x(3 * n);
    for (int i = 0; i < n; ++i) {
        const double f[3] = {u(rng), u(rng), u(rng)};
        for (int d = 0; d < 3; ++d)
            x[3 * i + d] = f[0] * cell[d] + f[1] * cell[3 + d] + f[2] * cell[6 + d];
    }
    return x;
}

// padding + neighbour list of one cell, checked for consistency; returns the
// flat layout for the oracle
struct Layout {
    std::vector<double> coords;
    std::vector<int> species, numneigh, neighlist;
    int n = 0;
};

static Layout layout(const std::vector<double> &pos, const std::vector<int> &sp, const double *cell,
                     const int *pbc, double rc)
{
    Layout L;
    const int n = (int)sp.size();
    L.n = n;
    ldsf_padding *pad = nullptr;
    CHECK(ldsf_nbl_create_paddings(n, rc, cell, pbc, pos.data(), sp.data(), &pad) == LDSF_OK);
    const int np = ldsf_padding_count(pad);
    std::vector<double> pc(3 * (size_t)np);
    std::vector<int> ps(np), pi(np);
    CHECK(ldsf_padding_copy(pad, np ? pc.data() : nullptr, np ? ps.data() : nullptr,
                            np ? pi.data() : nullptr) == LDSF_OK);
    for (int k = 0; k < np; ++k) CHECK(pi[k] >= 0 && pi[k] < n && ps[k] == sp[pi[k]]);
    ldsf_padding_free(pad);
    L.coords = pos;
    L.coords.insert(L.coords.end(), pc.begin(), pc.end());
    L.species = sp;
    L.species.insert(L.species.end(), ps.begin(), ps.end());
    const int rows = n + np;
    std::vector<int> need(rows, 0);
    for (int i = 0; i < n; ++i) need[i] = 1;
    ldsf_neighlist *nl = ldsf_nbl_create();
    const double cuts[1] = {rc};
    CHECK(ldsf_nbl_build(nl, rows, L.coords.data(), rc, 1, cuts, need.data()) == LDSF_OK);
    CHECK(ldsf_nbl_num_lists(nl) == 1 && ldsf_nbl_num_particles(nl) == rows);
    CHECK(ldsf_nbl_list_cutoff(nl, 0) == rc);
    const long total = ldsf_nbl_total(nl, 0, n);
    L.numneigh.assign(n, 0);
    L.neighlist.assign(total > 0 ? total : 1, 0);
    CHECK(ldsf_nbl_flatten(nl, 0, n, L.numneigh.data(), L.neighlist.data()) == LDSF_OK);
    L.neighlist.resize(total);
    long off = 0;
    for (int i = 0; i < n; ++i) {
        int m = 0;
        const int *nb = nullptr;
        CHECK(ldsf_nbl_get_neigh(nl, 1, cuts, 0, i, &m, &nb) == 0);
        CHECK(m == L.numneigh[i]);
        for (int k = 0; k < m; ++k) {
            CHECK(nb[k] == L.neighlist[off + k]);
            const int j = nb[k];
            CHECK(j >= 0 && j < rows && j != i);
            double r2 = 0;
            for (int d = 0; d < 3; ++d) {
                const double e = L.coords[3 * j + d] - L.coords[3 * i + d];
                r2 += e * e;
            }
            CHECK(r2 < rc * rc);
        }
        off += m;
    }
    int m = 0;
    const int *nb = nullptr;
    CHECK(ldsf_nbl_get_neigh(nl, 1, cuts, 0, rows + 5, &m, &nb) != 0);  // out of range
    ldsf_nbl_free(nl);
    return L;
}

static void neighbour_lists(std::vector<Layout> &keep)
{
    std::mt19937_64 rng(7);
    const double cubic[9] = {9, 0, 0, 0, 9, 0, 0, 0, 9};
    const double tri[9] = {8, 0, 0, 2.5, 7.5, 0, 1.0, -1.5, 9.0};
    const double small[9] = {3.1, 0, 0, 0, 3.1, 0, 0, 0, 3.1};
    struct Case {
        const double *cell;
        int pbc[3];
        int nat;
    } cases[] = {{cubic, {1, 1, 1}, 40}, {tri, {1, 1, 1}, 40}, {cubic, {1, 0, 1}, 40},
                 {cubic, {0, 0, 0}, 40}, {small, {1, 1, 1}, 6}};
    for (auto &c : cases) {
        std::vector<double> pos = random_atoms(c.nat, c.cell, rng);
        std::vector<int> sp(c.nat);
        for (int i = 0; i < c.nat; ++i) sp[i] = (int)(rng() % 2);
        keep.push_back(layout(pos, sp, c.cell, c.pbc, 4.1));
    }
    // a collision and a singular cell are errors
    std::vector<double> coll = {0, 0, 0, 1e-6, 0, 0, 3, 0, 0};
    ldsf_neighlist *nl = ldsf_nbl_create();
    const double cuts[1] = {4.0};
    const int need[3] = {1, 1, 1};
    CHECK(ldsf_nbl_build(nl, 3, coll.data(), 4.0, 1, cuts, need) != LDSF_OK);
    ldsf_nbl_free(nl);
    const double zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int pbc[3] = {1, 1, 1};
    const int sp3[3] = {0, 0, 0};
    ldsf_padding *pad = nullptr;
    CHECK(ldsf_nbl_create_paddings(3, 4.0, zero, pbc, coll.data(), sp3, &pad) ==
          LDSF_ERR_SINGULAR_CELL);
    if (pad) ldsf_padding_free(pad);
}

static void oracle(const std::vector<Layout> &cells)
{
    // every oracle entry point, two species, both layouts, on the periodic cells
    const double rcut[4] = {4.1, 4.0, 4.0, 3.9};
    for (int res = 0; res <= 1; ++res) {
        orc_sf *O = orc_sf_create(2, rcut, 5, kKinds, kRows, kCols, kParams, res);
        const int W = orc_sf_width(O);
        std::vector<int> n_atoms, ptr, species, numneigh, neighlist;
        std::vector<double> coords;
        int row = 0;
        for (const Layout &L : cells) {
            const int rows = (int)L.species.size();
            std::vector<double> desc(W * (size_t)L.n), dE(W * (size_t)L.n), g(3 * (size_t)rows);
            for (size_t k = 0; k < dE.size(); ++k) dE[k] = std::sin(0.37 * k);
            orc_compute(O, L.n, L.species.data(), L.neighlist.data(), L.numneigh.data(),
                        L.coords.data(), desc.data());
            orc_gradient(O, L.n, L.species.data(), L.neighlist.data(), L.numneigh.data(),
                         L.coords.data(), dE.data(), g.data());
            std::vector<int> offs(L.n + 1, 0);
            for (int i = 0; i < L.n; ++i) offs[i + 1] = offs[i] + L.numneigh[i];
            std::vector<double> d2(desc.size()), g2(g.size());
            orc_fwd_adj_omp(O, L.n, rows, L.species.data(), L.neighlist.data(), offs.data(),
                            L.coords.data(), dE.data(), d2.data(), g2.data(), 3);
            for (size_t k = 0; k < desc.size(); ++k) CHECK(d2[k] == desc[k]);
            std::vector<double> dp(orc_num_params(O)), drc(4);
            orc_param_gradient(O, L.n, L.species.data(), L.neighlist.data(), L.numneigh.data(),
                               L.coords.data(), dE.data(), dp.data(), drc.data());
            double fd[3] = {0, 0, 0};
            if (L.n > 0)
                orc_num_gradient_single_atom(O, 0, rows, L.species.data(), L.neighlist.data(),
                                             L.numneigh[0], L.coords.data(), dE.data(), fd);
            for (double v : g) CHECK(std::isfinite(v));
            // the batch form over all cells at once
            n_atoms.push_back(L.n);
            ptr.push_back(row);
            species.insert(species.end(), L.species.begin(), L.species.end());
            numneigh.insert(numneigh.end(), L.numneigh.begin(), L.numneigh.end());
            for (int j : L.neighlist) neighlist.push_back(j + row);
            coords.insert(coords.end(), L.coords.begin(), L.coords.end());
            row += rows;
        }
        int tot = 0;
        for (int a : n_atoms) tot += a;
        std::vector<double> db(W * (size_t)tot), dEb(W * (size_t)tot, 0.5), gb(coords.size());
        orc_compute_batch(O, (int)n_atoms.size(), n_atoms.data(), ptr.data(), species.data(),
                          neighlist.data(), numneigh.data(), coords.data(), db.data());
        orc_gradient_batch(O, (int)n_atoms.size(), n_atoms.data(), ptr.data(), species.data(),
                           neighlist.data(), numneigh.data(), coords.data(), dEb.data(), gb.data());
        orc_sf_destroy(O);
    }
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <tests/golden directory>\n", argv[0]);
        return 2;
    }
    parameter_sets(argv[1]);
    std::vector<Layout> cells;
    neighbour_lists(cells);
    oracle(cells);
    std::printf("sanitize_host: %d failed checks\n", failures);
    return failures ? 1 : 0;
}
