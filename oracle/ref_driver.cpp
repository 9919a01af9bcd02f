// ORACLE -- test infrastructure only.
//
// A thin extern "C" driver over the reference's own, unmodified sources,
// compiled where they lie by oracle/build_ref.sh into oracle/_ref/libsfref.so:
//   src/SymmetryFunctions.cpp, src/Descriptors.cpp, src/helper.cpp,
//   python_bindings/neighbor/neighbor_list.cpp.
// The reference objects are built with hidden visibility and linked with
// --gc-sections, so the Enzyme-generated gradient (which needs the Enzyme
// plugin, absent here) and the other descriptor kinds drop out; what remains
// is the reference forward, its parameter-file reader and its neighbour list.
// Used only by tests/ and tests/golden/ generators to pin the oracle and the
// product against the reference's own compiled behaviour.
#include <string>
#include <vector>
#include "Descriptors.hpp"
#include "SymmetryFunctions.hpp"
#include "neighbor_list.h"

#define REF_API extern "C" __attribute__((visibility("default")))

// DescriptorKind::initDescriptor, SF vector overload (src/Descriptors.cpp:579-590),
// as the pybind module calls it (python_bindings/python_bindings.cpp:73-75).
REF_API void *ref_sf_create(int n_species, const char *const *species_names, double cutoff,
                            int n_types, const char *const *type_names, const int *sizes,
                            const double *params)
{
    std::vector<std::string> species;
    for (int i = 0; i < n_species; ++i) species.emplace_back(species_names[i]);
    std::string cutoff_fn = "cos";
    std::vector<std::string> types;
    std::vector<int> sz;
    int n_par = 0;
    for (int t = 0; t < n_types; ++t) {
        types.emplace_back(type_names[t]);
        sz.push_back(sizes[2 * t]);
        sz.push_back(sizes[2 * t + 1]);
        n_par += sizes[2 * t] * sizes[2 * t + 1];
    }
    std::vector<double> par(params, params + n_par);
    double rc = cutoff;
    auto *k = Descriptor::DescriptorKind::initDescriptor(Descriptor::KindSymmetryFunctions, &species,
                                                         &cutoff_fn, &rc, &types, &sz, &par);
    k->descriptor_kind = Descriptor::KindSymmetryFunctions;
    return k;
}

// SymmetryFunctions(std::string&) -> initFromFile (src/SymmetryFunctions.cpp:262-407)
REF_API void *ref_sf_create_from_file(const char *path)
{
    std::string p(path);
    auto *sf = new SymmetryFunctions(p);
    sf->descriptor_kind = Descriptor::KindSymmetryFunctions;
    return sf;
}

REF_API int ref_sf_width(void *k) { return static_cast<Descriptor::DescriptorKind *>(k)->width; }

REF_API void ref_sf_destroy(void *k) { delete static_cast<SymmetryFunctions *>(k); }

// Descriptor::compute (src/Descriptors.cpp:121-136): desc +=
REF_API void ref_compute(void *k, int n_atoms, int *species, int *neighbor_list, int *numneigh,
                         double *coords, double *desc)
{
    Descriptor::compute(n_atoms, species, neighbor_list, numneigh, coords, desc,
                        static_cast<Descriptor::DescriptorKind *>(k));
}

// Descriptor::compute_single_atom (src/Descriptors.cpp:355-365)
REF_API void ref_compute_single_atom(void *k, int index, int n_atoms, int *species,
                                     int *neighbor_list, int numneigh, double *coords,
                                     double *desc)
{
    Descriptor::compute_single_atom(index, n_atoms, species, neighbor_list, numneigh, coords, desc,
                                    static_cast<Descriptor::DescriptorKind *>(k));
}

// Descriptor::compute_batch (src/Descriptors.cpp:138-163)
REF_API void ref_compute_batch(void *k, int n_conf, int *n_atoms, int *ptr, int *species,
                               int *neighbor_list, int *numneigh, double *coords, double *desc)
{
    Descriptor::compute_batch(n_conf, n_atoms, ptr, species, neighbor_list, numneigh, coords, desc,
                              static_cast<Descriptor::DescriptorKind *>(k));
}

// Descriptor::num_gradient_single_atom (src/Descriptors.cpp:488-518)
REF_API void ref_num_gradient_single_atom(void *k, int index, int n_atoms, int *species,
                                          int *neighbor_list, int numneigh, double *coords,
                                          double *d_coords, double *dE_dzeta)
{
    Descriptor::num_gradient_single_atom(index, n_atoms, species, neighbor_list, numneigh, coords,
                                         d_coords, dE_dzeta,
                                         static_cast<Descriptor::DescriptorKind *>(k));
}

// nbl_create_paddings (python_bindings/neighbor/neighbor_list.cpp:283-418).
// Two-call protocol: first call with out buffers NULL returns the count.
REF_API int ref_create_paddings(int natoms, double cutoff, const double *cell, const int *pbc,
                                const double *coords, const int *species, int cap,
                                double *pad_coords, int *pad_species, int *pad_image)
{
    int n = 0;
    std::vector<double> pc;
    std::vector<int> ps, pi;
    int err = nbl_create_paddings(natoms, cutoff, cell, pbc, coords, species, n, pc, ps, pi);
    if (err) return -1;
    if (pad_coords && n <= cap) {
        for (int i = 0; i < 3 * n; ++i) pad_coords[i] = pc[i];
        for (int i = 0; i < n; ++i) { pad_species[i] = ps[i]; pad_image[i] = pi[i]; }
    }
    return n;
}

// nbl_build + nbl_get_neigh (neighbor_list.cpp:78-280), flattened the way
// NeighborList.get_numneigh_and_neighlist_1D does (neighbor.py:108-154).
// Returns the total neighbour count, or -1 on error; fills the outputs when
// neighlist != NULL and the total fits in cap.
REF_API long ref_neighlist(int natoms, const double *coords, double infl, const int *need_neigh,
                           int n_query, int *numneigh, int *neighlist, long cap)
{
    NeighList *nl = nullptr;
    nbl_initialize(&nl);
    double cut = infl;
    int err = nbl_build(nl, natoms, coords, infl, 1, &cut, need_neigh);
    if (err) { nbl_clean(&nl); return -1; }
    long total = 0;
    for (int i = 0; i < n_query; ++i) {
        int nn = 0;
        const int *nb = nullptr;
        if (nbl_get_neigh(nl, 1, &cut, 0, i, &nn, &nb)) { nbl_clean(&nl); return -1; }
        if (neighlist && total + nn <= cap) {
            numneigh[i] = nn;
            for (int m = 0; m < nn; ++m) neighlist[total + m] = nb[m];
        }
        total += nn;
    }
    nbl_clean(&nl);
    return total;
}
