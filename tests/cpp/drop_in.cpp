// A reference-style C++ caller of the symmetry-function path: it includes the
// drop-in headers, builds the descriptor with DescriptorKind::initDescriptor
// (src/Descriptors.cpp:579-590) and calls Descriptor::compute and
// Descriptor::gradient with caller-owned, zeroed buffers, exactly as a KIM or
// TorchML model driver would against the reference library.
//
// It also builds an empty SymmetryFunctions and calls clone_empty on it
// (src/SymmetryFunctions.cpp:410-427: the other's tables, every parameter
// zero -- the Enzyme shadow's starting point), checks that the clone keeps
// the width, species, cutoffs and descriptor count, and computes with it.
//
// usage: drop_in <input.bin> <output.bin>
//   input : int32 n_atoms, n_total, n_neigh; int32 species[n_total],
//           numneigh[n_atoms], neighlist[n_neigh]; float64 coords[n_total*3],
//           dE_dzeta[n_atoms*24]
//   output: float64 desc[n_atoms*24], d_coords[n_total*3],
//           desc_single[24] (compute_single_atom of atom 0),
//           desc_clone[n_atoms*24] (Descriptor::compute with the clone)
//        drop_in --clone-check   (host only: the clone's tables, no GPU call)
// Exit status 0 on success; 3 if the unknown-kind error was not an
// std::invalid_argument; 5 if the clone's tables differ.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "Descriptors.hpp"
#include "SymmetryFunctions.hpp"

template <typename T>
static bool read_n(FILE *f, std::vector<T> &v, size_t n)
{
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

// the Behler set: 8 g2 rows (eta log-spaced 0.001..0.5, Rs = 0), 16 g4 rows
static Descriptor::DescriptorKind *behler_set()
{
    std::vector<std::string> sp{"Si"}, types{"g2", "g4"};
    std::string cutfn = "cos";
    double rc = 4.5;
    std::vector<int> sizes{8, 2, 16, 3};
    std::vector<double> params;
    for (int q = 0; q < 8; ++q) {
        params.push_back(0.001 * std::pow(500.0, q / 7.0));
        params.push_back(0.0);
    }
    for (double z : {1.0, 2.0, 4.0, 16.0})
        for (double lam : {1.0, -1.0})
            for (double eta : {0.001, 0.01}) {
                params.push_back(z);
                params.push_back(lam);
                params.push_back(eta);
            }
    return Descriptor::DescriptorKind::initDescriptor(Descriptor::KindSymmetryFunctions, &sp,
                                                      &cutfn, &rc, &types, &sizes, &params);
}

// clone_empty keeps everything but the parameter values
static bool clone_matches(Descriptor::DescriptorKind *dk, SymmetryFunctions &clone)
{
    auto *sf = dynamic_cast<SymmetryFunctions *>(dk);
    if (!sf) return false;
    clone.clone_empty(dk);
    std::vector<std::string> a, b;
    sf->get_species(a);
    clone.get_species(b);
    return clone.width == dk->width && a == b && clone.get_num_species() == sf->get_num_species() &&
           clone.get_num_descriptors() == sf->get_num_descriptors() &&
           clone.get_cutoff(0, 0) == sf->get_cutoff(0, 0);
}

int main(int argc, char **argv)
{
    if (argc == 2 && std::string(argv[1]) == "--clone-check") {
        Descriptor::DescriptorKind *dk = behler_set();
        SymmetryFunctions clone;
        const bool ok = clone_matches(dk, clone);
        delete dk;
        return ok ? 0 : 5;
    }
    if (argc != 3) {
        fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[3];
    if (fread(hdr, sizeof(int), 3, f) != 3) return 2;
    const int n = hdr[0], n_total = hdr[1], n_nl = hdr[2];
    std::vector<int> species, numneigh, nl;
    std::vector<double> coords, dE;
    if (!read_n(f, species, n_total) || !read_n(f, numneigh, n) || !read_n(f, nl, n_nl) ||
        !read_n(f, coords, 3 * (size_t)n_total) || !read_n(f, dE, 24 * (size_t)n))
        return 2;
    fclose(f);

    Descriptor::DescriptorKind *dk = behler_set();
    if (dk->width != 24) return 4;

    std::vector<double> desc(24 * (size_t)n, 0.0), dcoords(3 * (size_t)n_total, 0.0);
    std::vector<double> scratch(24 * (size_t)n, 0.0), single(24, 0.0);
    Descriptor::compute(n, species.data(), nl.data(), numneigh.data(), coords.data(),
                        desc.data(), dk);
    Descriptor::gradient(n, species.data(), nl.data(), numneigh.data(), coords.data(),
                         dcoords.data(), scratch.data(), dE.data(), dk);
    Descriptor::compute_single_atom(0, n_total, species.data(), nl.data(), numneigh[0],
                                    coords.data(), single.data(), dk);
    SymmetryFunctions clone;
    int status = clone_matches(dk, clone) ? 0 : 5;
    delete dk;
    std::vector<double> desc_clone(24 * (size_t)n, 0.0);
    Descriptor::compute(n, species.data(), nl.data(), numneigh.data(), coords.data(),
                        desc_clone.data(), &clone);

    std::vector<std::string> sp{"Si"};
    std::string cutfn = "cos";
    try {
        Descriptor::DescriptorKind::initDescriptor(Descriptor::KindSOAP, 4, 4, 5.0, sp,
                                                   cutfn, 1.0);
        status = 3;
    } catch (const std::invalid_argument &) {
    } catch (...) {
        status = 3;
    }

    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    fwrite(desc.data(), sizeof(double), desc.size(), o);
    fwrite(dcoords.data(), sizeof(double), dcoords.size(), o);
    fwrite(single.data(), sizeof(double), single.size(), o);
    fwrite(desc_clone.data(), sizeof(double), desc_clone.size(), o);
    fclose(o);
    return status;
}
