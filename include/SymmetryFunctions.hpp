// Drop-in SymmetryFunctions class (reference: include/SymmetryFunctions.hpp:25-86),
// holding its parameters on the host and as an ldsf_sf handle whose device
// tables feed the HIP kernels.
#ifndef LIBDESCRIPTOR_AMD_SYMMETRY_FUNCTIONS_HPP_
#define LIBDESCRIPTOR_AMD_SYMMETRY_FUNCTIONS_HPP_

#include <cstddef>
#include <string>
#include <vector>

#include "Descriptors.hpp"
#include "ldsf.h"

class SymmetryFunctions final : public Descriptor::DescriptorKind {
public:
    // SymmetryFunctions.cpp:262-264 (reads the parameter file)
    explicit SymmetryFunctions(std::string &file_name);
    void initFromFile(std::string &file_name);

    SymmetryFunctions();
    // SymmetryFunctions.cpp:429-450
    SymmetryFunctions(std::vector<std::string> *species, std::string *cutoff_function,
                      double *cutoff_matrix, std::vector<std::string> *symmetry_function_types,
                      std::vector<int> *symmetry_function_sizes,
                      std::vector<double> *symmetry_function_parameters);
    ~SymmetryFunctions() override;
    SymmetryFunctions(const SymmetryFunctions &) = delete;
    SymmetryFunctions &operator=(const SymmetryFunctions &) = delete;

    // one centre, on the GPU (SymmetryFunctions.cpp:141-259)
    void compute(int index, int n_atoms, int *species, int *neigh_list, int number_of_neigh,
                 double *coords, double *zeta) override;

    // copy of another SF's tables with all parameters set to zero
    // (SymmetryFunctions.cpp:410-427)
    void clone_empty(DescriptorKind *descriptorKind);

    void set_species(std::vector<std::string> &species);
    void get_species(std::vector<std::string> &species);
    int get_num_species();
    void set_cutoff(char const *name, std::size_t Nspecies, double const *rcut_2D);
    double get_cutoff(int iCode, int jCode);
    void add_descriptor(char const *name, double const *values, int row, int col);
    int get_num_descriptors();

    // ---- extensions ----
    // LDSF_FP64 (default) or LDSF_FP32_BULK for every entry point
    void set_precision(int precision);
    int get_precision() const { return precision_; }
    // per-neighbour-species / per-species-pair blocks (off = reference layout)
    void set_species_resolved(bool on);
    bool get_species_resolved() const { return species_resolved_; }
    // the C-ABI handle, rebuilt after any parameter change
    ldsf_sf *handle();

private:
    void invalidate();
    std::vector<std::string> species_;
    std::vector<double> rcut_2D_;
    std::vector<std::string> name_;
    std::vector<int> rows_, cols_;
    std::vector<std::vector<double>> params_;
    int precision_ = LDSF_FP64;
    bool species_resolved_ = false;
    ldsf_sf *h_ = nullptr;
    std::string file_;  // non-empty: the handle comes from this parameter file
};

#endif  // LIBDESCRIPTOR_AMD_SYMMETRY_FUNCTIONS_HPP_
