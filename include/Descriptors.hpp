// Drop-in C++ API of the symmetry-function path, with the reference's names,
// signatures and "+=" semantics (reference: include/Descriptors.hpp:19-271).
// Symmetry-function work runs on the GPU through the C-ABI in ldsf.h; a
// DescriptorKind subclass supplied by the user runs through its own virtual
// compute() on the host, as in the reference.  Errors from the C-ABI are
// rethrown as std::invalid_argument (bad arguments, unknown kind) or
// std::runtime_error (device, file and neighbour-list errors).
#ifndef LIBDESCRIPTOR_AMD_DESCRIPTORS_HPP_
#define LIBDESCRIPTOR_AMD_DESCRIPTORS_HPP_

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace Descriptor {

// include/Descriptors.hpp:21-26
enum AvailableDescriptor {
    KindSymmetryFunctions,
    KindBispectrum,
    KindSOAP,
    KindXi
};

class DescriptorKind;

// include/Descriptors.hpp:42-44
inline void fwd_diff(int, int *, int *, int *, double *, double *, DescriptorKind *)
{
    throw std::logic_error("Function not implemented yet.");
}

// include/Descriptors.hpp:58-60
void gradient(int n_atoms, int *species, int *neighbor_list, int *number_of_neighs,
              double *coordinates, double *d_coordinates, double *desc, double *dE_dzeta,
              DescriptorKind *descriptor_to_diff);

// include/Descriptors.hpp:78-80
void gradient_batch(int n_configurations, int *n_atoms, int *config_ptr, int *species,
                    int *neighbor_list, int *number_of_neighs, double *coordinates,
                    double *d_coordinates, double *desc, double *dE_dzeta,
                    DescriptorKind *descriptor_to_diff);

// include/Descriptors.hpp:97-99
void gradient_single_atom(int index, int n_atoms, int *species, int *neighbor_list,
                          int number_of_neighs, double *coordinates, double *d_coordinates,
                          double *desc, double *dE_dzeta, DescriptorKind *descriptor_to_diff);

// include/Descriptors.hpp:101-103
void num_gradient_single_atom(int index, int n_atoms, int *species, int *neighbor_list,
                              int number_of_neighs, double *coordinates, double *d_coordinates,
                              double *dE_dzeta, DescriptorKind *descriptor_to_diff);

// include/Descriptors.hpp:117-118
void jacobian(int n_atoms, int n_total_atoms, int *species, int *neighbor_list,
              int *number_of_neighs, double *coordinates, double *J_coordinates,
              DescriptorKind *descriptor_to_diff);

// include/Descriptors.hpp:130-132
void compute(int n_atoms, int *species, int *neighbor_list, int *number_of_neighs,
             double *coordinates, double *desc, DescriptorKind *descriptor_kind);

// include/Descriptors.hpp:147-149
void compute_batch(int n_configurations, int *n_atoms, int *config_ptr, int *species,
                   int *neighbor_list, int *number_of_neighs, double *coordinates, double *desc,
                   DescriptorKind *descriptor_kind);

// include/Descriptors.hpp:163-165
void compute_single_atom(int index, int n_atoms, int *species, int *neighbor_list,
                         int number_of_neighs, double *coordinates, double *desc,
                         DescriptorKind *descriptor_kind);

}  // namespace Descriptor

// include/Descriptors.hpp:179-271
class Descriptor::DescriptorKind {
public:
    AvailableDescriptor descriptor_kind = KindSymmetryFunctions;
    std::string descriptor_param_file;
    int width = -1;

    DescriptorKind() = default;

    // an empty descriptor of a kind (Descriptors.cpp:104-116)
    static DescriptorKind *initDescriptor(AvailableDescriptor availableDescriptorKind);

    // from a KIM-style parameter file (Descriptors.cpp:77-102)
    static DescriptorKind *initDescriptor(std::string &file_name,
                                          AvailableDescriptor availableDescriptorKind);

    // one centre's descriptor, added into desc[0:width]
    virtual void compute(int index, int n_atoms, int *species, int *neighbor_lists,
                         int number_of_neighbors, double *coordinates, double *desc) = 0;

    virtual ~DescriptorKind();

    // symmetry functions with one cutoff for every species pair (Descriptors.cpp:579-590)
    static DescriptorKind *initDescriptor(AvailableDescriptor availableDescriptorKind,
                                          std::vector<std::string> *species,
                                          std::string *cutoff_function, double *cutoff_matrix,
                                          std::vector<std::string> *symmetry_function_types,
                                          std::vector<int> *symmetry_function_sizes,
                                          std::vector<double> *symmetry_function_parameters);

    // Bispectrum, SOAP and Xi are not part of this build: these overloads keep
    // source compatibility and throw std::invalid_argument.
    static DescriptorKind *initDescriptor(AvailableDescriptor availableDescriptorKind,
                                          double rfac0_in, int twojmax_in, int diagonalstyle_in,
                                          int use_shared_arrays_in, double rmin0_in,
                                          int switch_flag_in, int bzero_flag_in,
                                          double *cutoff_array, std::vector<std::string> *species,
                                          std::vector<double> *weights);
    static DescriptorKind *initDescriptor(AvailableDescriptor availableDescriptorKind, int n_max,
                                          int l_max, double cutoff,
                                          std::vector<std::string> &species,
                                          std::string radial_basis, double eta);
    static DescriptorKind *initDescriptor(AvailableDescriptor availableDescriptorKind, int q,
                                          double cutoff, std::vector<std::string> &species,
                                          std::string &radial_basis);
};

#endif  // LIBDESCRIPTOR_AMD_DESCRIPTORS_HPP_
