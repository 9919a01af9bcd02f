// Drop-in C++ API (include/Descriptors.hpp, include/SymmetryFunctions.hpp)
// over the C-ABI.  The free functions keep the argument lists of
// src/Descriptors.cpp and dispatch on the descriptor's dynamic type:
// SymmetryFunctions goes to the GPU plans; any other DescriptorKind subclass
// is driven through its own virtual compute() on the host, exactly as
// Descriptor::compute does (src/Descriptors.cpp:121-163).
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/Descriptors.hpp"
#include "../../include/SymmetryFunctions.hpp"
#include "../../include/ldsf.h"
#include "ldsf_internal.h"

namespace {

void check(int rc)
{
    if (rc == LDSF_OK) return;
    std::string msg = ldsf_last_error();
    if (rc == LDSF_ERR_ARG || rc == LDSF_ERR_KIND || rc == LDSF_ERR_UNSUPPORTED)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

SymmetryFunctions *as_sf(Descriptor::DescriptorKind *k) { return dynamic_cast<SymmetryFunctions *>(k); }

[[noreturn]] void unsupported_gradient()
{
    throw std::invalid_argument("Descriptor kind not supported");
}

}  // namespace

// ---------------- SymmetryFunctions ----------------

SymmetryFunctions::SymmetryFunctions() { descriptor_kind = Descriptor::KindSymmetryFunctions; }

SymmetryFunctions::SymmetryFunctions(std::string &file_name) : SymmetryFunctions()
{
    initFromFile(file_name);
}

void SymmetryFunctions::initFromFile(std::string &file_name)
{
    ldsf_sf *h = nullptr;
    check(ldsf_sf_create_from_file(&h, file_name.c_str()));
    invalidate();
    h_ = h;
    file_ = file_name;
    // mirror the parsed parameter set on the host side
    species_ = h->species;
    rcut_2D_ = h->rcut;
    name_.clear(); rows_.clear(); cols_.clear(); params_.clear();
    static const char *names[] = {"", "g1", "g2", "g3", "g4", "g5"};
    for (size_t p = 0; p < h->kind.size(); ++p) {
        name_.emplace_back(names[h->kind[p]]);
        rows_.push_back(h->rows[p]);
        cols_.push_back(h->cols[p]);
        params_.push_back(h->par[p]);
    }
    width = h->width;
}

SymmetryFunctions::SymmetryFunctions(std::vector<std::string> *species, std::string *cutoff_function,
                                     double *cutoff_matrix,
                                     std::vector<std::string> *symmetry_function_types,
                                     std::vector<int> *symmetry_function_sizes,
                                     std::vector<double> *symmetry_function_parameters)
    : SymmetryFunctions()
{
    set_species(*species);
    set_cutoff(cutoff_function->c_str(), species->size(), cutoff_matrix);
    size_t off = 0;
    for (size_t i = 0; i < symmetry_function_types->size(); ++i) {
        int r = symmetry_function_sizes->at(2 * i), c = symmetry_function_sizes->at(2 * i + 1);
        add_descriptor(symmetry_function_types->at(i).c_str(),
                       symmetry_function_parameters->data() + off, r, c);
        off += (size_t)r * c;
    }
    handle();  // validate now, as the reference's constructor would fail early
}

SymmetryFunctions::~SymmetryFunctions() { invalidate(); }

void SymmetryFunctions::invalidate()
{
    if (h_) ldsf_sf_destroy(h_);
    h_ = nullptr;
    file_.clear();
}

ldsf_sf *SymmetryFunctions::handle()
{
    if (h_) return h_;
    std::vector<const char *> sp, ty;
    for (auto &s : species_) sp.push_back(s.c_str());
    for (auto &n : name_) ty.push_back(n.c_str());
    std::vector<int> sizes;
    std::vector<double> par;
    for (size_t p = 0; p < name_.size(); ++p) {
        sizes.push_back(rows_[p]);
        sizes.push_back(cols_[p]);
        par.insert(par.end(), params_[p].begin(), params_[p].end());
    }
    int ns = (int)species_.size();
    if (ns == 0 || (int)rcut_2D_.size() != ns * ns)
        throw std::invalid_argument("SymmetryFunctions: species and cutoffs must be set first");
    ldsf_sf *h = nullptr;
    check(ldsf_sf_create(&h, ns, sp.data(), rcut_2D_.data(), (int)name_.size(), ty.data(),
                         sizes.data(), par.empty() ? nullptr : par.data(), species_resolved_ ? 1 : 0));
    ldsf_sf_set_precision(h, precision_);
    h_ = h;
    width = ldsf_sf_width(h);
    return h_;
}

void SymmetryFunctions::compute(int index, int n_atoms, int *species, int *neigh_list,
                                int number_of_neigh, double *coords, double *zeta)
{
    check(ldsf_compute_single_atom(handle(), index, n_atoms, species, neigh_list, number_of_neigh,
                                   coords, zeta));
}

void SymmetryFunctions::clone_empty(DescriptorKind *descriptorKind)
{
    auto *o = as_sf(descriptorKind);
    if (!o) throw std::invalid_argument("clone_empty: not a SymmetryFunctions descriptor");
    invalidate();
    species_ = o->species_;
    rcut_2D_ = o->rcut_2D_;
    name_ = o->name_;
    rows_ = o->rows_;
    cols_ = o->cols_;
    params_ = o->params_;
    for (auto &p : params_) std::fill(p.begin(), p.end(), 0.0);
    precision_ = o->precision_;
    species_resolved_ = o->species_resolved_;
    width = o->width;
}

void SymmetryFunctions::set_species(std::vector<std::string> &species)
{
    invalidate();
    species_ = species;
}

void SymmetryFunctions::get_species(std::vector<std::string> &species) { species = species_; }

int SymmetryFunctions::get_num_species() { return (int)species_.size(); }

void SymmetryFunctions::set_cutoff(char const *name, std::size_t Nspecies, double const *rcut_2D)
{
    (void)name;  // only the cosine cutoff exists (SymmetryFunctions.cpp:35)
    invalidate();
    rcut_2D_.assign(rcut_2D, rcut_2D + Nspecies * Nspecies);
}

double SymmetryFunctions::get_cutoff(int iCode, int jCode)
{
    int ns = (int)species_.size();
    if (iCode < 0 || jCode < 0 || iCode >= ns || jCode >= ns)
        throw std::invalid_argument("get_cutoff: species code out of range");
    return rcut_2D_[iCode * ns + jCode];
}

void SymmetryFunctions::add_descriptor(char const *name, double const *values, int row, int col)
{
    std::string n(name);
    if (n != "g1" && n != "g2" && n != "g3" && n != "g4" && n != "g5")
        throw std::invalid_argument("unknown symmetry function type: " + n);
    invalidate();
    name_.push_back(n);
    rows_.push_back(row);
    cols_.push_back(col);
    params_.emplace_back(values, values + (size_t)row * col);
    int w = 0;
    for (int r : rows_) w += r;
    width = w;
}

int SymmetryFunctions::get_num_descriptors()
{
    int w = 0;
    for (int r : rows_) w += r;
    return w;
}

void SymmetryFunctions::set_precision(int precision)
{
    if (precision != LDSF_FP64 && precision != LDSF_FP32_BULK)
        throw std::invalid_argument("set_precision: LDSF_FP64 or LDSF_FP32_BULK");
    precision_ = precision;
    if (h_) ldsf_sf_set_precision(h_, precision);
}

void SymmetryFunctions::set_species_resolved(bool on)
{
    if (on == species_resolved_) return;
    invalidate();
    species_resolved_ = on;
    if (!species_.empty() && !rcut_2D_.empty()) handle();
}

// ---------------- DescriptorKind ----------------

Descriptor::DescriptorKind::~DescriptorKind() = default;

Descriptor::DescriptorKind *Descriptor::DescriptorKind::initDescriptor(AvailableDescriptor k)
{
    if (k == KindSymmetryFunctions) return new SymmetryFunctions();
    throw std::invalid_argument("Descriptor kind not implemented yet");
}

Descriptor::DescriptorKind *Descriptor::DescriptorKind::initDescriptor(std::string &file_name,
                                                                       AvailableDescriptor k)
{
    if (k != KindSymmetryFunctions) throw std::invalid_argument("Descriptor kind not implemented yet");
    auto *sf = new SymmetryFunctions(file_name);
    sf->descriptor_kind = k;
    sf->descriptor_param_file = file_name;
    return sf;
}

Descriptor::DescriptorKind *Descriptor::DescriptorKind::initDescriptor(
    AvailableDescriptor k, std::vector<std::string> *species, std::string *cutoff_function,
    double *cutoff, std::vector<std::string> *types, std::vector<int> *sizes,
    std::vector<double> *params)
{
    // one cutoff for every species pair (Descriptors.cpp:585)
    std::vector<double> cutoff_array(species->size() * species->size(), cutoff[0]);
    auto *sf = new SymmetryFunctions(species, cutoff_function, cutoff_array.data(), types, sizes, params);
    sf->descriptor_kind = k;
    return sf;
}

Descriptor::DescriptorKind *Descriptor::DescriptorKind::initDescriptor(
    AvailableDescriptor, double, int, int, int, double, int, int, double *,
    std::vector<std::string> *, std::vector<double> *)
{
    throw std::invalid_argument("Bispectrum is not part of this build");
}

Descriptor::DescriptorKind *Descriptor::DescriptorKind::initDescriptor(
    AvailableDescriptor, int, int, double, std::vector<std::string> &, std::string, double)
{
    throw std::invalid_argument("SOAP is not part of this build");
}

Descriptor::DescriptorKind *Descriptor::DescriptorKind::initDescriptor(
    AvailableDescriptor, int, double, std::vector<std::string> &, std::string &)
{
    throw std::invalid_argument("Xi is not part of this build");
}

// ---------------- free functions ----------------

void Descriptor::compute(int n_atoms, int *species, int *neighbor_list, int *number_of_neighs,
                         double *coordinates, double *desc, DescriptorKind *k)
{
    if (auto *sf = as_sf(k)) {
        check(ldsf_compute(sf->handle(), n_atoms, species, neighbor_list, number_of_neighs,
                           coordinates, desc));
        return;
    }
    int *nptr = neighbor_list;
    double *dptr = desc;
    for (int i = 0; i < n_atoms; ++i) {
        k->compute(i, n_atoms, species, nptr, number_of_neighs[i], coordinates, dptr);
        nptr += number_of_neighs[i];
        dptr += k->width;
    }
}

void Descriptor::compute_batch(int n_configurations, int *n_atoms, int *config_ptr, int *species,
                               int *neighbor_list, int *number_of_neighs, double *coordinates,
                               double *desc, DescriptorKind *k)
{
    if (auto *sf = as_sf(k)) {
        check(ldsf_compute_batch(sf->handle(), n_configurations, n_atoms, config_ptr, species,
                                 neighbor_list, number_of_neighs, coordinates, desc));
        return;
    }
    int total = 0;
    for (int c = 0; c < n_configurations; ++c) total += n_atoms[c];
    int *nptr = neighbor_list;
    double *dptr = desc;
    int idx = 0;
    for (int c = 0; c < n_configurations; ++c)
        for (int j = 0; j < n_atoms[c]; ++j) {
            k->compute(j + config_ptr[c], total, species, nptr, number_of_neighs[idx], coordinates, dptr);
            nptr += number_of_neighs[idx];
            dptr += k->width;
            ++idx;
        }
}

void Descriptor::compute_single_atom(int index, int n_atoms, int *species, int *neighbor_list,
                                     int number_of_neighs, double *coordinates, double *desc,
                                     DescriptorKind *k)
{
    k->compute(index, n_atoms, species, neighbor_list, number_of_neighs, coordinates, desc);
}

void Descriptor::gradient(int n_atoms, int *species, int *neighbor_list, int *number_of_neighs,
                          double *coordinates, double *d_coordinates, double *desc, double *dE_dzeta,
                          DescriptorKind *k)
{
    auto *sf = as_sf(k);
    if (!sf) unsupported_gradient();
    check(ldsf_gradient(sf->handle(), n_atoms, species, neighbor_list, number_of_neighs,
                        coordinates, d_coordinates, desc, dE_dzeta));
}

void Descriptor::gradient_batch(int n_configurations, int *n_atoms, int *config_ptr, int *species,
                                int *neighbor_list, int *number_of_neighs, double *coordinates,
                                double *d_coordinates, double *desc, double *dE_dzeta,
                                DescriptorKind *k)
{
    auto *sf = as_sf(k);
    if (!sf) unsupported_gradient();
    check(ldsf_gradient_batch(sf->handle(), n_configurations, n_atoms, config_ptr, species,
                              neighbor_list, number_of_neighs, coordinates, d_coordinates, desc,
                              dE_dzeta));
}

void Descriptor::gradient_single_atom(int index, int n_atoms, int *species, int *neighbor_list,
                                      int number_of_neighs, double *coordinates,
                                      double *d_coordinates, double *desc, double *dE_dzeta,
                                      DescriptorKind *k)
{
    auto *sf = as_sf(k);
    if (!sf) unsupported_gradient();
    check(ldsf_gradient_single_atom(sf->handle(), index, n_atoms, species, neighbor_list,
                                    number_of_neighs, coordinates, d_coordinates, desc, dE_dzeta));
}

void Descriptor::num_gradient_single_atom(int index, int n_atoms, int *species, int *neighbor_list,
                                          int number_of_neighs, double *coordinates,
                                          double *d_coordinates, double *dE_dzeta,
                                          DescriptorKind *k)
{
    auto *sf = as_sf(k);
    if (!sf) unsupported_gradient();
    check(ldsf_num_gradient_single_atom(sf->handle(), index, n_atoms, species, neighbor_list,
                                        number_of_neighs, coordinates, d_coordinates, dE_dzeta));
}

void Descriptor::jacobian(int n_atoms, int n_total_atoms, int *species, int *neighbor_list,
                          int *number_of_neighs, double *coordinates, double *J_coordinates,
                          DescriptorKind *k)
{
    auto *sf = as_sf(k);
    if (!sf) unsupported_gradient();
    check(ldsf_jacobian(sf->handle(), n_atoms, n_total_atoms, species, neighbor_list,
                        number_of_neighs, coordinates, J_coordinates));
}
