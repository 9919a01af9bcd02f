// pybind11 module `libdescriptor`: the Python ABI of the symmetry-function
// path with the names, positional orders and dtypes of the reference module
// (python_bindings/python_bindings.cpp:50-556).  Outputs are fresh, zeroed
// numpy arrays (:99,124,147,176,205,241,276) and nothing leaks.  The GIL is
// released around GPU work.  init_descriptor is bound as a static method, so
// both DescriptorKind.init_descriptor(kind, ...) (the documented call,
// python_package/docs/tutorial.rst:90-98) and instance calls work.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/Descriptors.hpp"
#include "../../include/SymmetryFunctions.hpp"
#include "../../include/ldsf.h"
#include "ldsf_internal.h"

namespace py = pybind11;
using namespace Descriptor;

using IArr = py::array_t<int, py::array::c_style | py::array::forcecast>;
using DArr = py::array_t<double, py::array::c_style | py::array::forcecast>;

namespace {

// Python subclasses of DescriptorKind override compute (python_bindings.cpp:32-47)
class PyDescriptorKind : public DescriptorKind {
public:
    using DescriptorKind::DescriptorKind;
    void compute(int index, int n_atoms, int *species, int *neighbor_lists, int number_of_neighbors,
                 double *coordinates, double *desc) override
    {
        PYBIND11_OVERRIDE_PURE(void, DescriptorKind, compute, index, n_atoms, species,
                               neighbor_lists, number_of_neighbors, coordinates, desc);
    }
};

SymmetryFunctions *sf_of(DescriptorKind &k) { return dynamic_cast<SymmetryFunctions *>(&k); }

int *ip(const IArr &a) { return const_cast<int *>(a.size() ? a.data() : nullptr); }
double *dp(const DArr &a) { return const_cast<double *>(a.size() ? a.data() : nullptr); }

py::array_t<double> zeros(std::vector<py::ssize_t> shape)
{
    py::array_t<double> a(shape);
    if (a.size()) std::memset(a.mutable_data(), 0, sizeof(double) * a.size());
    return a;
}

// the coordinate rows of the arrays check_layout last saw on this thread
thread_local long g_layout_rows = -1;

// runs f with the GIL released when the descriptor is a SymmetryFunctions,
// with the row count check_layout found as the limit of every index the
// call's host-pointer staging reads
template <typename F>
void run(DescriptorKind &k, F &&f)
{
    const long rows = g_layout_rows;
    g_layout_rows = -1;
    if (sf_of(k)) {
        ldsf::HostRowsLimit limit(rows);
        py::gil_scoped_release nogil;
        f();
    } else {
        f();
    }
}

// The reference trusts its inputs (out-of-range indices are undefined
// behaviour there); numpy arrays carry their sizes, so check the layout here
// and raise ValueError before any work is staged.  Neighbour indices are
// checked against the row count while the call stages the neighbour list
// (run passes it on), not in a separate pass here.
void check_layout(const IArr &species, const IArr &neighbors, long n_neigh_used, const DArr &coords,
                  long n_centers, const int *centers)
{
    long rows = coords.ndim() == 2 ? (long)coords.shape(0) : (long)coords.size() / 3;
    g_layout_rows = -1;
    if (coords.size() != rows * 3) throw std::invalid_argument("coordinates must have shape (n, 3)");
    if ((long)species.size() < rows) throw std::invalid_argument("species must cover every coordinate row");
    if (n_neigh_used > (long)neighbors.size())
        throw std::invalid_argument("neighbor list is shorter than the sum of the neighbour counts");
    g_layout_rows = rows;
    for (long t = 0; t < n_centers; ++t) {
        long c = centers ? centers[t] : t;
        if (c < 0 || c >= rows) throw std::invalid_argument("centre index out of range");
    }
}

long sum_counts(const IArr &nn, long n)
{
    if ((long)nn.size() < n) throw std::invalid_argument("number_of_neighbors is shorter than n_atoms");
    long s = 0;
    for (long i = 0; i < n; ++i) {
        if (nn.data()[i] < 0) throw std::invalid_argument("negative neighbour count");
        s += nn.data()[i];
    }
    return s;
}

std::vector<int> batch_centres(const IArr &n_atoms, const IArr &config_ptr)
{
    if (config_ptr.size() < n_atoms.size()) throw std::invalid_argument("config_ptr is shorter than n_atoms");
    std::vector<int> c;
    for (long k = 0; k < (long)n_atoms.size(); ++k)
        for (int j = 0; j < n_atoms.data()[k]; ++j) c.push_back(j + config_ptr.data()[k]);
    return c;
}

// a C-ABI status as the C++ API's exceptions (invalid_argument for argument
// errors, runtime_error otherwise), which pybind11 maps to ValueError and
// RuntimeError
void raise_status(int rc)
{
    if (rc == LDSF_OK) return;
    std::string msg = ldsf_last_error();
    if (rc == LDSF_ERR_ARG || rc == LDSF_ERR_KIND || rc == LDSF_ERR_UNSUPPORTED)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void check_width(const DArr &a, long rows, int width, const char *what)
{
    if ((long)a.size() < rows * (long)width)
        throw std::invalid_argument(std::string(what) + " must hold n_atoms x width entries");
}

struct NeighListHandle {
    ldsf_neighlist *nl = ldsf_nbl_create();
    ~NeighListHandle() { ldsf_nbl_free(nl); }
};

}  // namespace

PYBIND11_MODULE(libdescriptor, m)
{
    m.doc() = "Python interface to libdescriptor (MI355X symmetry-function path)";

    py::enum_<AvailableDescriptor>(m, "AvailableDescriptors")
        .value("Bispectrum", KindBispectrum)
        .value("SymmetryFunctions", KindSymmetryFunctions)
        .value("SOAP", KindSOAP)
        .value("Xi", KindXi);

    m.attr("FP64") = LDSF_FP64;
    m.attr("FP32_BULK") = LDSF_FP32_BULK;

    py::class_<DescriptorKind, PyDescriptorKind>(m, "DescriptorKind")
        .def(py::init<>())
        .def_static("init_descriptor", py::overload_cast<AvailableDescriptor>(&DescriptorKind::initDescriptor))
        .def_static("init_descriptor",
                    py::overload_cast<std::string &, AvailableDescriptor>(&DescriptorKind::initDescriptor))
        .def_static("init_descriptor",
                    py::overload_cast<AvailableDescriptor, std::vector<std::string> *, std::string *,
                                      double *, std::vector<std::string> *, std::vector<int> *,
                                      std::vector<double> *>(&DescriptorKind::initDescriptor))
        .def_static("init_descriptor",
                    py::overload_cast<AvailableDescriptor, double, int, int, int, double, int, int,
                                      double *, std::vector<std::string> *, std::vector<double> *>(
                        &DescriptorKind::initDescriptor))
        .def_static("init_descriptor",
                    py::overload_cast<AvailableDescriptor, int, double, std::vector<std::string> &,
                                      std::string &>(&DescriptorKind::initDescriptor))
        .def_static("init_descriptor",
                    py::overload_cast<AvailableDescriptor, int, int, double, std::vector<std::string> &,
                                      std::string, double>(&DescriptorKind::initDescriptor))
        // extension: full cutoff matrix, species-resolved layout and precision form
        .def_static(
            "init_symmetry_functions",
            [](std::vector<std::string> species, py::object cutoff, std::vector<std::string> types,
               std::vector<int> sizes, std::vector<double> params, bool species_resolved,
               int precision) {
                size_t ns = species.size();
                std::vector<double> rc;
                DArr arr = DArr::ensure(cutoff);
                if (!arr) throw std::invalid_argument("cutoff must be a number or an n_species^2 array");
                if (arr.size() == 1) rc.assign(ns * ns, arr.data()[0]);
                else if ((size_t)arr.size() == ns * ns) rc.assign(arr.data(), arr.data() + ns * ns);
                else throw std::invalid_argument("cutoff must be a number or an n_species^2 array");
                std::string fn = "cos";
                auto *sf = new SymmetryFunctions(&species, &fn, rc.data(), &types, &sizes, &params);
                sf->set_species_resolved(species_resolved);
                sf->set_precision(precision);
                return static_cast<DescriptorKind *>(sf);
            },
            py::arg("species"), py::arg("cutoff"), py::arg("types"), py::arg("sizes"),
            py::arg("params"), py::arg("species_resolved") = false, py::arg("precision") = LDSF_FP64)
        .def(
            "compute",
            [](DescriptorKind &ds, int index, IArr &species, IArr &neighbors, DArr &coords) {
                if (sf_of(ds)) check_layout(species, neighbors, neighbors.size(), coords, 1, &index);
                int n_atoms = (int)coords.shape(0);
                int n_neigh = (int)neighbors.shape(0);
                auto desc = zeros({ds.width});
                double *out = desc.mutable_data();
                run(ds, [&] { ds.compute(index, n_atoms, ip(species), ip(neighbors), n_neigh, dp(coords), out); });
                return desc;
            })
        .def_readwrite("param_file", &DescriptorKind::descriptor_param_file)
        .def_readwrite("kind", &DescriptorKind::descriptor_kind)
        .def_readwrite("width", &DescriptorKind::width)
        .def_property(
            "precision",
            [](DescriptorKind &k) {
                auto *sf = sf_of(k);
                return sf ? sf->get_precision() : LDSF_FP64;
            },
            [](DescriptorKind &k, int p) {
                auto *sf = sf_of(k);
                if (!sf) throw std::invalid_argument("precision applies to SymmetryFunctions only");
                sf->set_precision(p);
            })
        .def_property(
            "species_resolved",
            [](DescriptorKind &k) {
                auto *sf = sf_of(k);
                return sf ? sf->get_species_resolved() : false;
            },
            [](DescriptorKind &k, bool on) {
                auto *sf = sf_of(k);
                if (!sf) throw std::invalid_argument("species_resolved applies to SymmetryFunctions only");
                sf->set_species_resolved(on);
            })
        .def_property_readonly("species", [](DescriptorKind &k) {
            std::vector<std::string> s;
            if (auto *sf = sf_of(k)) sf->get_species(s);
            return s;
        })
        .def_property_readonly("_handle", [](DescriptorKind &k) {
            auto *sf = sf_of(k);
            if (!sf) throw std::invalid_argument("not a SymmetryFunctions descriptor");
            return reinterpret_cast<std::uintptr_t>(sf->handle());
        });

    m.def(
        "compute_single_atom",
        [](DescriptorKind &ds, int index, IArr &species, IArr &neighbors, DArr &coordinates) {
            if (sf_of(ds)) check_layout(species, neighbors, neighbors.size(), coordinates, 1, &index);
            int n_atoms = (int)coordinates.shape(0);
            int n_neigh = (int)neighbors.shape(0);
            auto desc = zeros({ds.width});
            double *out = desc.mutable_data();
            run(ds, [&] {
                compute_single_atom(index, n_atoms, ip(species), ip(neighbors), n_neigh,
                                    dp(coordinates), out, &ds);
            });
            return desc;
        },
        "Compute descriptor for single atom of configuration.");

    m.def(
        "compute",
        [](DescriptorKind &ds, int n_atoms, IArr &species, IArr &neighbors, IArr &number_of_neighbors,
           DArr &coordinates) {
            if (sf_of(ds))
                check_layout(species, neighbors, sum_counts(number_of_neighbors, n_atoms), coordinates,
                             n_atoms, nullptr);
            // the symmetry functions overwrite every entry: no zero fill
            auto desc = sf_of(ds) ? py::array_t<double>(std::vector<py::ssize_t>{n_atoms, ds.width})
                                  : zeros({n_atoms, ds.width});
            double *out = desc.mutable_data();
            if (auto *sf = sf_of(ds)) {
                // a new array: overwritten on the device, no upload of its zeros
                int rc = 0;
                run(ds, [&] {
                    rc = ldsf_compute_overwrite(sf->handle(), n_atoms, ip(species), ip(neighbors),
                                                ip(number_of_neighbors), dp(coordinates), out);
                });
                raise_status(rc);
                return desc;
            }
            run(ds, [&] {
                compute(n_atoms, ip(species), ip(neighbors), ip(number_of_neighbors), dp(coordinates),
                        out, &ds);
            });
            return desc;
        },
        "Compute descriptor for single atom of configuration.");

    m.def(
        "compute_batch",
        [](DescriptorKind &ds, IArr &n_atoms, IArr &config_ptr, IArr &species, IArr &neighbors,
           IArr &number_of_neighbors, DArr &coordinates) {
            int n_conf = (int)n_atoms.size();
            int total = 0;
            for (int i = 0; i < n_conf; ++i) total += n_atoms.at(i);
            if (sf_of(ds)) {
                auto ctr = batch_centres(n_atoms, config_ptr);
                check_layout(species, neighbors, sum_counts(number_of_neighbors, total), coordinates,
                             total, ctr.data());
            }
            auto desc = sf_of(ds) ? py::array_t<double>(std::vector<py::ssize_t>{total, ds.width})
                                  : zeros({total, ds.width});
            double *out = desc.mutable_data();
            if (auto *sf = sf_of(ds)) {
                int rc = 0;
                run(ds, [&] {
                    rc = ldsf_compute_batch_overwrite(sf->handle(), n_conf, ip(n_atoms),
                                                      ip(config_ptr), ip(species), ip(neighbors),
                                                      ip(number_of_neighbors), dp(coordinates), out);
                });
                raise_status(rc);
                return desc;
            }
            run(ds, [&] {
                compute_batch(n_conf, ip(n_atoms), ip(config_ptr), ip(species), ip(neighbors),
                              ip(number_of_neighbors), dp(coordinates), out, &ds);
            });
            return desc;
        },
        "Compute descriptor for multiple configurations.");

    m.def(
        "gradient_single_atom",
        [](DescriptorKind &ds, int index, IArr &species, IArr &neighbors, DArr &coordinates,
           DArr &desc, DArr &dE_ddesc) {
            if (sf_of(ds)) {
                check_layout(species, neighbors, neighbors.size(), coordinates, 1, &index);
                check_width(dE_ddesc, 1, ds.width, "dE_ddesc");
            }
            int n_atoms = (int)coordinates.shape(0);
            int n_neigh = (int)neighbors.shape(0);
            auto d = zeros({coordinates.shape(0), coordinates.shape(1)});
            double *out = d.mutable_data();
            std::vector<double> throwaway(dp(desc), dp(desc) + desc.size());
            run(ds, [&] {
                gradient_single_atom(index, n_atoms, ip(species), ip(neighbors), n_neigh,
                                     dp(coordinates), out, throwaway.data(), dp(dE_ddesc), &ds);
            });
            return d;
        },
        "Compute gradient of descriptor for single atom configuration.");

    m.def(
        "gradient",
        [](DescriptorKind &ds, int n_atoms, IArr &species, IArr &neighbors, IArr &number_of_neighbors,
           DArr &coordinates, DArr &desc, DArr &dE_ddesc) {
            if (sf_of(ds)) {
                check_layout(species, neighbors, sum_counts(number_of_neighbors, n_atoms), coordinates,
                             n_atoms, nullptr);
                check_width(dE_ddesc, n_atoms, ds.width, "dE_ddesc");
            }
            auto d = zeros({coordinates.shape(0), coordinates.shape(1)});
            double *out = d.mutable_data();
            if (auto *sf = sf_of(ds)) {
                // the reference hands Enzyme a throwaway copy of desc; the forward
                // it re-adds there is never seen, so no desc is passed at all
                int rc = 0;
                run(ds, [&] {
                    rc = ldsf_gradient(sf->handle(), n_atoms, ip(species), ip(neighbors),
                                       ip(number_of_neighbors), dp(coordinates), out, nullptr,
                                       dp(dE_ddesc));
                });
                raise_status(rc);
                return d;
            }
            std::vector<double> throwaway(dp(desc), dp(desc) + desc.size());
            run(ds, [&] {
                gradient(n_atoms, ip(species), ip(neighbors), ip(number_of_neighbors), dp(coordinates),
                         out, throwaway.data(), dp(dE_ddesc), &ds);
            });
            return d;
        },
        "Compute gradient of descriptor for complete configuration.");

    m.def(
        "param_gradient",
        [](DescriptorKind &ds, int n_atoms, IArr &species, IArr &neighbors, IArr &number_of_neighbors,
           DArr &coordinates, DArr &dE_ddesc) {
            auto *sf = sf_of(ds);
            if (!sf) throw std::invalid_argument("param_gradient: symmetry functions only");
            check_layout(species, neighbors, sum_counts(number_of_neighbors, n_atoms), coordinates,
                         n_atoms, nullptr);
            check_width(dE_ddesc, n_atoms, ds.width, "dE_ddesc");
            const int ns = ldsf_sf_num_species(sf->handle());
            auto dpar = zeros({(py::ssize_t)ldsf_sf_num_params(sf->handle())});
            auto drc = zeros({(py::ssize_t)ns, (py::ssize_t)ns});
            int rc = 0;
            run(ds, [&] {
                rc = ldsf_param_gradient(sf->handle(), n_atoms, ip(species), ip(neighbors),
                                         ip(number_of_neighbors), dp(coordinates), dp(dE_ddesc),
                                         dpar.mutable_data(), drc.mutable_data());
            });
            raise_status(rc);
            return py::make_tuple(dpar, drc);
        },
        "dE/d(parameters) (initDescriptor's flat layout) and dE/d(cutoff matrix): what the "
        "reference's Enzyme gradient computes into the descriptor's shadow and discards.");

    m.def(
        "gradient_batch",
        [](DescriptorKind &ds, IArr &n_atoms, IArr &config_ptr, IArr &species, IArr &neighbors,
           IArr &number_of_neighbors, DArr &coordinates, DArr &desc, DArr &dE_ddesc) {
            int n_conf = (int)n_atoms.size();
            if (sf_of(ds)) {
                auto ctr = batch_centres(n_atoms, config_ptr);
                check_layout(species, neighbors, sum_counts(number_of_neighbors, (long)ctr.size()),
                             coordinates, (long)ctr.size(), ctr.data());
                check_width(dE_ddesc, (long)ctr.size(), ds.width, "dE_ddesc");
            }
            auto d = zeros({coordinates.shape(0), coordinates.shape(1)});
            double *out = d.mutable_data();
            if (auto *sf = sf_of(ds)) {  // no throwaway desc, as in gradient
                int rc = 0;
                run(ds, [&] {
                    rc = ldsf_gradient_batch(sf->handle(), n_conf, ip(n_atoms), ip(config_ptr),
                                             ip(species), ip(neighbors), ip(number_of_neighbors),
                                             dp(coordinates), out, nullptr, dp(dE_ddesc));
                });
                raise_status(rc);
                return d;
            }
            std::vector<double> throwaway(dp(desc), dp(desc) + desc.size());
            run(ds, [&] {
                gradient_batch(n_conf, ip(n_atoms), ip(config_ptr), ip(species), ip(neighbors),
                               ip(number_of_neighbors), dp(coordinates), out, throwaway.data(),
                               dp(dE_ddesc), &ds);
            });
            return d;
        },
        "Compute gradient of descriptor for multiple configurations.");

    m.def(
        "jacobian",
        [](DescriptorKind &ds, int n_atoms, IArr &species, IArr &neighbors, IArr &number_of_neighbors,
           DArr &coordinates) {
            if (sf_of(ds))
                check_layout(species, neighbors, sum_counts(number_of_neighbors, n_atoms), coordinates,
                             n_atoms, nullptr);
            int n_total = (int)coordinates.shape(0);
            auto J = zeros({(py::ssize_t)ds.width * n_atoms, (py::ssize_t)coordinates.size()});
            double *out = J.mutable_data();
            run(ds, [&] {
                jacobian(n_atoms, n_total, ip(species), ip(neighbors), ip(number_of_neighbors),
                         dp(coordinates), out, &ds);
            });
            return J;
        },
        "Compute Jacobian of descriptor for complete configuration");

    m.def(
        "num_gradient_single_atom",
        [](DescriptorKind &ds, int index, IArr &species, IArr &neighbors, DArr &coordinates,
           DArr &dE_ddesc) {
            if (sf_of(ds)) {
                check_layout(species, neighbors, neighbors.size(), coordinates, 1, &index);
                check_width(dE_ddesc, 1, ds.width, "dE_ddesc");
            }
            int n_atoms = (int)coordinates.shape(0);
            int n_neigh = (int)neighbors.shape(0);
            auto d = zeros({3});
            double *out = d.mutable_data();
            run(ds, [&] {
                num_gradient_single_atom(index, n_atoms, ip(species), ip(neighbors), n_neigh,
                                         dp(coordinates), out, dp(dE_ddesc), &ds);
            });
            return d;
        },
        "Compute gradient of descriptor for single atom configuration.");

    // ---- neighbour list (python_bindings.cpp:348-555) ----
    py::class_<NeighListHandle, std::shared_ptr<NeighListHandle>>(m, "NeighList", py::module_local())
        .def(py::init<>())
        .def(
            "build",
            [](NeighListHandle &self, py::array_t<double> coords, double influence_distance,
               py::array_t<double> cutoffs, py::array_t<int> need_neigh) {
                int n1 = (int)(coords.size() / 3), n2 = (int)need_neigh.size();
                if (n1 != n2)
                    py::print("* Error (Neighbor List) : \"\"coords\" size and \"need_neigh\" size do not match!\"",
                              py::arg("file") = py::module_::import("sys").attr("stderr"));
                int n = n1 <= n2 ? n1 : n2;
                int rc = ldsf_nbl_build(self.nl, n, coords.data(), influence_distance,
                                        (int)cutoffs.size(), cutoffs.data(), need_neigh.data());
                if (rc)
                    throw std::runtime_error("Cell size too large! (partilces fly away) or\n"
                                             "Collision of atoms happened!");
            },
            "Build the neighbor list.", py::arg("coords").noconvert(), py::arg("influence_distance"),
            py::arg("cutoffs").noconvert(), py::arg("need_neigh").noconvert())
        .def(
            "get_neigh",
            [](NeighListHandle &self, py::array_t<double> cutoffs, int neighbor_list_index,
               int particle_number) {
                int nn = 0;
                const int *nb = nullptr;
                int rc = ldsf_nbl_get_neigh(self.nl, (int)cutoffs.size(), cutoffs.data(),
                                            neighbor_list_index, particle_number, &nn, &nb);
                if (rc) {
                    if (neighbor_list_index >= ldsf_nbl_num_lists(self.nl))
                        throw std::runtime_error("neighbor_list_index = " + std::to_string(neighbor_list_index) +
                                                 " >= self.numberOfNeighborLists = " +
                                                 std::to_string(ldsf_nbl_num_lists(self.nl)));
                    if (neighbor_list_index >= 0 &&
                        cutoffs.data()[neighbor_list_index] > ldsf_nbl_list_cutoff(self.nl, neighbor_list_index))
                        throw std::runtime_error("cutoffs_data[neighbor_list_index] > self.lists[neighbor_list_index].cutoff");
                    throw std::runtime_error("particle_number = " + std::to_string(particle_number) + " < 0!");
                }
                py::array_t<int> arr(nn);
                if (nn) std::memcpy(arr.mutable_data(), nb, sizeof(int) * nn);
                return py::make_tuple(nn, arr);
            },
            py::arg("cutoffs").noconvert(), py::arg("neighbor_list_index"), py::arg("particle_number"))
        // extension: the flat layout of get_numneigh_and_neighlist_1D without a Python loop
        .def(
            "flatten",
            [](NeighListHandle &self, int n_query, int neighbor_list_index) {
                long total = ldsf_nbl_total(self.nl, neighbor_list_index, n_query);
                if (total < 0) throw std::runtime_error("flatten: bad list index or atom count");
                py::array_t<int> nn(n_query), nl(total);
                int rc = ldsf_nbl_flatten(self.nl, neighbor_list_index, n_query, nn.mutable_data(),
                                          nl.mutable_data());
                if (rc) throw std::runtime_error(ldsf_last_error());
                return py::make_tuple(nn, nl);
            },
            py::arg("n_query"), py::arg("neighbor_list_index") = 0);

    m.def("create", []() { return std::make_shared<NeighListHandle>(); }, "Create a new NeighList object.");

    m.def("get_neigh_kim", []() { return (void const *)&ldsf_nbl_get_neigh; });

    m.def(
        "create_paddings",
        [](double influence_distance, py::array_t<double> cell, py::array_t<int> pbc,
           py::array_t<double> coords, py::array_t<int> species) {
            int n1 = (int)(coords.size() / 3), n2 = (int)species.size();
            int n = n1 <= n2 ? n1 : n2;
            ldsf_padding *p = nullptr;
            int rc = ldsf_nbl_create_paddings(n, influence_distance, cell.data(), pbc.data(),
                                              coords.data(), species.data(), &p);
            if (rc) throw std::runtime_error("In inverting the cell matrix, the determinant is 0!");
            std::unique_ptr<ldsf_padding, void (*)(ldsf_padding *)> guard(p, ldsf_padding_free);
            int np = ldsf_padding_count(p);
            py::array_t<double> pc({np, 3});
            py::array_t<int> ps(np), pi(np);
            ldsf_padding_copy(p, pc.mutable_data(), ps.mutable_data(), pi.mutable_data());
            return py::make_tuple(pc, ps, pi);
        },
        py::arg("influence_distance"), py::arg("cell").noconvert(), py::arg("pbc").noconvert(),
        py::arg("coords").noconvert(), py::arg("species").noconvert());
}
