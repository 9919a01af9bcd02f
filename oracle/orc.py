"""ORACLE -- test infrastructure only: ctypes wrapper of oracle/sf_oracle.c.

Every function here is a checker; none of it is the measured product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

KIND_CODE = {"g1": 1, "g2": 2, "g3": 3, "g4": 4, "g5": 5}


def build():
    import subprocess
    subprocess.check_call(["bash", os.path.join(_HERE, "build_oracle.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_sf_create.restype = vp
        L.orc_sf_create.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _i32p, _i32p, _i32p,
                                    _f64p, ctypes.c_int]
        L.orc_sf_destroy.argtypes = [vp]
        L.orc_sf_width.argtypes = [vp]
        L.orc_sf_width.restype = ctypes.c_int
        L.orc_compute_atom.argtypes = [vp, ctypes.c_int, _i32p, _i32p, ctypes.c_int, _f64p, _f64p]
        L.orc_gradient_atom.argtypes = [vp, ctypes.c_int, _i32p, _i32p, ctypes.c_int, _f64p,
                                        _f64p, _f64p]
        L.orc_compute.argtypes = [vp, ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p]
        L.orc_gradient.argtypes = [vp, ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_compute_batch.argtypes = [vp, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p,
                                        _f64p, _f64p]
        L.orc_gradient_batch.argtypes = [vp, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p,
                                         _f64p, _f64p, _f64p]
        L.orc_fwd_adj_omp.argtypes = [vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _f64p,
                                      _f64p, _f64p, _f64p, ctypes.c_int]
        L.orc_fwd_adj_batch_omp.argtypes = [vp, ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p,
                                            _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                            ctypes.c_int]
        L.orc_num_gradient_single_atom.argtypes = [vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                                   ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_num_params.argtypes = [vp]
        L.orc_num_params.restype = ctypes.c_int
        L.orc_param_gradient.argtypes = [vp, ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p,
                                         _f64p, _f64p]
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleSF:
    """FP64 CPU oracle of a symmetry-function descriptor.

    Parameters follow DescriptorKind.init_descriptor's SF overload
    (src/Descriptors.cpp:579-590): ``species`` names, one cutoff (or an
    n_s x n_s matrix), ``types`` ("g1".."g5"), flat ``sizes`` [rows0, cols0, ...]
    and flat row-major ``params``.  ``species_resolved`` selects the build-side
    per-species / per-species-pair block layout.
    """

    def __init__(self, species, cutoff, types, sizes, params, species_resolved=False):
        ns = len(species)
        rc = np.asarray(cutoff, dtype=np.float64)
        if rc.ndim == 0 or rc.size == 1:
            rc = np.full((ns, ns), float(rc.reshape(-1)[0]))
        rc = _f64(rc.reshape(ns, ns))
        kinds = _i32([KIND_CODE[t] for t in types])
        rows = _i32(sizes[0::2])
        cols = _i32(sizes[1::2])
        par = _f64(params) if len(params) else np.zeros(1)
        self._h = lib().orc_sf_create(ns, rc, len(types), kinds, rows, cols, par,
                                      int(bool(species_resolved)))
        self.width = lib().orc_sf_width(self._h)
        self.n_species = ns

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.orc_sf_destroy(h)
            self._h = None

    # Descriptor::compute (src/Descriptors.cpp:121-136); returns a fresh (n, W) array
    def compute(self, n_atoms, species, neighlist, numneigh, coords):
        desc = np.zeros((n_atoms, self.width))
        lib().orc_compute(self._h, n_atoms, _i32(species), _i32(neighlist), _i32(numneigh),
                          _f64(coords), desc)
        return desc

    # DescriptorKind::compute / compute_single_atom (src/Descriptors.cpp:355-365)
    def compute_single_atom(self, index, species, neighbors, coords):
        desc = np.zeros(self.width)
        nb = _i32(neighbors)
        lib().orc_compute_atom(self._h, index, _i32(species), nb, len(nb), _f64(coords), desc)
        return desc

    # Descriptor::gradient (src/Descriptors.cpp:165-190); returns fresh (n_total, 3)
    def gradient(self, n_atoms, species, neighlist, numneigh, coords, dE_dzeta):
        c = _f64(coords)
        out = np.zeros_like(c)
        lib().orc_gradient(self._h, n_atoms, _i32(species), _i32(neighlist), _i32(numneigh), c,
                           _f64(dE_dzeta), out)
        return out

    # the shadow of the descriptor object in Descriptor::gradient (src/Descriptors.cpp:176-188):
    # dE/d(parameters) in initDescriptor's flat layout and dE/d(rcut matrix)
    def param_gradient(self, n_atoms, species, neighlist, numneigh, coords, dE_dzeta):
        L = lib()
        dp = np.zeros(L.orc_num_params(self._h))
        ns = self.n_species
        drc = np.zeros(ns * ns)
        L.orc_param_gradient(self._h, n_atoms, _i32(species), _i32(neighlist), _i32(numneigh),
                             _f64(coords), _f64(dE_dzeta), dp, drc)
        return dp, drc.reshape(ns, ns)

    # Descriptor::gradient_single_atom (src/Descriptors.cpp:368-395)
    def gradient_single_atom(self, index, species, neighbors, coords, dE_dzeta):
        c = _f64(coords)
        out = np.zeros_like(c)
        nb = _i32(neighbors)
        lib().orc_gradient_atom(self._h, index, _i32(species), nb, len(nb), c, _f64(dE_dzeta),
                                out)
        return out

    # Descriptor::compute_batch (src/Descriptors.cpp:138-163)
    def compute_batch(self, n_atoms, config_ptr, species, neighlist, numneigh, coords):
        n_atoms = _i32(n_atoms)
        desc = np.zeros((int(n_atoms.sum()), self.width))
        lib().orc_compute_batch(self._h, len(n_atoms), n_atoms, _i32(config_ptr), _i32(species),
                                _i32(neighlist), _i32(numneigh), _f64(coords), desc)
        return desc

    # Descriptor::gradient_batch (src/Descriptors.cpp:263-292)
    def gradient_batch(self, n_atoms, config_ptr, species, neighlist, numneigh, coords, dE_dzeta):
        n_atoms = _i32(n_atoms)
        c = _f64(coords)
        out = np.zeros_like(c)
        lib().orc_gradient_batch(self._h, len(n_atoms), n_atoms, _i32(config_ptr), _i32(species),
                                 _i32(neighlist), _i32(numneigh), c, _f64(dE_dzeta), out)
        return out

    # Descriptor::num_gradient_single_atom (src/Descriptors.cpp:488-518)
    def num_gradient_single_atom(self, index, species, neighbors, coords, dE_dzeta):
        c = _f64(coords)
        out = np.zeros(3)
        nb = _i32(neighbors)
        lib().orc_num_gradient_single_atom(self._h, index, c.shape[0], _i32(species), nb, len(nb),
                                           c, _f64(dE_dzeta), out)
        return out

    def fwd_adj_batch_omp(self, n_atoms, config_ptr, species, neighlist, numneigh, coords,
                          dE_dzeta, n_threads):
        """compute_batch + gradient_batch, OpenMP over configurations."""
        n_atoms = _i32(n_atoms)
        c = _f64(coords)
        desc = np.zeros((int(n_atoms.sum()), self.width))
        out = np.zeros_like(c)
        lib().orc_fwd_adj_batch_omp(self._h, len(n_atoms), n_atoms, _i32(config_ptr), c.shape[0],
                                    _i32(species), _i32(neighlist), _i32(numneigh), c,
                                    _f64(dE_dzeta), desc, out, int(n_threads))
        return desc, out

    def fwd_adj_omp(self, n_centers, species, neighlist, offsets, coords, dE_dzeta, n_threads):
        """Forward + adjoint of the first ``n_centers`` centres on ``n_threads``
        host threads (CPU baseline timing)."""
        c = _f64(coords)
        desc = np.zeros((n_centers, self.width))
        out = np.zeros_like(c)
        lib().orc_fwd_adj_omp(self._h, n_centers, c.shape[0], _i32(species), _i32(neighlist),
                              _i32(offsets), c, _f64(dE_dzeta), desc, out, int(n_threads))
        return desc, out
