"""ORACLE -- test infrastructure only: ctypes wrapper of oracle/_ref/libsfref.so,
the reference's own SF forward, parameter-file reader and neighbour list
compiled from its sources by oracle/build_ref.sh.

``available()`` is False where the reference tree was not present at build
time; tests that need it skip there and rely on the committed golden fixtures.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "libsfref.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def available():
    return os.path.exists(_LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        cpp = ctypes.POINTER(ctypes.c_char_p)
        L.ref_sf_create.restype = vp
        L.ref_sf_create.argtypes = [ctypes.c_int, cpp, ctypes.c_double, ctypes.c_int, cpp, _i32p,
                                    _f64p]
        L.ref_sf_create_from_file.restype = vp
        L.ref_sf_create_from_file.argtypes = [ctypes.c_char_p]
        L.ref_sf_width.restype = ctypes.c_int
        L.ref_sf_width.argtypes = [vp]
        L.ref_sf_destroy.argtypes = [vp]
        L.ref_compute.argtypes = [vp, ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p]
        L.ref_compute_single_atom.argtypes = [vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                              ctypes.c_int, _f64p, _f64p]
        L.ref_compute_batch.argtypes = [vp, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _f64p,
                                        _f64p]
        L.ref_num_gradient_single_atom.argtypes = [vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                                   ctypes.c_int, _f64p, _f64p, _f64p]
        L.ref_create_paddings.restype = ctypes.c_int
        L.ref_create_paddings.argtypes = [ctypes.c_int, ctypes.c_double, _f64p, _i32p, _f64p, _i32p,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]
        L.ref_neighlist.restype = ctypes.c_long
        L.ref_neighlist.argtypes = [ctypes.c_int, _f64p, ctypes.c_double, _i32p, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
        _lib = L
    return _lib


def _cstrs(names):
    arr = (ctypes.c_char_p * len(names))()
    arr[:] = [n.encode() for n in names]
    return arr


class RefSF:
    """The reference's SymmetryFunctions (vector init with one cutoff, or a file)."""

    def __init__(self, species=None, cutoff=None, types=None, sizes=None, params=None, path=None):
        L = lib()
        if path is not None:
            self._h = L.ref_sf_create_from_file(path.encode())
        else:
            sz = np.ascontiguousarray(sizes, dtype=np.int32)
            par = np.ascontiguousarray(params if len(params) else [0.0], dtype=np.float64)
            self._h = L.ref_sf_create(len(species), _cstrs(species), float(cutoff), len(types),
                                      _cstrs(types), sz, par)
        self.width = L.ref_sf_width(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.ref_sf_destroy(h)
            self._h = None

    def compute(self, n_atoms, species, neighlist, numneigh, coords):
        desc = np.zeros((n_atoms, self.width))
        lib().ref_compute(self._h, n_atoms, np.ascontiguousarray(species, np.int32),
                          np.ascontiguousarray(neighlist, np.int32),
                          np.ascontiguousarray(numneigh, np.int32),
                          np.ascontiguousarray(coords, np.float64), desc)
        return desc

    def compute_single_atom(self, index, species, neighbors, coords):
        desc = np.zeros(self.width)
        c = np.ascontiguousarray(coords, np.float64)
        nb = np.ascontiguousarray(neighbors, np.int32)
        lib().ref_compute_single_atom(self._h, index, c.shape[0],
                                      np.ascontiguousarray(species, np.int32), nb, len(nb), c, desc)
        return desc

    def compute_batch(self, n_atoms, config_ptr, species, neighlist, numneigh, coords):
        n_atoms = np.ascontiguousarray(n_atoms, np.int32)
        desc = np.zeros((int(n_atoms.sum()), self.width))
        lib().ref_compute_batch(self._h, len(n_atoms), n_atoms,
                                np.ascontiguousarray(config_ptr, np.int32),
                                np.ascontiguousarray(species, np.int32),
                                np.ascontiguousarray(neighlist, np.int32),
                                np.ascontiguousarray(numneigh, np.int32),
                                np.ascontiguousarray(coords, np.float64), desc)
        return desc

    def num_gradient_single_atom(self, index, species, neighbors, coords, dE_dzeta):
        c = np.ascontiguousarray(coords, np.float64)
        nb = np.ascontiguousarray(neighbors, np.int32)
        out = np.zeros(3)
        lib().ref_num_gradient_single_atom(self._h, index, c.shape[0],
                                           np.ascontiguousarray(species, np.int32), nb, len(nb), c,
                                           out, np.ascontiguousarray(dE_dzeta, np.float64))
        return out


def create_paddings(cutoff, cell, pbc, coords, species):
    L = lib()
    cell = np.ascontiguousarray(cell, np.float64).reshape(9)
    pbc = np.ascontiguousarray(pbc, np.int32)
    coords = np.ascontiguousarray(coords, np.float64)
    species = np.ascontiguousarray(species, np.int32)
    n = L.ref_create_paddings(coords.shape[0], float(cutoff), cell, pbc, coords, species, 0,
                              None, None, None)
    if n < 0:
        raise RuntimeError("In inverting the cell matrix, the determinant is 0!")
    pc = np.zeros((n, 3))
    ps = np.zeros(n, np.int32)
    pi = np.zeros(n, np.int32)
    L.ref_create_paddings(coords.shape[0], float(cutoff), cell, pbc, coords, species, n,
                          pc.ctypes.data, ps.ctypes.data, pi.ctypes.data)
    return pc, ps, pi


def neighlist(coords, infl, need_neigh, n_query):
    L = lib()
    coords = np.ascontiguousarray(coords, np.float64)
    nn_flag = np.ascontiguousarray(need_neigh, np.int32)
    total = L.ref_neighlist(coords.shape[0], coords, float(infl), nn_flag, n_query, None, None, 0)
    if total < 0:
        raise RuntimeError("Cell size too large! (partilces fly away) or\nCollision of atoms happened!")
    numneigh = np.zeros(n_query, np.int32)
    nl = np.zeros(total, np.int32)
    L.ref_neighlist(coords.shape[0], coords, float(infl), nn_flag, n_query,
                    numneigh.ctypes.data, nl.ctypes.data, total)
    return numneigh, nl
