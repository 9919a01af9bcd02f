"""Neighbour list and ghost-gradient fold for the SF input layout, without ase.

Mirrors python_package/libdescriptor/neighbor.py:10-272.  ``NeighborList``
accepts either an ase.Atoms-like object (``get_positions()``,
``get_chemical_symbols()``, ``cell``, ``get_pbc()``) as the reference does, or
plain arrays through ``NeighborList.from_arrays``.  The flat int32
``numneigh``/``neighlist`` come from the C++ list in one call instead of the
reference's per-atom Python loop (neighbor.py:141-152); the order is the same.
"""
from typing import Dict, List, Tuple

import numpy as np

from .libdescriptor import create, create_paddings


class NeighborListError(Exception):
    def __init__(self, msg):
        super().__init__(msg)
        self.msg = msg

    def __expr__(self):
        return self.msg


class _Conf:
    """Minimal ase.Atoms stand-in built from arrays."""

    def __init__(self, positions, cell, pbc, symbols):
        self._pos = np.asarray(positions, dtype=np.double).reshape(-1, 3)
        self.cell = np.asarray(cell, dtype=np.double).reshape(3, 3)
        self._pbc = np.asarray(pbc, dtype=bool).reshape(3)
        self._sym = list(symbols)

    def get_positions(self):
        return self._pos

    def get_chemical_symbols(self):
        return list(self._sym)

    def get_pbc(self):
        return self._pbc

    def get_global_number_of_atoms(self):
        return self._pos.shape[0]


class NeighborList:
    """neighbor.py:10-224.  ``infl_dist`` is the cutoff used for both padding
    and the list."""

    def __init__(self, conf, infl_dist: float, padding_need_neigh: bool = False):
        self.conf = conf
        self.infl_dist = infl_dist
        self.padding_need_neigh = padding_need_neigh
        self.coords = None
        self.species = None
        self.image = None
        self.padding_coords = None
        self.padding_species = None
        self.padding_image = None
        self.neigh = create()
        self.create_neigh()

    @classmethod
    def from_arrays(cls, positions, cell, pbc, symbols, infl_dist, padding_need_neigh=False):
        return cls(_Conf(positions, cell, pbc, symbols), infl_dist, padding_need_neigh)

    def _n_contrib(self):
        c = self.conf
        if hasattr(c, "get_global_number_of_atoms"):
            return c.get_global_number_of_atoms()
        return len(c.get_positions())

    def create_neigh(self):
        coords_cb = np.ascontiguousarray(self.conf.get_positions(), dtype=np.double)
        species_cb = list(self.conf.get_chemical_symbols())
        cell = np.ascontiguousarray(np.asarray(self.conf.cell, dtype=np.double))
        pbc = np.ascontiguousarray(np.asarray(self.conf.get_pbc(), dtype=np.intc))
        # padding atoms carry the index of their symbol; the reference maps
        # atomic numbers back through ase.data (neighbor.py:42-52)
        uniq = sorted(set(species_cb))
        code = {s: k for k, s in enumerate(uniq)}
        species_code_cb = np.asarray([code[s] for s in species_cb], dtype=np.intc)
        try:
            coords_pd, species_code_pd, image_pd = create_paddings(
                self.infl_dist, cell, pbc, coords_cb, species_code_cb)
        except RuntimeError:
            raise NeighborListError("Calling `neighlist.create_paddings` failed.")
        species_pd = [uniq[c] for c in species_code_pd]
        self.padding_coords = np.asarray(coords_pd, dtype=np.double)
        self.padding_species = species_pd
        self.padding_image = np.asarray(image_pd, dtype=np.intc)
        num_cb = coords_cb.shape[0]
        num_pd = self.padding_coords.shape[0]
        self.coords = np.ascontiguousarray(np.concatenate((coords_cb, self.padding_coords)),
                                           dtype=np.double)
        self.species = np.concatenate((np.asarray(species_cb, dtype=object),
                                       np.asarray(species_pd, dtype=object)))
        self.image = np.asarray(np.concatenate((np.arange(num_cb), image_pd)), dtype=np.intc)
        need_neigh = np.ones(num_cb + num_pd, dtype=np.intc)
        if not self.padding_need_neigh:
            need_neigh[num_cb:] = 0
        cutoffs = np.asarray([self.infl_dist], dtype=np.double)
        try:
            self.neigh.build(self.coords, self.infl_dist, cutoffs, need_neigh)
        except RuntimeError:
            raise NeighborListError("Calling `neighlist.build` failed.")

    def get_neigh(self, index: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        cutoffs = np.asarray([self.infl_dist], dtype=np.double)
        try:
            _, neigh_indices = self.neigh.get_neigh(cutoffs, 0, index)
        except RuntimeError:
            raise NeighborListError("Calling `neighlist.get_neigh` failed.")
        return neigh_indices, self.coords[neigh_indices], self.species[neigh_indices]

    def get_numneigh_and_neighlist_1D(self, request_padding: bool = False):
        if request_padding:
            if not self.padding_need_neigh:
                raise NeighborListError(
                    "Request to get neighbors of padding atoms, but "
                    '"padding_need_neigh" is set to "False" at initialization.')
            n = len(self.coords)
        else:
            n = self._n_contrib()
        numneigh, neighlist = self.neigh.flatten(n, 0)
        return np.asarray(numneigh, dtype=np.intc), np.asarray(neighlist, dtype=np.intc)

    def get_coords(self) -> np.ndarray:
        return self.coords.copy()

    def get_species(self) -> List[str]:
        return list(self.species)

    def get_species_code(self, mapping: Dict[str, int]) -> np.ndarray:
        return np.asarray([mapping[s] for s in self.species], dtype=np.intc)

    def get_image(self) -> np.ndarray:
        return self.image.copy()

    def get_padding_coords(self) -> np.ndarray:
        return self.padding_coords.copy()

    def get_padding_species(self) -> List[str]:
        return list(self.padding_species)

    def get_padding_species_code(self, mapping: Dict[str, int]) -> np.ndarray:
        return np.asarray([mapping[s] for s in self.padding_species], dtype=np.intc)

    def get_padding_image(self) -> np.ndarray:
        return self.padding_image.copy()


def assemble_forces(forces: np.ndarray, n: int, padding_image: np.ndarray) -> np.ndarray:
    """Fold gradient rows of padding atoms onto their contributing atoms
    (neighbor.py:227-263): returns (n, 3), rounded as the reference rounds it.

    The reference has two summation orders.  With fewer contributing atoms
    than padding rows it adds, to every atom, numpy's axis-0 sum of its image
    rows -- a sequential sum from +0 in ascending padding order (+0 for an atom
    without images, so a -0.0 component becomes +0.0).  Otherwise it adds the
    padding rows one at a time in padding order.  Both are kept: the first as
    one vectorised step per image rank (the k-th image of every atom at once),
    the second with np.add.at, which applies repeated indices in order."""
    forces = np.asarray(forces)
    padding_image = np.asarray(padding_image, dtype=np.intp)
    total = np.array(forces[:n])
    n_pad = padding_image.size
    if n_pad == 0:
        return total
    pad = forces[n:n + n_pad]
    if n < n_pad:
        order = np.argsort(padding_image, kind="stable")  # each atom's images, ascending
        img = padding_image[order]
        counts = np.bincount(img, minlength=n)
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        rank = np.arange(n_pad) - first[img]  # k: the k-th image of its atom
        s = np.zeros((n, 3), dtype=total.dtype)
        for k in range(int(counts.max())):
            sel = rank == k
            s[img[sel]] += pad[order[sel]]
        total += s
    else:
        np.add.at(total, padding_image, pad)
    return total
