"""The product's padding + neighbour list (libdescriptor_amd, C++) against the
reference's compiled neighbour code (golden fixture / oracle/_ref), the numpy
restatement (oracle.nbl) and brute-force membership.  CPU only."""
import numpy as np
import pytest

import libdescriptor_amd as ld
from conftest import golden
from oracle import nbl, ref


def product_layout(pos, cell, pbc, species, rc):
    pc, ps, pi = ld.create_paddings(rc, np.ascontiguousarray(cell, dtype=np.double),
                                    np.ascontiguousarray(pbc, dtype=np.intc),
                                    np.ascontiguousarray(pos, dtype=np.double),
                                    np.ascontiguousarray(species, dtype=np.intc))
    coords = np.ascontiguousarray(np.concatenate([pos, pc]))
    need = np.r_[np.ones(len(pos)), np.zeros(len(pc))].astype(np.intc)
    n = ld.create()
    n.build(coords, rc, np.array([rc]), need)
    nn, nlist = n.flatten(len(pos), 0)
    return pc, ps, pi, coords, nn, nlist, n


def test_golden_c1_layout_bit_exact(c1):
    pc, ps, pi, coords, nn, nlist, _ = product_layout(c1["positions"], c1["cell"], [1, 1, 1],
                                                      np.zeros(64, np.int32), 4.5)
    np.testing.assert_array_equal(coords, c1["coords"])
    np.testing.assert_array_equal(pi, c1["image"])
    np.testing.assert_array_equal(nn, c1["numneigh"])
    np.testing.assert_array_equal(nlist, c1["neighlist"])


def test_golden_two_species_layout():
    t = golden("two_species_ref.npz")
    L = float(t["L"])
    pc, ps, pi, coords, nn, nlist, _ = product_layout(t["positions"], np.eye(3) * L, [1, 1, 1],
                                                      t["species0"], 5.0)
    np.testing.assert_array_equal(coords, t["coords"])
    np.testing.assert_array_equal(np.concatenate([t["species0"], ps]), t["species"])
    np.testing.assert_array_equal(nn, t["numneigh"])
    np.testing.assert_array_equal(nlist, t["neighlist"])


CELLS = [
    (np.eye(3) * 9.0, [1, 1, 1]),
    (np.array([[8.0, 0, 0], [2.5, 7.5, 0], [1.0, -1.5, 9.0]]), [1, 1, 1]),  # triclinic
    (np.eye(3) * 7.0, [1, 0, 1]),                                          # one open direction
    (np.eye(3) * 6.0, [0, 0, 0]),                                          # cluster
]


@pytest.mark.parametrize("case", range(len(CELLS)))
def test_random_cells_match_numpy_restatement(case):
    cell, pbc = CELLS[case]
    rng = np.random.default_rng(case)
    frac = rng.random((40, 3))
    pos = frac @ cell
    sp = rng.integers(0, 2, 40).astype(np.int32)
    rc = 4.1
    pc, ps, pi, coords, nn, nlist, _ = product_layout(pos, cell, pbc, sp, rc)
    pc2, ps2, pi2 = nbl.create_paddings(rc, cell, pbc, pos, sp)
    np.testing.assert_array_equal(pc, pc2)
    np.testing.assert_array_equal(ps, ps2)
    np.testing.assert_array_equal(pi, pi2)
    need = np.r_[np.ones(40), np.zeros(len(pc))].astype(np.int32)
    nn2, nl2 = nbl.neighlist(coords, rc, need, 40)
    np.testing.assert_array_equal(nn, nn2)
    np.testing.assert_array_equal(nlist, nl2)
    # membership by brute force
    off = np.r_[0, np.cumsum(nn)]
    for i, s in enumerate(nbl.brute_force_sets(coords, rc, 40)):
        np.testing.assert_array_equal(np.sort(nlist[off[i]:off[i + 1]]), s)
    if ref.available():
        pc3, ps3, pi3 = ref.create_paddings(rc, cell, pbc, pos, sp)
        np.testing.assert_array_equal(pc, pc3)
        np.testing.assert_array_equal(pi, pi3)
        nn3, nl3 = ref.neighlist(coords, rc, need, 40)
        np.testing.assert_array_equal(nlist, nl3)


def test_get_neigh_and_multiple_cutoffs():
    rng = np.random.default_rng(4)
    pos = rng.random((60, 3)) * 12.0
    n = ld.create()
    cut = np.array([3.0, 4.5])
    n.build(pos, 4.5, cut, np.ones(60, np.intc))
    for i in (0, 17, 59):
        c0, a0 = n.get_neigh(cut, 0, i)
        c1, a1 = n.get_neigh(cut, 1, i)
        d = np.linalg.norm(pos - pos[i], axis=1)
        assert set(a0) == set(np.flatnonzero((d < 3.0) & (d > 0)))
        assert set(a1) == set(np.flatnonzero((d < 4.5) & (d > 0)))
        assert c0 == len(a0) and c1 == len(a1)
    with pytest.raises(RuntimeError):
        n.get_neigh(cut, 2, 0)
    with pytest.raises(RuntimeError):
        n.get_neigh(np.array([5.0, 5.0]), 0, 0)
    with pytest.raises(RuntimeError):
        n.get_neigh(cut, 0, -1)


def test_errors_collision_and_singular_cell():
    pos = np.array([[0.0, 0, 0], [1e-6, 0, 0], [3.0, 0, 0]])
    n = ld.create()
    with pytest.raises(RuntimeError):
        n.build(pos, 4.0, np.array([4.0]), np.ones(3, np.intc))
    with pytest.raises(RuntimeError):
        ld.create_paddings(4.0, np.zeros((3, 3)), np.ones(3, np.intc), pos, np.zeros(3, np.intc))
    with pytest.raises(ld.NeighborListError):
        ld.NeighborList.from_arrays(pos, np.eye(3) * 5, [1, 1, 1], ["Si"] * 3, 4.0)


def test_neighborlist_class_and_fold(c1):
    nlc = ld.NeighborList.from_arrays(c1["positions"], c1["cell"], [1, 1, 1], ["Si"] * 64, 4.5)
    nn, nlist = nlc.get_numneigh_and_neighlist_1D()
    np.testing.assert_array_equal(nn, c1["numneigh"])
    np.testing.assert_array_equal(nlist, c1["neighlist"])
    np.testing.assert_array_equal(nlc.get_species_code({"Si": 0}), c1["species"])
    f = np.random.default_rng(0).normal(size=(len(nlc.get_coords()), 3))
    # 64 atoms, more padding rows: the per-atom np.sum branch, bitwise
    assert len(nlc.get_padding_image()) > 64
    np.testing.assert_array_equal(ld.assemble_forces(f, 64, nlc.get_padding_image()),
                                  nbl.assemble_forces(f, 64, nlc.get_padding_image()))
    with pytest.raises(ld.NeighborListError):
        nlc.get_numneigh_and_neighlist_1D(request_padding=True)


def _signed_forces(rng, rows):
    f = rng.normal(size=(rows, 3)) * 10.0 ** rng.integers(-6, 6, size=(rows, 3))
    f[rng.random((rows, 3)) < 0.2] = -0.0
    return f


def test_numpy_axis0_sum_is_sequential_from_plus_zero():
    """The reference's fold (neighbor.py:245-253) adds np.sum(rows, axis=0);
    the oracle restates it with the same call, and the device fold sums each
    atom's image rows one after another from +0.  Pin that the two agree
    bitwise: numpy's axis-0 sum of an (m, 3) array is the sequential sum from
    +0, the sign of zero included (a lone -0.0 sums to +0.0)."""
    rng = np.random.default_rng(3)
    for m in list(range(0, 40)) + [100, 1000]:
        x = _signed_forces(rng, m)
        seq = np.zeros(3)
        for k in range(m):
            seq = seq + x[k]
        s = np.sum(x, axis=0)
        assert np.array_equal(s, seq) and np.array_equal(np.signbit(s), np.signbit(seq)), m
    assert not np.signbit(np.sum(np.array([[-0.0, -0.0, -0.0]]), axis=0)).any()


@pytest.mark.parametrize("n,n_pad", [(7, 60), (60, 7), (3, 500), (20, 20), (5, 0)])
def test_fold_branches_bitwise_and_placement(n, n_pad):
    """Both branches of assemble_forces (neighbor.py:245-256: per-atom np.sum
    when n < n_padding, row by row otherwise): the product's host fold equals
    the oracle bitwise (signed zeros included), and on integer-valued forces,
    where every order is exact, both equal np.add.at -- which pins where each
    padding row is added."""
    rng = np.random.default_rng(n * 1000 + n_pad)
    img = rng.integers(0, n, n_pad).astype(np.int32)
    f = _signed_forces(rng, n + n_pad)
    a, b = ld.assemble_forces(f, n, img), nbl.assemble_forces(f, n, img)
    np.testing.assert_array_equal(a, b)
    assert np.array_equal(np.signbit(a), np.signbit(b))
    fi = rng.integers(-1000, 1000, size=(n + n_pad, 3)).astype(np.float64)
    ref = fi[:n].copy()
    np.add.at(ref, img, fi[n:])
    np.testing.assert_array_equal(nbl.assemble_forces(fi, n, img), ref)
    np.testing.assert_array_equal(ld.assemble_forces(fi, n, img), ref)
