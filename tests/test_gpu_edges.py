"""Edge cases of the reference's inputs through the GPU entry points, against
the oracle: an empty input, centres with no neighbours or with every listed
neighbour beyond the cutoff (ragged counts with zeros), a neighbour exactly at
rc, exactly collinear triplets (1 + lambda cos(theta) = 0, where the
reference's sym_g4 takes its base <= 0 branch, src/SymmetryFunctions.cpp:118-119),
a repeated neighbour entry, and neighbourhoods of more than 64 atoms (the
general kernels take them).

Tolerances: FP64 descriptors elementwise relative <= 1e-12 (rows that are
zero in the oracle must be exactly zero), FP64 gradients normwise <= 1e-12;
FP32-bulk descriptors elementwise relative <= 1e-5 with an absolute floor of
1e-6 of the largest entry, FP32-bulk gradients normwise <= 1e-5."""
import numpy as np
import pytest

import libdescriptor_amd as ld
from conftest import behler, elem_rel_err, rel_err
from oracle import orc
from test_oracle import cluster, full_list

pytestmark = pytest.mark.gpu
TOL64, TOL32 = 1e-12, 1e-5


def descriptor(rc, precision):
    types, sizes, params = behler()
    return ld.DescriptorKind.init_symmetry_functions(["Si"], rc, types, sizes, list(params),
                                                     precision=precision)


def check_both(rc, n, sp, nl, nn, x, seed=3, gradient=True):
    """FP64 and FP32-bulk compute (and gradient) against the oracle."""
    O = orc.OracleSF(["Si"], rc, *behler())
    d0 = O.compute(n, sp, nl, nn, x)
    g = np.random.default_rng(seed).normal(size=d0.shape)
    G0 = O.gradient(n, sp, nl, nn, x, g) if gradient else None
    for prec in (ld.FP64, ld.FP32_BULK):
        ds = descriptor(rc, prec)
        d = ld.compute(ds, n, sp, nl, nn, x)
        assert d.shape == d0.shape
        if prec == ld.FP64:
            assert elem_rel_err(d, d0, 1e-300) <= TOL64
        else:
            assert elem_rel_err(d, d0, 1e-6 * np.abs(d0).max()) <= TOL32
        if gradient:
            G = ld.gradient(ds, n, sp, nl, nn, x, d.copy(), g)
            assert G.shape == G0.shape
            assert rel_err(G, G0) <= (TOL64 if prec == ld.FP64 else TOL32)
        # rows the oracle leaves at zero stay exactly zero
        np.testing.assert_array_equal(d[np.all(d0 == 0, axis=1)], 0.0)
    return d0, G0


def test_empty_input():
    print("case: empty input", flush=True)
    for prec in (ld.FP64, ld.FP32_BULK):
        ds = descriptor(4.0, prec)
        e_i, e_x = np.zeros(0, np.int32), np.zeros((0, 3))
        d = ld.compute(ds, 0, e_i, e_i, e_i, e_x)
        assert d.shape == (0, ds.width)
        G = ld.gradient(ds, 0, e_i, e_i, e_i, e_x, d, np.zeros((0, ds.width)))
        assert G.shape == (0, 3)
        # no contributing atoms, two padding rows: a zero gradient of every row
        x2 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
        G = ld.gradient(ds, 0, np.zeros(2, np.int32), e_i, e_i, x2, d, np.zeros((0, ds.width)))
        np.testing.assert_array_equal(G, np.zeros((2, 3)))


def test_isolated_and_out_of_range_neighbours():
    """A cluster, three far-away atoms with no neighbours, and lists built with
    a larger radius than rc, so that some centres list only atoms past rc."""
    print("case: isolated atoms and neighbours past rc", flush=True)
    rc = 4.0
    pos = np.vstack([cluster(12, 7), [[40.0, 0, 0], [0, 40.0, 0], [0, 0, 40.0]],
                     [[20.0, 20.0, 20.0], [24.5, 20.0, 20.0]]])
    nn, nl = full_list(pos, rc + 1.0)
    assert (nn == 0).sum() == 3 and nn[-1] == 1 and nn[-2] == 1
    n = len(pos)
    d0, G0 = check_both(rc, n, np.zeros(n, np.int32), nl, nn, pos)
    assert np.all(d0[12:] == 0.0) and np.all(G0[12:] == 0.0)


def test_neighbour_exactly_at_cutoff():
    """r = rc exactly: fc = 0 and, as the reference's branch, no gradient."""
    print("case: neighbour at r = rc", flush=True)
    rc = 4.0
    pos = np.array([[0.0, 0.0, 0.0], [4.0, 0.0, 0.0], [0.0, 2.5, 0.0], [0.0, 0.0, 3.0]])
    nn, nl = full_list(pos, rc + 0.5)
    n = len(pos)
    check_both(rc, n, np.zeros(n, np.int32), nl, nn, pos)


def test_collinear_triplets():
    """Atoms on a line (and one off it): for the centre in the middle,
    cos(theta) = -1, so 1 + cos(theta) = 0 for lambda = +1."""
    print("case: collinear triplets", flush=True)
    rc = 4.5
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [-2.0, 0.0, 0.0], [0.0, 0.0, 2.5],
                    [4.0, 0.0, 0.0]])
    nn, nl = full_list(pos, rc)
    n = len(pos)
    check_both(rc, n, np.zeros(n, np.int32), nl, nn, pos)


def test_repeated_neighbour_entry():
    """A neighbour listed twice counts twice, as in the reference's loops: its
    pair terms twice, and triplets with itself at r_jk = 0 (fc_jk = 1).  The
    forward only: the derivative of r_jk = |x_k - x_j| at 0 is undefined."""
    print("case: repeated neighbour entry", flush=True)
    rc = 4.0
    pos = cluster(8, 9)
    nn, nl = full_list(pos, rc)
    nl = nl.tolist()
    first = nl[0]
    nl.insert(1, first)
    nn = nn.copy()
    nn[0] += 1
    n = len(pos)
    check_both(rc, n, np.zeros(n, np.int32), np.array(nl, np.int32), nn, pos, gradient=False)


def test_more_than_64_neighbours():
    """A dense cluster whose neighbourhoods exceed 64 atoms: the general kernels."""
    import torch
    from libdescriptor_amd import _capi
    print("case: more than 64 neighbours", flush=True)
    rc = 6.0
    pos = cluster(110, 4, box=9.0, dmin=1.2)
    nn, nl = full_list(pos, rc)
    assert nn.max() > 64
    n = len(pos)
    check_both(rc, n, np.zeros(n, np.int32), nl, nn, pos)
    dv = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.asarray(a), dtype=torch.int32, device=dv)  # noqa: E731
    plan = _capi.Plan(descriptor(rc, ld.FP64), n, t(np.zeros(n)), t(nl), t(nn))
    assert plan.kernel_family == 0
