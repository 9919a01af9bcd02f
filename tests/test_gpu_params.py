"""Hyperparameter gradients on the GPU (ldsf_param_gradient / ldsf_plan_param_gradient):
dE/d(parameters) and dE/d(cutoff matrix), what the reference's Enzyme gradient computes into
the shadow of the descriptor object and discards (src/Descriptors.cpp:176-188), against the
oracle (itself pinned by central finite differences in test_oracle.py).

Tolerance: normwise relative <= 1e-12 (FP64; the sums run in a different fixed order)."""
import numpy as np
import pytest

import libdescriptor_amd as ld
from conftest import behler, rel_err
from oracle import orc
from test_oracle import ALL_KINDS, cluster, full_list

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("case", ["all_kinds", "all_kinds_two_species"])
def test_param_gradient_all_kinds(case):
    print(f"case: param gradient {case}", flush=True)
    pos = cluster(14, 5)
    types, sizes, params = ALL_KINDS
    if case == "all_kinds":
        species, rc, resolved = ["Si"], np.array([[4.0]]), False
        sp = np.zeros(len(pos), np.int32)
    else:
        species, rc, resolved = ["A", "B"], np.array([[4.0, 4.2], [3.9, 3.8]]), True
        sp = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    nn, nl = full_list(pos, 4.2)
    n = len(pos)
    ds = ld.DescriptorKind.init_symmetry_functions(species, rc.reshape(-1), types, sizes, params,
                                                   species_resolved=resolved)
    O = orc.OracleSF(species, rc, types, sizes, params, species_resolved=resolved)
    g = np.random.default_rng(9).normal(size=(n, ds.width))
    dp, drc = ld.param_gradient(ds, n, sp, nl, nn, pos, g)
    dp0, drc0 = O.param_gradient(n, sp, nl, nn, pos, g)
    assert dp.shape == (len(params),) and drc.shape == rc.shape
    assert rel_err(dp, dp0) <= TOL
    assert rel_err(drc, drc0) <= TOL


def test_param_gradient_two_species_system_on_device():
    """A 6000-atom two-species periodic system through a device plan (the 64-entry
    species-resolved set), twice: equal to the oracle and bitwise the same on both runs."""
    import torch
    from libdescriptor_amd import _capi, workloads as w
    print("case: param gradient 6000-atom two-species plan", flush=True)
    pos, cell, pbc, spc = w.random_system(6000, 0.05, 1.8, seed=21)
    L = w.layout(pos, cell, pbc, spc, 5.0, symbols=["A", "B"])
    n = L["n_contrib"]
    types, sizes, params = behler()
    ds = ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, params,
                                                   species_resolved=True)
    d = torch.device("cuda:0")
    t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device=d)  # noqa: E731
    plan = _capi.Plan(ds, len(L["coords"]), t(L["species"]), t(L["neighlist"]),
                      t(L["numneigh"]))
    g = np.random.default_rng(4).normal(size=(n, ds.width))
    x, gd = t(L["coords"], torch.float64), t(g, torch.float64)
    outs = []
    for _ in range(2):
        dp = torch.zeros(len(params), dtype=torch.float64, device=d)
        drc = torch.zeros((2, 2), dtype=torch.float64, device=d)
        plan.param_gradient(x, gd, dp, drc)
        torch.cuda.synchronize()
        outs.append((dp.cpu().numpy(), drc.cpu().numpy()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    O = orc.OracleSF(["A", "B"], 5.0, types, sizes, params, species_resolved=True)
    dp0, drc0 = O.param_gradient(n, L["species"], L["neighlist"], L["numneigh"], L["coords"], g)
    assert rel_err(outs[0][0], dp0) <= TOL
    assert rel_err(outs[0][1], drc0) <= TOL


def test_param_gradient_accumulates(c1):
    """+= into caller buffers, like the gradient entry points."""
    import ctypes
    from libdescriptor_amd import _capi
    print("case: param gradient += (C1)", flush=True)
    ds = ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"], "cos",
                                           4.5, *behler())
    n_par = _capi.lib().ldsf_sf_num_params(ctypes.c_void_p(ds._handle))
    assert n_par == 8 * 2 + 16 * 3
    sp, nl, nn = (np.ascontiguousarray(c1[k], np.int32) for k in ("species", "neighlist", "numneigh"))
    x, g = np.ascontiguousarray(c1["coords"]), np.ascontiguousarray(c1["dE_dzeta"])
    dp, drc = np.ones(n_par), np.ones(1)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    _capi.check(_capi.lib().ldsf_param_gradient(ctypes.c_void_p(ds._handle), 64, p(sp), p(nl),
                                                p(nn), p(x), p(g), p(dp), p(drc)))
    dp0, drc0 = orc.OracleSF(["Si"], 4.5, *behler()).param_gradient(64, sp, nl, nn, x, g)
    assert rel_err(dp - 1.0, dp0) <= TOL
    assert rel_err(drc - 1.0, drc0.reshape(-1)) <= TOL
