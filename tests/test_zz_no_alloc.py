"""The allocation-free gradient entry point, ldsf_gradient_noalloc, through the
C-ABI against the entry point it is a variant of, ldsf_gradient (the
reference's Descriptor::gradient, src/Descriptors.cpp:165-190: d_coordinates
+= J^T dE/dzeta, desc += the descriptors).  Both run the same kernels over the
same centres, so the results must be bitwise equal (integer/index work and the
same floating-point operations in the same order), in both precision forms:

  two calls in a row on one GPU, the second with a smaller input (the scratch
  array keeps the first call's contents, so stale counters or records would
  show up here);
  a dense cluster whose neighbourhoods exceed 64 atoms (the general kernels),
  also checked against the FP64 oracle (descriptors <= 1e-12 elementwise,
  dE/dr <= 1e-12 normwise);
  rejected inputs (a neighbour index out of range, a count above
  max_neighbors, counts that do not sum to n_neigh) leave both outputs
  unchanged, and a valid call afterwards is still exact;
  an input beyond LDSF_NOALLOC_SCRATCH_BYTES is refused before any device work
  (CPU).
"""

import numpy as np
import pytest

from conftest import behler, elem_rel_err, rel_err

SCRATCH_BYTES = 64 << 20


def sf(prec):
    import libdescriptor_amd as ld
    types, sizes, params = behler()
    return ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, params,
                                                     species_resolved=True, precision=prec)


def system(n, seed, kind="bulk", density=0.05):
    from libdescriptor_amd import workloads as w
    if kind == "bulk":
        pos, cell, pbc, sp = w.random_system(n, density, 1.8, seed=seed)
    else:
        pos, cell, pbc, sp = w.cluster(n, density, 1.2 if density > 0.1 else 1.8, seed=seed)
    L = w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"])
    rng = np.random.default_rng(seed + 1)
    L["dE"] = rng.normal(size=(L["n_contrib"], 64))
    # outputs start from non-zero values: both entry points add (+=)
    L["desc0"] = rng.normal(size=(L["n_contrib"], 64))
    L["dx0"] = rng.normal(size=L["coords"].shape)
    return L


def host_gradient(ds, L):
    """ldsf_gradient on host arrays (the existing entry point)."""
    from libdescriptor_amd import _capi
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    sp, nl, nn, x, g = (i32(L["species"]), i32(L["neighlist"]), i32(L["numneigh"]),
                        f64(L["coords"]), f64(L["dE"]))
    desc, dx = L["desc0"].copy(), L["dx0"].copy()
    _capi.check(_capi.lib().ldsf_gradient(ds._handle, L["n_contrib"], sp.ctypes.data,
                                          nl.ctypes.data, nn.ctypes.data, x.ctypes.data,
                                          dx.ctypes.data, desc.ctypes.data, g.ctypes.data))
    return desc, dx


def dev(L):
    import torch
    t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device="cuda")  # noqa
    return dict(sp=t(L["species"]), nl=t(L["neighlist"]), nn=t(L["numneigh"]),
                x=t(L["coords"], torch.float64), dE=t(L["dE"], torch.float64),
                desc=t(L["desc0"], torch.float64), dx=t(L["dx0"], torch.float64))


def noalloc(ds, L, D, prec, n_neigh=None, max_nn=None):
    """ldsf_gradient_noalloc on device tensors; returns the status code."""
    import torch
    from libdescriptor_amd import _capi
    n = L["n_contrib"]
    nn = np.asarray(L["numneigh"])
    n_neigh = int(nn.sum()) if n_neigh is None else n_neigh
    max_nn = int(nn.max()) if max_nn is None else max_nn
    torch.cuda.synchronize()
    rc = _capi.lib().ldsf_gradient_noalloc(
        ds._handle, n, len(L["coords"]), n_neigh, max_nn, D["sp"].data_ptr(), D["nl"].data_ptr(),
        D["nn"].data_ptr(), D["x"].data_ptr(), D["dx"].data_ptr(), D["desc"].data_ptr(),
        D["dE"].data_ptr(), int(prec), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_two_calls_equal_gradient(prec):
    from libdescriptor_amd import _capi
    ds = sf(prec)
    # the second input is smaller: the scratch holds the first call's contents
    for n, kind, seed in ((20000, "bulk", 11), (1500, "cluster", 12)):
        print(f"case: noalloc prec={prec} n={n} {kind}", flush=True)
        L = system(n, seed, kind)
        D = dev(L)
        _capi.check(noalloc(ds, L, D, prec))
        desc, dx = host_gradient(ds, L)
        assert np.array_equal(D["desc"].cpu().numpy(), desc)
        assert np.array_equal(D["dx"].cpu().numpy(), dx)


@pytest.mark.gpu
def test_general_kernels_and_oracle():
    """Neighbourhoods above 64 atoms run the general kernels."""
    from libdescriptor_amd import _capi
    from oracle import orc
    L = system(400, 21, "cluster", density=0.15)
    assert int(np.asarray(L["numneigh"]).max()) > 64
    n = L["n_contrib"]
    # outputs from zero here, so that they are the oracle's values themselves
    L["desc0"] = np.zeros_like(L["desc0"])
    L["dx0"] = np.zeros_like(L["dx0"])
    types, sizes, params = behler()
    O = orc.OracleSF(["A", "B"], 5.0, types, sizes, params, species_resolved=True)
    d0 = O.compute(n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    g0 = O.gradient(n, L["species"], L["neighlist"], L["numneigh"], L["coords"], L["dE"])
    for prec in (0, 1):
        print(f"case: noalloc general kernels prec={prec}", flush=True)
        ds = sf(prec)
        D = dev(L)
        _capi.check(noalloc(ds, L, D, prec))
        desc, dx = host_gradient(ds, L)
        assert np.array_equal(D["desc"].cpu().numpy(), desc)
        assert np.array_equal(D["dx"].cpu().numpy(), dx)
        if prec == 0:
            assert elem_rel_err(desc, d0, 1e-300) <= 1e-12
            assert rel_err(dx, g0) <= 1e-12


@pytest.mark.gpu
def test_rejected_inputs_leave_outputs():
    import torch
    from libdescriptor_amd import _capi
    ds = sf(1)
    L = system(3000, 31)
    n, nn = L["n_contrib"], np.asarray(L["numneigh"])
    bad_index = dict(L, neighlist=np.array(L["neighlist"], copy=True))
    bad_index["neighlist"][7] = len(L["coords"]) + 3
    cases = [("index out of range", bad_index, {}),
             ("count above max_neighbors", L, dict(max_nn=int(nn.max()) - 1)),
             ("counts do not sum to n_neigh", L, dict(n_neigh=int(nn.sum()) - 1))]
    for name, LL, kw in cases:
        print(f"case: noalloc rejects {name}", flush=True)
        if "max_nn" in kw and kw["max_nn"] * n < int(nn.sum()):
            continue  # refused on the host already
        D = dev(LL)
        rc = noalloc(ds, LL, D, 1, **kw)
        assert rc == 1, name  # LDSF_ERR_ARG
        assert torch.equal(D["desc"].cpu(), torch.tensor(LL["desc0"]))
        assert torch.equal(D["dx"].cpu(), torch.tensor(LL["dx0"]))
    D = dev(L)
    _capi.check(noalloc(ds, L, D, 1))
    desc, dx = host_gradient(ds, L)
    assert np.array_equal(D["desc"].cpu().numpy(), desc)
    assert np.array_equal(D["dx"].cpu().numpy(), dx)


def test_oversize_refused_before_device_work():
    from libdescriptor_amd import _capi
    ds = sf(1)
    dummy = np.zeros(16, dtype=np.float64)  # never dereferenced: refused on the host
    p = dummy.ctypes.data
    n, n_neigh = 60000, 2_100_000  # 32 B per record: past 64 MiB
    rc = _capi.lib().ldsf_gradient_noalloc(ds._handle, n, n, n_neigh, 64, p, p, p, p, p, p, p, 1,
                                           None)
    assert rc == 3  # LDSF_ERR_UNSUPPORTED
    assert b"LDSF_NOALLOC_SCRATCH_BYTES" in _capi.lib().ldsf_last_error()
    assert SCRATCH_BYTES < 32 * (n_neigh + n)
