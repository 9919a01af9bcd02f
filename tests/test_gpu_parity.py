"""GPU parity: the HIP path (through the pybind module / C++ API / C-ABI, and
through the C-ABI plan path from torch) against the oracle and the reference's
golden outputs on the same inputs.

Tolerances (written here, used below):
  FP64 descriptors      : elementwise relative <= 1e-12
  FP64 gradients        : max|diff| <= 1e-12 * max|oracle| (normwise; entries
                          can cancel to ~0, so no elementwise relative bound)
  FP32-bulk descriptors : elementwise relative <= 1e-5
  FP32-bulk gradients   : normwise <= 1e-5
Neighbour indexing and descriptor ordering are bit-exact by construction: the
kernels consume the golden neighbour list and write the reference's slots.
"""
import numpy as np
import pytest

import libdescriptor_amd as ld
from conftest import behler, elem_rel_err, golden, rel_err
from oracle import orc
from test_oracle import ALL_KINDS, cluster, full_list

pytestmark = pytest.mark.gpu

SF = ld.AvailableDescriptors.SymmetryFunctions
TOL64 = 1e-12
TOL32 = 1e-5


def sf_desc(species=("Si",), rc=4.5, kinds=None, resolved=False, precision=0):
    types, sizes, params = kinds if kinds else behler()
    return ld.DescriptorKind.init_symmetry_functions(list(species), rc, types, sizes, list(params),
                                                     species_resolved=resolved, precision=precision)


def oracle_of(species=("Si",), rc=4.5, kinds=None, resolved=False):
    types, sizes, params = kinds if kinds else behler()
    return orc.OracleSF(list(species), rc, types, sizes, params, species_resolved=resolved)


def test_c1_forward_matches_reference_golden(c1):
    ds = ld.DescriptorKind.init_descriptor(SF, ["Si"], "cos", 4.5, *behler())
    d = ld.compute(ds, 64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"])
    assert d.shape == (64, 24)
    assert elem_rel_err(d, c1["desc"], 1e-300) <= TOL64


def test_c1_gradient_matches_oracle(c1):
    ds = ld.DescriptorKind.init_descriptor(SF, ["Si"], "cos", 4.5, *behler())
    g = c1["dE_dzeta"]
    G = ld.gradient(ds, 64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"],
                    c1["desc"], g)
    O = oracle_of()
    G0 = O.gradient(64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"], g)
    assert G.shape == c1["coords"].shape
    assert rel_err(G, G0) <= TOL64
    # padding rows carry gradient too (reference layout (n_total, 3))
    assert np.abs(G[64:]).max() > 0


def test_single_atom_entry_points(c1):
    ds = ld.DescriptorKind.init_descriptor(SF, ["Si"], "cos", 4.5, *behler())
    O = oracle_of()
    off = np.r_[0, np.cumsum(c1["numneigh"])]
    for i in (0, 13, 63):
        nb = c1["neighlist"][off[i]:off[i + 1]]
        d1 = ld.compute_single_atom(ds, i, c1["species"], nb, c1["coords"])
        d2 = ds.compute(i, c1["species"], nb, c1["coords"])
        assert elem_rel_err(d1, c1["desc"][i], 1e-300) <= TOL64
        np.testing.assert_array_equal(d1, d2)
        g = c1["dE_dzeta"][i]
        G = ld.gradient_single_atom(ds, i, c1["species"], nb, c1["coords"], c1["desc"][i], g)
        G0 = O.gradient_single_atom(i, c1["species"], nb, c1["coords"], g)
        assert rel_err(G, G0) <= TOL64


def test_reference_fd_routine_on_gpu(c1):
    ds = ld.DescriptorKind.init_descriptor(SF, ["Si"], "cos", 4.5, *behler())
    off = np.r_[0, np.cumsum(c1["numneigh"])]
    for a, i in enumerate(c1["fd_atoms"]):
        nb = c1["neighlist"][off[i]:off[i + 1]]
        fd = ld.num_gradient_single_atom(ds, int(i), c1["species"], nb, c1["coords"],
                                         c1["dE_dzeta"][i])
        # descriptor differences at +-k h (h ~ 0.037 A) divided by 105 h
        assert rel_err(fd, c1["fd"][a]) <= 1e-9


@pytest.mark.parametrize("seed", [0, 1])
def test_all_kinds_general_kernels(seed):
    pos = cluster(14, seed)
    nn, nl = full_list(pos, 4.0)
    sp = np.zeros(len(pos), np.int32)
    ds = sf_desc(rc=4.0, kinds=ALL_KINDS)
    O = oracle_of(rc=4.0, kinds=ALL_KINDS)
    d = ld.compute(ds, len(pos), sp, nl, nn, pos)
    d0 = O.compute(len(pos), sp, nl, nn, pos)
    assert elem_rel_err(d, d0, 1e-300) <= TOL64
    g = np.random.default_rng(seed).normal(size=d.shape)
    G = ld.gradient(ds, len(pos), sp, nl, nn, pos, d, g)
    assert rel_err(G, O.gradient(len(pos), sp, nl, nn, pos, g)) <= TOL64


def small_two_species(n=200, seed=3):
    from libdescriptor_amd import workloads as w
    pos, cell, pbc, sp = w.random_system(n, 0.05, 1.8, seed=seed)
    return w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"])


def test_species_resolved_two_species():
    L = small_two_species()
    n = L["n_contrib"]
    ds = sf_desc(("A", "B"), 5.0, resolved=True)
    assert ds.width == 64
    O = oracle_of(("A", "B"), 5.0, resolved=True)
    d = ld.compute(ds, n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    d0 = O.compute(n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    assert elem_rel_err(d, d0, 1e-300) <= TOL64
    g = np.random.default_rng(0).normal(size=d.shape)
    G = ld.gradient(ds, n, L["species"], L["neighlist"], L["numneigh"], L["coords"], d, g)
    assert rel_err(G, O.gradient(n, L["species"], L["neighlist"], L["numneigh"], L["coords"], g)) <= TOL64


def test_batch_entry_points():
    from libdescriptor_amd import workloads as w
    lays = []
    for k, n in enumerate([8, 30, 111, 17]):
        if k % 2 == 0:
            pos, cell, pbc, sp = w.random_system(n, 0.05, 1.8, seed=k)
        else:
            pos, cell, pbc, sp = w.cluster(n, 0.05, 1.8, seed=k)
        lays.append(w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"]))
    B = w.concat(lays)
    ds = sf_desc(("A", "B"), 5.0, resolved=True)
    O = oracle_of(("A", "B"), 5.0, resolved=True)
    args = (B["n_atoms"], B["config_ptr"], B["species"], B["neighlist"], B["numneigh"], B["coords"])
    d = ld.compute_batch(ds, *args)
    d0 = O.compute_batch(*args)
    assert elem_rel_err(d, d0, 1e-300) <= TOL64
    g = np.random.default_rng(1).normal(size=d.shape)
    G = ld.gradient_batch(ds, *args, d, g)
    assert rel_err(G, O.gradient_batch(*args, g)) <= TOL64
    # each configuration alone gives the same rows
    row = 0
    for k, L in enumerate(lays):
        n = L["n_contrib"]
        dk = ld.compute(ds, n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
        np.testing.assert_array_equal(dk, d[row:row + n])
        row += n


def test_jacobian_matches_one_hot_oracle():
    pos = cluster(7, 5, box=5.0)
    nn, nl = full_list(pos, 4.0)
    sp = np.zeros(len(pos), np.int32)
    ds = sf_desc(rc=4.0)
    O = oracle_of(rc=4.0)
    J = ld.jacobian(ds, len(pos), sp, nl, nn, pos)
    W = ds.width
    assert J.shape == (W * len(pos), pos.size)
    for row in (0, 5, W + 3, W * len(pos) - 1):
        g = np.zeros(W * len(pos)); g[row] = 1.0
        ref = O.gradient(len(pos), sp, nl, nn, pos, g.reshape(len(pos), W)).reshape(-1)
        assert rel_err(J[row], ref, floor=1e-300) <= TOL64


def test_file_initialised_descriptors():
    import os
    from conftest import GOLDEN
    t = golden("two_species_ref.npz")
    ds = ld.DescriptorKind.init_descriptor(os.path.join(GOLDEN, "sf_two_species.params"), SF)
    d = ld.compute(ds, 96, t["species"], t["neighlist"], t["numneigh"], t["coords"])
    assert elem_rel_err(d, t["desc_file"], 1e-300) <= TOL64
    c1 = golden("c1_ref.npz")
    pr = golden("params_ref.npz")
    ds = ld.DescriptorKind.init_descriptor(os.path.join(GOLDEN, "sf_g2g4.params"), SF)
    d = ld.compute(ds, 64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"])
    assert elem_rel_err(d, pr["desc_g2g4_file"], 1e-300) <= TOL64


def test_fp32_bulk_error_bounds(c1):
    for species, rc, resolved, L in [(("Si",), 4.5, False, None),
                                     (("A", "B"), 5.0, True, small_two_species())]:
        if L is None:
            L = dict(coords=c1["coords"], species=c1["species"], neighlist=c1["neighlist"],
                     numneigh=c1["numneigh"], n_contrib=64)
        n = L["n_contrib"]
        args = (n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
        d64 = ld.compute(sf_desc(species, rc, resolved=resolved), *args)
        ds32 = sf_desc(species, rc, resolved=resolved, precision=ld.FP32_BULK)
        d32 = ld.compute(ds32, *args)
        assert elem_rel_err(d32, d64, 1e-30) <= TOL32
        g = np.random.default_rng(2).normal(size=d64.shape)
        G64 = ld.gradient(sf_desc(species, rc, resolved=resolved), *args, d64, g)
        G32 = ld.gradient(ds32, *args, d64, g)
        assert rel_err(G32, G64) <= TOL32


def test_torch_plan_path_and_autograd(c1):
    import torch
    from libdescriptor_amd.autograd import SymmetryFunctionsLayer
    ds = sf_desc()
    layer = SymmetryFunctionsLayer(ds, c1["species"], c1["neighlist"], c1["numneigh"])
    x = torch.tensor(c1["coords"], device="cuda", requires_grad=True)
    d = layer(x)
    ref_d = ld.compute(ds, 64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"])
    np.testing.assert_array_equal(d.detach().cpu().numpy(), ref_d)
    w = torch.tensor(c1["dE_dzeta"], device="cuda")
    (d * w).sum().backward()
    G = ld.gradient(ds, 64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"],
                    ref_d, c1["dE_dzeta"])
    np.testing.assert_array_equal(x.grad.cpu().numpy(), G)


def test_gradient_is_deterministic():
    L = small_two_species(400, seed=9)
    n = L["n_contrib"]
    ds = sf_desc(("A", "B"), 5.0, resolved=True)
    args = (n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    g = np.random.default_rng(4).normal(size=(n, 64))
    G1 = ld.gradient(ds, *args, np.zeros((n, 64)), g)
    G2 = ld.gradient(ds, *args, np.zeros((n, 64)), g)
    np.testing.assert_array_equal(G1, G2)


def test_out_of_range_neighbour_is_rejected(c1):
    import torch
    from libdescriptor_amd import _capi
    ds = sf_desc()
    bad = c1["neighlist"].copy()
    bad[5] = len(c1["coords"])  # one past the last row
    with pytest.raises(ValueError):
        ld.compute(ds, 64, c1["species"], bad, c1["numneigh"], c1["coords"])
    bad[5] = -3
    with pytest.raises(ValueError):
        ld.gradient(ds, 64, c1["species"], bad, c1["numneigh"], c1["coords"], c1["desc"],
                    c1["dE_dzeta"])
    # the device plan validates indices on the GPU before any kernel reads them
    bad[5] = len(c1["coords"]) + 7
    t = lambda a: torch.tensor(a, dtype=torch.int32, device="cuda")  # noqa: E731
    with pytest.raises(_capi.LdsfError):
        _capi.Plan(ds, len(c1["coords"]), t(c1["species"]), t(bad), t(c1["numneigh"]))
    sp = c1["species"].copy()
    sp[100] = 4  # species code beyond the parameter set
    with pytest.raises(_capi.LdsfError):
        _capi.Plan(ds, len(c1["coords"]), t(sp), t(c1["neighlist"]), t(c1["numneigh"]))


def test_c2_full_size_properties():
    from libdescriptor_amd import workloads as w
    L = w.config_c2()
    n = L["n_contrib"]
    ds = sf_desc()
    args = (n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    d = ld.compute(ds, *args)
    O = oracle_of()
    m = 1500  # oracle on the first m centres
    d0 = O.compute(m, L["species"], L["neighlist"], L["numneigh"][:m], L["coords"])
    assert elem_rel_err(d[:m], d0, 1e-300) <= TOL64
    g = np.random.default_rng(6).normal(size=d.shape)
    G = ld.gradient(ds, *args, d, g)
    off = np.r_[0, np.cumsum(L["numneigh"])].astype(np.int32)
    _, G0 = O.fwd_adj_omp(n, L["species"], L["neighlist"], off, L["coords"], g, 8)
    assert rel_err(G, G0) <= TOL64
    assert np.abs(G.sum(axis=0)).max() <= 1e-11 * np.abs(G).max() * len(G) ** 0.5


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("resolved", [False, True])
def test_wave_and_general_kernels_agree(precision, resolved):
    """Both kernel families on the same plan: the wave kernels (used by
    default where they apply) and the thread-per-centre kernels."""
    import torch
    from libdescriptor_amd import _capi
    L = small_two_species(500, seed=11)
    species = ("A", "B")
    ds = sf_desc(species, 5.0, resolved=resolved)
    t = lambda a, dt=torch.int32: torch.tensor(a, dtype=dt, device="cuda")  # noqa: E731
    plan = _capi.Plan(ds, len(L["coords"]), t(L["species"]), t(L["neighlist"]),
                      t(L["numneigh"]))
    assert plan.kernel_family == 1
    n = L["n_contrib"]
    x = t(L["coords"], torch.float64)
    g = t(np.random.default_rng(5).normal(size=(n, ds.width)), torch.float64)
    out = {}
    for fam in (0, 1):
        plan.kernel_family = fam
        d = torch.zeros((n, ds.width), dtype=torch.float64, device="cuda")
        dx = torch.zeros_like(x)
        plan.compute(x, d, precision)
        plan.gradient(x, g, dx, precision)
        torch.cuda.synchronize()
        out[fam] = (d.cpu().numpy(), dx.cpu().numpy())
    tol = TOL64 if precision == 0 else TOL32
    assert elem_rel_err(out[1][0], out[0][0], 1e-30) <= tol
    assert rel_err(out[1][1], out[0][1]) <= tol
    O = oracle_of(species, 5.0, resolved=resolved)
    d0 = O.compute(n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    G0 = O.gradient(n, L["species"], L["neighlist"], L["numneigh"], L["coords"], g.cpu().numpy())
    assert elem_rel_err(out[1][0], d0, 1e-30) <= tol
    assert rel_err(out[1][1], G0) <= tol


def test_adjoint_edges_then_gather_equals_gradient():
    import torch
    from libdescriptor_amd import _capi
    L = small_two_species(300, seed=12)
    ds = sf_desc(("A", "B"), 5.0, resolved=True)
    t = lambda a, dt=torch.int32: torch.tensor(a, dtype=dt, device="cuda")  # noqa: E731
    plan = _capi.Plan(ds, len(L["coords"]), t(L["species"]), t(L["neighlist"]),
                      t(L["numneigh"]))
    n = L["n_contrib"]
    x = t(L["coords"], torch.float64)
    g = t(np.random.default_rng(8).normal(size=(n, ds.width)), torch.float64)
    with pytest.raises(_capi.LdsfError):
        plan.gather(torch.zeros_like(x))  # no adjoint records yet
    for prec in (0, 1):
        a = torch.zeros_like(x)
        b = torch.zeros_like(x)
        plan.gradient(x, g, a, prec)
        plan.adjoint_edges(x, g, prec)
        plan.gather(b)
        torch.cuda.synchronize()
        assert torch.equal(a, b)


@pytest.mark.parametrize("precision", [0, 1])
def test_host_path_workspace_reuse(precision):
    """The host-pointer calls stage through one workspace per descriptor and
    device, kept at the largest size used.  Calls on 32,768, 64 and 64,000-atom
    systems in turn (the largest arrays split over several copy threads, each
    over several pinned chunks; the last call larger than the first) equal,
    bitwise, the device plan's kernels on the same inputs, and ldsf_compute /
    ldsf_gradient keep the reference's += on the caller's arrays."""
    import torch
    from libdescriptor_amd import _capi, workloads as w
    ds = sf_desc(("A", "B"), 5.0, resolved=True, precision=precision)
    lib = _capi.lib()
    P = lambda a: a.ctypes.data  # noqa: E731
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.int32: torch.tensor(a, dtype=dt, device=dev)  # noqa: E731
    for ncell, seed in ((16, 1), (2, 2), (20, 3)):
        pos, cell, pbc = w.diamond_si(ncell, seed=seed)
        spc = (np.random.default_rng(seed).random(len(pos)) < 0.3).astype(np.int32)
        L = w.layout(pos, cell, pbc, spc, 5.0, symbols=["A", "B"])
        n = L["n_contrib"]
        print(f"case: host workspace {n} atoms precision {precision}", flush=True)
        sp = np.ascontiguousarray(L["species"], np.int32)
        nl = np.ascontiguousarray(L["neighlist"], np.int32)
        nn = np.ascontiguousarray(L["numneigh"], np.int32)
        x = np.ascontiguousarray(L["coords"], np.float64)
        g = np.random.default_rng(seed).normal(size=(n, 64))
        plan = _capi.Plan(ds, len(x), t(sp), t(nl), t(nn))
        xd, gd = t(x, torch.float64), t(g, torch.float64)
        zd = lambda: torch.zeros((n, 64), dtype=torch.float64, device=dev)  # noqa: E731
        zx = lambda: torch.zeros((len(x), 3), dtype=torch.float64, device=dev)  # noqa: E731
        ref_c, ref_g, ref_fd, ref_fx = zd(), zx(), zd(), zx()
        plan.compute(xd, ref_c, precision)
        plan.gradient(xd, gd, ref_g, precision)
        plan.forward_gradient(xd, gd, ref_fd, ref_fx, precision)
        ref_c, ref_g = ref_c.cpu().numpy(), ref_g.cpu().numpy()
        ref_fd, ref_fx = ref_fd.cpu().numpy(), ref_fx.cpu().numpy()
        # the module: overwrite compute, gradient without a forward
        np.testing.assert_array_equal(ld.compute(ds, n, sp, nl, nn, x), ref_c)
        np.testing.assert_array_equal(ld.gradient(ds, n, sp, nl, nn, x, ref_c, g), ref_g)
        # the C-ABI: the reference's gradient (forward too) into zeroed arrays
        d0, x0 = np.zeros((n, 64)), np.zeros_like(x)
        assert lib.ldsf_gradient(ds._handle, n, P(sp), P(nl), P(nn), P(x), P(x0), P(d0), P(g)) == 0
        np.testing.assert_array_equal(d0, ref_fd)
        np.testing.assert_array_equal(x0, ref_fx)
        # += onto values already there
        d1 = np.full((n, 64), 0.5)
        assert lib.ldsf_compute(ds._handle, n, P(sp), P(nl), P(nn), P(x), P(d1)) == 0
        assert rel_err(d1 - 0.5, ref_c) <= 1e-14
        x1 = np.full_like(x, -0.25)
        assert lib.ldsf_gradient(ds._handle, n, P(sp), P(nl), P(nn), P(x), P(x1), None, P(g)) == 0
        assert rel_err(x1 + 0.25, ref_g) <= 1e-14


def test_plan_past_the_bucketed_build():
    """A plan over more than 4096 x 256 rows builds its transposed list with the
    count/fill/sort path instead of buckets: a 300-atom system with 1.1 M
    unreferenced rows appended gives, bitwise, the gradient of the plain
    system on its rows and zero on the rest, in both precisions."""
    import torch
    from libdescriptor_amd import _capi
    print("case: plan over 1.1M rows", flush=True)
    L = small_two_species(300, seed=21)
    ds = sf_desc(("A", "B"), 5.0, resolved=True)
    d = torch.device("cuda:0")
    t = lambda a, dt=torch.int32: torch.tensor(a, dtype=dt, device=d)  # noqa: E731
    R = len(L["coords"])
    extra = 1_100_000
    sp = np.concatenate([L["species"], np.zeros(extra, np.int32)])
    x = np.concatenate([L["coords"], np.full((extra, 3), 1e6)])
    n = L["n_contrib"]
    g = t(np.random.default_rng(3).normal(size=(n, ds.width)), torch.float64)
    small = _capi.Plan(ds, R, t(L["species"]), t(L["neighlist"]), t(L["numneigh"]))
    big = _capi.Plan(ds, R + extra, t(sp), t(L["neighlist"]), t(L["numneigh"]))
    for prec in (0, 1):
        a = torch.zeros((R, 3), dtype=torch.float64, device=d)
        b = torch.zeros((R + extra, 3), dtype=torch.float64, device=d)
        small.gradient(t(L["coords"], torch.float64), g, a, prec)
        big.gradient(t(x, torch.float64), g, b, prec)
        torch.cuda.synchronize()
        assert torch.equal(a, b[:R])
        assert not torch.any(b[R:])


def test_cpp_drop_in_caller(c1, tmp_path):
    """A reference-style C++ program (tests/cpp/drop_in.cpp) through
    Descriptor::compute / gradient / compute_single_atom on the GPU."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "_build", "drop_in")
    n, nt, nnl = 64, len(c1["coords"]), len(c1["neighlist"])
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.asarray([n, nt, nnl], np.int32).tofile(f)
        np.asarray(c1["species"], np.int32).tofile(f)
        np.asarray(c1["numneigh"], np.int32).tofile(f)
        np.asarray(c1["neighlist"], np.int32).tofile(f)
        np.asarray(c1["coords"], np.float64).tofile(f)
        np.asarray(c1["dE_dzeta"], np.float64).tofile(f)
    print(f"running {exe} on C1", flush=True)
    r = subprocess.run([exe, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(outp, dtype=np.float64)
    desc = out[: n * 24].reshape(n, 24)
    grad = out[n * 24: n * 24 + nt * 3].reshape(nt, 3)
    single = out[n * 24 + nt * 3: n * 24 + nt * 3 + 24]
    clone = out[n * 24 + nt * 3 + 24:].reshape(n, 24)
    assert elem_rel_err(desc, c1["desc"], 1e-300) <= TOL64
    assert elem_rel_err(single, c1["desc"][0], 1e-300) <= TOL64
    # clone_empty (src/SymmetryFunctions.cpp:410-427): the same tables with every
    # parameter zero -- g2 rows become fc(r), g4 rows 2 fc fc fc per triplet
    types, sizes, params = behler()
    Oz = orc.OracleSF(["Si"], 4.5, types, sizes, np.zeros(len(params)))
    dz = Oz.compute(n, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"])
    assert elem_rel_err(clone, dz, 1e-300) <= TOL64
    O = oracle_of()
    G0 = O.gradient(n, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"],
                    c1["dE_dzeta"])
    assert rel_err(grad, G0) <= TOL64


GRIDS = {
    # 3 g2 rows + a 2 x 1 x 1 g4 grid
    "small_grid": (["g2", "g4"], [3, 2, 2, 3],
                   [0.01, 0.0, 0.1, 0.5, 0.3, 1.0, 1.0, -1.0, 0.005, 4.0, -1.0, 0.005]),
    # g2 only
    "g2_only": (["g2"], [5, 2], [0.001, 0.0, 0.01, 0.0, 0.05, 1.0, 0.2, 0.0, 0.5, 2.0]),
    # g4 only, lambda not +/-1, a non-integer zeta
    "g4_general": (["g4"], [4, 3], [1.0, 0.5, 0.002, 2.5, 0.5, 0.002, 1.0, 0.5, 0.02,
                                    2.5, 0.5, 0.02]),
    # 2 g2 rows + the power-of-two grid zeta (1, 2, 4, 8) x lambda +-1 x 2 eta (the
    # FP32-bulk kernels' compile-time squaring chain for this grid)
    "zeta_1248": (["g2", "g4"], [2, 2, 16, 3],
                  [0.01, 0.0, 0.2, 0.0] + [v for z in (1.0, 2.0, 4.0, 8.0) for lam in (1.0, -1.0)
                                           for e in (0.002, 0.02) for v in (z, lam, e)]),
}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("grid", sorted(GRIDS))
def test_wave_kernels_on_partial_grids(grid, precision):
    """Parameter sets that fill only part of the wave kernels' 8-row g2 and
    4 x 2 x 2 g4 grid (the rest is evaluated with neutral entries and has no
    slot): wave kernels against the thread-per-centre kernels and the oracle."""
    import torch
    from libdescriptor_amd import _capi
    L = small_two_species(400, seed=17)
    species = ("A", "B")
    kinds = GRIDS[grid]
    ds = sf_desc(species, 5.0, kinds=kinds, resolved=True)
    t = lambda a, dt=torch.int32: torch.tensor(a, dtype=dt, device="cuda")  # noqa: E731
    plan = _capi.Plan(ds, len(L["coords"]), t(L["species"]), t(L["neighlist"]),
                      t(L["numneigh"]))
    assert plan.kernel_family == 1
    n = L["n_contrib"]
    x = t(L["coords"], torch.float64)
    g = t(np.random.default_rng(21).normal(size=(n, ds.width)), torch.float64)
    out = {}
    for fam in (0, 1, "fused"):
        plan.kernel_family = 1 if fam == "fused" else fam
        d = torch.full((n, ds.width), 0.0, dtype=torch.float64, device="cuda")
        dx = torch.zeros_like(x)
        if fam == "fused":
            plan.forward_gradient(x, g, d, dx, precision)
        else:
            plan.compute(x, d, precision)
            plan.gradient(x, g, dx, precision)
        torch.cuda.synchronize()
        out[fam] = (d.cpu().numpy(), dx.cpu().numpy())
    tol = TOL64 if precision == 0 else TOL32
    O = oracle_of(species, 5.0, kinds=kinds, resolved=True)
    d0 = O.compute(n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    G0 = O.gradient(n, L["species"], L["neighlist"], L["numneigh"], L["coords"], g.cpu().numpy())
    for fam in (0, 1, "fused"):
        sig = np.abs(d0) >= 1e-6 * np.abs(d0).max(axis=0, keepdims=True)
        rel = np.abs(out[fam][0] - d0)[sig] / np.abs(d0)[sig]
        assert rel.max() <= tol, (fam, rel.max())
        assert rel_err(out[fam][1], G0) <= tol
