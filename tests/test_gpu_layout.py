"""The device layout producer (ldsf_layout_*: padding atoms, KLIFF-order
neighbour list, flat layout) and the device ghost fold (assemble_forces)
against the reference's compiled neighbour code -- through the golden
fixtures and the SHA-256 digests of its outputs in
tests/golden/layout_digests.json (written by tests/golden/make_golden.py from
the reference's own nbl_create_paddings / nbl_build, so the reference's build
is not needed here) -- the numpy restatement (oracle.nbl) and brute-force
membership.  Integer and index outputs and padding coordinates are compared
bit for bit; the fold is compared bit for bit with the reference's summation
order."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def device_layout(pos, cell, pbc, species, rc):
    from libdescriptor_amd import _capi
    d = _dev()
    x = torch.tensor(np.ascontiguousarray(pos, dtype=np.float64), device=d)
    s = torch.tensor(np.ascontiguousarray(species, dtype=np.int32), device=d)
    L = _capi.DeviceLayout(x, s, cell, pbc, rc)
    torch.cuda.synchronize()
    return L


def as_np(L):
    return (L.coords.cpu().numpy(), L.species.cpu().numpy(), L.image.cpu().numpy(),
            L.numneigh.cpu().numpy(), L.neighlist.cpu().numpy())


def digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def reference_digests():
    with open(os.path.join(GOLDEN, "layout_digests.json")) as f:
        return json.load(f)


def check_digests(ref, pos, sp, n, coords, spd, image, nn, nl):
    """Inputs as the fixture's, then every output as the reference's, bit for bit."""
    assert digest(pos, "<f8") == ref["positions"] and digest(sp, "<i4") == ref["species_in"]
    assert len(coords) - n == ref["n_padding"] and len(nl) == ref["n_neigh"]
    assert digest(coords[n:], "<f8") == ref["padding_coords"]
    assert digest(spd[n:], "<i4") == ref["padding_species"]
    assert digest(image[n:], "<i4") == ref["padding_image"]
    assert digest(nn, "<i4") == ref["numneigh"]
    assert digest(nl, "<i4") == ref["neighlist"]


def test_golden_c1_layout_bit_exact(c1):
    print("case: C1 64-atom diamond Si, rc 4.5", flush=True)
    L = device_layout(c1["positions"], c1["cell"], [1, 1, 1], np.zeros(64, np.int32), 4.5)
    coords, sp, image, nn, nl = as_np(L)
    np.testing.assert_array_equal(coords, c1["coords"])
    np.testing.assert_array_equal(image, np.r_[np.arange(64), c1["image"]])
    np.testing.assert_array_equal(sp, c1["species"])
    np.testing.assert_array_equal(nn, c1["numneigh"])
    np.testing.assert_array_equal(nl, c1["neighlist"])
    assert L.max_neighbors == int(c1["numneigh"].max())


def test_golden_two_species_layout():
    print("case: two-species golden, rc 5.0", flush=True)
    t = golden("two_species_ref.npz")
    L = float(t["L"])
    Ld = device_layout(t["positions"], np.eye(3) * L, [1, 1, 1], t["species0"], 5.0)
    coords, sp, image, nn, nl = as_np(Ld)
    np.testing.assert_array_equal(coords, t["coords"])
    np.testing.assert_array_equal(sp, t["species"])
    np.testing.assert_array_equal(nn, t["numneigh"])
    np.testing.assert_array_equal(nl, t["neighlist"])


CELLS = [
    (np.eye(3) * 9.0, [1, 1, 1]),
    (np.array([[8.0, 0, 0], [2.5, 7.5, 0], [1.0, -1.5, 9.0]]), [1, 1, 1]),  # triclinic
    (np.eye(3) * 7.0, [1, 0, 1]),                                          # one open direction
    (np.eye(3) * 6.0, [0, 0, 0]),                                          # cluster
    (np.eye(3) * 3.1, [1, 1, 1]),                                          # cell < cutoff: 2 image shells
]


@pytest.mark.parametrize("case", range(len(CELLS)))
def test_random_cells_match_reference_order(case):
    from oracle import nbl
    cell, pbc = CELLS[case]
    print(f"case: random cell {case}", flush=True)
    rng = np.random.default_rng(case)
    nat = 40 if case < 4 else 6
    pos = rng.random((nat, 3)) @ cell
    sp = rng.integers(0, 2, nat).astype(np.int32)
    rc = 4.1
    coords, spd, image, nn, nl = as_np(device_layout(pos, cell, pbc, sp, rc))
    pc, ps, pi = nbl.create_paddings(rc, cell, pbc, pos, sp)
    np.testing.assert_array_equal(coords, np.concatenate([pos, pc]))
    np.testing.assert_array_equal(spd, np.concatenate([sp, ps]))
    np.testing.assert_array_equal(image, np.concatenate([np.arange(nat), pi]))
    need = np.r_[np.ones(nat), np.zeros(len(pc))].astype(np.int32)
    nn2, nl2 = nbl.neighlist(coords, rc, need, nat)
    np.testing.assert_array_equal(nn, nn2)
    np.testing.assert_array_equal(nl, nl2)
    off = np.r_[0, np.cumsum(nn)]
    for i, s in enumerate(nbl.brute_force_sets(coords, rc, nat)):
        np.testing.assert_array_equal(np.sort(nl[off[i]:off[i + 1]]), s)
    # the reference's own producers, through the digests of their outputs
    check_digests(reference_digests()["cells"][case], pos, sp, nat, coords, spd, image, nn, nl)


def test_full_size_two_species_layout():
    """262,144 atoms (the benchmark's system): every padding row, species,
    image, count and list entry equal, bit for bit, to the reference's own
    producers (their digests), plus the numpy restatement on the first 3000
    centres."""
    from libdescriptor_amd import workloads
    from oracle import nbl
    print("case: 262144-atom two-species system, rc 5.0", flush=True)
    pos, cell, pbc, sp = workloads.random_system(262144, 0.05, 1.8, seed=0)
    coords, spd, image, nn, nl = as_np(device_layout(pos, cell, pbc, sp, 5.0))
    n = len(pos)
    check_digests(reference_digests()["system_262144"], pos, sp, n, coords, spd, image, nn, nl)
    k = 3000
    need = np.r_[np.ones(k), np.zeros(len(coords) - k)].astype(np.int32)
    nn3, nl3 = nbl.neighlist(coords, 5.0, need, k)
    np.testing.assert_array_equal(nn[:k], nn3)
    np.testing.assert_array_equal(nl[:int(nn3.sum())], nl3)


def test_errors_collision_and_singular_cell():
    from libdescriptor_amd import _capi
    print("case: collision and singular cell", flush=True)
    pos = np.array([[0.0, 0, 0], [1e-6, 0, 0], [3.0, 0, 0]])
    with pytest.raises(_capi.LdsfError, match="Collision of atoms 1 and 2"):
        device_layout(pos, np.eye(3) * 20.0, [0, 0, 0], np.zeros(3, np.int32), 4.0)
    with pytest.raises(_capi.LdsfError, match="determinant"):
        device_layout(pos + [5.0, 0, 0], np.zeros((3, 3)), [1, 1, 1], np.zeros(3, np.int32), 4.0)


@pytest.mark.parametrize("which", ["c1", "large"])
def test_assemble_forces_bitwise(which, c1):
    """Both summation orders of neighbor.py:227-263: per-atom np.sum when
    n < n_padding (C1: 64 atoms, 327 padding rows), one add per padding row
    otherwise (a 4000-atom periodic system)."""
    from libdescriptor_amd import _capi, workloads
    from oracle import nbl
    print(f"case: assemble_forces {which}", flush=True)
    if which == "c1":
        L = device_layout(c1["positions"], c1["cell"], [1, 1, 1], np.zeros(64, np.int32), 4.5)
    else:
        pos, cell, pbc, sp = workloads.random_system(4000, 0.05, 1.8, seed=3)
        L = device_layout(pos, cell, pbc, sp, 5.0)
    rows = L.n_contrib + L.n_padding
    assert (L.n_contrib < L.n_padding) == (which == "c1")
    rng = np.random.default_rng(7)
    # magnitudes over many decades, so that a different order rounds differently
    f = rng.normal(size=(rows, 3)) * 10.0 ** rng.integers(-6, 6, size=(rows, 3))
    f[5] = -0.0  # sign of zero follows the reference's order too
    ft = torch.tensor(f, device=_dev())
    out = L.assemble_forces(ft).cpu().numpy()
    img = L.image.cpu().numpy()[L.n_contrib:]
    exp = nbl.assemble_forces(f, L.n_contrib, img)
    np.testing.assert_array_equal(out.view(np.int64), exp.view(np.int64))
    out2 = _capi.assemble_forces_device(ft, L.n_contrib,
                                        torch.tensor(img, device=_dev())).cpu().numpy()
    np.testing.assert_array_equal(out2.view(np.int64), exp.view(np.int64))


@pytest.mark.parametrize("which", ["c1", "large"])
@pytest.mark.parametrize("prec", ["fp64", "fp32_bulk"])
def test_gather_fold_bitwise(which, prec, c1):
    """plan.gather_fold(layout) equals, bitwise, the gradient into zeros followed by
    layout.assemble_forces, for both summation orders of neighbor.py:227-263
    (C1: n < n_padding; a 4000-atom periodic system: n >= n_padding)."""
    import libdescriptor_amd as ld
    from libdescriptor_amd import _capi, workloads
    from libdescriptor_amd.workloads import behler_si_params
    print(f"case: gather_fold {which} {prec}", flush=True)
    d = _dev()
    types, sizes, params = behler_si_params()
    if which == "c1":
        L = device_layout(c1["positions"], c1["cell"], [1, 1, 1], np.zeros(64, np.int32), 4.5)
        ds = ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"],
                                               "cos", 4.5, types, sizes, params)
    else:
        pos, cell, pbc, sp = workloads.random_system(4000, 0.05, 1.8, seed=3)
        L = device_layout(pos, cell, pbc, sp, 5.0)
        ds = ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, params,
                                                       species_resolved=True)
    p = _capi.FP64 if prec == "fp64" else _capi.FP32_BULK
    rows = L.n_contrib + L.n_padding
    plan = _capi.Plan(ds, rows, L.species, L.neighlist, L.numneigh)
    g = torch.tensor(np.random.default_rng(11).normal(size=(L.n_contrib, ds.width)), device=d)
    dx = torch.zeros((rows, 3), dtype=torch.float64, device=d)
    plan.gradient(L.coords, g, dx, p)
    exp = L.assemble_forces(dx).cpu().numpy()
    plan.adjoint_edges(L.coords, g, p)
    out = plan.gather_fold(L).cpu().numpy()
    np.testing.assert_array_equal(out.view(np.int64), exp.view(np.int64))
    # a second fold into a caller's buffer, over stale contents
    buf = torch.full((L.n_contrib, 3), 7.0, dtype=torch.float64, device=d)
    assert plan.gather_fold(L, out=buf) is buf
    np.testing.assert_array_equal(buf.cpu().numpy().view(np.int64), exp.view(np.int64))


def test_gather_fold_rejects_mismatch(c1):
    """A plan whose row count is not the layout's, or that has no edge records yet,
    is refused."""
    import libdescriptor_amd as ld
    from libdescriptor_amd import _capi
    from libdescriptor_amd.workloads import behler_si_params
    print("case: gather_fold mismatch", flush=True)
    L = device_layout(c1["positions"], c1["cell"], [1, 1, 1], np.zeros(64, np.int32), 4.5)
    types, sizes, params = behler_si_params()
    ds = ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"],
                                           "cos", 4.5, types, sizes, params)
    rows = L.n_contrib + L.n_padding
    plan = _capi.Plan(ds, rows, L.species, L.neighlist, L.numneigh)
    with pytest.raises(_capi.LdsfError):
        plan.gather_fold(L)  # no adjoint_edges yet
    sp1 = torch.cat([L.species, torch.zeros(1, dtype=torch.int32, device=_dev())])
    other = _capi.Plan(ds, rows + 1, sp1, L.neighlist, L.numneigh)
    g = torch.zeros((64, ds.width), dtype=torch.float64, device=_dev())
    x1 = torch.cat([L.coords, torch.zeros((1, 3), dtype=torch.float64, device=_dev())])
    other.adjoint_edges(x1, g)
    with pytest.raises(_capi.LdsfError):
        other.gather_fold(L)


def test_assemble_forces_rejects_bad_image():
    from libdescriptor_amd import _capi
    print("case: assemble_forces bad image", flush=True)
    d = _dev()
    f = torch.zeros((5, 3), dtype=torch.float64, device=d)
    with pytest.raises(_capi.LdsfError):
        _capi.assemble_forces_device(f, 3, torch.tensor([0, 3], dtype=torch.int32, device=d))


def test_device_layout_feeds_the_plan(c1):
    """Positions on the device -> layout -> plan -> descriptors and dE/dr ->
    fold, all on the GPU, equal to the golden descriptors and the oracle's
    folded gradient."""
    import libdescriptor_amd as ld
    from libdescriptor_amd.autograd import SymmetryFunctionsLayer
    from libdescriptor_amd.workloads import behler_si_params
    from oracle import nbl, orc
    print("case: C1 device layout -> plan -> fold", flush=True)
    d = _dev()
    L = device_layout(c1["positions"], c1["cell"], [1, 1, 1], np.zeros(64, np.int32), 4.5)
    types, sizes, params = behler_si_params()
    ds = ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"], "cos",
                                           4.5, types, sizes, params)
    layer = SymmetryFunctionsLayer(ds, L.species, L.neighlist, L.numneigh, device=d)
    x = L.coords.clone().requires_grad_(True)
    desc = layer(x)
    w = torch.tensor(c1["dE_dzeta"], device=d)
    (desc * w).sum().backward()
    forces = L.assemble_forces(x.grad).cpu().numpy()
    assert float(np.abs(desc.detach().cpu().numpy() - c1["desc"]).max()
                 / np.abs(c1["desc"]).max()) <= 1e-12
    O = orc.OracleSF(["Si"], 4.5, types, sizes, params)
    G = O.gradient(64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"], c1["dE_dzeta"])
    F0 = nbl.assemble_forces(G, 64, c1["image"])
    assert float(np.abs(forces - F0).max() / np.abs(F0).max()) <= 1e-12


def test_positions_module_c1(c1):
    """SymmetryFunctionsFromPositions: positions -> device layout -> descriptors, and
    the backward -> adjoint -> gather -> fold, against the golden descriptors and the
    oracle's folded gradient."""
    import libdescriptor_amd as ld
    from libdescriptor_amd.autograd import SymmetryFunctionsFromPositions
    from libdescriptor_amd.workloads import behler_si_params
    from oracle import nbl, orc
    print("case: C1 positions module", flush=True)
    d = _dev()
    types, sizes, params = behler_si_params()
    ds = ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"], "cos",
                                           4.5, types, sizes, params)
    mod = SymmetryFunctionsFromPositions(ds, np.zeros(64, np.int32), c1["cell"], [1, 1, 1], 4.5)
    pos = torch.tensor(c1["positions"], device=d, requires_grad=True)
    desc = mod(pos)
    (desc * torch.tensor(c1["dE_dzeta"], device=d)).sum().backward()
    assert float(np.abs(desc.detach().cpu().numpy() - c1["desc"]).max()
                 / np.abs(c1["desc"]).max()) <= 1e-12
    O = orc.OracleSF(["Si"], 4.5, types, sizes, params)
    G = O.gradient(64, c1["species"], c1["neighlist"], c1["numneigh"], c1["coords"], c1["dE_dzeta"])
    F0 = nbl.assemble_forces(G, 64, c1["image"])
    F = pos.grad.cpu().numpy()
    assert float(np.abs(F - F0).max() / np.abs(F0).max()) <= 1e-12


def test_positions_module_central_fd():
    """d(sum w * desc)/d(positions) by autograd against central differences of the whole
    pipeline (layout rebuilt on the GPU at every displaced position), two species, FP64."""
    import libdescriptor_amd as ld
    from libdescriptor_amd import workloads
    from libdescriptor_amd.autograd import SymmetryFunctionsFromPositions
    print("case: positions module central FD", flush=True)
    d = _dev()
    pos0, cell, pbc, sp = workloads.random_system(150, 0.05, 1.8, seed=8)
    types, sizes, params = workloads.behler_si_params()
    ds = ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, params,
                                                   species_resolved=True)
    mod = SymmetryFunctionsFromPositions(ds, sp, cell, pbc, 5.0)
    w = torch.tensor(np.random.default_rng(2).normal(size=(150, ds.width)), device=d)
    pos = torch.tensor(pos0, device=d, requires_grad=True)
    (mod(pos) * w).sum().backward()
    G = pos.grad.cpu().numpy()
    h = 1e-5
    for a, c in [(0, 0), (17, 1), (88, 2), (149, 0)]:
        vals = []
        for sgn in (1, -1):
            p = pos0.copy()
            p[a, c] += sgn * h
            with torch.no_grad():
                vals.append(float((mod(torch.tensor(p, device=d)) * w).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - G[a, c]) <= 1e-6 * max(1.0, abs(G[a, c])), (a, c, fd, G[a, c])
