"""The fused forward + adjoint (ldsf_plan_forward_gradient, and the host
ldsf_gradient, which re-adds the forward into desc as the reference's Enzyme
call does, src/Descriptors.cpp:165-190) against the oracle and against the
separate forward and adjoint kernels.

Tolerances: FP64 descriptors elementwise relative <= 1e-12 against the oracle
and <= 1e-14 against the separate forward kernel (the fused kernel forms
B^zeta as B^(zeta-1) B); FP64 gradients normwise <= 1e-12; FP32-bulk
descriptors elementwise relative <= 1e-5 on entries carrying signal
(|ref| >= 1e-6 of the block maximum), gradients normwise <= 1e-5 and, at
full size, elementwise <= 1e-5 on the components >= 10 % of the largest."""
import ctypes

import numpy as np
import pytest

import libdescriptor_amd as ld
from conftest import behler, elem_rel_err, rel_err
from oracle import orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TOL64, TOL32 = 1e-12, 1e-5


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def signal_rel(a, ref, blocks):
    worst = 0.0
    for lo, hi in blocks:
        r = ref[:, lo:hi]
        scale = np.abs(r).max()
        if scale == 0:
            continue
        sig = np.abs(r) >= 1e-6 * scale
        worst = max(worst, float((np.abs(a[:, lo:hi] - r)[sig] / np.abs(r)[sig]).max()))
    return worst


def two_species(n, seed):
    from libdescriptor_amd import workloads as w
    pos, cell, pbc, sp = w.random_system(n, 0.05, 1.8, seed=seed)
    return w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"])


class Setup:
    def __init__(self, L, species=("A", "B"), rc=5.0, resolved=True):
        d = _dev()
        types, sizes, params = behler()
        self.ds = ld.DescriptorKind.init_symmetry_functions(list(species), rc, types, sizes,
                                                            list(params), species_resolved=resolved)
        self.O = orc.OracleSF(list(species), rc, types, sizes, params, species_resolved=resolved)
        from libdescriptor_amd import _capi
        t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device=d)  # noqa: E731
        self.L = L
        self.n = L["n_contrib"]
        self.x = t(L["coords"], torch.float64)
        self.plan = _capi.Plan(self.ds, len(L["coords"]), t(L["species"]),
                               t(L["neighlist"]), t(L["numneigh"]))
        self.W = self.ds.width
        self.g = t(np.random.default_rng(5).normal(size=(self.n, self.W)), torch.float64)

    def split(self, prec):
        desc = torch.zeros((self.n, self.W), dtype=torch.float64, device=self.x.device)
        dx = torch.zeros_like(self.x)
        self.plan.compute(self.x, desc, prec)
        self.plan.gradient(self.x, self.g, dx, prec)
        torch.cuda.synchronize()
        return desc.cpu().numpy(), dx.cpu().numpy()

    def fused(self, prec, desc0=None):
        desc = torch.zeros((self.n, self.W), dtype=torch.float64, device=self.x.device) \
            if desc0 is None else torch.tensor(desc0, device=self.x.device)
        dx = torch.zeros_like(self.x)
        self.plan.forward_gradient(self.x, self.g, desc, dx, prec)
        torch.cuda.synchronize()
        return desc.cpu().numpy(), dx.cpu().numpy()

    def oracle(self):
        L = self.L
        d0 = self.O.compute(self.n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
        G0 = self.O.gradient(self.n, L["species"], L["neighlist"], L["numneigh"], L["coords"],
                             self.g.cpu().numpy())
        return d0, G0


BLOCKS = [(0, 8), (8, 16), (16, 32), (32, 48), (48, 64)]


@pytest.mark.parametrize("family", [1, 0])
def test_fused_two_species_against_oracle(family):
    print(f"case: 400-atom two-species, family {family}", flush=True)
    S = Setup(two_species(400, 11))
    if family == 0:
        S.plan.kernel_family = 0
    d0, G0 = S.oracle()
    d, G = S.fused(0)
    assert elem_rel_err(d, d0, 1e-300) <= TOL64
    assert rel_err(G, G0) <= TOL64
    ds_, Gs = S.split(0)
    assert elem_rel_err(d, ds_, 1e-300) <= 1e-14
    assert rel_err(G, Gs) <= 1e-14
    d32, G32 = S.fused(1)
    assert signal_rel(d32, d0, BLOCKS) <= TOL32
    assert rel_err(G32, G0) <= TOL32
    # deterministic
    d2, G2 = S.fused(1)
    np.testing.assert_array_equal(d2, d32)
    np.testing.assert_array_equal(G2, G32)


def test_neighbourhoods_of_up_to_64():
    """A dense cluster whose largest neighbourhood has 63 entries (49 centres
    over 32, each a wave of its own, beside paired ones): the wave kernels'
    largest pair enumerations (up to 1,953 candidates, 31 rotational rows),
    fused and separate, both precisions, against the oracle."""
    from test_oracle import cluster, full_list
    print("case: neighbourhoods of up to 64 entries", flush=True)
    rc = 4.8
    pos = cluster(110, 4, box=9.0, dmin=1.2)
    nn, nl = full_list(pos, rc)
    assert 57 <= nn.max() <= 64 and (nn > 32).sum() >= 10 and (nn <= 32).sum() >= 10
    L = dict(n_contrib=len(pos), coords=pos, species=np.zeros(len(pos), np.int32), neighlist=nl,
             numneigh=nn)
    S = Setup(L, species=("Si",), rc=rc, resolved=False)
    assert S.plan.kernel_family == 1
    d0, G0 = S.oracle()
    d, G = S.fused(0)
    assert elem_rel_err(d, d0, 1e-300) <= TOL64
    assert rel_err(G, G0) <= TOL64
    ds_, Gs = S.split(0)
    assert elem_rel_err(ds_, d0, 1e-300) <= TOL64
    assert rel_err(Gs, G0) <= TOL64
    d32, G32 = S.fused(1)
    assert signal_rel(d32, d0, [(0, 8), (8, 24)]) <= TOL32
    assert rel_err(G32, G0) <= TOL32
    ds32, Gs32 = S.split(1)
    assert signal_rel(ds32, d0, [(0, 8), (8, 24)]) <= TOL32
    assert rel_err(Gs32, G0) <= TOL32


@pytest.mark.parametrize("resolved,rc,family", [
    (True, [[5.0, 4.6], [4.6, 4.2]], 1),    # three pair blocks, one cutoff each
    (False, [[5.0, 4.4], [4.7, 4.1]], 1),   # one block, the cutoff looked up per species pair
    (True, [[5.0, 4.4], [4.7, 4.1]], 0),    # resolved, rc(A, B) != rc(B, A): general kernels
])
def test_species_pair_cutoffs(resolved, rc, family):
    """Cutoff matrices that differ per species pair (the reference's rcut_2D_,
    src/SymmetryFunctions.cpp:150-157,203-207; asymmetric ones only through
    the vector init), fused and separate, both precisions, against the oracle.
    The wave kernels take the species-resolved blocks only with a symmetric
    matrix."""
    print(f"case: 400-atom two-species, cutoff matrix {rc}, resolved {resolved}", flush=True)
    rcm = np.array(rc)
    S = Setup(two_species(400, 12), rc=rcm.reshape(-1), resolved=resolved)
    assert S.plan.kernel_family == family
    d0, G0 = S.oracle()
    blocks = BLOCKS if resolved else [(0, 8), (8, 24)]
    for fused in (True, False):
        d, G = S.fused(0) if fused else S.split(0)
        assert elem_rel_err(d, d0, 1e-300) <= TOL64
        assert rel_err(G, G0) <= TOL64
        d32, G32 = S.fused(1) if fused else S.split(1)
        assert signal_rel(d32, d0, blocks) <= TOL32
        assert rel_err(G32, G0) <= TOL32


def test_fused_accumulates_into_desc(c1):
    """desc += descriptors (the reference's += semantics), single species."""
    print("case: C1 fused +=", flush=True)
    L = {k: c1[k] for k in ("coords", "species", "neighlist", "numneigh")}
    L["n_contrib"] = 64
    S = Setup(L, species=("Si",), rc=4.5, resolved=False)
    base = np.random.default_rng(2).normal(size=(64, 24))
    d, G = S.fused(0, base)
    assert elem_rel_err(d - base, c1["desc"], 1e-300) <= 1e-11
    d0, G0 = S.oracle()
    assert rel_err(G, G0) <= TOL64


def test_centre_same_in_a_shared_wave_and_alone():
    """The fused kernel runs two consecutive centres of at most 32 neighbours
    in one wave (a shared pair phase); a centre's descriptors and dE/dr are
    bitwise those of the same centre run alone.  A 40,000-atom system (more
    centres than the plan's 16,384 blocks, so runs hold pairs); dE/dzeta is
    zero except on the centre, so the whole-system dE/dr holds that centre's
    records only, summed with exact zeros; a plan over the one centre (centre
    list [c]) runs it alone."""
    from libdescriptor_amd import _capi
    print("case: 40,000-atom two-species, paired vs alone", flush=True)
    S = Setup(two_species(40000, seed=21))
    L, n, W, d = S.L, S.n, S.W, S.x.device
    nn = np.asarray(L["numneigh"])
    off = np.r_[0, np.cumsum(nn)]
    t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device=d)  # noqa: E731
    sp = t(L["species"])
    # centres whose neighbour and its own lists fit 32 lanes (most are), the
    # first and the last
    small = [c for c in range(1, n - 1) if nn[c] <= 32 and nn[c - 1] <= 32 and nn[c + 1] <= 32]
    cases = [0, small[0], small[len(small) // 2], small[-1], n - 1]
    for prec in (0, 1):
        for c in cases:
            g = torch.zeros((n, W), dtype=torch.float64, device=d)
            g[c] = S.g[c]
            desc = torch.zeros((n, W), dtype=torch.float64, device=d)
            dx = torch.zeros_like(S.x)
            S.plan.forward_gradient(S.x, g, desc, dx, prec)
            one = _capi.Plan(S.ds, len(L["coords"]), sp, t(L["neighlist"][off[c]:off[c + 1]]),
                             t(nn[c:c + 1]), centers=t([c]))
            d1 = torch.zeros((1, W), dtype=torch.float64, device=d)
            dx1 = torch.zeros_like(S.x)
            one.forward_gradient(S.x, g[c:c + 1].contiguous(), d1, dx1, prec)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(d1.cpu().numpy()[0], desc.cpu().numpy()[c])
            np.testing.assert_array_equal(dx1.cpu().numpy(), dx.cpu().numpy())
            assert np.abs(dx1.cpu().numpy()).max() > 0


@pytest.mark.parametrize("family", [1, 0])
@pytest.mark.parametrize("prec", [0, 1])
def test_forward_gradient_assign(family, prec):
    """ldsf_plan_forward_gradient_assign over stale outputs (NaN) equals,
    bitwise, ldsf_plan_forward_gradient into zeroed outputs; both kernel
    families (the wave kernels assign in the kernel, the general ones clear
    the rows first), both precisions, a 1,000-atom two-species system with
    padding rows that no edge points at."""
    print(f"case: forward_gradient_assign family {family} precision {prec}", flush=True)
    S = Setup(two_species(1000, seed=14))
    S.plan.kernel_family = family
    d0, g0 = S.fused(prec)
    desc = torch.full((S.n, S.W), float("nan"), dtype=torch.float64, device=S.x.device)
    dx = torch.full_like(S.x, float("nan"))
    for _ in range(2):  # twice: the second call over the first one's results
        S.plan.forward_gradient_assign(S.x, S.g, desc, dx, prec)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(desc.cpu().numpy().view(np.int64), d0.view(np.int64))
        np.testing.assert_array_equal(dx.cpu().numpy().view(np.int64), g0.view(np.int64))


def test_host_gradient_readds_forward(c1):
    """ldsf_gradient with a desc buffer: desc += descriptors, d_coords += J^T g."""
    from libdescriptor_amd import _capi
    print("case: C1 host ldsf_gradient with desc", flush=True)
    _dev()
    ds = ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"], "cos",
                                           4.5, *behler())
    sp = np.ascontiguousarray(c1["species"], np.int32)
    nl = np.ascontiguousarray(c1["neighlist"], np.int32)
    nn = np.ascontiguousarray(c1["numneigh"], np.int32)
    x = np.ascontiguousarray(c1["coords"])
    g = np.ascontiguousarray(c1["dE_dzeta"])
    desc = np.ones((64, 24))
    dx = np.zeros_like(x)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    _capi.check(_capi.lib().ldsf_gradient(ctypes.c_void_p(ds._handle), 64, p(sp), p(nl), p(nn),
                                          p(x), p(dx), p(desc), p(g)))
    assert elem_rel_err(desc - 1.0, c1["desc"], 1e-300) <= 1e-11
    G0 = orc.OracleSF(["Si"], 4.5, *behler()).gradient(64, sp, nl, nn, x, g)
    assert rel_err(dx, G0) <= TOL64


def test_host_compute_overwrite_and_accumulate(c1):
    """ldsf_compute adds into desc (the reference's +=); ldsf_compute_overwrite
    (and the batch form) overwrites it, whatever it held; the Python module's
    compute, which uses the overwrite form, equals both from zero."""
    from libdescriptor_amd import _capi
    print("case: C1 host compute += and overwrite", flush=True)
    _dev()
    ds = ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"], "cos",
                                           4.5, *behler())
    sp = np.ascontiguousarray(c1["species"], np.int32)
    nl = np.ascontiguousarray(c1["neighlist"], np.int32)
    nn = np.ascontiguousarray(c1["numneigh"], np.int32)
    x = np.ascontiguousarray(c1["coords"])
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    h = ctypes.c_void_p(ds._handle)
    base = np.random.default_rng(3).normal(size=(64, 24))
    acc = base.copy()
    _capi.check(_capi.lib().ldsf_compute(h, 64, p(sp), p(nl), p(nn), p(x), p(acc)))
    over = base.copy()
    _capi.check(_capi.lib().ldsf_compute_overwrite(h, 64, p(sp), p(nl), p(nn), p(x), p(over)))
    bover = base.copy()
    na, cp = np.array([64], np.int32), np.array([0], np.int32)
    _capi.check(_capi.lib().ldsf_compute_batch_overwrite(h, 1, p(na), p(cp), p(sp), p(nl), p(nn),
                                                         p(x), p(bover)))
    mod = ld.compute(ds, 64, sp, nl, nn, x)
    assert elem_rel_err(over, c1["desc"], 1e-300) <= TOL64
    np.testing.assert_array_equal(bover, over)
    np.testing.assert_array_equal(mod, over)
    assert elem_rel_err(acc - base, over, 1e-300) <= 1e-11


def test_caller_record_buffer():
    """ldsf_plan_set_records: the adjoint writes into the caller's buffer and the
    results equal the plan's own; a buffer too small is refused."""
    print("case: 400-atom two-species, caller record buffer", flush=True)
    S = Setup(two_species(400, 13))
    d0, G0 = S.fused(1)
    with pytest.raises(Exception):
        S.plan.set_records(torch.empty(3 * S.plan.num_records - 1, dtype=torch.float64,
                                       device=S.x.device))
    rec = torch.full((3 * S.plan.num_records,), float("nan"), dtype=torch.float64,
                     device=S.x.device)
    S.plan.set_records(rec)
    d1, G1 = S.fused(1)
    np.testing.assert_array_equal(d1, d0)
    np.testing.assert_array_equal(G1, G0)
    assert not torch.isnan(rec).any()


def test_fused_full_size_262144():
    """The benchmark's system: fused against the separate kernels, both forms."""
    from libdescriptor_amd import workloads as w
    print("case: 262144-atom two-species fused vs split", flush=True)
    S = Setup(w.config_c3(n=262144, seed=0))
    d64, G64 = S.split(0)
    df, Gf = S.fused(0)
    assert elem_rel_err(df, d64, 1e-300) <= 1e-13
    assert rel_err(Gf, G64) <= 1e-13
    # the FP64 form against the oracle at full size (its OpenMP forward + adjoint)
    L = S.L
    offs = np.r_[0, np.cumsum(L["numneigh"])].astype(np.int32)
    d0, G0 = S.O.fwd_adj_omp(S.n, L["species"], L["neighlist"], offs, L["coords"],
                             S.g.cpu().numpy(), 0)
    assert elem_rel_err(df, d0, 1e-300) <= TOL64
    assert rel_err(Gf, G0) <= TOL64
    d32, G32 = S.fused(1)
    assert signal_rel(d32, d64, BLOCKS) <= TOL32
    assert rel_err(G32, G64) <= TOL32
    # dE/dr elementwise on the components at least 10 % of the largest (a
    # component that cancels to far below the largest forces carries the
    # absolute error of its FP32 terms; the floors below are reported, not bounded)
    from libdescriptor_amd import errstat
    ge = errstat.gradient_errors(G32, G64)
    print(f"FP32-bulk dE/dr vs FP64: {ge}", flush=True)
    assert errstat.elem_at(ge, 0.1) <= TOL32
    # size-independent properties at full size, both forms: translation
    # invariance (every row, padding included: sum of dE/dr = 0), and
    # linearity in dE/dzeta, exact for a power-of-two scale
    for prec, G in ((0, Gf), (1, G32)):
        assert np.abs(G.sum(axis=0)).max() <= 1e-12 * np.abs(G).sum(axis=0).max()
        g1 = S.g
        S.g = 2.0 * g1
        _, G2 = S.fused(prec)
        S.g = g1
        np.testing.assert_array_equal(G2, 2.0 * G)
