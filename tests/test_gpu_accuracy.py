"""FP32-bulk accuracy beyond the benchmark's parameter grid, against the FP64
oracle (sym_g2 takes any eta and Rs, src/SymmetryFunctions.cpp:133-135; sym_g4
any zeta, lambda, eta, :101-127).

The wave kernels form each FP32-bulk term's exponent x in FP32 (relative
error ~3e-7 |x|); a parameter set whose exponents can exceed 32 over the
cutoff sphere (eta (r - Rs)^2 for g2 rows, eta 3 rc^2 for g4 rows) runs its
FP32-bulk calls on the general kernels instead, which split the exponent in
FP64.  Checks, on a 3,000-atom two-species system (rc = 5 A):

  * g2 rows with Rs != 0 and eta up to 1 (exponents <= 16: the wave kernels):
    FP32-bulk descriptors <= 1e-5 relative on the entries that carry signal
    and dE/dr <= 1e-5 normwise and elementwise on the components >= 10 % of
    the largest (libdescriptor_amd.errstat); the FP64 form <= 1e-12 of the
    oracle;
  * eta = 4 A^-2 g2 rows (exponents up to 100, past FP32's underflow at 87) and
    eta = 0.5 g4 rows: the FP32-bulk call equals, bitwise, the same call with
    the general kernels forced, and meets the same bounds.
"""
import numpy as np
import pytest

from conftest import elem_rel_err, rel_err

pytestmark = pytest.mark.gpu

G4_ZETA = (1.0, 2.0, 4.0, 16.0)


def params(case):
    if case == "rs_rows":
        g2 = [(eta, rs) for eta in (0.5, 1.0) for rs in (1.0, 2.0, 3.0, 4.0)]
        g4_eta = (0.001, 0.01)
    else:  # large_eta
        g2 = [(eta, rs) for eta in (4.0, 2.0) for rs in (0.0, 1.5, 3.0, 4.5)]
        g4_eta = (0.001, 0.5)
    g4 = [(z, lam, e) for z in G4_ZETA for lam in (1.0, -1.0) for e in g4_eta]
    flat = [v for row in g2 for v in row] + [v for row in g4 for v in row]
    return ["g2", "g4"], [len(g2), 2, len(g4), 3], flat


@pytest.mark.parametrize("case", ["rs_rows", "large_eta"])
def test_fp32_bulk_against_oracle(case):
    import torch
    import libdescriptor_amd as ld
    from libdescriptor_amd import _capi, errstat
    from libdescriptor_amd import workloads as w
    from oracle import orc
    print(f"case: FP32-bulk accuracy {case}", flush=True)
    types, sizes, flat = params(case)
    pos, cell, pbc, sp = w.random_system(3000, 0.05, 1.8, seed=5)
    L = w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"])
    n = L["n_contrib"]
    ds = ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, flat,
                                                   species_resolved=True)
    O = orc.OracleSF(["A", "B"], 5.0, types, sizes, flat, species_resolved=True)
    g = np.random.default_rng(6).normal(size=(n, ds.width))
    d0 = O.compute(n, L["species"], L["neighlist"], L["numneigh"], L["coords"])
    g0 = O.gradient(n, L["species"], L["neighlist"], L["numneigh"], L["coords"], g)
    dv = torch.device("cuda:0")
    t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device=dv)  # noqa: E731
    sp_d, nl_d, nn_d, x = t(L["species"]), t(L["neighlist"]), t(L["numneigh"]), \
        t(L["coords"], torch.float64)
    gd = t(g, torch.float64)

    def run(prec, family=None):
        plan = _capi.Plan(ds, len(L["coords"]), sp_d, nl_d, nn_d)
        if family is not None:
            plan.kernel_family = family
        assert plan.kernel_family == (1 if family is None else family)
        d = torch.zeros((n, ds.width), dtype=torch.float64, device=dv)
        dx = torch.zeros_like(x)
        plan.forward_gradient(x, gd, d, dx, prec)
        torch.cuda.synchronize()
        return d.cpu().numpy(), dx.cpu().numpy()

    d64, g64 = run(_capi.FP64)
    assert elem_rel_err(d64, d0, 1e-300) <= 1e-12
    assert rel_err(g64, g0) <= 1e-12
    d32, g32 = run(_capi.FP32_BULK)
    d32g, g32g = run(_capi.FP32_BULK, family=0)
    if case == "large_eta":  # routed to the general kernels' split exponent
        assert np.array_equal(d32, d32g) and np.array_equal(g32, g32g)
    else:  # the wave kernels ran (their arithmetic differs from the general kernels')
        assert not np.array_equal(d32, d32g)
    blocks = {"g2_A": (0, 8), "g2_B": (8, 16), "g4_AA": (16, 32), "g4_AB": (32, 48),
              "g4_BB": (48, 64)}
    for name, (d, gx) in (("auto", (d32, g32)), ("general", (d32g, g32g))):
        de = errstat.descriptor_errors(d, d0, blocks)
        ge = errstat.gradient_errors(gx, g0)
        print(f"{case} {name}: descriptors max_rel {errstat.worst(de, 'max_rel'):.2e}, "
              f"dE/dr {ge}", flush=True)
        assert errstat.worst(de, "max_rel") <= 1e-5
        assert ge["max_rel_normwise"] <= 1e-5
        assert errstat.elem_at(ge, 0.1) <= 1e-5
