"""The serving loop (ldsf_server): continuous batching of descriptor + dE/dr
requests over device-resident buffers into graph-replayed lanes.

Checks:
  every result of a request served in a mix equals the same request run alone
  through a device plan (ldsf_plan_forward_gradient) -- bitwise (each centre is computed on its own and
  each row's dE/dr records are summed in ascending edge order in both paths),
  in both precision forms;
  a request with an out-of-range neighbour index is rejected (LDSF_ERR_NEIGHBOR)
  without disturbing the requests batched with it;
  requests submitted from several threads while batches run are all served;
  admission limits: a request beyond the lane capacity and a budget below one
  lane are errors; graphs are replayed, not re-captured, across batches;
  the serving workload at its own scale: 200 requests of 16-20,000 atoms
  (at least three above 10,000) drawn like the benchmark's request pool, plus
  a dense cluster whose neighbourhoods exceed 64 atoms (the general-kernel
  bucket), submitted on a seeded Poisson schedule to 3 lanes in both
  precision forms -- every result bitwise equal to the same request alone,
  and a sample (the large requests, the dense one and a few others) checked
  against the FP64 oracle's compute + gradient (src/Descriptors.cpp:263-292):
  FP64 descriptors <= 1e-12 elementwise and dE/dr <= 1e-12 normwise,
  FP32-bulk <= 1e-5 in the error statement of libdescriptor_amd.errstat
  (descriptors per block on the entries that carry signal, dE/dr normwise
  and elementwise on the components >= 10 % of the largest).
"""
import threading

import numpy as np
import pytest

from conftest import behler, elem_rel_err, rel_err

pytestmark = pytest.mark.gpu


def make_requests(n_req, seed, nmax=1500):
    from libdescriptor_amd import workloads as w
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_req):
        n = int(np.exp(rng.uniform(np.log(16), np.log(nmax))))
        s = int(rng.integers(1 << 30))
        if k % 3 == 2:
            pos, cell, pbc, sp = w.cluster(n, 0.05, 1.8, seed=s)
        else:
            pos, cell, pbc, sp = w.random_system(n, 0.05, 1.8, seed=s)
        out.append(w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"]))
    return out


def to_dev(L, width, seed):
    import torch
    t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device="cuda")  # noqa
    n = L["n_contrib"]
    g = np.random.default_rng(seed).normal(size=(n, width))
    return dict(species=t(L["species"]), numneigh=t(L["numneigh"]), nl=t(L["neighlist"]),
                x=t(L["coords"], torch.float64), dE=t(g, torch.float64),
                max_nn=int(L["numneigh"].max()) if n else 0)


def alone(ds, prec, D):
    import torch
    from libdescriptor_amd import _capi
    plan = _capi.Plan(ds, D["x"].shape[0], D["species"], D["nl"], D["numneigh"])
    n = D["numneigh"].numel()
    d = torch.zeros((n, ds.width), dtype=torch.float64, device="cuda")
    g = torch.zeros_like(D["x"])
    # descriptors and dE/dr from one pass, as the server computes them
    plan.forward_gradient(D["x"], D["dE"], d, g, prec)
    torch.cuda.synchronize()
    return d, g


def sf(prec):
    import libdescriptor_amd as ld
    types, sizes, params = behler()
    return ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, params,
                                                     species_resolved=True, precision=prec)


def submit(srv, ds, D):
    import torch
    n = D["numneigh"].numel()
    D["desc"] = torch.full((n, ds.width), float("nan"), dtype=torch.float64, device="cuda")
    D["dx"] = torch.full_like(D["x"], float("nan"))
    torch.cuda.synchronize()  # the fills above run on torch's stream, the server on its own
    return srv.submit(D["species"], D["numneigh"], D["nl"], D["x"], D["dE"], D["desc"], D["dx"],
                      D["max_nn"])


@pytest.mark.parametrize("prec", [0, 1])
def test_mix_equals_alone(prec):
    import torch
    from libdescriptor_amd import _capi
    ds = sf(prec)
    reqs = [to_dev(L, ds.width, 100 + k) for k, L in enumerate(make_requests(40, seed=prec))]
    # small lanes: several batches, several lanes busy at once
    srv = _capi.Server(ds, prec, hbm_budget_bytes=1 << 30, lane_centers=4000,
                       lane_rows=16000, lane_edges=4000 * 40, max_requests=16, max_lanes=3)
    assert srv.num_lanes == 3
    tickets = [submit(srv, ds, D) for D in reqs]
    lat = [srv.wait(t, timeout=60.0) for t in tickets]
    assert all(v > 0 for v in lat)
    st = srv.stats()
    print(f"stats {st}", flush=True)
    assert st["requests"] == len(reqs) and st["batches"] >= 3
    for D in reqs:
        d, g = alone(ds, prec, D)
        assert torch.equal(D["desc"], d)
        assert torch.equal(D["dx"], g)
    # a second round replays the captured graphs (one per lane and bucket at most)
    tickets = [submit(srv, ds, D) for D in reqs[:10]]
    for t in tickets:
        srv.wait(t, timeout=60.0)
    st = srv.stats()
    assert st["graphs"] <= 3 * 9 and st["batches"] > st["graphs"]
    srv.close()


def test_bad_request_is_rejected_alone():
    import torch
    from libdescriptor_amd import _capi
    ds = sf(1)
    reqs = [to_dev(L, ds.width, 200 + k) for k, L in enumerate(make_requests(6, seed=5, nmax=300))]
    bad = reqs[2]
    nl = bad["nl"].clone()
    nl[3] = bad["x"].shape[0] + 5  # one past the request's rows
    bad["nl"] = nl
    srv = _capi.Server(ds, 1, hbm_budget_bytes=1 << 30, lane_centers=4000,
                       lane_rows=16000, lane_edges=4000 * 40, max_requests=16, max_lanes=1)
    tickets = [submit(srv, ds, D) for D in reqs]
    for k, t in enumerate(tickets):
        if k == 2:
            with pytest.raises(_capi.LdsfError) as ei:
                srv.wait(t, timeout=60.0)
            assert ei.value.code == _capi.LDSF_ERR_NEIGHBOR
        else:
            srv.wait(t, timeout=60.0)
            d, g = alone(ds, 1, reqs[k])
            assert torch.equal(reqs[k]["desc"], d)
            assert torch.equal(reqs[k]["dx"], g)
    # a neighbour count above the declared bound is rejected as well
    D = dict(reqs[0])
    n, ne = D["numneigh"].numel(), D["nl"].numel()
    if (D["max_nn"] - 1) * n >= ne:
        D["max_nn"] -= 1
        t = submit(srv, ds, D)
        with pytest.raises(_capi.LdsfError) as ei:
            srv.wait(t, timeout=60.0)
        assert ei.value.code == _capi.LDSF_ERR_NEIGHBOR
    srv.close()


def test_threads_submit_while_batches_run():
    import time

    import torch
    from libdescriptor_amd import _capi
    ds = sf(1)
    reqs = [to_dev(L, ds.width, 300 + k) for k, L in enumerate(make_requests(24, seed=9))]
    srv = _capi.Server(ds, 1, hbm_budget_bytes=1 << 30, lane_centers=3000,
                       lane_rows=12000, lane_edges=3000 * 40, max_requests=8, max_lanes=2)
    errors = []

    def client(part, delay):
        try:
            for D in part:
                t = submit(srv, ds, D)
                time.sleep(delay)
                srv.wait(t, timeout=60.0)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    th = [threading.Thread(target=client, args=(reqs[k::3], 0.001 * k)) for k in range(3)]
    for x in th:
        x.start()
    for x in th:
        x.join(120)
    assert not errors, errors
    for D in reqs:
        d, g = alone(ds, 1, D)
        assert torch.equal(D["desc"], d)
        assert torch.equal(D["dx"], g)
    srv.close()


def test_admission_limits():
    from libdescriptor_amd import _capi
    from libdescriptor_amd import workloads as w
    ds = sf(1)
    with pytest.raises(_capi.LdsfError):
        _capi.Server(ds, 1, hbm_budget_bytes=1 << 20)  # below one lane
    srv = _capi.Server(ds, 1, hbm_budget_bytes=1 << 30, lane_centers=100, lane_rows=400,
                       lane_edges=4000, max_requests=4, max_lanes=1)
    pos, cell, pbc, sp = w.cluster(150, 0.05, 1.8, seed=4)
    D = to_dev(w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"]), ds.width, 1)
    with pytest.raises(_capi.LdsfError):  # 150 centres > 100 per lane
        submit(srv, ds, D)
    srv.close()


def c5_stream(seed, n_req=200):
    """Requests drawn from a pool generated like the benchmark's serving pool
    (log-uniform 16..20,000 atoms, random periodic, 0.05 atoms/A^3), with the
    pool's three largest configurations (each above 10,000 atoms) always in,
    and one dense cluster (max_nn > 64)."""
    from libdescriptor_amd import workloads as w
    pool = w.config_c5_pool(n_pool=64, seed=seed)
    sizes = np.array([L["n_contrib"] for L in pool])
    big = list(np.argsort(sizes)[-3:])
    assert all(sizes[k] > 10000 for k in big)
    pos, cell, pbc, sp = w.cluster(400, 0.15, 1.2, seed=seed + 1)
    dense = w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"])
    assert int(dense["numneigh"].max()) > 64
    pool.append(dense)
    rng = np.random.default_rng(seed)
    picks = big + [len(pool) - 1] + list(rng.integers(0, len(pool) - 1, n_req - 4))
    rng.shuffle(picks)
    return pool, picks


@pytest.mark.parametrize("prec", [0, 1])
def test_c5_stream_against_oracle(prec):
    import time

    import torch
    from libdescriptor_amd import _capi, errstat
    from oracle import orc
    ds = sf(prec)
    pool, picks = c5_stream(seed=40 + prec)
    print(f"case: serving stream prec={prec}: {len(picks)} requests, "
          f"{sum(pool[k]['n_contrib'] for k in picks)} atoms", flush=True)
    devs = {}
    for k in set(picks):
        D = to_dev(pool[k], ds.width, 0)
        devs[k] = D
    rng = np.random.default_rng(prec)
    reqs = []
    for q, k in enumerate(picks):
        D = dict(devs[k])
        g = rng.normal(size=(pool[k]["n_contrib"], ds.width))
        D["dE"] = torch.tensor(g, dtype=torch.float64, device="cuda")
        reqs.append(D)
    max_rows = max(int(D["x"].shape[0]) for D in reqs)
    srv = _capi.Server(ds, prec, hbm_budget_bytes=3 << 30, lane_centers=32768,
                       lane_rows=max(65536, max_rows), lane_edges=32768 * 48, max_requests=64,
                       max_lanes=3, batch_window_us=200)
    assert srv.num_lanes == 3
    # seeded Poisson arrivals (mean gap 0.2 ms): requests arrive while others run
    gaps = np.random.default_rng(7).exponential(2e-4, len(reqs))
    tickets = []
    for D, gap in zip(reqs, gaps):
        tickets.append(submit(srv, ds, D))
        time.sleep(float(gap))
    for t in tickets:
        srv.wait(t, timeout=120.0)
    st = srv.stats()
    print(f"stats {st}", flush=True)
    assert st["requests"] == len(reqs) and st["batches"] >= 3
    for q, D in enumerate(reqs):
        d, g = alone(ds, prec, D)
        assert torch.equal(D["desc"], d), q
        assert torch.equal(D["dx"], g), q
    srv.close()
    # the oracle on a sample: the three large requests, the dense one, three more
    types, sizes, params = behler()
    O = orc.OracleSF(["A", "B"], 5.0, types, sizes, params, species_resolved=True)
    big = sorted(range(len(picks)), key=lambda q: -pool[picks[q]]["n_contrib"])[:3]
    dense = [q for q in range(len(picks)) if picks[q] == len(pool) - 1][:1]
    sample = big + dense + [q for q in range(len(picks)) if q not in big + dense][:3]
    for q in sample:
        L, D = pool[picks[q]], reqs[q]
        n = L["n_contrib"]
        offs = np.concatenate([[0], np.cumsum(L["numneigh"])]).astype(np.int32)
        g = D["dE"].cpu().numpy()
        d0, g0 = O.fwd_adj_omp(n, L["species"], L["neighlist"], offs, L["coords"], g, 16)
        d, gx = D["desc"].cpu().numpy(), D["dx"].cpu().numpy()
        if prec == 0:
            assert elem_rel_err(d, d0, 1e-300) <= 1e-12, q
            assert rel_err(gx, g0) <= 1e-12, q
        else:
            de = errstat.descriptor_errors(d, d0, errstat.BLOCKS_TWO_SPECIES)
            ge = errstat.gradient_errors(gx, g0)
            print(f"request {q} ({n} atoms): descriptors max_rel "
                  f"{errstat.worst(de, 'max_rel'):.2e}, dE/dr {ge}", flush=True)
            assert errstat.worst(de, "max_rel") <= 1e-5, q
            assert ge["max_rel_normwise"] <= 1e-5, q
            assert errstat.elem_at(ge, 0.1) <= 1e-5, q
