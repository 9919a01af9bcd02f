"""Benchmark of the symmetry-function forward + adjoint on one MI355X.

Workload: 262,144-atom two-species periodic system (uniform random sequential
addition at 0.05 atoms/A^3, 1.8 A minimum distance, species 80/20, seed 0),
cutoff 5.0 A, species-resolved G2 (2 x 8) + G4 (3 x 16) = 64 entries per atom.
A step is one forward (descriptors) plus one adjoint (dE/dr for a fixed
N(0,1) dE/dzeta) over every atom, inputs resident in HBM, neighbour-list
plan built beforehand.  The headline value is the FP32-bulk form; the plain
FP64 form is timed beside it, with the FP32-bulk error against it.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X published vector peaks, TFLOP/s
HBM_PEAK = 8000.0  # GB/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def triplet_counts(L, rc, device):
    """Exact per-centre counts on the GPU: valid neighbours N (r < rc),
    triplet tests N(N-1)/2 and active triplets (r_jk < rc) of the input."""
    import torch
    x = torch.tensor(L["coords"], device=device)
    nn = torch.tensor(L["numneigh"], device=device, dtype=torch.int64)
    nl = torch.tensor(L["neighlist"], device=device, dtype=torch.int64)
    off = torch.zeros(len(nn) + 1, dtype=torch.int64, device=device)
    off[1:] = torch.cumsum(nn, 0)
    nmax = int(nn.max())
    N = tests = active = 0
    chunk = 8192
    for s in range(0, len(nn), chunk):
        e = min(len(nn), s + chunk)
        cnt = nn[s:e]
        k = torch.arange(nmax, device=device)
        mask = k[None, :] < cnt[:, None]
        idx = torch.where(mask, off[s:e, None] + k[None, :], torch.zeros_like(k)[None, :])
        nb = nl[idx]
        d = x[nb] - x[torch.arange(s, e, device=device)][:, None, :]
        r = torch.sqrt((d * d).sum(-1))
        valid = mask & (r < rc)
        nv = valid.sum(1)
        N += int(nv.sum())
        tests += int((nv * (nv - 1) // 2).sum())
        djk = d[:, None, :, :] - d[:, :, None, :]
        rjk = torch.sqrt((djk * djk).sum(-1))
        pair = valid[:, :, None] & valid[:, None, :] & (rjk < rc)
        active += int(torch.triu(pair, diagonal=1).sum())
    return N, tests, active


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--family", type=int, default=-1, help="force kernel family (0 general, 1 wave)")
    ap.add_argument("--only", choices=["fp64", "fp32_bulk"], default=None,
                    help="time one form only and print its timings (for profiling runs)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched (512 configurations) and serving measurements")
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--serving-lanes", type=int, default=3)
    ap.add_argument("--serving-lane-centers", type=int, default=65536)
    ap.add_argument("--serving-min-centers", type=int, default=0)
    ap.add_argument("--serving-window-us", type=int, default=1000)
    ap.add_argument("--serving-load", type=float, default=0.8,
                    help="arrival rate as a fraction of the offline request rate")
    ap.add_argument("--serving-compare", type=int, default=1,
                    help="also run the immediate-dispatch policy (0 to skip)")
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--serving-dump", default=None, metavar="NPZ",
                    help="write request sizes, arrival times and latencies to NPZ")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path "
              "shards by configuration or by spatial slab, which this build does not run",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.serving_only:
        print(json.dumps({"serving": bench_serving(args)}), flush=True)
        return

    import torch
    import libdescriptor_amd as ld
    from libdescriptor_amd import _capi, workloads as w

    dev = torch.device("cuda:0")
    t0 = time.time()
    L = w.config_c3(n=args.n, seed=0)
    n = L["n_contrib"]
    n_total = len(L["coords"])
    log(f"workload ready: {n} atoms, {n_total} rows, {len(L['neighlist'])} edges "
        f"({time.time() - t0:.1f} s)")
    types, sizes, params = w.behler_si_params()
    ds = ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, params,
                                                   species_resolved=True)
    W = ds.width
    sp = torch.tensor(L["species"], dtype=torch.int32, device=dev)
    nl = torch.tensor(L["neighlist"], dtype=torch.int32, device=dev)
    nn = torch.tensor(L["numneigh"], dtype=torch.int32, device=dev)
    x = torch.tensor(L["coords"], dtype=torch.float64, device=dev)
    rng = np.random.default_rng(1)
    g = torch.tensor(rng.normal(size=(n, W)), dtype=torch.float64, device=dev)
    plan = _capi.Plan(ds, n_total, sp, nl, nn)
    if args.family >= 0:
        plan.kernel_family = args.family
    log(f"plan ready: family {plan.kernel_family}, max neighbours {plan.max_neighbors}")
    desc = torch.zeros((n, W), dtype=torch.float64, device=dev)
    dx = torch.zeros_like(x)

    # a step: descriptors and dE/dr of every atom, from the fused kernel (the
    # reference's gradient semantics: the forward runs with the adjoint) and
    # the gather, both writing their outputs (=, bitwise the reference's +=
    # into zeroed arrays); the separate forward and adjoint kernels (+=, after
    # zeroing) are timed beside it
    def step(prec):
        plan.forward_gradient_assign(x, g, desc, dx, prec)

    def step_split(prec):
        desc.zero_()
        dx.zero_()
        plan.compute(x, desc, prec)
        plan.gradient(x, g, dx, prec)

    res = {}
    outs = {}
    forms = (("fp64", _capi.FP64), ("fp32_bulk", _capi.FP32_BULK))
    if args.only:
        forms = tuple(f for f in forms if f[0] == args.only)
    def timed(fn, prec):
        for _ in range(args.warmup):
            fn(prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn(prec)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / args.steps

    for name, prec in forms:
        ms = timed(step, prec)
        split_ms = timed(step_split, prec)
        # per-kernel times with HIP events on the stream the kernels run on:
        # fused kernel | gather, and forward kernel | adjoint kernel | gather
        s = torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        fus_ms = fwd_ms = adj_ms = gat_ms = 0.0
        reps = max(3, min(args.steps, 10))
        for _ in range(reps):
            desc.zero_()
            dx.zero_()
            ev[0].record(s)
            plan.compute(x, desc, prec)
            ev[1].record(s)
            plan.adjoint_edges(x, g, prec)
            ev[2].record(s)
            ev[3].record(s)
            plan.forward_adjoint_edges(x, g, desc, prec, assign=True)
            ev[4].record(s)
            plan.gather(dx, assign=True)
            ev[5].record(s)
            torch.cuda.synchronize()
            fwd_ms += ev[0].elapsed_time(ev[1]) / reps
            adj_ms += ev[1].elapsed_time(ev[2]) / reps
            fus_ms += ev[3].elapsed_time(ev[4]) / reps
            gat_ms += ev[4].elapsed_time(ev[5]) / reps
        outs[name] = (desc.cpu().numpy(), dx.cpu().numpy())
        res[name] = dict(ms_per_step=ms, fused_ms=fus_ms, gather_ms=gat_ms,
                         split_ms_per_step=split_ms, fwd_ms=fwd_ms, adj_ms=adj_ms)
        log(f"{name}: {ms:.3f} ms/step (fused kernel {fus_ms:.3f} ms, gather {gat_ms:.3f} ms); "
            f"separate kernels {split_ms:.3f} ms/step (forward {fwd_ms:.3f} ms, adjoint "
            f"{adj_ms:.3f} ms)")

    if args.only:
        print(json.dumps({"only": args.only, **res[args.only]}), flush=True)
        return

    # FP32-bulk error against FP64: descriptors per block, dE/dr overall and
    # with dE/dzeta restricted to one block at a time (libdescriptor_amd.errstat)
    from libdescriptor_amd import errstat
    d64, g64 = outs["fp64"]
    d32, g32 = outs["fp32_bulk"]
    err = {"descriptors": errstat.descriptor_errors(d32, d64, errstat.BLOCKS_TWO_SPECIES),
           "dE_dr": errstat.gradient_errors(g32, g64), "dE_dr_per_block": {}}
    for b, (lo, hi) in errstat.BLOCKS_TWO_SPECIES.items():
        gb = torch.zeros_like(g)
        gb[:, lo:hi] = g[:, lo:hi]
        gx = {}
        for prec in (_capi.FP64, _capi.FP32_BULK):
            desc.zero_()
            dx.zero_()
            plan.forward_gradient(x, gb, desc, dx, prec)
            gx[prec] = dx.cpu().numpy()
        err["dE_dr_per_block"][b] = errstat.gradient_errors(gx[_capi.FP32_BULK], gx[_capi.FP64])
    log("FP32-bulk vs FP64: descriptors max_rel (signal) %.2e, dE/dr normwise %.2e, "
        "elementwise on components >= 10%% / 1%% of the largest %.2e / %.2e"
        % (errstat.worst(err["descriptors"], "max_rel"), err["dE_dr"]["max_rel_normwise"],
           errstat.elem_at(err["dE_dr"], 0.1), errstat.elem_at(err["dE_dr"], 0.01)))

    # algorithmic FLOPs from exact counts of this input
    Nv, tests, active = triplet_counts(L, 5.0, dev)
    fwd_flop = Nv * (14 + 8 * 6) + tests * 9 + active * 94
    adj_flop = 2 * fwd_flop
    r32 = res["fp32_bulk"]
    value = n / (r32["ms_per_step"] * 1e-3)
    value64 = n / (res["fp64"]["ms_per_step"] * 1e-3)
    # the dominant kernel is the fused one: forward + adjoint FLOPs per launch
    dom_flop = fwd_flop + adj_flop
    dom_ms = r32["fused_ms"]
    achieved = dom_flop / (dom_ms * 1e-3) / 1e12
    dom_kernel = "k_fused_pair" if plan.kernel_family == 1 else "k_fwd_general+k_adj_general"
    traffic = measured_traffic(dom_kernel, n) if args.n == 262144 else None
    # the same kernel in the FP64 form, against the FP64 vector peak
    dom64_ms = res["fp64"]["fused_ms"]
    achieved64 = dom_flop / (dom64_ms * 1e-3) / 1e12
    traffic64 = measured_traffic(dom_kernel, n, "_fp64") if args.n == 262144 else None

    batched = serving = layout = None
    if not args.no_extra:
        layout = bench_layout(L, dev)
        batched = bench_batched(args, ds)
        serving = bench_serving(args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(L, types, sizes, params, np.asarray(g.cpu()))

    line = {
        "metric": "atom-envs/s fwd+adjoint, FP32-bulk vs FP64 GPU at stated err; p50/p99 latency",
        "value": value,
        "unit": "atom-envs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 geometry and accumulation)",
        "data": "synthetic",
        "config": {"workload": "two-species random periodic system, species-resolved SF "
                               "(2x8 G2 + 3x16 G4 = 64), forward + adjoint",
                   "n_atoms": n, "n_rows_with_padding": n_total, "cutoff_A": 5.0,
                   "density_per_A3": 0.05, "width": W, "seed": 0},
        "fp64": {"value": value64, "ms_per_step": res["fp64"]["ms_per_step"],
                 "fused_kernel_ms": res["fp64"]["fused_ms"], "gather_ms": res["fp64"]["gather_ms"],
                 "separate_kernels": {"ms_per_step": res["fp64"]["split_ms_per_step"],
                                      "forward_ms": res["fp64"]["fwd_ms"],
                                      "adjoint_ms": res["fp64"]["adj_ms"]}},
        "fp32_bulk": {"fused_kernel_ms": r32["fused_ms"], "gather_ms": r32["gather_ms"],
                      "separate_kernels": {"ms_per_step": r32["split_ms_per_step"],
                                           "forward_ms": r32["fwd_ms"],
                                           "adjoint_ms": r32["adj_ms"]}},
        "speedup_fp32_bulk_vs_fp64": value / value64,
        "error_fp32_bulk_vs_fp64": err,
        "counts": {"valid_neighbours": Nv, "triplet_tests": tests, "active_triplets": active,
                   "fwd_flop_per_step": fwd_flop, "adj_flop_per_step": adj_flop},
        "roofline": {"bound": "valu", "kernel": dom_kernel + " (fused forward + adjoint)",
                     "achieved": achieved,
                     "peak": PEAK["fp32"], "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"],
                     "traffic": traffic,
                     "hbm_GBps": (traffic / (dom_ms * 1e-3) / 1e9) if traffic else None},
        "roofline_fp64": {"bound": "valu", "kernel": dom_kernel + " (fused, FP64 form)",
                          "achieved": achieved64, "peak": PEAK["fp64"], "unit": "TFLOP/s",
                          "frac": achieved64 / PEAK["fp64"], "traffic": traffic64,
                          "hbm_GBps": (traffic64 / (dom64_ms * 1e-3) / 1e9) if traffic64 else None},
        "cpu_baseline": cpu,
        "layout": layout,
        "batched": batched,
        "serving": serving,
        "kernel_family": plan.kernel_family,
    }
    print(json.dumps(line), flush=True)


def bench_layout(L, dev):
    """The input layout of the same system built on the GPU (padding rows and
    the KLIFF-order neighbour list of the contributing atoms, ldsf_layout_create)
    against the product's host producer (C++, OpenMP).  The reference's own
    single-threaded producers are timed in the build container, not here."""
    import torch
    from libdescriptor_amd import _capi, workloads as w
    n = L["n_contrib"]
    pos = np.ascontiguousarray(L["coords"][:n])
    sp = np.ascontiguousarray(L["species"][:n])
    cell = np.eye(3) * (n / 0.05) ** (1.0 / 3.0)
    pbc = np.array([1, 1, 1], np.int32)
    xd = torch.tensor(pos, device=dev)
    sd = torch.tensor(sp, device=dev)
    times = []
    for _ in range(4):
        torch.cuda.synchronize()
        t = time.perf_counter()
        D = _capi.DeviceLayout(xd, sd, cell, pbc, 5.0)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    dev_ms = float(np.median(times[1:]) * 1e3)
    same = (np.array_equal(D.coords.cpu().numpy(), L["coords"]) and
            np.array_equal(D.numneigh.cpu().numpy(), L["numneigh"]) and
            np.array_equal(D.neighlist.cpu().numpy(), L["neighlist"]))
    t = time.perf_counter()
    w.layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"])
    host_ms = (time.perf_counter() - t) * 1e3
    out = {"rows": D.n_contrib + D.n_padding, "edges": D.n_neigh, "device_ms": dev_ms,
           "equal_to_host_producer": bool(same), "host_product_ms": host_ms,
           "host_threads": min(len(os.sched_getaffinity(0)), 16)}
    log(f"layout: device {dev_ms:.1f} ms (equal to host: {same}), host C++ {host_ms:.0f} ms")
    return out


def bench_batched(args, ds):
    """512 configurations (atom counts log-uniform in [8, 4096], half periodic
    bulk at 0.04-0.06 atoms/A^3, half clusters) as one batch plan: one forward +
    adjoint launch (the fused kernel) and the gather over every centre, both
    precision forms."""
    import torch
    from libdescriptor_amd import _capi, workloads as w
    t0 = time.time()
    B = w.concat(w.config_c4())
    log(f"batched workload ready ({time.time() - t0:.1f} s)")
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device=dev)  # noqa: E731
    n = int(B["n_atoms"].sum())
    x = t(B["coords"], torch.float64)
    g = t(np.random.default_rng(2).normal(size=(n, ds.width)), torch.float64)
    plan = _capi.Plan(ds, len(B["coords"]), t(B["species"]), t(B["neighlist"]),
                      t(B["numneigh"]), n_atoms=B["n_atoms"], config_ptr=B["config_ptr"])
    desc = torch.zeros((n, ds.width), dtype=torch.float64, device=dev)
    dx = torch.zeros_like(x)
    out = {"configurations": int(len(B["n_atoms"])), "atoms": n, "rows": int(len(B["coords"])),
           "kernel_family": plan.kernel_family}
    for name, prec in (("fp64", _capi.FP64), ("fp32_bulk", _capi.FP32_BULK)):
        def step():
            plan.forward_gradient_assign(x, g, desc, dx, prec)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t1) * 1e3 / args.steps
        out[name] = {"ms_per_step": ms, "atom_envs_per_s": n / (ms * 1e-3)}
        log(f"batched {name}: {ms:.3f} ms/step, {n / (ms * 1e-3):.3e} atom-envs/s")
    return out


def bench_serving(args):
    """Poisson request stream through the serving loop (ldsf_server): each
    request one configuration like the 262k-atom system's (random periodic,
    0.05 atoms/A^3, two species, species-resolved 64-entry set), atom count
    log-uniform in [16, 20000], asking descriptors + dE/dr for a random
    dE/dzeta, FP32-bulk.  Inputs and outputs are device-resident; latency is
    submission to completion.  Offline throughput: every request submitted at
    once.  Online: Poisson arrivals at --serving-load (default 80 %) of the
    offline request rate of the same server configuration.  The headline
    policy forms batches over a 1 ms window on 3 lanes; the immediate-dispatch
    policy (4 lanes, no window) is measured beside it."""
    import threading

    import torch
    import libdescriptor_amd as ld
    from libdescriptor_amd import _capi, workloads as w
    types, sizes, params = w.behler_si_params()
    ds = ld.DescriptorKind.init_symmetry_functions(["A", "B"], 5.0, types, sizes, params,
                                                   species_resolved=True, precision=1)
    t0 = time.time()
    pool = w.config_c5_pool()
    log(f"serving pool ready: {len(pool)} configurations ({time.time() - t0:.1f} s)")
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.int32: torch.tensor(np.asarray(a), dtype=dt, device=dev)  # noqa: E731
    rng = np.random.default_rng(5)
    P = []
    for k, L in enumerate(pool):
        n = L["n_contrib"]
        P.append(dict(species=t(L["species"]), numneigh=t(L["numneigh"]), nl=t(L["neighlist"]),
                      x=t(L["coords"], torch.float64),
                      dE=t(rng.normal(size=(n, ds.width)), torch.float64),
                      max_nn=int(L["numneigh"].max()), n=n))
    nreq = args.serving_requests
    pick = rng.integers(len(P), size=nreq)
    outs = [(torch.empty((P[k]["n"], ds.width), dtype=torch.float64, device=dev),
             torch.empty_like(P[k]["x"])) for k in pick]
    torch.cuda.synchronize()
    atoms = int(sum(P[k]["n"] for k in pick))

    def run(lanes, lc, min_centers, window_us, load, dump=None):
        srv = _capi.Server(ds, _capi.FP32_BULK, hbm_budget_bytes=8 << 30,
                           lane_centers=lc, lane_rows=max(3 * lc, 65536), lane_edges=lc * 48,
                           max_requests=256, max_lanes=lanes, dispatch_min_centers=min_centers,
                           batch_window_us=window_us)

        def submit(i):
            D = P[pick[i]]
            return srv.submit(D["species"], D["numneigh"], D["nl"], D["x"], D["dE"],
                              outs[i][0], outs[i][1], D["max_nn"])

        for tk in [submit(i) for i in range(min(nreq, 64))]:  # warm-up
            srv.wait(tk)
        # offline: everything queued at once
        t1 = time.perf_counter()
        tickets = [submit(i) for i in range(nreq)]
        for tk in tickets:
            srv.wait(tk)
        off_s = time.perf_counter() - t1
        off_rate = atoms / off_s
        # online: Poisson arrivals at `load` of the offline request rate
        lam = load * nreq / off_s
        arrivals = np.cumsum(np.random.default_rng(6).exponential(1.0 / lam, size=nreq))
        lat = np.zeros(nreq)
        tick = [None] * nreq

        def waiter():
            for i in range(nreq):
                while tick[i] is None:
                    time.sleep(0)
                lat[i] = srv.wait(tick[i])

        th = threading.Thread(target=waiter)
        th.start()
        t2 = time.perf_counter()
        for i in range(nreq):
            while time.perf_counter() - t2 < arrivals[i]:
                pass
            tick[i] = submit(i)
        th.join()
        on_s = time.perf_counter() - t2
        st = srv.stats()
        n_lanes = srv.num_lanes
        srv.close()
        if dump:
            np.savez(dump, atoms=np.array([P[k]["n"] for k in pick]), arrival=arrivals,
                     latency=lat)
        p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
        log(f"serving lanes={n_lanes} window={window_us}us: offline {off_rate:.3e}, online "
            f"{atoms / on_s:.3e} atom-envs/s, p50 {p50 * 1e3:.3f} ms, p99 {p99 * 1e3:.3f} ms")
        return {"offline_atom_envs_per_s": off_rate, "online_atom_envs_per_s": atoms / on_s,
                "online_fraction_of_offline": (atoms / on_s) / off_rate,
                "arrival_rate_per_s": lam, "p50_ms": p50 * 1e3, "p99_ms": p99 * 1e3,
                "p99_over_p50": p99 / p50, "lanes": n_lanes, "lane_centers": lc,
                "dispatch_min_centers": min_centers, "batch_window_us": window_us,
                "server_stats": st}

    out = {"requests": nreq, "atoms": atoms, "pool_configurations": len(P)}
    # headline policy: batches formed over a window, several lanes in flight
    out.update(run(args.serving_lanes, args.serving_lane_centers, args.serving_min_centers,
                   args.serving_window_us, args.serving_load, args.serving_dump))
    if args.serving_compare:
        # dispatch as soon as a lane is free: lower median latency, wider tail
        out["immediate_dispatch"] = run(4, 49152, 0, 0, args.serving_load)
    return out


def measured_traffic(kernel, n, form=""):
    """HBM bytes per launch of `kernel` on this workload, from the PMC summary
    kept beside the kernel traces (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration recorded in the same file; form "_fp64" for the
    FP64 instance); None if absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}{form}_c3_{n}_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            if ln.startswith("hbm_bytes_per_launch"):
                return float(ln.split()[1])
    return None


def copy_calibration(mib):
    """A device-to-device copy of `mib` MiB, 5 times, for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 2 * n * 8, "launches": 5}), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(L, types, sizes, params, g):
    """The oracle's FP64 forward + adjoint (reference loop structure, per-row
    recomputation kept) on the host cores, over bounded samples of centres:
    on the GPU's share of host threads (16 per GPU on these machines, capped
    by the cores this process may use) and on one thread, the reference's own
    mode (its OpenMP is commented out, src/Descriptors.cpp:10,466)."""
    from oracle import orc
    O = orc.OracleSF(["A", "B"], 5.0, types, sizes, params, species_resolved=True)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(cores, 16))
    off = np.r_[0, np.cumsum(L["numneigh"])].astype(np.int32)
    n = L["n_contrib"]

    def timed(k, nt):
        t = time.perf_counter()
        O.fwd_adj_omp(k, L["species"], L["neighlist"], off, L["coords"], g, nt)
        return time.perf_counter() - t

    def rate(nt, seconds):
        # grow a prefix until one takes >= 1 s, then size the sample for `seconds`
        k = 256
        while True:
            dt = timed(k, nt)
            if dt >= 1.0 or k >= n:
                break
            k = min(n, k * 4)
        k = int(min(n, max(256, seconds * k / dt)))
        dt = timed(k, nt)
        return k, dt

    k, dt = rate(threads, 10.0)
    k1, dt1 = rate(1, 6.0)
    log(f"cpu baseline: {k} centres in {dt:.2f} s on {threads} threads; {k1} centres in "
        f"{dt1:.2f} s on 1 thread")
    return {"value": k / dt, "unit": "atom-envs/s", "cores": threads, "kind": "port",
            "sample": f"first {k} of {n} centres of the same workload, FP64 forward + adjoint, "
                      f"OpenMP over centres",
            "single_thread_value": k1 / dt1, "single_thread_sample": f"first {k1} centres",
            "nproc": os.cpu_count(), "usable_cores": cores, "cpu_model": cpu_model(),
            "threads_note": "16 host threads per GPU is this machine's share (capped by the "
                            "usable cores); the reference itself is single-threaded"}


if __name__ == "__main__":
    main()
