"""Seeded synthetic workloads for the symmetry-function path.

* ``behler_si_params()``  -- G2: 8 eta log-spaced 0.001..0.5 A^-2 (eta_q =
  0.001 * 500^(q/7)), Rs = 0; G4: zeta in (1, 2, 4, 16) x lambda in (+1, -1)
  x eta in (0.001, 0.01), rows in that nesting order (zeta outermost).
  24 entries per atom; 64 in the two-species species-resolved layout
  (2 x 8 G2 + 3 x 16 G4).
* ``diamond_si(ncell)``  -- diamond-cubic Si, a = 5.431 A, ncell^3 conventional
  cells, N(0, sigma=0.05 A) jitter per coordinate, wrapped into the box.
* ``random_system(n, density, dmin)`` -- uniform random sequential addition in
  a cubic periodic box with a minimum pair distance, species drawn 80/20.
* ``cluster(n, density, dmin)`` -- the same inside a sphere, non-periodic.
* ``layout(...)`` -- padding + neighbour list (the library's C++ producer) into
  the flat arrays every entry point consumes; ``concat(...)`` joins layouts
  into the compute_batch / gradient_batch form (src/Descriptors.cpp:138-163).
"""
import numpy as np

from .neighbor import NeighborList


def behler_si_params():
    eta2 = 0.001 * 500.0 ** (np.arange(8) / 7.0)
    g2 = np.stack([eta2, np.zeros(8)], axis=1).reshape(-1)
    g4 = np.array([[z, lam, e] for z in (1.0, 2.0, 4.0, 16.0) for lam in (1.0, -1.0)
                   for e in (0.001, 0.01)]).reshape(-1)
    return ["g2", "g4"], [8, 2, 16, 3], list(np.concatenate([g2, g4]))


def diamond_si(ncell, a=5.431, sigma=0.05, seed=0):
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=float)
    basis = np.concatenate([basis, basis + 0.25])
    g = np.arange(ncell)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    pos = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    rng = np.random.default_rng(seed)
    L = ncell * a
    pos = np.mod(pos + rng.normal(0.0, sigma, pos.shape), L)
    return pos, np.eye(3) * L, np.array([1, 1, 1], dtype=np.int32)


def _rsa(n, rng, dmin, sample, periodic_L=None, batch=4096, max_tries=200):
    """Random sequential addition: candidates from ``sample(k)`` are accepted
    in order when no accepted point lies closer than dmin."""
    from scipy.spatial import cKDTree
    pts = np.zeros((0, 3))
    batch = min(batch, max(64, 2 * n))  # small systems: few candidates per round
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > max_tries * (1 + n // batch):
            raise RuntimeError("random sequential addition did not converge")
        cand = sample(batch)
        if len(pts):
            tree = cKDTree(pts, boxsize=periodic_L) if periodic_L else cKDTree(pts)
            d, _ = tree.query(cand, k=1, distance_upper_bound=dmin)
            cand = cand[~(d < dmin)]
        if len(cand) > 1:
            ct = cKDTree(cand, boxsize=periodic_L) if periodic_L else cKDTree(cand)
            pairs = ct.query_pairs(dmin, output_type="ndarray")
            if len(pairs):
                keep = np.ones(len(cand), dtype=bool)
                pairs = pairs[np.lexsort((pairs[:, 0], pairs[:, 1]))]
                for a, b in pairs:  # earlier candidate wins
                    if keep[a] and keep[b]:
                        keep[b] = False
                cand = cand[keep]
        pts = np.concatenate([pts, cand[: n - len(pts)]])
    return pts


def random_system(n, density=0.05, dmin=1.8, frac_minor=0.2, seed=0):
    """n atoms uniform in a periodic cube of side (n/density)^(1/3); species 0
    (80%) / 1 (20%)."""
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    L_box = L * (1.0 - 1e-12)  # keep wrapped samples strictly inside [0, L)
    pos = _rsa(n, rng, dmin, lambda k: rng.random((k, 3)) * L_box, periodic_L=L)
    species = (rng.random(n) < frac_minor).astype(np.int32)
    return pos, np.eye(3) * L, np.array([1, 1, 1], dtype=np.int32), species


def cluster(n, density=0.05, dmin=1.8, frac_minor=0.2, seed=0):
    """n atoms uniform in a sphere of volume n/density, non-periodic."""
    rng = np.random.default_rng(seed)
    R = (3.0 * n / (4.0 * np.pi * density)) ** (1.0 / 3.0)

    def sample(k):
        v = rng.normal(size=(k, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        return v * (R * rng.random((k, 1)) ** (1.0 / 3.0))

    pos = _rsa(n, rng, dmin, sample) + R
    species = (rng.random(n) < frac_minor).astype(np.int32)
    cell = np.eye(3) * (2.0 * R + 10.0)  # non-singular, unused without pbc
    return pos, cell, np.array([0, 0, 0], dtype=np.int32), species


def layout(positions, cell, pbc, species_codes, rc, symbols=None):
    """Padded coordinates, species codes of every row, numneigh and neighlist
    of the contributing atoms, and the image map."""
    species_codes = np.asarray(species_codes, dtype=np.int32)
    if symbols is None:
        symbols = [f"S{c}" for c in range(int(species_codes.max()) + 1)]
    sym = [symbols[c] for c in species_codes]
    nl = NeighborList.from_arrays(positions, cell, pbc, sym, rc)
    numneigh, neighlist = nl.get_numneigh_and_neighlist_1D()
    mapping = {s: k for k, s in enumerate(symbols)}
    return dict(coords=nl.get_coords(), species=nl.get_species_code(mapping),
                numneigh=numneigh, neighlist=neighlist, image=nl.get_image(),
                n_contrib=len(positions))


def concat(layouts):
    """Batch layout: n_atoms[c], config_ptr[c] (first row of c), global
    neighbour indices, rows of all configurations concatenated."""
    coords, species, numneigh, neighlist, n_atoms, ptr = [], [], [], [], [], []
    row = 0
    for L in layouts:
        n_atoms.append(L["n_contrib"])
        ptr.append(row)
        coords.append(L["coords"])
        species.append(L["species"])
        numneigh.append(L["numneigh"])
        neighlist.append(L["neighlist"] + row)
        row += len(L["coords"])
    return dict(coords=np.concatenate(coords), species=np.concatenate(species).astype(np.int32),
                numneigh=np.concatenate(numneigh).astype(np.int32),
                neighlist=np.concatenate(neighlist).astype(np.int32),
                n_atoms=np.asarray(n_atoms, dtype=np.int32),
                config_ptr=np.asarray(ptr, dtype=np.int32))


def config_c1(seed=0):
    pos, cell, pbc = diamond_si(2, seed=seed)
    return layout(pos, cell, pbc, np.zeros(len(pos), np.int32), 4.5, symbols=["Si"])


def config_c2(seed=0):
    pos, cell, pbc = diamond_si(16, seed=seed)
    return layout(pos, cell, pbc, np.zeros(len(pos), np.int32), 4.5, symbols=["Si"])


def config_c3(n=262144, seed=0, cache_dir=None):
    """262,144-atom two-species system; with ``cache_dir`` (or $LDSF_WORKLOAD_CACHE)
    the layout is stored there as .npz and reused."""
    import os
    cache_dir = cache_dir or os.environ.get("LDSF_WORKLOAD_CACHE")
    path = os.path.join(cache_dir, f"c3_{n}_{seed}.npz") if cache_dir else None
    if path and os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        return {k: (int(z[k]) if k == "n_contrib" else z[k]) for k in z.files}
    pos, cell, pbc, sp = random_system(n, 0.05, 1.8, seed=seed)
    L = layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"])
    if path:
        os.makedirs(cache_dir, exist_ok=True)
        np.savez(path, **L)
    return L


def config_c4(n_configs=512, seed=0, nmin=8, nmax=4096):
    """Half periodic bulk (density U[0.04, 0.06]), half clusters; log-uniform sizes."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_configs)).astype(int)
    layouts = []
    for c, n in enumerate(sizes):
        rho = rng.uniform(0.04, 0.06)
        s = int(rng.integers(1 << 30))
        if c % 2 == 0:
            pos, cell, pbc, sp = random_system(int(n), rho, 1.8, seed=s)
        else:
            pos, cell, pbc, sp = cluster(int(n), rho, 1.8, seed=s)
        layouts.append(layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"]))
    return layouts


def config_c5_pool(n_pool=192, seed=0, nmin=16, nmax=20000, cache_dir=None):
    """Configurations a serving request stream draws from: atom counts
    log-uniform in [nmin, nmax], each generated like config_c3 (random
    periodic, 0.05 atoms/A^3, 1.8 A minimum distance, species 80/20,
    cutoff 5.0 A).  Cached like config_c3."""
    import os
    cache_dir = cache_dir or os.environ.get("LDSF_WORKLOAD_CACHE")
    path = os.path.join(cache_dir, f"c5_{n_pool}_{seed}.npz") if cache_dir else None
    if path and os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        return [{k: (int(z[f"{c}_{k}"]) if k == "n_contrib" else z[f"{c}_{k}"])
                 for k in ("coords", "species", "numneigh", "neighlist", "image", "n_contrib")}
                for c in range(n_pool)]
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_pool)).astype(int)
    pool = []
    for n in sizes:
        pos, cell, pbc, sp = random_system(int(n), 0.05, 1.8, seed=int(rng.integers(1 << 30)))
        pool.append(layout(pos, cell, pbc, sp, 5.0, symbols=["A", "B"]))
    if path:
        os.makedirs(cache_dir, exist_ok=True)
        np.savez(path, **{f"{c}_{k}": v for c, L in enumerate(pool) for k, v in L.items()})
    return pool
