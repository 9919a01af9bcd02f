"""Error statement of the FP32-bulk form against FP64, per descriptor block and
for dE/dr, as the benchmark line and the tests report it.

Descriptors, per block (the block's columns of every centre):
  * ``max_rel`` / ``rms_rel``: elementwise relative error over the entries that
    carry signal, |ref| >= SIGNAL x the block's largest |ref|;
  * ``n_signal``, ``n_excluded``: how many non-zero entries are in and out of
    that filter, and ``excluded_max_err`` the largest |error| among the
    excluded ones as a fraction of the block's largest |ref| (they are
    negligible against the block: both they and their errors sit below
    SIGNAL of its scale);
  * ``max_rel_normwise``: max |error| / the block's largest |ref|;
  * ``max_rel_all_nonzero``: elementwise relative error over every non-zero
    entry, unfiltered.  Entries far below the filter come from
    (1 + lambda cos theta)^zeta with the base within ~1e-3 of 0 and zeta = 16:
    there the FP64 value is itself rounding noise of a cancelling base, and
    the FP32-bulk form gives 0 or a value of the same tiny size, so this
    figure can reach 1.

dE/dr (every coordinate row, padding rows included):
  * ``max_rel_normwise``, ``rms_rel_normwise``: |error| over max|ref|;
  * ``elem``: for each significance floor f in GRAD_FLOORS, the elementwise
    figure max |error_i| / max(|ref_i|, f max|ref|) and the number of
    components below the floor.  A component whose contributions cancel to
    far below the largest forces carries the absolute error of its terms
    (about 1e-7 of them in FP32), not one relative to its own size, so the
    elementwise figure grows as the floor drops: no form that evaluates its
    terms in FP32 has a small elementwise error on such components.
"""
import numpy as np

SIGNAL = 1e-6
GRAD_FLOORS = (1e-1, 1e-2, 1e-3, 1e-4, 1e-6)

BLOCKS_TWO_SPECIES = {"g2_A": (0, 8), "g2_B": (8, 16), "g4_AA": (16, 32), "g4_AB": (32, 48),
                      "g4_BB": (48, 64)}


def descriptor_errors(d, ref, blocks=None):
    """Per-block figures of d against ref, both (n, width)."""
    d = np.asarray(d, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if blocks is None:
        blocks = {"all": (0, ref.shape[1])}
    out = {}
    for name, (lo, hi) in blocks.items():
        a, r = d[:, lo:hi], ref[:, lo:hi]
        ar = np.abs(r)
        scale = float(ar.max()) if ar.size else 0.0
        diff = np.abs(a - r)
        sig = ar >= SIGNAL * scale if scale > 0 else np.zeros(ar.shape, bool)
        nz = ar > 0
        exc = nz & ~sig
        rel = diff[sig] / ar[sig]
        out[name] = {
            "max_rel": float(rel.max()) if rel.size else 0.0,
            "rms_rel": float(np.sqrt((rel ** 2).mean())) if rel.size else 0.0,
            "n_signal": int(sig.sum()),
            "n_excluded": int(exc.sum()),
            "excluded_max_err": float(diff[exc].max() / scale) if exc.any() else 0.0,
            "max_rel_normwise": float(diff.max() / scale) if scale > 0 else 0.0,
            "max_rel_all_nonzero": float((diff[nz] / ar[nz]).max()) if nz.any() else 0.0,
        }
    return out


def gradient_errors(g, ref):
    """Figures of dE/dr g against ref, both (rows, 3)."""
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    ref = np.asarray(ref, dtype=np.float64).reshape(-1)
    scale = float(np.abs(ref).max()) if ref.size else 0.0
    out = {"max_rel_normwise": 0.0, "rms_rel_normwise": 0.0, "n": int(ref.size),
           "elem": {f"{f:g}": {"max_rel": 0.0, "n_below": 0} for f in GRAD_FLOORS}}
    if scale == 0:
        return out
    diff = np.abs(g - ref)
    aref = np.abs(ref)
    out["max_rel_normwise"] = float(diff.max() / scale)
    out["rms_rel_normwise"] = float(np.sqrt((diff ** 2).mean()) / scale)
    for f in GRAD_FLOORS:
        out["elem"][f"{f:g}"] = {"max_rel": float((diff / np.maximum(aref, f * scale)).max()),
                                 "n_below": int((aref < f * scale).sum())}
    return out


def elem_at(table, floor):
    """The elementwise dE/dr figure of a gradient_errors table at one floor."""
    return table["elem"][f"{floor:g}"]["max_rel"]


def worst(table, key):
    """Largest figure `key` over the blocks of a descriptor_errors table."""
    return max((v[key] for v in table.values()), default=0.0)
