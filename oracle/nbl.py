"""ORACLE -- test infrastructure only: numpy restatement of the KLIFF neighbour
list, periodic padding and ghost-gradient fold that produce/consume the SF
path's input layout.

* ``create_paddings``  restates nbl_create_paddings
  (python_bindings/neighbor/neighbor_list.cpp:283-418, helper.hpp:19-93).
* ``neighlist``        restates nbl_build + nbl_get_neigh + the 1-D flattening
  (neighbor_list.cpp:78-280, python_package/libdescriptor/neighbor.py:108-154).
* ``assemble_forces``  restates neighbor.py:227-263.

Pinned against oracle/_ref (the reference's neighbor_list.cpp compiled from
its own source) for order and values, and by brute-force O(n^2) membership.
"""
import numpy as np

TOL = 1.0e-10  # neighbor_list.cpp:14 and helper.hpp:10 (SMALL)


def _dot(a, b0, b1, b2):
    # helper.hpp:26-29, evaluated left to right like the C expression
    return a[0] * b0 + a[1] * b1 + a[2] * b2


def _inverse(m):
    # helper.hpp:73-93 (cofactors / det), same operation order
    def det2(a11, a12, a21, a22):
        return (a11 * a22) - (a12 * a21)
    inv = np.empty(9)
    inv[0] = det2(m[4], m[5], m[7], m[8])
    inv[1] = det2(m[2], m[1], m[8], m[7])
    inv[2] = det2(m[1], m[2], m[4], m[5])
    inv[3] = det2(m[5], m[3], m[8], m[6])
    inv[4] = det2(m[0], m[2], m[6], m[8])
    inv[5] = det2(m[2], m[0], m[5], m[3])
    inv[6] = det2(m[3], m[4], m[6], m[7])
    inv[7] = det2(m[1], m[0], m[7], m[6])
    inv[8] = det2(m[0], m[1], m[3], m[4])
    dd = (m[0] * m[4] * m[8] - m[0] * m[5] * m[7] - m[1] * m[3] * m[8] + m[1] * m[5] * m[6]
          + m[2] * m[3] * m[7] - m[2] * m[4] * m[6])
    if abs(dd) < TOL:
        raise RuntimeError("In inverting the cell matrix, the determinant is 0!")
    return inv / dd


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]])


def _norm(a):
    return np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def create_paddings(cutoff, cell, pbc, coords, species):
    """Returns (pad_coords (n_pad,3), pad_species (n_pad,), pad_image (n_pad,))."""
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    species = np.asarray(species, dtype=np.int32)
    tcell = cell.reshape(3, 3).T.reshape(9).copy()
    fcell = _inverse(tcell)
    x = _dot(fcell[0:3], coords[:, 0], coords[:, 1], coords[:, 2])
    y = _dot(fcell[3:6], coords[:, 0], coords[:, 1], coords[:, 2])
    z = _dot(fcell[6:9], coords[:, 0], coords[:, 1], coords[:, 2])
    mn = [min(1e10, x.min()) - TOL, min(1e10, y.min()) - TOL, min(1e10, z.min()) - TOL]
    mx = [max(-1e10, x.max()) + TOL, max(-1e10, y.max()) + TOL, max(-1e10, z.max()) + TOL]
    c0, c1, c2 = cell[0:3], cell[3:6], cell[6:9]
    xp = _cross(c1, c2)
    volume = abs(c0[0] * xp[0] + c0[1] * xp[1] + c0[2] * xp[2])
    dist = [volume / _norm(_cross(c1, c2)), volume / _norm(_cross(c2, c0)),
            volume / _norm(_cross(c0, c1))]
    ratio = [cutoff / dist[0], cutoff / dist[1], cutoff / dist[2]]
    size = [int(np.ceil(r)) for r in ratio]
    diff = [float(size[d]) - ratio[d] for d in range(3)]
    pc, ps, pi = [], [], []
    for i in range(-size[0], size[0] + 1):
        for j in range(-size[1], size[1] + 1):
            for k in range(-size[2], size[2] + 1):
                if i == 0 and j == 0 and k == 0:
                    continue
                if pbc[0] == 0 and i != 0:
                    continue
                if pbc[1] == 0 and j != 0:
                    continue
                if pbc[2] == 0 and k != 0:
                    continue
                keep = np.ones(len(x), dtype=bool)
                if i == -size[0]:
                    keep &= ~(x - mn[0] < diff[0])
                if i == size[0]:
                    keep &= ~(mx[0] - x < diff[0])
                if j == -size[1]:
                    keep &= ~(y - mn[1] < diff[1])
                if j == size[1]:
                    keep &= ~(mx[1] - y < diff[1])
                if k == -size[2]:
                    keep &= ~(z - mn[2] < diff[2])
                if k == size[2]:
                    keep &= ~(mx[2] - z < diff[2])
                idx = np.nonzero(keep)[0]
                fx, fy, fz = i + x[idx], j + y[idx], k + z[idx]
                pc.append(np.stack([_dot(tcell[0:3], fx, fy, fz), _dot(tcell[3:6], fx, fy, fz),
                                    _dot(tcell[6:9], fx, fy, fz)], axis=1))
                ps.append(species[idx])
                pi.append(idx.astype(np.int32))
    if not pc:
        return np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros(0, np.int32)
    return np.concatenate(pc), np.concatenate(ps).astype(np.int32), np.concatenate(pi)


def _coords_to_index(x, size, mx, mn):
    # helper.hpp:96-108
    idx = np.empty((x.shape[0], 3), dtype=np.int64)
    for d in range(3):
        v = ((x[:, d] - mn[d]) / (mx[d] - mn[d])) * size[d]
        idx[:, d] = np.minimum(np.trunc(v).astype(np.int64), size[d] - 1)
    return idx


def neighlist(coords, infl, need_neigh=None, n_query=None):
    """Full neighbour list in KLIFF cell order; returns (numneigh, neighlist) int32
    over the first ``n_query`` atoms (default: the atoms with need_neigh)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    n = coords.shape[0]
    if need_neigh is None:
        need_neigh = np.ones(n, dtype=np.int32)
    if n_query is None:
        n_query = n
    mn = coords[0].copy()
    mx = coords[0] + 1.0
    mx = np.maximum(mx, coords.max(axis=0))
    mn = np.minimum(mn, coords.min(axis=0))
    size = [max(1, int((mx[d] - mn[d]) / infl)) for d in range(3)]
    if size[0] * size[1] * size[2] > 1000000000:
        raise RuntimeError("Cell size too large! (partilces fly away) or\nCollision of atoms happened!")
    cidx = _coords_to_index(coords, size, mx, mn)
    flat = cidx[:, 0] + cidx[:, 1] * size[0] + cidx[:, 2] * size[0] * size[1]
    order = np.argsort(flat, kind="stable")
    cells = {}
    bounds = np.flatnonzero(np.diff(flat[order])) + 1
    for grp in np.split(order, bounds):
        if len(grp):
            cells[int(flat[grp[0]])] = grp
    cutsq = infl * infl
    numneigh = np.zeros(n_query, dtype=np.int32)
    out = []
    for i in range(n_query):
        if not need_neigh[i]:
            continue
        ci = cidx[i]
        cand = []
        for ii in range(max(0, ci[0] - 1), min(ci[0] + 1, size[0] - 1) + 1):
            for jj in range(max(0, ci[1] - 1), min(ci[1] + 1, size[1] - 1) + 1):
                for kk in range(max(0, ci[2] - 1), min(ci[2] + 1, size[2] - 1) + 1):
                    g = cells.get(ii + jj * size[0] + kk * size[0] * size[1])
                    if g is not None:
                        cand.append(g)
        cand = np.concatenate(cand) if cand else np.zeros(0, dtype=np.int64)
        cand = cand[cand != i]
        d = coords[cand] - coords[i]
        rsq = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        if np.any(rsq < TOL):
            raise RuntimeError("Cell size too large! (partilces fly away) or\nCollision of atoms happened!")
        sel = cand[rsq < cutsq]
        numneigh[i] = len(sel)
        out.append(sel.astype(np.int32))
    nl = np.concatenate(out).astype(np.int32) if out else np.zeros(0, dtype=np.int32)
    return numneigh, nl


def brute_force_sets(coords, infl, n_query):
    """Membership by O(n^2) distances: list of sorted neighbour index arrays."""
    coords = np.asarray(coords, dtype=np.float64)
    res = []
    for i in range(n_query):
        d = coords - coords[i]
        rsq = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        m = rsq < infl * infl
        m[i] = False
        res.append(np.flatnonzero(m))
    return res


def assemble_forces(forces, n, padding_image):
    """python_package/libdescriptor/neighbor.py:227-263, both branches kept."""
    forces = np.asarray(forces)
    padding_image = np.asarray(padding_image)
    total = np.array(forces[:n])
    if padding_image.size != 0:
        pad = forces[n:]
        if n < pad.shape[0]:
            for i in range(n):
                ind = np.where(padding_image == i)
                total[i] += np.sum(pad[ind], axis=0)
        else:
            for f, org in zip(pad, padding_image):
                total[org] += f
    return total
