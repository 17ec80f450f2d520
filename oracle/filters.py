"""Cone (smoothing) filter and the OC update (oracle; test infrastructure only).

Restates:
  * ``createSmoothingFilter`` src/topopt_filters.py:8-39: r = rel * h_x
    (``:24``, only elem_size[0]); candidates are element centres with
    |c_i - c_e|^2 <= r^2 (the hash-grid query, src/hex_mesher.py:1162-1200),
    then pairs with ``dist > r`` are skipped (``:29-31``); weight (r-d)/r;
    H is CSC, Hs = rowsum(H) (``:37-38``; src/topopt_common.py:158);
  * the OC loop's sensitivity filter (H (x g))/Hs/x, src/topopt_ocm.py:239,
    force-element pinning ``:241-245``, bisection update ``:253-269``;
  * the density filter and its chain rule, src/topopt_mma.py:255-256,342-345.
"""
import numpy as np
import scipy.sparse as sp


def cone_filter(mesh, rel_filter_radius=1.5):
    """Brute-force restatement of src/topopt_filters.py:8-39.

    The candidate set is every element centre within r (a box search that is a
    superset of the reference's hash-grid cells), then the same distance test.
    """
    c = mesh.elem_centers
    r = rel_filter_radius * mesh.elem_size[0]
    ne = c.shape[0]
    h = np.array(mesh.elem_size)
    reach = np.ceil(r / h).astype(int)
    nelx, nely, nelz = mesh.grid
    ix = np.arange(ne) % nelx
    iy = (np.arange(ne) // nelx) % nely
    iz = np.arange(ne) // (nelx * nely)
    rows, cols, vals = [], [], []
    for dz in range(-reach[2], reach[2] + 1):
        for dy in range(-reach[1], reach[1] + 1):
            for dx in range(-reach[0], reach[0] + 1):
                jx, jy, jz = ix + dx, iy + dy, iz + dz
                ok = ((jx >= 0) & (jx < nelx) & (jy >= 0) & (jy < nely) &
                      (jz >= 0) & (jz < nelz))
                e = np.nonzero(ok)[0]
                j = jx[ok] + nelx * jy[ok] + nelx * nely * jz[ok]
                d2 = np.sum((c[e] - c[j]) ** 2, axis=1)
                keep = d2 <= r * r
                e, j = e[keep], j[keep]
                dist = np.linalg.norm(c[e] - c[j], axis=1)
                keep = ~(dist > r)
                rows.append(e[keep])
                cols.append(j[keep])
                vals.append((r - dist[keep]) / r)
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ne, ne)).tocsc()
    Hs = np.array(H.sum(1)).squeeze()
    return H, Hs


def cone_filter_all_pairs(mesh, rel_filter_radius=1.5):
    """O(N^2) all-pairs version (tiny grids only), used to pin ``cone_filter``."""
    c = mesh.elem_centers
    r = rel_filter_radius * mesh.elem_size[0]
    d = np.linalg.norm(c[:, None, :] - c[None, :, :], axis=2)
    e, j = np.nonzero(~(d > r))
    H = sp.coo_matrix(((r - d[e, j]) / r, (e, j)), shape=(len(c), len(c))).tocsc()
    return H, np.array(H.sum(1)).squeeze()


def sensitivity_filter(H, Hs, x, g):
    """src/topopt_ocm.py:239."""
    return (H @ (x * g)) / Hs / x


def density_filter(H, Hs, x):
    """src/topopt_mma.py:255-256: xPhys = (H x)/Hs."""
    return (H @ x) / Hs


def density_filter_chain(H, Hs, dphys):
    """src/topopt_mma.py:342-345: dx = H (dphys / Hs)."""
    return H @ (dphys / Hs)


def oc_update(x, grad, volfrac, move=0.2, xmin=0.001, xmax=1.0, return_info=False):
    """Bisection OC update, src/topopt_ocm.py:253-269.  With return_info,
    also the last lmid evaluated and the number of bisection steps."""
    dv = np.ones_like(x) / volfrac / x.size
    l1, l2 = 0.0, 1e12
    lmid = 0.5 * (l2 + l1)
    steps = 0
    xnew = x.copy()
    while (l2 - l1) / ((l1 + l2) / 2 + 1e-10) > 1e-7:
        steps += 1
        lmid = 0.5 * (l2 + l1)
        D = np.clip(-grad / (dv * lmid), 0.1, 10)
        xnew = np.maximum(xmin, np.maximum(x - move, np.minimum(xmax, np.minimum(x + move, x * np.sqrt(D)))))
        vc = np.mean(xnew) / volfrac - 1.0
        if vc > 0:
            l1 = lmid
        else:
            l2 = lmid
    if return_info:
        return xnew, lmid, steps
    return xnew
