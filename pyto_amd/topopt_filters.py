"""Cone (smoothing) filter as a fixed GPU stencil (src/topopt_filters.py:8-39).

``createSmoothingFilter`` returns ``(H, Hs)`` like the reference, where H is a
``ConeFilter`` (``H * v`` / ``H @ v`` apply the filter, ``H.T`` is H since the
weights are symmetric) and Hs is the row-sum vector.  The index structure is
the reference's: every element centre within r = rel*h_x, d <= r, truncated
at the grid boundary; weight (r - d)/r.
"""
import numpy as np
import scipy.sparse as sp
import torch

from .operator import GridOperator


class ConeFilter:
    def __init__(self, operator, rel_filter_radius):
        self.op = operator
        self.rel = float(rel_filter_radius)
        self.shape = (operator.num_elems, operator.num_elems)

    def _apply(self, v, mode, aux=None):
        as_numpy = not isinstance(v, torch.Tensor)
        out = self.op.filter(v, self.rel, mode, aux)
        return out.cpu().numpy() if as_numpy else out

    def __mul__(self, v):
        return self._apply(v, 0)

    __matmul__ = __mul__

    @property
    def T(self):
        return self

    def rowsum(self):
        return self.op.filter(None, self.rel, 4)

    def sum(self, axis=None):
        if axis in (0, 1):
            return self.rowsum().cpu().numpy()[:, None] if axis == 1 else self.rowsum().cpu().numpy()[None, :]
        return float(self.rowsum().sum())

    def density_filter(self, x):
        """xPhys = (H x)/Hs, src/topopt_mma.py:255-256."""
        return self._apply(x, 1)

    def density_chain(self, dphys):
        """dx = H (dphys / Hs), src/topopt_mma.py:342-345."""
        return self._apply(dphys, 3)

    def sensitivity_filter(self, x, g):
        """(H (x g))/Hs/x, src/topopt_ocm.py:239."""
        as_numpy = not isinstance(g, torch.Tensor)
        out = self.op.filter(g, self.rel, 2, aux=x)
        return out.cpu().numpy() if as_numpy else out

    def stencil(self):
        return self.op.filter_stencil(self.rel)

    def to_scipy(self):
        """Assemble H as CSC from the stencil (small grids; inspection only)."""
        off, w = self.stencil()
        nelx, nely, nelz = self.op.grid
        ne = self.op.num_elems
        e = np.arange(ne)
        ix, iy, iz = e % nelx, (e // nelx) % nely, e // (nelx * nely)
        rows, cols, vals = [], [], []
        for (dx, dy, dz), wk in zip(off, w):
            jx, jy, jz = ix + dx, iy + dy, iz + dz
            ok = (jx >= 0) & (jx < nelx) & (jy >= 0) & (jy < nely) & (jz >= 0) & (jz < nelz)
            rows.append(e[ok])
            cols.append((jx + nelx * (jy + nely * jz))[ok])
            vals.append(np.full(int(ok.sum()), wk))
        return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(ne, ne)).tocsc()


def createSmoothingFilter(mesh, rel_filter_radius=1.5, operator=None):
    """src/topopt_filters.py:8-39 -> (H, Hs)."""
    if operator is None:
        operator = GridOperator(mesh.grid, mesh.elem_size, 1.0, 0.3, [])
    H = ConeFilter(operator, rel_filter_radius)
    return H, H.rowsum()
