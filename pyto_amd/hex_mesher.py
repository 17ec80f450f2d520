"""Structured voxel grid (src/hex_mesher.py, the ``grid_mesh`` part).

Numbering follows the reference exactly: node id = ix + (nx+1) iy +
(nx+1)(ny+1) iz (src/hex_mesher.py:104-109), element id = ix + nx iy + nx ny iz
with corners [n0, n0+1, n0+1+sx, n0+sx, ...+sx*sy] (:117-124), DOF 3*node+c
(:1080-1095).  The large per-node / per-element arrays are built on first
access (vectorised instead of the reference's Python loops); the GPU kernels
never read them, they use the closed-form indices.
"""
import numpy as np


class BoundingBox:
    def __init__(self, lo, hi):
        self.lo = np.asarray(lo, float)
        self.hi = np.asarray(hi, float)
        self.lx, self.ly, self.lz = (self.hi - self.lo)


class HexMesher:
    """src/hex_mesher.py:62-169 (grid part)."""

    def __init__(self):
        self.num_nodes = 0
        self.num_elems = 0

    def grid_mesh(self, num_elems, elem_size):
        nelx, nely, nelz = (int(v) for v in num_elems)
        if min(nelx, nely, nelz) < 1:
            raise ValueError("grid must have at least one element per axis")
        self.grid = [nelx, nely, nelz]
        self.elem_size = tuple(float(v) for v in elem_size)
        self.origin = [0, 0, 0]
        self.num_elems = nelx * nely * nelz
        self.num_nodes = (nelx + 1) * (nely + 1) * (nelz + 1)
        self.bbox = BoundingBox([0, 0, 0], [nelx * self.elem_size[0], nely * self.elem_size[1],
                                            nelz * self.elem_size[2]])
        self._cache = {}
        return self

    # -- lazily built arrays (reference attribute names) --------------------
    def _node_ijk(self):
        nelx, nely, nelz = self.grid
        n = np.arange(self.num_nodes, dtype=np.int64)
        sx, sy = nelx + 1, nely + 1
        return n % sx, (n // sx) % sy, n // (sx * sy)

    @property
    def node_indices(self):
        if "node_indices" not in self._cache:
            ix, iy, iz = self._node_ijk()
            a = np.zeros((self.num_nodes, 4), dtype=np.int32)
            a[:, 0], a[:, 1], a[:, 2] = ix, iy, iz
            self._cache["node_indices"] = a
        return self._cache["node_indices"]

    @property
    def node_xyz(self):
        if "node_xyz" not in self._cache:
            ix, iy, iz = self._node_ijk()
            h = self.elem_size
            self._cache["node_xyz"] = np.stack([ix * h[0], iy * h[1], iz * h[2]], axis=1).astype(np.float64)
        return self._cache["node_xyz"]

    @property
    def elemArray(self):
        if "elemArray" not in self._cache:
            nelx, nely, _ = self.grid
            e = np.arange(self.num_elems, dtype=np.int64)
            ex, ey, ez = e % nelx, (e // nelx) % nely, e // (nelx * nely)
            sx, sy = nelx + 1, nely + 1
            n0 = ez * sx * sy + ey * sx + ex
            self._cache["elemArray"] = np.stack(
                [n0, n0 + 1, n0 + 1 + sx, n0 + sx, n0 + sx * sy, n0 + 1 + sx * sy,
                 n0 + 1 + sx + sx * sy, n0 + sx + sx * sy], axis=1).astype(np.int32)
        return self._cache["elemArray"]

    @property
    def elem_centers(self):
        if "elem_centers" not in self._cache:
            self._cache["elem_centers"] = self.node_xyz[self.elemArray].sum(axis=1) / 8.0
        return self._cache["elem_centers"]

    def createEdofMatStructural(self):
        """src/hex_mesher.py:1080-1095 (built on first access of edofMatStructural)."""
        self.dofs_per_node = 3

    @property
    def edofMatStructural(self):
        if "edof" not in self._cache:
            e = self.elemArray.astype(np.int64)
            self._cache["edof"] = np.repeat(3 * e, 3, axis=1) + np.tile(np.arange(3), 8)
        return self._cache["edof"]

    def createEdofMatThermal(self):
        """src/hex_mesher.py:1120-1135: one DOF per node, the element's nodes."""
        self.dofs_per_node = 1

    @property
    def edofMatThermal(self):
        return self.elemArray.astype(np.int64)

    def getNodesOnBoundingBoxPlane(self, axis, minLimit):
        """src/hex_mesher.py:732-760."""
        if axis not in (0, 1, 2):
            raise ValueError("Invalid axis. Must be 0, 1, or 2.")
        nelx, nely, nelz = self.grid
        sx, sy, sz = nelx + 1, nely + 1, nelz + 1
        target = 0 if minLimit else self.grid[axis]
        iz, iy, ix = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        sel = [ix, iy, iz][axis] == target
        return (ix + sx * (iy + sy * iz))[sel].astype(np.int64)

    def setPseudoDensity(self, rho):
        self.elemPseudoDensity = rho
