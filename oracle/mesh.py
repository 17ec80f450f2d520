"""Structured voxel grid numbering (oracle; test infrastructure only).

Restates ``src/hex_mesher.py``:
  * node order x fastest, then y, then z ``:104-109``; coordinates ``:111-113``;
  * element order x fastest with corners [n0, n0+1, n0+1+sx, n0+sx, +sx*sy ...]
    ``:114-124``; element centres = mean of the 8 corners ``:154-157``;
  * ``createEdofMatStructural`` DOFs 3*node + {0,1,2} ``:1080-1095``;
  * ``getNodesOnBoundingBoxPlane`` ``:732-760``.
"""
import numpy as np


class GridMesh:
    """Vectorised equivalent of ``HexMesher.grid_mesh`` (src/hex_mesher.py:72-169)."""

    def __init__(self, num_elems, elem_size):
        nelx, nely, nelz = (int(v) for v in num_elems)
        self.grid = [nelx, nely, nelz]
        self.elem_size = tuple(float(v) for v in elem_size)
        self.num_elems = nelx * nely * nelz
        self.num_nodes = (nelx + 1) * (nely + 1) * (nelz + 1)
        sx, sy = nelx + 1, nely + 1
        iz, iy, ix = np.meshgrid(np.arange(nelz + 1), np.arange(nely + 1),
                                 np.arange(nelx + 1), indexing="ij")
        self.node_indices = np.zeros((self.num_nodes, 4), dtype=np.int32)
        self.node_indices[:, 0] = ix.ravel()
        self.node_indices[:, 1] = iy.ravel()
        self.node_indices[:, 2] = iz.ravel()
        self.node_xyz = self.node_indices[:, :3] * np.array(self.elem_size)
        ez, ey, ex = np.meshgrid(np.arange(nelz), np.arange(nely), np.arange(nelx),
                                 indexing="ij")
        n0 = (ez * sx * sy + ey * sx + ex).ravel()
        self.elemArray = np.stack([n0, n0 + 1, n0 + 1 + sx, n0 + sx,
                                   n0 + sx * sy, n0 + 1 + sx * sy,
                                   n0 + 1 + sx + sx * sy, n0 + sx + sx * sy],
                                  axis=1).astype(np.int32)
        self.elem_centers = self.node_xyz[self.elemArray].sum(axis=1) / 8.0

    def createEdofMatStructural(self):
        """src/hex_mesher.py:1080-1095."""
        e = self.elemArray.astype(np.int64)
        self.edofMatStructural = np.repeat(3 * e, 3, axis=1) + np.tile(np.arange(3), 8)
        return self.edofMatStructural

    def getNodesOnBoundingBoxPlane(self, axis, minLimit):
        """src/hex_mesher.py:732-760."""
        if axis not in (0, 1, 2):
            raise ValueError("Invalid axis. Must be 0, 1, or 2.")
        target = 0 if minLimit else self.grid[axis]
        return np.where(self.node_indices[:, axis] == target)[0]


def grid_mesh_loops(num_elems, elem_size):
    """Literal restatement of the reference's nested loops (tiny grids only).

    src/hex_mesher.py:104-124 and :1084-1094.  Used to pin ``GridMesh``.
    """
    nelx, nely, nelz = num_elems
    num_nodes = (nelx + 1) * (nely + 1) * (nelz + 1)
    node_indices = np.zeros((num_nodes, 4), dtype=np.int32)
    node = 0
    for iz in range(nelz + 1):
        for iy in range(nely + 1):
            for ix in range(nelx + 1):
                node_indices[node] = [ix, iy, iz, 0]
                node += 1
    elemArray = np.zeros((nelx * nely * nelz, 8), dtype=np.int32)
    elem = 0
    sx, sy = nelx + 1, nely + 1
    for iz in range(nelz):
        for iy in range(nely):
            for ix in range(nelx):
                n0 = iz * sx * sy + iy * sx + ix
                elemArray[elem] = [n0, n0 + 1, n0 + 1 + sx, n0 + sx,
                                   n0 + sx * sy, n0 + 1 + sx * sy,
                                   n0 + 1 + sx + sx * sy, n0 + sx + sx * sy]
                elem += 1
    edof = np.zeros((elem, 24), dtype=np.int64)
    for el in range(elem):
        row = []
        for n in elemArray[el]:
            row += [3 * n, 3 * n + 1, 3 * n + 2]
        edof[el] = row
    return node_indices, elemArray, edof
