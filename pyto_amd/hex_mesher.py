"""Structured voxel grid (src/hex_mesher.py, the ``grid_mesh`` part).

Numbering follows the reference exactly: node id = ix + (nx+1) iy +
(nx+1)(ny+1) iz (src/hex_mesher.py:104-109), element id = ix + nx iy + nx ny iz
with corners [n0, n0+1, n0+1+sx, n0+sx, ...+sx*sy] (:117-124), DOF 3*node+c
(:1080-1095).  The large per-node / per-element arrays are built on first
access (vectorised instead of the reference's Python loops); the GPU kernels
never read them, they use the closed-form indices.

``export_vtu_mesh`` writes the reference's VTU (src/hex_mesher.py:1446-1499)
without pyvista: a VTK XML UnstructuredGrid of hexahedra with the element
field averaged onto the points, base64-encoded binary arrays.
"""
import base64
import os

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

    def find_connected_components(self, threshold=0.01):
        """src/hex_mesher.py:1285-1339: element-index arrays of the connected
        sets of elements with elemPseudoDensity > threshold (face, edge and
        corner neighbours), ordered by their smallest element index as the
        reference lists them.  Labelled on the GPU; each array is sorted
        (the reference keeps its search's visiting order)."""
        from .builders import element_components
        labels = element_components(self.grid, self.elemPseudoDensity, threshold).cpu().numpy()
        present = np.nonzero(labels >= 0)[0]
        if present.size == 0:
            return []
        order = np.argsort(labels[present], kind="stable")
        lab_sorted = labels[present][order]
        cuts = np.nonzero(np.diff(lab_sorted))[0] + 1
        return [np.sort(c) for c in np.split(present[order], cuts)]

    def export_vtu_mesh(self, elem_field, mask_low_pseudodensity=False, density_field="density",
                        file_name=None):
        """src/hex_mesher.py:1446-1499: hexahedral cells (VTK type 12) over all
        grid nodes, optionally only the elements with pseudo-density > 0.5;
        the element field becomes point data, each point the mean over the
        kept elements that share it (cells without point use get 0)."""
        field = np.asarray(elem_field.cpu() if hasattr(elem_field, "cpu") else elem_field,
                           dtype=np.float64).ravel()
        elems = self.elemArray
        if mask_low_pseudodensity:
            rho = np.asarray(self.elemPseudoDensity, dtype=np.float64).ravel()
            keep = rho > 0.5
            elems, field = elems[keep], field[keep]
        if len(elems) == 0:
            print("No elements to plot after filtering")
            return
        file_name = "./temp.vtu" if file_name is None else file_name
        write_vtu_hexes(file_name, self.node_xyz, elems, {density_field: field})


def _vtu_array(name, arr, ncomp=1):
    a = np.ascontiguousarray(arr)
    vtk_type = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
                np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32",
                np.dtype(np.uint8): "UInt8"}[a.dtype]
    raw = a.tobytes()
    payload = base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")
    return ('<DataArray type="%s" Name="%s" NumberOfComponents="%d" format="binary">%s'
            '</DataArray>' % (vtk_type, name, ncomp, payload))


def write_vtu_hexes(file_name, points, cells, cell_fields_to_points):
    """VTK XML UnstructuredGrid of 8-node hexahedra; each field (one value per
    cell) is written as point data, the mean over the cells sharing a point."""
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    cells = np.asarray(cells, dtype=np.int64).reshape(-1, 8)
    npts, ncells = points.shape[0], cells.shape[0]
    count = np.bincount(cells.ravel(), minlength=npts).astype(np.float64)
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
             'header_type="UInt64">',
             '<UnstructuredGrid>',
             '<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (npts, ncells),
             '<PointData>']
    for name, f in cell_fields_to_points.items():
        f = np.asarray(f, dtype=np.float64).ravel()
        s = np.bincount(cells.ravel(), weights=np.repeat(f, 8), minlength=npts)
        parts.append(_vtu_array(name, np.divide(s, count, out=np.zeros(npts), where=count > 0)))
    parts += ['</PointData>', '<Points>', _vtu_array("Points", points, 3), '</Points>',
              '<Cells>', _vtu_array("connectivity", cells.ravel()),
              _vtu_array("offsets", 8 * np.arange(1, ncells + 1, dtype=np.int64)),
              _vtu_array("types", np.full(ncells, 12, dtype=np.uint8)), '</Cells>',
              '</Piece>', '</UnstructuredGrid>', '</VTKFile>']
    d = os.path.dirname(os.path.abspath(file_name))
    os.makedirs(d, exist_ok=True)
    with open(file_name, "w") as fh:
        fh.write("\n".join(parts) + "\n")
