"""Voxel problem builders on the GPU.

The per-node and per-element set-up that the reference does with Python
loops or whole-grid numpy passes, built on the device and bit-identical to
the host results:

* ``grid_arrays``: node coordinates, element corner nodes and structural
  element DOFs (src/hex_mesher.py:104-152,1080-1095);
* ``fixed_mask``: the per-node fixed-DOF mask of boundary planes
  (getNodesOnBoundingBoxPlane, src/hex_mesher.py:732-760, and the clamp of
  src/hex_structural_examples.py:353-356);
* ``face_load``: the face traction with interior/edge/corner weights 4/2/1
  (src/hex_structural_examples.py:360-395);
* ``box_groups``: the deflation groups of create_deflation_groups
  (src/deflation.py:86-229).  On a grid the box index of a node is a
  function of one axis at a time, so the floor and clamp of the reference
  run on the nx+1, ny+1, nz+1 axis coordinates, the box counts are products
  of per-axis counts, and the node-order coordinate sums of the box centres
  run on the device; the merge of boxes under ``min_nodes`` nodes into the
  nearest valid box is decided on the host over the boxes, as the reference
  does it.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _dev(device):
    return torch.device("cuda") if device is None else torch.device(device)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _grid(grid):
    nx, ny, nz = (int(v) for v in grid)
    if min(nx, ny, nz) < 1:
        raise ValueError("grid must have at least one element per axis")
    return nx, ny, nz


def grid_arrays(grid, elem_size, device=None, node_xyz=True, elemArray=True,
                edofMatStructural=True):
    """Device arrays of HexMesher.grid_mesh: ``node_xyz`` (N x 3 float64),
    ``elemArray`` (E x 8 int32) and ``edofMatStructural`` (E x 24 int64)."""
    nx, ny, nz = _grid(grid)
    dev = _dev(device)
    nn, ne = (nx + 1) * (ny + 1) * (nz + 1), nx * ny * nz
    out = {}
    if node_xyz:
        out["node_xyz"] = torch.empty((nn, 3), dtype=torch.float64, device=dev)
    if elemArray:
        out["elemArray"] = torch.empty((ne, 8), dtype=torch.int32, device=dev)
    if edofMatStructural:
        out["edofMatStructural"] = torch.empty((ne, 24), dtype=torch.int64, device=dev)
    h = [float(v) for v in elem_size]
    with torch.cuda.device(dev):
        _lib.check(_lib.lib().pyto_grid_arrays(nx, ny, nz, h[0], h[1], h[2],
                                               _p(out.get("node_xyz")), _p(out.get("elemArray")),
                                               _p(out.get("edofMatStructural")), _stream(dev)),
                   "pyto_grid_arrays")
    return out


def fixed_mask(grid, planes, device=None, out=None):
    """Per-node uint8 mask (bit c = DOF c fixed) of the boundary planes
    ``planes`` = [(axis, at_max, comp_bits), ...]; e.g. the x = 0 clamp of
    all three components is [(0, False, 7)]."""
    nx, ny, nz = _grid(grid)
    dev = _dev(device)
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    mask = torch.zeros(nn, dtype=torch.uint8, device=dev) if out is None else out
    with torch.cuda.device(dev):
        for axis, at_max, bits in planes:
            _lib.check(_lib.lib().pyto_plane_mask(nx, ny, nz, int(axis), int(bool(at_max)),
                                                  int(bits), _p(mask), _stream(dev)),
                       "pyto_plane_mask")
    return mask


def face_load(grid, axis=0, at_max=True, comp=2, total=1.0, dofs_per_node=3, device=None,
              out=None):
    """Load vector of total magnitude ``total`` in direction -comp on a grid
    face, nodal weights 4/2/1 (the reference's cantilever tip traction is
    axis 0, at_max, comp 2, total 1)."""
    nx, ny, nz = _grid(grid)
    dev = _dev(device)
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    f = torch.zeros(dofs_per_node * nn, dtype=torch.float64, device=dev) if out is None else out
    with torch.cuda.device(dev):
        _lib.check(_lib.lib().pyto_face_load(nx, ny, nz, int(axis), int(bool(at_max)), int(comp),
                                             int(dofs_per_node), float(total), _p(f),
                                             _stream(dev)), "pyto_face_load")
    return f


def reference_boxes(grid, elem_size, nGroupsDesired, minNodesPerGroup=15):
    """nX, nY, nZ of create_deflation_groups (src/deflation.py:108-126)."""
    nx, ny, nz = _grid(grid)
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    n = min(nGroupsDesired, int(nn / (1 + minNodesPerGroup)))
    L = [nx * float(elem_size[0]), ny * float(elem_size[1]), nz * float(elem_size[2])]
    alpha = (n / (L[0] * L[1] * L[2])) ** (1.0 / 3)
    return tuple(max(round(alpha * v), 1) for v in L)


def _axis_table(n_el, h, nb):
    """Box index of every node along one axis, as the reference computes it
    from the node coordinates: floor((x - min) / (L / nb)), clamped."""
    x = np.arange(n_el + 1, dtype=np.int64) * h
    mn = x.min()
    size = (x.max() - mn) / nb
    t = np.floor((x - mn) / size).astype(np.int32)
    return np.minimum(t, nb - 1).astype(np.int32)


def _ranges(t, nb):
    """Half-open node range of each box along one axis (t is non-decreasing)."""
    lo = np.searchsorted(t, np.arange(nb), side="left")
    hi = np.searchsorted(t, np.arange(nb), side="right")
    return lo, hi


def box_groups(grid, elem_size, boxes, min_nodes=15, device=None):
    """Deflation groups of an nX x nY x nZ box partition with the reference's
    merge of small boxes.  Returns (node_group int32 CUDA tensor, centres
    (G, 3) float64, counts (G,) int64), identical to the host
    DeflationSolver's arrays."""
    nx, ny, nz = _grid(grid)
    nbx, nby, nbz = (int(b) for b in boxes)
    h = [float(v) for v in elem_size]
    dev = _dev(device)
    L = _lib.lib()
    tabs = [_axis_table(n, hh, nb) for n, hh, nb in zip((nx, ny, nz), h, (nbx, nby, nbz))]
    cnt_axis = [np.bincount(t, minlength=nb) for t, nb in zip(tabs, (nbx, nby, nbz))]
    rng = [_ranges(t, nb) for t, nb in zip(tabs, (nbx, nby, nbz))]
    ntent = nbx * nby * nbz
    b = np.arange(ntent)
    bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
    count = (cnt_axis[0][bx] * cnt_axis[1][by] * cnt_axis[2][bz]).astype(np.int64)
    bbox = np.stack([rng[0][0][bx], rng[0][1][bx], rng[1][0][by], rng[1][1][by],
                     rng[2][0][bz], rng[2][1][bz]], axis=1).astype(np.int32)

    def sums(node_group, bb):
        bb_t = torch.from_numpy(np.ascontiguousarray(bb, dtype=np.int32)).to(dev)
        out = torch.empty(3 * bb.shape[0], dtype=torch.float64, device=dev)
        with torch.cuda.device(dev):
            _lib.check(L.pyto_group_sums(nx, ny, nz, h[0], h[1], h[2], _p(node_group),
                                         int(bb.shape[0]), _p(bb_t), _p(out), _stream(dev)),
                       "pyto_group_sums")
        return out.cpu().numpy().reshape(-1, 3)

    center = sums(None, bbox)
    nzc = count > 0
    center[nzc] /= count[nzc, None]
    valid = count >= min_nodes
    n_groups = int(valid.sum())
    if n_groups == 0:
        raise ValueError("no deflation group has %d nodes" % min_nodes)
    mapping = np.full(ntent, -1, dtype=np.int32)
    mapping[valid] = np.arange(n_groups, dtype=np.int32)
    merged = not np.all(valid)
    if merged:
        small = np.where(~valid & (count > 0))[0]
        vidx = np.where(valid)[0]
        if len(small):
            d = np.linalg.norm(center[small][:, None, :] - center[vidx][None, :, :], axis=2)
            mapping[small] = mapping[vidx[np.argmin(d, axis=1)]]
    tt = [torch.from_numpy(t).to(dev) for t in tabs]
    m_t = torch.from_numpy(mapping).to(dev)
    node_group = torch.empty((nx + 1) * (ny + 1) * (nz + 1), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        _lib.check(L.pyto_node_groups(nx, ny, nz, _p(tt[0]), _p(tt[1]), _p(tt[2]), nbx, nby, nbz,
                                      _p(m_t), _p(node_group), _stream(dev)), "pyto_node_groups")
    if merged:
        # counts and node-order centres of the merged groups; each group's
        # sums visit the bounding box of its member boxes
        box_count = count
        member = (mapping >= 0) & (box_count > 0)
        count = np.bincount(mapping[member], weights=box_count[member],
                            minlength=n_groups).astype(np.int64)
        gb = np.zeros((n_groups, 6), dtype=np.int32)
        gb[:, 0::2] = np.iinfo(np.int32).max
        for k in np.nonzero(member)[0]:
            g = mapping[k]
            gb[g, 0::2] = np.minimum(gb[g, 0::2], bbox[k, 0::2])
            gb[g, 1::2] = np.maximum(gb[g, 1::2], bbox[k, 1::2])
        center = sums(node_group, gb) / count[:, None]
    if np.any(count < min_nodes):
        raise ValueError("deflation group with fewer than %d nodes (smallest %d)"
                         % (min_nodes, int(count.min())))
    return node_group, center, count


def element_components(grid, x, threshold=0.01, device=None):
    """Connected components of the elements with x > threshold on the
    device (find_connected_components, src/hex_mesher.py:1285-1339): an
    int32 label per element, -1 for an absent element, otherwise the
    smallest element index of its component."""
    nx, ny, nz = _grid(grid)
    dev = _dev(device)
    xt = torch.as_tensor(x, dtype=torch.float64).to(dev).contiguous().reshape(-1)
    if xt.numel() != nx * ny * nz:
        raise ValueError("x has %d entries, the grid %d elements" % (xt.numel(), nx * ny * nz))
    labels = torch.empty(nx * ny * nz, dtype=torch.int32, device=dev)
    rounds = ctypes.c_int()
    with torch.cuda.device(dev):
        _lib.check(_lib.lib().pyto_element_components(nx, ny, nz, _p(xt), float(threshold),
                                                      _p(labels), ctypes.byref(rounds),
                                                      _stream(dev)),
                   "pyto_element_components")
    return labels


def largest_component(grid, x, threshold=0.01, device=None):
    """(number of components, mask of the largest one) of element_components;
    among equally large components the one with the smallest element index
    wins, as max(components, key=len) picks the first of the reference's
    list (src/topopt_ocm.py:318-325)."""
    labels = element_components(grid, x, threshold, device)
    present = labels >= 0
    if not bool(present.any()):
        return 0, present
    counts = torch.bincount(labels[present].long(), minlength=labels.numel())
    ncomp = int((counts > 0).sum())
    best = int(torch.argmax(counts))      # first maximal count: smallest label
    return ncomp, labels == best
