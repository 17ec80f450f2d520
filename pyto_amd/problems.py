"""Voxel cantilever workloads: the grids, loads, supports and seeded density
fields of the benchmark configurations.

Every workload clamps all DOFs of the x = 0 face (src/hex_structural_examples.py:353-356)
and uses unit bricks, E0 = 1, nu = 0.3, SIMP p = 3.  Loads: a unit -z point
load, or a unit total -z traction on the x = max face with the reference's
interior/edge/corner weights 4/2/1 (src/hex_structural_examples.py:360-395).
Densities come from numpy's PCG64 ``default_rng(seed)`` in element order.
"""
import numpy as np

from . import bound_cond, mat_lib
from .hex_mesher import HexMesher


def cantilever(num_elems, elem_size=(1.0, 1.0, 1.0), load="traction", point=None,
               E0=1.0, nu=0.3, total=1.0):
    mesh = HexMesher().grid_mesh(num_elems, elem_size)
    mesh.createEdofMatStructural()
    nelx, nely, nelz = mesh.grid
    sx, sy = nelx + 1, nely + 1
    fixed_nodes = mesh.getNodesOnBoundingBoxPlane(0, True)
    fixed = np.sort(np.concatenate([3 * fixed_nodes, 3 * fixed_nodes + 1, 3 * fixed_nodes + 2]))
    f = np.zeros(3 * mesh.num_nodes)
    if load == "traction":
        nodes = mesh.getNodesOnBoundingBoxPlane(0, False)
        iy = (nodes // sx) % sy
        iz = nodes // (sx * sy)
        ext = ((iy == 0) | (iy == nely)).astype(int) + ((iz == 0) | (iz == nelz)).astype(int)
        w = np.where(ext == 2, 1.0, np.where(ext == 1, 2.0, 4.0))
        f[3 * nodes + 2] = -w * (total / w.sum())
    elif load == "point":
        ix, iy, iz = point if point is not None else (nelx, nely // 2, nelz // 2)
        f[3 * (ix + sx * (iy + sy * iz)) + 2] = -total
    else:
        raise ValueError("load must be 'traction' or 'point'")
    bc = bound_cond.BC(force=f, fixed_dofs=fixed, dirichlet_values=np.zeros(fixed.size))
    mat = mat_lib.Material("Unit", E0, nu)
    return mesh, mat, bc


def density_uniform(num_elems, lo, hi, seed):
    return np.random.default_rng(seed).uniform(lo, hi, num_elems)


def density_bumps(grid, seed=2, n_bumps=16, lo=0.001):
    """Sum of Gaussian bumps on the element centres, normalised by its max and
    clipped to [lo, 1]: centres ~ U(box), sigma ~ U[0.05,0.25]*min(L),
    amplitudes ~ U[0.5,1]."""
    nelx, nely, nelz = grid
    rng = np.random.default_rng(seed)
    L = np.array([nelx, nely, nelz], float)
    cen = rng.uniform(0, 1, (n_bumps, 3)) * L
    sig = rng.uniform(0.05, 0.25, n_bumps) * L.min()
    amp = rng.uniform(0.5, 1.0, n_bumps)
    cx = np.arange(nelx) + 0.5
    cy = np.arange(nely) + 0.5
    cz = np.arange(nelz) + 0.5
    field = np.zeros((nelz, nely, nelx))
    for k in range(n_bumps):
        gx = np.exp(-(cx - cen[k, 0]) ** 2 / (2 * sig[k] ** 2))
        gy = np.exp(-(cy - cen[k, 1]) ** 2 / (2 * sig[k] ** 2))
        gz = np.exp(-(cz - cen[k, 2]) ** 2 / (2 * sig[k] ** 2))
        field += amp[k] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    field /= field.max()
    return np.clip(field, lo, 1.0).ravel()


# Benchmark configurations (grid, load, density recipe)
CONFIGS = {
    "c1": dict(grid=(60, 20, 10), load="point", point=(60, 10, 5)),
    "c2": dict(grid=(192, 64, 64), load="traction"),
    "c3": dict(grid=(256, 128, 128), load="traction"),
    "c4": dict(grid=(128, 64, 64), load="traction"),
}


def config_density(name, grid):
    ne = grid[0] * grid[1] * grid[2]
    if name == "c1":
        return np.full(ne, 0.5)
    if name == "c2":
        return density_uniform(ne, 0.01, 1.0, 0)
    if name == "c3":
        return density_bumps(grid, seed=2)
    if name == "c4":
        return density_uniform(ne, 0.2, 1.0, 1)
    raise KeyError(name)
