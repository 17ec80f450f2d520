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


def batch_designs(n_designs=32, seed=5, nx_choices=(32, 48, 64, 80, 96, 112, 128),
                  nz_choices=(16, 32), max_loads=6):
    """The batching workload: n_designs cantilevers with grids (nx, nx/2, nz),
    nx and nz drawn from the choices, each with 1..max_loads unit -z point
    loads on distinct random nodes of the x = max face, densities U[0.01, 1].
    One generator (PCG64, ``seed``) draws, per design in order: nx, nz, the
    number of load cases, the load nodes, then the densities.  Returns a list
    of dicts (grid, mesh, bc, x, loads)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_designs):
        nx = int(rng.choice(nx_choices))
        nz = int(rng.choice(nz_choices))
        grid = (nx, nx // 2, nz)
        mesh, mat, bc = cantilever(grid, load="point")
        nl = int(rng.integers(1, max_loads + 1))
        sx, sy = grid[0] + 1, grid[1] + 1
        face = np.arange(sy * (grid[2] + 1))
        pick = rng.choice(face, nl, replace=False)
        loads = []
        for fn in pick:
            iy, iz = fn % sy, fn // sy
            f = np.zeros(3 * mesh.num_nodes)
            f[3 * (grid[0] + sx * (iy + sy * iz)) + 2] = -1.0
            loads.append(f)
        x = rng.uniform(0.01, 1.0, mesh.num_elems)
        out.append(dict(grid=grid, mesh=mesh, mat=mat, bc=bc, x=x, loads=loads))
    return out
