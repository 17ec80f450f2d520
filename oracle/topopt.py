"""Optimality-criteria loop and its clean-up tail (oracle; test infrastructure
only).

Restates:
  * ``find_elements_with_forces``, src/topopt_common.py:79-100: elements with
    a corner node that carries a non-zero load component;
  * the 27-entry element neighbour array, src/hex_mesher.py:126-151 (the
    element itself and its face, edge and corner neighbours; -1 outside the
    grid), and ``find_connected_components``, src/hex_mesher.py:1285-1339:
    elements with density > threshold, joined through that array, searched
    depth-first from each unvisited element in element order (so components
    come ordered by their lowest element index);
  * ``topopt_optimality_criteria``, src/topopt_ocm.py:107-345, for the
    compliance objective with one volume constraint: objective scaled by its
    first value, sensitivity filter (H (x g))/Hs/x, force elements pinned to
    min(g), the bisection OC update, the two stopping tests, then (after the
    loop) the binary threshold at the volume fraction, the largest connected
    component and a final solve whose compliance is appended to the history.

``solver`` is a callable make_solve(x) -> solve(b) -> u (full length, u = 0
on the fixed DOFs): a sparse direct solve (oracle.sensitivities.
direct_solver) or the reference-form DPCG (``reference_dpcg_solver``).
"""
import numpy as np
import scipy.sparse as sp

from . import deflation as ODF
from . import fea
from . import filters as OFl
from . import sensitivities as OS


def elements_with_forces(mesh, force, ndof_per_node):
    """src/topopt_common.py:79-100."""
    forced_nodes = np.unique(np.nonzero(np.asarray(force) != 0)[0] // ndof_per_node)
    return np.nonzero(np.isin(mesh.elemArray, forced_nodes).any(axis=1))[0]


def element_neighbours(dims):
    """src/hex_mesher.py:126-151: (N_el, 27) int32, offsets dz, dy, dx in
    {-1, 0, 1} with dx fastest; -1 where the neighbour is outside the grid."""
    nx, ny, nz = dims
    e = np.arange(nx * ny * nz)
    ix, iy, iz = e % nx, (e // nx) % ny, e // (nx * ny)
    out = np.empty((e.size, 27), dtype=np.int64)
    k = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                jx, jy, jz = ix + dx, iy + dy, iz + dz
                inside = (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < ny) & (jz >= 0) & (jz < nz)
                out[:, k] = np.where(inside, e + dz * nx * ny + dy * nx + dx, -1)
                k += 1
    return out.astype(np.int32)


def find_connected_components(dims, density, threshold=0.01):
    """src/hex_mesher.py:1285-1339: list of element-index arrays, one per
    connected set of elements with density > threshold, in the order the
    reference's depth-first search finds them (each array in visiting
    order)."""
    nbr = element_neighbours(dims)
    present = np.asarray(density) > threshold
    visited = np.zeros(present.size, dtype=bool)
    comps = []
    for e0 in range(present.size):
        if not present[e0] or visited[e0]:
            continue
        comp = []
        stack = [e0]
        while stack:
            e = stack.pop()
            if visited[e]:
                continue
            visited[e] = True
            comp.append(e)
            for n in nbr[e]:
                if n >= 0 and n != e and present[n] and not visited[n]:
                    stack.append(int(n))
        comps.append(np.array(comp))
    return comps


def reference_dpcg_solver(mesh, KE, fixed, penal=3.0, eps=1e-8, rtol=1e-8, maxIters=2000,
                          n_groups=None, shift=1e-10, coarse="spsolve"):
    """make_solve(x) in the reference's form: COO assembly and CSR
    conversion per solve (src/hex_structural_fea.py:87-102), Dirichlet
    elimination by slicing (src/bound_cond.py:56-69), Jacobi, the reference
    group heuristic and its W, and the two-SpMV DPCG with a SuperLU
    factorisation of the coarse matrix on every coarse solve
    (src/deflation.py:504-584, ``coarse="spsolve"``).  shift = 0 and
    coarse = "splu" with a tight rtol give the same iteration with the exact
    Galerkin matrix, factorised once per solve."""
    nn = mesh.num_nodes
    if n_groups is None:
        n_groups = ODF.recommended_groups_ocm(nn)
    ngn, _, cen = ODF.create_groups(mesh.node_xyz, n_groups)
    W_all = ODF.deflation_matrix(mesh.node_xyz, ngn, cen)
    fixed = np.asarray(fixed)
    gv = np.zeros(fixed.size)
    stats = {"iterations": []}

    def make_solve(x):
        K = fea.assemble_K(mesh, KE, fea.simp_scaling(x, penal, eps))

        def solve(b):
            A, bb, free = fea.impose_dirichlet(K, b, fixed, gv)
            Minv = fea.jacobi_inverse_diag(A)
            W = W_all[free, :]
            xf, info = ODF.deflated_pcg(A, bb, W, Minv, rtol=rtol, maxIters=maxIters,
                                        shift=shift, coarse=coarse)
            stats["iterations"].append(info["iterations"])
            return fea.scatter_solution(len(b), fixed, gv, free, xf)
        return solve
    make_solve.stats = stats
    return make_solve


def oc_loop(mesh, KE, f, fixed, volfrac, make_solve, max_iterations=50, rel_filter_radius=1.5,
            penal=3.0, eps=1e-8, move=0.2, move_tol=0.05, rel_conv_tol=1e-4,
            binarize_topology=True):
    """src/topopt_ocm.py:200-345 for the compliance objective.  Returns a
    dict with the history (objective, volfrac, change; one more entry for
    the final solve), the densities before the tail (x_loop), the final
    design (x_final), the number of components found and nFEAs."""
    mesh.createEdofMatStructural()
    edof = mesh.edofMatStructural
    H, Hs = OFl.cone_filter(mesh, rel_filter_radius)
    H = H.tocsr()
    forced = elements_with_forces(mesh, f, 3)
    n = mesh.num_elems
    x = np.full(n, float(volfrac))
    hist = {"objective": [], "volfrac": [], "change": []}
    scale = None
    change = 0.0
    it = 0
    for it in range(max_iterations):
        u = make_solve(x)(f)
        C, g = OS.compliance_and_gradient(u, x, edof, KE, penal, eps)
        if scale is None:
            scale = abs(C)
        obj, g = C / scale, g / scale
        g = OFl.sensitivity_filter(H, Hs, x, g)
        if forced.size:
            g[forced] = g.min()
        xold = x
        x = OFl.oc_update(x, g, volfrac, move=move)
        change = float(np.max(np.abs(x - xold)))
        vol = float(np.mean(x)) / volfrac - 1.0
        hist["objective"].append(obj * scale)
        hist["volfrac"].append(float(np.mean(x)))
        hist["change"].append(change)
        grey = np.sum((x > 0.05) & (x < 0.95)) / n
        if np.isnan(obj):
            break
        if len(hist["objective"]) >= 2:
            dJ = abs((hist["objective"][-1] - hist["objective"][-2]) / hist["objective"][-2])
            if dJ < rel_conv_tol and vol < rel_conv_tol and change < 0.2 and grey < 0.1:
                break
            if dJ < rel_conv_tol and vol < rel_conv_tol and change < move_tol:
                break
    x_loop = x.copy()
    if binarize_topology:
        threshold = np.sort(x)[int((1 - np.mean(x)) * len(x))]
        x = np.where(x < threshold, 0.0, 1.0)
    volfrac_final = float(np.mean(x))
    comps = find_connected_components(mesh.grid, x)
    if len(comps) > 1:
        largest = max(comps, key=len)
        x = np.zeros(n)
        x[largest] = 1.0
        volfrac_final = float(np.mean(x))
    u = make_solve(x)(f)
    C, _ = OS.compliance_and_gradient(u, x, edof, KE, penal, eps)
    hist["objective"].append(C)
    hist["volfrac"].append(volfrac_final)
    hist["change"].append(change)
    return {"history": hist, "x_loop": x_loop, "x_final": x, "components": len(comps),
            "nFEAs": it + 1, "u_final": u}


def direct_make_solve(mesh, KE, fixed, penal=3.0, eps=1e-8):
    """make_solve(x) with a sparse direct solve (exact up to rounding)."""
    fixed = np.asarray(fixed)
    gv = np.zeros(fixed.size)

    def make_solve(x):
        solve, _ = OS.direct_solver(mesh, KE, x, fixed, gv, penal, eps)
        return solve
    return make_solve


__all__ = ["elements_with_forces", "element_neighbours", "find_connected_components",
           "reference_dpcg_solver", "direct_make_solve", "oc_loop", "sp"]
