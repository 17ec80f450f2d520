"""Assembled-matrix structural FEA (oracle; test infrastructure only).

Restates:
  * SIMP scaling E = eps + x^p (1-eps) and its derivative,
    src/topopt_material_model.py:76-95,119-139 (eps, p parametrised; the
    reference hard-codes eps = 1e-8 at :18 and p = 3 at :13);
  * COO assembly of KE (x) E_e with node_idx = kron(edofMat, ones),
    src/hex_structural_fea.py:47-50,87-102;
  * Dirichlet elimination b -= A[:,fixed] g; A[free][:,free],
    src/bound_cond.py:56-69; u[fixed] = g, u[free] = x,
    src/linear_solvers.py:199-202;
  * Jacobi diag^{-1} with |d| < 1e-12 -> 1, src/linear_solvers.py:71-88;
  * centre-point stresses, von Mises and p-norm,
    src/hex_structural_fea.py:152-234 (relaxation q at
    src/topopt_material_model.py:15,199-201; p at :26);
  * the thermal conductivity law, K_T assembly and centre temperature
    gradient, src/topopt_material_model.py:98-157,
    src/hex_thermal_fea.py:47-50,85-112,256-270.
"""
import numpy as np
import scipy.sparse as sp

from . import ke as ke_mod


def simp_scaling(x, penal=3.0, eps=1e-8):
    """src/topopt_material_model.py:88."""
    return eps + (x ** penal) * (1.0 - eps)


def simp_sensitivity(x, penal=3.0, eps=1e-8):
    """src/topopt_material_model.py:131."""
    return penal * (x ** (penal - 1.0)) * (1.0 - eps)


def thermal_scaling(x, penal=1.0, kvoid=1e-8):
    """src/topopt_material_model.py:106-109 (SIMP; _KVOID_RELATIVE at :19,
    _SIMP_THERMAL_PENALTY at :14)."""
    return kvoid + (x ** penal) * (1.0 - kvoid)


def thermal_sensitivity(x, penal=1.0, kvoid=1e-8):
    """src/topopt_material_model.py:150-151."""
    return penal * (x ** (penal - 1.0)) * (1.0 - kvoid)


def assemble_KT(mesh, KE_T, k_e):
    """Global K_T as CSR; src/hex_thermal_fea.py:47-50,85-112 + .tocsr()."""
    edof = mesh.elemArray.astype(np.int64)
    nn = mesh.num_nodes
    iK = np.kron(edof, np.ones((8, 1), dtype=np.int64)).ravel()
    jK = np.kron(edof, np.ones((1, 8), dtype=np.int64)).ravel()
    vals = np.einsum("ij,e->eij", KE_T, k_e).ravel(order="C")
    return sp.coo_matrix((vals, (iK, jK)), shape=(nn, nn)).tocsr()


def thermal_gradient(mesh, T):
    """Centre gradient in reference coordinates, gradN @ T[edof].T with
    gradN = (1/8) [+-1 signs]; src/hex_thermal_fea.py:256-270."""
    gradN = 0.125 * np.array([[-1, 1, 1, -1, -1, 1, 1, -1],
                              [-1, -1, 1, 1, -1, -1, 1, 1],
                              [-1, -1, -1, -1, 1, 1, 1, 1]], dtype=np.float64)
    return gradN @ T[mesh.elemArray].T


def assemble_K(mesh, KE, E_e):
    """Global K as CSR; src/hex_structural_fea.py:47-50,89-102 + .tocsr()."""
    edof = mesh.edofMatStructural
    ndof = 3 * mesh.num_nodes
    iK = np.kron(edof, np.ones((24, 1), dtype=np.int64)).ravel()
    jK = np.kron(edof, np.ones((1, 24), dtype=np.int64)).ravel()
    vals = np.einsum("ij,e->eij", KE, E_e).ravel(order="C")
    return sp.coo_matrix((vals, (iK, jK)), shape=(ndof, ndof)).tocsr()


def free_dofs(ndof, fixed_dofs):
    """src/bound_cond.py:31-32."""
    return np.setdiff1d(np.arange(ndof), fixed_dofs)


def impose_dirichlet(A, b, fixed_dofs, dirichlet_values):
    """src/bound_cond.py:56-69."""
    free = free_dofs(A.shape[0], fixed_dofs)
    bm = b.copy()
    if len(fixed_dofs):
        bm -= A[:, fixed_dofs] @ dirichlet_values
    Am = A[free, :][:, free]
    return Am.tocsr(), bm[free], free


def jacobi_inverse_diag(A, eps_tol=1e-12):
    """src/linear_solvers.py:84-88 (returned as the diagonal vector)."""
    d = A.diagonal().copy()
    d[np.abs(d) < eps_tol] = 1.0
    return 1.0 / d


def scatter_solution(ndof, fixed_dofs, dirichlet_values, free, x):
    """src/linear_solvers.py:199-202."""
    u = np.zeros(ndof)
    u[fixed_dofs] = dirichlet_values
    u[free] = x
    return u


def centre_gradN(elem_size):
    """src/hex_structural_fea.py:163-169."""
    g = (1.0 / 8.0) * np.array([[-1, 1, 1, -1, -1, 1, 1, -1],
                                [-1, -1, 1, 1, -1, -1, 1, 1],
                                [-1, -1, -1, -1, 1, 1, 1, 1]], dtype=np.float64)
    for i in range(3):
        g[i, :] = 2 * g[i, :] / elem_size[i]
    return g


def postprocess(mesh, sol, x, E, nu, q=0.5, pexp=6.0):
    """Centre-point strains/stresses/von Mises/p-norm.

    src/hex_structural_fea.py:163-232.  Strain order [xx, yy, zz, xy, xz, yz]
    (:178-183).  Returns a dict with the attribute names the reference sets.
    """
    gradN = centre_gradN(mesh.elem_size)
    edof = mesh.edofMatStructural
    uG = gradN @ sol[edof[:, 0::3]].T
    vG = gradN @ sol[edof[:, 1::3]].T
    wG = gradN @ sol[edof[:, 2::3]].T
    strain = np.stack([uG[0], vG[1], wG[2], uG[1] + vG[0], uG[2] + wG[0],
                       vG[2] + wG[1]], axis=1)
    D = ke_mod.isotropic_D(E, nu)
    stress = np.einsum("ij,ej->ei", D, strain)
    stress *= (x ** q).reshape(-1, 1)
    s = stress
    vm = np.sqrt(0.5 * ((s[:, 0] - s[:, 1]) ** 2 + (s[:, 1] - s[:, 2]) ** 2 +
                        (s[:, 2] - s[:, 0]) ** 2) +
                 3 * (s[:, 3] ** 2 + s[:, 4] ** 2 + s[:, 5] ** 2))
    pn = np.sum(vm ** pexp) ** (1.0 / pexp)
    energy = 0.5 * np.sum(strain * stress, axis=1)
    return {"strainComponents": strain, "stressComponents": stress,
            "vonMisesStress": vm, "pNormStress": pn, "elemStrainEnergy": energy}


# ---------------------------------------------------------------------------
# Load / BC recipes
# ---------------------------------------------------------------------------
def clamp_x0(mesh):
    """All DOFs of the x = 0 plane; src/hex_structural_examples.py:353-356."""
    nodes = mesh.getNodesOnBoundingBoxPlane(0, True)
    fixed = np.array([3 * nodes, 3 * nodes + 1, 3 * nodes + 2]).ravel().astype(np.int64)
    return np.sort(fixed), np.zeros(fixed.size)


def face_traction_load(mesh, total=1.0):
    """-z load on the x = max face, weights interior 4 / edge 2 / corner 1,
    normalised to |total|; src/hex_structural_examples.py:360-395."""
    nelx, nely, nelz = mesh.grid
    nodes = mesh.getNodesOnBoundingBoxPlane(0, False)
    iy = mesh.node_indices[nodes, 1]
    iz = mesh.node_indices[nodes, 2]
    ext = ((iy == 0) | (iy == nely)).astype(int) + ((iz == 0) | (iz == nelz)).astype(int)
    w = np.where(ext == 2, 1.0, np.where(ext == 1, 2.0, 4.0))
    f = np.zeros(3 * mesh.num_nodes)
    f[3 * nodes + 2] = -w
    f[3 * nodes + 2] *= total / abs(np.sum(f[3 * nodes + 2]))
    return f


def point_load(mesh, ix, iy, iz, value=-1.0, comp=2):
    """Unit point load at grid node (ix, iy, iz)."""
    nelx, nely, _ = mesh.grid
    n = ix + (nelx + 1) * iy + (nelx + 1) * (nely + 1) * iz
    f = np.zeros(3 * mesh.num_nodes)
    f[3 * n + comp] = value
    return f
