"""Write the golden fixtures of tests/golden/ from the oracle (CPU, FP64).

c1.npz  : the 60x20x10 cantilever of config 1 (x = 0.5, unit -z point load
          at node (60,10,5), x=0 clamped, E0 = 1, nu = 0.3, SIMP p = 3,
          eps = 1e-8): displacement from a sparse direct solve, the
          reference-algorithm DPCG iteration count at rtol 1e-8, compliance,
          its gradient, p-norm (p = 6, q = 0.5), von Mises, the cone-filter
          pattern (r = 1.5) and Hs; plus the reference-algorithm DPCG
          iterate at rtol 1e-8 (u_dpcg, C_dpcg): the reference stops there,
          1.5e-6 short of the direct-solve compliance.
tiny.npz: 4x3x2 grid, densities U[0.1,1] seed 3: K (dense), KE, f, u.
c1_oc.npz: the 50-iteration OC loop of config 1 with its tail (see c1_oc).

Provenance: the reference cannot be imported here (numba, pyvista, pyamg,
numpy-stl absent), so these vectors come from the restatement in oracle/,
which tests/test_oracle_*.py pin against closed forms, brute force and FD.
Run: python tests/golden/make_golden.py [c1] [tiny] [c1_oc]
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import deflation as ODF  # noqa: E402
from oracle import fea as OF  # noqa: E402
from oracle import filters as OFl  # noqa: E402
from oracle import ke as OK  # noqa: E402
from oracle import mesh as OM  # noqa: E402
from oracle import sensitivities as OS  # noqa: E402


def c1():
    g = OM.GridMesh((60, 20, 10), (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    x = np.full(g.num_elems, 0.5)
    K = OF.assemble_K(g, KE, OF.simp_scaling(x))
    fixed, gv = OF.clamp_x0(g)
    f = OF.point_load(g, 60, 10, 5)
    A, b, free = OF.impose_dirichlet(K, f, fixed, gv)
    u = OF.scatter_solution(len(f), fixed, gv, free, spla.spsolve(A.tocsc(), b))
    ngn, cnt, cen = ODF.create_groups(g.node_xyz, ODF.recommended_groups_ocm(g.num_nodes))
    W = ODF.deflation_matrix(g.node_xyz, ngn, cen)[free, :]
    xd, info = ODF.deflated_pcg(A, b, W, OF.jacobi_inverse_diag(A), rtol=1e-8)
    u_dpcg = OF.scatter_solution(len(f), fixed, gv, free, xd)
    C_dpcg = f @ u_dpcg
    C, grad = OS.compliance_and_gradient(u, x, g.edofMatStructural, KE)
    post = OF.postprocess(g, u, x, 1.0, 0.3, q=0.5, pexp=6.0)
    H, Hs = OFl.cone_filter(g, 1.5)
    H = H.tocsr()
    H.sort_indices()
    np.savez_compressed(os.path.join(HERE, "c1.npz"), u=u, C=C, grad=grad,
                        pnorm=post["pNormStress"], vm=post["vonMisesStress"],
                        dpcg_iterations=info["iterations"], u_dpcg=u_dpcg, C_dpcg=C_dpcg,
                        node_group=ngn, group_center=cen,
                        H_indptr=H.indptr, H_indices=H.indices, H_data=H.data, Hs=Hs)


def c1_oc():
    """The 50-iteration OC loop of config 1 (60x20x10, x0 = 0.5, unit -z
    point load at node (60,10,5), x = 0 clamped, E0 = 1, nu = 0.3, p = 3,
    Emin = 1e-9, sensitivity filter r = 1.5, volume fraction 0.5, move 0.2)
    with its tail (binarise, largest connected component, final solve):
    once with sparse direct solves ("exact"), once with the reference's
    DPCG iteration to rtol 1e-8 and the exact Galerkin coarse matrix,
    refactorised by spsolve at every coarse solve ("dpcg").  With the
    reference's own coarse shift 1e-10 the DPCG of this workload diverges
    once void elements (E = 2e-9) appear (it runs to maxIters from the third
    OC iteration on), so the shift is left out here."""
    from oracle import topopt as OT
    g = OM.GridMesh((60, 20, 10), (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    fixed, _ = OF.clamp_x0(g)
    f = OF.point_load(g, 60, 10, 5)
    out = {}
    for name in ("exact", "dpcg"):
        if name == "exact":
            ms = OT.direct_make_solve(g, KE, fixed, eps=1e-9)
        else:
            ms = OT.reference_dpcg_solver(g, KE, fixed, eps=1e-9, rtol=1e-8, shift=0.0,
                                          coarse="spsolve")
        r = OT.oc_loop(g, KE, f, fixed, 0.5, ms, max_iterations=50, eps=1e-9)
        h = r["history"]
        out[name + "_objective"] = np.array(h["objective"])
        out[name + "_volfrac"] = np.array(h["volfrac"])
        out[name + "_change"] = np.array(h["change"])
        out[name + "_x_loop"] = r["x_loop"]
        out[name + "_x_final"] = r["x_final"].astype(np.uint8)
        out[name + "_components"] = r["components"]
        out[name + "_nFEAs"] = r["nFEAs"]
        if name == "dpcg":
            out["dpcg_iterations"] = np.array(ms.stats["iterations"])
        print(name, "final objective", h["objective"][-1], "components", r["components"],
              flush=True)
    np.savez_compressed(os.path.join(HERE, "c1_oc.npz"), **out)


def tiny():
    g = OM.GridMesh((4, 3, 2), (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    x = np.random.default_rng(3).uniform(0.1, 1.0, g.num_elems)
    K = OF.assemble_K(g, KE, OF.simp_scaling(x))
    fixed, gv = OF.clamp_x0(g)
    f = OF.face_traction_load(g)
    A, b, free = OF.impose_dirichlet(K, f, fixed, gv)
    u = OF.scatter_solution(len(f), fixed, gv, free, spla.spsolve(A.tocsc(), b))
    np.savez_compressed(os.path.join(HERE, "tiny.npz"), K=K.toarray(), KE=KE, x=x, f=f, u=u,
                        fixed=fixed)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c1", "tiny", "c1_oc"]
    for name in which:
        {"c1": c1, "tiny": tiny, "c1_oc": c1_oc}[name]()
    print("wrote", os.listdir(HERE))
