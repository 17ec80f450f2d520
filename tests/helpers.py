"""Shared builders for the GPU parity tests: the same seeded problem built
once through the oracle (assembled CSR, FP64) and once as a GridOperator."""
import numpy as np
import scipy.sparse as sp

from oracle import deflation as ODF
from oracle import fea as OF
from oracle import ke as OK
from oracle import mesh as OM


class Case:
    def __init__(self, dims, seed=0, lo=0.01, hi=1.0, load="traction", h=(1.0, 1.0, 1.0),
                 E0=1.0, nu=0.3, penal=3.0, eps=1e-8, x=None, groups="reference"):
        self.dims = dims
        self.mesh = OM.GridMesh(dims, h)
        self.mesh.createEdofMatStructural()
        g = self.mesh
        self.KE = OK.hex8_stiffness(E0, nu, h)
        self.x = np.random.default_rng(seed).uniform(lo, hi, g.num_elems) if x is None else x
        self.E = OF.simp_scaling(self.x, penal, eps)
        self.fixed, self.gv = OF.clamp_x0(g)
        if load == "traction":
            self.f = OF.face_traction_load(g)
        else:
            self.f = OF.point_load(g, *load)
        self.K = OF.assemble_K(g, self.KE, self.E)
        self.free = OF.free_dofs(3 * g.num_nodes, self.fixed)
        self.mask = np.zeros(3 * g.num_nodes)
        self.mask[self.free] = 1.0
        self.E0, self.nu, self.penal, self.eps, self.h = E0, nu, penal, eps, h
        self.ngn = None
        if groups == "reference" and g.num_nodes >= 2 * 16:
            ng = ODF.recommended_groups_ocm(g.num_nodes)
            self.ngn, self.cnt, self.cen = ODF.create_groups(g.node_xyz, ng)

    def Kff_full(self, u):
        """K_ff u_f embedded in full length (fixed entries 0)."""
        return self.mask * (self.K @ (self.mask * u))

    def W_free(self):
        W = ODF.deflation_matrix(self.mesh.node_xyz, self.ngn, self.cen)
        return sp.diags(self.mask) @ W

    def operator(self, device, groups=True, shift=1e-10):
        from pyto_amd.operator import GridOperator
        op = GridOperator(self.dims, self.h, self.E0, self.nu, self.fixed, penal=self.penal,
                          eps_void=self.eps)
        op.set_coarse_shift(shift, 0.0)
        if groups and self.ngn is not None:
            op.set_groups(self.ngn, self.cen)
        op.set_density(self.x)
        return op


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
