"""Differentiable structural responses on the GPU.

The FEA solve sits behind ``torch.autograd.Function`` with hand-written
adjoints, so topology-optimisation objectives compose with ordinary torch
code:

* ``responses(x, fe, pnorm=True) -> (C, PN)``: compliance and the von Mises
  p-norm from ONE forward solve.  Backward: compliance is self-adjoint
  (lambda = u, no extra solve, src/topopt_obj_cons_sensitivities.py:148-179)
  when every Dirichlet value is zero; with prescribed displacements it takes
  one adjoint solve K_ff lambda_f = f_f;
  the p-norm runs one adjoint deflated-PCG solve K lambda = g with the same
  operator and boundary conditions (:181-287, intended math with g_e as the
  24-vector).  Both terms share one adjoint when both are requested:
  dL/dx = gC dC/dx + gP (T1 + T2).
* ``displacement(x, fe) -> u``: the solve itself; backward solves
  K lambda = dL/du and returns dL/dx_e = -E'(x_e) lambda_e^T KE u_e.

Nothing element-wise is stored between forward and backward apart from u,
x and the von Mises field: the backward kernels recompute the element
strains from u.  Every adjoint solve uses the forward problem's fixed DOFs
with zero values.  ``fe`` is a ``HexStructuralFEA``; x is a float64 CUDA
tensor of element densities.
"""
import torch

from . import bound_cond, linear_solvers
from . import topopt_material_model as tmm


def _ensure_density(op, x):
    """The operator holds one density state; restore x if another solve
    replaced it between forward and backward."""
    if op.x is None or op.x.data_ptr() != x.data_ptr() or not torch.equal(op.x, x):
        op.set_density(x)


def _adjoint_solve(fe, rhs):
    """K_ff lambda_f = rhs_f with lambda = 0 on every fixed DOF (homogeneous
    Dirichlet values whatever the forward problem prescribes)."""
    op = fe.stiff_mtrx
    return linear_solvers.solve(op, rhs, fe.solver, bound_cond.homogeneous(fe.bc),
                                dsolver=fe.dsolver, **fe.kwargs)


def _prescribed(fe):
    """True when the forward problem prescribes non-zero displacements."""
    return bound_cond.dirichlet_vector(fe.bc, fe.stiff_mtrx.num_dofs) is not None


class _Responses(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, fe, want_pnorm):
        xd = x.detach().contiguous()
        u = fe.solve(xd)
        op = fe.stiff_mtrx
        C, _, _ = op.compliance(u, xd, want_grad=False)
        q = tmm.get_stress_relaxation_factor()
        p = tmm.get_pNorm_exponent()
        if want_pnorm:
            st = op.stress(u, xd, q=q, pexp=p, want_strain=False, want_stress=False,
                           want_energy=False)
            vm, svp = st["vm"], st["sum_vm_p"]
            PN = svp[0] ** (1.0 / p)
        else:
            vm = svp = None
            PN = torch.zeros((), dtype=torch.float64, device=x.device)
        ctx.fe, ctx.q, ctx.p, ctx.want_pnorm = fe, q, p, want_pnorm
        ctx.save_for_backward(xd, u, vm if vm is not None else u[:0],
                              svp if svp is not None else u[:0])
        return C.reshape(()), PN.reshape(())

    @staticmethod
    def backward(ctx, gC, gP):
        xd, u, vm, svp = ctx.saved_tensors
        fe = ctx.fe
        op = fe.stiff_mtrx
        _ensure_density(op, xd)
        grad = torch.zeros_like(xd)
        if gC is not None and bool(gC != 0):
            _, dC, _ = op.compliance(u, xd)           # -E'(x) u_e^T KE u_e
            if _prescribed(fe):
                # C = u^T K u with u = g on the fixed DOFs: the implicit term
                # is 2 (K u)_f^T du_f/dx = -2 lam^T K' u with K_ff lam_f = f_f
                # (zero on the fixed DOFs), so dC/dx_e = E'(x_e) (ce_e -
                # 2 lam_e^T KE u_e); with g = 0, lam = u and it is -E' ce.
                lam = _adjoint_solve(fe, op.as_dof(fe.total_force))
                dE = tmm.get_structural_material_model_sensitivity(xd, tmm.MaterialModel.SIMP)
                dC = -dC - 2.0 * dE * op.bilinear(lam, u)
            grad += gC * dC
        if ctx.want_pnorm and gP is not None and bool(gP != 0):
            g, T1 = op.pnorm_adjoint_rhs(u, xd, vm, svp, q=ctx.q, pexp=ctx.p)
            lam = _adjoint_solve(fe, g)
            xc = torch.clamp(xd, min=1e-12)
            T2 = -tmm.get_structural_material_model_sensitivity(xc, tmm.MaterialModel.SIMP) \
                * op.bilinear(lam, u)
            grad += gP * (T1 + T2)
        return grad, None, None


class _Displacement(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, fe):
        xd = x.detach().contiguous()
        u = fe.solve(xd).clone()
        ctx.fe = fe
        ctx.save_for_backward(xd, u)
        return u

    @staticmethod
    def backward(ctx, gu):
        xd, u = ctx.saved_tensors
        fe = ctx.fe
        op = fe.stiff_mtrx
        _ensure_density(op, xd)
        lam = _adjoint_solve(fe, gu.contiguous())
        dE = tmm.get_structural_material_model_sensitivity(xd, tmm.MaterialModel.SIMP)
        return -dE * op.bilinear(lam, u), None


def responses(x, fe, pnorm=True):
    """(C, PN) of density x; differentiable in x (see the module docstring)."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float64:
        raise TypeError("x must be a float64 CUDA tensor of element densities")
    return _Responses.apply(x, fe, bool(pnorm))


def compliance(x, fe):
    """C(x) = u^T K(x) u (= f^T u when every Dirichlet value is zero);
    differentiable in x (self-adjoint backward, or one adjoint solve with
    prescribed displacements)."""
    return responses(x, fe, pnorm=False)[0]


def pnorm_stress(x, fe):
    """PN(x) = (sum vm_e^p)^(1/p) with relaxed stresses; differentiable in x
    (one adjoint solve in backward)."""
    return responses(x, fe, pnorm=True)[1]


def displacement(x, fe):
    """u(x) = K(x)^{-1} f; differentiable in x (one adjoint solve in backward)."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float64:
        raise TypeError("x must be a float64 CUDA tensor of element densities")
    return _Displacement.apply(x, fe)
