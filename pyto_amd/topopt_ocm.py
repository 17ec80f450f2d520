"""Optimality-criteria topology optimisation, device resident
(src/topopt_ocm.py:107-345).

``topopt_optimality_criteria`` keeps the reference's signature, loop and
return value ``(sol, history, success, errorMsg, nFEAs)``.  The densities,
the solve, the sensitivities, the sensitivity filter (H (x g))/Hs/x, the
pinning of force elements and the OC update with its volume bisection all
stay on the GPU; the host reads back one objective value and the change per
iteration for the history and the stopping tests.

Extension (not in the reference): ``objective=callable`` replaces the
compliance objective by any differentiable function of the densities built
from :mod:`pyto_amd.autograd` (for example C/C0 + w PN/PN0); its gradient
comes from the hand-written adjoints through ``torch.autograd``.  The OC
update then uses that gradient exactly as it uses dC/dx.
"""
import time

import numpy as np
import torch

from . import topopt_material_model as tmm
from .builders import largest_component
from .topopt_common import FEA_MODE, TO_QOI, createFilters, find_elements_with_forces
from .topopt_obj_cons_sensitivities import compute_objective_and_gradient


def topopt_optimality_criteria(feaMode, fe_solver, to_params, maxIterations=250, move=0.2,
                               move_tol=0.05, rel_conv_tol=1.e-4, print_progress=True,
                               plot_progress=False, debug=False, binarize_topology=True,
                               use_continuation=False, progress_callback=None, plotter=None,
                               objective=None, x0=None):
    """src/topopt_ocm.py:107-345."""
    if objective is None and to_params.Objective[0] != TO_QOI.COMPLIANCE:
        raise ValueError(f"Unsupported objective type: {to_params.Objective[0]}")
    if to_params.Constraints[0][0] != TO_QOI.VOLUME_FRACTION:
        raise ValueError(f"Unsupported constraint type: {to_params.Constraints[0][0]}")
    if feaMode not in (FEA_MODE.STRUCTURAL, FEA_MODE.THERMAL):
        raise ValueError("only the structural and thermal modes are part of the GPU path")
    if plot_progress:
        raise NotImplementedError("plotting is not part of the GPU path")
    volFractionConstraint = to_params.Constraints[0][2]

    def log_message(msg):
        if progress_callback:
            progress_callback(str(msg))
        else:
            print(msg)

    material_model = to_params.materialModel
    tStart = time.time()
    num_elems = fe_solver.mesh.num_elems
    if print_progress:
        log_message("Computing Filters ...")
    H, Hs = createFilters(fe_solver, to_params)
    dev = fe_solver.stiff_mtrx.device
    nDOFPerNode = 3 if feaMode == FEA_MODE.STRUCTURAL else 1
    elemsWithForces = torch.as_tensor(
        find_elements_with_forces(fe_solver.mesh, fe_solver.bc.force, nDOFPerNode),
        dtype=torch.long, device=dev)
    keep = None if to_params.ElemsToKeep is None else torch.as_tensor(
        np.asarray(to_params.ElemsToKeep), dtype=torch.long, device=dev)

    if x0 is None:
        x = torch.full((num_elems,), float(volFractionConstraint), dtype=torch.float64, device=dev)
    else:
        x = torch.as_tensor(x0, dtype=torch.float64).to(dev).clone()
    history = {'objective': [], 'volfrac': [], 'change': []}
    xmin, xmax = 0.001, 1.0
    tmm.initialize_SIMP_STRUCTURAL_PENALTY(1 if use_continuation else 3)

    success = True
    errorMsg = "No errors."
    objScaling = None
    change = 0.0
    iteration = 0
    sol = None
    for iteration in range(maxIterations):
        if progress_callback is not None:
            progress_callback()
        if objective is None:
            sol = fe_solver.solve(x, material_model)
            obj_t, grad_obj = compute_objective_and_gradient(feaMode, to_params, sol, x, fe_solver,
                                                             fe_solver.elem_stiff[0])
            obj = float(obj_t)
        else:
            xr = x.clone().requires_grad_(True)
            J = objective(xr, fe_solver)
            (grad_obj,) = torch.autograd.grad(J, xr)
            obj = float(J.detach())
            sol = fe_solver.sol
        if objScaling is None:
            objScaling = abs(obj)
        obj = obj / objScaling
        grad_obj = grad_obj / objScaling

        # sensitivity filter (src/topopt_ocm.py:239) and pinning (:241-245)
        grad_obj = H.sensitivity_filter(x, grad_obj)
        if elemsWithForces.numel() > 0:
            grad_obj[elemsWithForces] = grad_obj.min()
        if keep is not None:
            grad_obj[keep] = grad_obj.min()

        xold = x
        # the handle may have been rebuilt by the solve (a new material law)
        x = fe_solver.stiff_mtrx.oc_update(x, grad_obj, volFractionConstraint, move=move,
                                           xmin=xmin, xmax=xmax)
        change = float((x - xold).abs().max())
        volfrac_now = float(x.mean())
        volConstraint = volfrac_now / volFractionConstraint - 1.0

        history['objective'].append(obj * objScaling)
        history['volfrac'].append(volfrac_now)
        history['change'].append(change)
        fraction_grey = float(((x > 0.05) & (x < 0.95)).sum()) / num_elems
        if print_progress:
            log_message(f"Iteration: {iteration:d}, obj.: {(obj * objScaling):.3g}, "
                        f"vol.: {volfrac_now:.3g}, grey: {fraction_grey:.3f}")
        if np.isnan(obj):
            log_message("Objective function became NaN. Exiting optimization.")
            errorMsg = "Objective is diverging"
            success = False
            break
        if len(history['objective']) >= 2:
            dJ = abs((history['objective'][-1] - history['objective'][-2]) / history['objective'][-2])
            if (dJ < rel_conv_tol) and (volConstraint < rel_conv_tol) and (change < 0.2) and \
                    (fraction_grey < 0.1):
                log_message("OC optimization converged.")
                break
            if (dJ < rel_conv_tol) and (volConstraint < rel_conv_tol) and (change < move_tol):
                log_message("OC optimization converged.")
                break
        if use_continuation and (iteration % 10 == 0):
            tmm.increment_SIMP_THERMAL_PENALTY(0.5)
            tmm.increment_SIMP_STRUCTURAL_PENALTY(0.5)
    if iteration == maxIterations - 1:
        errorMsg = "Maximum iterations reached"
        log_message(errorMsg)
        success = False
    totalTime = time.time() - tStart

    if binarize_topology:
        # extract a binary topology that keeps the volume fraction
        # (src/topopt_ocm.py:312-316)
        x_sorted = torch.sort(x).values
        threshold = x_sorted[int((1 - float(x.mean())) * num_elems)]
        x = torch.where(x < threshold, torch.zeros_like(x), torch.ones_like(x))
    volfrac = float(x.mean())
    # keep only the largest connected component (src/topopt_ocm.py:317-327),
    # labelled on the device
    fe_solver.mesh.setPseudoDensity(x)
    ncomp, largest = largest_component(fe_solver.mesh.grid, x, 0.01, dev)
    if ncomp > 1:
        if print_progress:
            log_message("Disconnected topology detected. Removing hanging elements.")
        x = largest.to(torch.float64)
        fe_solver.mesh.setPseudoDensity(x)
        volfrac = float(x.mean())
    sol = fe_solver.solve(x, material_model)
    if objective is None:
        obj_t, _ = compute_objective_and_gradient(feaMode, to_params, sol, x, fe_solver,
                                                  fe_solver.elem_stiff[0])
        obj = float(obj_t)
    else:
        with torch.no_grad():
            obj = float(objective(x, fe_solver))
    history['objective'].append(obj)
    history['volfrac'].append(volfrac)
    history['change'].append(change)
    if volfrac > 1.1 * volFractionConstraint:
        errorMsg = f"vf {volFractionConstraint:0.3f} not reached"
        success = False
    nFEAs = iteration + 1
    if print_progress:
        log_message(f"Final objective: {obj:.4g}, vf: {volfrac:.3f}")
        log_message(f"Total Time: {totalTime:.2f} s")
        log_message(f"Error: {errorMsg}")
    return sol, history, success, errorMsg, nFEAs
