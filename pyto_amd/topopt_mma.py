"""Density-based topology optimisation with MMA, device resident
(src/topopt_mma.py:146-469).

``topopt_mma`` keeps the reference's signature, its optimisation function
(density filter, solve, objective and constraints with their sensitivities,
normalisation by the first objective, pinning of force and keep elements,
the filter chain rule back to the design variables) and its return value
``(sol, history, success, errorMsg, nFEAs)``.  The MMA step runs on the GPU
(:mod:`pyto_amd.mmaWrapper`); densities, gradients and constraint Jacobians
never leave the device, only the objective and the m constraint values do.

Not on this path: plotting and progress plots (``plot_progress`` raises),
the post-optimisation removal of hanging elements
(``Eliminate_Hanging_Elements`` binarises like ``binarize_topology``; the
connected-component search is a host mesh utility), element body forces
(the structural solver rejects them), and ``postprocess`` inside the loop
(stress fields for plots; the p-norm sensitivities compute what they need).
"""
import time

import numpy as np
import torch

from . import topopt_material_model as tmm
from .mmaWrapper import runMMA
from .topopt_common import FEA_MODE, TO_QOI, createFilters, find_elements_with_forces
from .topopt_obj_cons_sensitivities import (compute_constraint_and_gradient,
                                            compute_objective_and_gradient,
                                            compute_thermoelastic_compliance_and_gradient)


def topopt_mma(feaMode, fe_structural_solver, fe_thermal_solver, to_params, maxiteration=300,
               timeLimitSecs=36000, move_limit=0.05, kkt_tol=1.0e-6, objective_tol=1.0e-3,
               constraint_tol=1.0e-3, print_progress=True, plot_progress=False,
               use_continuation=False, binarize_topology=True, progress_callback=None,
               plotter=None, x0=None):
    """src/topopt_mma.py:146-469.  ``x0`` (not in the reference) starts from
    given design variables instead of the volume fraction."""
    def log_message(msg):
        if progress_callback:
            progress_callback(str(msg))
        else:
            print(msg)

    if feaMode == FEA_MODE.STRUCTURAL:
        fe_solver, nDOFPerNode = fe_structural_solver, 3
    elif feaMode == FEA_MODE.THERMAL:
        fe_solver, nDOFPerNode = fe_thermal_solver, 1
    elif feaMode == FEA_MODE.THERMO_STRUCTURAL:
        fe_solver, nDOFPerNode = fe_structural_solver, 3
        if fe_structural_solver.mesh.num_elems != fe_thermal_solver.mesh.num_elems:
            raise ValueError("Structural and thermal meshes must have the same number of elements.")
        if not use_continuation:
            print("Continuation is recommended for thermo-structural problems.")
    else:
        raise ValueError("Invalid FEA mode.")
    if plot_progress:
        raise NotImplementedError("plotting is not part of the GPU path")
    mesh = fe_solver.mesh
    mat_prop = fe_solver.mat_prop
    material_model = to_params.materialModel
    tStart = time.time()
    num_elems = mesh.num_elems

    history = {"objective": []}
    for idx, _ in enumerate(to_params.Constraints):
        history[f"constraint_{idx + 1}"] = []
    history["volfrac"] = []

    if print_progress:
        log_message("Computing Filters ...")
    H, _Hs = createFilters(fe_solver, to_params)
    dev = fe_solver.stiff_mtrx.device if fe_solver.stiff_mtrx is not None else torch.device("cuda")
    if print_progress:
        log_message("Finding elements with forces ...")
    elemsWithForces = torch.as_tensor(find_elements_with_forces(mesh, fe_solver.bc.force,
                                                                nDOFPerNode),
                                      dtype=torch.long, device=dev)
    keep = None if to_params.ElemsToKeep is None else torch.as_tensor(
        np.asarray(to_params.ElemsToKeep), dtype=torch.long, device=dev)
    fe_solver.set_material(mat_prop)
    KE = fe_solver.elem_stiff[0]
    if getattr(fe_solver, "elem_body_force", None) is not None:
        raise NotImplementedError("element body forces are not part of the GPU path")

    success = True
    errorMsg = "No errors."
    state = {"obj0": None, "iteration": 0, "nFEAs": 0}
    if use_continuation:
        tmm.initialize_SIMP_STRUCTURAL_PENALTY(1)
        tmm.initialize_SIMP_THERMAL_PENALTY(1)
    else:
        tmm.initialize_SIMP_STRUCTURAL_PENALTY(3)
    objective_name = getattr(to_params.Objective[0], "name", str(to_params.Objective[0]))
    density_filter = to_params.APPLY_FILTER_TO_DENSITY
    objective_kind = to_params.Objective[0]

    def evaluate(rho):
        if feaMode in (FEA_MODE.STRUCTURAL, FEA_MODE.THERMAL):
            sol = fe_solver.solve(rho, material_model)
            obj, grad = compute_objective_and_gradient(feaMode, to_params, sol, rho, fe_solver, KE)
        else:
            temperature = fe_thermal_solver.solve(rho, material_model)
            f_te = fe_thermal_solver.get_thermoelastic_force(rho, material_model)
            fe_structural_solver.set_thermal_forces(f_te)
            sol = fe_structural_solver.solve(rho, material_model)
            obj, grad = compute_thermoelastic_compliance_and_gradient(
                rho, temperature, sol, to_params, fe_thermal_solver, fe_structural_solver)
        c, dcdx = compute_constraint_and_gradient(feaMode, to_params, sol, rho, fe_solver, KE)
        return sol, float(obj), grad, c, dcdx

    def optimizationFunction(x):
        """src/topopt_mma.py:246-375."""
        tIterationStart = time.time()
        x = x.reshape(-1)
        rho = H.density_filter(x) if density_filter else x
        if print_progress:
            grey = float(((rho > 0.1) & (rho < 0.9)).sum()) / num_elems
            log_message(f"Percentage grey elements (rho): {grey * 100:.2f}%")
        mesh.setPseudoDensity(rho)
        if progress_callback is not None:
            progress_callback()
        _, obj, grad_obj, c, dcdx = evaluate(rho)
        if state["obj0"] is None:
            state["obj0"] = obj
        obj0 = state["obj0"]
        if any(c0 > 0.5 for c0 in c.flatten()):
            grad_obj = grad_obj * 0.0
        obj = obj / obj0
        grad_obj = grad_obj / obj0
        if not density_filter and objective_kind is TO_QOI.COMPLIANCE:
            grad_obj = H.density_filter(rho * grad_obj) / (rho + 1e-12)
        elif not density_filter and objective_kind is not TO_QOI.VOLUME_FRACTION:
            grad_obj = H.density_filter(grad_obj)
        if elemsWithForces.numel() > 0:
            grad_obj[elemsWithForces] = grad_obj.min()
        if keep is not None:
            grad_obj[keep] = grad_obj.min()
        if not density_filter:
            for m in range(len(to_params.Constraints)):
                kind = to_params.Constraints[m][0]
                if kind is TO_QOI.COMPLIANCE:
                    dcdx[m, :] = H.density_filter(rho * dcdx[m, :]) / (rho + 1e-12)
                elif kind is not TO_QOI.VOLUME_FRACTION:
                    dcdx[m, :] = H.density_filter(dcdx[m, :])
        else:
            # chain rule back to the design variables: df/dx = H^T (df/drho / Hs)
            grad_obj = H.density_chain(grad_obj)
            for m in range(dcdx.shape[0]):
                dcdx[m, :] = H.density_chain(dcdx[m, :])
        history["objective"].append(obj * obj0)
        history["volfrac"].append(float(rho.mean()))
        for idx, val in enumerate(c.flatten()):
            history[f"constraint_{idx + 1}"].append(float(val))
        if print_progress:
            lines = ["-" * 50, f"Iteration: {state['iteration']}",
                     f"Min. Objective ({objective_name}): {obj * obj0:.3g}"]
            for idx, val in enumerate(c.flatten()):
                lim = to_params.Constraints[idx][2]
                lines.append(f"Constraint {idx + 1}: {(val + 1) * lim:.3g} <= {lim:.3g}?")
            lines.append(f"Time for iteration {state['iteration']}: "
                         f"{time.time() - tIterationStart:.2f} s")
            log_message("\n".join(lines))
        state["iteration"] += 1
        state["nFEAs"] += 1
        if use_continuation and (state["iteration"] % 10 == 0):
            tmm.increment_SIMP_THERMAL_PENALTY(0.25)
            tmm.increment_SIMP_STRUCTURAL_PENALTY(0.25)
        return obj, grad_obj, c, dcdx

    initialDensity = 0.5
    for constraint in to_params.Constraints:
        if constraint[0] == TO_QOI.VOLUME_FRACTION:
            initialDensity = constraint[2]
            break
    if x0 is None:
        x0 = torch.full((num_elems,), float(initialDensity), dtype=torch.float64, device=dev)
    else:
        x0 = torch.as_tensor(x0, dtype=torch.float64).reshape(-1).to(dev)
    lowerBound = torch.full((num_elems,), 1e-6, dtype=torch.float64, device=dev)
    upperBound = torch.full((num_elems,), 1 - 1e-6, dtype=torch.float64, device=dev)
    if print_progress:
        log_message("Calling MMA ...")
    xOptimal, f0val, df0dx, gval, dgdx, outit = runMMA(
        num_elems, len(to_params.Constraints), optimizationFunction, x0, lowerBound, upperBound,
        maxIterations=maxiteration, timeLimitSecs=timeLimitSecs, move_limit=move_limit,
        fTolerance=objective_tol, gTolerance=constraint_tol, kktTol=kkt_tol, verbose=False,
        progress_callback=progress_callback)
    nFEAs = outit   # the reference returns runMMA's iteration count (:393)

    d_opt = xOptimal.reshape(-1)
    rho = H.density_filter(d_opt) if density_filter else d_opt
    if to_params.Eliminate_Hanging_Elements or binarize_topology:
        # binary topology that keeps the volume fraction (:418-436)
        x_sorted = torch.sort(rho).values
        threshold = x_sorted[int((1 - float(rho.mean())) * len(rho))]
        rho = torch.where(rho < threshold, torch.zeros_like(rho), torch.ones_like(rho))
    mesh.setPseudoDensity(rho)
    sol, obj, _, c, _ = evaluate(rho)
    grey = float(((rho > 0.1) & (rho < 0.9)).sum()) / num_elems
    history["objective"].append(obj)
    history["volfrac"].append(float(rho.mean()))
    for idx, val in enumerate(c.flatten()):
        history[f"constraint_{idx + 1}"].append(float(val))
    if print_progress:
        print("-" * 50)
        log_message(f"Final objective: {obj:.4g}, vf: {float(rho.mean()):.3f}, grey: {100 * grey:.2f}%")
        log_message(f"Total Time: {time.time() - tStart:.2f} s")
        log_message(f"Error: {errorMsg}")
    return sol, history, success, errorMsg, nFEAs
