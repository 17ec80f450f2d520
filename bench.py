"""Benchmark: FEA solves/s of the 256x128x128 voxel cantilever (12.8M DOF).

A step is one full solve with inputs resident in HBM: set the densities (SIMP
moduli, Jacobi diagonal, Galerkin coarse matrix + FP64 Cholesky and inverse)
and solve K u = f to sqrt(r^T M r / r0^T M r0) <= 1e-8 with FP32 deflated PCG
inside FP64 iterative refinement.  The same solve in plain FP64 is timed
beside it, and the dominant kernel (the FP32 matrix-free K*u) is timed with
events on its stream for the roofline.  The CPU baseline is the oracle in the
reference's form (assembled CSR, two-SpMV DPCG, SuperLU refactorisation per
coarse solve) on a bounded sample of config 1 (60x20x10), timed beside the
complete config-1 OC loop on the GPU; the bench workload itself does not fit
the reference's form on a host (its COO assembly needs ~100 GB).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FEA solves/s at 12.6M DOF (FP32+refine vs FP64, rel res 1e-8); HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def config1_lines(torch, dev, budget_s=20.0, run_cpu=True):
    """Config 1, the reference's own CPU-runnable case (60x20x10, 42,273 DOF,
    x0 = 0.5, unit -z point load at node (60,10,5), Emin = 1e-9, sensitivity
    filter r = 1.5, volume fraction 0.5, DPCG rtol 1e-8 with the reference's
    group heuristic, 96 groups) on both sides of the same run.

    GPU: the complete device OC loop, 50 iterations plus the tail (binary
    threshold, largest connected component, final solve), wall clock.
    CPU: the oracle in the reference's form (COO assembly and CSR per solve,
    Dirichlet slicing, Jacobi, two-SpMV DPCG with scipy spsolve
    refactorising the coarse matrix at every coarse solve) for as many OC
    iterations as fit in about budget_s, one BLAS thread (scipy's sparse
    kernels are single-threaded)."""
    from pyto_amd import deflation, hex_structural_fea, linear_solvers, problems
    from pyto_amd import topopt_common, topopt_material_model as tmm, topopt_ocm
    mesh, mat, bc = problems.cantilever((60, 20, 10), load="point", point=(60, 10, 5))
    ng = min(2000, max(10, round(3 * mesh.num_nodes / 500)))   # requested (85): 96 boxes
    tmm.set_EVOID_RELATIVE(1e-9)
    try:
        ds = deflation.DeflationSolver()
        ds.create_deflation_groups(mesh, ng)
        fe = hex_structural_fea.HexStructuralFEA(mesh, mat, bc, linear_solvers.Solvers.DPCG,
                                                 dsolver=ds, rtol=1e-8, maxIters=2000)
        # exact Galerkin matrix: with the reference's 1e-10 coarse shift this
        # workload's DPCG diverges once void elements (E = 2e-9) appear
        fe.stiff_mtrx.set_coarse_shift(0.0, 0.0)
        params = topopt_common.TOParams()
        topopt_ocm.topopt_optimality_criteria(topopt_common.FEA_MODE.STRUCTURAL, fe, params,
                                              maxIterations=2, print_progress=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hist, _, _, nfea = topopt_ocm.topopt_optimality_criteria(
            topopt_common.FEA_MODE.STRUCTURAL, fe, params, maxIterations=50,
            print_progress=False)
        torch.cuda.synchronize()
        t_gpu = time.perf_counter() - t0
    finally:
        tmm.set_EVOID_RELATIVE(1e-8)
    gpu = {"oc_iterations": nfea, "solves": nfea + 1, "seconds": t_gpu,
           "solves_per_s": (nfea + 1) / t_gpu, "final_objective": hist["objective"][-1],
           "groups": ds.ws_nGroups}
    if not run_cpu:
        return gpu, None
    from threadpoolctl import threadpool_limits
    from oracle import fea as OF
    from oracle import ke as OK
    from oracle import mesh as OM
    from oracle import topopt as OT
    g = OM.GridMesh((60, 20, 10), (1.0, 1.0, 1.0))
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    fixed, _ = OF.clamp_x0(g)
    f = OF.point_load(g, 60, 10, 5)
    with threadpool_limits(limits=1):
        ms = OT.reference_dpcg_solver(g, KE, fixed, eps=1e-9, rtol=1e-8, shift=0.0)
        t0 = time.perf_counter()
        OT.oc_loop(g, KE, f, fixed, 0.5, ms, max_iterations=1, eps=1e-9)
        t_two = time.perf_counter() - t0           # one iteration + the tail's solve
        k = max(2, int(budget_s / (t_two / 2)) - 1)
        ms = OT.reference_dpcg_solver(g, KE, fixed, eps=1e-9, rtol=1e-8, shift=0.0)
        t0 = time.perf_counter()
        r = OT.oc_loop(g, KE, f, fixed, 0.5, ms, max_iterations=k, eps=1e-9)
        t_cpu = time.perf_counter() - t0
    solves = r["nFEAs"] + 1
    env = {k_: os.environ.get(k_) for k_ in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                             "MKL_NUM_THREADS")}
    cpu = {"value": solves / t_cpu, "unit": "solves/s (config 1: 60x20x10, 42,273 DOF)",
           "cores": 1, "kind": "port",
           "sample": "config 1 OC loop, first %d iterations plus the tail's final solve (%d "
                     "solves, %.1f s) of the oracle in the reference's form: COO assembly and "
                     "CSR per solve, Dirichlet slicing, Jacobi, %d reference-heuristic groups, "
                     "two-SpMV DPCG to rtol 1e-8 with spsolve refactorising the coarse matrix "
                     "at every coarse solve (DPCG iterations per solve %d..%d), sensitivity "
                     "filter and OC update; one thread; both sides use the exact Galerkin "
                     "matrix (the reference's 1e-10 shift makes this workload diverge)"
                     % (r["nFEAs"], solves, t_cpu, ds.ws_nGroups, min(ms.stats["iterations"]),
                        max(ms.stats["iterations"])),
           "host_cores_available": len(os.sched_getaffinity(0)),
           "blas_threads_used": 1, "thread_env": env,
           "gpu_same_workload": gpu,
           "gpu_over_cpu_same_workload": gpu["solves_per_s"] / (solves / t_cpu),
           "bench_workload": "N/A on the CPU in the reference's form: its COO assembly of the "
                             "256x128x128 grid needs about 100 GB of host memory"}
    return gpu, cpu


def reference_groups_line(op, mesh, f, u, step, dev):
    """The bench workload with the reference's own coarse space:
    create_deflation_groups(mesh, min(2000, max(10, round(3 N / 500)))), 20 x
    10 x 10 = 2000 groups (12,000 coarse unknowns, a 1.15 GB dense FP64
    inverse applied every iteration).  Reported beside `value`."""
    from pyto_amd import deflation
    ds = deflation.DeflationSolver()
    ng = min(2000, max(10, round(3 * mesh.num_nodes / 500)))
    t0 = time.perf_counter()
    ds.create_deflation_groups(mesh, ng, device=dev)
    ds.attach(op)
    step("mixed")
    import torch
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    out = {"groups": op.ngroups, "set_groups_and_first_solve_s": t_setup}
    for prec in ("mixed", "fp64"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = step(prec)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        # the same solve without the density set-up (Galerkin matrix,
        # Cholesky and inverse of the 12,000-unknown coarse matrix)
        t0 = time.perf_counter()
        st2 = op.solve(f, rtol=1e-8, max_iters=5000, precision=prec, out=u)[1]
        torch.cuda.synchronize()
        t2 = time.perf_counter() - t0
        out[prec] = {"ms_per_solve": t * 1e3, "ms_solve_only": t2 * 1e3,
                     "ms_density_setup": (t - t2) * 1e3, "iterations": st["iterations"],
                     "ms_per_iteration": t2 * 1e3 / max(st2["iterations"], 1),
                     "converged": st["converged"], "rel": st["rel"]}
    return out


def sensitivity_errors(op, x, u_mixed, u64, inner_rtol):
    """Errors of the FP32+refine path against the FP64 path in the
    sensitivities: the compliance gradient -E'(x) u_e^T KE u_e, and the p-norm
    stress (p = 6, q = 0.5) with its gradient T1 + T2, whose adjoint is solved
    in the same precision as the forward solve (one solve each)."""
    import torch
    Ep = 3.0 * x * x * (1.0 - 1e-9)   # SIMP p = 3, Emin = 1e-9
    out = {}
    _, gm, _ = op.compliance(u_mixed, x)
    _, g6, _ = op.compliance(u64, x)
    out["compliance_gradient_max_rel"] = float((gm - g6).abs().max() / g6.abs().max())
    pn, grads = {}, {}
    for name, u, prec in (("mixed", u_mixed, "mixed"), ("fp64", u64, "fp64")):
        st = op.stress(u, x, q=0.5, pexp=6.0, want_strain=False, want_stress=False,
                       want_energy=False)
        pn[name] = float(st["sum_vm_p"][0]) ** (1.0 / 6.0)
        g, T1 = op.pnorm_adjoint_rhs(u, x, st["vm"], st["sum_vm_p"], q=0.5, pexp=6.0)
        lam, _ = op.solve(g, rtol=1e-8, max_iters=5000, precision=prec, inner_rtol=inner_rtol)
        grads[name] = T1 - Ep * op.bilinear(lam, u)
    out["pnorm_rel"] = abs(pn["mixed"] - pn["fp64"]) / pn["fp64"]
    out["pnorm_gradient_max_rel"] = float((grads["mixed"] - grads["fp64"]).abs().max()
                                          / grads["fp64"].abs().max())
    return out


def mma_problem(n, seed=7):
    """Synthetic data of one MMA step (iteration 3, one volume constraint
    v_f = 0.3): densities in [1e-6, 1 - 1e-6], compliance-like negative
    sensitivities, the two previous iterates a few per cent away."""
    import numpy as np
    rng = np.random.default_rng(seed)
    x = np.clip(rng.beta(0.6, 1.2, n), 1e-6, 1 - 1e-6)
    df0dx = -3.0 * x * x * rng.lognormal(0.0, 1.0, n) - 1e-9
    xo1 = np.clip(x + rng.normal(0, 0.02, n), 1e-6, 1 - 1e-6)
    xo2 = np.clip(xo1 + rng.normal(0, 0.02, n), 1e-6, 1 - 1e-6)
    low = x - rng.uniform(0.1, 0.5, n)
    upp = x + rng.uniform(0.1, 0.5, n)
    dfdx = np.full((1, n), 1.0 / (0.3 * n))
    fval = np.array([x.mean() / 0.3 - 1.0])
    return x, xo1, xo2, low, upp, df0dx, dfdx, fval


def mma_cpu(n=262144):
    """The oracle's mmasub on n variables of the same synthetic step."""
    import numpy as np
    from oracle import mma as OMMA
    x, xo1, xo2, low, upp, df0dx, dfdx, fval = mma_problem(n)
    t0 = time.perf_counter()
    OMMA.mmasub(1, n, 3, x, np.full(n, 1e-6), np.full(n, 1 - 1e-6), xo1, xo2, 0.0, df0dx, fval,
                dfdx, low, upp, 1.0, np.zeros(1), 100 * np.ones(1), np.ones(1), 0.05)
    t = time.perf_counter() - t0
    nw, ev, lv = OMMA.subsolv.last_counts
    return {"variables": n, "ms": t * 1e3, "newton_steps": nw, "evaluations": ev,
            "mvars_per_s": n / t / 1e6, "cores": 1, "kind": "port"}


def mma_line(torch, dev, n):
    """One device MMA step (mmasub) on n variables of the synthetic step:
    time per call, Newton steps, and the algorithmic bandwidth of its element
    passes, m = 1: setup 8 (20 + 3m) n bytes, each residual evaluation
    8 (12 + 2m) n, each Newton direction 8 (18 + 2m) n, the final point 72 n."""
    import numpy as np
    from pyto_amd import mmaWrapper as W
    x, xo1, xo2, low, upp, df0dx, dfdx, fval = mma_problem(n)
    T = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    args = [T(x), T(np.full(n, 1e-6)), T(np.full(n, 1 - 1e-6)), T(xo1), T(xo2), 0.0, T(df0dx),
            fval, T(dfdx), T(low), T(upp), 1.0, np.zeros(1), 100 * np.ones(1), np.ones(1), 0.05]
    W.mmasub(1, n, 3, *args)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        W.mmasub(1, n, 3, *args)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    st = W.mmasub.last_stats
    nbytes = 8 * n * (23 + 14 * st["evaluations"] + 20 * st["newton_steps"] + 9)
    return {"variables": n, "constraints": 1, "ms": t * 1e3, "newton_steps": st["newton_steps"],
            "evaluations": st["evaluations"], "levels": st["levels"],
            "algorithmic_GBs": nbytes / t / 1e9,
            "workload": "one MMA step (iteration 3, move 0.05, volume constraint) on the "
                        "256x128x128 element count, synthetic compliance-like gradient"}


def traffic_bytes():
    """HBM bytes per k_matvec<float> launch at this workload, from the two
    rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) summarised in the profile
    file, each scaled by the same counter's reading for a copy of known size."""
    path = os.path.join(ROOT, "profiles", "k_matvec_f32_cantilever_256x128x128.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def batch_line(torch, dev, inner_rtol=None, boxes=12):
    """The 32-design batch; ``boxes``: the deflation box partition of every
    design, or a number (a box edge in elements: each design gets
    max(1, round(n / edge)) boxes per axis), or None (Jacobi PCG only)."""
    from pyto_amd import deflation, problems
    from pyto_amd.batch import BatchSolver, measure_grid_barrier
    from pyto_amd.operator import GridOperator
    designs = problems.batch_designs()
    ops = []
    for d in designs:
        op = GridOperator(d["grid"], (1, 1, 1), 1.0, 0.3, d["bc"].fixed_dofs, penal=3.0,
                          eps_void=1e-8)
        if boxes is not None:
            op.set_coarse_shift(0.0, 0.0)
            ds = deflation.DeflationSolver()
            nb = boxes if isinstance(boxes, (tuple, list)) else \
                tuple(max(1, int(round(n / boxes))) for n in d["grid"])
            ds.create_box_groups(d["mesh"], nb)
            ds.attach(op)
        op.set_density(torch.from_numpy(d["x"]).to(dev))
        ops.append(op)
    bs = BatchSolver(ops, [d["loads"] for d in designs])
    tune = bs.autotune()
    kw = {} if inner_rtol is None else {"inner_rtol": inner_rtol}
    bs.solve(rtol=1e-8, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, stats = bs.solve(rtol=1e-8, **kw)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    # the per-design-update set-up: E_e, the Jacobi diagonal and (deflated)
    # the Galerkin coarse matrix and its factor, for every design
    xs = [op.x.clone() for op in ops]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for op, x in zip(ops, xs):
        op.set_density(x)
    torch.cuda.synchronize()
    t_set_seq = time.perf_counter() - t0
    # the same update with one batched coarse factorisation for all designs
    bs.set_densities(xs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bs.set_densities(xs)
    torch.cuda.synchronize()
    t_set = time.perf_counter() - t0
    it = [s["iterations"] for s in stats]
    ndof = sum(3 * d["mesh"].num_nodes * len(d["loads"]) for d in designs)
    out = {"designs": len(designs), "systems": bs.nsys, "dof_total": ndof,
           "seconds": t, "systems_per_s": bs.nsys / t, "density_update_seconds": t_set,
           "density_update_seconds_per_design_calls": t_set_seq,
           "all_converged": all(s["converged"] for s in stats),
           "inner_iterations_max": max(it), "inner_iterations_median": float(np.median(it)),
           "outer_iterations_max": max(s["outer_iterations"] for s in stats),
           "deflation_boxes": (None if boxes is None else list(boxes)
                               if isinstance(boxes, (tuple, list)) else "edge %g" % boxes),
           "zchunk": tune["zchunk"], "tile_rows": tune["rows"],
           "ms_per_lockstep_iteration": t * 1e3 / max(it),
           # the persistent-kernel alternative: one grid-wide barrier per
           # dependent phase of an iteration instead of a graph node
           "grid_barrier_us": {str(b): measure_grid_barrier(b, rounds=2000) for b in (1, 2)}}
    bs.close()
    out["block_pcg"] = block_pcg_probe(torch, ops, designs)
    for op in ops:
        op.close()
    return out


def block_pcg_probe(torch, ops, designs, count=2):
    """Block deflated PCG (one Krylov space per design's load cases) against
    the per-load-case form of the same GPU code (s = 1 per column), FP64 to
    rel res 1e-8, on the `count` designs with the most load cases (largest
    grid first on ties).  The lockstep batch runs as many iterations as its
    hardest system, so the figure of merit is the block iteration count over
    the per-case maximum."""
    from pyto_amd.batch import block_dpcg
    order = sorted(range(len(designs)), key=lambda i: (-len(designs[i]["loads"]),
                                                       -int(np.prod(designs[i]["grid"]))))
    res = []
    for i in order[:count]:
        op, loads = ops[i], designs[i]["loads"]
        _, sb = block_dpcg(op, loads, rtol=1e-8)
        its = [block_dpcg(op, [f], rtol=1e-8)[1]["iterations"] for f in loads]
        res.append({"grid": list(designs[i]["grid"]), "load_cases": len(loads),
                    "groups": op.ngroups, "block_iterations": sb["iterations"],
                    "per_case_iterations": its, "block_converged": sb["converged"],
                    "block_over_per_case_max": sb["iterations"] / max(its)})
    return res


def thermal_line(torch, dev, grid, x):
    """The thermal operator on the same grid and densities (conductivity law
    k = x, one temperature per node): FP32/FP64 launch times against the
    algorithmic bytes, one Jacobi PCG solve to rel res 1e-8 (sink T = 0 on
    x = 0, uniform heat flux into the x = max face) in FP64 and mixed, and the
    same solves deflated by 8x4x4 and 16x8x8 box groups (one constant per
    group)."""
    import numpy as np
    from pyto_amd.operator import ThermalOperator
    nelx, nely, nelz = grid
    sx, sy, sz = nelx + 1, nely + 1, nelz + 1
    nn, ne = sx * sy * sz, nelx * nely * nelz
    idx = np.arange(nn)
    ix = idx % sx
    op = ThermalOperator(grid, (1.0, 1.0, 1.0), 1.0, idx[ix == 0], penal=1.0, k_void=1e-3)
    op.set_density(x)
    q = torch.zeros(nn, dtype=torch.float64, device=dev)
    q[torch.from_numpy(idx[ix == nelx]).to(dev)] = 1.0 / (sy * sz)
    res = {}
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        u = torch.randn(nn, dtype=dt, device=dev)
        y = torch.empty_like(u)
        for _ in range(3):
            op.matvec(u, y)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            op.matvec(u, y)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / 20 * 1e-3
        s = 4 if dt == torch.float32 else 8
        nbytes = s * 2 * nn + s * ne
        res["matvec_" + name + "_us"] = t * 1e6
        res["matvec_" + name + "_GBs"] = nbytes / t / 1e9
    op.solve(q, rtol=1e-8, max_iters=100000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    T, st = op.solve(q, rtol=1e-8, max_iters=100000)
    torch.cuda.synchronize()
    res.update({"grid": list(grid), "nodes": nn, "solve_ms": (time.perf_counter() - t0) * 1e3,
                "iterations": st["iterations"], "converged": bool(st["converged"]),
                "solver": "Jacobi PCG FP64 (no deflation groups)"})
    # the same solve in FP32 inside FP64 refinement
    op.solve(q, rtol=1e-8, max_iters=100000, precision="mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Tm, sm = op.solve(q, rtol=1e-8, max_iters=100000, precision="mixed")
    torch.cuda.synchronize()
    C64, Cm = float(torch.dot(q, T)), float(torch.dot(q, Tm))
    res.update({"mixed_solve_ms": (time.perf_counter() - t0) * 1e3,
                "mixed_iterations": sm["iterations"], "mixed_refinements": sm["outer_iterations"],
                "mixed_converged": bool(sm["converged"]),
                "mixed_compliance_rel_vs_fp64": abs(Cm - C64) / abs(C64)})
    # the deflated solves: one constant per box group (the thermal coarse
    # space), density set-up (Galerkin matrix and its inverse) timed apart
    from pyto_amd import builders
    op.set_coarse_shift(0.0, 0.0)
    deflated = []
    for boxes in ((8, 4, 4), (16, 8, 8)):
        ngn, cen, _ = builders.box_groups(grid, (1.0, 1.0, 1.0), boxes, 15, dev)
        op.set_groups(ngn, cen)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.set_density(x)
        torch.cuda.synchronize()
        t_set = time.perf_counter() - t0
        row = {"boxes": list(boxes), "groups": op.ngroups, "density_setup_ms": t_set * 1e3}
        for prec in ("fp64", "mixed"):
            op.solve(q, rtol=1e-8, max_iters=100000, precision=prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            Td, sd = op.solve(q, rtol=1e-8, max_iters=100000, precision=prec)
            torch.cuda.synchronize()
            row[prec] = {"solve_ms": (time.perf_counter() - t0) * 1e3,
                         "iterations": sd["iterations"], "converged": bool(sd["converged"]),
                         "compliance_rel_vs_jacobi_fp64":
                             abs(float(torch.dot(q, Td)) - C64) / abs(C64)}
        deflated.append(row)
    res["deflated"] = deflated
    op.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--boxes", default="4,4,4")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inner-rtol", type=float, default=0.1)
    ap.add_argument("--tuned-boxes", default="12,6,6",
                    help="second aggregate layout timed beside the named one ('' to skip)")
    ap.add_argument("--no-thermal", action="store_true",
                    help="skip the thermal-operator line")
    ap.add_argument("--no-mma", action="store_true", help="skip the MMA step line")
    ap.add_argument("--no-reference-groups", action="store_true",
                    help="skip the solves with the reference's 2000-group coarse space")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 32-design line reported beside `value`")
    ap.add_argument("--thermal-only", action="store_true",
                    help="only the thermal line (operator launches, Jacobi and deflated "
                         "solves) on the workload's grid and densities (for kernel traces)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched 32-design line (for kernel traces)")
    ap.add_argument("--solve-only", action="store_true",
                    help="only the warm-up, --steps mixed solves and --fp64-steps FP64 solves "
                         "of the workload (for kernel traces); prints their times")
    ap.add_argument("--reference-groups-only", action="store_true",
                    help="only the workload with the reference's own coarse space (2000 "
                         "groups) after one warm-up solve (for kernel traces)")
    ap.add_argument("--tile-rows", type=int, default=16,
                    help="operator tile rows for --operator-only (8, 16 or 32)")
    ap.add_argument("--operator-only", action="store_true",
                    help="only 20 launches each of the FP32 operator, the FP64 one and the "
                         "FP32 matrix-core variant on the workload (for counter collection); "
                         "prints their HIP-event averages")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from pyto_amd import deflation, problems
    from pyto_amd.operator import GridOperator
    if args.batch_only:
        print(json.dumps({"batch_only": True, "batch_32_designs": batch_line(torch,
                                                                              torch.device("cuda"))}))
        return

    from pyto_amd import builders

    cfg = problems.CONFIGS[args.config]
    if args.thermal_only:
        x_t = torch.from_numpy(problems.config_density(args.config, cfg["grid"])).to("cuda")
        print(json.dumps({"thermal_only": True,
                          "thermal": thermal_line(torch, torch.device("cuda"), cfg["grid"], x_t)}))
        return
    grid = cfg["grid"]
    boxes = tuple(int(v) for v in args.boxes.split(","))
    dev = torch.device("cuda")
    # problem set-up, host builders (vectorised numpy) against the device
    # builders; both give the same bits, the device arrays are used
    t0 = time.perf_counter()
    mesh, mat, bc = problems.cantilever(grid, load=cfg["load"], point=cfg.get("point"))
    ds_host = deflation.DeflationSolver()
    ds_host.create_box_groups(mesh, boxes)
    t_setup_host = time.perf_counter() - t0
    builders.fixed_mask((2, 2, 2), [(0, False, 7)], dev)     # first launch of the builders
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mask = builders.fixed_mask(grid, [(0, False, 7)], dev)
    if cfg["load"] == "traction":
        f = builders.face_load(grid, device=dev)
    else:
        f = torch.from_numpy(bc.force).to(dev)
    ds = deflation.DeflationSolver()
    ds.create_box_groups(mesh, boxes, device=dev)
    torch.cuda.synchronize()
    t_setup_dev = time.perf_counter() - t0
    del ds_host
    log("set-up: host builders %.1f ms, device builders %.1f ms"
        % (t_setup_host * 1e3, t_setup_dev * 1e3))
    x_np = problems.config_density(args.config, grid)
    if args.operator_only:
        # 20 back-to-back launches each of the FP32 operator, the FP64 one
        # and the FP32 variant with the dense KE on the matrix cores
        op = GridOperator(grid, (1, 1, 1), 1.0, 0.3, None, penal=3.0, eps_void=1e-9,
                          fixed_mask=mask)
        op.set_density(torch.from_numpy(x_np).to(dev))
        op.set_tile_rows(args.tile_rows)
        p32 = torch.randn(op.num_dofs, dtype=torch.float32, device=dev)
        p64 = p32.double()
        res = {"operator_only": True}
        for name, fn, u in (("k_matvec<float>", op.matvec, p32),
                            ("k_matvec<double>", op.matvec, p64),
                            ("k_matvec_mfma", op.matvec_mfma, p32)):
            q = torch.empty_like(u)
            for _ in range(3):
                fn(u, q)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                fn(u, q)
            e1.record()
            torch.cuda.synchronize()
            res[name + "_us"] = e0.elapsed_time(e1) / 20 * 1e3
            res[name + "_out"] = q
        y32, y64, ym = (res.pop(k + "_out") for k in ("k_matvec<float>", "k_matvec<double>",
                                                      "k_matvec_mfma"))
        res["f32_vs_f64_relmax"] = float((y32.double() - y64).abs().max() / y64.abs().max())
        res["mfma_vs_f64_relmax"] = float((ym.double() - y64).abs().max() / y64.abs().max())
        res["tile_rows"] = op.tile_rows
        # the PMC calibration: three copies of one FP32 DOF vector (known bytes)
        cal_src = torch.randn(op.num_dofs, dtype=torch.float32, device=dev)
        cal_dst = torch.empty_like(cal_src)
        for _ in range(3):
            cal_dst.copy_(cal_src)
        torch.cuda.synchronize()
        res["calibration_copy_bytes"] = 4 * op.num_dofs
        print(json.dumps(res))
        return
    log("grid", grid, "dof", 3 * mesh.num_nodes)
    op = GridOperator(grid, (1, 1, 1), 1.0, 0.3, None, penal=3.0, eps_void=1e-9,
                      fixed_mask=mask)
    op.set_coarse_shift(0.0, 0.0)
    ds.attach(op)
    x = torch.from_numpy(x_np).to(dev)
    u = torch.empty_like(f)
    ngroups = op.ngroups
    log("groups", ngroups, "workspace MB", op.workspace_bytes() / 1e6)

    def step(precision):
        op.set_density(x)
        _, st = op.solve(f, rtol=1e-8, max_iters=5000, precision=precision,
                         inner_rtol=args.inner_rtol, out=u)
        return st

    if args.reference_groups_only:
        print(json.dumps({"reference_groups_only": True,
                          "reference_groups": reference_groups_line(op, mesh, f, u, step, dev)}))
        return

    for _ in range(args.warmup):
        st = step("mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.set_density(x)
    torch.cuda.synchronize()
    log("set_density (SIMP, Jacobi, Galerkin + Cholesky + inverse): %.2f ms"
        % ((time.perf_counter() - t0) * 1e3))
    # each solve returns after reading its scalars back, so the per-step
    # clock brackets a finished solve
    t0 = time.perf_counter()
    stats, per_mixed = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        stats.append(step("mixed"))
        per_mixed.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    u_mixed = u.clone()
    log("mixed: %.3f ms/solve" % (t_mixed * 1e3), stats[-1])
    if args.solve_only:
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            st64 = step("fp64")
        torch.cuda.synchronize()
        t_fp64 = (time.perf_counter() - t0) / max(1, args.fp64_steps)
        print(json.dumps({"solve_only": True, "mixed_ms": t_mixed * 1e3,
                          "fp64_ms": t_fp64 * 1e3, "mixed": stats[-1], "fp64": st64,
                          "groups": ngroups}))
        return

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per_fp64 = []
    for _ in range(args.fp64_steps):
        t1 = time.perf_counter()
        st64 = step("fp64")
        per_fp64.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    t_fp64 = (time.perf_counter() - t0) / args.fp64_steps
    u64 = u.clone()
    log("fp64: %.3f ms/solve" % (t_fp64 * 1e3), st64)
    C_mixed = float(torch.dot(f, u_mixed))
    C_64 = float(torch.dot(f, u64))
    err_u = float((u_mixed - u64).abs().max() / u64.abs().max())
    err_C = abs(C_mixed - C_64) / abs(C_64)
    err_sens = sensitivity_errors(op, x, u_mixed, u64, args.inner_rtol)
    log("errors vs fp64: displacement %.2e, compliance %.2e," % (err_u, err_C), err_sens)

    # roofline of the dominant kernel: FP32 K*u, events on the launch stream
    nd, ne = op.num_dofs, op.num_elems
    p32 = torch.randn(nd, dtype=torch.float32, device=dev)
    q32 = torch.empty_like(p32)
    for _ in range(3):
        op.matvec(p32, q32)
    reps = 20
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        op.matvec(p32, q32)
    e1.record()
    torch.cuda.synchronize()
    t_mv32 = e0.elapsed_time(e1) / reps * 1e-3
    bytes32 = 4 * 2 * nd + 4 * ne
    p64 = p32.double()
    q64 = torch.empty_like(p64)
    e0.record()
    for _ in range(reps):
        op.matvec(p64, q64)
    e1.record()
    torch.cuda.synchronize()
    t_mv64 = e0.elapsed_time(e1) / reps * 1e-3
    bytes64 = 8 * 2 * nd + 8 * ne
    gbs32 = bytes32 / t_mv32 / 1e9
    # calibration for the PMC traffic counters: a plain device copy of one
    # FP32 DOF vector (known bytes) under the same counters
    cal_src = torch.randn(nd, dtype=torch.float32, device=dev)
    cal_dst = torch.empty_like(cal_src)
    for _ in range(3):
        cal_dst.copy_(cal_src)
    torch.cuda.synchronize()
    log("matvec f32 %.1f us %.0f GB/s; f64 %.1f us %.0f GB/s" %
        (t_mv32 * 1e6, gbs32, t_mv64 * 1e6, bytes64 / t_mv64 / 1e9))

    # the same solves with a tuned aggregate layout (reported beside `value`,
    # which stays on the named 64-aggregate configuration)
    tuned = None
    if args.tuned_boxes:
        ds2 = deflation.DeflationSolver()
        ds2.create_box_groups(mesh, tuple(int(v) for v in args.tuned_boxes.split(",")), device=dev)
        ds2.attach(op)
        step("mixed")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st2 = step("mixed")
        torch.cuda.synchronize()
        t2m = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        st2d = step("fp64")
        torch.cuda.synchronize()
        t2d = time.perf_counter() - t0
        tuned = {"boxes": args.tuned_boxes, "groups": op.ngroups,
                 "fp32_refined_solves_per_s": 1.0 / t2m, "inner_iterations": st2["iterations"],
                 "fp64_plain_solves_per_s": 1.0 / t2d, "fp64_iterations": st2d["iterations"],
                 "speedup_vs_fp64": t2d / t2m}
        log("tuned aggregates %s: mixed %.3f ms, fp64 %.3f ms" % (args.tuned_boxes, t2m * 1e3,
                                                                   t2d * 1e3))

    # the reference's own coarse space (2000 groups) on the bench workload
    ref_groups = None
    if not args.no_reference_groups:
        ref_groups = reference_groups_line(op, mesh, f, u, step, dev)
        log("reference groups", ref_groups)

    # the batching workload, reported beside `value`: 32 designs (seed 5)
    # with 1-6 point-load cases each, all systems to rtol 1e-8 in one batch
    batch = None
    if not args.no_batch:
        del op
        torch.cuda.empty_cache()
        batch = batch_line(torch, dev)
        log("batch", batch)

    mma = None
    if not args.no_mma:
        mma = mma_line(torch, dev, int(np.prod(grid)))
        log("mma", mma)

    thermal = None
    if not args.no_thermal:
        thermal = thermal_line(torch, dev, grid, x)
        log("thermal", thermal)

    c1_gpu, cpu = config1_lines(torch, dev, run_cpu=not args.no_cpu)
    log("config 1 GPU OC loop", c1_gpu)
    if cpu is not None:
        if mma is not None:
            cpu["mma_step_oracle"] = mma_cpu()
        log("cpu baseline", cpu)

    s = stats[-1]
    line = {
        "metric": METRIC, "value": 1.0 / t_mixed, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "voxel cantilever %dx%dx%d (%d DOF), x=0 clamped, unit -z traction "
                               "on x=max, SIMP p=3 Emin=1e-9, 16-bump density seed 2; %d box "
                               "deflation groups; FP32 DPCG in FP64 refinement to rel res 1e-8"
                               % (grid + (nd, ngroups)),
                   "grid": list(grid), "dof": nd},
        "fp32_refined": {"solves_per_s": 1.0 / t_mixed, "inner_iterations": s["iterations"],
                         "outer_iterations": s["outer_iterations"], "rel": s["rel"],
                         "rel_true": s["rel_true"],
                         "pcg_iterations_per_s": s["iterations"] / t_mixed,
                         "ms_median": float(np.median(per_mixed)) * 1e3,
                         "ms_min": min(per_mixed) * 1e3},
        "fp64_plain": {"solves_per_s": 1.0 / t_fp64, "iterations": st64["iterations"],
                       "rel": st64["rel"], "rel_true": st64["rel_true"],
                       "pcg_iterations_per_s": st64["iterations"] / t_fp64,
                       "ms_median": float(np.median(per_fp64)) * 1e3,
                       "ms_min": min(per_fp64) * 1e3},
        "speedup_vs_fp64": t_fp64 / t_mixed,
        "error_vs_fp64": dict({"displacement_max_rel": err_u, "compliance_rel": err_C},
                              **err_sens),
        "roofline": {"bound": "hbm", "achieved": gbs32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs32 / HBM_PEAK_GBS, "traffic": traffic_bytes(),
                     "kernel": "k_matvec<float>", "kernel_us": t_mv32 * 1e6,
                     "algorithmic_bytes": bytes32,
                     "fp64_kernel_us": t_mv64 * 1e6, "fp64_GBs": bytes64 / t_mv64 / 1e9},
        "setup_ms": {"host_builders": t_setup_host * 1e3, "device_builders": t_setup_dev * 1e3,
                     "what": "fixed-DOF mask of the x=0 clamp, tip traction, %dx%dx%d box "
                             "deflation groups (ids, counts, centres)" % boxes},
        "tuned_aggregates": tuned,
        "reference_groups": ref_groups,
        "config1_oc_loop_gpu": c1_gpu,
        "batch_32_designs": batch,
        "thermal": thermal,
        "mma_step": mma,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
