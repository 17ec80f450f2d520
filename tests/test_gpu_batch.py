"""Batched solves of many designs and load cases against single-system solves."""
import numpy as np
import pytest
import torch

from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _design(grid, seed, nloads, gpu, boxes=None, h=(1.0, 1.0, 1.0)):
    from pyto_amd import deflation, problems
    from pyto_amd.operator import GridOperator
    mesh, mat, bc = problems.cantilever(grid, elem_size=h, load="point")
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.01, 1.0, mesh.num_elems)
    sx, sy = grid[0] + 1, grid[1] + 1
    loads = []
    for fn in rng.choice(sy * (grid[2] + 1), nloads, replace=False):
        f = np.zeros(3 * mesh.num_nodes)
        f[3 * (grid[0] + sx * (fn % sy + sy * (fn // sy))) + 2] = -1.0
        loads.append(f)
    op = GridOperator(grid, h, 1.0, 0.3, bc.fixed_dofs, penal=3.0, eps_void=1e-8)
    if boxes is not None:
        op.set_coarse_shift(0.0, 0.0)
        ds = deflation.DeflationSolver()
        ds.create_box_groups(mesh, boxes)
        ds.attach(op)
    op.set_density(torch.from_numpy(x).to(gpu))
    return op, loads


def _check_against_single(op, f, u, tol_u=1e-6):
    ft = torch.from_numpy(f).to(u.device)
    # true residual of the batched solution
    r = ft - op.matvec(u)
    free = op.diagonal() != 0
    res = float(torch.linalg.norm(r[free]) / torch.linalg.norm(ft[free]))
    ref, st = op.solve(ft, rtol=1e-11, max_iters=20000)
    C, C_ref = float(torch.dot(ft, u)), float(torch.dot(ft, ref))
    return res, rel_err(u.cpu().numpy(), ref.cpu().numpy()), abs(C - C_ref) / abs(C_ref)


def test_batch_matches_single_solves(gpu):
    from pyto_amd.batch import BatchSolver
    print("batch vs single", flush=True)
    specs = [((16, 8, 8), 1, 2), ((24, 12, 8), 2, 1), ((32, 16, 16), 3, 3), ((12, 6, 4), 4, 2),
             ((20, 10, 12), 5, 1)]
    ops, loads = zip(*[_design(g, s, n, gpu) for g, s, n in specs])
    bs = BatchSolver(ops, loads)
    us, stats = bs.solve(rtol=1e-8)
    assert len(us) == sum(n for _, _, n in specs) == bs.nsys
    for (d, k), u, st in zip(bs.systems, us, stats):
        assert st["converged"] and st["rel"] <= 1e-8
        res, eu, eC = _check_against_single(ops[d], loads[d][k], u)
        print(d, k, st["iterations"], st["outer_iterations"], "res %.1e err u %.1e C %.1e" % (res, eu, eC),
              flush=True)
        assert res <= 1e-6 and eu <= 1e-5 and eC <= 1e-7


def test_deflated_batch_matches_single_solves(gpu):
    """Designs with box deflation groups (and one without) in one batch:
    every system against a tight single-system solve, and far fewer inner
    iterations than the same batch without groups."""
    from pyto_amd.batch import BatchSolver
    print("deflated batch vs single", flush=True)
    specs = [((32, 16, 16), 11, 3, (4, 2, 2)), ((48, 24, 16), 12, 2, (6, 3, 2)),
             ((24, 12, 8), 13, 1, None), ((40, 20, 12), 14, 2, (5, 2, 1)),
             ((20, 10, 10), 15, 1, (2, 1, 1))]
    ops, loads = zip(*[_design(g, s, n, gpu, bx) for g, s, n, bx in specs])
    bs = BatchSolver(ops, loads)
    us, stats = bs.solve(rtol=1e-8)
    plain_ops, _ = zip(*[_design(g, s, n, gpu) for g, s, n, _ in specs])
    bp = BatchSolver(plain_ops, loads)
    _, stats_p = bp.solve(rtol=1e-8)
    for (d, k), u, st, sp in zip(bs.systems, us, stats, stats_p):
        assert st["converged"] and st["rel"] <= 1e-8
        res, eu, eC = _check_against_single(plain_ops[d], loads[d][k], u)
        print(d, k, st["iterations"], sp["iterations"], st["outer_iterations"],
              "res %.1e err u %.1e C %.1e" % (res, eu, eC), flush=True)
        assert res <= 1e-6 and eu <= 1e-5 and eC <= 1e-7
        if specs[d][3] is not None:
            assert st["iterations"] < 0.6 * sp["iterations"]
        else:
            assert st["iterations"] == sp["iterations"]
    # new densities keep the batch valid; new groups need a new batch
    x_new = ops[0].x * 0.9
    ops[0].set_density(x_new)
    plain_ops[0].set_density(x_new)
    us2, stats2 = bs.solve(rtol=1e-8)
    assert all(st["converged"] for st in stats2)
    res, eu, eC = _check_against_single(plain_ops[0], loads[0][0], us2[0])
    assert res <= 1e-6 and eu <= 1e-5 and eC <= 1e-7
    from pyto_amd import deflation, problems
    mesh0, _, _ = problems.cantilever(specs[0][0], load="point")
    ds = deflation.DeflationSolver()
    ds.create_box_groups(mesh0, (2, 2, 2))
    ds.attach(ops[0])
    with pytest.raises(RuntimeError, match="groups changed"):
        bs.solve(rtol=1e-8)
    bs.close()
    bp.close()


def test_deflated_batch_many_load_cases(gpu):
    """A design with 10 load cases (the coarse GEMV serves at most 8 systems
    of one design per run) and the same design listed again after another
    one: every system against single solves."""
    from pyto_amd.batch import BatchSolver
    print("deflated batch, 10 load cases", flush=True)
    op_a, loads_a = _design((32, 16, 16), 31, 10, gpu, (4, 2, 2))
    op_b, loads_b = _design((24, 12, 8), 32, 2, gpu, (3, 1, 1))
    bs = BatchSolver([op_a, op_b, op_a], [loads_a, loads_b, loads_a[:3]])
    us, stats = bs.solve(rtol=1e-8)
    assert bs.nsys == 15
    ops, loads = [op_a, op_b, op_a], [loads_a, loads_b, loads_a[:3]]
    for (d, k), u, st in zip(bs.systems, us, stats):
        assert st["converged"] and st["rel"] <= 1e-8
        res, eu, eC = _check_against_single(ops[d], loads[d][k], u)
        print(d, k, st["iterations"], "res %.1e err u %.1e C %.1e" % (res, eu, eC), flush=True)
        assert res <= 1e-6 and eu <= 1e-5 and eC <= 1e-7
    # the third design repeats the first one's first three load cases
    for k in range(3):
        assert torch.equal(us[12 + k], us[k])
    bs.close()


def test_deflated_batch_more_groups_than_the_lds_table(gpu):
    """A design with 640 groups (beyond the 512 staged in LDS by the batched
    direction update) beside a small one: both against single solves."""
    from pyto_amd.batch import BatchSolver
    print("deflated batch, 640 groups", flush=True)
    op_a, loads_a = _design((48, 24, 24), 41, 2, gpu, (10, 8, 8))
    op_b, loads_b = _design((24, 12, 8), 42, 1, gpu, (3, 1, 1))
    assert op_a.ngroups > 512
    bs = BatchSolver([op_a, op_b], [loads_a, loads_b])
    us, stats = bs.solve(rtol=1e-8)
    ops, loads = [op_a, op_b], [loads_a, loads_b]
    for (d, k), u, st in zip(bs.systems, us, stats):
        assert st["converged"] and st["rel"] <= 1e-8
        res, eu, eC = _check_against_single(ops[d], loads[d][k], u)
        print(d, k, st["iterations"], "res %.1e err u %.1e C %.1e" % (res, eu, eC), flush=True)
        assert res <= 1e-6 and eu <= 1e-5 and eC <= 1e-7
    bs.close()


def test_deflated_batch_brick_elements(gpu):
    """Brick elements (the non-cube element form in the batched restriction
    and operator): deflated and plain designs against single solves."""
    from pyto_amd.batch import BatchSolver
    print("deflated batch, bricks", flush=True)
    h = (0.5, 1.0, 1.5)
    specs = [((24, 12, 8), 21, 2, (3, 2, 1)), ((16, 8, 8), 22, 1, None),
             ((32, 16, 8), 23, 2, (4, 2, 2))]
    ops, loads = zip(*[_design(g, s, n, gpu, bx, h) for g, s, n, bx in specs])
    bs = BatchSolver(ops, loads)
    us, stats = bs.solve(rtol=1e-8)
    for (d, k), u, st in zip(bs.systems, us, stats):
        assert st["converged"] and st["rel"] <= 1e-8
        res, eu, eC = _check_against_single(ops[d], loads[d][k], u)
        print(d, k, st["iterations"], st["outer_iterations"],
              "res %.1e err u %.1e C %.1e" % (res, eu, eC), flush=True)
        assert res <= 1e-6 and eu <= 1e-5 and eC <= 1e-7
    bs.close()


def test_batch_autotune_and_resolve(gpu):
    """The autotuner picks a tile shape (rows x depth) among the candidates;
    solutions after re-tuning (new graph) and a repeated solve (graph reuse)
    agree, and every shape gives the same operator bits."""
    from pyto_amd.batch import BatchSolver
    print("batch autotune", flush=True)
    specs = [((32, 16, 16), 7, 2), ((48, 24, 16), 8, 1)]
    ops, loads = zip(*[_design(g, s, n, gpu) for g, s, n in specs])
    bs = BatchSolver(ops, loads)
    us0, _ = bs.solve(rtol=1e-9)
    y16 = [op.matvec(u) for op, u in zip(ops, us0[::2])]
    tune = bs.autotune(candidates=(4, 8, 16, 32))
    assert tune["zchunk"] in (4, 8, 16, 32) and bs.info()["zchunk"] == tune["zchunk"]
    assert tune["rows"] in (8, 16, 32) and bs.tile_rows == tune["rows"]
    assert all(op.tile_rows == tune["rows"] for op in ops)
    assert sorted(tune["ms_per_matvec"]) == [8, 16, 32]
    for op, u, y in zip(ops, us0[::2], y16):
        assert torch.equal(op.matvec(u), y)
    us1, st1 = bs.solve(rtol=1e-9)
    us2, st2 = bs.solve(rtol=1e-9)
    for a, b, c in zip(us0, us1, us2):
        assert rel_err(b.cpu().numpy(), a.cpu().numpy()) <= 1e-7
        assert torch.equal(b, c)
    assert [s["iterations"] for s in st1] == [s["iterations"] for s in st2]


def test_batch_32_designs(gpu):
    """The batching workload: 32 designs (seed 5) with 1-6 point-load cases
    each, all systems to rtol 1e-8 in one batch; a sample checked against
    single-system solves."""
    from pyto_amd import problems
    from pyto_amd.batch import BatchSolver
    from pyto_amd.operator import GridOperator
    print("batch 32 designs", flush=True)
    designs = problems.batch_designs()
    ops = []
    for d in designs:
        op = GridOperator(d["grid"], (1, 1, 1), 1.0, 0.3, d["bc"].fixed_dofs, penal=3.0, eps_void=1e-8)
        op.set_density(torch.from_numpy(d["x"]).to(gpu))
        ops.append(op)
    bs = BatchSolver(ops, [d["loads"] for d in designs])
    bs.autotune()
    us, stats = bs.solve(rtol=1e-8, max_iters=200000, max_outer=60)
    assert all(s["converged"] for s in stats), [(i, s) for i, s in enumerate(stats) if not s["converged"]][:3]
    print("systems", bs.nsys, "max inner iterations", max(s["iterations"] for s in stats), flush=True)
    for i in (0, 17, 60, bs.nsys - 1):
        d, k = bs.systems[i]
        res, eu, eC = _check_against_single(ops[d], designs[d]["loads"][k], us[i])
        assert res <= 1e-6 and eu <= 1e-5 and eC <= 1e-7, (i, res, eu, eC)


def test_batch_refuses_thermal_handle(gpu):
    """A thermal handle (one value per node) cannot join a batch: creation
    fails before any kernel runs."""
    from pyto_amd.batch import BatchSolver
    from pyto_amd.operator import ThermalOperator
    print("batch refuses thermal handle", flush=True)
    op = ThermalOperator((8, 4, 4), (1.0, 1.0, 1.0), 1.0, [0, 1, 2], penal=1.0, k_void=1e-3)
    op.set_density(np.full(op.num_elems, 0.5))
    f = np.zeros(op.num_dofs)
    f[-1] = 1.0
    with pytest.raises(ValueError, match="structural handles only"):
        BatchSolver([op], [[f]])


def test_batched_density_update_matches_sequential(gpu):
    """pyto_set_density_batch (one batched Cholesky and inverse of all the
    designs' coarse matrices, padded with identity to the largest order)
    against per-design pyto_set_density: the same Galerkin matrices bit for
    bit, and FP64 solves that agree to rounding."""
    from pyto_amd.batch import set_density_batch
    print("batched density update", flush=True)
    specs = [((32, 16, 16), 41, 1, (4, 2, 2)), ((24, 12, 8), 42, 1, (3, 1, 1)),
             ((40, 20, 12), 43, 1, (5, 2, 1)), ((16, 8, 8), 44, 1, None)]
    ops, loads = zip(*[_design(g, s, n, gpu, bx) for g, s, n, bx in specs])
    rng = np.random.default_rng(45)
    xs = [torch.from_numpy(rng.uniform(0.05, 1.0, op.num_elems)).to(gpu) for op in ops]
    ref = []
    for op, x, fl in zip(ops, xs, loads):
        op.set_density(x)
        E = op.coarse_matrix().clone() if op.ngroups else None
        u, st = op.solve(torch.from_numpy(fl[0]).to(gpu), rtol=1e-11, max_iters=20000)
        ref.append((E, u, st["iterations"]))
    for op in ops:
        op.set_density(torch.full((op.num_elems,), 0.5, dtype=torch.float64, device=gpu))
    set_density_batch(ops, xs)
    for op, x, fl, (E, u0, it0) in zip(ops, xs, loads, ref):
        assert torch.equal(op.x, x)
        if E is not None:
            assert torch.equal(op.coarse_matrix(), E)
        u, st = op.solve(torch.from_numpy(fl[0]).to(gpu), rtol=1e-11, max_iters=20000)
        print(op.grid, st["iterations"], it0, float((u - u0).abs().max() / u0.abs().max()),
              flush=True)
        assert abs(st["iterations"] - it0) <= 2
        assert float((u - u0).abs().max() / u0.abs().max()) <= 1e-9
    with pytest.raises(ValueError):
        set_density_batch([ops[0], ops[0]], [xs[0], xs[0]])


def test_batch_matches_oracle_direct_solves(gpu):
    """Three small designs (two deflated, one plain; one with two load
    cases) solved in one batch to rtol 1e-10 against the oracle's sparse
    direct solves of the assembled K (same densities, fixed DOFs and SIMP
    law)."""
    from oracle import ke as OK
    from oracle import mesh as OM
    from oracle import sensitivities as OS
    from pyto_amd.batch import BatchSolver
    print("batch vs oracle direct", flush=True)
    specs = [((12, 6, 4), 51, 2, (2, 1, 1)), ((16, 8, 4), 52, 1, None),
             ((10, 6, 6), 53, 1, (2, 1, 1))]
    ops, loads = zip(*[_design(g, s, n, gpu, bx) for g, s, n, bx in specs])
    bs = BatchSolver(ops, loads)
    us, stats = bs.solve(rtol=1e-10, inner_rtol=1e-2)
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    for (d, k), u, st in zip(bs.systems, us, stats):
        assert st["converged"]
        op = ops[d]
        m = OM.GridMesh(op.grid, (1, 1, 1))
        m.createEdofMatStructural()
        fixed = np.nonzero(np.repeat(op.fixed_mask.cpu().numpy(), 3) &
                           np.tile([1, 2, 4], op.num_nodes).astype(np.uint8))[0]
        solve, _ = OS.direct_solver(m, KE, op.x.cpu().numpy(), fixed, np.zeros(fixed.size),
                                    3.0, 1e-8)
        u_ref = solve(loads[d][k])
        ug = u.cpu().numpy()
        f = loads[d][k]
        eu = rel_err(ug, u_ref)
        eC = abs(f @ ug - f @ u_ref) / abs(f @ u_ref)
        print(d, k, st["iterations"], "err u %.1e C %.1e" % (eu, eC), flush=True)
        assert eu <= 1e-7 and eC <= 1e-9
    bs.close()


def test_grid_barrier_probe(gpu):
    """The persistent-kernel measurement: a grid-wide barrier costs a finite,
    positive number of microseconds at 1 and 2 workgroups per CU (every wait
    bounded; a timed-out wait would raise), and bad arguments are refused."""
    from pyto_amd.batch import measure_grid_barrier
    print("grid barrier probe", flush=True)
    for bpc in (1, 2):
        us = measure_grid_barrier(bpc, rounds=500)
        print("blocks per CU", bpc, "us per barrier", us, flush=True)
        assert 0.0 < us < 1000.0
    with pytest.raises(ValueError):
        measure_grid_barrier(0, rounds=10)


def test_block_dpcg_matches_single_solves(gpu):
    """Block deflated PCG over a design's load cases (shared Krylov space):
    every column agrees with the design's own single-system FP64 solve and
    meets rtol, in no more iterations than the per-column form of the same
    code (s = 1) takes for the hardest column; Jacobi-only as well."""
    from pyto_amd.batch import block_dpcg
    print("block dpcg", flush=True)
    for boxes in ((3, 1, 1), None):
        op, loads = _design((32, 16, 16), 51, 5, gpu, boxes)
        U, st = block_dpcg(op, loads, rtol=1e-9)
        assert st["converged"] and st["rel"] <= 1e-9
        single = [block_dpcg(op, [f], rtol=1e-9)[1]["iterations"] for f in loads]
        print(boxes, "block", st["iterations"], "per column", single, flush=True)
        assert st["iterations"] <= max(single)
        for f, u in zip(loads, U):
            ref, _ = op.solve(torch.from_numpy(f).to(gpu), rtol=1e-11, max_iters=20000)
            fd = torch.from_numpy(f).to(gpu)
            C, Cr = float(torch.dot(fd, u)), float(torch.dot(fd, ref))
            assert abs(C - Cr) <= 1e-7 * abs(Cr)
