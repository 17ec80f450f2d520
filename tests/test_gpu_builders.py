"""Device problem builders (pyto_amd.builders) against the oracle's mesh,
loads and deflation groups and the host mirror: every output bit-identical."""
import numpy as np
import pytest
import torch

from oracle import deflation as ODF
from oracle import fea as OF
from oracle import mesh as OM

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid,h", [((7, 5, 3), (0.7, 1.3, 0.9)), ((33, 17, 9), (1.0, 1.0, 1.0))])
def test_grid_arrays_match_oracle(gpu, grid, h):
    from pyto_amd import builders
    print("grid arrays", grid, h, flush=True)
    m = OM.GridMesh(grid, h)
    m.createEdofMatStructural()
    out = builders.grid_arrays(grid, h, gpu)
    assert np.array_equal(out["node_xyz"].cpu().numpy(), m.node_xyz)
    assert np.array_equal(out["elemArray"].cpu().numpy(), np.asarray(m.elemArray, np.int32))
    assert np.array_equal(out["edofMatStructural"].cpu().numpy(),
                          np.asarray(m.edofMatStructural, np.int64))


def test_fixed_mask_matches_clamp_and_brute_force(gpu):
    from pyto_amd import builders
    from pyto_amd.operator import fixed_mask_from_dofs
    print("fixed mask", flush=True)
    grid = (9, 6, 4)
    m = OM.GridMesh(grid, (1, 1, 1))
    fixed, _ = OF.clamp_x0(m)
    want = fixed_mask_from_dofs(m.num_nodes, fixed)
    got = builders.fixed_mask(grid, [(0, False, 7)], gpu).cpu().numpy()
    assert np.array_equal(got, want)
    planes = [(0, True, 4), (1, False, 1), (1, True, 2), (2, False, 3), (2, True, 6)]
    got = builders.fixed_mask(grid, planes, gpu).cpu().numpy()
    ref = np.zeros(m.num_nodes, np.uint8)
    ijk = m.node_indices[:, :3]
    for axis, at_max, bits in planes:
        sel = ijk[:, axis] == (grid[axis] if at_max else 0)
        ref[sel] |= bits
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("grid", [(60, 20, 10), (5, 1, 3), (256, 128, 128)])
def test_face_load_matches_oracle(gpu, grid):
    from pyto_amd import builders, problems
    print("face load", grid, flush=True)
    m = OM.GridMesh(grid, (1, 1, 1))
    got = builders.face_load(grid, 0, True, 2, 1.0, device=gpu).cpu().numpy()
    assert np.array_equal(got, OF.face_traction_load(m))
    if grid[0] <= 60:
        _, _, bc = problems.cantilever(grid)
        assert np.array_equal(got, bc.force)
    # another face and component against a brute-force restatement
    got = builders.face_load(grid, 2, False, 0, -3.5, device=gpu).cpu().numpy()
    ijk = m.node_indices[:, :3]
    sel = np.nonzero(ijk[:, 2] == 0)[0]
    ext = ((ijk[sel, 0] == 0) | (ijk[sel, 0] == grid[0])).astype(int) + \
        ((ijk[sel, 1] == 0) | (ijk[sel, 1] == grid[1])).astype(int)
    w = np.where(ext == 2, 1.0, np.where(ext == 1, 2.0, 4.0))
    ref = np.zeros(3 * m.num_nodes)
    ref[3 * sel] = -w * (-3.5 / w.sum())
    assert np.array_equal(got, ref)


def _host_groups(grid, h, boxes=None, ngroups=None):
    from pyto_amd import deflation, hex_mesher
    mesh = hex_mesher.HexMesher().grid_mesh(grid, h)
    ds = deflation.DeflationSolver()
    if boxes is None:
        ds.create_deflation_groups(mesh, ngroups)
    else:
        ds.create_box_groups(mesh, boxes)
    return mesh, ds


@pytest.mark.parametrize("grid,h,boxes,ngroups", [
    ((60, 20, 10), (1, 1, 1), None, 96),           # the reference heuristic, 12 x 4 x 2
    ((33, 17, 9), (0.7, 1.3, 0.9), (4, 3, 2), None),
    ((6, 2, 2), (1, 1, 1), (5, 1, 1), None),        # small boxes merged
    ((10, 3, 3), (0.5, 1.0, 2.0), (9, 2, 1), None),  # small boxes merged, brick elements
    ((256, 128, 128), (1, 1, 1), (4, 4, 4), None),
    ((256, 128, 128), (1, 1, 1), None, 2000),       # 20 x 10 x 10
    # the hand-built merge cases whose groups tests/test_oracle_groups.py
    # writes out: nearest valid box, distance tie to the first valid box,
    # empty box skipped
    ((6, 2, 3), (1, 1, 1), None, 5),
    ((4, 2, 3), (1.0, 0.5, 0.5), None, 3),
    ((2, 5, 5), (1.0, 0.125, 0.125), None, 6),
])
def test_box_groups_match_host_and_oracle(gpu, grid, h, boxes, ngroups):
    from pyto_amd import deflation
    print("groups", grid, h, boxes, ngroups, flush=True)
    mesh, ds = _host_groups(grid, h, boxes, ngroups)
    dd = deflation.DeflationSolver()
    if boxes is None:
        dd.create_deflation_groups(mesh, ngroups, device=gpu)
    else:
        dd.create_box_groups(mesh, boxes, device=gpu)
    assert dd.ws_nGroups == ds.ws_nGroups
    assert np.array_equal(dd.ws_nodeGroupNumber.cpu().numpy(), ds.ws_nodeGroupNumber)
    assert np.array_equal(dd.ws_groupCount, ds.ws_groupCount)
    assert np.array_equal(dd.ws_groupCenter, ds.ws_groupCenter)   # bitwise
    if boxes is None and grid[0] <= 60:
        ngn, cnt, cen = ODF.create_groups(OM.GridMesh(grid, h).node_xyz, ngroups)
        assert np.array_equal(dd.ws_nodeGroupNumber.cpu().numpy(), ngn)
        assert np.array_equal(dd.ws_groupCenter, cen)


def test_solve_from_device_built_problem_is_identical(gpu):
    """A DPCG solve whose mask, load and groups come from the device builders
    returns the same bits as the one built on the host."""
    from pyto_amd import builders, problems
    from pyto_amd.operator import GridOperator
    print("device-built solve", flush=True)
    grid = (24, 8, 8)
    mesh, mat, bc = problems.cantilever(grid)
    x = problems.density_uniform(mesh.num_elems, 0.1, 1.0, 3)
    _, ds = _host_groups(grid, (1, 1, 1), (3, 1, 1))
    op_h = GridOperator(grid, (1, 1, 1), 1.0, 0.3, bc.fixed_dofs)
    op_h.set_coarse_shift(0.0, 0.0)
    op_h.set_groups(ds.ws_nodeGroupNumber, ds.ws_groupCenter)
    op_h.set_density(x)
    u_h, _ = op_h.solve(bc.force, rtol=1e-10)
    mask = builders.fixed_mask(grid, [(0, False, 7)], gpu)
    f = builders.face_load(grid, device=gpu)
    ngn, cen, _ = builders.box_groups(grid, (1, 1, 1), (3, 1, 1), device=gpu)
    op_d = GridOperator(grid, (1, 1, 1), 1.0, 0.3, None, fixed_mask=mask)
    op_d.set_coarse_shift(0.0, 0.0)
    op_d.set_groups(ngn, cen)
    op_d.set_density(x)
    u_d, _ = op_d.solve(f, rtol=1e-10)
    assert torch.equal(u_h, u_d)
