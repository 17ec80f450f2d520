"""Host-side logic of the package (no GPU): grid numbering, BC masks,
deflation groups, workload builders, against the oracle."""
import numpy as np
import pytest

from oracle import deflation as ODF
from oracle import fea as OF
from oracle import mesh as OM
from pyto_amd import deflation, hex_mesher, problems
from pyto_amd.operator import fixed_mask_from_dofs


@pytest.mark.parametrize("dims,h", [((3, 2, 2), (1, 1, 1)), ((5, 4, 3), (0.5, 1.0, 2.0))])
def test_mesher_equals_oracle(dims, h):
    m = hex_mesher.HexMesher().grid_mesh(dims, h)
    o = OM.GridMesh(dims, h)
    o.createEdofMatStructural()
    assert np.array_equal(m.node_indices, o.node_indices)
    assert np.allclose(m.node_xyz, o.node_xyz)
    assert np.array_equal(m.elemArray, o.elemArray)
    assert np.array_equal(m.edofMatStructural, o.edofMatStructural)
    assert np.allclose(m.elem_centers, o.elem_centers)
    for ax in range(3):
        for lo in (True, False):
            assert np.array_equal(m.getNodesOnBoundingBoxPlane(ax, lo), o.getNodesOnBoundingBoxPlane(ax, lo))


@pytest.mark.parametrize("dims,ng,h", [((60, 20, 10), None, (1, 1, 1)), ((12, 6, 4), None, (1, 1, 1)),
                                       ((30, 7, 5), 40, (1, 1, 1)), ((8, 8, 8), 300, (1, 1, 1)),
                                       # the hand-built merge cases of test_oracle_groups
                                       ((6, 2, 3), 5, (1, 1, 1)), ((4, 2, 3), 3, (1.0, 0.5, 0.5)),
                                       ((2, 5, 5), 6, (1.0, 0.125, 0.125))])
def test_deflation_groups_equal_oracle(dims, ng, h):
    m = hex_mesher.HexMesher().grid_mesh(dims, h)
    if ng is None:
        ng = ODF.recommended_groups_ocm(m.num_nodes)
    d = deflation.DeflationSolver()
    d.create_deflation_groups(m, ng)
    ngn, cnt, cen = ODF.create_groups(m.node_xyz, ng)
    assert np.array_equal(d.ws_nodeGroupNumber, ngn)
    assert np.array_equal(d.ws_groupCount, cnt)
    assert np.allclose(d.ws_groupCenter, cen, rtol=0, atol=1e-12)


def test_c1_group_count():
    m = hex_mesher.HexMesher().grid_mesh((60, 20, 10), (1, 1, 1))
    d = deflation.DeflationSolver()
    d.create_deflation_groups(m, min(2000, max(10, round(3 * m.num_nodes / 500))))
    assert d.ws_nGroups == 96


def test_fixed_mask():
    m = fixed_mask_from_dofs(4, [0, 1, 2, 4, 11])
    assert list(m) == [7, 2, 0, 4]
    with pytest.raises(ValueError):
        fixed_mask_from_dofs(4, [12])


def test_workloads_match_oracle_recipes():
    mesh, mat, bc = problems.cantilever((6, 3, 4), load="traction")
    o = OM.GridMesh((6, 3, 4), (1, 1, 1))
    fixed, _ = OF.clamp_x0(o)
    assert np.array_equal(np.sort(bc.fixed_dofs), fixed)
    assert np.allclose(bc.force, OF.face_traction_load(o))
    mesh, mat, bc = problems.cantilever((60, 20, 10), load="point", point=(60, 10, 5))
    o = OM.GridMesh((60, 20, 10), (1, 1, 1))
    assert np.allclose(bc.force, OF.point_load(o, 60, 10, 5))
    x = problems.density_bumps((16, 8, 8), seed=2)
    assert x.min() >= 0.001 and x.max() == 1.0
