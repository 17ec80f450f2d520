"""Pins of the oracle's grid numbering (oracle/mesh.py) against the literal
restatement of the reference's loops."""
import numpy as np
import pytest

from oracle import mesh as M


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 2), (4, 3, 5)])
def test_vectorised_equals_loops(dims):
    g = M.GridMesh(dims, (1.0, 0.5, 2.0))
    g.createEdofMatStructural()
    ni, ea, edof = M.grid_mesh_loops(dims, (1.0, 0.5, 2.0))
    assert np.array_equal(g.node_indices, ni)
    assert np.array_equal(g.elemArray, ea)
    assert np.array_equal(g.edofMatStructural, edof)


def test_planes_and_centres():
    g = M.GridMesh((3, 2, 2), (1, 1, 1))
    assert np.array_equal(g.getNodesOnBoundingBoxPlane(0, True),
                          np.where(g.node_indices[:, 0] == 0)[0])
    assert len(g.getNodesOnBoundingBoxPlane(2, False)) == 4 * 3
    assert np.allclose(g.elem_centers[0], [0.5, 0.5, 0.5])
    assert np.allclose(g.elem_centers[-1], [2.5, 1.5, 1.5])
