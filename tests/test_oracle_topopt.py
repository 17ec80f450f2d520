"""Pins of oracle/topopt.py: the connected components against
scipy.ndimage.label with full 26-connectivity (an independent labelling
algorithm; its labels run in raster order of first occurrence, which is the
element order) and hand-built grids; the force elements against a brute
force over the element corners; and the OC loop's invariants on a small
cantilever (every update meets the volume fraction, the tail leaves one
connected binary design near the volume fraction, and the final objective is
the direct-solve compliance of that design)."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import fea as OF
from oracle import ke as OK
from oracle import mesh as OM
from oracle import sensitivities as OS
from oracle import topopt as OT


def _scipy_components(dims, x, thr):
    nx, ny, nz = dims
    present = (np.asarray(x) > thr).reshape(nz, ny, nx)
    lab, n = ndi.label(present, structure=np.ones((3, 3, 3), dtype=int))
    lab = lab.ravel()
    return [np.nonzero(lab == k)[0] for k in range(1, n + 1)]


@pytest.mark.parametrize("dims,seed,thr", [((9, 7, 5), 0, 0.85), ((12, 5, 8), 1, 0.88),
                                           ((6, 6, 6), 2, 0.8), ((20, 3, 2), 3, 0.7)])
def test_components_match_scipy_label(dims, seed, thr):
    x = np.random.default_rng(seed).uniform(0, 1, int(np.prod(dims)))
    ours = OT.find_connected_components(dims, x, thr)
    ref = _scipy_components(dims, x, thr)
    assert len(ours) == len(ref) and len(ref) > 1
    for a, b in zip(ours, ref):
        assert np.array_equal(np.sort(a), b)
        assert a[0] == b.min()        # the search starts at the smallest index


def test_components_hand_built():
    dims = (5, 4, 3)
    x = np.zeros(60)

    def at(ix, iy, iz):
        return ix + 5 * (iy + 4 * iz)
    # two elements sharing only a corner: one component (corner neighbours)
    x[at(0, 0, 0)] = 1.0
    x[at(1, 1, 1)] = 1.0
    # a block two cells away: a second component
    x[at(3, 3, 2)] = x[at(4, 3, 2)] = 1.0
    # below the threshold (0.01): absent
    x[at(4, 0, 0)] = 0.01
    comps = OT.find_connected_components(dims, x)
    assert [sorted(c.tolist()) for c in comps] == [[at(0, 0, 0), at(1, 1, 1)],
                                                   [at(3, 3, 2), at(4, 3, 2)]]


def test_neighbour_array_brute_force():
    dims = (4, 3, 2)
    nbr = OT.element_neighbours(dims)
    for e in range(24):
        ix, iy, iz = e % 4, (e // 4) % 3, e // 12
        k = 0
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    j = (ix + dx, iy + dy, iz + dz)
                    inside = 0 <= j[0] < 4 and 0 <= j[1] < 3 and 0 <= j[2] < 2
                    assert nbr[e, k] == (j[0] + 4 * (j[1] + 3 * j[2]) if inside else -1)
                    k += 1


def test_elements_with_forces_brute_force():
    g = OM.GridMesh((5, 3, 2), (1, 1, 1))
    f = np.zeros(3 * g.num_nodes)
    f[3 * 7 + 1] = 2.0
    f[3 * (g.num_nodes - 1) + 2] = -1.0
    got = OT.elements_with_forces(g, f, 3)
    want = [e for e in range(g.num_elems)
            if any(f[3 * n + c] != 0 for n in g.elemArray[e] for c in range(3))]
    assert got.tolist() == want


def test_oc_loop_invariants():
    dims = (16, 6, 4)
    g = OM.GridMesh(dims, (1, 1, 1))
    g.createEdofMatStructural()
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    fixed, gv = OF.clamp_x0(g)
    f = OF.point_load(g, 16, 3, 2)
    ms = OT.direct_make_solve(g, KE, fixed, eps=1e-9)
    r = OT.oc_loop(g, KE, f, fixed, 0.4, ms, max_iterations=12, eps=1e-9)
    h = r["history"]
    assert len(h["objective"]) == r["nFEAs"] + 1
    # the bisection meets the volume fraction at every update
    assert np.all(np.abs(np.array(h["volfrac"][:-1]) - 0.4) <= 1e-6)
    assert h["objective"][r["nFEAs"] - 1] < h["objective"][0]
    xf = r["x_final"]
    assert set(np.unique(xf)) <= {0.0, 1.0}
    assert len(OT.find_connected_components(dims, xf)) == 1
    assert abs(xf.mean() - 0.4) <= 0.05
    u = ms(xf)(f)
    C, _ = OS.compliance_and_gradient(u, xf, g.edofMatStructural, KE, 3.0, 1e-9)
    assert h["objective"][-1] == C
