"""Pins of the oracle's deflation-group heuristic (oracle.deflation.
create_groups, restating src/deflation.py:86-229) by what the mathematics
fixes, independently of the restatement:

* box counts per axis: n = round(alpha L) with alpha^3 = G / (Lx Ly Lz),
  decided in exact rational arithmetic ((n - 1/2)^3 <= G L^3 / V <= (n + 1/2)^3,
  ties to even as Python's round()),
  giving the survey's figures 12x4x2, 26x9x9, 20x10x10 and 20x10x10 at the
  four workload grids;
* node -> box assignment by brute force in integer arithmetic
  (box = min(i n // N, n - 1) per axis on a unit grid);
* box centres as closed-form lattice means ((first + last) / 2 node index
  per axis);
* the small-box merge (src/deflation.py:178-211) on hand-built grids whose
  expected groups are written out by hand: a merge into the nearest valid
  box, a distance tie resolved to the first valid box (numpy argmin), an
  empty box skipped, and the cap G <= int(N / 16).
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import deflation as ODF
from oracle import mesh as OM


def _exact_round_cuberoot(num, den):
    """round((num/den)^(1/3)) for num/den >= 1/8, in exact arithmetic: the
    n with (n - 1/2)^3 <= num/den <= (n + 1/2)^3, an exact tie going to the
    even neighbour as Python's round() does."""
    q = Fraction(num, den)
    n = 1
    while Fraction(2 * n + 1, 2) ** 3 < q or (Fraction(2 * n + 1, 2) ** 3 == q and n % 2 == 1):
        n += 1
    return n


def _boxes_exact(dims, G):
    V = dims[0] * dims[1] * dims[2]
    return tuple(max(_exact_round_cuberoot(G * L ** 3, V), 1) for L in dims)


def _requested(num_nodes):
    # src/topopt_ocm.py:35, then the cap of src/deflation.py:108-109
    G = min(2000, max(10, round(3 * num_nodes / 500)))
    return G, min(G, num_nodes // 16)


@pytest.mark.parametrize("dims,expected", [((60, 20, 10), (12, 4, 2)),
                                           ((192, 64, 64), (26, 9, 9)),
                                           ((256, 128, 128), (20, 10, 10)),
                                           ((128, 64, 64), (20, 10, 10))])
def test_box_counts_closed_form(dims, expected):
    nn = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    G, Gc = _requested(nn)
    assert _boxes_exact(dims, Gc) == expected


def _lattice_groups(dims, boxes):
    """Exact integer assignment and closed-form centres of an unmerged box
    partition of a unit grid (node index = ix + sx (iy + sy iz))."""
    per_axis = []
    for L, n in zip(dims, boxes):
        i = np.arange(L + 1)
        per_axis.append(np.minimum(i * n // L, n - 1))
    bx, by, bz = per_axis
    sx, sy, sz = (d + 1 for d in dims)
    ix = np.arange(sx * sy * sz) % sx
    iy = (np.arange(sx * sy * sz) // sx) % sy
    iz = np.arange(sx * sy * sz) // (sx * sy)
    ngn = bx[ix] + boxes[0] * (by[iy] + boxes[1] * bz[iz])
    cen = np.zeros((boxes[0] * boxes[1] * boxes[2], 3))
    cnt = np.zeros(len(cen), dtype=np.int64)
    for gz in range(boxes[2]):
        for gy in range(boxes[1]):
            for gx in range(boxes[0]):
                g = gx + boxes[0] * (gy + boxes[1] * gz)
                rng = [np.flatnonzero(b == k) for b, k in ((bx, gx), (by, gy), (bz, gz))]
                cen[g] = [(r[0] + r[-1]) / 2.0 for r in rng]
                cnt[g] = len(rng[0]) * len(rng[1]) * len(rng[2])
    return ngn, cnt, cen


@pytest.mark.parametrize("dims", [(60, 20, 10), (192, 64, 64), (128, 64, 64), (64, 32, 32)])
def test_assignment_and_centres_brute_force(dims):
    m = OM.GridMesh(dims, (1, 1, 1))
    G, Gc = _requested(m.num_nodes)
    ngn, cnt, cen = ODF.create_groups(m.node_xyz, G)
    boxes = _boxes_exact(dims, Gc)
    ngn_x, cnt_x, cen_x = _lattice_groups(dims, boxes)
    assert cnt_x.min() >= 15            # no merge at these sizes
    assert np.array_equal(ngn, ngn_x)
    assert np.array_equal(cnt, cnt_x)
    assert np.abs(cen - cen_x).max() <= 1e-12 * max(dims)


def test_merge_into_nearest_valid_box():
    """6x2x3 unit grid, 5 groups: boxes 3 (x) x 1 (y) x 2 (z); the x slabs
    {0,1}, {2,3} hold 12 nodes per z half (< 15), {4,5,6} 18.  Each small box
    is nearest to the valid box of its own z half (distance 4.5 or 2.5
    against sqrt(4.5^2 + 2^2) or more), so the groups are the two z halves."""
    m = OM.GridMesh((6, 2, 3), (1, 1, 1))
    ngn, cnt, cen = ODF.create_groups(m.node_xyz, 5)
    iz = np.arange(m.num_nodes) // (7 * 3)
    assert np.array_equal(ngn, (iz >= 2).astype(np.int32))
    assert list(cnt) == [42, 42]
    assert np.array_equal(cen, np.array([[3.0, 1.0, 0.5], [3.0, 1.0, 2.5]]))


def test_merge_tie_goes_to_first_valid_box():
    """4x2x3 grid of 1 x 0.5 x 0.5 bricks, 3 groups: 3 x-boxes of width 4/3
    hold node slabs {0,1} (24 nodes), {2} (12) and {3,4} (24).  The small
    middle box's centre x = 2 is 1.5 from both valid centres (0.5 and 3.5,
    same y, z): the tie goes to the first valid box, as numpy's argmin does."""
    m = OM.GridMesh((4, 2, 3), (1.0, 0.5, 0.5))
    ngn, cnt, cen = ODF.create_groups(m.node_xyz, 3)
    ix = np.arange(m.num_nodes) % 5
    assert np.array_equal(ngn, (ix >= 3).astype(np.int32))
    assert list(cnt) == [36, 24]
    assert np.array_equal(cen, np.array([[1.0, 0.5, 0.75], [3.5, 0.5, 0.75]]))


def test_empty_box_is_skipped_and_groups_renumbered():
    """2x5x5 grid of 1 x 0.125 x 0.125 bricks, 6 groups: 4 x-boxes of width
    0.5 for node x in {0, 1, 2}; box 1 holds no node, boxes 0, 2, 3 hold one
    36-node slab each and become groups 0, 1, 2."""
    m = OM.GridMesh((2, 5, 5), (1.0, 0.125, 0.125))
    ngn, cnt, cen = ODF.create_groups(m.node_xyz, 6)
    ix = np.arange(m.num_nodes) % 3
    assert np.array_equal(ngn, ix.astype(np.int32))
    assert list(cnt) == [36, 36, 36]
    assert np.allclose(cen[:, 0], [0.0, 1.0, 2.0], rtol=0, atol=0)


def test_requested_groups_capped_by_node_count():
    """4x2x3 unit grid: 60 nodes cap the request at int(60/16) = 3 groups;
    alpha = (3/24)^(1/3) = 1/2 gives boxes 2 x 1 x 2 (an exact tie, 1.5 rounds
    to 2); the x slabs {0,1} and {2,3,4} times the z halves hold 12 and 18
    nodes, so the two small boxes merge and 2 groups of 30 remain."""
    m = OM.GridMesh((4, 2, 3), (1, 1, 1))
    ngn, cnt, cen = ODF.create_groups(m.node_xyz, 500)
    ngn3, cnt3, cen3 = ODF.create_groups(m.node_xyz, 3)
    assert np.array_equal(ngn, ngn3) and np.array_equal(cnt, cnt3)
    assert _boxes_exact((4, 2, 3), 3) == (2, 1, 2)
    assert _exact_round_cuberoot(5 ** 3, 8) == 2      # round(2.5) == 2
    assert list(cnt) == [30, 30]
