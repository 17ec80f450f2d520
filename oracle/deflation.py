"""Deflation groups, rigid-body W and the deflated PCG (oracle; test infra only).

Restates ``src/deflation.py``:
  * ``getRecommendedNumberOfGroups`` ``:70-84`` and the OC driver's
    ``min(2000, max(10, round(3*N_nodes/500)))`` (src/topopt_ocm.py:35);
  * ``create_deflation_groups`` box partition + small-group merge ``:86-229``
    (the blocking ``input()`` at ``:219-222`` becomes a ValueError);
  * ``create_deflation_matrix`` rigid-body W with column map
    u:{0,4,5}, v:{1,3,5}, w:{2,3,4} ``:386-458``;
  * ``deflatedPCG_CPU`` ``:504-584``: E = sym(W^T K W) + 1e-10 I, deflated
    start, two products with K per iteration, stop when
    sqrt(rz_new/rz0) <= rtol, maxIters 2000.

By default the coarse matrix is factorised once with SuperLU (``splu``)
instead of being refactorised by ``spsolve`` on every call
(src/deflation.py:533,547,575): the factors, and so the solves, are the same
up to rounding.  ``coarse="spsolve"`` keeps the reference's per-call
refactorisation (the CPU baseline times that form).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def recommended_groups_ocm(num_nodes, max_groups=2000, min_groups=10, dof_per_group=500):
    """src/topopt_ocm.py:35."""
    return min(max_groups, max(min_groups, round(3 * num_nodes / dof_per_group)))


def create_groups(node_xyz, n_groups_desired, min_nodes_per_group=15):
    """src/deflation.py:86-229.  Returns (nodeGroupNumber, groupCount, groupCenter)."""
    num_nodes = node_xyz.shape[0]
    n_groups_desired = min(n_groups_desired, int(num_nodes / (1 + min_nodes_per_group)))
    xyz = node_xyz
    mn = xyz.min(axis=0)
    L = xyz.max(axis=0) - mn
    alpha = (n_groups_desired / (L[0] * L[1] * L[2])) ** (1.0 / 3)
    nX = max(round(alpha * L[0]), 1)
    nY = max(round(alpha * L[1]), 1)
    nZ = max(round(alpha * L[2]), 1)
    ntent = nX * nY * nZ
    size = np.array([L[0] / nX, L[1] / nY, L[2] / nZ])
    idx = np.floor((xyz - mn) / size).astype(np.int32)
    idx = np.minimum(idx, np.array([nX - 1, nY - 1, nZ - 1]))
    ngn = (idx[:, 0] + nX * idx[:, 1] + nX * nY * idx[:, 2]).astype(np.int32)
    count = np.bincount(ngn, minlength=ntent)
    center = np.zeros((ntent, 3))
    for i in range(3):
        np.add.at(center[:, i], ngn, xyz[:, i])
    nz = count > 0
    center[nz] /= count[nz, None]
    mapping = np.zeros(ntent, dtype=np.int32)
    cur = 0
    for g in range(ntent):
        if count[g] < min_nodes_per_group:
            mapping[g] = -1
        else:
            mapping[g] = cur
            cur += 1
    n_groups = cur
    if np.any(mapping == -1):
        small = np.where((mapping == -1) & (count > 0))[0]
        valid = np.where(count >= min_nodes_per_group)[0]
        if len(small) > 0:
            d = np.linalg.norm(center[small][:, None, :] - center[valid][None, :, :], axis=2)
            mapping[small] = mapping[valid[np.argmin(d, axis=1)]]
        ngn = mapping[ngn]
        count = np.bincount(ngn, minlength=n_groups)
        center = np.zeros((n_groups, 3))
        for i in range(3):
            np.add.at(center[:, i], ngn, xyz[:, i])
        center /= count[:, None]
    if np.any(count < min_nodes_per_group):
        raise ValueError("deflation group with fewer than %d nodes" % min_nodes_per_group)
    if np.sum(count) != num_nodes:
        raise ValueError("invalid assignment of nodes to groups")
    return ngn.astype(np.int32), count, center


def deflation_matrix(node_xyz, ngn, center):
    """src/deflation.py:386-458; W is (3 N_nodes) x (6 G) CSR."""
    nn = node_xyz.shape[0]
    G = center.shape[0]
    rel = node_xyz - center[ngn]
    nodes = np.arange(nn)
    g6 = 6 * ngn.astype(np.int64)
    rows = np.concatenate([np.repeat(3 * nodes + c, 3) for c in range(3)])
    cols = np.concatenate([
        np.stack([g6, g6 + 4, g6 + 5], axis=1).ravel(),
        np.stack([g6 + 1, g6 + 3, g6 + 5], axis=1).ravel(),
        np.stack([g6 + 2, g6 + 3, g6 + 4], axis=1).ravel()])
    one = np.ones(nn)
    vals = np.concatenate([
        np.stack([one, rel[:, 2], -rel[:, 1]], axis=1).ravel(),
        np.stack([one, -rel[:, 2], rel[:, 0]], axis=1).ravel(),
        np.stack([one, rel[:, 1], -rel[:, 0]], axis=1).ravel()])
    return sp.coo_matrix((vals, (rows, cols)), shape=(3 * nn, 6 * G)).tocsr()


def deflated_pcg(K, f, W, Minv, rtol=1e-8, maxIters=2000, shift=1e-10, coarse="splu"):
    """src/deflation.py:504-584.  Minv is the Jacobi diagonal (vector).

    Returns (x, info) where info has the iteration count, the final
    sqrt(rz/rz0) and whether the loop stopped on the tolerance.  ``shift`` is
    the reference's 1e-10 diagonal shift of the coarse matrix (:521); it
    makes the deflation inexact, and below rtol ~1e-9 the iteration stalls
    and diverges (tests/test_oracle_fea.py pins both behaviours).
    ``coarse="splu"`` factorises the coarse matrix once; ``"spsolve"`` calls
    scipy's spsolve on the CSR coarse matrix at every coarse solve, which
    refactorises it each time as the reference does (:533,547,575).
    """
    WT = W.T.tocsr()
    KW = K @ W
    WKW = (WT @ KW).toarray()
    WKW = (WKW + WKW.T) / 2 + shift * np.eye(WKW.shape[0])
    if coarse == "splu":
        lu = spla.splu(sp.csc_matrix(WKW))
    elif coarse == "spsolve":
        WKW_csr = sp.csr_matrix(WKW)

        class _Refactor:
            @staticmethod
            def solve(b):
                return spla.spsolve(WKW_csr, b)
        lu = _Refactor
    else:
        raise ValueError("coarse must be 'splu' or 'spsolve'")
    mu = lu.solve(WT @ f)
    x = W @ mu
    r = f - K @ x
    z = Minv * r
    Kz = K @ z
    mu = lu.solve(WT @ Kz)
    p = z - W @ mu
    rz = r.dot(z)
    rz0 = rz
    converged = False
    rel = 1.0
    it = 0
    for it in range(maxIters):
        Kp = K @ p
        pKp = p.dot(Kp)
        alpha = rz / pKp
        x += alpha * p
        r -= alpha * Kp
        z = Minv * r
        rz_new = r.dot(z)
        rel = np.sqrt(rz_new / rz0)
        if rel <= rtol:
            converged = True
            break
        beta = rz_new / rz
        Kz = K @ z
        mu = lu.solve(WT @ Kz)
        p = z + beta * p - W @ mu
        rz = rz_new
    return x, {"iterations": it + 1, "rel": float(rel), "converged": converged,
               "coarse_dim": WKW.shape[0]}
