"""Element stiffness of the 8-node hex (oracle; test infrastructure only).

Restates ``src/hex_element_stiffness.py``:
  * 2-point Gauss rule ``:34-36`` as a 3D tensor product ``:52-55``;
  * node ordering (bottom face CCW 0-1-2-3, top 4-7) ``:63-71,103-105``;
  * shape-function gradients ``:126-140``; Jacobian ``:159-186``;
  * physical gradients ``:205-218``; isotropic D ``:220-228``;
  * B rows [xx, yy, zz, xy, yz, xz] and KE = sum B^T D B detJ w ``:254-286``;
  * thermal KE_T = sum k B^T B detJ w ``:339-385``;
  * the thermo-elastic coupling matrix of ``src/hex_thermal_fea.py:119-196``.
"""
from fractions import Fraction

import numpy as np

# Local node corner signs, src/hex_element_stiffness.py:103-105.
XIS = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=np.float64)
ETAS = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=np.float64)
ZETAS = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=np.float64)


def isotropic_D(E, nu):
    """src/hex_element_stiffness.py:220-228 (engineering shear, Voigt)."""
    c = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
    g = (1.0 - 2.0 * nu) / 2.0
    return c * np.array([[1 - nu, nu, nu, 0, 0, 0],
                         [nu, 1 - nu, nu, 0, 0, 0],
                         [nu, nu, 1 - nu, 0, 0, 0],
                         [0, 0, 0, g, 0, 0],
                         [0, 0, 0, 0, g, 0],
                         [0, 0, 0, 0, 0, g]], dtype=np.float64)


def gauss_points_3d():
    """src/hex_element_stiffness.py:34-36,52-55 (order 2, x slowest)."""
    x = np.array([-0.5773502691896257, 0.5773502691896257])
    w = np.array([1.0, 1.0])
    pts = np.array([[a, b, c] for a in x for b in x for c in x])
    wts = np.array([a * b * c for a in w for b in w for c in w])
    return pts, wts


def _iso_gradients(p):
    """dN/d(xi,eta,zeta) at one point, shape (8,3); src/...stiffness.py:135-140."""
    xi, eta, zeta = p
    d = np.empty((8, 3))
    d[:, 0] = 0.125 * XIS * (1 + eta * ETAS) * (1 + zeta * ZETAS)
    d[:, 1] = 0.125 * (1 + xi * XIS) * ETAS * (1 + zeta * ZETAS)
    d[:, 2] = 0.125 * (1 + xi * XIS) * (1 + eta * ETAS) * ZETAS
    return d


def hex8_stiffness(E, nu, elem_size):
    """24x24 KE by 2x2x2 Gauss; src/hex_element_stiffness.py:231-286.

    Row/column index = 3*local_node + component.
    """
    dx, dy, dz = elem_size
    nodes = np.array([[0, dx, dx, 0, 0, dx, dx, 0],
                      [0, 0, dy, dy, 0, 0, dy, dy],
                      [0, 0, 0, 0, dz, dz, dz, dz]], dtype=np.float64).T
    D = isotropic_D(E, nu)
    ke = np.zeros((24, 24))
    pts, wts = gauss_points_3d()
    for p, w in zip(pts, wts):
        diso = _iso_gradients(p)
        # jac[:, j] = dN/dxi_j @ xyz  (src/...stiffness.py:176-178)
        jac = np.stack([diso[:, j] @ nodes for j in range(3)], axis=1)
        jinv = np.linalg.inv(jac)
        dphys = np.array([jinv @ diso[n] for n in range(8)])  # (8,3), :214-217
        b = np.zeros((6, 24))
        for l in range(8):
            b[0, 3 * l] = dphys[l, 0]
            b[1, 3 * l + 1] = dphys[l, 1]
            b[2, 3 * l + 2] = dphys[l, 2]
            b[3, 3 * l] = dphys[l, 1]
            b[3, 3 * l + 1] = dphys[l, 0]
            b[4, 3 * l + 1] = dphys[l, 2]
            b[4, 3 * l + 2] = dphys[l, 1]
            b[5, 3 * l] = dphys[l, 2]
            b[5, 3 * l + 2] = dphys[l, 0]
        ke += b.T @ D @ b * np.linalg.det(jac) * w
    return ke


# ---------------------------------------------------------------------------
# Independent pin: exact rational integration of B^T D B over the brick.
# The integrand is a polynomial of degree <= 2 in each coordinate, so this is
# the exact stiffness; the Gauss rule above must reproduce it to rounding.
# ---------------------------------------------------------------------------
def _poly_mul(a, b):
    out = {}
    for ea, ca in a.items():
        for eb, cb in b.items():
            e = (ea[0] + eb[0], ea[1] + eb[1], ea[2] + eb[2])
            out[e] = out.get(e, 0) + ca * cb
    return out


def _poly_integrate_unit_cube(p):
    """Integral over [0,1]^3 of sum c x^i y^j z^k."""
    s = Fraction(0)
    for (i, j, k), c in p.items():
        s += Fraction(c) / ((i + 1) * (j + 1) * (k + 1))
    return s


def hex8_stiffness_exact(E, nu, elem_size):
    """Exact KE by symbolic integration on the unit cube, scaled to the brick.

    Shape functions on [0,1]^3: N = X(a) Y(b) Z(c) with X(0) = 1-x, X(1) = x.
    Physical gradient dN/dx = (1/hx) dN/dx_hat; volume element hx*hy*hz.
    """
    hx, hy, hz = elem_size
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
              (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]

    def lin(flag, axis):
        # 1-t or t as polynomial dict in the given axis
        e1 = [0, 0, 0]
        e1[axis] = 1
        if flag:
            return {tuple(e1): Fraction(1)}
        return {(0, 0, 0): Fraction(1), tuple(e1): Fraction(-1)}

    def dlin(flag):
        return {(0, 0, 0): Fraction(1 if flag else -1)}

    grads = []  # grads[n][j] polynomial of dN_n/dx_hat_j
    for (a, b, c) in corner:
        gx = _poly_mul(_poly_mul(dlin(a), lin(b, 1)), lin(c, 2))
        gy = _poly_mul(_poly_mul(lin(a, 0), dlin(b)), lin(c, 2))
        gz = _poly_mul(_poly_mul(lin(a, 0), lin(b, 1)), dlin(c))
        grads.append([gx, gy, gz])
    # integrals I[n][m][j][k] = int dN_n/dxj * dN_m/dxk over unit cube
    I = np.zeros((8, 8, 3, 3))
    for n in range(8):
        for m in range(8):
            for j in range(3):
                for k in range(3):
                    I[n, m, j, k] = float(_poly_integrate_unit_cube(
                        _poly_mul(grads[n][j], grads[m][k])))
    h = np.array([hx, hy, hz], dtype=np.float64)
    V = hx * hy * hz
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    ke = np.zeros((24, 24))
    # a(u,v) = int lam div u div v + mu sum_ij d_j u_i (d_j v_i + d_i v_j)
    for n in range(8):
        for m in range(8):
            for i in range(3):       # component of u at node n (row)
                for k in range(3):   # component of v at node m (col)
                    val = lam * I[n, m, i, k] / (h[i] * h[k])
                    if i == k:
                        val += mu * sum(I[n, m, j, j] / (h[j] * h[j]) for j in range(3))
                    val += mu * I[n, m, k, i] / (h[k] * h[i])
                    ke[3 * n + i, 3 * m + k] = V * val
    return ke


def hex8_stiffness_thermal(k, elem_size):
    """8x8 KE_T by 2x2x2 Gauss; src/hex_element_stiffness.py:339-385:
    b = dN/dx (3 x 8) at each point, KE_T = sum k b^T b detJ w."""
    dx, dy, dz = elem_size
    nodes = np.array([[0, dx, dx, 0, 0, dx, dx, 0],
                      [0, 0, dy, dy, 0, 0, dy, dy],
                      [0, 0, 0, 0, dz, dz, dz, dz]], dtype=np.float64).T
    ke = np.zeros((8, 8))
    pts, wts = gauss_points_3d()
    for p, w in zip(pts, wts):
        diso = _iso_gradients(p)
        jac = np.stack([diso[:, j] @ nodes for j in range(3)], axis=1)
        jinv = np.linalg.inv(jac)
        b = np.array([jinv @ diso[n] for n in range(8)]).T   # (3, 8)
        ke += k * b.T @ b * np.linalg.det(jac) * w
    return ke


def _corner_polys():
    """Shape functions on [0,1]^3 and their reference gradients as exact
    polynomials: N_n = X(a) Y(b) Z(c), X(0) = 1-x, X(1) = x."""
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
              (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]

    def lin(flag, axis):
        e1 = [0, 0, 0]
        e1[axis] = 1
        if flag:
            return {tuple(e1): Fraction(1)}
        return {(0, 0, 0): Fraction(1), tuple(e1): Fraction(-1)}

    def dlin(flag):
        return {(0, 0, 0): Fraction(1 if flag else -1)}

    N, grads = [], []
    for (a, b, c) in corner:
        N.append(_poly_mul(_poly_mul(lin(a, 0), lin(b, 1)), lin(c, 2)))
        grads.append([_poly_mul(_poly_mul(dlin(a), lin(b, 1)), lin(c, 2)),
                      _poly_mul(_poly_mul(lin(a, 0), dlin(b)), lin(c, 2)),
                      _poly_mul(_poly_mul(lin(a, 0), lin(b, 1)), dlin(c))])
    return N, grads


def hex8_stiffness_thermal_exact(k, elem_size):
    """Exact KE_T = k int grad N_n . grad N_m dV (rational integration on
    the unit cube, scaled to the brick)."""
    h = np.asarray(elem_size, dtype=np.float64)
    V = float(np.prod(h))
    _, grads = _corner_polys()
    ke = np.zeros((8, 8))
    for n in range(8):
        for m in range(8):
            ke[n, m] = k * V * sum(float(_poly_integrate_unit_cube(
                _poly_mul(grads[n][j], grads[m][j]))) / (h[j] * h[j]) for j in range(3))
    return ke


def thermoelastic_H(elem_size, nu, layout="reference"):
    """24 x 8 coupling matrix H with f_e = E0 alpha E(x) H (T_e - T_ref):
    H = (1/(1-2nu)) int B^T m N dV, m = [1,1,1,0,0,0] (the thermal strain
    alpha dT m through the isotropic D, D m = E/(1-2nu) m), i.e.
    H[(c, j), k] = (1/(1-2nu)) int dN_j/dx_c N_k dV (exact integration).

    layout "physical": row 3j + c (the node-major DOF order of edofMat).
    layout "reference": row 8c + j, the row order the reference's getHMatrix
    returns (src/hex_thermal_fea.py:119-196, reshaped column-major), which
    the reference then indexes as 3j + c (:224-237, and
    src/topopt_obj_cons_sensitivities.py:358-366,439-447).
    """
    h = np.asarray(elem_size, dtype=np.float64)
    V = float(np.prod(h))
    N, grads = _corner_polys()
    H = np.zeros((24, 8))
    for j in range(8):
        for c in range(3):
            row = 3 * j + c if layout == "physical" else 8 * c + j
            for k in range(8):
                H[row, k] = V / h[c] * float(_poly_integrate_unit_cube(
                    _poly_mul(grads[j][c], N[k]))) / (1.0 - 2.0 * nu)
    return H


def rigid_body_modes(elem_size):
    """Six 24-vectors spanning the rigid motions of one element."""
    dx, dy, dz = elem_size
    xyz = np.array([[0, 0, 0], [dx, 0, 0], [dx, dy, 0], [0, dy, 0],
                    [0, 0, dz], [dx, 0, dz], [dx, dy, dz], [0, dy, dz]], float)
    R = np.zeros((24, 6))
    for n in range(8):
        x, y, z = xyz[n]
        R[3 * n:3 * n + 3, 0] = [1, 0, 0]
        R[3 * n:3 * n + 3, 1] = [0, 1, 0]
        R[3 * n:3 * n + 3, 2] = [0, 0, 1]
        R[3 * n:3 * n + 3, 3] = [0, -z, y]
        R[3 * n:3 * n + 3, 4] = [z, 0, -x]
        R[3 * n:3 * n + 3, 5] = [-y, x, 0]
    return R
