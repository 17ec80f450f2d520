// Element-level arithmetic of the 8-node brick in the Hadamard ("modal") basis.
//
// Nodal values u[n][i] (n = reference local node 0..7, i = x,y,z component)
// are transformed by a 3-stage butterfly into the sums c_m = sum_n s_m(n) u_n
// for the monomials m in {x, y, z, xy, xz, yz, xyz} of the reference
// coordinates (the constant mode is a rigid translation and drops out).  The
// displacement gradient of each monomial gives one isotropic stress
// evaluation; because the monomials are orthogonal over the element, KE*u is
// the transposed butterfly of those stresses.  This reproduces KE of
// src/hex_element_stiffness.py:231-286 (2x2x2 Gauss, exact for these
// polynomials) to rounding.
#pragma once

#include "common.hpp"

namespace pyto {

// Hadamard sums per component (NC = 3: displacement, 1: temperature).
template <typename T, int NC>
struct ModalN {
  T cx[NC], cy[NC], cz[NC], cxy[NC], cxz[NC], cyz[NC], cxyz[NC];
};
template <typename T>
using Modal = ModalN<T, 3>;

// Local node order (a,b,c): 0:(000) 1:(100) 2:(110) 3:(010) 4:(001) 5:(101)
// 6:(111) 7:(011), src/hex_mesher.py:121-124.
template <typename T, int NC>
__host__ __device__ __forceinline__ void modal_forward(const T u[8][NC], ModalN<T, NC>& m) {
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const T v000 = u[0][i], v100 = u[1][i], v110 = u[2][i], v010 = u[3][i];
    const T v001 = u[4][i], v101 = u[5][i], v111 = u[6][i], v011 = u[7][i];
    // x stage
    const T sx00 = v100 + v000, dx00 = v100 - v000;
    const T sx10 = v110 + v010, dx10 = v110 - v010;
    const T sx01 = v101 + v001, dx01 = v101 - v001;
    const T sx11 = v111 + v011, dx11 = v111 - v011;
    // y stage
    const T s0_0 = sx10 + sx00, sy_0 = sx10 - sx00;
    const T s0_1 = sx11 + sx01, sy_1 = sx11 - sx01;
    const T dxs_0 = dx10 + dx00, dxy_0 = dx10 - dx00;
    const T dxs_1 = dx11 + dx01, dxy_1 = dx11 - dx01;
    // z stage
    m.cz[i] = s0_1 - s0_0;
    m.cy[i] = sy_1 + sy_0;
    m.cyz[i] = sy_1 - sy_0;
    m.cx[i] = dxs_1 + dxs_0;
    m.cxz[i] = dxs_1 - dxs_0;
    m.cxy[i] = dxy_1 + dxy_0;
    m.cxyz[i] = dxy_1 - dxy_0;
  }
}

// Modal forces F from modal sums (linear in lam, mu which already carry E_e
// and the V/64 scaling).  Output uses the same Modal layout.
template <typename T>
__host__ __device__ __forceinline__ void modal_stress(const Modal<T>& c, T gx, T gy, T gz, T lam,
                                                      T mu, Modal<T>& F) {
  const T two_mu = mu + mu;
  const T third = T(1) / T(3);
  // mode 1 (full gradient)
  {
    const T ax0 = gx * c.cx[0], ax1 = gx * c.cx[1], ax2 = gx * c.cx[2];
    const T ay0 = gy * c.cy[0], ay1 = gy * c.cy[1], ay2 = gy * c.cy[2];
    const T az0 = gz * c.cz[0], az1 = gz * c.cz[1], az2 = gz * c.cz[2];
    const T ltr = lam * (ax0 + ay1 + az2);
    const T s00 = ltr + two_mu * ax0, s11 = ltr + two_mu * ay1, s22 = ltr + two_mu * az2;
    const T s01 = mu * (ay0 + ax1), s02 = mu * (az0 + ax2), s12 = mu * (az1 + ay2);
    F.cx[0] = gx * s00; F.cx[1] = gx * s01; F.cx[2] = gx * s02;
    F.cy[0] = gy * s01; F.cy[1] = gy * s11; F.cy[2] = gy * s12;
    F.cz[0] = gz * s02; F.cz[1] = gz * s12; F.cz[2] = gz * s22;
  }
  // modes xi (columns y,z from cxy,cxz), eta (x,z from cxy,cyz),
  // zeta (x,y from cxz,cyz)
  const T by0 = gy * c.cxy[0], by1 = gy * c.cxy[1], by2 = gy * c.cxy[2];
  const T bz0 = gz * c.cxz[0], bz1 = gz * c.cxz[1], bz2 = gz * c.cxz[2];
  const T cx0 = gx * c.cxy[0], cx1 = gx * c.cxy[1], cx2 = gx * c.cxy[2];
  const T czz0 = gz * c.cyz[0], czz1 = gz * c.cyz[1], czz2 = gz * c.cyz[2];
  const T dx0 = gx * c.cxz[0], dx1 = gx * c.cxz[1], dx2 = gx * c.cxz[2];
  const T dy0 = gy * c.cyz[0], dy1 = gy * c.cyz[1], dy2 = gy * c.cyz[2];
  // S_xi[i][1], S_xi[i][2]
  const T lxi = lam * (by1 + bz2);
  const T sxi01 = mu * by0, sxi11 = lxi + two_mu * by1, sxi21 = mu * (by2 + bz1);
  const T sxi02 = mu * bz0, sxi22 = lxi + two_mu * bz2;  // sxi12 == sxi21
  // S_eta[i][0], S_eta[i][2]
  const T leta = lam * (cx0 + czz2);
  const T seta00 = leta + two_mu * cx0, seta10 = mu * cx1, seta20 = mu * (cx2 + czz0);
  const T seta12 = mu * czz1, seta22 = leta + two_mu * czz2;  // seta02 == seta20
  // S_zeta[i][0], S_zeta[i][1]
  const T lze = lam * (dx0 + dy1);
  const T sze00 = lze + two_mu * dx0, sze10 = mu * (dx1 + dy0), sze20 = mu * dx2;
  const T sze11 = lze + two_mu * dy1, sze21 = mu * dy2;  // sze01 == sze10
  F.cxy[0] = third * (gy * sxi01 + gx * seta00);
  F.cxy[1] = third * (gy * sxi11 + gx * seta10);
  F.cxy[2] = third * (gy * sxi21 + gx * seta20);
  F.cxz[0] = third * (gz * sxi02 + gx * sze00);
  F.cxz[1] = third * (gz * sxi21 + gx * sze10);
  F.cxz[2] = third * (gz * sxi22 + gx * sze20);
  F.cyz[0] = third * (gz * seta20 + gy * sze10);
  F.cyz[1] = third * (gz * seta12 + gy * sze11);
  F.cyz[2] = third * (gz * seta22 + gy * sze21);
  // mode xyz: diagonal
  const T gx2 = gx * gx, gy2 = gy * gy, gz2 = gz * gz;
  const T ninth = T(1) / T(9);
  const T base = mu * (gx2 + gy2 + gz2);
  const T lm = lam + mu;
  F.cxyz[0] = ninth * (base + lm * gx2) * c.cxyz[0];
  F.cxyz[1] = ninth * (base + lm * gy2) * c.cxyz[1];
  F.cxyz[2] = ninth * (base + lm * gz2) * c.cxyz[2];
}

// Cube elements (hx = hy = hz): every gradient coefficient carries the same
// g = 2/h, so g^2 folds into the Lame constants and the mode weights 1/3 and
// 1/9 into per-mode copies of them; this saves ~50 of the ~250 operations.
// L = lam g^2 E, M = mu g^2 E (lam, mu already scaled by V/64).
template <typename T>
__host__ __device__ __forceinline__ void modal_stress_cube(const Modal<T>& c, T L, T M, T L3,
                                                           T M3, T K9, Modal<T>& F) {
  const T M2 = M + M;
  {
    const T ltr = L * (c.cx[0] + c.cy[1] + c.cz[2]);
    const T s01 = M * (c.cy[0] + c.cx[1]), s02 = M * (c.cz[0] + c.cx[2]);
    const T s12 = M * (c.cz[1] + c.cy[2]);
    F.cx[0] = ltr + M2 * c.cx[0]; F.cx[1] = s01; F.cx[2] = s02;
    F.cy[0] = s01; F.cy[1] = ltr + M2 * c.cy[1]; F.cy[2] = s12;
    F.cz[0] = s02; F.cz[1] = s12; F.cz[2] = ltr + M2 * c.cz[2];
  }
  // b = cxy (as d/dy in mode xi), cxz (d/dz in mode xi); cxy (d/dx in mode
  // eta), cyz (d/dz in mode eta); cxz (d/dx in mode zeta), cyz (d/dy in zeta).
  // A diagonal term collects the shear M/3 and the normal 2M/3 of its two
  // modes: (M/3 + 2M/3) b = M b.
  const T lxi = L3 * (c.cxy[1] + c.cxz[2]);
  const T leta = L3 * (c.cxy[0] + c.cyz[2]);
  const T lze = L3 * (c.cxz[0] + c.cyz[1]);
  const T sxi21 = M3 * (c.cxy[2] + c.cxz[1]);
  const T seta20 = M3 * (c.cxy[2] + c.cyz[0]);
  const T sze10 = M3 * (c.cxz[1] + c.cyz[0]);
  F.cxy[0] = M * c.cxy[0] + leta;
  F.cxy[1] = M * c.cxy[1] + lxi;
  F.cxy[2] = sxi21 + seta20;
  F.cxz[0] = M * c.cxz[0] + lze;
  F.cxz[1] = sxi21 + sze10;
  F.cxz[2] = M * c.cxz[2] + lxi;
  F.cyz[0] = seta20 + sze10;
  F.cyz[1] = M * c.cyz[1] + lze;
  F.cyz[2] = M * c.cyz[2] + leta;
  F.cxyz[0] = K9 * c.cxyz[0];
  F.cxyz[1] = K9 * c.cxyz[1];
  F.cxyz[2] = K9 * c.cxyz[2];
}

// The cube constants of modal_stress_cube for relative modulus E: the unit
// values (ModalConsts cl, cm, cl3, cm3, ck9, host-computed in FP64) times E.
template <typename T>
__host__ __device__ __forceinline__ void cube_consts(T E, const ModalConsts<T>& mc, T& L, T& M,
                                                     T& L3, T& M3, T& K9) {
  L = E * mc.cl;
  M = E * mc.cm;
  L3 = E * mc.cl3;
  M3 = E * mc.cm3;
  K9 = E * mc.ck9;
}

// Transposed butterfly: f[n][i] = sum_m s_m(n) F_m,i (no constant mode).
template <typename T, int NC>
__host__ __device__ __forceinline__ void modal_backward(const ModalN<T, NC>& F, T f[8][NC]) {
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    // z stage
    const T s0_0 = -F.cz[i], s0_1 = F.cz[i];
    const T sy_0 = F.cy[i] - F.cyz[i], sy_1 = F.cy[i] + F.cyz[i];
    const T dxs_0 = F.cx[i] - F.cxz[i], dxs_1 = F.cx[i] + F.cxz[i];
    const T dxy_0 = F.cxy[i] - F.cxyz[i], dxy_1 = F.cxy[i] + F.cxyz[i];
    // y stage: sx[b][c], dx[b][c]
    const T sx00 = s0_0 - sy_0, sx10 = s0_0 + sy_0;
    const T sx01 = s0_1 - sy_1, sx11 = s0_1 + sy_1;
    const T dx00 = dxs_0 - dxy_0, dx10 = dxs_0 + dxy_0;
    const T dx01 = dxs_1 - dxy_1, dx11 = dxs_1 + dxy_1;
    // x stage
    f[0][i] = sx00 - dx00;  // (0,0,0)
    f[1][i] = sx00 + dx00;  // (1,0,0)
    f[2][i] = sx10 + dx10;  // (1,1,0)
    f[3][i] = sx10 - dx10;  // (0,1,0)
    f[4][i] = sx01 - dx01;  // (0,0,1)
    f[5][i] = sx01 + dx01;  // (1,0,1)
    f[6][i] = sx11 + dx11;  // (1,1,1)
    f[7][i] = sx11 - dx11;  // (0,1,1)
  }
}

// f = E * KE * u for one element (E relative to E0; KE carries E0).
template <typename T>
__host__ __device__ __forceinline__ void element_apply(const T u[8][3], T E,
                                                       const ModalConsts<T>& mc, T f[8][3]) {
  Modal<T> c, F;
  modal_forward(u, c);
  modal_stress(c, mc.gx, mc.gy, mc.gz, E * mc.lam, E * mc.mu, F);
  modal_backward(F, f);
}

// f = E * KE * u for a cube element (hx = hy = hz); mc as for element_apply.
template <typename T>
__host__ __device__ __forceinline__ void element_apply_cube(const T u[8][3], T E,
                                                            const ModalConsts<T>& mc,
                                                            T f[8][3]) {
  Modal<T> c, F;
  modal_forward(u, c);
  T L, M, L3, M3, K9;
  cube_consts(E, mc, L, M, L3, M3, K9);
  modal_stress_cube(c, L, M, L3, M3, K9, F);
  modal_backward(F, f);
}

// x and y butterfly stages of one face of the element (corners (0,0), (1,0),
// (1,1), (0,1) of that face).  Adjacent element layers share a face, so an
// operator marching in z computes each face once: the top face of one layer
// is the bottom face of the next.
template <typename T, int NC = 3>
struct Face {
  T s[NC], sy[NC], dx[NC], dxy[NC];
};

template <typename T, int NC>
__host__ __device__ __forceinline__ void face_xy(const T c0[NC], const T c1[NC], const T c2[NC],
                                                 const T c3[NC], Face<T, NC>& f) {
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const T sx0 = c1[i] + c0[i], dx0 = c1[i] - c0[i];
    const T sx1 = c2[i] + c3[i], dx1 = c2[i] - c3[i];
    f.s[i] = sx1 + sx0;
    f.sy[i] = sx1 - sx0;
    f.dx[i] = dx1 + dx0;
    f.dxy[i] = dx1 - dx0;
  }
}

// The z stage of modal_forward from the two faces' x/y stages.
template <typename T, int NC>
__host__ __device__ __forceinline__ void modal_from_faces(const Face<T, NC>& b,
                                                          const Face<T, NC>& t,
                                                          ModalN<T, NC>& m) {
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    m.cz[i] = t.s[i] - b.s[i];
    m.cy[i] = t.sy[i] + b.sy[i];
    m.cyz[i] = t.sy[i] - b.sy[i];
    m.cx[i] = t.dx[i] + b.dx[i];
    m.cxz[i] = t.dx[i] - b.dx[i];
    m.cxy[i] = t.dxy[i] + b.dxy[i];
    m.cxyz[i] = t.dxy[i] - b.dxy[i];
  }
}

// element_apply / element_apply_cube from the faces' butterfly stages.
template <typename T, bool CUBE>
__host__ __device__ __forceinline__ void element_apply_faces(const Face<T>& b, const Face<T>& t,
                                                             T E, const ModalConsts<T>& mc,
                                                             T f[8][3]) {
  Modal<T> c, F;
  modal_from_faces(b, t, c);
  if (CUBE) {
    T L, M, L3, M3, K9;
    cube_consts(E, mc, L, M, L3, M3, K9);
    modal_stress_cube(c, L, M, L3, M3, K9, F);
  } else {
    modal_stress(c, mc.gx, mc.gy, mc.gz, E * mc.lam, E * mc.mu, F);
  }
  modal_backward(F, f);
}

// Modal forces F of one element from its faces' butterfly stages (the
// element_apply_faces work before the transposed butterfly).
template <typename T, bool CUBE>
__host__ __device__ __forceinline__ void element_modal_faces(const Face<T>& b, const Face<T>& t,
                                                             T E, const ModalConsts<T>& mc,
                                                             Modal<T>& F) {
  Modal<T> c;
  modal_from_faces(b, t, c);
  if (CUBE) {
    T L, M, L3, M3, K9;
    cube_consts(E, mc, L, M, L3, M3, K9);
    modal_stress_cube(c, L, M, L3, M3, K9, F);
  } else {
    modal_stress(c, mc.gx, mc.gy, mc.gz, E * mc.lam, E * mc.mu, F);
  }
}

// Modal forces F of one element from its 8 corner values.
template <typename T, bool CUBE>
__host__ __device__ __forceinline__ void element_modal(const T u[8][3], T E,
                                                       const ModalConsts<T>& mc, Modal<T>& F) {
  Modal<T> c;
  modal_forward(u, c);
  if (CUBE) {
    T L, M, L3, M3, K9;
    cube_consts(E, mc, L, M, L3, M3, K9);
    modal_stress_cube(c, L, M, L3, M3, K9, F);
  } else {
    modal_stress(c, mc.gx, mc.gy, mc.gz, E * mc.lam, E * mc.mu, F);
  }
}

// Transposed butterfly of one node plane shared by two element layers.
// modal_backward's z stage splits an element's modal forces into its bottom
// face (s = -F_z, sy = F_y - F_yz, dx = F_x - F_xz, dxy = F_xy - F_xyz) and
// its top face (the same with + signs); the y and x stages are linear, so
// the plane between element layers k-1 and k is finished once from the sum
// of layer k-1's top face (carried in `top`) and layer k's bottom face.  On
// return c holds the plane's corner forces (corners (0,0), (1,0), (1,1),
// (0,1) of the element's face) and `top` this element's top face.
template <typename T, int NC>
__host__ __device__ __forceinline__ void face_backward(const ModalN<T, NC>& F, Face<T, NC>& top,
                                                       T c[4][NC]) {
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const T s = top.s[i] - F.cz[i];
    const T sy = top.sy[i] + (F.cy[i] - F.cyz[i]);
    const T dx = top.dx[i] + (F.cx[i] - F.cxz[i]);
    const T dxy = top.dxy[i] + (F.cxy[i] - F.cxyz[i]);
    top.s[i] = F.cz[i];
    top.sy[i] = F.cy[i] + F.cyz[i];
    top.dx[i] = F.cx[i] + F.cxz[i];
    top.dxy[i] = F.cxy[i] + F.cxyz[i];
    const T sx0 = s - sy, sx1 = s + sy;
    const T dx0 = dx - dxy, dx1 = dx + dxy;
    c[0][i] = sx0 - dx0;
    c[1][i] = sx0 + dx0;
    c[2][i] = sx1 + dx1;
    c[3][i] = sx1 - dx1;
  }
}

// Thermal brick (one temperature per node): the conduction energy of each
// monomial is a single gradient norm, so KE_T is diagonal in the modal basis,
// F_m = k w_m c_m with w = (gx^2, gy^2, gz^2, (gx^2+gy^2)/3, (gx^2+gz^2)/3,
// (gy^2+gz^2)/3, (gx^2+gy^2+gz^2)/9) times k0 V/64 (2x2x2 Gauss is exact for
// these products).  Reproduces hex8_stiffness_matrix_thermal,
// src/hex_element_stiffness.py:339-385.
template <typename T>
__host__ __device__ __forceinline__ void thermal_modal(const ModalN<T, 1>& c, T k,
                                                       const ModalConsts<T>& mc,
                                                       ModalN<T, 1>& F) {
  F.cx[0] = (k * mc.tw[0]) * c.cx[0];
  F.cy[0] = (k * mc.tw[1]) * c.cy[0];
  F.cz[0] = (k * mc.tw[2]) * c.cz[0];
  F.cxy[0] = (k * mc.tw[3]) * c.cxy[0];
  F.cxz[0] = (k * mc.tw[4]) * c.cxz[0];
  F.cyz[0] = (k * mc.tw[5]) * c.cyz[0];
  F.cxyz[0] = (k * mc.tw[6]) * c.cxyz[0];
}

// f = k * KE_T * T_e for one thermal element (k relative to k0).
template <typename T>
__host__ __device__ __forceinline__ void element_apply_thermal(const T u[8][1], T k,
                                                               const ModalConsts<T>& mc,
                                                               T f[8][1]) {
  ModalN<T, 1> c, F;
  modal_forward(u, c);
  thermal_modal(c, k, mc, F);
  modal_backward(F, f);
}

// Modal fluxes of one thermal element from its faces' butterfly stages.
template <typename T>
__host__ __device__ __forceinline__ void thermal_modal_faces(const Face<T, 1>& b,
                                                             const Face<T, 1>& t, T k,
                                                             const ModalConsts<T>& mc,
                                                             ModalN<T, 1>& F) {
  ModalN<T, 1> c;
  modal_from_faces(b, t, c);
  thermal_modal(c, k, mc, F);
}

// element_apply_thermal from the faces' butterfly stages.
template <typename T>
__host__ __device__ __forceinline__ void element_apply_thermal_faces(const Face<T, 1>& b,
                                                                     const Face<T, 1>& t, T k,
                                                                     const ModalConsts<T>& mc,
                                                                     T f[8][1]) {
  ModalN<T, 1> c, F;
  modal_from_faces(b, t, c);
  thermal_modal(c, k, mc, F);
  modal_backward(F, f);
}

template <typename T>
inline ModalConsts<T> make_thermal_consts(double k0, double hx, double hy, double hz) {
  const double gx2 = 4.0 / (hx * hx), gy2 = 4.0 / (hy * hy), gz2 = 4.0 / (hz * hz);
  const double s = k0 * hx * hy * hz / 64.0;
  ModalConsts<T> m{};
  m.gx = T(2.0 / hx);
  m.gy = T(2.0 / hy);
  m.gz = T(2.0 / hz);
  m.tw[0] = T(s * gx2);
  m.tw[1] = T(s * gy2);
  m.tw[2] = T(s * gz2);
  m.tw[3] = T(s * (gx2 + gy2) / 3.0);
  m.tw[4] = T(s * (gx2 + gz2) / 3.0);
  m.tw[5] = T(s * (gy2 + gz2) / 3.0);
  m.tw[6] = T(s * (gx2 + gy2 + gz2) / 9.0);
  return m;
}

template <typename T>
inline ModalConsts<T> make_modal_consts(double E0, double nu, double hx, double hy, double hz) {
  const double lam = E0 * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E0 / (2.0 * (1.0 + nu));
  const double V = hx * hy * hz;
  ModalConsts<T> m{};
  m.gx = T(2.0 / hx);
  m.gy = T(2.0 / hy);
  m.gz = T(2.0 / hz);
  m.lam = T(lam * V / 64.0);
  m.mu = T(mu * V / 64.0);
  // cube elements (hx = hy = hz): g^2 folded into the Lame constants, and the
  // 1/3 and 1/9 mode weights into per-mode copies (modal_stress_cube)
  const double g2 = (2.0 / hx) * (2.0 / hx);
  const double cl = lam * V / 64.0 * g2, cm = mu * V / 64.0 * g2;
  m.cl = T(cl);
  m.cm = T(cm);
  m.cl3 = T(cl / 3.0);
  m.cm3 = T(cm / 3.0);
  m.ck9 = T((cl + 4.0 * cm) / 9.0);
  return m;
}

// Centre-point engineering strain [xx, yy, zz, xy, xz, yz] of an element,
// src/hex_structural_fea.py:163-183: gradients at the centre are the
// mode-1 coefficients c_x, c_y, c_z times g/8.
template <typename T>
__host__ __device__ __forceinline__ void centre_strain(const T u[8][3], T gx, T gy, T gz,
                                                       T eps[6]) {
  Modal<T> c;
  modal_forward(u, c);
  const T k = T(0.125);
  const T ux = k * gx * c.cx[0], uy = k * gy * c.cy[0], uz = k * gz * c.cz[0];
  const T vx = k * gx * c.cx[1], vy = k * gy * c.cy[1], vz = k * gz * c.cz[1];
  const T wx = k * gx * c.cx[2], wy = k * gy * c.cy[2], wz = k * gz * c.cz[2];
  eps[0] = ux;
  eps[1] = vy;
  eps[2] = wz;
  eps[3] = uy + vx;
  eps[4] = uz + wx;
  eps[5] = vz + wy;
}

}  // namespace pyto
