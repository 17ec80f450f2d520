// Element kernels of the sensitivities and the cone filter.
//
//  * compliance: ce_e = u_e^T KE u_e, C = sum E(x) ce, dC/dx = -E'(x) ce
//    (src/topopt_obj_cons_sensitivities.py:148-179);
//  * bilinear a_e^T KE b_e (the T2 factor, :277-282);
//  * centre-point stress relaxed by x^q, von Mises, sum vm^p
//    (src/hex_structural_fea.py:163-229);
//  * p-norm adjoint right-hand side g and direct term T1
//    (src/topopt_obj_cons_sensitivities.py:229-266, g_e as a 24-vector): an
//    element pass writes the 6-vector w_e = x^q dPN/dvm vm^{p-1} D dvm/dsigma,
//    then a node pass gathers g_n = sum_e B_n^T w_e without atomics;
//  * cone filter H as a fixed stencil with boundary truncation
//    (src/topopt_filters.py:8-39).
#include <algorithm>
#include <cmath>
#include <vector>

#include "element.hpp"

namespace pyto {

template <int NC = 3>
__device__ __forceinline__ void gather_elem(const double* __restrict__ u, int64_t e,
                                            const GridDims& g, double ue[8][NC]) {
  int ix, iy, iz;
  elem_coords(e, g, ix, iy, iz);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t n = (ix + corner_a(j)) + (int64_t)g.sx * ((iy + corner_b(j)) +
                                                            (int64_t)g.sy * (iz + corner_c(j)));
#pragma unroll
    for (int c = 0; c < NC; ++c) ue[j][c] = u[NC * n + c];
  }
}

// f = KE u_e (NC = 3) or KE_T T_e (NC = 1), unit modulus
template <int NC>
__device__ __forceinline__ void apply_unit(const double ue[8][NC], const ModalConsts<double>& mc,
                                           double f[8][NC]) {
  if constexpr (NC == 3)
    element_apply<double>(ue, 1.0, mc, f);
  else
    element_apply_thermal<double>(ue, 1.0, mc, f);
}

// NC = 1: the thermal compliance T^T K_T T with the conductivity law
// (src/topopt_obj_cons_sensitivities.py:160-176, FEA_MODE.THERMAL).
template <int NC>
__global__ __launch_bounds__(256) void k_compliance(const double* __restrict__ u,
                                                    const double* __restrict__ x, GridDims g,
                                                    ModalConsts<double> mc, double penal,
                                                    double eps, double* __restrict__ ce_out,
                                                    double* __restrict__ grad,
                                                    double* __restrict__ C,
                                                    double* __restrict__ partials,
                                                    unsigned* __restrict__ counter) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.nelems;
       e += (int64_t)gridDim.x * blockDim.x) {
    double ue[8][NC], f[8][NC];
    gather_elem<NC>(u, e, g, ue);
    apply_unit<NC>(ue, mc, f);
    double ce = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int c = 0; c < NC; ++c) ce += ue[j][c] * f[j][c];
    const double xe = x[e];
    const double Ee = eps + pow(xe, penal) * (1.0 - eps);
    acc += Ee * ce;
    if (ce_out) ce_out[e] = ce;
    if (grad) grad[e] = -penal * pow(xe, penal - 1.0) * (1.0 - eps) * ce;
  }
  const double bs = block_reduce_sum256(acc, red);
  grid_reduce_finish(bs, partials, counter, gridDim.x, blockIdx.x,
                     [=](double tot) { *C = tot; });
}

template <int NC>
__global__ __launch_bounds__(256) void k_bilinear(const double* __restrict__ a,
                                                  const double* __restrict__ b, GridDims g,
                                                  ModalConsts<double> mc,
                                                  double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.nelems;
       e += (int64_t)gridDim.x * blockDim.x) {
    double ae[8][NC], be[8][NC], f[8][NC];
    gather_elem<NC>(a, e, g, ae);
    gather_elem<NC>(b, e, g, be);
    apply_unit<NC>(be, mc, f);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int c = 0; c < NC; ++c) s += ae[j][c] * f[j][c];
    out[e] = s;
  }
}

// Centre gradient of the temperature in reference coordinates, (1/8) sum_j
// s_k(j) T_j for k = x, y, z: the three mode-1 Hadamard sums over 8
// (src/hex_thermal_fea.py:250-270, gradN @ sol[edof].T); out is 3 x N_el.
__global__ __launch_bounds__(256) void k_thermal_grad(const double* __restrict__ T, GridDims g,
                                                      double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.nelems;
       e += (int64_t)gridDim.x * blockDim.x) {
    double te[8][1];
    gather_elem<1>(T, e, g, te);
    ModalN<double, 1> c;
    modal_forward(te, c);
    out[e] = 0.125 * c.cx[0];
    out[g.nelems + e] = 0.125 * c.cy[0];
    out[2 * g.nelems + e] = 0.125 * c.cz[0];
  }
}

struct DMat {
  double d11, d12, d44;  // isotropic D: diag normal, off normal, shear
};

__device__ __forceinline__ void apply_D(const DMat& D, const double eps[6], double sig[6]) {
  sig[0] = D.d11 * eps[0] + D.d12 * (eps[1] + eps[2]);
  sig[1] = D.d11 * eps[1] + D.d12 * (eps[0] + eps[2]);
  sig[2] = D.d11 * eps[2] + D.d12 * (eps[0] + eps[1]);
  sig[3] = D.d44 * eps[3];
  sig[4] = D.d44 * eps[4];
  sig[5] = D.d44 * eps[5];
}

__device__ __forceinline__ double von_mises(const double s[6]) {
  return sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                     (s[2] - s[0]) * (s[2] - s[0])) +
              3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
}

__global__ __launch_bounds__(256) void k_stress(const double* __restrict__ u,
                                                const double* __restrict__ x, GridDims g,
                                                ModalConsts<double> mc, DMat D, double q,
                                                double pexp, double* __restrict__ strain,
                                                double* __restrict__ stress,
                                                double* __restrict__ vm_out,
                                                double* __restrict__ energy,
                                                double* __restrict__ sum_out,
                                                double* __restrict__ partials,
                                                unsigned* __restrict__ counter) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.nelems;
       e += (int64_t)gridDim.x * blockDim.x) {
    double ue[8][3], eps[6], sig[6];
    gather_elem(u, e, g, ue);
    centre_strain<double>(ue, mc.gx, mc.gy, mc.gz, eps);
    apply_D(D, eps, sig);
    const double relax = pow(x[e], q);
#pragma unroll
    for (int k = 0; k < 6; ++k) sig[k] *= relax;
    const double vm = von_mises(sig);
    if (stress)
#pragma unroll
      for (int k = 0; k < 6; ++k) stress[6 * e + k] = sig[k];
    if (strain)
#pragma unroll
      for (int k = 0; k < 6; ++k) strain[6 * e + k] = eps[k];
    if (energy) {
      double w = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) w += eps[k] * sig[k];
      energy[e] = 0.5 * w;
    }
    vm_out[e] = vm;
    acc += pow(vm, pexp);
  }
  const double bs = block_reduce_sum256(acc, red);
  grid_reduce_finish(bs, partials, counter, gridDim.x, blockIdx.x,
                     [=](double tot) { *sum_out = tot; });
}

// w_e (6) and T1_e.  dvm/dsigma * vm^{p-1} is formed as vm^{p-2} * (...) so
// that a zero stress state gives 0 instead of 0/0.
__global__ __launch_bounds__(256) void k_pnorm_elem(const double* __restrict__ u,
                                                    const double* __restrict__ x, GridDims g,
                                                    ModalConsts<double> mc, DMat D, double q,
                                                    double pexp,
                                                    const double* __restrict__ sum_vm_p,
                                                    double* __restrict__ w_out,
                                                    double* __restrict__ T1) {
  const double dpn = pow(*sum_vm_p, 1.0 / pexp - 1.0);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.nelems;
       e += (int64_t)gridDim.x * blockDim.x) {
    double ue[8][3], eps[6], s0[6];
    gather_elem(u, e, g, ue);
    centre_strain<double>(ue, mc.gx, mc.gy, mc.gz, eps);
    apply_D(D, eps, s0);  // unrelaxed D B u_e
    const double xr = x[e];
    const double relax = pow(xr, q);
    double s[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = relax * s0[k];
    const double vm = von_mises(s);
    // h = vm^{p-1} dvm/dsigma
    const double vp2 = vm > 0.0 ? pow(vm, pexp - 2.0) : 0.0;
    double hv[6];
    hv[0] = vp2 * 0.5 * (2 * s[0] - s[1] - s[2]);
    hv[1] = vp2 * 0.5 * (2 * s[1] - s[0] - s[2]);
    hv[2] = vp2 * 0.5 * (2 * s[2] - s[0] - s[1]);
    hv[3] = vp2 * 3.0 * s[3];
    hv[4] = vp2 * 3.0 * s[4];
    hv[5] = vp2 * 3.0 * s[5];
    const double xc = fmax(xr, 1e-12);  // src/topopt_obj_cons_sensitivities.py:252
    double t1 = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) t1 += hv[k] * s0[k];
    T1[e] = dpn * q * pow(xc, q - 1.0) * t1;
    double Dh[6];
    apply_D(D, hv, Dh);  // D symmetric, D^T h
    const double sc = pow(xc, q) * dpn;
#pragma unroll
    for (int k = 0; k < 6; ++k) w_out[6 * e + k] = sc * Dh[k];
  }
}

// g_n = sum over the node's elements of B_n^T w_e (centre-point B,
// rows xx,yy,zz,xy,xz,yz; src/topopt_obj_cons_sensitivities.py:204-225).
__global__ __launch_bounds__(256) void k_pnorm_gather(const double* __restrict__ w, GridDims g,
                                                      double* __restrict__ gout) {
  const double qx = 0.25 / g.hx, qy = 0.25 / g.hy, qz = 0.25 / g.hz;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nnodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    node_coords(n, g, ix, iy, iz);
    double o0 = 0.0, o1 = 0.0, o2 = 0.0;
#pragma unroll
    for (int cz = 0; cz < 2; ++cz)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cx = 0; cx < 2; ++cx) {
          // element whose local corner (cx,cy,cz) is this node
          const int ex = ix - cx, ey = iy - cy, ez = iz - cz;
          if (ex < 0 || ex >= g.nx || ey < 0 || ey >= g.ny || ez < 0 || ez >= g.nz) continue;
          const double* we = w + 6 * (ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez));
          const double dx = (cx ? qx : -qx), dy = (cy ? qy : -qy), dz = (cz ? qz : -qz);
          o0 += dx * we[0] + dy * we[3] + dz * we[4];
          o1 += dy * we[1] + dx * we[3] + dz * we[5];
          o2 += dz * we[2] + dx * we[4] + dy * we[5];
        }
    gout[3 * n] = o0;
    gout[3 * n + 1] = o1;
    gout[3 * n + 2] = o2;
  }
}

// ---------------------------------------------------------------------------
// Cone filter
// ---------------------------------------------------------------------------
constexpr int kMaxStencil = 343;  // (2*3+1)^3: radius up to 3 elements per axis

struct Stencil {
  int n;
  int reach[3];
  int8_t off[kMaxStencil][3];
  double w[kMaxStencil];
};

// Offsets with distance <= r between element centres (centres are the mean
// of the 8 corners, src/hex_mesher.py:154-157); weight (r - d)/r.
static int build_stencil(const GridDims& g, double rel, Stencil& st) {
  const double r = rel * g.hx;
  const double h[3] = {g.hx, g.hy, g.hz};
  int reach[3];
  for (int k = 0; k < 3; ++k) reach[k] = (int)std::ceil(r / h[k]);
  PYTO_REQUIRE(reach[0] <= 3 && reach[1] <= 3 && reach[2] <= 3, PYTO_EINVAL,
               "filter radius %.3g reaches more than 3 elements per axis", rel);
  st.n = 0;
  for (int k = 0; k < 3; ++k) st.reach[k] = reach[k];
  for (int dz = -reach[2]; dz <= reach[2]; ++dz)
    for (int dy = -reach[1]; dy <= reach[1]; ++dy)
      for (int dx = -reach[0]; dx <= reach[0]; ++dx) {
        const double ddx = dx * g.hx, ddy = dy * g.hy, ddz = dz * g.hz;
        const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (d2 > r * r) continue;
        const double d = std::sqrt(d2);
        if (d > r) continue;
        st.off[st.n][0] = (int8_t)dx;
        st.off[st.n][1] = (int8_t)dy;
        st.off[st.n][2] = (int8_t)dz;
        st.w[st.n] = (r - d) / r;
        st.n++;
      }
  return PYTO_OK;
}

// The stencil lives in a per-handle device buffer (st), read with
// wave-uniform indices.
template <int MODE>
__global__ __launch_bounds__(256) void k_filter(const double* __restrict__ v,
                                                const double* __restrict__ aux, GridDims g,
                                                const Stencil* __restrict__ st,
                                                double* __restrict__ out) {
  const int ns = st->n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.nelems;
       e += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    elem_coords(e, g, ix, iy, iz);
    double hv = 0.0, hs = 0.0;
    for (int k = 0; k < ns; ++k) {
      const int jx = ix + st->off[k][0], jy = iy + st->off[k][1], jz = iz + st->off[k][2];
      if (jx < 0 || jx >= g.nx || jy < 0 || jy >= g.ny || jz < 0 || jz >= g.nz) continue;
      const int64_t j = jx + (int64_t)g.nx * (jy + (int64_t)g.ny * jz);
      const double w = st->w[k];
      hs += w;
      if (MODE == 0 || MODE == 1) hv += w * v[j];
      if (MODE == 2) hv += w * aux[j] * v[j];
      if (MODE == 3) {
        // H (v / Hs): needs Hs of the neighbour; recompute it
        double hsj = 0.0;
        for (int l = 0; l < ns; ++l) {
          const int kx = jx + st->off[l][0], ky = jy + st->off[l][1], kz = jz + st->off[l][2];
          if (kx < 0 || kx >= g.nx || ky < 0 || ky >= g.ny || kz < 0 || kz >= g.nz) continue;
          hsj += st->w[l];
        }
        hv += w * v[j] / hsj;
      }
    }
    double o;
    if (MODE == 0 || MODE == 3) o = hv;
    else if (MODE == 1) o = hv / hs;
    else if (MODE == 2) o = hv / hs / aux[e];
    else o = hs;
    out[e] = o;
  }
}

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_compliance(pyto_handle* h, const double* u, const double* x, double* ce, double* grad,
                    double* C, void* stream) {
  PYTO_REQUIRE(h && u && x && C, PYTO_EINVAL, "null argument");
  const int blocks = (int)std::min<int64_t>((h->g.nelems + 255) / 256, 2048);
  if (h->ncomp == 1)
    hipLaunchKernelGGL(k_compliance<1>, dim3(blocks), dim3(256), 0, as_stream(stream), u, x,
                       h->g, h->mc64, h->penal, h->eps_void, ce, grad, C, h->partials,
                       h->counters + C_ELEM);
  else
    hipLaunchKernelGGL(k_compliance<3>, dim3(blocks), dim3(256), 0, as_stream(stream), u, x,
                       h->g, h->mc64, h->penal, h->eps_void, ce, grad, C, h->partials,
                       h->counters + C_ELEM);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_thermal_gradient(pyto_handle* h, const double* T, double* grad, void* stream) {
  PYTO_REQUIRE(h && T && grad, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->ncomp == 1, PYTO_EINVAL, "the temperature gradient needs a thermal handle");
  const int blocks = (int)std::min<int64_t>((h->g.nelems + 255) / 256, 4096);
  hipLaunchKernelGGL(k_thermal_grad, dim3(blocks), dim3(256), 0, as_stream(stream), T, h->g,
                     grad);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_element_bilinear(pyto_handle* h, const double* a, const double* b, double* be,
                          void* stream) {
  PYTO_REQUIRE(h && a && b && be, PYTO_EINVAL, "null argument");
  const int blocks = (int)std::min<int64_t>((h->g.nelems + 255) / 256, 4096);
  if (h->ncomp == 1)
    hipLaunchKernelGGL(k_bilinear<1>, dim3(blocks), dim3(256), 0, as_stream(stream), a, b, h->g,
                       h->mc64, be);
  else
    hipLaunchKernelGGL(k_bilinear<3>, dim3(blocks), dim3(256), 0, as_stream(stream), a, b, h->g,
                       h->mc64, be);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

static DMat make_D(const Handle* h) {
  const double E = h->E0, nu = h->nu;
  const double c = E / ((1 + nu) * (1 - 2 * nu));
  return DMat{c * (1 - nu), c * nu, c * (1 - 2 * nu) / 2};
}

int pyto_stress(pyto_handle* h, const double* u, const double* x, double q, double pexp,
                double* strain, double* stress, double* vm, double* energy, double* sum_vm_p,
                void* stream) {
  PYTO_REQUIRE(h && u && x && vm && sum_vm_p, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->ncomp == 3, PYTO_EINVAL, "stress needs a structural handle");
  const int blocks = (int)std::min<int64_t>((h->g.nelems + 255) / 256, 2048);
  hipLaunchKernelGGL(k_stress, dim3(blocks), dim3(256), 0, as_stream(stream), u, x, h->g,
                     h->mc64, make_D(h), q, pexp, strain, stress, vm, energy, sum_vm_p, h->partials,
                     h->counters + C_ELEM);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_pnorm_adjoint_rhs(pyto_handle* h, const double* u, const double* x, const double* vm,
                           const double* sum_vm_p, double q, double pexp, double* g,
                           double* T1, void* stream) {
  PYTO_REQUIRE(h && u && x && sum_vm_p && g && T1, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->ncomp == 3, PYTO_EINVAL, "p-norm stress needs a structural handle");
  (void)vm;
  if (!h->elem_buf) PYTO_TRY(device_alloc(h, (void**)&h->elem_buf, (size_t)h->g.nelems * 6 * 8));
  hipStream_t s = as_stream(stream);
  int blocks = (int)std::min<int64_t>((h->g.nelems + 255) / 256, 4096);
  hipLaunchKernelGGL(k_pnorm_elem, dim3(blocks), dim3(256), 0, s, u, x, h->g, h->mc64,
                     make_D(h), q, pexp, sum_vm_p, h->elem_buf, T1);
  blocks = (int)std::min<int64_t>((h->g.nnodes + 255) / 256, 4096);
  hipLaunchKernelGGL(k_pnorm_gather, dim3(blocks), dim3(256), 0, s, h->elem_buf, h->g, g);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_filter_stencil(pyto_handle* h, double rel_radius, int capacity, int* count,
                        int32_t* offsets_host, double* weights_host) {
  PYTO_REQUIRE(h && count, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(rel_radius > 0, PYTO_EINVAL, "filter radius must be positive");
  Stencil st;
  PYTO_TRY(build_stencil(h->g, rel_radius, st));
  *count = st.n;
  if (offsets_host && weights_host) {
    PYTO_REQUIRE(capacity >= st.n, PYTO_EINVAL, "capacity %d < stencil size %d", capacity, st.n);
    for (int k = 0; k < st.n; ++k) {
      for (int c = 0; c < 3; ++c) offsets_host[3 * k + c] = st.off[k][c];
      weights_host[k] = st.w[k];
    }
  }
  return PYTO_OK;
}

int pyto_filter_apply(pyto_handle* h, double rel_radius, int mode, const double* v,
                      const double* aux, double* out, void* stream) {
  PYTO_REQUIRE(h && out, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(mode >= 0 && mode <= 4, PYTO_EINVAL, "filter mode must be 0..4");
  PYTO_REQUIRE(mode == 4 || v, PYTO_EINVAL, "null input");
  PYTO_REQUIRE(mode != 2 || aux, PYTO_EINVAL, "mode 2 needs the density (aux)");
  PYTO_REQUIRE(rel_radius > 0, PYTO_EINVAL, "filter radius must be positive");
  hipStream_t s = as_stream(stream);
  // The handle keeps the stencil of the last radius on the device; a new
  // radius is uploaded in stream order (the call then waits for the copy,
  // so the host staging struct may go out of scope).
  if (!h->stencil) PYTO_TRY(device_alloc(h, &h->stencil, sizeof(Stencil)));
  if (h->stencil_rel != rel_radius) {
    Stencil st;
    PYTO_TRY(build_stencil(h->g, rel_radius, st));
    h->stencil_rel = 0.0;
    PYTO_CHECK_HIP(hipMemcpyAsync(h->stencil, &st, sizeof(Stencil), hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipStreamSynchronize(s));
    h->stencil_rel = rel_radius;
  }
  const Stencil* st = (const Stencil*)h->stencil;
  const int blocks = (int)std::min<int64_t>((h->g.nelems + 255) / 256, 4096);
  switch (mode) {
    case 0: hipLaunchKernelGGL(k_filter<0>, dim3(blocks), dim3(256), 0, s, v, aux, h->g, st, out); break;
    case 1: hipLaunchKernelGGL(k_filter<1>, dim3(blocks), dim3(256), 0, s, v, aux, h->g, st, out); break;
    case 2: hipLaunchKernelGGL(k_filter<2>, dim3(blocks), dim3(256), 0, s, v, aux, h->g, st, out); break;
    case 3: hipLaunchKernelGGL(k_filter<3>, dim3(blocks), dim3(256), 0, s, v, aux, h->g, st, out); break;
    default: hipLaunchKernelGGL(k_filter<4>, dim3(blocks), dim3(256), 0, s, v, aux, h->g, st, out); break;
  }
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

}  // extern "C"
