// Thermo-elastic coupling on the voxel grid: the element coupling matrix H
// (24 x 8, rows as the caller lays them out) between nodal temperature
// rises and structural nodal forces, f_e = s_e H (T_e - T_ref).
//
//  * force: f_n = sum over the node's elements of s_e (H (T_e - T_ref)) at
//    the node's three rows (src/hex_thermal_fea.py:198-237), gathered per
//    node (no atomics), s_e = E alpha E(x_e) with the structural law;
//  * adjoint right-hand side: r_n = sum_e c E(x_e) (H^T d_e) at the node's
//    column (src/topopt_obj_cons_sensitivities.py:352-366, c = -2 E alpha);
//  * element bilinear d_e^T H (T_e - T_ref) (the direct term of
//    src/topopt_obj_cons_sensitivities.py:455-458).
// Rows of H are indexed as 3 j + c for node j, component c, which is how the
// reference indexes the matrix it builds; the layout of the values is the
// caller's (see the Python mirror).
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace pyto {

struct HMat {
  double v[24 * 8];   // row-major, row 3 j + c
};

// Local index (0..7) of the element corner at offsets (a, b, c).
__device__ __forceinline__ int corner_index(int a, int b, int c) {
  return (c ? 4 : 0) + (b ? (a ? 2 : 3) : (a ? 1 : 0));
}

__device__ __forceinline__ int64_t elem_node(const GridDims& g, int ex, int ey, int ez, int j) {
  return (ex + corner_a(j)) + (int64_t)g.sx * ((ey + corner_b(j)) + (int64_t)g.sy * (ez + corner_c(j)));
}

__device__ __forceinline__ double simp_law(double x, double penal, double eps) {
  return eps + pow(x, penal) * (1.0 - eps);
}

__global__ __launch_bounds__(256) void k_te_force(const double* __restrict__ T,
                                                  const double* __restrict__ x, GridDims g,
                                                  double penal, double eps, double e_alpha,
                                                  double t_ref, HMat H, double* __restrict__ f) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nnodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    node_coords(n, g, ix, iy, iz);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 2; ++c)
      for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
          const int ex = ix - a, ey = iy - b, ez = iz - c;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          const int64_t e = ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez);
          const double s = e_alpha * simp_law(x[e], penal, eps);
          const int j = corner_index(a, b, c);
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const double dt = s * (T[elem_node(g, ex, ey, ez, k)] - t_ref);
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[q] += H.v[(3 * j + q) * 8 + k] * dt;
          }
        }
    f[3 * n] = acc[0];
    f[3 * n + 1] = acc[1];
    f[3 * n + 2] = acc[2];
  }
}

__global__ __launch_bounds__(256) void k_te_adjoint_rhs(const double* __restrict__ d,
                                                        const double* __restrict__ x, GridDims g,
                                                        double penal, double eps, double coef,
                                                        HMat H, double* __restrict__ r) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nnodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    node_coords(n, g, ix, iy, iz);
    double acc = 0.0;
    for (int c = 0; c < 2; ++c)
      for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
          const int ex = ix - a, ey = iy - b, ez = iz - c;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
          const int64_t e = ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez);
          const int k = corner_index(a, b, c);   // the node's column of H
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int64_t m = elem_node(g, ex, ey, ez, j);
#pragma unroll
            for (int q = 0; q < 3; ++q) s += H.v[(3 * j + q) * 8 + k] * d[3 * m + q];
          }
          acc += coef * simp_law(x[e], penal, eps) * s;
        }
    r[n] = acc;
  }
}

__global__ __launch_bounds__(256) void k_te_bilinear(const double* __restrict__ d,
                                                     const double* __restrict__ T, GridDims g,
                                                     double t_ref, HMat H,
                                                     double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.nelems;
       e += (int64_t)gridDim.x * blockDim.x) {
    int ex, ey, ez;
    elem_coords(e, g, ex, ey, ez);
    double dt[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) dt[k] = T[elem_node(g, ex, ey, ez, k)] - t_ref;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t m = elem_node(g, ex, ey, ez, j);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double hr = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) hr += H.v[(3 * j + q) * 8 + k] * dt[k];
        s += d[3 * m + q] * hr;
      }
    }
    out[e] = s;
  }
}

static HMat load_H(const double* H_host) {
  HMat H;
  std::copy(H_host, H_host + 24 * 8, H.v);
  return H;
}

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_thermoelastic_force(pyto_handle* h, const double* T, const double* x, double penal,
                             double eps, double e_alpha, double t_ref, const double* H_host,
                             double* f, void* stream) {
  PYTO_REQUIRE(h && T && x && H_host && f, PYTO_EINVAL, "null argument");
  const int blocks = (int)std::min<int64_t>((h->g.nnodes + 255) / 256, 4096);
  hipLaunchKernelGGL(k_te_force, dim3(blocks), dim3(256), 0, as_stream(stream), T, x, h->g, penal,
                     eps, e_alpha, t_ref, load_H(H_host), f);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_thermoelastic_adjoint_rhs(pyto_handle* h, const double* d, const double* x,
                                   double penal, double eps, double coef, const double* H_host,
                                   double* r, void* stream) {
  PYTO_REQUIRE(h && d && x && H_host && r, PYTO_EINVAL, "null argument");
  const int blocks = (int)std::min<int64_t>((h->g.nnodes + 255) / 256, 4096);
  hipLaunchKernelGGL(k_te_adjoint_rhs, dim3(blocks), dim3(256), 0, as_stream(stream), d, x, h->g,
                     penal, eps, coef, load_H(H_host), r);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_thermoelastic_bilinear(pyto_handle* h, const double* d, const double* T, double t_ref,
                                const double* H_host, double* out, void* stream) {
  PYTO_REQUIRE(h && d && T && H_host && out, PYTO_EINVAL, "null argument");
  const int blocks = (int)std::min<int64_t>((h->g.nelems + 255) / 256, 4096);
  hipLaunchKernelGGL(k_te_bilinear, dim3(blocks), dim3(256), 0, as_stream(stream), d, T, h->g,
                     t_ref, load_H(H_host), out);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

}  // extern "C"
