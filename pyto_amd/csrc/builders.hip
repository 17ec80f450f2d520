// Problem builders of the voxel grid on the device: what HexMesher.grid_mesh,
// getNodesOnBoundingBoxPlane, the cantilever load recipe and
// DeflationSolver.create_deflation_groups compute with per-node Python loops
// or numpy passes in the reference (src/hex_mesher.py:72-169,732-760,
// 1080-1095; src/hex_structural_examples.py:353-395; src/deflation.py:86-229).
//
// All outputs are bit-identical to the reference arithmetic:
//  * node coordinates are (double)i * h, as numpy forms ix * h;
//  * a face load is -w * (total / sum w) with the weights 4/2/1 of interior,
//    edge and corner nodes; sum w = 4 a b exactly for an a x b face;
//  * group coordinate sums run sequentially in node order, each coordinate
//    rounded before it is added (floating-point contraction is off in that
//    kernel), the order in which numpy's add.at accumulates them.
#include <algorithm>

#include "common.hpp"

namespace pyto {

__global__ __launch_bounds__(256) void k_grid_arrays(int nx, int ny, int nz, double hx, double hy,
                                                     double hz, double* __restrict__ xyz,
                                                     int32_t* __restrict__ elem,
                                                     int64_t* __restrict__ edof) {
  const int64_t sx = nx + 1, sy = ny + 1;
  const int64_t nn = sx * sy * (nz + 1), ne = (int64_t)nx * ny * nz;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (xyz)
    for (int64_t n = t0; n < nn; n += stride) {
      const int64_t ix = n % sx, iy = (n / sx) % sy, iz = n / (sx * sy);
      xyz[3 * n] = __dmul_rn((double)ix, hx);
      xyz[3 * n + 1] = __dmul_rn((double)iy, hy);
      xyz[3 * n + 2] = __dmul_rn((double)iz, hz);
    }
  if (elem || edof)
    for (int64_t e = t0; e < ne; e += stride) {
      const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / ((int64_t)nx * ny);
      const int64_t n0 = ez * sx * sy + ey * sx + ex;
      const int64_t c[8] = {n0, n0 + 1, n0 + 1 + sx, n0 + sx, n0 + sx * sy, n0 + 1 + sx * sy,
                            n0 + 1 + sx + sx * sy, n0 + sx + sx * sy};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (elem) elem[8 * e + j] = (int32_t)c[j];
        if (edof) {
          edof[24 * e + 3 * j] = 3 * c[j];
          edof[24 * e + 3 * j + 1] = 3 * c[j] + 1;
          edof[24 * e + 3 * j + 2] = 3 * c[j] + 2;
        }
      }
    }
}

// Nodes of the plane index[axis] == at: a face of (s_u x s_v) nodes, u, v the
// other two axes in increasing order.
struct Plane {
  int axis, at, su, sv;
  int64_t sx, sy;
  __device__ int64_t node(int u, int v) const {
    int64_t i[3];
    i[axis] = at;
    i[axis == 0 ? 1 : 0] = u;
    i[axis == 2 ? 1 : 2] = v;
    return i[0] + sx * (i[1] + sy * i[2]);
  }
};

static Plane make_plane(int nx, int ny, int nz, int axis, int at_max) {
  const int n[3] = {nx, ny, nz};
  Plane p;
  p.axis = axis;
  p.at = at_max ? n[axis] : 0;
  p.su = n[axis == 0 ? 1 : 0] + 1;
  p.sv = n[axis == 2 ? 1 : 2] + 1;
  p.sx = nx + 1;
  p.sy = ny + 1;
  return p;
}

__global__ __launch_bounds__(256) void k_plane_mask(Plane p, uint8_t bits, uint8_t* __restrict__ mask) {
  const int64_t cnt = (int64_t)p.su * p.sv;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < cnt;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = p.node((int)(k % p.su), (int)(k / p.su));
    mask[n] |= bits;
  }
}

__global__ __launch_bounds__(256) void k_face_load(Plane p, int comp, int dpn, double scale,
                                                   double* __restrict__ f) {
  const int64_t cnt = (int64_t)p.su * p.sv;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < cnt;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int u = (int)(k % p.su), v = (int)(k / p.su);
    const int ext = (u == 0 || u == p.su - 1) + (v == 0 || v == p.sv - 1);
    const double w = ext == 2 ? 1.0 : (ext == 1 ? 2.0 : 4.0);
    f[dpn * p.node(u, v) + comp] = -__dmul_rn(w, scale);
  }
}

// node_group[n] = box_to_group[tx[ix] + nbx (ty[iy] + nby tz[iz])] (-1 for a
// box index outside [0, nbox))
__global__ __launch_bounds__(256) void k_node_groups(int nx, int ny, int nz,
                                                     const int32_t* __restrict__ tx,
                                                     const int32_t* __restrict__ ty,
                                                     const int32_t* __restrict__ tz, int nbx,
                                                     int nby, int nbox,
                                                     const int32_t* __restrict__ box_to_group,
                                                     int32_t* __restrict__ node_group) {
  const int64_t sx = nx + 1, sy = ny + 1, nn = sx * sy * (nz + 1);
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(n % sx), iy = (int)((n / sx) % sy), iz = (int)(n / (sx * sy));
    const int b = tx[ix] + nbx * (ty[iy] + nby * tz[iz]);
    node_group[n] = (b >= 0 && b < nbox) ? box_to_group[b] : -1;
  }
}

// sums[3 g + c] = sum over the nodes of group g in node order of coordinate c,
// visiting the node-index box bbox[6 g ..] = (x0, x1, y0, y1, z0, z1), half
// open and clamped to the grid; with node_group == null every node of the box
// belongs to g.  One thread per (group, coordinate), a sequential chain like
// add.at's.
__global__ __launch_bounds__(64) void k_group_sums(int nx, int ny, int nz, int ngroups, double hx,
                                                   double hy, double hz,
                                                   const int32_t* __restrict__ node_group,
                                                   const int32_t* __restrict__ bbox,
                                                   double* __restrict__ sums) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * ngroups) return;
  const int g = t / 3, c = t % 3;
  const int64_t sx = nx + 1, sy = ny + 1;
  const int x0 = max(bbox[6 * g], 0), x1 = min(bbox[6 * g + 1], nx + 1);
  const int y0 = max(bbox[6 * g + 2], 0), y1 = min(bbox[6 * g + 3], ny + 1);
  const int z0 = max(bbox[6 * g + 4], 0), z1 = min(bbox[6 * g + 5], nz + 1);
  const double h = c == 0 ? hx : (c == 1 ? hy : hz);
  double s = 0.0;
  for (int iz = z0; iz < z1; ++iz)
    for (int iy = y0; iy < y1; ++iy) {
      const int64_t row = sx * (iy + sy * iz);
      for (int ix = x0; ix < x1; ++ix) {
        if (node_group && node_group[row + ix] != g) continue;
        const int i = c == 0 ? ix : (c == 1 ? iy : iz);
        const double v = (double)i * h;   // the node coordinate, rounded
        s = s + v;
      }
    }
  sums[t] = s;
}

static int grid_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_grid_arrays(int nx, int ny, int nz, double hx, double hy, double hz, double* node_xyz,
                     int32_t* elem_nodes, int64_t* edof, void* stream) {
  PYTO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, PYTO_EINVAL, "grid needs >= 1 element per axis");
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  PYTO_REQUIRE(nn < (int64_t)INT32_MAX, PYTO_EINVAL, "node ids must fit int32");
  const int64_t n = std::max<int64_t>(nn, (int64_t)nx * ny * nz);
  hipLaunchKernelGGL(k_grid_arrays, dim3(grid_blocks(n)), dim3(256), 0, as_stream(stream), nx, ny,
                     nz, hx, hy, hz, node_xyz, elem_nodes, edof);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_plane_mask(int nx, int ny, int nz, int axis, int at_max, int comp_bits, uint8_t* mask,
                    void* stream) {
  PYTO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && mask, PYTO_EINVAL, "bad grid or null mask");
  PYTO_REQUIRE(axis >= 0 && axis <= 2 && comp_bits > 0 && comp_bits < 8, PYTO_EINVAL,
               "axis must be 0..2 and comp_bits 1..7");
  const Plane p = make_plane(nx, ny, nz, axis, at_max);
  hipLaunchKernelGGL(k_plane_mask, dim3(grid_blocks((int64_t)p.su * p.sv)), dim3(256), 0,
                     as_stream(stream), p, (uint8_t)comp_bits, mask);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_face_load(int nx, int ny, int nz, int axis, int at_max, int comp, int dofs_per_node,
                   double total, double* f, void* stream) {
  PYTO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && f, PYTO_EINVAL, "bad grid or null f");
  PYTO_REQUIRE(axis >= 0 && axis <= 2 && dofs_per_node >= 1 && comp >= 0 && comp < dofs_per_node,
               PYTO_EINVAL, "axis 0..2, 0 <= comp < dofs_per_node");
  const Plane p = make_plane(nx, ny, nz, axis, at_max);
  // sum of the weights over an (a+1) x (b+1) face: 4 a b
  const double wsum = 4.0 * (double)(p.su - 1) * (double)(p.sv - 1);
  hipLaunchKernelGGL(k_face_load, dim3(grid_blocks((int64_t)p.su * p.sv)), dim3(256), 0,
                     as_stream(stream), p, comp, dofs_per_node, total / wsum, f);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_node_groups(int nx, int ny, int nz, const int32_t* tx, const int32_t* ty,
                     const int32_t* tz, int nbx, int nby, int nbz, const int32_t* box_to_group,
                     int32_t* node_group, void* stream) {
  PYTO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && tx && ty && tz && box_to_group && node_group,
               PYTO_EINVAL, "bad grid or null argument");
  PYTO_REQUIRE(nbx >= 1 && nby >= 1 && nbz >= 1, PYTO_EINVAL, "box counts must be positive");
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(k_node_groups, dim3(grid_blocks(nn)), dim3(256), 0, as_stream(stream), nx,
                     ny, nz, tx, ty, tz, nbx, nby, nbx * nby * nbz, box_to_group, node_group);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_group_sums(int nx, int ny, int nz, double hx, double hy, double hz,
                    const int32_t* node_group, int ngroups, const int32_t* bbox, double* sums,
                    void* stream) {
  PYTO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && ngroups >= 1 && bbox && sums, PYTO_EINVAL,
               "bad grid, group count or null argument");
  const int threads = 3 * ngroups;
  hipLaunchKernelGGL(k_group_sums, dim3((threads + 63) / 64), dim3(64), 0, as_stream(stream), nx,
                     ny, nz, ngroups, hx, hy, hz, node_group, bbox, sums);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

}  // extern "C"
