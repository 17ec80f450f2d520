// Deflation coarse space on the GPU: implicit rigid-body W, restriction and
// prolongation, the Galerkin matrix E = sym(W^T K W) + shift assembled from
// the elements that are not rigid under W, and E^{-1} (rocSOLVER Cholesky and
// inverse) applied as a dense FP64 GEMV inside the iteration.
//
// W (src/deflation.py:386-458): node n of group g contributes rows 3n+c and
// columns 6g..6g+5 with the rigid-body modes about the group centre, column
// map u:{0,4,5} v:{1,3,5} w:{2,3,4}.  Rows of fixed DOFs are dropped
// (W = dsolver.W[free_dofs], src/linear_solvers.py:184).
//
// "Special" elements: an element whose 8 nodes are in one group and carry no
// fixed DOF sees W as a rigid motion, so KE * (W restricted to it) = 0.  Only
// the other elements contribute to W^T K v and to W^T K W; they are listed
// once per set of groups, sorted by that set, cut into chunks of up to 64
// elements (one wave each), and the per-chunk partial sums are gathered per
// group in a fixed order (no atomics in the iteration).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "element.hpp"

namespace pyto {

int density_elements(Handle* h, const double* x, hipStream_t s);  // operator.hip

// rigid-mode weights of node (ix,iy,iz) relative to centre c, per component:
// returns the 6 column coefficients for component comp.
__device__ __forceinline__ void w_row(int comp, double rx, double ry, double rz, double w[6]) {
#pragma unroll
  for (int j = 0; j < 6; ++j) w[j] = 0.0;
  if (comp == 0) { w[0] = 1.0; w[4] = rz; w[5] = -ry; }
  else if (comp == 1) { w[1] = 1.0; w[3] = -rz; w[5] = rx; }
  else { w[2] = 1.0; w[3] = ry; w[4] = -rx; }
}

__global__ void k_coarse_gather(const double* __restrict__ cpart, const int32_t* __restrict__ gptr,
                                const int32_t* __restrict__ gidx, int ngroups,
                                double* __restrict__ t, const int* __restrict__ done);
__global__ void k_gather_scalars_t(const double* __restrict__ cpart,
                                   const int32_t* __restrict__ gptr,
                                   const int32_t* __restrict__ gidx, int ngroups,
                                   double* __restrict__ t, const double* __restrict__ part_pq,
                                   int npq, const double* __restrict__ part_rz, int nrz,
                                   double* __restrict__ scal, int* __restrict__ flags);

// ---------------------------------------------------------------------------
// t = W^T v over all nodes (fixed DOFs excluded): one block per run of <= 256
// nodes of one group (nodes sorted by group on the host), a block reduction
// into per-run partials, then the per-group gather.  Deterministic.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_restrict_full(const T* __restrict__ v,
                                                       const int32_t* __restrict__ gnodes,
                                                       const int2* __restrict__ nchunk,
                                                       const int32_t* __restrict__ nchunk_grp,
                                                       const double* __restrict__ centers,
                                                       const uint8_t* __restrict__ mask,
                                                       GridDims g, double* __restrict__ npart) {
  __shared__ double red[4][6];
  const int2 ch = nchunk[blockIdx.x];
  const int gid = nchunk_grp[blockIdx.x];
  double c[6] = {0, 0, 0, 0, 0, 0};
  if ((int)threadIdx.x < ch.y) {
    const int64_t n = gnodes[ch.x + threadIdx.x];
    int ix, iy, iz;
    node_coords(n, g, ix, iy, iz);
    const double rx = ix * g.hx - centers[3 * gid];
    const double ry = iy * g.hy - centers[3 * gid + 1];
    const double rz = iz * g.hz - centers[3 * gid + 2];
    const uint8_t m = mask[n];
    const double vx = (m & 1) ? 0.0 : (double)v[3 * n];
    const double vy = (m & 2) ? 0.0 : (double)v[3 * n + 1];
    const double vz = (m & 4) ? 0.0 : (double)v[3 * n + 2];
    c[0] = vx; c[1] = vy; c[2] = vz;
    c[3] = -rz * vy + ry * vz;
    c[4] = rz * vx - rx * vz;
    c[5] = -ry * vx + rx * vy;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  {
    const double sj = wave_reduce6(c);
    if ((lane & 7) == 0 && lane < 48) red[w][lane >> 3] = sj;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int j = threadIdx.x;
    npart[(int64_t)blockIdx.x * 48 + j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
  }
}

// out = W mu (fixed DOFs 0), or p = z + beta p - W mu (UPDATE).
template <typename T, bool UPDATE>
__global__ __launch_bounds__(256) void k_prolong(T* __restrict__ out, const T* __restrict__ z,
                                                 const double* __restrict__ mu,
                                                 const int32_t* __restrict__ grp,
                                                 const double* __restrict__ centers,
                                                 const uint8_t* __restrict__ mask, GridDims g,
                                                 const double* __restrict__ beta_ptr,
                                                 bool first, const int* __restrict__ done,
                                                 double* __restrict__ zero_vec, int zero_n) {
  if (done && *done) return;
  if (zero_vec && blockIdx.x == 0)
    for (int i = threadIdx.x; i < zero_n; i += blockDim.x) zero_vec[i] = 0.0;
  const double beta = (UPDATE && !first) ? *beta_ptr : 0.0;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nnodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    node_coords(n, g, ix, iy, iz);
    double wm[3] = {0, 0, 0};
    if (grp) {
      const int gid = grp[n];
      const double rx = ix * g.hx - centers[3 * gid];
      const double ry = iy * g.hy - centers[3 * gid + 1];
      const double rz = iz * g.hz - centers[3 * gid + 2];
      const double* m6 = mu + 6 * gid;
      wm[0] = m6[0] + rz * m6[4] - ry * m6[5];
      wm[1] = m6[1] - rz * m6[3] + rx * m6[5];
      wm[2] = m6[2] + ry * m6[3] - rx * m6[4];
    }
    const uint8_t m = mask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T val;
      if (UPDATE) {
        const double pold = first ? 0.0 : (double)out[3 * n + c];
        val = (T)((double)z[3 * n + c] + beta * pold - wm[c]);
      } else {
        val = (T)wm[c];
      }
      out[3 * n + c] = ((m >> c) & 1) ? T(0) : val;
    }
  }
}

// Per-chunk partials of W^T K v over the special elements (one wave per
// chunk; k_coarse_gather sums them into t), where v = z + beta p (the new
// search direction before its coarse correction; HAS_P, with beta = rz_new /
// rz formed from the update kernel's block partials) or v = z, z = M^-1 r as
// the solver stores it.  The element arithmetic is in the solver's precision
// T (the cube form for cube elements) and the per-group sums in FP64; in FP32
// that halves the element registers, so more waves fit per SIMD.  Group
// centres are wave-uniform per slot.
// A workgroup serves one block of up to kSpecialBlockChunks chunks: it first
// stages v at the block's sorted corner nodes in LDS (an element's corners
// would otherwise be 24 scattered loads, each wave instruction touching up to
// 64 cache lines on interfaces normal to x), then each wave applies the
// element operator of one chunk reading its corners from LDS.
// The block tables are padded (pwave, pslot, pnodes, pcen, pE), so every
// load but the staged values depends on the block index alone and they are
// all in flight at once; the stop flag is tested after they are issued, and
// before the value loads.
// WPE: minimum waves per SIMD the registers must allow (5 for FP32, which
// fits 96 VGPRs without spilling: the 4-wave-per-SIMD grid of about a
// thousand blocks would otherwise run a short second round).
template <typename T, bool HAS_P, bool CUBE, int WPE = sizeof(T) == 4 ? 5 : 4>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void
k_restrict_special(
    const T* __restrict__ z, const T* __restrict__ p,
    const double* __restrict__ part_rz, int nrz, const double* __restrict__ rz_old,
    const int2* __restrict__ pwave, const SpecialSlot* __restrict__ pslot,
    const int32_t* __restrict__ pnodes, const double* __restrict__ pcen,
    const T* __restrict__ pE, GridDims g, ModalConsts<T> mc, double* __restrict__ cpart,
    const int* __restrict__ done) {
  constexpr int kStage = (3 * kSpecialBlockNodes + 255) / 256;   // staged DOFs per thread
  __shared__ T vs[3 * kSpecialBlockNodes];
  __shared__ double csh[kSpecialBlockChunks][8 * 3];   // group centres per chunk and slot
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int stopped = done ? *done : 0;
  const int2 pw = pwave[b * kSpecialBlockChunks + w];   // (chunk or -1, ng)
  const int64_t slot = ((int64_t)b * kSpecialBlockChunks + w) * 64 + lane;
  const SpecialSlot se = pslot[slot];
  const T Ev = pE[slot];
  // the block's corner nodes (padded to kSpecialBlockNodes by repeating the
  // last one): consecutive lanes read consecutive DOFs wherever the nodes are
  // consecutive.  (32-bit DOF indices: the handle limits 3 * nodes below 2^32)
  uint32_t dof[kStage];
#pragma unroll
  for (int t = 0; t < kStage; ++t) {
    const int i = (int)threadIdx.x + 256 * t;
    const int q = i / 3;
    dof[t] = 3u * (uint32_t)pnodes[(int64_t)b * kSpecialBlockNodes + q] + (uint32_t)(i - 3 * q);
  }
  if (threadIdx.x < kSpecialBlockChunks * 24)
    (&csh[0][0])[threadIdx.x] = pcen[(int64_t)b * kSpecialBlockChunks * 24 + threadIdx.x];
  if (stopped) return;   // the same flag for the whole block
  T za[kStage], pa[kStage];
#pragma unroll
  for (int t = 0; t < kStage; ++t) {
    za[t] = z[dof[t]];
    pa[t] = HAS_P ? p[dof[t]] : T(0);
  }
  // beta = rz_new / rz from the update kernel's partials, summed here in the
  // same order as by the scalar logic (no scalar kernel needed before this);
  // formed after the staging loads are issued, so their latencies overlap
  __shared__ double redb[5];
  const double beta = HAS_P ? block_sum_partials(part_rz, nrz, redb) / *rz_old : 0.0;
#pragma unroll
  for (int t = 0; t < kStage; ++t)
    vs[threadIdx.x + 256 * t] = (T)((double)za[t] + (HAS_P ? beta * (double)pa[t] : 0.0));
  __syncthreads();
  if (pw.x < 0) return;
  const bool active = se.e >= 0;
  T f[8][3];
  uint64_t info = 0;
  int ix = 0, iy = 0, iz = 0;
  if (active) {
    const uint32_t lw[4] = {se.li.x, se.li.y, se.li.z, se.li.w};
    info = se.info;
    elem_coords(se.e, g, ix, iy, iz);
    T u[8][3];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int l = (int)((lw[j >> 1] >> (16 * (j & 1))) & 0xffffu);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        u[j][c] = ((info >> (24 + 3 * j + c)) & 1) ? T(0) : vs[3 * l + c];
    }
    if (CUBE)
      element_apply_cube<T>(u, Ev, mc, f);
    else
      element_apply<T>(u, Ev, mc, f);
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if ((info >> (24 + 3 * j + c)) & 1) f[j][c] = T(0);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) f[j][0] = f[j][1] = f[j][2] = T(0);
  }
  // corner positions relative to the group centre and the moments r x f in
  // the solver precision T, the sums in FP64
  const double* cen = csh[w];
  for (int sl = 0; sl < pw.y; ++sl) {
    const double cx = cen[3 * sl], cy = cen[3 * sl + 1], cz = cen[3 * sl + 2];
    const T x0 = (T)(ix * g.hx - cx), y0 = (T)(iy * g.hy - cy), z0 = (T)(iz * g.hz - cz);
    const T hx = (T)g.hx, hy = (T)g.hy, hz = (T)g.hz;
    double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (!active || (int)((info >> (3 * j)) & 7) != sl) continue;
      const T rx = corner_a(j) ? x0 + hx : x0;
      const T ry = corner_b(j) ? y0 + hy : y0;
      const T rz = corner_c(j) ? z0 + hz : z0;
      const T fx = f[j][0], fy = f[j][1], fz = f[j][2];
      acc[0] += (double)fx;
      acc[1] += (double)fy;
      acc[2] += (double)fz;
      acc[3] += (double)(ry * fz - rz * fy);
      acc[4] += (double)(rz * fx - rx * fz);
      acc[5] += (double)(rx * fy - ry * fx);
    }
    const double tot = wave_reduce6(acc);
    if ((lane & 7) == 0 && lane < 48) cpart[(int64_t)pw.x * 48 + 6 * sl + (lane >> 3)] = tot;
  }
}

// E_e of every lane of the padded restriction blocks (0 without an element),
// both precisions; refreshed with every density.
__global__ void k_special_E(const SpecialSlot* __restrict__ pslot, int64_t n,
                            const double* __restrict__ E64, const float* __restrict__ E32,
                            double* __restrict__ pE64, float* __restrict__ pE32) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int e = pslot[i].e;
  pE64[i] = e >= 0 ? E64[e] : 0.0;
  pE32[i] = e >= 0 ? E32[e] : 0.f;
}

// acc = this lane's share of a group's chunk partials: cpart[6 gidx[k] ..]
// for k = k0 + lane, k0 + lane + 64, ... below k1, added in that order.  The
// index and partial loads go out 4 trips at a time (a plain loop waits for
// both loads of every trip in turn).
__device__ __forceinline__ void group_partials_sum(const double* __restrict__ cpart,
                                                   const int32_t* __restrict__ gidx, int k0,
                                                   int k1, int lane, double acc[6]) {
#pragma unroll
  for (int j = 0; j < 6; ++j) acc[j] = 0.0;
  for (int k = k0 + lane; k < k1; k += 4 * 64) {
    int id[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) id[u] = k + 64 * u < k1 ? gidx[k + 64 * u] : -1;
    double c[4][6];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 6; ++j) c[u][j] = id[u] >= 0 ? cpart[(int64_t)id[u] * 6 + j] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (id[u] >= 0)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[j] += c[u][j];
  }
}

// t[6g..6g+5] = sum of the chunk partials of group g: one wave per group,
// lanes stride over the group's partials, then a wave reduction (fixed
// order).  Replaces atomics: tens of thousands of FP64 atomics on a few
// hundred addresses serialise.
__global__ __launch_bounds__(256) void k_coarse_gather(const double* __restrict__ cpart,
                                                       const int32_t* __restrict__ gptr,
                                                       const int32_t* __restrict__ gidx, int ngroups,
                                                       double* __restrict__ t,
                                                       const int* __restrict__ done) {
  if (done && *done) return;
  const int g = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (g >= ngroups) return;
  double acc[6];
  group_partials_sum(cpart, gidx, gptr[g], gptr[g + 1], lane, acc);
  const double s = wave_reduce6(acc);
  if ((lane & 7) == 0 && lane < 48) t[6 * g + (lane >> 3)] = s;
}

static int round_up64(int n) { return (n + 63) / 64 * 64; }

// t = per-group sums of the chunk partials (one wave per group, as
// k_coarse_gather) and, in one extra workgroup running beside them, the
// iteration's scalars (pcg_scalar_step) from the operator's and the update's
// block partials.
__global__ __launch_bounds__(256) void k_gather_scalars(
    const double* __restrict__ cpart, const int32_t* __restrict__ gptr,
    const int32_t* __restrict__ gidx, int ngroups, double* __restrict__ t,
    const double* __restrict__ part_pq, int npq, const double* __restrict__ part_rz, int nrz,
    double* __restrict__ scal, int* __restrict__ flags) {
  if (blockIdx.x == gridDim.x - 1) {
    __shared__ int done_s;
    __shared__ double red[10];
    if (threadIdx.x == 0) done_s = flags[F_DONE];
    __syncthreads();
    if (done_s) return;
    double pkp, rz_new;
    block_sum_partials2(part_pq, npq, part_rz, nrz, red, pkp, rz_new);
    if (threadIdx.x == 0) pcg_scalar_step(scal, flags, pkp, rz_new, scal[S_FLOOR]);
    return;
  }
  // F_DONE may be set by the scalar workgroup meanwhile: a wave that sees it
  // skips (t is then unused)
  if (flags[F_DONE]) return;
  const int g = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (g >= ngroups) return;
  double acc[6];
  group_partials_sum(cpart, gidx, gptr[g], gptr[g + 1], lane, acc);
  const double sj = wave_reduce6(acc);
  if ((lane & 7) == 0 && lane < 48) t[6 * g + (lane >> 3)] = sj;
}

int launch_gather_scalars(Handle* h, double* t, const double* part_pq, int npq,
                          const double* part_rz, int nrz, hipStream_t s) {
  const int gb = (h->ngroups * 64 + 255) / 256;
  if (h->ncomp == 1) {
    hipLaunchKernelGGL(k_gather_scalars_t, dim3(gb + 1), dim3(256), 0, s, h->cpart, h->gptr,
                       h->gidx, h->ngroups, t, part_pq, npq, part_rz, nrz, h->scal, h->flags);
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  }
  hipLaunchKernelGGL(k_gather_scalars, dim3(gb + 1), dim3(256), 0, s, h->cpart, h->gptr, h->gidx,
                     h->ngroups, t, part_pq, npq, part_rz, nrz, h->scal, h->flags);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

// Direction update of the deflated PCG (z-form):
//   x += alpha p_old  (alpha = scal[S_ALPHA_X], published by the CG update;
//                      skipped on the first call or when alpha is 0),
//   p  = z + beta p_old - W mu  (first: z - W mu; fixed DOFs 0),
// with z = M^-1 r as stored by the CG update.  z and p_old are 0 on fixed
// DOFs; the fixed mask drops W mu there.  Vectorised over DOFs: a thread owns
// one 16-byte vector (4 FP32 / 2 FP64 DOFs), which spans at most two nodes;
// W mu is formed for both nodes (all three components) and each DOF selects
// its own, so the loop has no per-DOF branches.  W mu of group q at position
// X is a_q + w_q x X with w = (mu3, mu4, mu5) and a = (mu0, mu1, mu2) - w x c_q,
// staged in LDS (6 values per group) for up to kLdsGroups groups and formed
// from mu and the centre otherwise.
template <typename T>
struct DirVec;
template <>
struct DirVec<float> {
  using V = float4;
  static constexpr int W = 4;
  __device__ static float get(const V& v, int j) { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }
  __device__ static void set(V& v, int j, float a) {
    if (j == 0) v.x = a; else if (j == 1) v.y = a; else if (j == 2) v.z = a; else v.w = a;
  }
};
template <>
struct DirVec<double> {
  using V = double2;
  static constexpr int W = 2;
  __device__ static double get(const V& v, int j) { return j == 0 ? v.x : v.y; }
  __device__ static void set(V& v, int j, double a) {
    if (j == 0) v.x = a; else v.y = a;
  }
};

constexpr int kLdsGroups = 512;
// Largest dense coarse space: 6 x 4096 = 24576 unknowns, whose FP64 matrix
// and inverse take 4.8 GB each.  The reference's own heuristic asks for at
// most 2000 groups and its box rounding can exceed that (26 x 9 x 9 = 2106
// on a 192 x 64 x 64 grid).
constexpr int kMaxCoarseGroups = 4096;

// INPLACE: pin == pout (one restrict-qualified pointer; otherwise the two
// are distinct buffers and both are restrict-qualified).
template <typename T, bool FLUSH, bool INPLACE>
__global__ __launch_bounds__(256) void k_dir_update(
    DirIO<T> io, const double* __restrict__ mu, const int32_t* __restrict__ grp,
    const double* __restrict__ centers, int ngroups, const uint16_t* __restrict__ gm,
    GridDims g, int nc, const double* __restrict__ scal, bool first,
    const int* __restrict__ done) {
  if (done && *done) return;
  // 6 values per group, sized by the launch (ngroups * 48 bytes when staged)
  extern __shared__ double aw[];
  const bool lds = grp && ngroups <= kLdsGroups;
  if (lds) {
    for (int q = threadIdx.x; q < ngroups; q += blockDim.x) {
      const double* m6 = mu + 6 * q;
      const double wx = m6[3], wy = m6[4], wz = m6[5];
      const double cx = centers[3 * q], cy = centers[3 * q + 1], cz = centers[3 * q + 2];
      aw[6 * q + 0] = m6[0] - (wy * cz - wz * cy);
      aw[6 * q + 1] = m6[1] - (wz * cx - wx * cz);
      aw[6 * q + 2] = m6[2] - (wx * cy - wy * cx);
      aw[6 * q + 3] = wx;
      aw[6 * q + 4] = wy;
      aw[6 * q + 5] = wz;
    }
    __syncthreads();
  }
  // the three components of W mu at node (px, py, pz) of group q
  auto wmu3 = [&](int px, int py, int pz, int q, double* w) {
    const double X = px * g.hx, Y = py * g.hy, Z = pz * g.hz;
    double e[6];
    if (lds) {
#pragma unroll
      for (int k = 0; k < 6; ++k) e[k] = aw[6 * q + k];
    } else {
      const double* m6 = mu + 6 * q;
      const double wx = m6[3], wy = m6[4], wz = m6[5];
      const double cx = centers[3 * q], cy = centers[3 * q + 1], cz = centers[3 * q + 2];
      e[0] = m6[0] - (wy * cz - wz * cy);
      e[1] = m6[1] - (wz * cx - wx * cz);
      e[2] = m6[2] - (wx * cy - wy * cx);
      e[3] = wx;
      e[4] = wy;
      e[5] = wz;
    }
    w[0] = e[0] + e[4] * Z - e[5] * Y;
    w[1] = e[1] + e[5] * X - e[3] * Z;
    w[2] = e[2] + e[3] * Y - e[4] * X;
  };
  using DV = DirVec<T>;
  using V = typename DV::V;
  constexpr int W = DV::W;
  const double beta = first ? 0.0 : scal[S_BETA];
  const double alpha = (first || !io.xupd) ? 0.0 : scal[S_ALPHA_X];
  const int64_t n = nc * g.nnodes;   // groups (grp) only with nc = 3
  const int64_t nv = n / W;
  T* __restrict__ x = io.x;
  T* __restrict__ pio = io.pout;   // INPLACE: the one direction buffer
  const T* __restrict__ p = INPLACE ? pio : io.pin;
  const T* __restrict__ z = io.z;
  const V* pv = reinterpret_cast<const V*>(p);
  V* pov = reinterpret_cast<V*>(pio);
  V* xv = reinterpret_cast<V*>(x);
  const V* zv = reinterpret_cast<const V*>(z);

  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  // the vector's first node n0, its first component c0 and (with groups) the
  // node's coordinates: formed once, then advanced by the constant step of
  // the grid-stride loop (stride * W DOFs = kn nodes + kc components, kn at
  // (sxn, syn, szn)) instead of divided out on every trip
  int64_t n0 = (i * W) / 3;
  int c0 = (int)(i * W - 3 * n0);
  const int64_t kn = (stride * W) / 3;
  const int kc = (int)(stride * W - 3 * kn);
  int ix = 0, iy = 0, iz = 0, sxn = 0, syn = 0, szn = 0;
  if (grp && i < nv) {
    node_coords(n0, g, ix, iy, iz);
    node_coords(kn, g, sxn, syn, szn);
  }
  for (; i < nv; i += stride) {
    // W mu of the vector's first node (0..2) and of the next node (3..5);
    // fixed bits likewise (bit c of the first node, 3 + c of the next)
    double w6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    uint32_t fx = 0;
    if (grp) {
      const bool wx = ix + 1 == g.sx;
      const int jx = wx ? 0 : ix + 1;
      const bool wy = wx && iy + 1 == g.sy;
      const int jy = wx ? (wy ? 0 : iy + 1) : iy;
      const int jz = wy ? iz + 1 : iz;
      const bool two = c0 + W > 3 && n0 + 1 < g.nnodes;
      const uint32_t a0 = gm[n0];
      const uint32_t a1 = two ? (uint32_t)gm[n0 + 1] : a0;
      const int g0 = (int)(a0 & 0xfffu);
      const int g1 = (int)(a1 & 0xfffu);
      fx = (a0 >> 12) | (two ? (a1 >> 12) << 3 : 0u);
      wmu3(ix, iy, iz, g0, w6);
      wmu3(jx, jy, jz, g1, w6 + 3);
    }
    V pp = first ? V{} : pv[i];
    const V zz = zv[i];
    if (FLUSH) {
      // x += the deferred alpha_s p_s of the window (FP64 sum)
      V xx = load_nt(xv + i);
      double acc[W];
#pragma unroll
      for (int j = 0; j < W; ++j) acc[j] = (double)DV::get(xx, j);
      const V* rv = reinterpret_cast<const V*>(io.ring);
      const int64_t ldv = io.ld / W;
#pragma unroll
      for (int sl = 0; sl < kDirSlots; ++sl) {
        if (sl < io.nflush) {
          const V r = rv[sl * ldv + i];
          const double a = io.al[sl];
#pragma unroll
          for (int j = 0; j < W; ++j) acc[j] += a * (double)DV::get(r, j);
        }
      }
#pragma unroll
      for (int j = 0; j < W; ++j) DV::set(xx, j, (T)acc[j]);
      store_nt(xv + i, xx);
      // the window's first slot is the new direction's: its loads above stay
      // before the direction's store below (the pointers are restrict-qualified)
      asm volatile("" ::: "memory");
    } else if (alpha != 0.0) {
      V xx = load_nt(xv + i);
#pragma unroll
      for (int j = 0; j < W; ++j)
        DV::set(xx, j, (T)((double)DV::get(xx, j) + alpha * (double)DV::get(pp, j)));
      store_nt(xv + i, xx);
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
      // component c0 + j of the pair of nodes; c0 is 0, 1 or 2
      const double w = c0 == 0 ? w6[j] : (c0 == 1 ? w6[j + 1] : w6[j + 2]);
      const double wm = ((fx >> (c0 + j)) & 1u) ? 0.0 : w;
      DV::set(pp, j, (T)((double)DV::get(zz, j) + beta * (double)DV::get(pp, j) - wm));
    }
    pov[i] = pp;
    // next trip: c0 + kc may carry one node (so may x and then y)
    c0 += kc;
    const int extra = c0 >= 3 ? 1 : 0;
    c0 -= 3 * extra;
    n0 += kn + extra;
    if (grp) {
      ix += sxn + extra;
      iy += syn;
      iz += szn;
      if (ix >= g.sx) {
        ix -= g.sx;
        ++iy;
      }
      if (iy >= g.sy) {
        iy -= g.sy;
        ++iz;
      }
    }
  }
  // tail: DOFs past the last whole vector, one at a time by block 0
  if (blockIdx.x == 0)
    for (int64_t d = nv * W + threadIdx.x; d < n; d += blockDim.x) {
      const int64_t nd = d / 3;
      const int c = (int)(d - 3 * nd);
      double wm = 0.0;
      if (grp && !((gm[nd] >> (12 + c)) & 1u)) {
        int ix, iy, iz;
        node_coords(nd, g, ix, iy, iz);
        double w3[3];
        wmu3(ix, iy, iz, (int)(gm[nd] & 0xfffu), w3);
        wm = c == 0 ? w3[0] : (c == 1 ? w3[1] : w3[2]);
      }
      const T pd = first ? T(0) : p[d];
      if (FLUSH) {
        double acc = (double)x[d];
        for (int sl = 0; sl < io.nflush; ++sl) acc += io.al[sl] * (double)io.ring[sl * io.ld + d];
        x[d] = (T)acc;
      } else if (alpha != 0.0) {
        x[d] = (T)((double)x[d] + alpha * (double)pd);
      }
      pio[d] = (T)((double)z[d] + beta * (double)pd - wm);
    }
}

// The same direction update with one node (its three DOFs) per thread, for
// the structural FP64 solver: W mu and the fixed bits are formed once per node
// instead of for both nodes a 16-byte vector can span, and the per-DOF
// arithmetic is the vector kernel's.  (In FP32 a node is 12 bytes, loaded as
// separate dwords; the vector kernel is faster there.)
template <typename T, bool FLUSH, bool INPLACE>
__global__ __launch_bounds__(256) void k_dir_update_nodes(
    const double* __restrict__ mu, const int32_t* __restrict__ grp,
    const double* __restrict__ centers, int ngroups, const uint16_t* __restrict__ gm,
    GridDims g, const double* __restrict__ scal, bool first, const int* __restrict__ done,
    DirIO<T> io) {
  if (done && *done) return;
  // 6 values per group, sized by the launch (ngroups * 48 bytes when staged)
  extern __shared__ double aw[];
  const bool lds = grp && ngroups <= kLdsGroups;
  if (lds) {
    for (int q = threadIdx.x; q < ngroups; q += blockDim.x) {
      const double* m6 = mu + 6 * q;
      const double wx = m6[3], wy = m6[4], wz = m6[5];
      const double cx = centers[3 * q], cy = centers[3 * q + 1], cz = centers[3 * q + 2];
      aw[6 * q + 0] = m6[0] - (wy * cz - wz * cy);
      aw[6 * q + 1] = m6[1] - (wz * cx - wx * cz);
      aw[6 * q + 2] = m6[2] - (wx * cy - wy * cx);
      aw[6 * q + 3] = wx;
      aw[6 * q + 4] = wy;
      aw[6 * q + 5] = wz;
    }
    __syncthreads();
  }
  const double beta = first ? 0.0 : scal[S_BETA];
  const double alpha = (first || !io.xupd) ? 0.0 : scal[S_ALPHA_X];

  // the node's coordinates, formed once and advanced by the constant step of
  // the grid-stride loop (stride nodes, at (sxn, syn, szn))
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t nd = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int ix = 0, iy = 0, iz = 0, sxn = 0, syn = 0, szn = 0;
  if (grp && nd < g.nnodes) {
    node_coords(nd, g, ix, iy, iz);
    node_coords(stride, g, sxn, syn, szn);
  }
  for (; nd < g.nnodes; nd += stride) {
    double w[3] = {0.0, 0.0, 0.0};
    uint32_t fx = 0;
    if (grp) {
      const uint32_t a = gm[nd];
      fx = a >> 12;
      const int q = (int)(a & 0xfffu);
      const double X = ix * g.hx, Y = iy * g.hy, Z = iz * g.hz;
      double e[6];
      if (lds) {
#pragma unroll
        for (int k = 0; k < 6; ++k) e[k] = aw[6 * q + k];
      } else {
        const double* m6 = mu + 6 * q;
        const double wx = m6[3], wy = m6[4], wz = m6[5];
        const double cx = centers[3 * q], cy = centers[3 * q + 1], cz = centers[3 * q + 2];
        e[0] = m6[0] - (wy * cz - wz * cy);
        e[1] = m6[1] - (wz * cx - wx * cz);
        e[2] = m6[2] - (wx * cy - wy * cx);
        e[3] = wx;
        e[4] = wy;
        e[5] = wz;
      }
      w[0] = e[0] + e[4] * Z - e[5] * Y;
      w[1] = e[1] + e[5] * X - e[3] * Z;
      w[2] = e[2] + e[3] * Y - e[4] * X;
    }
    T* __restrict__ po = io.pout + 3 * nd;
    const T* __restrict__ pn = INPLACE ? po : io.pin + 3 * nd;
    T* __restrict__ xn = io.x + 3 * nd;
    const T* __restrict__ zn = io.z + 3 * nd;
    T pp[3], zz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pp[c] = first ? T(0) : pn[c];
      zz[c] = zn[c];
    }
    if (FLUSH) {
      // x += the deferred alpha_s p_s of the window (FP64 sum)
      double acc[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] = (double)__builtin_nontemporal_load(xn + c);
#pragma unroll
      for (int sl = 0; sl < kDirSlots; ++sl) {
        if (sl < io.nflush) {
          const T* r = io.ring + sl * io.ld + 3 * nd;
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[c] += io.al[sl] * (double)r[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) __builtin_nontemporal_store((T)acc[c], xn + c);
      // the window's first slot is the new direction's (po): its loads above
      // stay before the stores below (the pointers are restrict-qualified)
      asm volatile("" ::: "memory");
    } else if (alpha != 0.0) {
      T xx[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) xx[c] = __builtin_nontemporal_load(xn + c);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        __builtin_nontemporal_store((T)((double)xx[c] + alpha * (double)pp[c]), xn + c);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double wm = ((fx >> c) & 1u) ? 0.0 : w[c];
      po[c] = (T)((double)zz[c] + beta * (double)pp[c] - wm);
    }
    if (grp) {
      ix += sxn;
      iy += syn;
      iz += szn;
      if (ix >= g.sx) {
        ix -= g.sx;
        ++iy;
      }
      if (iy >= g.sy) {
        iy -= g.sy;
        ++iz;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Thermal coarse space: one constant mode per group (W[n][g] = 1 for node n
// of group g, rows of fixed temperatures dropped), so W^T v is a per-group
// sum, W mu takes mu at the node's group, and E = W^T K_T W is G x G.  The
// special elements are those of the structural space (corners in more than
// one group, or a fixed corner): K_T annihilates constants, so an element
// whose corners are free and in one group adds nothing to W^T K_T v.  The
// tables (chunks, padded restriction blocks, group and pair lists) are the
// structural ones; partials keep one value per (chunk, slot) at index
// chunk * 8 + slot (per node run: run * 8).
// ---------------------------------------------------------------------------

// t = W^T v over all nodes: one block per run of <= 256 nodes of one group,
// a block sum per run (k_gather_t then sums the runs of each group).
template <typename T>
__global__ __launch_bounds__(256) void k_restrict_full_t(const T* __restrict__ v,
                                                         const int32_t* __restrict__ gnodes,
                                                         const int2* __restrict__ nchunk,
                                                         const uint8_t* __restrict__ mask,
                                                         double* __restrict__ npart) {
  __shared__ double red[4];
  const int2 ch = nchunk[blockIdx.x];
  double c = 0.0;
  if ((int)threadIdx.x < ch.y) {
    const int64_t n = gnodes[ch.x + threadIdx.x];
    c = (mask[n] & 1) ? 0.0 : (double)v[n];
  }
  const double tot = block_reduce_sum256(c, red);
  if (threadIdx.x == 0) npart[(int64_t)blockIdx.x * 8] = tot;
}

// t[g] = sum of part[idx[k]] over k in [ptr[g], ptr[g+1]): one wave per group,
// lanes striding over the list, then a wave sum (fixed order).
__device__ __forceinline__ double group_sum_t(const double* __restrict__ part,
                                              const int32_t* __restrict__ idx, int k0, int k1,
                                              int lane) {
  double acc = 0.0;
  for (int k = k0 + lane; k < k1; k += 4 * 64) {
    int id[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) id[u] = k + 64 * u < k1 ? idx[k + 64 * u] : -1;
    double c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = id[u] >= 0 ? part[id[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += c[u];
  }
  return wave_reduce_sum(acc);
}

__global__ __launch_bounds__(256) void k_gather_t(const double* __restrict__ part,
                                                  const int32_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ idx, int ngroups,
                                                  double* __restrict__ t,
                                                  const int* __restrict__ done) {
  if (done && *done) return;
  const int g = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (g >= ngroups) return;
  const double s = group_sum_t(part, idx, ptr[g], ptr[g + 1], lane);
  if (lane == 0) t[g] = s;
}

// k_gather_t and, in one extra workgroup, the iteration's scalars (as
// k_gather_scalars).
__global__ __launch_bounds__(256) void k_gather_scalars_t(
    const double* __restrict__ cpart, const int32_t* __restrict__ gptr,
    const int32_t* __restrict__ gidx, int ngroups, double* __restrict__ t,
    const double* __restrict__ part_pq, int npq, const double* __restrict__ part_rz, int nrz,
    double* __restrict__ scal, int* __restrict__ flags) {
  if (blockIdx.x == gridDim.x - 1) {
    __shared__ int done_s;
    __shared__ double red[10];
    if (threadIdx.x == 0) done_s = flags[F_DONE];
    __syncthreads();
    if (done_s) return;
    double pkp, rz_new;
    block_sum_partials2(part_pq, npq, part_rz, nrz, red, pkp, rz_new);
    if (threadIdx.x == 0) pcg_scalar_step(scal, flags, pkp, rz_new, scal[S_FLOOR]);
    return;
  }
  if (flags[F_DONE]) return;
  const int g = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (g >= ngroups) return;
  const double sj = group_sum_t(cpart, gidx, gptr[g], gptr[g + 1], lane);
  if (lane == 0) t[g] = sj;
}

// out = W mu: mu of the node's group, 0 at fixed temperatures.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_t(T* __restrict__ out,
                                                   const double* __restrict__ mu,
                                                   const uint16_t* __restrict__ gm, int64_t nn,
                                                   const int* __restrict__ done) {
  if (done && *done) return;
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t a = gm[n];
    out[n] = ((a >> 12) & 1u) ? T(0) : (T)mu[a & 0xfffu];
  }
}

// Per-(chunk, slot) partials of W^T K_T v over the special elements, v = z or
// z + beta p (beta from the update's partials), in the padded blocks of the
// structural restriction: the block stages v at its sorted corner nodes in
// LDS, each wave applies the thermal element (modal form, precision T) to its
// chunk and sums the nodal fluxes per group slot in FP64.
template <typename T, bool HAS_P>
__global__ __launch_bounds__(256) void k_restrict_special_t(
    const T* __restrict__ z, const T* __restrict__ p, const double* __restrict__ part_rz,
    int nrz, const double* __restrict__ rz_old, const int2* __restrict__ pwave,
    const SpecialSlot* __restrict__ pslot, const int32_t* __restrict__ pnodes,
    const T* __restrict__ pE, ModalConsts<T> mc, double* __restrict__ cpart,
    const int* __restrict__ done) {
  constexpr int kStage = kSpecialBlockNodes / 256;
  __shared__ T vs[kSpecialBlockNodes];
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int stopped = done ? *done : 0;
  const int2 pw = pwave[b * kSpecialBlockChunks + w];   // (chunk or -1, ng)
  const int64_t slot = ((int64_t)b * kSpecialBlockChunks + w) * 64 + lane;
  const SpecialSlot se = pslot[slot];
  const T Ev = pE[slot];
  uint32_t node[kStage];
#pragma unroll
  for (int t = 0; t < kStage; ++t)
    node[t] = (uint32_t)pnodes[(int64_t)b * kSpecialBlockNodes + threadIdx.x + 256 * t];
  if (stopped) return;
  T za[kStage], pa[kStage];
#pragma unroll
  for (int t = 0; t < kStage; ++t) {
    za[t] = z[node[t]];
    pa[t] = HAS_P ? p[node[t]] : T(0);
  }
  __shared__ double redb[5];
  const double beta = HAS_P ? block_sum_partials(part_rz, nrz, redb) / *rz_old : 0.0;
#pragma unroll
  for (int t = 0; t < kStage; ++t)
    vs[threadIdx.x + 256 * t] = (T)((double)za[t] + (HAS_P ? beta * (double)pa[t] : 0.0));
  __syncthreads();
  if (pw.x < 0) return;
  const bool active = se.e >= 0;
  T f[8][1];
  const uint64_t info = active ? se.info : 0;
  if (active) {
    const uint32_t lw[4] = {se.li.x, se.li.y, se.li.z, se.li.w};
    T u[8][1];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int l = (int)((lw[j >> 1] >> (16 * (j & 1))) & 0xffffu);
      u[j][0] = ((info >> (24 + 3 * j)) & 1) ? T(0) : vs[l];
    }
    element_apply_thermal<T>(u, Ev, mc, f);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) f[j][0] = T(0);
  }
  for (int sl = 0; sl < pw.y; ++sl) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (active && (int)((info >> (3 * j)) & 7) == sl && !((info >> (24 + 3 * j)) & 1))
        acc += (double)f[j][0];
    const double tot = wave_reduce_sum(acc);
    if (lane == 0) cpart[(int64_t)pw.x * 8 + sl] = tot;
  }
}

// Galerkin partials of E = W^T K_T W: one wave per chunk; for each slot sa,
// the element operator applied to the indicator of sa's free corners, then
// for every slot sb the sum of the fluxes at sb's free corners, summed over
// the chunk's elements, at part[off[chunk] + sa ng + sb].
__global__ __launch_bounds__(256) void k_galerkin_parts_t(
    const SpecialElem* __restrict__ sinfo, const SpecialChunk* __restrict__ chunks, int nchunks,
    const int32_t* __restrict__ off, const double* __restrict__ Ee, ModalConsts<double> mc,
    double* __restrict__ part) {
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= nchunks) return;
  const SpecialChunk* ch = chunks + wave;
  const int ng = ch->ng;
  const bool active = lane < ch->count;
  uint64_t info = 0;
  double Ev = 0.0;
  if (active) {
    const SpecialElem se = sinfo[ch->start + lane];
    info = se.info;
    Ev = Ee[se.e];
  }
  double* out = part + off[wave];
  for (int sa = 0; sa < ng; ++sa) {
    double u[8][1], f[8][1];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      u[j][0] = (active && (int)((info >> (3 * j)) & 7) == sa && !((info >> (24 + 3 * j)) & 1))
                    ? 1.0 : 0.0;
    element_apply_thermal<double>(u, Ev, mc, f);
    for (int sb = 0; sb < ng; ++sb) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (active && (int)((info >> (3 * j)) & 7) == sb && !((info >> (24 + 3 * j)) & 1))
          acc += f[j][0];
      const double tot = wave_reduce_sum(acc);
      if (lane == 0) out[sa * ng + sb] = tot;
    }
  }
}

// E[gb][ga] = sum of the pair's chunk partials in list order (one thread per
// group pair sharing a chunk; fixed order, so E is the same bits each time).
__global__ __launch_bounds__(256) void k_galerkin_reduce_t(const int2* __restrict__ pairs,
                                                           const int32_t* __restrict__ pptr,
                                                           const int32_t* __restrict__ pidx,
                                                           int npairs,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ E, int ld) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  double acc = 0.0;
  for (int k = pptr[p]; k < pptr[p + 1]; ++k) acc += part[pidx[k]];
  const int2 gp = pairs[p];
  E[(int64_t)gp.y * ld + gp.x] = acc;
}

// Direction update of the thermal deflated PCG: x += alpha p_old (or the
// window's deferred alpha_s p_s, FLUSH), p = z + beta p_old - W mu with W mu
// = mu of the node's group (0 at fixed temperatures), one 16-byte vector of
// nodes per thread, mu staged in LDS (8 bytes per group, sized by the
// launch).
template <typename T, bool FLUSH, bool INPLACE>
__global__ __launch_bounds__(256) void k_dir_update_t(
    DirIO<T> io, const double* __restrict__ mu, int ngroups, const uint16_t* __restrict__ gm,
    int64_t nn, const double* __restrict__ scal, bool first, const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ double amu[];
  for (int q = threadIdx.x; q < ngroups; q += blockDim.x) amu[q] = mu[q];
  __syncthreads();
  using DV = DirVec<T>;
  using V = typename DV::V;
  constexpr int W = DV::W;
  const double beta = first ? 0.0 : scal[S_BETA];
  const double alpha = (first || !io.xupd) ? 0.0 : scal[S_ALPHA_X];
  const int64_t nv = nn / W;
  T* __restrict__ x = io.x;
  T* __restrict__ pio = io.pout;
  const T* __restrict__ p = INPLACE ? pio : io.pin;
  const T* __restrict__ z = io.z;
  const V* pv = reinterpret_cast<const V*>(p);
  V* pov = reinterpret_cast<V*>(pio);
  V* xv = reinterpret_cast<V*>(x);
  const V* zv = reinterpret_cast<const V*>(z);
  auto wmu = [&](int64_t n) -> double {
    const uint32_t a = gm[n];
    return ((a >> 12) & 1u) ? 0.0 : amu[a & 0xfffu];
  };
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv;
       i += (int64_t)gridDim.x * blockDim.x) {
    V pp = first ? V{} : pv[i];
    const V zz = zv[i];
    if (FLUSH) {
      V xx = load_nt(xv + i);
      double acc[W];
#pragma unroll
      for (int j = 0; j < W; ++j) acc[j] = (double)DV::get(xx, j);
      const V* rv = reinterpret_cast<const V*>(io.ring);
      const int64_t ldv = io.ld / W;
#pragma unroll
      for (int sl = 0; sl < kDirSlots; ++sl) {
        if (sl < io.nflush) {
          const V r = rv[sl * ldv + i];
          const double a = io.al[sl];
#pragma unroll
          for (int j = 0; j < W; ++j) acc[j] += a * (double)DV::get(r, j);
        }
      }
#pragma unroll
      for (int j = 0; j < W; ++j) DV::set(xx, j, (T)acc[j]);
      store_nt(xv + i, xx);
      // the window's first slot is the new direction's: its loads above stay
      // before the direction's store below (the pointers are restrict-qualified)
      asm volatile("" ::: "memory");
    } else if (alpha != 0.0) {
      V xx = load_nt(xv + i);
#pragma unroll
      for (int j = 0; j < W; ++j)
        DV::set(xx, j, (T)((double)DV::get(xx, j) + alpha * (double)DV::get(pp, j)));
      store_nt(xv + i, xx);
    }
#pragma unroll
    for (int j = 0; j < W; ++j)
      DV::set(pp, j, (T)((double)DV::get(zz, j) + beta * (double)DV::get(pp, j) - wmu(i * W + j)));
    pov[i] = pp;
  }
  // nodes past the last whole vector, one at a time by block 0
  if (blockIdx.x == 0)
    for (int64_t d = nv * W + threadIdx.x; d < nn; d += blockDim.x) {
      const T pd = first ? T(0) : p[d];
      if (FLUSH) {
        double acc = (double)x[d];
        for (int sl = 0; sl < io.nflush; ++sl) acc += io.al[sl] * (double)io.ring[sl * io.ld + d];
        x[d] = (T)acc;
      } else if (alpha != 0.0) {
        x[d] = (T)((double)x[d] + alpha * (double)pd);
      }
      pio[d] = (T)((double)z[d] + beta * (double)pd - wmu(d));
    }
}

template <typename T>
int launch_dir_update(Handle* h, const DirIO<T>& io, bool first, const int* done, hipStream_t s) {
  // 2048 workgroups (8 per CU) striding over the vectors: each stages the
  // group maps once, so fewer, longer-lived blocks keep those L2 reads small
  // the group table in LDS takes only what the groups need, so that all
  // workgroups fit on the chip at once
  PYTO_REQUIRE(io.nflush >= 0 && io.nflush <= kDirSlots && (io.nflush == 0 || io.ld % 4 == 0),
               PYTO_EINVAL, "bad flush window");
  if (h->ncomp == 1 && h->ngroups > 0) {
    // thermal groups: one value per group in LDS (at most 32 KB)
    const int64_t nn = h->g.nnodes;
    const int blocks = (int)std::min<int64_t>((nn / (16 / (int64_t)sizeof(T)) + 255) / 256, 2048);
    const bool inp = io.pin == io.pout;
    auto kernel = io.nflush > 0 ? (inp ? k_dir_update_t<T, true, true> : k_dir_update_t<T, true, false>)
                                : (inp ? k_dir_update_t<T, false, true>
                                       : k_dir_update_t<T, false, false>);
    hipLaunchKernelGGL(kernel, dim3(std::max(blocks, 1)), dim3(256), (size_t)h->ngroups * 8, s, io,
                       (const double*)h->cmu, h->ngroups, (const uint16_t*)h->gmask, nn,
                       (const double*)h->scal, first, done);
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  }
  const size_t lds = (h->ngroups > 0 && h->ngroups <= kLdsGroups) ? (size_t)h->ngroups * 48 : 0;
  if (h->ncomp == 3 && sizeof(T) == 8) {
    const int blocks = (int)std::min<int64_t>((h->g.nnodes + 255) / 256, 2048);
    const bool inp = io.pin == io.pout;
    auto kernel = io.nflush > 0 ? (inp ? k_dir_update_nodes<T, true, true>
                                       : k_dir_update_nodes<T, true, false>)
                                : (inp ? k_dir_update_nodes<T, false, true>
                                       : k_dir_update_nodes<T, false, false>);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), lds, s, (const double*)h->cmu,
                       h->ngroups > 0 ? h->node_group : (const int32_t*)nullptr, h->centers,
                       h->ngroups, (const uint16_t*)h->gmask, h->g, (const double*)h->scal, first,
                       done, io);
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  }
  const int64_t nvec = h->ncomp * h->g.nnodes / (16 / (int64_t)sizeof(T));
  const int blocks = (int)std::min<int64_t>((nvec + 255) / 256, 2048);
  const bool inp = io.pin == io.pout;
  auto kernel = io.nflush > 0 ? (inp ? k_dir_update<T, true, true> : k_dir_update<T, true, false>)
                              : (inp ? k_dir_update<T, false, true> : k_dir_update<T, false, false>);
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), lds, s, io,
                     (const double*)h->cmu,
                     h->ngroups > 0 ? h->node_group : (const int32_t*)nullptr, h->centers,
                     h->ngroups, (const uint16_t*)h->gmask, h->g, h->ncomp,
                     (const double*)h->scal, first, done);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}
template int launch_dir_update<double>(Handle*, const DirIO<double>&, bool, const int*,
                                       hipStream_t);
template int launch_dir_update<float>(Handle*, const DirIO<float>&, bool, const int*,
                                      hipStream_t);

// Galerkin matrix E = W^T K W over the special elements, without atomics.
// k_galerkin_parts: one wave per chunk and coarse column ja (blockIdx.y) of
// each of the chunk's groups: the element operator applied to column ja of
// group ga (W restricted to the corners in ga, fixed DOFs 0), then for every
// group gb of the chunk the 6 products with gb's columns, summed over the
// chunk's elements.  The 6x6 block of the pair (ga, gb) lands at
// part[off[chunk] + (sa ng + sb) 36 + 6 ja + jb].  The node positions
// relative to the group centres and the slot and fixed bits come from the
// packed special-element records.
template <bool CUBE>
__global__ __launch_bounds__(256) void k_galerkin_parts(
    const SpecialElem* __restrict__ sinfo, const SpecialChunk* __restrict__ chunks, int nchunks,
    const int32_t* __restrict__ off, const double* __restrict__ Ee,
    const double* __restrict__ centers, GridDims g, ModalConsts<double> mc,
    double* __restrict__ part) {
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ja = blockIdx.y;
  if (wave >= nchunks) return;
  const SpecialChunk* ch = chunks + wave;
  const int ng = ch->ng;
  const bool active = lane < ch->count;
  uint64_t info = 0;
  double Ev = 0.0;
  int ix = 0, iy = 0, iz = 0;
  if (active) {
    const SpecialElem se = sinfo[ch->start + lane];
    info = se.info;
    Ev = Ee[se.e];
    elem_coords(se.e, g, ix, iy, iz);
  }
  double* out = part + off[wave];
  for (int sa = 0; sa < ng; ++sa) {
    const int ga = ch->groups[sa];
    const double ax = centers[3 * ga], ay = centers[3 * ga + 1], az = centers[3 * ga + 2];
    double u[8][3], f[8][3];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const bool in = active && (int)((info >> (3 * j)) & 7) == sa;
      double w[6];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        w_row(c, (ix + corner_a(j)) * g.hx - ax, (iy + corner_b(j)) * g.hy - ay,
              (iz + corner_c(j)) * g.hz - az, w);
        u[j][c] = (in && !((info >> (24 + 3 * j + c)) & 1)) ? w[ja] : 0.0;
      }
    }
    if (CUBE)
      element_apply_cube<double>(u, Ev, mc, f);
    else
      element_apply<double>(u, Ev, mc, f);
    for (int sb = 0; sb < ng; ++sb) {
      const int gb = ch->groups[sb];
      const double bx = centers[3 * gb], by = centers[3 * gb + 1], bz = centers[3 * gb + 2];
      double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (!active || (int)((info >> (3 * j)) & 7) != sb) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if ((info >> (24 + 3 * j + c)) & 1) continue;
          double w[6];
          w_row(c, (ix + corner_a(j)) * g.hx - bx, (iy + corner_b(j)) * g.hy - by,
                (iz + corner_c(j)) * g.hz - bz, w);
#pragma unroll
          for (int jb = 0; jb < 6; ++jb) acc[jb] += w[jb] * f[j][c];
        }
      }
      const double tot = wave_reduce6(acc);
      if ((lane & 7) == 0 && lane < 48) out[(sa * ng + sb) * 36 + 6 * ja + (lane >> 3)] = tot;
    }
  }
}

// E's 6x6 block of each group pair (ga, gb) that shares a chunk: one wave per
// pair, lane 6 ja + jb sums the pair's chunk blocks in list order (fixed
// order, so E is the same bits on every run) into E[6 gb + jb][6 ga + ja].
__global__ __launch_bounds__(256) void k_galerkin_reduce(const int2* __restrict__ pairs,
                                                         const int32_t* __restrict__ pptr,
                                                         const int32_t* __restrict__ pidx,
                                                         int npairs, const double* __restrict__ part,
                                                         double* __restrict__ E, int ld) {
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= npairs || lane >= 36) return;
  double acc = 0.0;
  for (int k = pptr[p]; k < pptr[p + 1]; ++k) acc += part[pidx[k] + lane];
  const int2 gp = pairs[p];
  const int ja = lane / 6, jb = lane - 6 * ja;
  E[(int64_t)(6 * gp.y + jb) * ld + (6 * gp.x + ja)] = acc;
}

// E <- (E + E^T)/2 + (sa + sr |E_ii|) I on the nc x nc block, identity on the
// padding.  A diagonal that is exactly zero (a group whose nodes are all
// fixed, so W has no rows for it) becomes 1: that unknown decouples.
// One workgroup per pair of 64 x 64 tiles (I, J), J <= I.  Both tiles are
// read with whole rows into LDS, each entry pair (i, j), i >= j, is formed
// once as 0.5 (E_ij + E_ji), and both tiles are written back with whole rows.
// ld is a multiple of 64.
__global__ __launch_bounds__(256) void k_symmetrize_tiles(double* __restrict__ E, int nc, int ld,
                                                          double sa, double sr) {
  const int I = blockIdx.x, J = blockIdx.y;
  if (J > I) return;
  __shared__ double A[64][65], B[64][65];   // tile (I, J) and tile (J, I)
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int r = e >> 6, c = e & 63;
    A[r][c] = E[(int64_t)(I * 64 + r) * ld + J * 64 + c];
    B[r][c] = E[(int64_t)(J * 64 + r) * ld + I * 64 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int r = e >> 6, c = e & 63;
    if (I == J && c > r) continue;   // a diagonal tile's pairs from its lower half
    const int i = I * 64 + r, j = J * 64 + c;   // i >= j
    double v;
    if (i < nc && j < nc) {
      v = 0.5 * (A[r][c] + B[c][r]);
      if (i == j) v = (v == 0.0) ? 1.0 : v + sa + sr * fabs(v);
    } else {
      v = (i == j) ? 1.0 : 0.0;
    }
    A[r][c] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int r = e >> 6, c = e & 63;
    E[(int64_t)(I * 64 + r) * ld + J * 64 + c] = (I == J && c > r) ? A[c][r] : A[r][c];
    if (I != J) E[(int64_t)(J * 64 + r) * ld + I * 64 + c] = A[c][r];
  }
}

// ---------------------------------------------------------------------------
// Dense FP64 blocks (n multiple of 64, row-major, leading dimension ld).
// ---------------------------------------------------------------------------

// y = b + sign * A (x1 + beta x2) (one wave per row; x2/beta optional).
__global__ __launch_bounds__(256) void k_gemv(const double* __restrict__ A, int ld, int n,
                                              const double* __restrict__ x,
                                              const double* __restrict__ x2,
                                              const double* __restrict__ beta_ptr,
                                              const double* __restrict__ b, double sign,
                                              double* __restrict__ y,
                                              const int* __restrict__ done) {
  if (done && *done) return;
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double beta = (x2 && beta_ptr) ? *beta_ptr : 0.0;
  double s = 0.0;
  const double* a = A + (int64_t)row * ld;
  // the loads of 8 trips go out together (a plain loop waits for each trip's
  // loads in turn); each lane still adds its terms in the order of j
  // (indices past the row are clamped to its last entry, and their terms
  // left out)
  constexpr int U = 8;
  for (int j0 = lane; j0 < n; j0 += 64 * U) {
    double av[U], xv[U], yv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = min(j0 + 64 * u, n - 1);
      av[u] = a[j];
      xv[u] = x[j];
      yv[u] = x2 ? x2[j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double t = av[u] * (x2 ? xv[u] + beta * yv[u] : xv[u]);
      if (j0 + 64 * u < n) s += t;
    }
  }
  s = wave_reduce_sum(s);
  if (lane == 0) y[row] = (b ? b[row] : 0.0) + sign * s;
}

__global__ void k_axpy_small(double* __restrict__ y, const double* __restrict__ x, int n,
                             const int* __restrict__ done) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}


// Symmetric GEMV for large coarse spaces.  E^-1 is symmetric, so its upper
// triangle of 64 x 64 blocks (nb (nb + 1) / 2 blocks, block (I, J), I <= J,
// at index I nb - I (I - 1) / 2 + J - I, row-major inside, entries past the
// order n zero) holds all of it: the GEMV reads half the bytes of the dense
// one.  One workgroup per stored block writes A_IJ x_J to spart[I][J] and,
// off the diagonal, A_IJ^T x_I to spart[J][I]; a second launch sums
// spart[I][J] over J in order for every row: no atomics, fixed order.
constexpr int kSymBlock = 64;
// coarse orders from which the packed form is used.  Measured on the
// 256 x 128 x 128 workload (mixed solve): 432 groups (2592 unknowns) 46.9 ->
// 46.2 ms, 1024 groups 51.7 -> 46.4 ms, 2000 groups 66.4 -> 53.2 ms; below
// that the dense GEMV's single launch wins.
constexpr int kSymMinOrder = 2048;

__device__ __forceinline__ int64_t sym_block_index(int I, int J, int nb) {
  return (int64_t)I * nb - (int64_t)I * (I - 1) / 2 + (J - I);
}

__global__ __launch_bounds__(256) void k_pack_upper(const double* __restrict__ A, int ld, int n,
                                                    int nb, double* __restrict__ P) {
  // one workgroup per stored block; blockIdx.x = I, blockIdx.y = J - I
  const int I = blockIdx.x, J = I + blockIdx.y;
  if (J >= nb) return;
  double* dst = P + sym_block_index(I, J, nb) * (kSymBlock * kSymBlock);
  for (int e = threadIdx.x; e < kSymBlock * kSymBlock; e += blockDim.x) {
    const int r = I * kSymBlock + e / kSymBlock, c = J * kSymBlock + e % kSymBlock;
    dst[e] = (r < n && c < n) ? A[(int64_t)r * ld + c] : 0.0;
  }
}

// Row and column of stored block b (b = I nb - I (I - 1) / 2 + J - I).
__device__ __forceinline__ void sym_block_rc(int64_t b, int nb, int& I, int& J) {
  const double m = 2.0 * nb + 1.0;
  int i = (int)((m - sqrt(m * m - 8.0 * (double)b)) * 0.5);
  i = max(0, min(i, nb - 1));
  while (i > 0 && sym_block_index(i, i, nb) > b) --i;
  while (i + 1 < nb && sym_block_index(i + 1, i + 1, nb) <= b) ++i;
  I = i;
  J = i + (int)(b - sym_block_index(i, i, nb));
}

// spart[I][J][0..63] = A_IJ x_J; spart[J][I][0..63] = A_IJ^T x_I (I < J).
// A fixed grid of workgroups strides over the stored blocks (a launch that
// returns at once after the stop flag then dispatches only that grid).  Each
// block is staged in LDS (rows padded to 65 doubles) with x_I and x_J beside
// it; threads 0..63 form the row sums, threads 64..127 the column sums.
constexpr int kSymvGrid = 1024;
__global__ __launch_bounds__(256) void k_symv_blocks(const double* __restrict__ P, int nb, int n,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ spart,
                                                     const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double a[kSymBlock][kSymBlock + 1];
  __shared__ double xi[kSymBlock], xj[kSymBlock];
  const int64_t nblk = (int64_t)nb * (nb + 1) / 2;
  for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    int I, J;
    sym_block_rc(b, nb, I, J);
    const double* src = P + b * (kSymBlock * kSymBlock);
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = src[threadIdx.x + 256 * k];
    __syncthreads();   // the previous block's LDS reads are done
    if (threadIdx.x < kSymBlock) {
      const int ri = I * kSymBlock + threadIdx.x, rj = J * kSymBlock + threadIdx.x;
      xi[threadIdx.x] = ri < n ? x[ri] : 0.0;
      xj[threadIdx.x] = rj < n ? x[rj] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int e = threadIdx.x + 256 * k;
      a[e / kSymBlock][e % kSymBlock] = v[k];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < kSymBlock) {
      double sr = 0.0;
      for (int c = 0; c < kSymBlock; ++c) sr += a[t][c] * xj[c];
      spart[((int64_t)I * nb + J) * kSymBlock + t] = sr;
    } else if (t < 2 * kSymBlock && I != J) {
      const int c = t - kSymBlock;
      double sc = 0.0;
      for (int r = 0; r < kSymBlock; ++r) sc += a[r][c] * xi[r];
      spart[((int64_t)J * nb + I) * kSymBlock + c] = sc;
    }
  }
}

// y[r] = sum over J of spart[r / 64][J][r % 64] (rows < n): four lanes per
// row, lane k summing the k-th quarter of J in order (loads of 8 columns in
// flight), then ((q0 + q1) + (q2 + q3)).
__global__ __launch_bounds__(256) void k_symv_sum(const double* __restrict__ spart, int nb, int n,
                                                  double* __restrict__ y,
                                                  const int* __restrict__ done) {
  if (done && *done) return;
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = gt >> 2, k = gt & 3;
  const bool row = r < n;
  const int I = (row ? r : 0) / kSymBlock, l = (row ? r : 0) % kSymBlock;
  const double* p = spart + (int64_t)I * nb * kSymBlock + l;
  const int q = (nb + 3) / 4, j0 = k * q, j1 = min(nb, j0 + q);
  double s = 0.0;
  constexpr int U = 8;
  for (int J0 = j0; J0 < j1; J0 += U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = p[(int64_t)min(J0 + u, j1 - 1) * kSymBlock];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (J0 + u < j1) s += v[u];
  }
  // lanes 4m .. 4m + 3 of a wave hold one row's quarters
  const double s1 = __shfl_xor(s, 1, 64);
  const double h = (k & 1) ? s1 + s : s + s1;   // (q0 + q1) on lanes 0, 1; (q2 + q3) on 2, 3
  const double h2 = __shfl_xor(h, 2, 64);
  if (row && k == 0) y[r] = h + h2;
}

// The packed copy of a large coarse inverse (after every factorisation).
// PYTO_SYMV_MIN_ORDER=n moves the threshold (tests use it on small spaces).
static int pack_coarse_inverse(Handle* h, hipStream_t s) {
  const char* env = getenv("PYTO_SYMV_MIN_ORDER");
  const int min_order = env ? std::max(1, atoi(env)) : kSymMinOrder;
  if (h->nc < min_order) {
    h->pk_nb = 0;   // the dense GEMV (a kept packed copy is not refreshed)
    return PYTO_OK;
  }
  const int ld = round_up64(h->nc), nb = ld / kSymBlock;
  if (!h->Epk || h->pk_nb != nb) {
    if (h->Epk) hipFree(h->Epk);
    if (h->spart) hipFree(h->spart);
    h->Epk = h->spart = nullptr;
    h->pk_nb = 0;
    const size_t nblk = (size_t)nb * (nb + 1) / 2;
    PYTO_TRY(device_alloc(h, (void**)&h->Epk, nblk * kSymBlock * kSymBlock * sizeof(double)));
    PYTO_TRY(device_alloc(h, (void**)&h->spart, (size_t)nb * nb * kSymBlock * sizeof(double)));
    h->pk_nb = nb;
  }
  hipLaunchKernelGGL(k_pack_upper, dim3(nb, nb), dim3(256), 0, s, (const double*)h->Einv, ld,
                     h->nc, nb, h->Epk);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

// mu = E^{-1} (t1 + beta t2) with the FP64 inverse (t2/beta optional).  With
// `refine`, one step of refinement on the coarse residual follows:
// r = t - E mu; mu += E^{-1} r (used by the public coarse solve; the PCG
// re-projects every direction, so it does not need it).
int coarse_apply(Handle* h, const double* t1, const double* t2, const double* beta,
                 double* mu, const int* done, hipStream_t s, bool refine) {
  const int n = h->nc, ld = round_up64(h->nc);
  const int blocks = (n * 64 + 255) / 256;
  if (h->pk_nb > 0 && t2 == nullptr && !refine) {
    const int nb = h->pk_nb;
    const int64_t nblk = (int64_t)nb * (nb + 1) / 2;
    hipLaunchKernelGGL(k_symv_blocks, dim3((unsigned)std::min<int64_t>(nblk, kSymvGrid)),
                       dim3(256), 0, s, (const double*)h->Epk, nb, n, t1, h->spart, done);
    hipLaunchKernelGGL(k_symv_sum, dim3((4 * n + 255) / 256), dim3(256), 0, s,
                       (const double*)h->spart, nb, n, mu, done);
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  }
  hipLaunchKernelGGL(k_gemv, dim3(blocks), dim3(256), 0, s, h->Einv, ld, n, t1, t2, beta,
                     (const double*)nullptr, 1.0, mu, done);
  if (refine) {
    PYTO_REQUIRE(t2 == nullptr, PYTO_EINVAL, "refined coarse solve takes one right-hand side");
    hipLaunchKernelGGL(k_gemv, dim3(blocks), dim3(256), 0, s, h->Ec, ld, n, mu,
                       (const double*)nullptr, (const double*)nullptr, t1, -1.0, h->cres, done);
    hipLaunchKernelGGL(k_gemv, dim3(blocks), dim3(256), 0, s, h->Einv, ld, n, h->cres,
                       (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                       1.0, h->cwork, done);
    hipLaunchKernelGGL(k_axpy_small, dim3((n + 255) / 256), dim3(256), 0, s, mu, h->cwork, n,
                       done);
  }
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

template <typename T>
int launch_restrict_full(Handle* h, const T* v, double* t, hipStream_t s) {
  if (h->ncomp == 1) {
    hipLaunchKernelGGL(k_restrict_full_t<T>, dim3(h->nnchunks), dim3(256), 0, s, v, h->gnodes,
                       h->nchunk, h->fixmask, h->npart);
    hipLaunchKernelGGL(k_gather_t, dim3((h->ngroups * 64 + 255) / 256), dim3(256), 0, s,
                       (const double*)h->npart, h->nptr, h->nidx, h->ngroups, t,
                       (const int*)nullptr);
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  }
  hipLaunchKernelGGL(k_restrict_full<T>, dim3(h->nnchunks), dim3(256), 0, s, v, h->gnodes,
                     h->nchunk, h->nchunk_grp, h->centers, h->fixmask, h->g, h->npart);
  hipLaunchKernelGGL(k_coarse_gather, dim3((h->ngroups * 64 + 255) / 256), dim3(256), 0, s,
                     h->npart, h->nptr, h->nidx, h->ngroups, t, (const int*)nullptr);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

template <typename T>
static const T* pe_of(const Handle* h);
template <>
const double* pe_of<double>(const Handle* h) { return h->pE64; }
template <>
const float* pe_of<float>(const Handle* h) { return h->pE32; }
template <typename T>
static ModalConsts<T> mc_of(const Handle* h);
template <>
ModalConsts<double> mc_of<double>(const Handle* h) { return h->mc64; }
template <>
ModalConsts<float> mc_of<float>(const Handle* h) { return h->mc32; }

// Chunk partials of W^T K v over the special elements (no gather): v = z, or
// with p, v = z + beta p with beta from part_rz (nrz partials) and *rz_old.
template <typename T>
int launch_restrict_special_parts(Handle* h, const T* z, const T* p, const double* part_rz,
                                  int nrz, const double* rz_old, const int* done,
                                  hipStream_t s) {
  PYTO_REQUIRE(z && (p == nullptr || (part_rz != nullptr && rz_old != nullptr)),
               PYTO_EINVAL, "special restriction needs z and (with p) the rz partials");
  if (h->nsblocks > 0 && h->ncomp == 1) {
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(h->nsblocks), dim3(256), 0, s, z, p, part_rz, nrz, rz_old,
                         h->pwave, h->pslot, h->pnodes, pe_of<T>(h), mc_of<T>(h), h->cpart, done);
    };
    if (p)
      go(k_restrict_special_t<T, true>);
    else
      go(k_restrict_special_t<T, false>);
  } else if (h->nsblocks > 0) {
    const bool cube = h->g.hx == h->g.hy && h->g.hy == h->g.hz;
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(h->nsblocks), dim3(256), 0, s, z, p, part_rz, nrz, rz_old,
                         h->pwave, h->pslot, h->pnodes, h->pcen, pe_of<T>(h), h->g, mc_of<T>(h),
                         h->cpart, done);
    };
    if (p)
      cube ? go(k_restrict_special<T, true, true>) : go(k_restrict_special<T, true, false>);
    else
      cube ? go(k_restrict_special<T, false, true>) : go(k_restrict_special<T, false, false>);
  }
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int launch_coarse_gather(Handle* h, double* t, const int* done, hipStream_t s) {
  if (h->ncomp == 1) {
    hipLaunchKernelGGL(k_gather_t, dim3((h->ngroups * 64 + 255) / 256), dim3(256), 0, s,
                       (const double*)h->cpart, h->gptr, h->gidx, h->ngroups, t, done);
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  }
  hipLaunchKernelGGL(k_coarse_gather, dim3((h->ngroups * 64 + 255) / 256), dim3(256), 0, s,
                     h->cpart, h->gptr, h->gidx, h->ngroups, t, done);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

// t = W^T K z over the special elements (overwrites t).
template <typename T>
int launch_restrict_special(Handle* h, const T* z, double* t, const int* done, hipStream_t s) {
  PYTO_TRY(launch_restrict_special_parts<T>(h, z, (const T*)nullptr, nullptr, 0, nullptr, done,
                                            s));
  return launch_coarse_gather(h, t, done, s);
}

template <typename T, bool UPDATE>
int launch_prolong(Handle* h, T* out, const T* z, const double* mu, const double* beta,
                   bool first, const int* done, hipStream_t s, double* zero_vec, int zero_n) {
  const int blocks = (int)std::min<int64_t>((h->g.nnodes + 255) / 256, 8192);
  if (h->ncomp == 1) {
    PYTO_REQUIRE(!UPDATE && zero_vec == nullptr && h->ngroups > 0, PYTO_EINVAL,
                 "thermal prolongation is W mu only");
    hipLaunchKernelGGL(k_prolong_t<T>, dim3(blocks), dim3(256), 0, s, out, mu,
                       (const uint16_t*)h->gmask, h->g.nnodes, done);
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  }
  hipLaunchKernelGGL((k_prolong<T, UPDATE>), dim3(blocks), dim3(256), 0, s, out, z, mu,
                     h->ngroups > 0 ? h->node_group : (const int32_t*)nullptr, h->centers,
                     h->fixmask, h->g, beta, first, done, zero_vec, zero_n);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

template int launch_restrict_full<double>(Handle*, const double*, double*, hipStream_t);
template int launch_restrict_full<float>(Handle*, const float*, double*, hipStream_t);
template int launch_restrict_special<double>(Handle*, const double*, double*, const int*,
                                             hipStream_t);
template int launch_restrict_special<float>(Handle*, const float*, double*, const int*,
                                            hipStream_t);
template int launch_restrict_special_parts<double>(Handle*, const double*, const double*,
                                                   const double*, int, const double*,
                                                   const int*, hipStream_t);
template int launch_restrict_special_parts<float>(Handle*, const float*, const float*,
                                                  const double*, int, const double*,
                                                  const int*, hipStream_t);
template int launch_prolong<double, false>(Handle*, double*, const double*, const double*,
                                           const double*, bool, const int*, hipStream_t,
                                           double*, int);
template int launch_prolong<double, true>(Handle*, double*, const double*, const double*,
                                          const double*, bool, const int*, hipStream_t,
                                          double*, int);
template int launch_prolong<float, false>(Handle*, float*, const float*, const double*,
                                          const double*, bool, const int*, hipStream_t,
                                          double*, int);
template int launch_prolong<float, true>(Handle*, float*, const float*, const double*,
                                         const double*, bool, const int*, hipStream_t,
                                         double*, int);

// Full symmetric matrix from the column-major lower triangle that the
// factorisation leaves (the row-major upper triangle): A[i][j] = A[j][i], i > j.
__global__ void k_fill_symmetric(double* __restrict__ A, int n, int ld) {
  const int i = blockIdx.y;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < i; j += gridDim.x * blockDim.x)
    A[(int64_t)i * ld + j] = A[(int64_t)j * ld + i];
}

// Einv = Ec^{-1} by rocSOLVER: Cholesky (dpotrf) then the inverse from the
// factor (dpotri), both in place on a copy of Ec, then the mirror fill.  Ec
// is symmetric, so its row-major storage is read as column-major.
static int factorize_coarse(Handle* h, hipStream_t s) {
  const int n = h->nc, ld = round_up64(h->nc);
  if (!h->blas) {
    rocblas_handle bh = nullptr;
    PYTO_REQUIRE(rocblas_create_handle(&bh) == rocblas_status_success, PYTO_EHIP,
                 "rocblas_create_handle failed");
    h->blas = bh;
  }
  rocblas_handle bh = (rocblas_handle)h->blas;
  PYTO_REQUIRE(rocblas_set_stream(bh, s) == rocblas_status_success, PYTO_EHIP,
               "rocblas_set_stream failed");
  int* dinfo = h->flags + F_COARSE_PIVOT;
  PYTO_CHECK_HIP(hipMemcpyAsync(h->Einv, h->Ec, sizeof(double) * ld * ld, hipMemcpyDeviceToDevice, s));
  PYTO_REQUIRE(rocsolver_dpotrf(bh, rocblas_fill_lower, n, h->Einv, ld, dinfo) ==
                   rocblas_status_success, PYTO_EHIP, "rocsolver_dpotrf failed");
  int info = 0;
  PYTO_CHECK_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  PYTO_REQUIRE(info == 0, PYTO_ENUMERIC,
               "coarse matrix W^T K W is not positive definite (leading minor %d of %d); a "
               "coarse diagonal shift (pyto_set_coarse_shift) regularises it", info, n);
  PYTO_REQUIRE(rocsolver_dpotri(bh, rocblas_fill_lower, n, h->Einv, ld, dinfo) ==
                   rocblas_status_success, PYTO_EHIP, "rocsolver_dpotri failed");
  hipLaunchKernelGGL(k_fill_symmetric, dim3((n + 255) / 256, n), dim3(256), 0, s, h->Einv, n, ld);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

void release_blas(Handle* h) {
  if (h->blas) rocblas_destroy_handle((rocblas_handle)h->blas);
  h->blas = nullptr;
}

// The Galerkin matrix of the current densities into h->Ec (symmetrised and
// shifted, identity on the padding); no factorisation.
static int coarse_assemble(Handle* h, hipStream_t s) {
  const int ld = round_up64(h->nc);
  if (h->nsblocks > 0) {
    const int64_t n = (int64_t)h->nsblocks * kSpecialBlockChunks * 64;
    hipLaunchKernelGGL(k_special_E, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, h->pslot,
                       n, (const double*)h->Ee64, (const float*)h->Ee32, h->pE64, h->pE32);
  }
  PYTO_CHECK_HIP(hipMemsetAsync(h->Ec, 0, sizeof(double) * ld * ld, s));
  if (h->nchunks > 0 && h->ncomp == 1) {
    hipLaunchKernelGGL(k_galerkin_parts_t, dim3((h->nchunks + 3) / 4), dim3(256), 0, s, h->sinfo,
                       h->chunks, h->nchunks, h->gal_off, h->Ee64, h->mc64, h->gal_part);
    hipLaunchKernelGGL(k_galerkin_reduce_t, dim3((h->npairs + 255) / 256), dim3(256), 0, s,
                       h->gal_pairs, h->gal_pptr, h->gal_pidx, h->npairs,
                       (const double*)h->gal_part, h->Ec, ld);
    PYTO_CHECK_HIP(hipGetLastError());
  } else if (h->nchunks > 0) {
    const bool cube = h->g.hx == h->g.hy && h->g.hy == h->g.hz;
    const dim3 grid((h->nchunks + 3) / 4, 6);
    if (cube)
      hipLaunchKernelGGL(k_galerkin_parts<true>, grid, dim3(256), 0, s, h->sinfo, h->chunks,
                         h->nchunks, h->gal_off, h->Ee64, h->centers, h->g, h->mc64,
                         h->gal_part);
    else
      hipLaunchKernelGGL(k_galerkin_parts<false>, grid, dim3(256), 0, s, h->sinfo, h->chunks,
                         h->nchunks, h->gal_off, h->Ee64, h->centers, h->g, h->mc64,
                         h->gal_part);
    hipLaunchKernelGGL(k_galerkin_reduce, dim3((h->npairs * 64 + 255) / 256), dim3(256), 0, s,
                       h->gal_pairs, h->gal_pptr, h->gal_pidx, h->npairs,
                       (const double*)h->gal_part, h->Ec, ld);
    PYTO_CHECK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_symmetrize_tiles, dim3(ld / 64, ld / 64), dim3(256), 0, s, h->Ec, h->nc,
                     ld, h->shift_abs, h->shift_rel);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int coarse_setup(Handle* h, hipStream_t s) {
  h->coarse_ready = false;   // until the new factor is complete
  PYTO_TRY(coarse_assemble(h, s));
  PYTO_CHECK_HIP(hipMemsetAsync(h->flags + F_COARSE_PIVOT, 0, sizeof(int), s));
  PYTO_TRY(factorize_coarse(h, s));
  PYTO_TRY(pack_coarse_inverse(h, s));
  h->coarse_ready = true;
  return PYTO_OK;
}

// Copy an nc x nc symmetric matrix (leading dimension ld) into an n x n
// matrix with identity padding: blockdiag(E, I) has inverse blockdiag(E^-1, I).
__global__ void k_pad_identity(const double* __restrict__ E, int ld, int nc,
                               double* __restrict__ P, int n) {
  const int r = blockIdx.y;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    P[(int64_t)r * n + c] = (r < nc && c < nc) ? E[(int64_t)r * ld + c] : (r == c ? 1.0 : 0.0);
}

// Einv (nc x nc of leading dimension ld, full symmetric) from the lower
// triangle of a column-major n x n inverse: A(i, j), i >= j, at P[i + j n].
__global__ void k_unpad_symmetric(const double* __restrict__ P, int n, double* __restrict__ Einv,
                                  int ld, int nc) {
  const int r = blockIdx.y;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const int i = max(r, c), j = min(r, c);
    Einv[(int64_t)r * ld + c] = P[i + (int64_t)j * n];
  }
}

// Density update of many handles: every handle's element kernels and Galerkin
// matrix, then ONE batched Cholesky factorisation and inverse of all the
// coarse matrices (padded with identity to the largest order), one host
// synchronisation for the pivot check.
int set_density_many(int n, Handle* const* hs, const double* const* x, hipStream_t s) {
  std::vector<Handle*> defl;
  int nmax = 0;
  for (int i = 0; i < n; ++i) {
    Handle* h = hs[i];
    PYTO_TRY(density_elements(h, x[i], s));
    if (h->ngroups > 0) {
      h->coarse_ready = false;
      PYTO_TRY(coarse_assemble(h, s));
      defl.push_back(h);
      nmax = std::max(nmax, h->nc);
    }
  }
  if (defl.empty()) return PYTO_OK;
  const int nb = (int)defl.size();
  Handle* h0 = defl[0];
  if (!h0->blas) {
    rocblas_handle bh = nullptr;
    PYTO_REQUIRE(rocblas_create_handle(&bh) == rocblas_status_success, PYTO_EHIP,
                 "rocblas_create_handle failed");
    h0->blas = bh;
  }
  rocblas_handle bh = (rocblas_handle)h0->blas;
  PYTO_REQUIRE(rocblas_set_stream(bh, s) == rocblas_status_success, PYTO_EHIP,
               "rocblas_set_stream failed");
  // workspace: nb padded matrices, their pointer table and pivot flags
  double* P = nullptr;
  double** ptrs = nullptr;
  int* info = nullptr;
  const size_t msz = (size_t)nmax * nmax;
  auto release = [&]() {
    if (P) hipFree(P);
    if (ptrs) hipFree(ptrs);
    if (info) hipFree(info);
  };
  int rc = PYTO_OK;
  if (hipMalloc(&P, msz * nb * sizeof(double)) != hipSuccess ||
      hipMalloc(&ptrs, nb * sizeof(double*)) != hipSuccess ||
      hipMalloc(&info, nb * sizeof(int)) != hipSuccess) {
    release();
    set_error("device allocation of the batched coarse workspace (%d x %d^2) failed", nb, nmax);
    return PYTO_ENOMEM;
  }
  std::vector<double*> hp(nb);
  for (int i = 0; i < nb; ++i) hp[i] = P + msz * i;
  std::vector<int> hinfo(nb, 0);
  do {
    if (hipMemcpyAsync(ptrs, hp.data(), nb * sizeof(double*), hipMemcpyHostToDevice, s) !=
        hipSuccess) { rc = PYTO_EHIP; set_error("pointer table copy failed"); break; }
    for (int i = 0; i < nb; ++i)
      hipLaunchKernelGGL(k_pad_identity, dim3((nmax + 255) / 256, nmax), dim3(256), 0, s,
                         (const double*)defl[i]->Ec, round_up64(defl[i]->nc), defl[i]->nc, hp[i],
                         nmax);
    if (hipGetLastError() != hipSuccess) { rc = PYTO_EHIP; set_error("pad kernel failed"); break; }
    if (rocsolver_dpotrf_batched(bh, rocblas_fill_lower, nmax, ptrs, nmax, info, nb) !=
        rocblas_status_success) { rc = PYTO_EHIP; set_error("rocsolver_dpotrf_batched failed"); break; }
    if (hipMemcpyAsync(hinfo.data(), info, nb * sizeof(int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { rc = PYTO_EHIP; set_error("pivot read-back failed"); break; }
    for (int i = 0; i < nb && rc == PYTO_OK; ++i)
      if (hinfo[i] != 0) {
        rc = PYTO_ENUMERIC;
        set_error("coarse matrix W^T K W of handle %d of the batch is not positive definite "
                  "(leading minor %d of %d); a coarse diagonal shift "
                  "(pyto_set_coarse_shift) regularises it", i, hinfo[i], defl[i]->nc);
      }
    if (rc != PYTO_OK) break;
    if (rocsolver_dpotri_batched(bh, rocblas_fill_lower, nmax, ptrs, nmax, info, nb) !=
        rocblas_status_success) { rc = PYTO_EHIP; set_error("rocsolver_dpotri_batched failed"); break; }
    for (int i = 0; i < nb; ++i)
      hipLaunchKernelGGL(k_unpad_symmetric, dim3((defl[i]->nc + 255) / 256, defl[i]->nc),
                         dim3(256), 0, s, (const double*)hp[i], nmax, defl[i]->Einv,
                         round_up64(defl[i]->nc), defl[i]->nc);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      rc = PYTO_EHIP;
      set_error("batched coarse inverse failed");
      break;
    }
    for (Handle* h : defl) {
      if ((rc = pack_coarse_inverse(h, s)) != PYTO_OK) break;
      h->coarse_ready = true;
    }
    if (rc != PYTO_OK) break;
  } while (false);
  release();
  return rc;
}

// Free a handle's coarse space and return it to the no-groups state.
static void release_coarse(Handle* h) {
  void* olds[] = {h->node_group, h->gmask, h->centers, h->sinfo, h->chunks, h->cpart,
                  h->gptr, h->gidx,
                  h->gnodes, h->nchunk, h->nchunk_grp, h->npart, h->nptr, h->nidx,
                  h->Ec, h->Einv, h->cwork, h->ct, h->cmu, h->cres, h->Epk, h->spart,
                  h->gal_off, h->gal_part, h->gal_pairs, h->gal_pptr, h->gal_pidx,
                  h->pwave, h->pslot, h->pnodes, h->pcen, h->pE64, h->pE32};
  for (void* p : olds)
    if (p) hipFree(p);
  h->node_group = nullptr; h->gmask = nullptr; h->centers = nullptr; h->chunks = nullptr;
  h->sinfo = nullptr;
  h->cpart = nullptr; h->gptr = nullptr; h->gidx = nullptr;
  h->nsblocks = 0;
  h->gnodes = nullptr; h->nchunk = nullptr; h->nchunk_grp = nullptr; h->npart = nullptr;
  h->nptr = nullptr; h->nidx = nullptr; h->nnchunks = 0;
  h->Ec = h->Einv = h->cwork = h->ct = h->cmu = h->cres = nullptr;
  h->Epk = h->spart = nullptr;
  h->pk_nb = 0;
  h->gal_off = nullptr; h->gal_part = nullptr; h->gal_pairs = nullptr; h->gal_pptr = nullptr;
  h->gal_pidx = nullptr; h->npairs = 0;
  h->pwave = nullptr; h->pslot = nullptr; h->pnodes = nullptr; h->pcen = nullptr;
  h->pE64 = nullptr; h->pE32 = nullptr;
  h->ngroups = 0; h->nc = 0; h->nspecial = 0; h->nchunks = 0; h->coarse_ready = false;
  h->groups_gen++;
}

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_set_groups(pyto_handle* h, const int32_t* node_group, int ngroups,
                    const double* centers, void* stream) {
  PYTO_REQUIRE(h, PYTO_EINVAL, "null handle");
  PYTO_REQUIRE(ngroups >= 0, PYTO_EINVAL, "ngroups must be >= 0");
  // structural handles: six rigid-body modes per group; thermal handles
  // (one temperature per node): one constant per group
  const bool th = h->ncomp == 1;
  const int modes = th ? 1 : 6;
  PYTO_REQUIRE(ngroups <= kMaxCoarseGroups, PYTO_EINVAL,
               "coarse space of %d groups (%d unknowns) exceeds the dense GPU limit of %d "
               "groups; use fewer, larger deflation groups", ngroups, modes * ngroups,
               kMaxCoarseGroups);
  hipStream_t s = as_stream(stream);
  release_coarse(h);
  if (ngroups == 0) return PYTO_OK;
  PYTO_REQUIRE(node_group && centers, PYTO_EINVAL, "null group arrays");
  const int64_t nn = h->g.nnodes;
  std::vector<int32_t> grp(nn);
  std::vector<uint8_t> mask(nn);
  PYTO_CHECK_HIP(hipMemcpyAsync(grp.data(), node_group, nn * 4, hipMemcpyDeviceToHost, s));
  PYTO_CHECK_HIP(hipMemcpyAsync(mask.data(), h->fixmask, nn, hipMemcpyDeviceToHost, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  for (int64_t n = 0; n < nn; ++n)
    PYTO_REQUIRE(grp[n] >= 0 && grp[n] < ngroups, PYTO_EINVAL,
                 "node_group[%lld] = %d out of range [0,%d)", (long long)n, grp[n], ngroups);
  // the direction updates read a node's group and fixed bits as one 16-bit
  // word (groups < kMaxCoarseGroups = 4096 take 12 bits)
  static_assert(kMaxCoarseGroups <= 4096, "group ids must fit 12 bits");
  std::vector<uint16_t> gmask(nn);
  for (int64_t n = 0; n < nn; ++n) gmask[n] = (uint16_t)(grp[n] | ((mask[n] & 7) << 12));
  // special elements and their group sets
  const GridDims& g = h->g;
  std::vector<std::pair<std::array<int, 8>, int32_t>> list;
  for (int iz = 0; iz < g.nz; ++iz)
    for (int iy = 0; iy < g.ny; ++iy)
      for (int ix = 0; ix < g.nx; ++ix) {
        std::array<int, 8> gs;
        bool fixed = false;
        for (int j = 0; j < 8; ++j) {
          const int64_t n = (ix + corner_a(j)) + (int64_t)g.sx * ((iy + corner_b(j)) +
                                                                  (int64_t)g.sy * (iz + corner_c(j)));
          gs[j] = grp[n];
          fixed |= mask[n] != 0;
        }
        std::sort(gs.begin(), gs.end());
        const int nu = (int)(std::unique(gs.begin(), gs.end()) - gs.begin());
        if (nu == 1 && !fixed) continue;
        for (int j = nu; j < 8; ++j) gs[j] = -1;
        list.push_back({gs, (int32_t)(ix + (int64_t)g.nx * (iy + (int64_t)g.ny * iz))});
      }
  std::sort(list.begin(), list.end());
  std::vector<int32_t> special(list.size());
  std::vector<SpecialChunk> chunks;
  for (size_t i = 0; i < list.size();) {
    size_t j = i;
    while (j < list.size() && list[j].first == list[i].first) ++j;
    for (size_t c = i; c < j; c += 64) {
      SpecialChunk ch;
      ch.start = (int)c;
      ch.count = (int)std::min<size_t>(64, j - c);
      ch.ng = 0;
      for (int q = 0; q < 8; ++q) {
        ch.groups[q] = list[i].first[q];
        if (ch.groups[q] >= 0) ch.ng++;
      }
      chunks.push_back(ch);
    }
    i = j;
  }
  for (size_t i = 0; i < list.size(); ++i) special[i] = list[i].second;
  std::vector<SpecialElem> sinfo(list.size());
  for (size_t i = 0; i < list.size(); ++i) {
    const int32_t e = list[i].second;
    const int ix = (int)(e % g.nx), iy = (int)((e / g.nx) % g.ny), iz = (int)(e / ((int64_t)g.nx * g.ny));
    uint64_t info = 0;
    for (int j = 0; j < 8; ++j) {
      const int64_t n = (ix + corner_a(j)) + (int64_t)g.sx * ((iy + corner_b(j)) +
                                                              (int64_t)g.sy * (iz + corner_c(j)));
      int slot = 0;
      while (list[i].first[slot] != grp[n]) ++slot;
      info |= (uint64_t)slot << (3 * j);
      info |= (uint64_t)(mask[n] & 7) << (24 + 3 * j);
    }
    sinfo[i].e = e;
    sinfo[i].pad = 0;
    sinfo[i].info = info;
  }
  // LDS staging blocks: up to kSpecialBlockChunks consecutive chunks whose
  // sorted unique corner nodes fit kSpecialBlockNodes, and per element the
  // block-local index of each corner
  std::vector<SpecialBlock> sblocks;
  std::vector<int32_t> sbnodes;
  std::vector<uint16_t> sidx(list.size() * 8);
  {
    auto corner_node = [&](int32_t e, int j) {
      const int ix = (int)(e % g.nx), iy = (int)((e / g.nx) % g.ny);
      const int iz = (int)(e / ((int64_t)g.nx * g.ny));
      return (int32_t)((ix + corner_a(j)) + (int64_t)g.sx * ((iy + corner_b(j)) +
                                                           (int64_t)g.sy * (iz + corner_c(j))));
    };
    auto chunk_nodes = [&](const SpecialChunk& ch, std::vector<int32_t>& out) {
      for (int i = ch.start; i < ch.start + ch.count; ++i)
        for (int j = 0; j < 8; ++j) out.push_back(corner_node(list[i].second, j));
    };
    std::vector<int32_t> cur, trial;
    size_t c = 0;
    while (c < chunks.size()) {
      SpecialBlock sb;
      sb.chunk0 = (int)c;
      sb.nch = 0;
      cur.clear();
      while (c < chunks.size() && sb.nch < kSpecialBlockChunks) {
        trial = cur;
        chunk_nodes(chunks[c], trial);
        std::sort(trial.begin(), trial.end());
        trial.erase(std::unique(trial.begin(), trial.end()), trial.end());
        if (sb.nch > 0 && (int)trial.size() > kSpecialBlockNodes) break;
        cur.swap(trial);
        ++sb.nch;
        ++c;
      }
      sb.node0 = (int)sbnodes.size();
      sb.nnodes = (int)cur.size();
      for (int q = 0; q < sb.nch; ++q) {
        const SpecialChunk& ch = chunks[sb.chunk0 + q];
        for (int i = ch.start; i < ch.start + ch.count; ++i)
          for (int j = 0; j < 8; ++j)
            sidx[(size_t)i * 8 + j] = (uint16_t)(
                std::lower_bound(cur.begin(), cur.end(), corner_node(list[i].second, j)) -
                cur.begin());
      }
      sbnodes.insert(sbnodes.end(), cur.begin(), cur.end());
      sblocks.push_back(sb);
    }
  }
  // the blocks padded: per (block, wave) its chunk and group count, per lane
  // its element record, per block kSpecialBlockNodes node ids (the last one
  // repeated), per (block, wave) the centres of the chunk's groups
  const size_t nsb = sblocks.size();
  std::vector<int2> pwave(nsb * kSpecialBlockChunks, make_int2(-1, 0));
  std::vector<SpecialSlot> pslot(nsb * kSpecialBlockChunks * 64);
  std::vector<int32_t> pnodes(nsb * kSpecialBlockNodes);
  std::vector<double> pcen(nsb * kSpecialBlockChunks * 24, 0.0);
  {
    std::vector<double> cen_h((size_t)ngroups * 3);
    PYTO_CHECK_HIP(hipMemcpyAsync(cen_h.data(), centers, cen_h.size() * 8, hipMemcpyDeviceToHost, s));
    PYTO_CHECK_HIP(hipStreamSynchronize(s));
    for (SpecialSlot& sl : pslot) {
      sl.e = -1;
      sl.pad = 0;
      sl.info = 0;
      sl.li = make_uint4(0, 0, 0, 0);
    }
    for (size_t bi = 0; bi < nsb; ++bi) {
      const SpecialBlock& sb = sblocks[bi];
      for (int q = 0; q < kSpecialBlockNodes; ++q)
        pnodes[bi * kSpecialBlockNodes + q] = sbnodes[sb.node0 + std::min(q, sb.nnodes - 1)];
      for (int w = 0; w < sb.nch; ++w) {
        const int c = sb.chunk0 + w;
        const SpecialChunk& ch = chunks[c];
        const size_t bw = bi * kSpecialBlockChunks + w;
        pwave[bw] = make_int2(c, ch.ng);
        for (int sl = 0; sl < ch.ng; ++sl)
          for (int k = 0; k < 3; ++k) pcen[bw * 24 + 3 * sl + k] = cen_h[3 * ch.groups[sl] + k];
        for (int l = 0; l < ch.count; ++l) {
          const size_t i = (size_t)ch.start + l;
          SpecialSlot& o = pslot[bw * 64 + l];
          o.e = sinfo[i].e;
          o.info = sinfo[i].info;
          const uint16_t* x = &sidx[i * 8];
          o.li = make_uint4(x[0] | ((uint32_t)x[1] << 16), x[2] | ((uint32_t)x[3] << 16),
                            x[4] | ((uint32_t)x[5] << 16), x[6] | ((uint32_t)x[7] << 16));
        }
      }
    }
  }
  // per group: the (chunk, slot) partials that belong to it
  std::vector<int32_t> gptr(ngroups + 1, 0), gidx;
  {
    std::vector<std::vector<int32_t>> per(ngroups);
    for (size_t c = 0; c < chunks.size(); ++c)
      for (int q = 0; q < chunks[c].ng; ++q) per[chunks[c].groups[q]].push_back((int32_t)(c * 8 + q));
    for (int gi = 0; gi < ngroups; ++gi) {
      gptr[gi + 1] = gptr[gi] + (int32_t)per[gi].size();
      gidx.insert(gidx.end(), per[gi].begin(), per[gi].end());
    }
  }
  // Galerkin partial blocks: per chunk ng^2 blocks of modes^2 at gal_off[chunk];
  // per group pair (ga, gb) sharing a chunk, the list of its blocks in chunk
  // order
  std::vector<int32_t> gal_off(chunks.size());
  std::vector<int2> gal_pairs;
  std::vector<int32_t> gal_pptr(1, 0), gal_pidx;
  int64_t galsz = 0;
  {
    std::vector<std::pair<int64_t, int32_t>> ent;   // (pair key, block offset)
    for (size_t c = 0; c < chunks.size(); ++c) {
      PYTO_REQUIRE(galsz < (int64_t)1 << 31, PYTO_EINVAL, "Galerkin partials exceed 2^31 values");
      gal_off[c] = (int32_t)galsz;
      const int ng = chunks[c].ng;
      for (int sa = 0; sa < ng; ++sa)
        for (int sb = 0; sb < ng; ++sb)
          ent.push_back({(int64_t)chunks[c].groups[sa] * ngroups + chunks[c].groups[sb],
                         (int32_t)(galsz + (sa * ng + sb) * modes * modes)});
      galsz += (int64_t)modes * modes * ng * ng;
    }
    std::stable_sort(ent.begin(), ent.end(),
                     [](const std::pair<int64_t, int32_t>& x,
                        const std::pair<int64_t, int32_t>& y) { return x.first < y.first; });
    for (size_t k = 0; k < ent.size(); ++k) {
      if (k == 0 || ent[k].first != ent[k - 1].first) {
        if (k > 0) gal_pptr.push_back((int32_t)k);
        gal_pairs.push_back(make_int2((int)(ent[k].first / ngroups), (int)(ent[k].first % ngroups)));
      }
      gal_pidx.push_back(ent[k].second);
    }
    if (!ent.empty()) gal_pptr.push_back((int32_t)ent.size());
  }
  // nodes sorted by group, in runs of <= 256 of one group
  std::vector<int32_t> gnodes(nn);
  {
    std::vector<int64_t> cnt(ngroups + 1, 0);
    for (int64_t n = 0; n < nn; ++n) cnt[grp[n] + 1]++;
    for (int gi = 0; gi < ngroups; ++gi) cnt[gi + 1] += cnt[gi];
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t n = 0; n < nn; ++n) gnodes[pos[grp[n]]++] = (int32_t)n;
  }
  std::vector<int2> nchunk;
  std::vector<int32_t> nchunk_grp;
  std::vector<int32_t> nptr(ngroups + 1, 0), nidx;
  for (int64_t i = 0; i < nn;) {
    const int gi = grp[gnodes[i]];
    int64_t j = i;
    while (j < nn && grp[gnodes[j]] == gi) ++j;
    for (int64_t c = i; c < j; c += 256) {
      nchunk.push_back(make_int2((int)c, (int)std::min<int64_t>(256, j - c)));
      nchunk_grp.push_back(gi);
    }
    i = j;
  }
  for (size_t c = 0; c < nchunk.size(); ++c) nptr[nchunk_grp[c] + 1]++;
  for (int gi = 0; gi < ngroups; ++gi) nptr[gi + 1] += nptr[gi];
  nidx.resize(nchunk.size());
  {
    std::vector<int32_t> pos(nptr.begin(), nptr.end() - 1);
    for (size_t c = 0; c < nchunk.size(); ++c) nidx[pos[nchunk_grp[c]]++] = (int32_t)(c * 8);
  }
  // install the new space; a failure part-way leaves the handle without groups
  auto install = [&]() -> int {
    h->ngroups = ngroups;
    h->nc = modes * ngroups;
    h->nnchunks = (int)nchunk.size();
    h->nspecial = (int)special.size();
    h->nchunks = (int)chunks.size();
    h->nsblocks = (int)sblocks.size();
    const int ld = round_up64(h->nc);
    PYTO_TRY(device_alloc(h, (void**)&h->node_group, nn * 4));
    PYTO_TRY(device_alloc(h, (void**)&h->gmask, nn * 2));
    PYTO_TRY(device_alloc(h, (void**)&h->centers, (size_t)ngroups * 3 * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->sinfo, sinfo.size() * sizeof(SpecialElem)));
    PYTO_TRY(device_alloc(h, (void**)&h->chunks, chunks.size() * sizeof(SpecialChunk)));
    // partials per (chunk, slot): 6 values (structural) or 1 (thermal) per slot
    const size_t pstride = th ? 8 : 48;
    PYTO_TRY(device_alloc(h, (void**)&h->cpart, std::max<size_t>(chunks.size(), 1) * pstride * sizeof(double)));
    PYTO_TRY(device_alloc(h, (void**)&h->gnodes, nn * sizeof(int32_t)));
    PYTO_TRY(device_alloc(h, (void**)&h->nchunk, nchunk.size() * sizeof(int2)));
    PYTO_TRY(device_alloc(h, (void**)&h->nchunk_grp, nchunk.size() * sizeof(int32_t)));
    PYTO_TRY(device_alloc(h, (void**)&h->npart, nchunk.size() * pstride * sizeof(double)));
    PYTO_TRY(device_alloc(h, (void**)&h->nptr, nptr.size() * sizeof(int32_t)));
    PYTO_TRY(device_alloc(h, (void**)&h->nidx, nidx.size() * sizeof(int32_t)));
    PYTO_TRY(device_alloc(h, (void**)&h->gptr, gptr.size() * sizeof(int32_t)));
    PYTO_TRY(device_alloc(h, (void**)&h->gidx, std::max<size_t>(gidx.size(), 1) * sizeof(int32_t)));
    h->npairs = (int)gal_pairs.size();
    PYTO_TRY(device_alloc(h, (void**)&h->pwave, std::max<size_t>(pwave.size(), 1) * sizeof(int2)));
    PYTO_TRY(device_alloc(h, (void**)&h->pslot, std::max<size_t>(pslot.size(), 1) * sizeof(SpecialSlot)));
    PYTO_TRY(device_alloc(h, (void**)&h->pnodes, std::max<size_t>(pnodes.size(), 1) * 4));
    PYTO_TRY(device_alloc(h, (void**)&h->pcen, std::max<size_t>(pcen.size(), 1) * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->pE64, std::max<size_t>(pslot.size(), 1) * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->pE32, std::max<size_t>(pslot.size(), 1) * 4));
    PYTO_TRY(device_alloc(h, (void**)&h->gal_off, std::max<size_t>(gal_off.size(), 1) * 4));
    PYTO_TRY(device_alloc(h, (void**)&h->gal_part, std::max<int64_t>(galsz, 1) * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->gal_pairs, std::max<size_t>(gal_pairs.size(), 1) * sizeof(int2)));
    PYTO_TRY(device_alloc(h, (void**)&h->gal_pptr, gal_pptr.size() * 4));
    PYTO_TRY(device_alloc(h, (void**)&h->gal_pidx, std::max<size_t>(gal_pidx.size(), 1) * 4));
    PYTO_TRY(device_alloc(h, (void**)&h->Ec, (size_t)ld * ld * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->Einv, (size_t)ld * ld * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->cwork, (size_t)ld * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->ct, (size_t)2 * ld * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->cmu, (size_t)ld * 8));
    PYTO_TRY(device_alloc(h, (void**)&h->cres, (size_t)ld * 8));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->node_group, node_group, nn * 4, hipMemcpyDeviceToDevice, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->gmask, gmask.data(), nn * 2, hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->centers, centers, (size_t)ngroups * 24, hipMemcpyDeviceToDevice, s));
    if (!special.empty()) {
      PYTO_CHECK_HIP(hipMemcpyAsync(h->sinfo, sinfo.data(), sinfo.size() * sizeof(SpecialElem),
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->chunks, chunks.data(), chunks.size() * sizeof(SpecialChunk),
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->gal_off, gal_off.data(), gal_off.size() * 4,
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->pwave, pwave.data(), pwave.size() * sizeof(int2),
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->pslot, pslot.data(), pslot.size() * sizeof(SpecialSlot),
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->pnodes, pnodes.data(), pnodes.size() * 4,
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->pcen, pcen.data(), pcen.size() * 8, hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->gal_pairs, gal_pairs.data(), gal_pairs.size() * sizeof(int2),
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->gal_pptr, gal_pptr.data(), gal_pptr.size() * 4,
                                    hipMemcpyHostToDevice, s));
      PYTO_CHECK_HIP(hipMemcpyAsync(h->gal_pidx, gal_pidx.data(), gal_pidx.size() * 4,
                                    hipMemcpyHostToDevice, s));
    }
    PYTO_CHECK_HIP(hipMemcpyAsync(h->gnodes, gnodes.data(), nn * sizeof(int32_t),
                                  hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->nchunk, nchunk.data(), nchunk.size() * sizeof(int2),
                                  hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->nchunk_grp, nchunk_grp.data(),
                                  nchunk_grp.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->nptr, nptr.data(), nptr.size() * sizeof(int32_t),
                                  hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->nidx, nidx.data(), nidx.size() * sizeof(int32_t),
                                  hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->gptr, gptr.data(), gptr.size() * sizeof(int32_t),
                                  hipMemcpyHostToDevice, s));
    if (!gidx.empty())
      PYTO_CHECK_HIP(hipMemcpyAsync(h->gidx, gidx.data(), gidx.size() * sizeof(int32_t),
                                    hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipMemsetAsync(h->ct, 0, (size_t)2 * ld * 8, s));
    PYTO_CHECK_HIP(hipMemsetAsync(h->cmu, 0, (size_t)ld * 8, s));
    PYTO_CHECK_HIP(hipStreamSynchronize(s));
    return PYTO_OK;
  };
  const int rc = install();
  if (rc != PYTO_OK) {
    release_coarse(h);
    return rc;
  }
  if (h->density_set) PYTO_TRY(coarse_setup(h, s));
  return PYTO_OK;
}

int pyto_set_density_batch(int n, pyto_handle* const* handles, const double* const* x,
                           void* stream) {
  PYTO_REQUIRE(n >= 1 && handles && x, PYTO_EINVAL, "null argument or n < 1");
  for (int i = 0; i < n; ++i) {
    PYTO_REQUIRE(handles[i] && x[i], PYTO_EINVAL, "handle or density %d is null", i);
    for (int j = 0; j < i; ++j)
      PYTO_REQUIRE(handles[j] != handles[i], PYTO_EINVAL, "handle %d is listed twice", i);
  }
  std::vector<Handle*> hs(handles, handles + n);
  return set_density_many(n, hs.data(), x, as_stream(stream));
}

int pyto_set_coarse_shift(pyto_handle* h, double abs_shift, double rel_shift) {
  PYTO_REQUIRE(h, PYTO_EINVAL, "null handle");
  PYTO_REQUIRE(abs_shift >= 0 && rel_shift >= 0, PYTO_EINVAL, "shifts must be >= 0");
  h->shift_abs = abs_shift;
  h->shift_rel = rel_shift;
  return PYTO_OK;
}

int pyto_restrict_f64(pyto_handle* h, const double* v, double* t, void* stream) {
  PYTO_REQUIRE(h && v && t, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->ngroups > 0, PYTO_ESTATE, "no deflation groups set");
  return launch_restrict_full<double>(h, v, t, as_stream(stream));
}

int pyto_prolong_f64(pyto_handle* h, const double* mu, double* v, void* stream) {
  PYTO_REQUIRE(h && mu && v, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->ngroups > 0, PYTO_ESTATE, "no deflation groups set");
  return launch_prolong<double, false>(h, v, nullptr, mu, nullptr, true, nullptr,
                                       as_stream(stream), nullptr, 0);
}

int pyto_coarse_solve_f64(pyto_handle* h, const double* t, double* mu, void* stream) {
  PYTO_REQUIRE(h && t && mu, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->coarse_ready, PYTO_ESTATE, "coarse space not assembled (set groups and density)");
  return coarse_apply(h, t, nullptr, nullptr, mu, nullptr, as_stream(stream), true);
}

int pyto_coarse_matrix_f64(pyto_handle* h, double* E, void* stream) {
  PYTO_REQUIRE(h && E, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->coarse_ready, PYTO_ESTATE, "coarse space not assembled (set groups and density)");
  const int ld = round_up64(h->nc);
  PYTO_CHECK_HIP(hipMemcpy2DAsync(E, sizeof(double) * h->nc, h->Ec, sizeof(double) * ld,
                                  sizeof(double) * h->nc, h->nc, hipMemcpyDeviceToDevice,
                                  as_stream(stream)));
  return PYTO_OK;
}

}  // extern "C"
