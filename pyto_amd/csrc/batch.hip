// Batched Jacobi PCG over many independent systems (designs x load cases)
// in FP32 inside FP64 iterative refinement.
//
// Every kernel is ONE launch for all systems of the batch: a device table of
// systems (grid, E_e, Jacobi diagonal and fixed mask of the system's operator,
// the system's vectors) and prefix sums of per-system block counts map each
// workgroup to its system (binary search) and to its share of that system.
// Each system has its own scalars and flags, so each stops on its own test;
// launches of finished systems return at once.  The inner iteration
// (operator, update, scalars, direction) is captured once as a hipGraph of
// kBatchGraphIters iterations and replayed until every system has stopped.
//
// Per system the algorithm is the plain Jacobi PCG of the reference's PCG
// branch (src/linear_solvers.py:157-176 with _jacobi_preconditioner,
// :71-88), with the FP64 refinement of the single-system mixed solver:
// r = f - K u in FP64, inner FP32 PCG on K d = r / sigma to
// min(0.1, max(inner_rtol, 0.3 rtol / rel)), u += sigma d, until the FP64
// sqrt(r^T M r / r0^T M r0) <= rtol.  sigma = sqrt(r0^T M r0) is fixed for the
// solve, so an inner run can continue the previous run's search direction
// (reliable update: beta = r^T M r / rz_prev) while that run converged and its
// recurrence residual still tracks the new true one (within 1.25 in norm);
// otherwise it restarts from p = M r.  As in the single-system solver, the
// inner run stores z = M^-1 r instead of r: the CG update forms z -= alpha
// M^-1 q with the inverse taken from each node's sum of E_e, and the
// direction updates and restrictions read z alone.
//
// Deflation: a system whose handle has deflation groups runs the deflated PCG
// of the single-system solver inside the same launches.  Each inner run starts
// from x0 = W E^-1 W^T b, and every direction is projected, p = M r + beta p -
// W E^-1 W^T K (M r + beta p), with W^T K v restricted over the handle's
// special elements.  The restrictions, the per-group gathers, the coarse GEMV
// and the prolongation are one launch each for all deflated systems.
#include <algorithm>
#include <vector>

#include "matvec.hpp"

namespace pyto {

constexpr int kBatchGraphIters = 16;   // inner iterations per captured graph

struct BatchSys {
  GridDims g;
  ModalConsts<float> mc32;
  ModalConsts<double> mc64;
  const float* E32;
  const double* E64;
  const float* invd32;
  const double* invd64;
  const uint8_t* mask;
  const uint8_t* tileflags;
  // FP32 Jacobi from the node sums of E_e (jacobi_inv32): the handle's sums
  // and the element's diagonal entry per component
  const float* esum32;
  float kd[3];
  // the inner DPCG stores z = M^-1 r, as the single-system solver does
  float *z, *p, *q, *x, *b;
  double *r64, *q64;
  const double* f;
  double* u;
  int64_t ndof;
  int tiles_x, ntiles, zchunk;
  // deflation (ngroups == 0: none); the handle's coarse space and the
  // system's own coarse vectors
  int ngroups, nc, ld;
  const int32_t* node_group;
  const uint16_t* gmask;   // group | fixed bits << 12 (the handle's)
  const double* centers;
  // the handle's block-padded restriction tables
  const int2* pwave;
  const SpecialSlot* pslot;
  const int32_t* pnodes;
  const double* pcen;
  const float* pE32;
  const int32_t* gptr;
  const int32_t* gidx;
  const double* Einv;
  double *cpart, *ct, *cmu;
  // the full restriction W^T v of the start of an inner run: node blocks of
  // one group each, their partials, per-group lists of them
  const int32_t* gnodes;
  const int2* nchunk;
  const int32_t* nchunk_grp;
  const int32_t* nptr;
  const int32_t* nidx;
  double* npart;
};

// Outer-refinement flags: the system has met rtol (or has a zero load); the
// next inner run continues the previous run's direction.
constexpr int F_ODONE = 7;
constexpr int F_RESUME = 8;
// largest ratio of true to recurrence residual norm at which an inner run
// continues the previous one's direction
constexpr double kBatchResumeDrift = 1.25;

// Every prefix-sum array pre[0..n] (blocks, groups or rows per system) is
// followed on the device by its map: pre[n + 1 + b] is the system of item b
// (pre[s] <= b < pre[s + 1]); one load instead of a search.
__device__ __forceinline__ int find_sys(const int* __restrict__ pre, int n, int b) {
  return pre[n + 1 + b];
}

// ---- operator --------------------------------------------------------------
template <bool CUBE, int TY>
__global__ __launch_bounds__(kTile * TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_bmatvec_inner(const BatchSys* __restrict__ sys,
                                                       const int* __restrict__ pre, int nsys,
                                                       const int* __restrict__ flags,
                                                       double* __restrict__ partials) {
  const int gb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int s = find_sys(pre, nsys, gb);
  const BatchSys& S = sys[s];
  MatvecArgs<float> a;
  a.u = S.p;
  a.y = S.q;
  a.Ee = S.E32;
  a.mask = S.mask;
  a.tileflags = S.tileflags;
  a.g = S.g;
  a.mc = S.mc32;
  a.zchunk = S.zchunk;
  a.tiles_x = S.tiles_x;
  a.ntiles = S.ntiles;
  a.partials = partials;
  a.done = flags + s * kNumFlags + F_DONE;
  a.zero_vec = nullptr;
  a.zero_n = 0;
  matvec_tile<float, CUBE, 3, true, 1, false, TY>(a, gb - pre[s], gb);
}

// q = K x (FP32) for the systems still refining: the start of a deflated
// inner run (x = x0).
template <bool CUBE, int TY>
__global__ __launch_bounds__(kTile * TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_bmatvec_start(const BatchSys* __restrict__ sys,
                                                       const int* __restrict__ pre, int nsys,
                                                       const int* __restrict__ flags) {
  const int gb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int s = find_sys(pre, nsys, gb);
  const BatchSys& S = sys[s];
  MatvecArgs<float> a;
  a.u = S.x;
  a.y = S.q;
  a.Ee = S.E32;
  a.mask = S.mask;
  a.tileflags = S.tileflags;
  a.g = S.g;
  a.mc = S.mc32;
  a.zchunk = S.zchunk;
  a.tiles_x = S.tiles_x;
  a.ntiles = S.ntiles;
  a.partials = nullptr;
  a.done = flags + s * kNumFlags + F_ODONE;
  a.zero_vec = nullptr;
  a.zero_n = 0;
  matvec_tile<float, CUBE, 3, true, 1, false, TY>(a, gb - pre[s], gb);
}

template <bool CUBE, int TY>
__global__ __launch_bounds__(kTile * TY) __attribute__((amdgpu_waves_per_eu(CUBE ? 3 : 2, 8))) void k_bmatvec_outer(const BatchSys* __restrict__ sys,
                                                       const int* __restrict__ pre, int nsys,
                                                       const int* __restrict__ flags) {
  const int gb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int s = find_sys(pre, nsys, gb);
  const BatchSys& S = sys[s];
  MatvecArgs<double> a;
  a.u = S.u;
  a.y = S.q64;
  a.Ee = S.E64;
  a.mask = S.mask;
  a.tileflags = S.tileflags;
  a.g = S.g;
  a.mc = S.mc64;
  a.zchunk = S.zchunk;
  a.tiles_x = S.tiles_x;
  a.ntiles = S.ntiles;
  a.partials = nullptr;
  a.done = flags + s * kNumFlags + F_ODONE;
  a.zero_vec = nullptr;
  a.zero_n = 0;
  matvec_tile<double, CUBE, 3, true, 1, false, TY>(a, gb - pre[s], gb);
}

// ---- vector kernels: a block works on a contiguous share of one system ----
struct VecShare {
  int s;        // system
  int lb, nb;   // block within the system, blocks of the system
};
__device__ __forceinline__ VecShare vec_share(const int* __restrict__ vpre, int nsys) {
  VecShare v;
  v.s = find_sys(vpre, nsys, blockIdx.x);
  v.lb = blockIdx.x - vpre[v.s];
  v.nb = vpre[v.s + 1] - vpre[v.s];
  return v;
}

// z = M^-1 r and the r^T M^-1 r term of one DOF from r (FP64) and the
// stored inverse di: z = di r, term r di r.
__device__ __forceinline__ float z_of_r(double r, float di, double& term) {
  term = r * (double)di * r;
  return (float)((double)di * r);
}

// x = 0, z = M^-1 b, partials of b^T M^-1 b (systems still refining).
__global__ __launch_bounds__(256) void k_binit(const BatchSys* __restrict__ sys,
                                               const int* __restrict__ vpre, int nsys,
                                               const int* __restrict__ flags,
                                               double* __restrict__ part) {
  const VecShare v = vec_share(vpre, nsys);
  __shared__ double red[4];
  if (flags[v.s * kNumFlags + F_ODONE]) return;
  const BatchSys& S = sys[v.s];
  const int64_t nv = S.ndof / 4;
  const float4* bv = reinterpret_cast<const float4*>(S.b);
  const float4* dv = reinterpret_cast<const float4*>(S.invd32);
  float4* zv = reinterpret_cast<float4*>(S.z);
  float4* xv = reinterpret_cast<float4*>(S.x);
  double acc = 0.0, t;
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < nv; i += (int64_t)v.nb * 256) {
    const float4 bb = bv[i], dd = dv[i];
    float4 zz;
    zz.x = z_of_r((double)bb.x, dd.x, t);
    acc += t;
    zz.y = z_of_r((double)bb.y, dd.y, t);
    acc += t;
    zz.z = z_of_r((double)bb.z, dd.z, t);
    acc += t;
    zz.w = z_of_r((double)bb.w, dd.w, t);
    acc += t;
    zv[i] = zz;
    xv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (v.lb == 0)
    for (int64_t i = nv * 4 + threadIdx.x; i < S.ndof; i += 256) {
      S.z[i] = z_of_r((double)S.b[i], S.invd32[i], t);
      acc += t;
      S.x[i] = 0.f;
    }
  const double bs = block_reduce_sum256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// Start of an inner run in a batch with deflated systems: x0 = W mu (fixed
// DOFs 0) for a deflated system, with mu = E^-1 W^T b from k_bgemv<true>;
// x0 = 0 for the others.
__global__ __launch_bounds__(256) void k_bstart_x(const BatchSys* __restrict__ sys,
                                                  const int* __restrict__ vpre, int nsys,
                                                  const int* __restrict__ flags) {
  const VecShare v = vec_share(vpre, nsys);
  if (flags[v.s * kNumFlags + F_ODONE]) return;
  const BatchSys& S = sys[v.s];
  if (S.ngroups == 0) {
    for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < S.ndof; i += (int64_t)v.nb * 256)
      S.x[i] = 0.f;
    return;
  }
  const int64_t nn = S.ndof / 3;
  for (int64_t n = (int64_t)v.lb * 256 + threadIdx.x; n < nn; n += (int64_t)v.nb * 256) {
    int ix, iy, iz;
    node_coords(n, S.g, ix, iy, iz);
    const int gid = S.node_group[n];
    const double rx = ix * S.g.hx - S.centers[3 * gid];
    const double ry = iy * S.g.hy - S.centers[3 * gid + 1];
    const double rz = iz * S.g.hz - S.centers[3 * gid + 2];
    const double* m6 = S.cmu + 6 * gid;
    const double wm[3] = {m6[0] + rz * m6[4] - ry * m6[5], m6[1] - rz * m6[3] + rx * m6[5],
                          m6[2] + ry * m6[3] - rx * m6[4]};
    const uint8_t m = S.mask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) S.x[3 * n + c] = ((m >> c) & 1) ? 0.f : (float)wm[c];
  }
}

// Start of a deflated inner run: z = M^-1 (b - q) (q = K x0, r formed in
// FP64), partials of r^T M^-1 r, in k_binit's order (a system without groups
// has q = 0 and gets k_binit's bits).
__global__ __launch_bounds__(256) void k_binit_q(const BatchSys* __restrict__ sys,
                                                 const int* __restrict__ vpre, int nsys,
                                                 const int* __restrict__ flags,
                                                 double* __restrict__ part) {
  const VecShare v = vec_share(vpre, nsys);
  __shared__ double red[4];
  if (flags[v.s * kNumFlags + F_ODONE]) return;
  const BatchSys& S = sys[v.s];
  const int64_t nv = S.ndof / 4;
  const float4* bv = reinterpret_cast<const float4*>(S.b);
  const float4* qv = reinterpret_cast<const float4*>(S.q);
  const float4* dv = reinterpret_cast<const float4*>(S.invd32);
  float4* zv = reinterpret_cast<float4*>(S.z);
  double acc = 0.0, t;
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < nv; i += (int64_t)v.nb * 256) {
    const float4 bb = bv[i], qq = qv[i], dd = dv[i];
    float4 zz;
    zz.x = z_of_r((double)bb.x - (double)qq.x, dd.x, t);
    acc += t;
    zz.y = z_of_r((double)bb.y - (double)qq.y, dd.y, t);
    acc += t;
    zz.z = z_of_r((double)bb.z - (double)qq.z, dd.z, t);
    acc += t;
    zz.w = z_of_r((double)bb.w - (double)qq.w, dd.w, t);
    acc += t;
    zv[i] = zz;
  }
  if (v.lb == 0)
    for (int64_t i = nv * 4 + threadIdx.x; i < S.ndof; i += 256) {
      S.z[i] = z_of_r((double)S.b[i] - (double)S.q[i], S.invd32[i], t);
      acc += t;
    }
  const double bs = block_reduce_sum256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// One block per system: rz0, inner tolerance and flags of a new inner solve.
__global__ __launch_bounds__(256) void k_bsetup(const int* __restrict__ vpre, int nsys,
                                                double* __restrict__ scal,
                                                int* __restrict__ flags,
                                                const double* __restrict__ part, double rtol,
                                                double inner_rtol) {
  __shared__ double red[5];
  const int s = blockIdx.x;
  int* fl = flags + s * kNumFlags;
  double* sc = scal + s * kNumScalars;
  if (fl[F_ODONE]) {
    if (threadIdx.x == 0) {
      fl[F_DONE] = 1;
      fl[F_ITERS] = 0;
      fl[F_BREAKDOWN] = 0;
      fl[F_STALL] = 0;
    }
    return;
  }
  const double rz = block_sum_partials(part + vpre[s], vpre[s + 1] - vpre[s], red);
  if (threadIdx.x == 0) {
    sc[S_RZ] = rz;
    sc[S_RZ0] = rz;
    const double want = 0.3 * rtol / fmax(sc[S_TMP2], 1e-300);
    sc[S_RTOL] = fmin(0.1, fmax(inner_rtol, want));
    sc[S_RELMIN] = 1e300;
    fl[F_DONE] = rz == 0.0 ? 1 : 0;
    fl[F_ITERS] = 0;
    fl[F_BREAKDOWN] = 0;
    fl[F_STALL] = 0;
  }
}

// alpha = rz / p^T q (the system's matvec partials), z -= alpha M^-1 q with
// the inverse formed from the node sums (cg_z_node32, the single-system FP32
// update's arithmetic), r^T M^-1 r partials.  A thread owns 16-byte vectors
// (4 DOFs on nodes n0 and n0 + 1); n0 and the component c0 are formed once
// and advanced by the loop's constant step.
__global__ __launch_bounds__(256) void k_bcg_update(const BatchSys* __restrict__ sys,
                                                    const int* __restrict__ vpre,
                                                    const int* __restrict__ mpre, int nsys,
                                                    const double* __restrict__ scal,
                                                    const int* __restrict__ flags,
                                                    const double* __restrict__ part_mv,
                                                    double* __restrict__ part_v) {
  const VecShare v = vec_share(vpre, nsys);
  __shared__ double red[5];
  __shared__ double red4[4];
  if (flags[v.s * kNumFlags + F_DONE]) return;
  const BatchSys& S = sys[v.s];
  const double pkp = block_sum_partials(part_mv + mpre[v.s], mpre[v.s + 1] - mpre[v.s], red);
  // pKp <= 0 or NaN: z unchanged (the scalar step flags the breakdown)
  const double alpha = pkp > 0.0 ? scal[v.s * kNumScalars + S_RZ] / pkp : 0.0;
  const int64_t nv = S.ndof / 4;
  float4* zv = reinterpret_cast<float4*>(S.z);
  const float4* qv = reinterpret_cast<const float4*>(S.q);
  const float kd[3] = {S.kd[0], S.kd[1], S.kd[2]};
  double acc = 0.0, t;
  const int64_t stride = (int64_t)v.nb * 256;
  int64_t i = (int64_t)v.lb * 256 + threadIdx.x;
  int64_t n0 = (4 * i) / 3;
  int c0 = (int)(4 * i - 3 * n0);
  const int64_t kn = (4 * stride) / 3;
  const int kc = (int)(4 * stride - 3 * kn);
  for (; i < nv; i += stride) {
    float4 zz = zv[i];
    const float4 qq = qv[i];
    const float s0 = S.esum32[n0], s1 = S.esum32[n0 + 1];
    const uint8_t m0 = S.mask[n0], m1 = S.mask[n0 + 1];
    float* zl[4] = {&zz.x, &zz.y, &zz.z, &zz.w};
    const float ql[4] = {qq.x, qq.y, qq.z, qq.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + j;
      const bool second = c >= 3;
      const int cl = second ? c - 3 : c;
      *zl[j] = cg_z_node32(*zl[j], ql[j], kd[cl], second ? s1 : s0,
                           ((second ? m1 : m0) >> cl) & 1, alpha, t);
      acc += t;
    }
    zv[i] = zz;
    c0 += kc;
    const int carry = c0 >= 3 ? 1 : 0;
    c0 -= 3 * carry;
    n0 += kn + carry;
  }
  if (v.lb == 0)
    for (int64_t d = nv * 4 + threadIdx.x; d < S.ndof; d += 256) {
      const int64_t nd = d / 3;
      const int c = (int)(d - 3 * nd);
      S.z[d] = cg_z_node32(S.z[d], S.q[d], kd[c], S.esum32[nd], (S.mask[nd] >> c) & 1, alpha, t);
      acc += t;
    }
  const double bs = block_reduce_sum256(acc, red4);
  if (threadIdx.x == 0) part_v[blockIdx.x] = bs;
}

// One block per system: the iteration's scalars, stop and stagnation tests
// (as the single-system solver's k_scalars).
__global__ __launch_bounds__(256) void k_bscalars(const int* __restrict__ vpre,
                                                  const int* __restrict__ mpre, int nsys,
                                                  double* __restrict__ scal,
                                                  int* __restrict__ flags,
                                                  const double* __restrict__ part_mv,
                                                  const double* __restrict__ part_v) {
  __shared__ double red[5];
  const int s = blockIdx.x;
  int* fl = flags + s * kNumFlags;
  if (fl[F_DONE]) return;
  double* sc = scal + s * kNumScalars;
  const double pkp = block_sum_partials(part_mv + mpre[s], mpre[s + 1] - mpre[s], red);
  const double rz_new = block_sum_partials(part_v + vpre[s], vpre[s + 1] - vpre[s], red);
  if (threadIdx.x == 0) pcg_scalar_step(sc, fl, pkp, rz_new, kFloor32);
}

// x += alpha p (unless FIRST), p = z + beta p.  FIRST: the start of an inner
// run, p = z, or with F_RESUME p = z + (rz / rz_prev) p (x is 0 then).  Also
// the direction update of the systems without groups in a batch that has
// deflated ones, so that they take the same bits.
template <bool FIRST>
__device__ __forceinline__ void bdir_plain(const BatchSys& S, const VecShare& v,
                                           const double* sc, bool resume) {
  const float al = FIRST ? 0.f : (float)sc[S_ALPHA];
  const float be = FIRST ? (resume ? (float)(sc[S_RZ] / sc[S_RZPREV]) : 0.f) : (float)sc[S_BETA];
  const int64_t nv = S.ndof / 4;
  float4* pv = reinterpret_cast<float4*>(S.p);
  float4* xv = reinterpret_cast<float4*>(S.x);
  const float4* zv = reinterpret_cast<const float4*>(S.z);
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < nv; i += (int64_t)v.nb * 256) {
    const float4 zz = zv[i];
    float4 pp = (FIRST && !resume) ? make_float4(0.f, 0.f, 0.f, 0.f) : pv[i];
    if (!FIRST) {
      float4 xx = load_nt(xv + i);
      xx.x += al * pp.x;
      xx.y += al * pp.y;
      xx.z += al * pp.z;
      xx.w += al * pp.w;
      store_nt(xv + i, xx);
    }
    pp.x = zz.x + be * pp.x;
    pp.y = zz.y + be * pp.y;
    pp.z = zz.z + be * pp.z;
    pp.w = zz.w + be * pp.w;
    pv[i] = pp;
  }
  if (v.lb == 0)
    for (int64_t i = nv * 4 + threadIdx.x; i < S.ndof; i += 256) {
      const float pp = (FIRST && !resume) ? 0.f : S.p[i];
      if (!FIRST) S.x[i] += al * pp;
      S.p[i] = S.z[i] + be * pp;
    }
}

template <bool FIRST>
__global__ __launch_bounds__(256) void k_bdir(const BatchSys* __restrict__ sys,
                                              const int* __restrict__ vpre, int nsys,
                                              const double* __restrict__ scal,
                                              const int* __restrict__ flags) {
  const VecShare v = vec_share(vpre, nsys);
  if (flags[v.s * kNumFlags + F_DONE]) return;
  bdir_plain<FIRST>(sys[v.s], v, scal + v.s * kNumScalars,
                    FIRST && flags[v.s * kNumFlags + F_RESUME]);
}

// The direction update of deflated systems: x += alpha p (unless FIRST),
// p = z + beta p - W mu (FIRST: beta as in k_bdir), formed in FP64 as the
// single-system solver does (z and W mu cancel in near-void groups).  A block
// works on one system; W mu of group q at position X is a_q + w_q x X with
// w = (mu3, mu4, mu5), a = (mu0, mu1, mu2) - w x c_q, staged in LDS.
constexpr int kBLdsGroups = 512;
template <bool FIRST>
__global__ __launch_bounds__(256) void k_bdir_defl(const BatchSys* __restrict__ sys,
                                                   const int* __restrict__ vpre, int nsys,
                                                   const double* __restrict__ scal,
                                                   const int* __restrict__ flags) {
  const VecShare v = vec_share(vpre, nsys);
  if (flags[v.s * kNumFlags + F_DONE]) return;
  const BatchSys& S = sys[v.s];
  const double* sc = scal + v.s * kNumScalars;
  const bool resume = FIRST && flags[v.s * kNumFlags + F_RESUME];
  if (S.ngroups == 0) {
    bdir_plain<FIRST>(S, v, sc, resume);
    return;
  }
  // 6 values per group, sized by the launch for the batch's largest staged
  // group count (at most kBLdsGroups)
  extern __shared__ double aw[];
  const bool lds = S.ngroups <= kBLdsGroups;
  auto coeffs = [&](int q, double* e) {
    const double* m6 = S.cmu + 6 * q;
    const double wx = m6[3], wy = m6[4], wz = m6[5];
    const double cx = S.centers[3 * q], cy = S.centers[3 * q + 1], cz = S.centers[3 * q + 2];
    e[0] = m6[0] - (wy * cz - wz * cy);
    e[1] = m6[1] - (wz * cx - wx * cz);
    e[2] = m6[2] - (wx * cy - wy * cx);
    e[3] = wx;
    e[4] = wy;
    e[5] = wz;
  };
  if (lds) {
    for (int q = threadIdx.x; q < S.ngroups; q += blockDim.x) coeffs(q, aw + 6 * q);
    __syncthreads();
  }
  auto wmu3 = [&](int q, int ix, int iy, int iz, double* w) {
    double e[6];
    if (lds) {
#pragma unroll
      for (int k = 0; k < 6; ++k) e[k] = aw[6 * q + k];
    } else {
      coeffs(q, e);
    }
    const double X = ix * S.g.hx, Y = iy * S.g.hy, Z = iz * S.g.hz;
    w[0] = e[0] + e[4] * Z - e[5] * Y;
    w[1] = e[1] + e[5] * X - e[3] * Z;
    w[2] = e[2] + e[3] * Y - e[4] * X;
  };
  const double al = FIRST ? 0.0 : sc[S_ALPHA];
  const double be = FIRST ? (resume ? sc[S_RZ] / sc[S_RZPREV] : 0.0) : sc[S_BETA];
  const bool read_p = !FIRST || resume;
  // A thread owns 16-byte vectors (4 DOFs on nodes n0 and n0 + 1, the
  // single-system FP32 direction update's layout): W mu and the fixed bits
  // for both nodes, each DOF selecting its own; the node index, component and
  // coordinates formed once and advanced by the loop's constant step.
  const int64_t nv = S.ndof / 4;
  const int64_t stride = (int64_t)v.nb * 256;
  int64_t i = (int64_t)v.lb * 256 + threadIdx.x;
  int64_t n0 = (4 * i) / 3;
  int c0 = (int)(4 * i - 3 * n0);
  const int64_t kn = (4 * stride) / 3;
  const int kc = (int)(4 * stride - 3 * kn);
  int ix = 0, iy = 0, iz = 0, sxn = 0, syn = 0, szn = 0;
  if (i < nv) {
    node_coords(n0, S.g, ix, iy, iz);
    node_coords(kn, S.g, sxn, syn, szn);
  }
  float4* pv = reinterpret_cast<float4*>(S.p);
  float4* xv = reinterpret_cast<float4*>(S.x);
  const float4* zv = reinterpret_cast<const float4*>(S.z);
  for (; i < nv; i += stride) {
    // the next node's coordinates (n0 + 1 exists: DOF 4i + 3 lies on it)
    const bool wx = ix + 1 == S.g.sx;
    const int jx = wx ? 0 : ix + 1;
    const bool wy = wx && iy + 1 == S.g.sy;
    const int jy = wx ? (wy ? 0 : iy + 1) : iy;
    const int jz = wy ? iz + 1 : iz;
    // each node's group and fixed bits in one 16-bit word
    const uint32_t a0 = S.gmask[n0], a1 = S.gmask[n0 + 1];
    double w6[6];
    wmu3((int)(a0 & 0xfffu), ix, iy, iz, w6);
    wmu3((int)(a1 & 0xfffu), jx, jy, jz, w6 + 3);
    const uint32_t fx = (a0 >> 12) | ((a1 >> 12) << 3);
    float4 pp = read_p ? pv[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 zz = zv[i];
    if (!FIRST) {
      float4 xx = load_nt(xv + i);
      xx.x = (float)((double)xx.x + al * (double)pp.x);
      xx.y = (float)((double)xx.y + al * (double)pp.y);
      xx.z = (float)((double)xx.z + al * (double)pp.z);
      xx.w = (float)((double)xx.w + al * (double)pp.w);
      store_nt(xv + i, xx);
    }
    float* pl[4] = {&pp.x, &pp.y, &pp.z, &pp.w};
    const float zl[4] = {zz.x, zz.y, zz.z, zz.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // component c0 + j of the pair of nodes; c0 is 0, 1 or 2
      const double w = c0 == 0 ? w6[j] : (c0 == 1 ? w6[j + 1] : w6[j + 2]);
      const double wm = ((fx >> (c0 + j)) & 1u) ? 0.0 : w;
      *pl[j] = (float)((double)zl[j] + be * (double)*pl[j] - wm);
    }
    pv[i] = pp;
    // next trip: c0 + kc may carry one node (so may x and then y)
    c0 += kc;
    const int extra = c0 >= 3 ? 1 : 0;
    c0 -= 3 * extra;
    n0 += kn + extra;
    ix += sxn + extra;
    iy += syn;
    iz += szn;
    if (ix >= S.g.sx) {
      ix -= S.g.sx;
      ++iy;
    }
    if (iy >= S.g.sy) {
      iy -= S.g.sy;
      ++iz;
    }
  }
  // DOFs past the last whole vector, one at a time by the system's block 0
  if (v.lb == 0)
    for (int64_t d = nv * 4 + threadIdx.x; d < S.ndof; d += 256) {
      const int64_t nd = d / 3;
      const int c = (int)(d - 3 * nd);
      int jx, jy, jz;
      node_coords(nd, S.g, jx, jy, jz);
      const uint32_t a = S.gmask[nd];
      double w3[3];
      wmu3((int)(a & 0xfffu), jx, jy, jz, w3);
      const double wm = (((a >> 12) >> c) & 1u) ? 0.0 : w3[c];
      const float pd = read_p ? S.p[d] : 0.f;
      if (!FIRST) S.x[d] = (float)((double)S.x[d] + al * (double)pd);
      S.p[d] = (float)((double)S.z[d] + be * (double)pd - wm);
    }
}

// The x += alpha p that the stopping iteration leaves pending (converged
// systems that iterated in this inner solve only).
__global__ __launch_bounds__(256) void k_bfinal(const BatchSys* __restrict__ sys,
                                                const int* __restrict__ vpre, int nsys,
                                                const double* __restrict__ scal,
                                                const int* __restrict__ flags) {
  const VecShare v = vec_share(vpre, nsys);
  const int* fl = flags + v.s * kNumFlags;
  if (fl[F_ODONE] || !fl[F_DONE] || fl[F_BREAKDOWN] || fl[F_ITERS] == 0) return;
  const BatchSys& S = sys[v.s];
  const float al = (float)scal[v.s * kNumScalars + S_ALPHA];
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < S.ndof; i += (int64_t)v.nb * 256)
    S.x[i] += al * S.p[i];
}

// ---- deflation ----------------------------------------------------------------
// Chunk partials of W^T K v over a deflated system's special elements, v =
// z + beta p (z = M^-1 r as stored) (beta: the iteration's, or at the start of an inner run the
// resumed direction's rz / rz_prev, 0 when it restarts): one block per special
// block of each deflated system (prefix sums rpre), the single-system
// restriction's staging and element arithmetic.
template <bool CUBE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_brestrict(
    const BatchSys* __restrict__ sys, const int* __restrict__ rpre, int nsys,
    const double* __restrict__ scal, const int* __restrict__ flags, bool start) {
  constexpr int kStage = (3 * kSpecialBlockNodes + 255) / 256;
  __shared__ float vs[3 * kSpecialBlockNodes];
  __shared__ double csh[kSpecialBlockChunks][8 * 3];
  const int gb = blockIdx.x;
  const int s = find_sys(rpre, nsys, gb);
  const int* fl = flags + s * kNumFlags;
  if (fl[F_DONE]) return;
  const BatchSys& S = sys[s];
  const double* sc = scal + s * kNumScalars;
  const int b = gb - rpre[s];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  // the handle's padded tables: every load but the staged values depends on
  // the block alone (as in the single-system restriction)
  const int2 pw = S.pwave[b * kSpecialBlockChunks + w];   // (chunk or -1, ng)
  const int64_t slot = ((int64_t)b * kSpecialBlockChunks + w) * 64 + lane;
  const SpecialSlot se = S.pslot[slot];
  const float Ev = S.pE32[slot];
  uint32_t dof[kStage];
#pragma unroll
  for (int t = 0; t < kStage; ++t) {
    const int i = (int)threadIdx.x + 256 * t;
    const int q = i / 3;
    dof[t] = 3u * (uint32_t)S.pnodes[(int64_t)b * kSpecialBlockNodes + q] + (uint32_t)(i - 3 * q);
  }
  if (threadIdx.x < kSpecialBlockChunks * 24)
    (&csh[0][0])[threadIdx.x] = S.pcen[(int64_t)b * kSpecialBlockChunks * 24 + threadIdx.x];
  const double beta = start ? (fl[F_RESUME] ? sc[S_RZ] / sc[S_RZPREV] : 0.0) : sc[S_BETA];
  const bool has_p = beta != 0.0;
  float za[kStage], pa[kStage];
#pragma unroll
  for (int t = 0; t < kStage; ++t) {
    za[t] = S.z[dof[t]];
    pa[t] = has_p ? S.p[dof[t]] : 0.f;
  }
#pragma unroll
  for (int t = 0; t < kStage; ++t)
    vs[threadIdx.x + 256 * t] = (float)((double)za[t] + beta * (double)pa[t]);
  __syncthreads();
  if (pw.x < 0) return;
  const bool active = se.e >= 0;
  float f[8][3];
  uint64_t info = 0;
  int ix = 0, iy = 0, iz = 0;
  if (active) {
    const uint32_t lw[4] = {se.li.x, se.li.y, se.li.z, se.li.w};
    info = se.info;
    elem_coords(se.e, S.g, ix, iy, iz);
    float u[8][3];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int l = (int)((lw[j >> 1] >> (16 * (j & 1))) & 0xffffu);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        u[j][c] = ((info >> (24 + 3 * j + c)) & 1) ? 0.f : vs[3 * l + c];
    }
    if (CUBE)
      element_apply_cube<float>(u, Ev, S.mc32, f);
    else
      element_apply<float>(u, Ev, S.mc32, f);
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if ((info >> (24 + 3 * j + c)) & 1) f[j][c] = 0.f;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) f[j][0] = f[j][1] = f[j][2] = 0.f;
  }
  const double* cen = csh[w];
  for (int sl = 0; sl < pw.y; ++sl) {
    const double cx = cen[3 * sl], cy = cen[3 * sl + 1], cz = cen[3 * sl + 2];
    const float x0 = (float)(ix * S.g.hx - cx), y0 = (float)(iy * S.g.hy - cy),
                z0 = (float)(iz * S.g.hz - cz);
    const float hx = (float)S.g.hx, hy = (float)S.g.hy, hz = (float)S.g.hz;
    double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (!active || (int)((info >> (3 * j)) & 7) != sl) continue;
      const float rx = corner_a(j) ? x0 + hx : x0;
      const float ry = corner_b(j) ? y0 + hy : y0;
      const float rz = corner_c(j) ? z0 + hz : z0;
      const float fx = f[j][0], fy = f[j][1], fz = f[j][2];
      acc[0] += (double)fx;
      acc[1] += (double)fy;
      acc[2] += (double)fz;
      acc[3] += (double)(ry * fz - rz * fy);
      acc[4] += (double)(rz * fx - rx * fz);
      acc[5] += (double)(rx * fy - ry * fx);
    }
    const double tot = wave_reduce6(acc);
    if ((lane & 7) == 0 && lane < 48) S.cpart[(int64_t)pw.x * 48 + 6 * sl + (lane >> 3)] = tot;
  }
}

// Block partials of t = W^T b over all nodes (fixed DOFs excluded) at the
// start of an inner run: one block per node block of one group of each
// deflated system (prefix sums npre), the single-system full restriction's
// arithmetic and order.
__global__ __launch_bounds__(256) void k_brestrict_full(const BatchSys* __restrict__ sys,
                                                        const int* __restrict__ npre, int nsys,
                                                        const int* __restrict__ flags) {
  __shared__ double red[4][6];
  const int s = find_sys(npre, nsys, blockIdx.x);
  if (flags[s * kNumFlags + F_ODONE]) return;
  const BatchSys& S = sys[s];
  const int blk = blockIdx.x - npre[s];
  const int2 ch = S.nchunk[blk];
  const int gid = S.nchunk_grp[blk];
  double c[6] = {0, 0, 0, 0, 0, 0};
  if ((int)threadIdx.x < ch.y) {
    const int64_t n = S.gnodes[ch.x + threadIdx.x];
    int ix, iy, iz;
    node_coords(n, S.g, ix, iy, iz);
    const double rx = ix * S.g.hx - S.centers[3 * gid];
    const double ry = iy * S.g.hy - S.centers[3 * gid + 1];
    const double rz = iz * S.g.hz - S.centers[3 * gid + 2];
    const uint8_t m = S.mask[n];
    const double vx = (m & 1) ? 0.0 : (double)S.b[3 * n];
    const double vy = (m & 2) ? 0.0 : (double)S.b[3 * n + 1];
    const double vz = (m & 4) ? 0.0 : (double)S.b[3 * n + 2];
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
    S.npart[(int64_t)blk * 48 + j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
  }
}

// t = per-group sums of the partials: one wave per group of each deflated
// system (prefix sums gpre over groups), fixed order.  START: the full
// restriction's node-block partials of a new inner run (systems still
// refining); otherwise the special restriction's chunk partials (systems
// still iterating).
template <bool START>
__global__ __launch_bounds__(256) void k_bgather(const BatchSys* __restrict__ sys,
                                                 const int* __restrict__ gpre, int nsys,
                                                 const int* __restrict__ flags) {
  const int G = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (G >= gpre[nsys]) return;
  const int s = find_sys(gpre, nsys, G);
  if (flags[s * kNumFlags + (START ? F_ODONE : F_DONE)]) return;
  const BatchSys& S = sys[s];
  const int g = G - gpre[s];
  const int32_t* ptr = START ? S.nptr : S.gptr;
  const int32_t* idx = START ? S.nidx : S.gidx;
  const double* part = START ? S.npart : S.cpart;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int k = ptr[g] + lane; k < ptr[g + 1]; k += 64) {
    const double* c = part + (int64_t)idx[k] * 6;
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] += c[j];
  }
  const double t = wave_reduce6(acc);
  if ((lane & 7) == 0 && lane < 48) S.ct[6 * g + (lane >> 3)] = t;
}

// mu = E^-1 t for the deflated systems: one wave per coarse row of each run
// of consecutive systems on one handle (the load cases of a design, at most
// kBatchRhs; prefix sums dpre over the runs' rows), so that each row of the
// shared E^-1 is read once for all of them.  START as in k_bgather.  Per
// system the order of the sums is that of a single GEMV.
constexpr int kBatchRhs = 8;
template <bool START>
__global__ __launch_bounds__(256) void k_bgemv(const BatchSys* __restrict__ sys,
                                               const int2* __restrict__ drun,
                                               const int* __restrict__ dpre, int nruns,
                                               const int* __restrict__ flags) {
  const int R = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (R >= dpre[nruns]) return;
  const int d = find_sys(dpre, nruns, R);
  const int2 run = drun[d];   // (first system, count)
  const int row = R - dpre[d];
  bool act[kBatchRhs];
  const double* t[kBatchRhs];
  bool any = false;
#pragma unroll
  for (int k = 0; k < kBatchRhs; ++k) {
    act[k] = k < run.y && !flags[(run.x + k) * kNumFlags + (START ? F_ODONE : F_DONE)];
    t[k] = act[k] ? sys[run.x + k].ct : nullptr;
    any = any || act[k];
  }
  if (!any) return;
  const BatchSys& S = sys[run.x];
  const double* a = S.Einv + (int64_t)row * S.ld;
  double acc[kBatchRhs];
#pragma unroll
  for (int k = 0; k < kBatchRhs; ++k) acc[k] = 0.0;
  // the row's loads of 4 trips go out together; each lane adds its terms in
  // the order of j
  // (indices past the row are clamped to its last entry, their terms left out)
  constexpr int U = 4;
  for (int j0 = lane; j0 < S.nc; j0 += 64 * U) {
    double av[U];
    int jj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      jj[u] = min(j0 + 64 * u, S.nc - 1);
      av[u] = a[jj[u]];
    }
#pragma unroll
    for (int k = 0; k < kBatchRhs; ++k) {
      if (!act[k]) continue;
      double tv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) tv[u] = t[k][jj[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double p = av[u] * tv[u];
        if (j0 + 64 * u < S.nc) acc[k] += p;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kBatchRhs; ++k) {
    if (!act[k]) continue;
    const double v = wave_reduce_sum(acc[k]);
    if (lane == 0) sys[run.x + k].cmu[row] = v;
  }
}

// ---- FP64 outer refinement --------------------------------------------------
// u = 0 (start of a solve).
__global__ __launch_bounds__(256) void k_bzero_u(const BatchSys* __restrict__ sys,
                                                 const int* __restrict__ vpre, int nsys) {
  const VecShare v = vec_share(vpre, nsys);
  const BatchSys& S = sys[v.s];
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < S.ndof; i += (int64_t)v.nb * 256)
    S.u[i] = 0.0;
}

// r64 = f - q64 (fixed DOFs carry invd = 0 and drop out of r^T M r), partials.
__global__ __launch_bounds__(256) void k_bresid64(const BatchSys* __restrict__ sys,
                                                  const int* __restrict__ vpre, int nsys,
                                                  const int* __restrict__ flags, bool first,
                                                  double* __restrict__ part) {
  const VecShare v = vec_share(vpre, nsys);
  __shared__ double red[4];
  if (flags[v.s * kNumFlags + F_ODONE]) return;
  const BatchSys& S = sys[v.s];
  double acc = 0.0;
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < S.ndof; i += (int64_t)v.nb * 256) {
    const double ri = S.f[i] - (first ? 0.0 : S.q64[i]);
    S.r64[i] = ri;
    acc += ri * S.invd64[i] * ri;
  }
  const double bs = block_reduce_sum256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// One block per system: outer relative residual, stop test, scale.
__global__ __launch_bounds__(256) void k_bouter(const int* __restrict__ vpre, int nsys,
                                                double* __restrict__ scal,
                                                int* __restrict__ flags,
                                                const double* __restrict__ part, bool first,
                                                double rtol) {
  __shared__ double red[5];
  const int s = blockIdx.x;
  int* fl = flags + s * kNumFlags;
  if (fl[F_ODONE]) return;
  double* sc = scal + s * kNumScalars;
  const double rmr = block_sum_partials(part + vpre[s], vpre[s + 1] - vpre[s], red);
  if (threadIdx.x == 0) {
    if (first) {
      sc[S_TMP1] = rmr;
      sc[S_SCALE] = sqrt(rmr);   // sigma: fixed for the solve
      fl[F_RESUME] = 0;
    }
    const double rel = sc[S_TMP1] > 0.0 ? sqrt(rmr / sc[S_TMP1]) : 0.0;
    sc[S_TMP2] = rel;
    sc[S_TMP0] = rmr;
    if (rel <= rtol) {
      fl[F_ODONE] = 1;
    } else if (!first) {
      // the last inner run's recurrence r^T M r (S_RZ) against the new true
      // one, both in the inner units (divided by sigma^2)
      const double rz_true = rmr / (sc[S_SCALE] * sc[S_SCALE]);
      const bool clean = fl[F_DONE] && !fl[F_BREAKDOWN] && fl[F_ITERS] > 0;
      fl[F_RESUME] = clean && rz_true <= kBatchResumeDrift * kBatchResumeDrift * sc[S_RZ];
    }
  }
}

// b32 = r64 / scale.
__global__ __launch_bounds__(256) void k_bscale(const BatchSys* __restrict__ sys,
                                                const int* __restrict__ vpre, int nsys,
                                                const double* __restrict__ scal,
                                                const int* __restrict__ flags) {
  const VecShare v = vec_share(vpre, nsys);
  if (flags[v.s * kNumFlags + F_ODONE]) return;
  const BatchSys& S = sys[v.s];
  const double inv = 1.0 / scal[v.s * kNumScalars + S_SCALE];
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < S.ndof; i += (int64_t)v.nb * 256)
    S.b[i] = (float)(S.r64[i] * inv);
}

// u += scale * x32.
__global__ __launch_bounds__(256) void k_baxpy(const BatchSys* __restrict__ sys,
                                               const int* __restrict__ vpre, int nsys,
                                               const double* __restrict__ scal,
                                               const int* __restrict__ flags) {
  const VecShare v = vec_share(vpre, nsys);
  if (flags[v.s * kNumFlags + F_ODONE]) return;
  const BatchSys& S = sys[v.s];
  const double sc = scal[v.s * kNumScalars + S_SCALE];
  for (int64_t i = (int64_t)v.lb * 256 + threadIdx.x; i < S.ndof; i += (int64_t)v.nb * 256)
    S.u[i] += sc * (double)S.x[i];
}

// All flags of the batch to F_DONE (timing runs of the operator use DONE = 0).
__global__ void k_bset_flag(int* flags, int nsys, int slot, int value) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nsys; s += gridDim.x * blockDim.x)
    flags[s * kNumFlags + slot] = value;
}

// Grid-barrier probe (a measurement for the persistent-kernel alternative to
// replaying captured iteration graphs): `rounds` barriers of all gridDim.x
// workgroups, which a cooperative launch keeps resident together.  Barrier
// r: each block's thread 0 publishes its writes (release fence), adds 1 to
// the counter and waits until it reaches (r + 1) * blocks.  The wait is
// bounded: after spin_limit polls it sets *timed_out and every block leaves.
__global__ __launch_bounds__(256) void k_grid_barrier_probe(unsigned* __restrict__ counter,
                                                            int rounds, unsigned spin_limit,
                                                            int* __restrict__ timed_out) {
  __shared__ int quit;
  if (threadIdx.x == 0) quit = 0;
  const unsigned nb = gridDim.x;
  for (int r = 0; r < rounds; ++r) {
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      atomicAdd(counter, 1u);
      const unsigned target = (unsigned)(r + 1) * nb;
      unsigned polls = 0;
      while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++polls > spin_limit || __hip_atomic_load(timed_out, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)) {
          atomicExch(timed_out, 1);
          quit = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    if (quit) return;
  }
}

// ---- host side ----------------------------------------------------------------
struct Batch {
  int device = 0;
  int nsys = 0;
  bool cube = true;
  std::vector<pyto_handle*> handles;
  std::vector<int> groups_gen;     // each handle's groups generation at creation
  std::vector<BatchSys> hs;        // host copy of the table
  BatchSys* sys = nullptr;
  std::vector<int> h_mpre, h_vpre;
  int* mpre = nullptr;
  int* vpre = nullptr;
  double* scal = nullptr;          // nsys x kNumScalars, then the flags
  int* flags = nullptr;            // nsys x kNumFlags, inside scal's allocation
  double* host_scal = nullptr;     // pinned mirror of both (nsys x kCtlBytes)
  int* host_flags = nullptr;
  double* part_mv = nullptr;
  double* part_v = nullptr;
  int cap_mv = 0;
  std::vector<void*> bufs;
  hipStream_t gstream = nullptr;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  hipGraphExec_t graph = nullptr;
  int zchunk = 16;
  int rows = 16;                   // operator tile rows of every system
  // deflation: prefix sums over the deflated systems' special blocks, groups
  // and coarse rows (host and device)
  bool defl = false;
  size_t dir_lds = 0;              // dynamic LDS of k_bdir_defl
  std::vector<int> h_rpre, h_gpre, h_npre, h_dpre;
  std::vector<int2> h_drun;        // runs of systems sharing one deflated handle
  int *rpre = nullptr, *gpre = nullptr, *npre = nullptr, *dpre = nullptr;
  int2* drun = nullptr;
};

}  // namespace pyto

struct pyto_batch : public pyto::Batch {};

namespace pyto {

// Upload prefix sums pre[0..n] with the item -> system map appended (the
// layout find_sys reads), replacing *dev.  Synchronises `s`.
static int upload_prefix(const std::vector<int>& pre, int** dev, hipStream_t s) {
  const int n = (int)pre.size() - 1;
  std::vector<int> buf(pre);
  buf.reserve(pre.size() + (size_t)pre[n]);
  for (int i = 0; i < n; ++i)
    for (int k = pre[i]; k < pre[i + 1]; ++k) buf.push_back(i);
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  if (*dev) PYTO_CHECK_HIP(hipFree(*dev));
  *dev = nullptr;
  PYTO_CHECK_HIP(hipMalloc((void**)dev, buf.size() * sizeof(int)));
  PYTO_CHECK_HIP(hipMemcpyAsync(*dev, buf.data(), buf.size() * sizeof(int), hipMemcpyHostToDevice, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  return PYTO_OK;
}

// Tile rows (the systems' handles are set to them) and z-chunk of the
// batched operator; rebuilds the block tables.
static int batch_geometry(Batch* b, int zc, int rows, hipStream_t s) {
  for (int i = 0; i < b->nsys; ++i) {
    PYTO_TRY(set_tile_rows(b->handles[i], rows, s));
    BatchSys& S = b->hs[i];
    S.ntiles = (int)tiles_for(S.g, rows, &S.tiles_x);
  }
  b->rows = rows;
  b->h_mpre.assign(b->nsys + 1, 0);
  b->h_vpre.assign(b->nsys + 1, 0);
  for (int i = 0; i < b->nsys; ++i) {
    BatchSys& S = b->hs[i];
    const int z = std::max(1, std::min({zc, 62, S.g.sz}));
    S.zchunk = z;
    const int mb = S.ntiles * ((S.g.sz + z - 1) / z);
    b->h_mpre[i + 1] = b->h_mpre[i] + mb;
    const int64_t nv = (S.ndof / 4 + 1023) / 1024;   // ~4 float4 per thread
    b->h_vpre[i + 1] = b->h_vpre[i] + (int)std::max<int64_t>(1, std::min<int64_t>(nv, 2048));
  }
  b->zchunk = zc;
  const int nm = b->h_mpre[b->nsys];
  if (nm > b->cap_mv) {
    if (b->part_mv) hipFree(b->part_mv);
    b->part_mv = nullptr;
    PYTO_CHECK_HIP(hipMalloc((void**)&b->part_mv, (size_t)nm * sizeof(double)));
    b->cap_mv = nm;
  }
  PYTO_CHECK_HIP(hipMemcpyAsync(b->sys, b->hs.data(), b->hs.size() * sizeof(BatchSys),
                                hipMemcpyHostToDevice, s));
  PYTO_TRY(upload_prefix(b->h_mpre, &b->mpre, s));
  PYTO_TRY(upload_prefix(b->h_vpre, &b->vpre, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  if (b->graph) {
    hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return PYTO_OK;
}

// f(cube, rows) with both as compile-time constants for the batch's shape.
template <typename F>
static void with_shape(const Batch* b, F&& f) {
  using std::integral_constant;
  auto rows = [&](auto cube) {
    switch (b->rows) {
      case 8: f(cube, integral_constant<int, 8>{}); break;
      case 32: f(cube, integral_constant<int, 32>{}); break;
      default: f(cube, integral_constant<int, 16>{}); break;
    }
  };
  if (b->cube)
    rows(integral_constant<bool, true>{});
  else
    rows(integral_constant<bool, false>{});
}

static void launch_bmatvec_inner(Batch* b, hipStream_t s) {
  const int nm = b->h_mpre[b->nsys];
  with_shape(b, [&](auto cube, auto ty) {
    hipLaunchKernelGGL((k_bmatvec_inner<decltype(cube)::value, decltype(ty)::value>), dim3(nm),
                       dim3(kTile * decltype(ty)::value), 0, s, b->sys, b->mpre, b->nsys,
                       (const int*)b->flags, b->part_mv);
  });
}

// mu = E^-1 W^T K (z + beta p) for every deflated system still iterating:
// special restriction, per-group gather, coarse GEMV.
static void launch_bcoarse(Batch* b, hipStream_t s, bool start) {
  const int nr = b->h_rpre[b->nsys], ng = b->h_gpre[b->nsys];
  const int nruns = (int)b->h_drun.size(), nc = nruns ? b->h_dpre[nruns] : 0;
  if (nr > 0) {
    if (b->cube)
      hipLaunchKernelGGL(k_brestrict<true>, dim3(nr), dim3(256), 0, s, b->sys, b->rpre, b->nsys,
                         (const double*)b->scal, (const int*)b->flags, start);
    else
      hipLaunchKernelGGL(k_brestrict<false>, dim3(nr), dim3(256), 0, s, b->sys, b->rpre,
                         b->nsys, (const double*)b->scal, (const int*)b->flags, start);
  }
  if (ng > 0)
    hipLaunchKernelGGL(k_bgather<false>, dim3((ng * 64 + 255) / 256), dim3(256), 0, s, b->sys,
                       b->gpre, b->nsys, (const int*)b->flags);
  if (nc > 0)
    hipLaunchKernelGGL(k_bgemv<false>, dim3((nc * 64 + 255) / 256), dim3(256), 0, s, b->sys,
                       b->drun, b->dpre, nruns, (const int*)b->flags);
}

// x0 = W E^-1 W^T b for every deflated system still refining, x0 = 0 for the
// others: full restriction, per-group gather, coarse GEMV, prolongation.
static void launch_bstart_x(Batch* b, hipStream_t s) {
  const int nn = b->h_npre[b->nsys], ng = b->h_gpre[b->nsys];
  const int nruns = (int)b->h_drun.size(), nc = nruns ? b->h_dpre[nruns] : 0;
  if (nn > 0)
    hipLaunchKernelGGL(k_brestrict_full, dim3(nn), dim3(256), 0, s, b->sys, b->npre, b->nsys,
                       (const int*)b->flags);
  if (ng > 0)
    hipLaunchKernelGGL(k_bgather<true>, dim3((ng * 64 + 255) / 256), dim3(256), 0, s, b->sys,
                       b->gpre, b->nsys, (const int*)b->flags);
  if (nc > 0)
    hipLaunchKernelGGL(k_bgemv<true>, dim3((nc * 64 + 255) / 256), dim3(256), 0, s, b->sys,
                       b->drun, b->dpre, nruns, (const int*)b->flags);
  hipLaunchKernelGGL(k_bstart_x, dim3(b->h_vpre[b->nsys]), dim3(256), 0, s, b->sys, b->vpre,
                     b->nsys, (const int*)b->flags);
}

static int batch_iteration(Batch* b, hipStream_t s) {
  const int nv = b->h_vpre[b->nsys];
  launch_bmatvec_inner(b, s);
  hipLaunchKernelGGL(k_bcg_update, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, b->mpre, b->nsys,
                     (const double*)b->scal, (const int*)b->flags, (const double*)b->part_mv,
                     b->part_v);
  hipLaunchKernelGGL(k_bscalars, dim3(b->nsys), dim3(256), 0, s, b->vpre, b->mpre, b->nsys,
                     b->scal, b->flags, (const double*)b->part_mv, (const double*)b->part_v);
  if (b->defl) {
    launch_bcoarse(b, s, false);
    hipLaunchKernelGGL(k_bdir_defl<false>, dim3(nv), dim3(256), b->dir_lds, s, b->sys, b->vpre, b->nsys,
                       (const double*)b->scal, (const int*)b->flags);
  } else {
    hipLaunchKernelGGL(k_bdir<false>, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, b->nsys,
                       (const double*)b->scal, (const int*)b->flags);
  }
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

static int batch_read(Batch* b, hipStream_t s) {
  // every system's scalars, then every system's flags: one block
  PYTO_CHECK_HIP(hipMemcpyAsync(b->host_scal, b->scal, (size_t)b->nsys * kCtlBytes,
                                hipMemcpyDeviceToHost, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  return PYTO_OK;
}

static bool all_flag(const Batch* b, int slot) {
  for (int i = 0; i < b->nsys; ++i)
    if (!b->host_flags[i * kNumFlags + slot]) return false;
  return true;
}

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_batch_destroy(pyto_batch* b);

int pyto_batch_create(pyto_batch** out, int nsys, pyto_handle* const* handles, void* stream) {
  PYTO_REQUIRE(out && handles && nsys > 0, PYTO_EINVAL, "null argument or nsys <= 0");
  pyto_batch* b = new pyto_batch();
  *out = nullptr;
  hipGetDevice(&b->device);
  b->nsys = nsys;
  b->handles.assign(handles, handles + nsys);
  auto fail = [&](int rc) { pyto_batch_destroy(b); return rc; };
  hipStream_t s = as_stream(stream);
  for (int i = 0; i < nsys; ++i) {
    const Handle* h = handles[i];
    if (!h) {
      set_error("handle %d is null", i);
      return fail(PYTO_EINVAL);
    }
    if (h->ncomp != 3) {
      // the batched operator and vector kernels assume 3 DOFs per node
      set_error("a batch holds structural handles only (system %d has %d value(s) per node)",
                i, h->ncomp);
      return fail(PYTO_EINVAL);
    }
    const bool cube = h->g.hx == h->g.hy && h->g.hy == h->g.hz;
    if (i == 0) b->cube = cube;
    if (cube != b->cube) {
      set_error("a batch holds either cube or brick elements, not both (system %d)", i);
      return fail(PYTO_EINVAL);
    }
    BatchSys S{};
    S.g = h->g;
    S.mc32 = h->mc32;
    S.mc64 = h->mc64;
    S.E32 = h->Ee32;
    S.E64 = h->Ee64;
    S.invd32 = h->invd32;
    S.invd64 = h->invd64;
    S.mask = h->fixmask;
    S.tileflags = h->tileflags;
    S.ndof = 3 * h->g.nnodes;
    const size_t n32 = (size_t)S.ndof * 4, n64 = (size_t)S.ndof * 8;
    float* f32[5];
    for (int k = 0; k < 5; ++k) {
      void* p = nullptr;
      if (hipMalloc(&p, n32) != hipSuccess) {
        set_error("device allocation of %zu bytes failed", n32);
        return fail(PYTO_ENOMEM);
      }
      b->bufs.push_back(p);
      hipMemsetAsync(p, 0, n32, s);
      f32[k] = (float*)p;
    }
    S.z = f32[0];
    S.esum32 = h->esum32;
    for (int c = 0; c < 3; ++c) S.kd[c] = (float)h->kediag[c];
    S.p = f32[1];
    S.q = f32[2];
    S.x = f32[3];
    S.b = f32[4];
    double* f64[2];
    for (int k = 0; k < 2; ++k) {
      void* p = nullptr;
      if (hipMalloc(&p, n64) != hipSuccess) {
        set_error("device allocation of %zu bytes failed", n64);
        return fail(PYTO_ENOMEM);
      }
      b->bufs.push_back(p);
      hipMemsetAsync(p, 0, n64, s);
      f64[k] = (double*)p;
    }
    S.r64 = f64[0];
    S.q64 = f64[1];
    if (h->ngroups > 0) {
      if (!h->pslot || !h->Einv || h->nsblocks <= 0) {
        set_error("system %d: the handle's deflation groups have no coarse space", i);
        return fail(PYTO_ESTATE);
      }
      S.ngroups = h->ngroups;
      S.nc = h->nc;
      S.ld = (h->nc + 63) / 64 * 64;
      S.node_group = h->node_group;
      S.gmask = h->gmask;
      S.centers = h->centers;
      S.pwave = h->pwave;
      S.pslot = h->pslot;
      S.pnodes = h->pnodes;
      S.pcen = h->pcen;
      S.pE32 = h->pE32;
      S.gptr = h->gptr;
      S.gidx = h->gidx;
      S.Einv = h->Einv;
      if (!h->gnodes || h->nnchunks <= 0) {
        set_error("system %d: the handle's deflation groups have no node blocks", i);
        return fail(PYTO_ESTATE);
      }
      S.gnodes = h->gnodes;
      S.nchunk = h->nchunk;
      S.nchunk_grp = h->nchunk_grp;
      S.nptr = h->nptr;
      S.nidx = h->nidx;
      const size_t sizes[4] = {(size_t)h->nchunks * 48, (size_t)S.ld, (size_t)S.ld,
                               (size_t)h->nnchunks * 48};
      double* cb[4];
      for (int k = 0; k < 4; ++k) {
        void* p = nullptr;
        if (hipMalloc(&p, sizes[k] * sizeof(double)) != hipSuccess) {
          set_error("device allocation of %zu bytes failed", sizes[k] * sizeof(double));
          return fail(PYTO_ENOMEM);
        }
        b->bufs.push_back(p);
        hipMemsetAsync(p, 0, sizes[k] * sizeof(double), s);
        cb[k] = (double*)p;
      }
      S.cpart = cb[0];
      S.ct = cb[1];
      S.cmu = cb[2];
      S.npart = cb[3];
      b->defl = true;
    }
    b->hs.push_back(S);
    b->groups_gen.push_back(h->groups_gen);
  }
  // prefix sums of the deflated systems' special blocks, groups, coarse rows
  b->h_rpre.assign(nsys + 1, 0);
  b->h_gpre.assign(nsys + 1, 0);
  b->h_npre.assign(nsys + 1, 0);
  for (int i = 0; i < nsys; ++i) {
    const bool d = b->hs[i].ngroups > 0;
    b->h_rpre[i + 1] = b->h_rpre[i] + (d ? handles[i]->nsblocks : 0);
    b->h_gpre[i + 1] = b->h_gpre[i] + b->hs[i].ngroups;
    b->h_npre[i + 1] = b->h_npre[i] + (d ? handles[i]->nnchunks : 0);
    if (!d) continue;
    // runs of consecutive deflated systems on one handle share E^-1
    if (!b->h_drun.empty() && handles[b->h_drun.back().x] == handles[i] &&
        b->h_drun.back().x + b->h_drun.back().y == i && b->h_drun.back().y < kBatchRhs)
      b->h_drun.back().y++;
    else
      b->h_drun.push_back(make_int2(i, 1));
  }
  for (int i = 0; i < nsys; ++i)
    if (b->hs[i].ngroups > 0 && b->hs[i].ngroups <= kBLdsGroups)
      b->dir_lds = std::max(b->dir_lds, (size_t)b->hs[i].ngroups * 48);
  b->h_dpre.assign(b->h_drun.size() + 1, 0);
  for (size_t r = 0; r < b->h_drun.size(); ++r)
    b->h_dpre[r + 1] = b->h_dpre[r] + b->hs[b->h_drun[r].x].nc;
  const size_t nruns = std::max<size_t>(b->h_drun.size(), 1);
  if (hipMalloc((void**)&b->drun, nruns * sizeof(int2)) != hipSuccess) {
    set_error("batch allocation failed");
    return fail(PYTO_ENOMEM);
  }
  if (!b->h_drun.empty())
    hipMemcpyAsync(b->drun, b->h_drun.data(), b->h_drun.size() * sizeof(int2),
                   hipMemcpyHostToDevice, s);
  {
    int rc = upload_prefix(b->h_rpre, &b->rpre, s);
    if (!rc) rc = upload_prefix(b->h_gpre, &b->gpre, s);
    if (!rc) rc = upload_prefix(b->h_npre, &b->npre, s);
    if (!rc) rc = upload_prefix(b->h_dpre, &b->dpre, s);
    if (rc) return fail(rc);
  }
  if (hipMalloc((void**)&b->sys, nsys * sizeof(BatchSys)) != hipSuccess ||
      hipMalloc((void**)&b->scal, (size_t)nsys * kCtlBytes) != hipSuccess ||
      hipHostMalloc((void**)&b->host_scal, (size_t)nsys * kCtlBytes, hipHostMallocDefault) !=
          hipSuccess) {
    set_error("batch allocation failed");
    return fail(PYTO_ENOMEM);
  }
  // the flags of all systems follow their scalars in the same allocations
  b->flags = reinterpret_cast<int*>(b->scal + (size_t)nsys * kNumScalars);
  b->host_flags = reinterpret_cast<int*>(b->host_scal + (size_t)nsys * kNumScalars);
  hipMemsetAsync(b->scal, 0, (size_t)nsys * kNumScalars * sizeof(double), s);
  hipMemsetAsync(b->flags, 0, (size_t)nsys * kNumFlags * sizeof(int), s);
  int rc = batch_geometry(b, 16, kTile, s);
  if (rc) return fail(rc);
  int64_t nvtot = b->h_vpre[nsys];
  if (hipMalloc((void**)&b->part_v, (size_t)nvtot * sizeof(double)) != hipSuccess) {
    set_error("batch allocation failed");
    return fail(PYTO_ENOMEM);
  }
  if (hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_a, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_b, hipEventDisableTiming) != hipSuccess) {
    set_error("stream/event creation failed");
    return fail(PYTO_EHIP);
  }
  *out = b;
  return PYTO_OK;
}

int pyto_batch_destroy(pyto_batch* b) {
  if (!b) return PYTO_OK;
  for (void* p : b->bufs) hipFree(p);
  void* ptrs[] = {b->sys, b->mpre, b->vpre, b->scal /* and flags */, b->part_mv, b->part_v,
                  b->rpre, b->gpre, b->npre, b->dpre, b->drun};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (b->graph) hipGraphExecDestroy(b->graph);
  if (b->ev_a) hipEventDestroy(b->ev_a);
  if (b->ev_b) hipEventDestroy(b->ev_b);
  if (b->gstream) hipStreamDestroy(b->gstream);
  if (b->host_scal) hipHostFree(b->host_scal);   // holds host_flags too
  delete b;
  return PYTO_OK;
}

int pyto_batch_set_zchunk(pyto_batch* b, int zchunk, void* stream) {
  PYTO_REQUIRE(b && zchunk >= 1, PYTO_EINVAL, "bad arguments");
  return batch_geometry(b, zchunk, b->rows, as_stream(stream));
}

int pyto_batch_set_tile_rows(pyto_batch* b, int rows, void* stream) {
  PYTO_REQUIRE(b, PYTO_EINVAL, "null batch");
  PYTO_REQUIRE(rows == 8 || rows == 16 || rows == 32, PYTO_EINVAL,
               "operator tile rows must be 8, 16 or 32 (got %d)", rows);
  return batch_geometry(b, b->zchunk, rows, as_stream(stream));
}

int pyto_batch_get_tile_rows(const pyto_batch* b) { return b ? b->rows : PYTO_EINVAL; }

int pyto_batch_info(pyto_batch* b, int* zchunk, int* matvec_blocks, int* vector_blocks) {
  PYTO_REQUIRE(b, PYTO_EINVAL, "null batch");
  if (zchunk) *zchunk = b->zchunk;
  if (matvec_blocks) *matvec_blocks = b->h_mpre[b->nsys];
  if (vector_blocks) *vector_blocks = b->h_vpre[b->nsys];
  return PYTO_OK;
}

int pyto_batch_time_matvec(pyto_batch* b, int reps, float* ms_per_launch, void* stream) {
  PYTO_REQUIRE(b && reps > 0 && ms_per_launch, PYTO_EINVAL, "bad arguments");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_bset_flag, dim3(1), dim3(256), 0, s, b->flags, b->nsys, (int)F_DONE, 0);
  launch_bmatvec_inner(b, s);
  hipEvent_t e0, e1;
  PYTO_CHECK_HIP(hipEventCreate(&e0));
  PYTO_CHECK_HIP(hipEventCreate(&e1));
  hipEventRecord(e0, s);
  for (int k = 0; k < reps; ++k) launch_bmatvec_inner(b, s);
  hipEventRecord(e1, s);
  PYTO_CHECK_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms_per_launch = ms / reps;
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_measure_grid_barrier(int blocks_per_cu, int rounds, float* us_per_barrier,
                              void* stream) {
  PYTO_REQUIRE(blocks_per_cu >= 1 && blocks_per_cu <= 4 && rounds >= 1 && rounds <= 100000 &&
                   us_per_barrier,
               PYTO_EINVAL, "bad arguments");
  hipStream_t s = as_stream(stream);
  int dev = 0, coop = 0, per_cu = 0;
  hipDeviceProp_t prop;
  PYTO_CHECK_HIP(hipGetDevice(&dev));
  PYTO_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
  PYTO_CHECK_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  PYTO_REQUIRE(coop, PYTO_EINVAL, "the device has no cooperative launch");
  PYTO_CHECK_HIP(
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_grid_barrier_probe, 256, 0));
  PYTO_REQUIRE(per_cu >= blocks_per_cu, PYTO_EINVAL,
               "%d workgroups per CU do not fit (at most %d)", blocks_per_cu, per_cu);
  const int nb = prop.multiProcessorCount * blocks_per_cu;
  unsigned* counter = nullptr;
  PYTO_CHECK_HIP(hipMalloc((void**)&counter, 2 * sizeof(unsigned)));
  int* timed_out = reinterpret_cast<int*>(counter + 1);
  const unsigned spin_limit = 1u << 22;
  float ms[2] = {0.f, 0.f};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  // warm-up, then the same launch with 0 rounds (the launch itself) and
  // with `rounds` barriers
  const int plan[3] = {8, 0, rounds};
  for (int k = 0; k < 3 && err == hipSuccess; ++k) {
    err = hipMemsetAsync(counter, 0, 2 * sizeof(unsigned), s);
    int rr = plan[k];
    void* args[] = {(void*)&counter, (void*)&rr, (void*)&spin_limit, (void*)&timed_out};
    if (err == hipSuccess) err = hipEventRecord(e0, s);
    if (err == hipSuccess)
      err = hipLaunchCooperativeKernel((const void*)k_grid_barrier_probe, dim3(nb), dim3(256),
                                       args, 0, s);
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess && k > 0) err = hipEventElapsedTime(&ms[k - 1], e0, e1);
  }
  int flag = 0;
  if (err == hipSuccess) err = hipMemcpy(&flag, timed_out, sizeof(int), hipMemcpyDeviceToHost);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(counter);
  PYTO_CHECK_HIP(err);
  PYTO_REQUIRE(!flag, PYTO_ENOCONV, "grid barrier probe: a wait exceeded its poll limit");
  *us_per_barrier = (ms[1] - ms[0]) * 1e3f / (float)rounds;
  return PYTO_OK;
}

int pyto_batch_solve(pyto_batch* b, const double* const* f, double* const* u, double rtol,
                     double inner_rtol, int max_iters, int max_outer, pyto_solve_stats* stats,
                     void* stream) {
  PYTO_REQUIRE(b && f && u, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(rtol > 0 && inner_rtol > 0 && max_iters > 0 && max_outer > 0, PYTO_EINVAL,
               "rtol, inner_rtol, max_iters and max_outer must be positive");
  hipStream_t s = as_stream(stream);
  for (int i = 0; i < b->nsys; ++i) {
    PYTO_REQUIRE(f[i] && u[i], PYTO_EINVAL, "null load or solution of system %d", i);
    PYTO_REQUIRE(b->handles[i]->density_set, PYTO_ESTATE, "system %d: density not set", i);
    const Handle* h = b->handles[i];
    PYTO_REQUIRE(h->groups_gen == b->groups_gen[i], PYTO_ESTATE,
                 "system %d: its handle's deflation groups changed after the batch was "
                 "created; create the batch again", i);
    PYTO_REQUIRE(b->hs[i].ngroups == 0 || h->coarse_ready, PYTO_ESTATE,
                 "system %d: the coarse matrix of its groups is not factorised", i);
    b->hs[i].f = f[i];
    b->hs[i].u = u[i];
  }
  PYTO_CHECK_HIP(hipMemcpyAsync(b->sys, b->hs.data(), b->hs.size() * sizeof(BatchSys),
                                hipMemcpyHostToDevice, s));
  const int nsys = b->nsys;
  const int nv = b->h_vpre[nsys];
  const int nm = b->h_mpre[nsys];
  std::vector<int> inner_iters(nsys, 0), outer_iters(nsys, 0);
  hipLaunchKernelGGL(k_bset_flag, dim3(1), dim3(256), 0, s, b->flags, nsys, (int)F_ODONE, 0);
  hipLaunchKernelGGL(k_bzero_u, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys);
  int launched_total = 0;
  for (int outer = 0; outer < max_outer; ++outer) {
    const bool first = outer == 0;
    if (!first) {
      with_shape(b, [&](auto cube, auto ty) {
        hipLaunchKernelGGL((k_bmatvec_outer<decltype(cube)::value, decltype(ty)::value>),
                           dim3(nm), dim3(kTile * decltype(ty)::value), 0, s, b->sys, b->mpre,
                           nsys, (const int*)b->flags);
      });
    }
    hipLaunchKernelGGL(k_bresid64, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys,
                       (const int*)b->flags, first, b->part_v);
    hipLaunchKernelGGL(k_bouter, dim3(nsys), dim3(256), 0, s, b->vpre, nsys, b->scal, b->flags,
                       (const double*)b->part_v, first, rtol);
    PYTO_CHECK_HIP(hipGetLastError());
    PYTO_TRY(batch_read(b, s));
    if (all_flag(b, F_ODONE) || launched_total >= max_iters) break;
    for (int i = 0; i < nsys; ++i)
      if (!b->host_flags[i * kNumFlags + F_ODONE]) outer_iters[i]++;
    hipLaunchKernelGGL(k_bscale, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys,
                       (const double*)b->scal, (const int*)b->flags);
    if (b->defl) {
      // deflated start: x0 = W E^-1 W^T b per deflated system (x0 = 0 for the
      // others), r = b - K x0, then the projected first direction
      launch_bstart_x(b, s);
      with_shape(b, [&](auto cube, auto ty) {
        hipLaunchKernelGGL((k_bmatvec_start<decltype(cube)::value, decltype(ty)::value>),
                           dim3(nm), dim3(kTile * decltype(ty)::value), 0, s, b->sys, b->mpre,
                           nsys, (const int*)b->flags);
      });
      hipLaunchKernelGGL(k_binit_q, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys,
                         (const int*)b->flags, b->part_v);
      hipLaunchKernelGGL(k_bsetup, dim3(nsys), dim3(256), 0, s, b->vpre, nsys, b->scal,
                         b->flags, (const double*)b->part_v, rtol, inner_rtol);
      launch_bcoarse(b, s, true);
      hipLaunchKernelGGL(k_bdir_defl<true>, dim3(nv), dim3(256), b->dir_lds, s, b->sys, b->vpre, nsys,
                         (const double*)b->scal, (const int*)b->flags);
    } else {
      hipLaunchKernelGGL(k_binit, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys,
                         (const int*)b->flags, b->part_v);
      hipLaunchKernelGGL(k_bsetup, dim3(nsys), dim3(256), 0, s, b->vpre, nsys, b->scal,
                         b->flags, (const double*)b->part_v, rtol, inner_rtol);
      hipLaunchKernelGGL(k_bdir<true>, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys,
                         (const double*)b->scal, (const int*)b->flags);
    }
    PYTO_CHECK_HIP(hipGetLastError());
    // inner iterations: replays of the captured batch of kGraphBatch
    if (!b->graph) {
      hipGraph_t g;
      PYTO_CHECK_HIP(hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal));
      int rc = PYTO_OK;
      for (int k = 0; k < kBatchGraphIters && rc == PYTO_OK; ++k) rc = batch_iteration(b, b->gstream);
      hipError_t e = hipStreamEndCapture(b->gstream, &g);
      if (rc != PYTO_OK) return rc;
      PYTO_CHECK_HIP(e);
      e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      PYTO_CHECK_HIP(e);
    }
    PYTO_CHECK_HIP(hipEventRecord(b->ev_a, s));
    PYTO_CHECK_HIP(hipStreamWaitEvent(b->gstream, b->ev_a, 0));
    int launched = 0;
    while (launched_total + launched < max_iters) {
      PYTO_CHECK_HIP(hipGraphLaunch(b->graph, b->gstream));
      launched += kBatchGraphIters;
      PYTO_TRY(batch_read(b, b->gstream));
      if (all_flag(b, F_DONE)) break;
    }
    launched_total += launched;
    PYTO_CHECK_HIP(hipEventRecord(b->ev_b, b->gstream));
    PYTO_CHECK_HIP(hipStreamWaitEvent(s, b->ev_b, 0));
    for (int i = 0; i < nsys; ++i) inner_iters[i] += b->host_flags[i * kNumFlags + F_ITERS];
    hipLaunchKernelGGL(k_bfinal, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys,
                       (const double*)b->scal, (const int*)b->flags);
    hipLaunchKernelGGL(k_baxpy, dim3(nv), dim3(256), 0, s, b->sys, b->vpre, nsys,
                       (const double*)b->scal, (const int*)b->flags);
    PYTO_CHECK_HIP(hipGetLastError());
  }
  PYTO_TRY(batch_read(b, s));
  bool all = true;
  for (int i = 0; i < nsys; ++i) {
    const bool conv = b->host_flags[i * kNumFlags + F_ODONE] != 0;
    all = all && conv;
    if (stats) {
      pyto_solve_stats& st = stats[i];
      st.iterations = inner_iters[i];
      st.outer_iterations = outer_iters[i];
      st.converged = conv;
      st.matvecs64 = outer_iters[i] + 1;
      st.matvecs32 = inner_iters[i];
      st.rel = b->host_scal[i * kNumScalars + S_TMP2];
      st.rel_true = -1.0;
      st.setup_ms = 0.0;
    }
  }
  if (!all) {
    set_error("batch: not every system met rtol within max_iters / max_outer");
    return PYTO_ENOCONV;
  }
  return PYTO_OK;
}

}  // extern "C"
