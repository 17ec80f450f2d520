// Deflated Jacobi-preconditioned CG on the device (src/deflation.py:504-584)
// and its FP32-inner / FP64-outer iterative refinement.
//
// One iteration is six launches on one stream; every launch first reads the
// device flag F_DONE and returns when it is set, so the host can enqueue a
// batch of iterations and read the flag back once per batch while the
// stopping rule sqrt(rz_new/rz0) <= rtol is applied on the device at exactly
// the reference's iteration:
//   1. q = K p, p.q block partials
//   2. alpha = rz/pKp; z -= alpha M^-1 q (z = M^-1 r is stored); r^T M^-1 r
//      partials; alpha published for step 6
//   3. beta = rz_new/rz; partials of W^T K (z + beta p) (non-rigid elements)
//   4. per-group gather of those partials; scalars and convergence test
//   5. mu = E^-1 t (FP64 coarse inverse as a dense GEMV)
//   6. x += alpha p; p = z + beta p - W mu
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "element.hpp"

namespace pyto {

template <typename T>
int launch_matvec(const Handle* h, const T* u, T* y, const T* Ee, const ModalConsts<T>& mc,
                  double* dot_out, const int* done, double* zero_vec, int zero_n,
                  hipStream_t s);
template <typename T>
int launch_restrict_full(Handle* h, const T* v, double* t, hipStream_t s);
template <typename T>
int launch_restrict_special(Handle* h, const T* z, double* t, const int* done, hipStream_t s);
template <typename T>
int launch_restrict_special_parts(Handle* h, const T* z, const T* p, const double* part_rz,
                                  int nrz, const double* rz_old, const int* done,
                                  hipStream_t s);
int launch_coarse_gather(Handle* h, double* t, const int* done, hipStream_t s);
int launch_gather_scalars(Handle* h, double* t, const double* part_pq, int npq,
                          const double* part_rz, int nrz, hipStream_t s);
template <typename T, bool UPDATE>
int launch_prolong(Handle* h, T* out, const T* z, const double* mu, const double* beta,
                   bool first, const int* done, hipStream_t s, double* zero_vec, int zero_n);
int coarse_apply(Handle* h, const double* t1, const double* t2, const double* beta,
                 double* mu, const int* done, hipStream_t s, bool refine);
template <typename T>
int matvec_blocks(const Handle* h);
template <typename T>
int launch_dir_update(Handle* h, const DirIO<T>& io, bool first, const int* done, hipStream_t s);

constexpr int kVecBlocks = 2048;

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
  static constexpr int width = 4;
};
template <>
struct Vec4<double> {
  using type = double2;
  static constexpr int width = 2;
};

// z-form update: alpha = rz/pKp with pKp summed from the operator's block
// partials (every block sums them in the same order); z -= alpha invd o q
// (formed in FP64, stored in T); per-block partials of
// r^T M^-1 r = sum z_i^2 / invd_i over free DOFs (invd_i > 0) to part_rz.
// Block 0 publishes alpha in scal[S_ALPHA_X] for the x update, which the
// direction update applies (or, for the iteration that stops the solve, the
// host after the loop).  With pKp <= 0 or NaN the block leaves z untouched
// and alpha is published as 0 (the scalar step flags the breakdown).  Vector
// loads of 16 bytes; block 0 takes the tail.  NODE (3 values per node):
// invd is not read; each DOF's inverse comes from its node's sum of E_e and
// fixed bits.  FP32: jacobi_inv32 (the bits of invd32), 5 bytes per node
// instead of 12, and r^T M^-1 r sums z^2 d.  FP64: the operations that form
// invd64 (the same bits, and the same rz terms), 9 bytes per node instead of
// 24.
template <typename T, bool NODE>
__global__ __launch_bounds__(256) void k_cg_update_z(
    T* __restrict__ z, const T* __restrict__ q, const T* __restrict__ invd, JacobiNodes jn,
    int64_t n, double* __restrict__ scal, const int* __restrict__ flags,
    const double* __restrict__ part_pq, int npq, double* __restrict__ part_rz,
    double* __restrict__ al_slot) {

  if (flags[F_DONE]) return;
  __shared__ double red[5];
  const double pkp = block_sum_partials(part_pq, npq, red);
  double acc = 0.0;
  const bool ok = pkp > 0.0;   // false for NaN
  const double alpha = ok ? scal[S_RZ] / pkp : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_ALPHA_X] = alpha;
    if (al_slot) *al_slot = alpha;   // the direction ring's alpha of this iteration
  }
  if (ok) {
    auto lane = [&](T& zi, const T qi, const T di) {
      zi = (T)((double)zi - alpha * ((double)di * (double)qi));
      if (di > T(0)) acc += (double)zi * (double)(zi / di);
    };
    // node form: the DOF's inverse and d from its node's sum and fixed bits
    auto nlane = [&](T& zi, const T qi, float s, uint8_t m, int c) {
      const float kd = c == 0 ? jn.kd0 : (c == 1 ? jn.kd1 : jn.kd2);
      float d;
      const float di = jacobi_inv32(kd, s, (m >> c) & 1, d);
      zi = (T)((double)zi - alpha * ((double)di * (double)qi));
      acc += (double)zi * (double)zi * (double)d;
    };
    // FP64 node form: the inverse with invd64's operations (k_inv_diag), so
    // the same bits as reading it, and the same rz terms as `lane`
    auto nlane64 = [&](T& zi, const T qi, double s, uint8_t m, int c) {
      double d = jn.kd64[c] * s;
      if (fabs(d) < 1e-12) d = 1.0;
      const double di = ((m >> c) & 1) ? 0.0 : 1.0 / d;
      zi = (T)((double)zi - alpha * (di * (double)qi));
      if (di > 0.0) acc += (double)zi * ((double)zi / di);
    };
    using V = typename Vec4<T>::type;
    constexpr int W = Vec4<T>::width;
    const int64_t nv = n / W;
    V* zv = reinterpret_cast<V*>(z);
    const V* qv = reinterpret_cast<const V*>(q);
    const V* dv = reinterpret_cast<const V*>(invd);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    // node form: DOFs 4i..4i+3 lie on nodes n0 and n0 + 1 (components c0..
    // and 0..); n0 and c0 are formed once and advanced by the loop's
    // constant step (4 stride DOFs = kn nodes + kc components)
    int64_t n0 = (W * i) / 3;
    int c0 = (int)(W * i - 3 * n0);
    const int64_t kn = (W * stride) / 3;
    const int kc = (int)(W * stride - 3 * kn);
    for (; i < nv; i += stride) {
      V zz = zv[i];
      const V qq = qv[i];
      if constexpr (NODE && W == 4) {
        const float s0 = jn.esum[n0], s1 = jn.esum[n0 + 1];
        const uint8_t m0 = jn.mask[n0], m1 = jn.mask[n0 + 1];
        T* zl[4] = {&zz.x, &zz.y, &zz.z, &zz.w};
        const T ql[4] = {qq.x, qq.y, qq.z, qq.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int c = c0 + j;
          const bool second = c >= 3;
          nlane(*zl[j], ql[j], second ? s1 : s0, second ? m1 : m0, second ? c - 3 : c);
        }
        c0 += kc;
        const int carry = c0 >= 3 ? 1 : 0;
        c0 -= 3 * carry;
        n0 += kn + carry;
      } else if constexpr (NODE && W == 2) {
        // DOFs 2i, 2i+1: component c0 of node n0, then c0 + 1 (node n0 + 1
        // when c0 is 2)
        const double s0 = jn.esum64[n0];
        const uint8_t m0 = jn.mask[n0];
        const bool two = c0 == 2;
        const double s1 = two ? jn.esum64[n0 + 1] : s0;
        const uint8_t m1 = two ? jn.mask[n0 + 1] : m0;
        nlane64(zz.x, qq.x, s0, m0, c0);
        nlane64(zz.y, qq.y, s1, m1, two ? 0 : c0 + 1);
        c0 += kc;
        const int carry = c0 >= 3 ? 1 : 0;
        c0 -= 3 * carry;
        n0 += kn + carry;
      } else {
        const V dd = dv[i];
        lane(zz.x, qq.x, dd.x);
        lane(zz.y, qq.y, dd.y);
        if constexpr (W == 4) {
          lane(zz.z, qq.z, dd.z);
          lane(zz.w, qq.w, dd.w);
        }
      }
      zv[i] = zz;
    }
    if (blockIdx.x == 0)
      for (int64_t i = nv * W + threadIdx.x; i < n; i += blockDim.x) {
        if constexpr (NODE && sizeof(T) == 4)
          nlane(z[i], q[i], jn.esum[i / 3], jn.mask[i / 3], (int)(i % 3));
        else if constexpr (NODE)
          nlane64(z[i], q[i], jn.esum64[i / 3], jn.mask[i / 3], (int)(i % 3));
        else
          lane(z[i], q[i], invd[i]);
      }
  }
  __shared__ double red4[4];
  const double bs = block_reduce_sum256(acc, red4);
  if (threadIdx.x == 0) part_rz[blockIdx.x] = bs;
}

// x += sum over j < count of al[a] ring[r * ld + i], a = (a0 + j) mod
// kDirSlots, r = (r0 + j) mod kDirSlots (r0 < 0: slot 0 throughout), summed
// in FP64: the updates still pending when a DPCG run ends (the last, partial
// window when x is deferred; the stopping iteration's alpha p otherwise).
template <typename T>
__global__ __launch_bounds__(256) void k_flush_x(T* __restrict__ x, const T* __restrict__ ring,
                                                int64_t ld, const double* __restrict__ al,
                                                int a0, int r0, int count, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = (double)x[i];
    for (int j = 0; j < count; ++j) {
      const int r = r0 < 0 ? 0 : (r0 + j) % kDirSlots;
      acc += al[(a0 + j) % kDirSlots] * (double)ring[r * ld + i];
    }
    x[i] = (T)acc;
  }
}

// z = invd o (f - q) (in T); per-block partials of r^T (invd o r).
template <typename T>
__global__ __launch_bounds__(256) void k_residual_z(const T* __restrict__ f,
                                                    const T* __restrict__ q, T* __restrict__ z,
                                                    const T* __restrict__ invd, int64_t n,
                                                    double* __restrict__ part) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = (double)f[i] - (double)q[i];
    const double di = (double)invd[i];
    z[i] = (T)(di * ri);
    acc += ri * di * ri;
  }
  const double bs = block_reduce_sum256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// One block: finishes the iteration's scalars from the partials: alpha (as
// in k_cg_update), rz_new, the stop test sqrt(rz_new/rz0) <= rtol
// (src/deflation.py:567), beta = rz_new/rz, the iteration count and the
// breakdown flag (p^T K p <= 0 or NaN) and the stagnation test.
__global__ __launch_bounds__(256) void k_scalars(double* __restrict__ scal, int* __restrict__ flags,
                                                 const double* __restrict__ part_pq, int npq,
                                                 const double* __restrict__ part_rz, int nrz) {
  if (flags[F_DONE]) return;
  __shared__ double red[10];
  double pkp, rz_new;
  block_sum_partials2(part_pq, npq, part_rz, nrz, red, pkp, rz_new);
  if (threadIdx.x == 0) pcg_scalar_step(scal, flags, pkp, rz_new, scal[S_FLOOR]);
}

// One block: scal[slot_a] (and slot_b) = sum of n partials.
__global__ __launch_bounds__(256) void k_sum_to(double* __restrict__ scal,
                                                const double* __restrict__ part, int n,
                                                int slot_a, int slot_b) {
  __shared__ double red[5];
  const double v = block_sum_partials(part, n, red);
  if (threadIdx.x == 0) {
    scal[slot_a] = v;
    if (slot_b >= 0) scal[slot_b] = v;
  }
}

// r = f - q (in T); per-block partials of r^T (invd o r) (k_sum_to finishes).
template <typename T, typename TF>
__global__ __launch_bounds__(256) void k_residual(const TF* __restrict__ f, const T* __restrict__ q,
                                                  T* __restrict__ r, const T* __restrict__ invd,
                                                  int64_t n, double* __restrict__ part) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = (T)((double)f[i] - (double)q[i]);
    r[i] = ri;
    acc += (double)ri * (double)invd[i] * (double)ri;
  }
  const double bs = block_reduce_sum256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// out = (T)(in * (1/scal[slot]))  (scal[slot] > 0)
template <typename T>
__global__ void k_scale_convert(const double* __restrict__ in, T* __restrict__ out, int64_t n,
                                const double* __restrict__ scal, int slot) {
  const double inv = 1.0 / sqrt(scal[slot]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (T)(in[i] * inv);
}

// u += sqrt(scal[slot]) * d
__global__ void k_axpy_scaled(double* __restrict__ u, const float* __restrict__ d, int64_t n,
                              const double* __restrict__ scal, int slot) {
  const double s = sqrt(scal[slot]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    u[i] += s * (double)d[i];
}

// scal[slot] = sum over free DOFs (free_ind != 0) of (a_i - b_i)^2 (b may be null)
__global__ __launch_bounds__(256) void k_norm2_diff(const double* __restrict__ a,
                                                    const double* __restrict__ b,
                                                    const double* __restrict__ free_ind,
                                                    int64_t n, double* __restrict__ part) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double d = a[i] - (b ? b[i] : 0.0);
    if (free_ind[i] != 0.0) acc += d * d;
  }
  const double bs = block_reduce_sum256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

__global__ void k_resume_scalars(double* scal) {
  scal[S_BETA] = scal[S_RZ] / scal[S_RZPREV];
  scal[S_ALPHA_X] = 0.0;
}

template <typename T>
struct Vecs {
  T *r, *q;
  const T* invd;
  const T* Ee;
  ModalConsts<T> mc;
  T* ring;   // kDirSlots directions: slot k mod kDirSlots holds p_k
};

static Vecs<double> vecs64(Handle* h) {
  return {h->r64, h->q64, h->invd64, h->Ee64, h->mc64, h->pring64};
}
static Vecs<float> vecs32(Handle* h) {
  return {h->r32, h->q32, h->invd32, h->Ee32, h->mc32, h->pring32};
}

template <typename T>
Vecs<T> get_vecs(Handle* h);
template <>
Vecs<double> get_vecs<double>(Handle* h) { return vecs64(h); }
template <>
Vecs<float> get_vecs<float>(Handle* h) { return vecs32(h); }

struct DpcgResult {
  int iterations = 0;
  bool converged = false;
  bool breakdown = false;
  bool stagnated = false;
  double rel = 0.0;
  int matvecs = 0;
};

static int read_back(Handle* h, hipStream_t s) {
  // scalars and flags are one block (kCtlBytes)
  PYTO_CHECK_HIP(hipMemcpyAsync(h->host_scal, h->scal, kCtlBytes, hipMemcpyDeviceToHost, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  return PYTO_OK;
}

// Iteration batches.  The host enqueues the launches of a batch of
// iterations directly on the caller's stream, then a copy of the control
// block into pinned memory and an event, and keeps the next batch queued
// while it waits for the previous one's copy, so the GPU does not idle at
// batch boundaries.  Far from the stop a batch is kFarBatch iterations; once
// the convergence rate seen in the read-backs puts the stop within
// kNearWindow iterations, batches shrink to kNearBatch, so at most about two
// such batches run after the stopping iteration (their launches return at
// once on the device flag, a few microseconds each).
constexpr int kFarBatch = 16;
constexpr int kNearBatch = 2;
constexpr int kNearWindow = 24;
constexpr size_t kCtlStride = 512;   // bytes between the two read-back slots
static_assert(kCtlBytes <= kCtlStride, "control block fits a read-back slot");

// The handle's pinned read-back slots and events, created on first use.
static int ensure_control(Handle* h) {
  if (!h->host_ctl) {
    PYTO_CHECK_HIP(hipHostMalloc((void**)&h->host_ctl, 2 * kCtlStride, hipHostMallocDefault));
    for (hipEvent_t& e : h->ev_ctl)
      PYTO_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  return PYTO_OK;
}
static double* ctl_slot(Handle* h, int slot) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(h->host_ctl) + slot * kCtlStride);
}

// largest ratio of true to recurrence residual norm at which a mixed
// refinement step continues the previous inner run's search direction
constexpr double kResumeDrift = 1.25;

// First launch of a run: its tolerance, the stagnation state and the flags;
// a run whose r0^T M r0 is 0 is done before its first iteration.
__global__ void k_start_run(double* scal, int* flags, double rtol, double floor) {
  const bool zero = scal[S_RZ0] == 0.0;
  scal[S_RTOL] = rtol;
  scal[S_RELMIN] = 1e300;
  scal[S_FLOOR] = floor;
  if (zero) scal[S_REL] = 0.0;
  flags[F_DONE] = zero ? 1 : 0;
  flags[F_ITERS] = 0;
  flags[F_BREAKDOWN] = 0;
  flags[F_STALL] = 0;
}

// Deflated PCG on K x = b (b and x full-length in precision T; fixed DOFs of
// b ignored, of x zero).  The deflated start x0 = W E^-1 W^T b and the first
// direction p0 = z0 - W E^-1 W^T K z0 follow src/deflation.py:532-551.
//
// Direction update.  The reference forms p = z + beta p - W E^-1 W^T K z,
// which relies on W^T K p_old = 0 holding exactly from the previous step;
// rounding (an FP32 operator, the coarse solve) then accumulates and the
// iteration stalls or diverges.  Here the whole new direction is projected,
// mu = E^-1 W^T K (z + beta p_old), so W^T K p_new = 0 is re-established
// every step.  In exact arithmetic the two forms coincide.
//
// The preconditioned residual z = M^-1 r is stored instead of r (r = z/invd
// on free DOFs), so neither the restriction nor the direction update reads
// the diagonal.  Per iteration: q = K p with p^T q partials; z -= alpha M^-1 q
// with r^T M^-1 r partials; chunk partials of W^T K (z + beta p) over the
// non-rigid elements (beta formed there); the per-group gather with the
// scalars and stop test beside it; mu = E^-1 t; x += alpha p,
// p = z + beta p - W mu.
//
// resume: b is the new residual of a system whose previous DPCG run on this
// handle (same precision) stopped; the deflated start is formed for b as
// usual, but the first direction continues the previous run's instead of
// restarting: beta = r^T M r / rz_prev (rz_prev: r^T M r before that run's
// last update) and p = z + beta p_prev - W mu, projected like every later
// direction.  With b the true residual this is a reliable update (residual
// replacement) that keeps the Krylov information across refinement steps.
template <typename T>
static int dpcg(Handle* h, const T* b, T* x, double rtol, int max_iters, DpcgResult& res,
                hipStream_t s, bool resume = false) {
  Vecs<T> v = get_vecs<T>(h);
  const int64_t n = h->ncomp * h->g.nnodes;
  const bool defl = h->ngroups > 0;
  double* ct = h->ct;
  // Directions and alphas.  x is deferred over a ring of R directions: p_k
  // goes to ring slot k mod R, alpha_k to alring[k mod R] (k counted from the
  // run's first iteration), and x takes a window's alpha_k p_k at once, from
  // the direction update of the window's last iteration.  FP64: R = 8 (its
  // update moves 40 B per DOF and the x pass is 16 of them).  FP32: R =
  // kDirSlots32 = 1, p updated in place and x on every iteration; its vectors
  // (51 MB each at 12.8M DOF) stay largely in the 256 MB last-level cache,
  // which a ring of 2 or 8 defeats (both measured slower).
  const int R = sizeof(T) == 8 ? kDirSlots : kDirSlots32;
  const bool defer = R > 1;
  const int tslot = sizeof(T) == 8 ? 0 : 1;
  const int64_t ld = ring_stride(n);   // slots start on 16-byte boundaries
  auto slot = [&](int k) { return defer ? v.ring + (int64_t)(k % R) * ld : v.ring; };
  T* p_prev = slot(h->ring_last[tslot]);
  PYTO_TRY(ensure_control(h));
  if (defl) {
    PYTO_TRY(launch_restrict_full<T>(h, b, ct, s));
    PYTO_TRY(coarse_apply(h, ct, nullptr, nullptr, h->cmu, nullptr, s, false));
    PYTO_TRY(launch_prolong<T, false>(h, x, nullptr, h->cmu, nullptr, true, nullptr, s, nullptr, 0));
  } else {
    PYTO_CHECK_HIP(hipMemsetAsync(x, 0, n * sizeof(T), s));
  }
  double* part_pq = h->partials;
  double* part_rz = h->partials + kMaxPartials;
  const int npq = matvec_blocks<T>(h);
  PYTO_TRY(launch_matvec<T>(h, x, v.q, v.Ee, v.mc, nullptr, nullptr, nullptr, 0, s));
  res.matvecs += 1;
  // z0 = M^-1 (b - K x0)
  hipLaunchKernelGGL((k_residual_z<T>), dim3(kVecBlocks), dim3(256), 0, s, b, v.q, v.r, v.invd,
                     n, part_rz);
  hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, h->scal, (const double*)part_rz,
                     kVecBlocks, (int)S_RZ, (int)S_RZ0);
  if (resume) {
    // beta = rz / rz_prev; no x update
    hipLaunchKernelGGL(k_resume_scalars, dim3(1), dim3(1), 0, s, h->scal);
    if (defl) {
      PYTO_TRY(launch_restrict_special_parts<T>(h, v.r, p_prev, part_rz, kVecBlocks,
                                                h->scal + S_RZPREV, nullptr, s));
      PYTO_TRY(launch_coarse_gather(h, ct, nullptr, s));
      PYTO_TRY(coarse_apply(h, ct, nullptr, nullptr, h->cmu, nullptr, s, false));
    }
  } else if (defl) {
    PYTO_TRY(launch_restrict_special<T>(h, v.r, ct, nullptr, s));
    PYTO_TRY(coarse_apply(h, ct, nullptr, nullptr, h->cmu, nullptr, s, false));
  }
  // p0 = z0 - W mu, or (resume) z0 + beta p_prev - W mu, into slot 0 (alpha
  // is 0 here: no x update)
  {
    const DirIO<T> io0{x, p_prev, slot(0), v.r, false, nullptr, 0, nullptr, 0};
    PYTO_TRY(launch_dir_update<T>(h, io0, !resume, nullptr, s));
  }
  hipLaunchKernelGGL(k_start_run, dim3(1), dim3(1), 0, s, h->scal, h->flags, rtol,
                     sizeof(T) == 8 ? kFloor64 : kFloor32);
  PYTO_CHECK_HIP(hipGetLastError());
  const int* done = h->flags + F_DONE;
  // FP32 with 3 values per node: the CG update forms the Jacobi from the
  // node sums (k_cg_update_z NODE)
  const bool node_jacobi = sizeof(T) == 4 ? h->esum32 != nullptr : h->esum64 != nullptr;
  const JacobiNodes jn{h->esum32, h->fixmask, (float)h->kediag[0], (float)h->kediag[1],
                       (float)h->kediag[2], h->esum64,
                       {h->kediag[0], h->kediag[1], h->kediag[2]}};
  // one iteration's launches on stream `st`
  // iteration k (from 0 in each run): p_k from slot k, p_{k+1} to slot k+1
  auto one_iteration = [&](hipStream_t st, int k) -> int {
    T* pk = slot(k);
    T* pk1 = slot(k + 1);
    const int pos = k % R;
    PYTO_TRY(launch_matvec<T>(h, pk, v.q, v.Ee, v.mc, part_pq, done, nullptr, 0, st));
    bool updated = false;
    if (node_jacobi) {
      hipLaunchKernelGGL((k_cg_update_z<T, true>), dim3(kVecBlocks), dim3(256), 0, st, v.r,
                         v.q, v.invd, jn, n, h->scal, (const int*)h->flags,
                         (const double*)part_pq, npq, part_rz, h->alring + pos);
      updated = true;
    }
    if (!updated)
      hipLaunchKernelGGL((k_cg_update_z<T, false>), dim3(kVecBlocks), dim3(256), 0, st, v.r, v.q,
                         v.invd, jn, n, h->scal, (const int*)h->flags, (const double*)part_pq,
                         npq, part_rz, h->alring + pos);
    if (defl) {
      // beta is formed inside the restriction from the update's partials, so
      // the scalar logic runs beside the per-group gather in one launch
      PYTO_TRY(launch_restrict_special_parts<T>(h, v.r, pk, part_rz, kVecBlocks,
                                                h->scal + S_RZ, done, st));
      PYTO_TRY(launch_gather_scalars(h, ct, part_pq, npq, part_rz, kVecBlocks, st));
      PYTO_TRY(coarse_apply(h, ct, nullptr, nullptr, h->cmu, done, st, false));
    } else {
      hipLaunchKernelGGL(k_scalars, dim3(1), dim3(256), 0, st, h->scal, h->flags,
                         (const double*)part_pq, npq, (const double*)part_rz, kVecBlocks);
    }
    // deferred: x takes the window's alpha p at its last position
    const DirIO<T> io{x, pk, pk1, v.r, !defer, v.ring, ld, h->alring,
                      (defer && pos == R - 1) ? R : 0};
    PYTO_TRY(launch_dir_update<T>(h, io, false, done, st));
    PYTO_CHECK_HIP(hipGetLastError());
    return PYTO_OK;
  };
  // Batches of iterations with one read-back each, the next batch queued
  // behind the one being waited for (see kFirstBatch).  PYTO_FIXED_BATCH=n
  // (n >= 1) fixes every batch at n iterations; the iterates do not depend
  // on the batching.
  const char* env = getenv("PYTO_FIXED_BATCH");
  const int fixed = env ? std::max(0, atoi(env)) : 0;
  int enq = 0;             // iterations enqueued
  int slots[2], nflight = 0, next_slot = 0;
  int est_end = -1;        // predicted stopping iteration (-1: no estimate)
  int it_prev = -1;
  double rel_prev = 0.0;
  const double* last = nullptr;
  for (;;) {
    while (nflight < 2 && enq < max_iters) {
      const bool near = est_end >= 0 && est_end - enq <= kNearWindow;
      int nb = fixed ? fixed : near ? kNearBatch : kFarBatch;
      nb = std::min(nb, max_iters - enq);
      for (int k = enq; k < enq + nb; ++k) PYTO_TRY(one_iteration(s, k));
      const int sl = next_slot;
      next_slot ^= 1;
      PYTO_CHECK_HIP(hipMemcpyAsync(ctl_slot(h, sl), h->scal, kCtlBytes, hipMemcpyDeviceToHost, s));
      PYTO_CHECK_HIP(hipEventRecord(h->ev_ctl[sl], s));
      slots[nflight++] = sl;
      enq += nb;
    }
    if (nflight == 0) break;   // max_iters enqueued and read back
    const int sl = slots[0];
    slots[0] = slots[1];
    nflight--;
    PYTO_CHECK_HIP(hipEventSynchronize(h->ev_ctl[sl]));
    last = ctl_slot(h, sl);
    const int* lf = reinterpret_cast<const int*>(last + kNumScalars);
    if (lf[F_DONE]) break;
    // log-linear rate of the relative residual between the last two
    // read-backs; without a decrease, no estimate
    const int it = lf[F_ITERS];
    const double rel = last[S_REL];
    est_end = -1;
    if (it_prev >= 0 && it > it_prev && rel > 0.0 && rel < rel_prev) {
      const double need = std::log(rtol / rel) / (std::log(rel / rel_prev) / (it - it_prev));
      if (need < 1e6) est_end = it + std::max(1, (int)std::ceil(need));
    }
    it_prev = it;
    rel_prev = rel;
  }
  // the read-back that showed the stop (or the last one) holds the final
  // control block: later batches of the run returned at once
  if (last) std::memcpy(h->host_scal, last, kCtlBytes);
  res.iterations = h->host_flags[F_ITERS];
  res.matvecs += res.iterations;
  res.breakdown = h->host_flags[F_BREAKDOWN] != 0;
  res.stagnated = h->host_flags[F_BREAKDOWN] == 2;
  res.converged = h->host_flags[F_DONE] != 0 && !res.breakdown;
  res.rel = h->host_scal[S_REL];
  // the pending alpha_k p_k: alphas were formed by `a` iterations, of which
  // all but a stopping one ran their direction update.  Deferred, the window
  // flushes ran at every kDirSlots-th of those and the last, partial window
  // is pending; otherwise only the stopping iteration's.  A breakdown
  // publishes alpha 0.
  const int a = h->host_flags[F_ITERS];
  const int dirs = h->host_flags[F_DONE] ? std::max(0, a - 1) : a;
  const int a0 = defer ? 0 : (a - 1 + R) % R;
  const int count = defer ? a - (dirs / R) * R : a - dirs;
  if (count > 0)
    hipLaunchKernelGGL(k_flush_x<T>, dim3(kVecBlocks), dim3(256), 0, s, x, (const T*)v.ring, ld,
                       (const double*)h->alring, a0, defer ? 0 : -1, count, n);
  // p_dirs, the last direction formed, for a resumed run
  h->ring_last[tslot] = defer ? dirs % R : 0;
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

// ||f - K u|| / ||f|| over free DOFs (invd64 != 0 marks them), FP64.
static int true_residual(Handle* h, const double* f, const double* u, double* out,
                         hipStream_t s) {
  const int64_t n = h->ncomp * h->g.nnodes;
  PYTO_TRY(launch_matvec<double>(h, u, h->q64, h->Ee64, h->mc64, nullptr, nullptr, nullptr, 0, s));
  hipLaunchKernelGGL(k_norm2_diff, dim3(kVecBlocks), dim3(256), 0, s, f, h->q64, h->invd64, n,
                     h->partials);
  hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, h->scal, (const double*)h->partials,
                     kVecBlocks, (int)S_DOT0, -1);
  hipLaunchKernelGGL(k_norm2_diff, dim3(kVecBlocks), dim3(256), 0, s, f, (const double*)nullptr,
                     h->invd64, n, h->partials);
  hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, h->scal, (const double*)h->partials,
                     kVecBlocks, (int)S_DOT1, -1);
  PYTO_CHECK_HIP(hipGetLastError());
  PYTO_TRY(read_back(h, s));
  const double nf = h->host_scal[S_DOT1];
  *out = nf > 0 ? sqrt(h->host_scal[S_DOT0] / nf) : 0.0;
  return PYTO_OK;
}

static int check_ready(Handle* h) {
  PYTO_REQUIRE(h->density_set, PYTO_ESTATE, "pyto_set_density must be called first");
  PYTO_REQUIRE(h->ngroups == 0 || h->coarse_ready, PYTO_ESTATE, "coarse space not assembled");
  return ensure_solver_workspace(h);
}

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_solve_dpcg_f64(pyto_handle* h, const double* f, double* u, double rtol,
                        int max_iters, pyto_solve_stats* stats, void* stream) {
  PYTO_REQUIRE(h && f && u, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(f != u, PYTO_EINVAL, "f and u must not alias");
  PYTO_REQUIRE(rtol > 0 && max_iters >= 1, PYTO_EINVAL, "rtol must be > 0 and max_iters >= 1");
  PYTO_TRY(check_ready(h));
  hipStream_t s = as_stream(stream);
  DpcgResult r;
  PYTO_TRY(dpcg<double>(h, f, u, rtol, max_iters, r, s));
  double rel_true = 0.0;
  PYTO_TRY(true_residual(h, f, u, &rel_true, s));
  if (stats) {
    stats->iterations = r.iterations;
    stats->outer_iterations = 1;
    stats->converged = r.converged;
    stats->matvecs64 = r.matvecs + 1;
    stats->matvecs32 = 0;
    stats->rel = r.rel;
    stats->rel_true = rel_true;
    stats->setup_ms = 0.0;
  }
  if (r.stagnated) {
    set_error("DPCG stagnated at relative residual %.3e after %d iterations (best %.3e "
              "below %.0e; then no 1%% improvement in max(%d, k/2) iterations, or growth by "
              "%.0fx): the attainable accuracy is above rtol", r.rel, r.iterations,
              h->host_scal[S_RELMIN], kFloor64, kStallIters, kDivergeFactor);
    return PYTO_ENOCONV;
  }
  if (r.breakdown) {
    set_error("DPCG breakdown (p^T K p <= 0 or NaN) after %d iterations", r.iterations);
    return PYTO_ENOCONV;
  }
  if (!r.converged) {
    set_error("maximum iterations reached in DPCG; relative residual %.3e", r.rel);
    return PYTO_ENOCONV;
  }
  return PYTO_OK;
}

int pyto_solve_dpcg_mixed(pyto_handle* h, const double* f, double* u, double rtol,
                          double inner_rtol, int max_iters, int max_outer,
                          pyto_solve_stats* stats, void* stream) {
  PYTO_REQUIRE(h && f && u, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(f != u, PYTO_EINVAL, "f and u must not alias");
  PYTO_REQUIRE(rtol > 0 && inner_rtol > 0 && inner_rtol < 1 && max_iters >= 1 && max_outer >= 1,
               PYTO_EINVAL, "bad tolerances or iteration limits");
  PYTO_TRY(check_ready(h));
  hipStream_t s = as_stream(stream);
  const int64_t n = h->ncomp * h->g.nnodes;
  const bool defl = h->ngroups > 0;
  // FP64 deflated start and outer residual
  if (defl) {
    PYTO_TRY(launch_restrict_full<double>(h, f, h->ct, s));
    PYTO_TRY(coarse_apply(h, h->ct, nullptr, nullptr, h->cmu, nullptr, s, false));
    PYTO_TRY(launch_prolong<double, false>(h, u, nullptr, h->cmu, nullptr, true, nullptr, s,
                                           nullptr, 0));
  } else {
    PYTO_CHECK_HIP(hipMemsetAsync(u, 0, n * sizeof(double), s));
  }
  int mv64 = 0, mv32 = 0, inner_total = 0, outer = 0;
  PYTO_TRY(launch_matvec<double>(h, u, h->q64, h->Ee64, h->mc64, nullptr, nullptr, nullptr, 0, s));
  mv64++;
  hipLaunchKernelGGL((k_residual<double, double>), dim3(kVecBlocks), dim3(256), 0, s, f, h->q64,
                     h->r64, h->invd64, n, h->partials);
  hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, h->scal, (const double*)h->partials,
                     kVecBlocks, (int)S_TMP1, (int)S_TMP0);
  PYTO_CHECK_HIP(hipGetLastError());
  PYTO_TRY(read_back(h, s));
  const double rz0 = h->host_scal[S_TMP0];
  double rel = rz0 > 0 ? 1.0 : 0.0;
  bool converged = rz0 == 0.0;
  bool inner_fail = false, resume = false;
  // The correction systems K e = r_k are all scaled by the same factor
  // 1/sqrt(r0^T M r0) (slot S_TMP0), so that each inner run can continue the
  // previous run's search direction (reliable update): the FP32 iteration
  // keeps its Krylov space while its residual is replaced by the true FP64
  // one at every refinement step.  After a breakdown or stagnation the next
  // inner run restarts from z - W mu.
  while (!converged && outer < max_outer && inner_total < max_iters) {
    hipLaunchKernelGGL(k_scale_convert<float>, dim3(kVecBlocks), dim3(256), 0, s, h->r64, h->b32,
                       n, h->scal, (int)S_TMP0);
    // inner tolerance: what is left to reach rtol (with a margin), but no
    // tighter than inner_rtol, the FP32 floor, and no looser than 0.1
    const double want = 0.3 * rtol / rel;
    const double tol_k = std::min(0.1, std::max(inner_rtol, want));
    DpcgResult ir;
    PYTO_TRY(dpcg<float>(h, h->b32, h->x32, tol_k, max_iters - inner_total, ir, s, resume));
    inner_total += ir.iterations;
    mv32 += ir.matvecs;
    if (ir.breakdown && !ir.stagnated) inner_fail = true;
    // recurrence r^T M r at the stop, in the scaled units of the inner run
    const double rz_rec = h->host_scal[S_RZ];
    hipLaunchKernelGGL(k_axpy_scaled, dim3(kVecBlocks), dim3(256), 0, s, u, h->x32, n, h->scal,
                       (int)S_TMP0);
    PYTO_TRY(launch_matvec<double>(h, u, h->q64, h->Ee64, h->mc64, nullptr, nullptr, nullptr, 0,
                                   s));
    mv64++;
    hipLaunchKernelGGL((k_residual<double, double>), dim3(kVecBlocks), dim3(256), 0, s, f,
                       h->q64, h->r64, h->invd64, n, h->partials);
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, h->scal, (const double*)h->partials,
                       kVecBlocks, (int)S_TMP1, -1);
    PYTO_CHECK_HIP(hipGetLastError());
    PYTO_TRY(read_back(h, s));
    outer++;
    rel = sqrt(fabs(h->host_scal[S_TMP1] / rz0));
    converged = rel <= rtol;
    // Continue the search direction only while the FP32 recurrence still
    // tracks the true residual (within a factor 1.25 in norm); once rounding
    // drift dominates, the old direction is no longer conjugate to the
    // replaced residual and the next run restarts.
    resume = !ir.breakdown && ir.converged &&
             h->host_scal[S_TMP1] / rz0 <= kResumeDrift * kResumeDrift * rz_rec;
    if (inner_fail) break;
  }
  double rel_true = 0.0;
  PYTO_TRY(true_residual(h, f, u, &rel_true, s));
  if (stats) {
    stats->iterations = inner_total;
    stats->outer_iterations = outer;
    stats->converged = converged;
    stats->matvecs64 = mv64 + 1;
    stats->matvecs32 = mv32;
    stats->rel = rel;
    stats->rel_true = rel_true;
    stats->setup_ms = 0.0;
  }
  if (!converged) {
    set_error("mixed-precision DPCG did not reach rtol %.1e (relative residual %.3e after %d "
              "refinement steps, %d inner iterations)", rtol, rel, outer, inner_total);
    return PYTO_ENOCONV;
  }
  return PYTO_OK;
}

}  // extern "C"
