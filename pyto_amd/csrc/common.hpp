// Shared definitions of libpytohip: the handle, error plumbing, launch
// helpers and the wave/block reduction primitives used by every kernel file.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/pyto_hip.h"

namespace pyto {

// ---------------------------------------------------------------------------
// Errors: a thread-local message and status codes; nothing throws across the
// C ABI.
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);

#define PYTO_CHECK_HIP(...)                                                    \
  do {                                                                         \
    hipError_t e_ = (__VA_ARGS__);                                             \
    if (e_ != hipSuccess) {                                                    \
      ::pyto::set_error("%s failed: %s (%s:%d)", #__VA_ARGS__, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                   \
      return PYTO_EHIP;                                                        \
    }                                                                          \
  } while (0)

#define PYTO_REQUIRE(cond, code, ...) \
  do {                                \
    if (!(cond)) {                    \
      ::pyto::set_error(__VA_ARGS__); \
      return (code);                  \
    }                                 \
  } while (0)

#define PYTO_TRY(...)         \
  do {                        \
    int rc_ = (__VA_ARGS__);  \
    if (rc_ != PYTO_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// Constants of the element operator in the Hadamard ("modal") basis.
//
// For a trilinear brick the nodal values transform into the 8 coefficients of
// {1, xi, eta, zeta, xi eta, xi zeta, eta zeta, xi eta zeta}; these monomials
// are orthogonal over the element, so the energy a(u,u) splits into one
// isotropic stress evaluation per monomial and KE*u costs ~270 operations
// instead of the 576 FMAs of a dense 24x24 product.  The formulation is exact
// (2x2x2 Gauss integrates the same polynomials exactly).
// ---------------------------------------------------------------------------
template <typename T>
struct ModalConsts {
  T gx, gy, gz;   // 2/h per axis
  T lam, mu;      // Lame constants of E0 times V/64 (Hadamard scaling folded)
  T tw[7];        // thermal: k0 V/64 times the modal weights (element.hpp)
  T cl, cm, cl3, cm3, ck9;   // cube elements: lam g^2, mu g^2, their thirds, (cl + 4 cm)/9
};

struct GridDims {
  int nx, ny, nz;   // elements per axis
  int sx, sy, sz;   // nodes per axis (n+1)
  int64_t nnodes, nelems;
  double hx, hy, hz;
};

// Work chunk of the special-element lists: up to 64 elements (one wave) that
// share the same set of deflation groups.
// One special element: its id and, per local node j, the slot (3 bits at
// 3j) of its group in the chunk's group list and its fixed-DOF bits
// (3 bits at 24+3j).
struct SpecialElem {
  int32_t e;
  int32_t pad;
  uint64_t info;
};

struct SpecialChunk {
  int start;      // offset into special element array
  int count;      // <= 64
  int ng;         // number of groups in the set (1..8)
  int groups[8];  // the groups (sorted), -1 padded
};

// A workgroup of the special-element restriction: up to kSpecialBlockChunks
// consecutive chunks whose corner nodes (sorted, unique, at most
// kSpecialBlockNodes) are staged together in LDS.
struct SpecialBlock {
  int chunk0;     // first chunk
  int nch;        // chunks in the block
  int node0;      // offset into the block node list
  int nnodes;     // unique corner nodes of the block
};
constexpr int kSpecialBlockChunks = 4;
constexpr int kSpecialBlockNodes = 1024;

// One lane of the single-system restriction in block-padded order
// (block b, wave w, lane l at (b kSpecialBlockChunks + w) 64 + l): the
// element (-1: no element), its packed slot and fixed bits, and its 8
// block-local corner indices (u16).
struct SpecialSlot {
  int32_t e;
  int32_t pad;
  uint64_t info;
  uint4 li;
};

struct Handle {
  int device = 0;
  GridDims g{};
  double E0 = 1.0, nu = 0.3, penal = 3.0, eps_void = 1e-8;
  ModalConsts<double> mc64{};
  ModalConsts<float> mc32{};
  double kediag[3] = {0, 0, 0};
  int ncomp = 3;                // values per node: 3 (structural) or 1 (thermal)
  int tile_rows = 16;           // operator tile: 16 x tile_rows element columns (8, 16 or 32)

  // constants / per-density state
  uint8_t* fixmask = nullptr;   // per node
  uint8_t* tileflags = nullptr; // per (matvec tile, node plane): any fixed node in its window,
                                // sized for the smallest tile (8 rows)
  double* Ee64 = nullptr;       // per element E_e/E0 (SIMP)
  float* Ee32 = nullptr;
  double* x64 = nullptr;        // copy of densities
  double* invd64 = nullptr;     // per DOF
  float* invd32 = nullptr;
  double* esum64 = nullptr;     // per node (3 values per node only): the FP64 sum of E_e
                                // from which the FP64 CG update forms invd64 on the fly
  float* esum32 = nullptr;      // per node (3 values per node only): FP32 sum of E_e over
                                // the node's elements, from which the FP32 CG update
                                // forms invd32 on the fly (jacobi_inv32)
  bool density_set = false;

  // deflation
  int ngroups = 0, nc = 0;      // nc = 6*ngroups
  int32_t* node_group = nullptr;
  uint16_t* gmask = nullptr;    // per node: group (bits 0-11) | fixed bits << 12
  double* centers = nullptr;
  int nspecial = 0, nchunks = 0;
  SpecialElem* sinfo = nullptr; // same order, with packed slots and fixed bits
  SpecialChunk* chunks = nullptr;
  double* cpart = nullptr;      // per-chunk partial W^T K v (nchunks x 8 x 6)
  int nsblocks = 0;
  // the same blocks padded, so that every load of a restriction block except
  // the staged vector values depends on the block index alone
  int2* pwave = nullptr;        // per (block, wave): (chunk index or -1, ng)
  SpecialSlot* pslot = nullptr; // per (block, wave, lane)
  int32_t* pnodes = nullptr;    // per block: kSpecialBlockNodes node ids (padded)
  double* pcen = nullptr;       // per (block, wave): 8 slots x 3 group-centre coordinates
  double* pE64 = nullptr;       // per (block, wave, lane): E_e (0 without an element)
  float* pE32 = nullptr;
  int32_t* gptr = nullptr;      // per group: range of (chunk*8+slot) in gidx
  int32_t* gidx = nullptr;
  // nodes sorted by group, cut into blocks of <= 256 nodes of one group
  int nnchunks = 0;
  int32_t* gnodes = nullptr;    // node ids, sorted by (group, id)
  int2* nchunk = nullptr;       // (start, count) per node block
  int32_t* nchunk_grp = nullptr;
  double* npart = nullptr;      // per node block partials (nnchunks x 8 x 6, slot 0 used)
  int32_t* nptr = nullptr;      // per group: range in nidx
  int32_t* nidx = nullptr;
  // Galerkin assembly: per chunk ng^2 6x6 partial blocks (at gal_off), per
  // group pair the list of its blocks
  int npairs = 0;
  int32_t* gal_off = nullptr;
  double* gal_part = nullptr;
  int2* gal_pairs = nullptr;    // (ga, gb)
  int32_t* gal_pptr = nullptr;
  int32_t* gal_pidx = nullptr;
  double* Ec = nullptr;         // coarse matrix (nc x nc), then Cholesky factor
  double* Einv = nullptr;       // coarse inverse (nc x nc)
  // large coarse spaces: the inverse's upper-triangle 64 x 64 blocks packed
  // (kSymBlock), and the block partials of the symmetric GEMV
  double* Epk = nullptr;
  double* spart = nullptr;
  int pk_nb = 0;                // block rows of Epk (0: not packed)
  double* cwork = nullptr;      // ld workspace (coarse refinement correction)
  void* blas = nullptr;         // rocblas_handle of the dense coarse factorisation
  double* ct = nullptr;         // coarse rhs: W^T K z (ld), then W^T K p (ld) at ct + ld
  double* cmu = nullptr;        // coarse solution (nc)
  double* cres = nullptr;       // coarse residual (nc)
  bool coarse_ready = false;
  int groups_gen = 0;           // counts pyto_set_groups calls (batches check it)
  double shift_abs = 1e-10;     // coarse diagonal shift (reference default)
  double shift_rel = 0.0;

  // solver workspace (allocated on first use); in the DPCG, r holds
  // z = M^-1 r
  double *r64 = nullptr, *q64 = nullptr;
  float *r32 = nullptr, *q32 = nullptr, *x32 = nullptr, *b32 = nullptr;
  double* partials = nullptr;   // block partial sums: 2 x kMaxPartials (p.q, then r.Mr)
  // direction rings (kDirSlots DOF vectors in FP64, one in FP32), the alphas
  // of a window and the slot of each precision's last direction (resumed runs)
  double* pring64 = nullptr;
  float* pring32 = nullptr;
  double* alring = nullptr;
  int ring_last[2] = {0, 0};
  unsigned* counters = nullptr; // last-block counters (kNumCounters)
  double* scal = nullptr;       // device scalars (kNumScalars), then the flags
  int* flags = nullptr;         // device flags (kNumFlags), inside scal's allocation
  double* host_scal = nullptr;  // pinned mirror of scal and flags (kCtlBytes)
  int* host_flags = nullptr;    // inside host_scal's allocation
  double* elem_buf = nullptr;   // N_el x 6 FP64 element workspace
  void* stencil = nullptr;      // cone-filter stencil on the device (postproc.hip)
  float* ke32 = nullptr;        // dense FP32 KE for the MFMA operator variant (lazy)
  double stencil_rel = 0.0;     // its radius in units of hx (0: none built yet)
  // the DPCG's read-backs of the control block, one per batch of iterations
  // in flight: two pinned copies (2 x kCtlBytes) and their events
  double* host_ctl = nullptr;
  hipEvent_t ev_ctl[2] = {nullptr, nullptr};
  int64_t workspace_bytes = 0;
};

constexpr int kMaxPartials = 8192;

// The DPCG keeps its last kDirSlots directions in a ring (slot k mod
// kDirSlots holds p_k) and adds alpha_k p_k to x once per kDirSlots
// iterations instead of on every iteration (FP64).  The FP32 solver keeps
// one slot (kDirSlots32) and updates x on every iteration: its vectors stay
// largely in the last-level cache, which even a second slot defeats.
constexpr int kDirSlots = 8;
constexpr int kDirSlots32 = 1;
static_assert(kDirSlots32 >= 1 && kDirSlots32 <= kDirSlots, "the FP32 ring fits the alphas");
// Elements between ring slots for n-element DOF vectors: n rounded up to a
// multiple of 4, so that every slot starts on a 16-byte boundary (the vector
// kernels read 16 bytes per thread).
inline int64_t ring_stride(int64_t n) { return (n + 3) / 4 * 4; }

// Direction-update I/O: p_out = z + beta p_in - W mu; with xupd, x += alpha
// p_in (alpha = scal[S_ALPHA_X]); with nflush > 0, x += sum over s <
// nflush of al[s] ring[s * ld + i] (the deferred updates of a window).
template <typename T>
struct DirIO {
  T* x;
  const T* pin;
  T* pout;
  const T* z;
  bool xupd;
  const T* ring;
  int64_t ld;
  const double* al;
  int nflush;
};
constexpr int kNumCounters = 16;
constexpr int kNumScalars = 32;
constexpr int kNumFlags = 10;
// a handle's device scalars followed by its flags, one allocation
constexpr size_t kCtlBytes = kNumScalars * sizeof(double) + kNumFlags * sizeof(int);

// Device scalar slots.
enum ScalarSlot {
  S_RZ = 0, S_RZ0 = 1, S_PKP = 2, S_ALPHA = 3, S_BETA = 4, S_RZNEW = 5, S_REL = 6,
  S_TMP0 = 7, S_TMP1 = 8, S_TMP2 = 9, S_RTOL = 10, S_DOT0 = 11, S_DOT1 = 12,
  S_SUMVMP = 13, S_COMPLIANCE = 14, S_SCALE = 15,
  // OC bisection: two (l1, l2, lmid, done) states alternating by step
  // parity, and the last multiplier at which the volume was evaluated
  S_OC_STATE = 16, S_OC_LAST = 24,
  S_RELMIN = 25,  // best relative residual of the running solve (stagnation test)
  S_FLOOR = 26,   // the stagnation test applies once S_RELMIN is below this
  S_RZPREV = 27,  // r^T M r before the last iteration's update (resumed directions)
  S_ALPHA_X = 28  // alpha of the running iteration for x += alpha p
};
// Device flag slots.
// F_BREAKDOWN: 1 = p^T K p <= 0 or NaN, 2 = stagnation (see k_scalars)
enum FlagSlot { F_DONE = 0, F_ITERS = 1, F_BREAKDOWN = 2, F_COARSE_PIVOT = 3, F_OC_DONE = 4,
                F_OC_STEPS = 5, F_STALL = 6 };
// Past its attainable accuracy the deflated iteration's recursive residual
// stops falling and then grows geometrically (rounding leaves coarse-space
// components in r that no K-orthogonal direction removes; p^T K p shrinks,
// alpha and the iterate grow).  Once its best relative residual is below a
// floor near that accuracy (kFloor64 / kFloor32), a solve therefore stops as
// stagnated when its relative residual exceeds kDivergeFactor times its best
// value, or has not improved on it by 1% for max(kStallIters, k/2)
// iterations (k = iterations so far).  Above the floor nothing is tested:
// CG residuals of hard problems plateau for hundreds of iterations and rise
// by large factors on the way down.
constexpr int kStallIters = 200;
constexpr double kDivergeFactor = 1e3;
constexpr double kFloor64 = 1e-11;
constexpr double kFloor32 = 1e-5;
// Last-block counter slots (one per reduction site in flight).
enum CounterSlot { C_MATVEC = 0, C_UPDATE = 1, C_DOT = 2, C_ELEM = 3, C_NORM = 4 };

hipStream_t as_stream(void* s);
int ensure_solver_workspace(Handle* h);
int device_alloc(Handle* h, void** p, size_t bytes);
// Operator tile rows (8, 16 or 32) and the tiles' fixed-node flags for them.
int set_tile_rows(Handle* h, int rows, hipStream_t s);

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b mod 8), and each XCD has its own L2.  Map hardware block b to a logical
// block so that every XCD receives one contiguous range of logical blocks:
// neighbouring tiles then share halo cache lines in one L2 instead of
// fetching them once per XCD.  A bijection on [0, nb) for any nb.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_logical_block(int b, int nb) {
  const int q = nb / kNumXcd, rem = nb % kNumXcd;
  const int x = b % kNumXcd, i = b / kNumXcd;
  return x * q + min(x, rem) + i;
}
__device__ __forceinline__ double wave_reduce_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Wave sums of six per-lane values at once: every lane l ends with the total
// over the wave of value l >> 3 (lanes 48..63 with a sum of zeros).  Three
// exchange steps (lane offsets 32, 16, 8) halve the values a lane carries,
// three more sum the last one within groups of 8 lanes: 12 cross-lane steps
// instead of 36 for six separate reductions, in a fixed order.
__device__ __forceinline__ double wave_reduce6(const double a[6]) {
  const int lane = threadIdx.x & 63;
  const double v[8] = {a[0], a[1], a[2], a[3], a[4], a[5], 0.0, 0.0};
  const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0, h3 = (lane & 8) != 0;
  double k4[4], k2[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double give = h5 ? v[i] : v[4 + i];
    k4[i] = (h5 ? v[4 + i] : v[i]) + __shfl_xor(give, 32, 64);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double give = h4 ? k4[i] : k4[2 + i];
    k2[i] = (h4 ? k4[2 + i] : k4[i]) + __shfl_xor(give, 16, 64);
  }
  const double give = h3 ? k2[0] : k2[1];
  double k1 = (h3 ? k2[1] : k2[0]) + __shfl_xor(give, 8, 64);
  k1 += __shfl_xor(k1, 4, 64);
  k1 += __shfl_xor(k1, 2, 64);
  k1 += __shfl_xor(k1, 1, 64);
  return k1;
}

// Block-wide FP64 sum for blockDim.x == 256; result valid in thread 0.
__device__ __forceinline__ double block_reduce_sum256(double v, double* lds4) {
  v = wave_reduce_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  return r;
}

// Sum of n per-block partials written by an earlier kernel, in a fixed order,
// returned to every thread (blockDim.x == 256).  Kernel boundaries make the
// partials visible, so no fence or atomic is involved; every block that
// calls this gets bit-identical results.
// One thread's strided share of n partials, part[t], part[t + stride], ...,
// summed in that order.  The loads go out 8 at a time before they are added
// (a loop of load-then-add with a run-time trip count waits for each load,
// and for every load issued before it, in turn).
__device__ __forceinline__ double strided_sum(const double* __restrict__ part, int n, int t,
                                              int stride) {
  double s = 0.0;
  int i = t;
  for (; i + 7 * stride < n; i += 8 * stride) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = part[i + k * stride];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; i < n; i += stride) s += part[i];
  return s;
}

__device__ __forceinline__ double block_sum_partials(const double* __restrict__ part, int n,
                                                     double* lds5) {
  double s = strided_sum(part, n, threadIdx.x, 256);
  s = wave_reduce_sum(s);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds5[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) lds5[4] = (lds5[0] + lds5[1]) + (lds5[2] + lds5[3]);
  __syncthreads();
  return lds5[4];
}

// The scalar step of one PCG iteration, run by one thread: alpha, the
// relative residual sqrt(rz_new/rz0) with the stop test (src/deflation.py:567),
// the breakdown test (p^T K p <= 0 or NaN), the stagnation test (see
// kFloor64), beta and rz.  sc/fl are the system's scalar and flag slots.
__device__ __forceinline__ void pcg_scalar_step(double* sc, int* fl, double pkp, double rz_new,
                                                double floor) {
  const double rz = sc[S_RZ];
  const double rel = sqrt(fabs(rz_new / sc[S_RZ0]));
  sc[S_PKP] = pkp;
  sc[S_ALPHA] = rz / pkp;
  sc[S_RZNEW] = rz_new;
  sc[S_REL] = rel;
  fl[F_ITERS] += 1;
  if (!(pkp > 0.0) || !(rz_new == rz_new)) {
    fl[F_BREAKDOWN] = 1;
    fl[F_DONE] = 1;
  } else if (rel <= sc[S_RTOL]) {
    fl[F_DONE] = 1;
  } else if (rel < 0.99 * sc[S_RELMIN]) {
    sc[S_RELMIN] = rel;
    fl[F_STALL] = 0;
  } else if (sc[S_RELMIN] < floor &&
             (rel > kDivergeFactor * sc[S_RELMIN] ||
              ++fl[F_STALL] >= max(kStallIters, fl[F_ITERS] / 2))) {
    fl[F_BREAKDOWN] = 2;
    fl[F_DONE] = 1;
  }
  sc[S_BETA] = rz_new / rz;
  sc[S_RZPREV] = rz;
  sc[S_RZ] = rz_new;
}

// Two sums of block partials in one pass (as block_sum_partials, fixed
// order), returned to every thread; lds10 holds 10 doubles.
__device__ __forceinline__ void block_sum_partials2(const double* __restrict__ pa, int na,
                                                    const double* __restrict__ pb, int nb,
                                                    double* lds10, double& sa, double& sb) {
  double a = strided_sum(pa, na, threadIdx.x, 256);
  double b = strided_sum(pb, nb, threadIdx.x, 256);
  a = wave_reduce_sum(a);
  b = wave_reduce_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    lds10[w] = a;
    lds10[4 + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    lds10[8] = (lds10[0] + lds10[1]) + (lds10[2] + lds10[3]);
    lds10[9] = (lds10[4] + lds10[5]) + (lds10[6] + lds10[7]);
  }
  __syncthreads();
  sa = lds10[8];
  sb = lds10[9];
}

// Deterministic grid reduction: every block stores its partial, the last
// block to finish sums the partials in index order and calls `fin` with the
// total.  `counter` returns to zero, so the pattern can be replayed.
template <typename Fin>
__device__ __forceinline__ void grid_reduce_finish(double block_sum, double* partials,
                                                   unsigned* counter, int nblocks,
                                                   int block_id, Fin fin) {
  __shared__ bool am_last;
  if (threadIdx.x == 0) {
    partials[block_id] = block_sum;
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == (unsigned)(nblocks - 1));
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  double s = strided_sum(partials, nblocks, threadIdx.x, blockDim.x);
  __shared__ double red[4];
  double tot = block_reduce_sum256(s, red);
  if (threadIdx.x == 0) {
    fin(tot);
    *counter = 0u;
  }
}

// Wave-level segmented flush of per-lane 6-vectors keyed by group id (gid < 0:
// nothing): lanes sharing a group are summed and one lane adds the sum to
// t[6*gid .. 6*gid+5].  Must be called by all lanes of the wave.
__device__ __forceinline__ void wave_flush6(const double acc[6], int gid, double* t) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(gid >= 0);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lg = __shfl(gid, leader, 64);
    const bool mine = gid == lg;
    const uint64_t members = __ballot(mine);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double s = wave_reduce_sum(mine ? acc[j] : 0.0);
      if (lane == leader) atomicAdd(&t[6 * lg + j], s);
    }
    pending &= ~members;
  }
}

// Node index -> (ix, iy, iz) without integer division by runtime values.
// The FP32 Jacobi inverse of one DOF from its node's FP32 sum of E_e and
// the element diagonal entry kd of its component: 0 for a fixed DOF,
// d = kd s with |d| < 1e-12 -> 1 (src/linear_solvers.py:84-85), 1/d.  d
// (0 when fixed) is returned too.  invd32 is written with it and the FP32 CG
// update forms it again, so both hold the same bits.
__device__ __forceinline__ float jacobi_inv32(float kd, float s, bool fixed, float& d) {
  float dd = kd * s;
  if (fabsf(dd) < 1e-12f) dd = 1.0f;
  d = fixed ? 0.0f : dd;
  return fixed ? 0.0f : 1.0f / dd;
}

// Per-node source of the FP32 Jacobi (jacobi_inv32).
struct JacobiNodes {
  const float* esum;      // per node (FP32 solver)
  const uint8_t* mask;    // per node, fixed bits
  float kd0, kd1, kd2;
  const double* esum64;   // per node (FP64 solver)
  double kd64[3];
};

// Non-temporal 16-byte accesses, for x: the direction updates read and
// write it once per iteration (FP64: once per window) and nothing reads it
// in between, so its lines are not kept; the vectors the next kernels read
// (p, q, z, the element moduli) then stay in the last-level cache.
typedef float pyto_f4 __attribute__((ext_vector_type(4)));
typedef double pyto_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 load_nt(const float4* a) {
  const pyto_f4 t = __builtin_nontemporal_load(reinterpret_cast<const pyto_f4*>(a));
  return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ double2 load_nt(const double2* a) {
  const pyto_d2 t = __builtin_nontemporal_load(reinterpret_cast<const pyto_d2*>(a));
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ void store_nt(float4* a, const float4& v) {
  const pyto_f4 t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, reinterpret_cast<pyto_f4*>(a));
}
__device__ __forceinline__ void store_nt(double2* a, const double2& v) {
  const pyto_d2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<pyto_d2*>(a));
}

// One DOF of the z-form CG update z - alpha M^-1 q with the FP32 Jacobi
// inverse formed from the node's sum of E_e (jacobi_inv32), and its
// r^T M^-1 r term z^2 d (d the diagonal entry, 0 on fixed DOFs).
__device__ __forceinline__ float cg_z_node32(float z, float q, float kd, float s, bool fixed,
                                             double alpha, double& term) {
  float d;
  const float di = jacobi_inv32(kd, s, fixed, d);
  const float zn = (float)((double)z - alpha * ((double)di * (double)q));
  term = (double)zn * (double)zn * (double)d;
  return zn;
}


__device__ __forceinline__ void node_coords(int64_t n, const GridDims& g, int& ix, int& iy,
                                            int& iz) {
  const double plane = (double)g.sx * (double)g.sy;
  int64_t q = (int64_t)((double)n / plane);
  int64_t r = n - q * (int64_t)g.sx * g.sy;
  if (r < 0) { q -= 1; r += (int64_t)g.sx * g.sy; }
  if (r >= (int64_t)g.sx * g.sy) { q += 1; r -= (int64_t)g.sx * g.sy; }
  int64_t q2 = (int64_t)((double)r / (double)g.sx);
  int64_t r2 = r - q2 * g.sx;
  if (r2 < 0) { q2 -= 1; r2 += g.sx; }
  if (r2 >= g.sx) { q2 += 1; r2 -= g.sx; }
  iz = (int)q;
  iy = (int)q2;
  ix = (int)r2;
}

__device__ __forceinline__ void elem_coords(int64_t e, const GridDims& g, int& ix, int& iy,
                                            int& iz) {
  const int64_t plane = (int64_t)g.nx * g.ny;
  int64_t q = (int64_t)((double)e / (double)plane);
  int64_t r = e - q * plane;
  if (r < 0) { q -= 1; r += plane; }
  if (r >= plane) { q += 1; r -= plane; }
  int64_t q2 = (int64_t)((double)r / (double)g.nx);
  int64_t r2 = r - q2 * g.nx;
  if (r2 < 0) { q2 -= 1; r2 += g.nx; }
  if (r2 >= g.nx) { q2 += 1; r2 -= g.nx; }
  iz = (int)q;
  iy = (int)q2;
  ix = (int)r2;
}

// Local node (0..7) of an element -> corner offsets (a,b,c), reference order
// (src/hex_mesher.py:121-124).
__host__ __device__ constexpr int corner_a(int n) { return (n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0; }
__host__ __device__ constexpr int corner_b(int n) { return (n == 2 || n == 3 || n == 6 || n == 7) ? 1 : 0; }
__host__ __device__ constexpr int corner_c(int n) { return n >= 4 ? 1 : 0; }

}  // namespace pyto

// The opaque C handle is the C++ Handle.
struct pyto_handle : public pyto::Handle {};
