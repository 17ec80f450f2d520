// Optimality-criteria density update with the volume bisection on the device
// (src/topopt_ocm.py:247-269):
//
//   l1 = 0, l2 = 1e12
//   while (l2 - l1) / ((l1 + l2)/2 + 1e-10) > 1e-7:
//     lmid = (l1 + l2)/2
//     D = clip(-dc / (dv lmid), 0.1, 10)
//     xnew = max(xmin, max(x - move, min(xmax, min(x + move, x sqrt(D)))))
//     if mean(xnew)/volfrac - 1 > 0: l1 = lmid  else: l2 = lmid
//   x = xnew            (the xnew of the last lmid evaluated)
//
// One launch of k_oc_step per bisection step.  A launch first finishes the
// previous step (sums its block partials of sum(xnew) in a fixed order,
// which every block does identically, and moves l1 or l2), then tests the
// loop condition and evaluates the new lmid.  The (l1, l2, lmid) state
// alternates between two slots by step parity so that no block reads a state
// another block of the same launch has already advanced.  When the loop
// condition fails, block 0 marks the next state done and later launches
// return at once, so the host enqueues a bounded number of steps without reading
// anything back.  k_oc_apply then writes xnew at the last evaluated lmid.
#include <mutex>

#include "common.hpp"

namespace pyto {

constexpr int kOcMaxBlocks = kMaxPartials / 2;
constexpr int kOcMaxSteps = 4096;

__device__ __forceinline__ double oc_xnew(double x, double dc, double dv, double lmid,
                                          double move, double xmin, double xmax) {
  double D = -dc / (dv * lmid);
  D = fmin(fmax(D, 0.1), 10.0);
  return fmax(xmin, fmax(x - move, fmin(xmax, fmin(x + move, x * sqrt(D)))));
}

__global__ __launch_bounds__(256) void k_oc_step(const double* __restrict__ x,
                                                 const double* __restrict__ dc, int64_t n,
                                                 double dv, double volfrac, double move,
                                                 double xmin, double xmax, int step,
                                                 double* __restrict__ scal,
                                                 int* __restrict__ flags,
                                                 double* __restrict__ partials, int nblocks,
                                                 int pstride) {
  __shared__ double lds[5];
  // state written by block 0 of the previous launch (never by this launch)
  const double* st = scal + S_OC_STATE + 4 * (step & 1);
  if (st[3] != 0.0) {
    // carry the done marker into the state the next launch reads (no block
    // of this launch reads that slot)
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_OC_STATE + 4 * ((step + 1) & 1) + 3] = 1.0;
    return;
  }
  double l1 = st[0], l2 = st[1], lmid = st[2];
  if (step > 0) {
    const double sum = block_sum_partials(partials + pstride * ((step - 1) & 1), nblocks, lds);
    const double vol = (sum / (double)n) / volfrac - 1.0;
    if (vol > 0) l1 = lmid;
    else l2 = lmid;
  }
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!((l2 - l1) / ((l1 + l2) / 2 + 1e-10) > 1e-7)) {
    if (first) {
      double* nx = scal + S_OC_STATE + 4 * ((step + 1) & 1);
      nx[3] = 1.0;
      flags[F_OC_DONE] = 1;   // read by the host only
      flags[F_OC_STEPS] = step;
    }
    return;
  }
  lmid = 0.5 * (l2 + l1);
  double s = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += oc_xnew(x[i], dc[i], dv, lmid, move, xmin, xmax);
  const double bs = block_reduce_sum256(s, lds);
  if (threadIdx.x == 0) partials[pstride * (step & 1) + blockIdx.x] = bs;
  if (first) {
    double* nx = scal + S_OC_STATE + 4 * ((step + 1) & 1);
    nx[0] = l1;
    nx[1] = l2;
    nx[2] = lmid;
    nx[3] = 0.0;
    scal[S_OC_LAST] = lmid;
    flags[F_OC_STEPS] = step + 1;
  }
}

__global__ void k_oc_init(double* __restrict__ scal, int* __restrict__ flags) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    scal[S_OC_STATE + 0] = 0.0;
    scal[S_OC_STATE + 1] = 1e12;
    scal[S_OC_STATE + 2] = 0.5 * 1e12;
    scal[S_OC_STATE + 3] = 0.0;
    scal[S_OC_LAST] = 0.5 * 1e12;
    flags[F_OC_DONE] = 0;
    flags[F_OC_STEPS] = 0;
  }
}

__global__ __launch_bounds__(256) void k_oc_apply(const double* __restrict__ x,
                                                  const double* __restrict__ dc, int64_t n,
                                                  double dv, double move, double xmin,
                                                  double xmax, const double* __restrict__ scal,
                                                  double* __restrict__ xnew) {
  const double lmid = scal[S_OC_LAST];
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    xnew[i] = oc_xnew(x[i], dc[i], dv, lmid, move, xmin, xmax);
}

// The launch sequence of one OC update: init, max_steps bisection steps,
// the final apply; scal/flags/partials are the caller's scratch (the
// handle's, or the allocation-free variant's static array).
static int oc_launch(const double* x, const double* dc, int64_t n, double dv, double volfrac,
                     double move, double xmin, double xmax, int max_steps, double* xnew,
                     double* scal, int* flags, double* partials, int pstride,
                     double* lmid_host, int* steps_host, hipStream_t s) {
  const int nblocks = (int)std::min<int64_t>((n + 255) / 256, kOcMaxBlocks);
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(64), 0, s, scal, flags);
  for (int k = 0; k < max_steps; ++k)
    hipLaunchKernelGGL(k_oc_step, dim3(nblocks), dim3(256), 0, s, x, dc, n, dv, volfrac, move,
                       xmin, xmax, k, scal, flags, partials, nblocks, pstride);
  hipLaunchKernelGGL(k_oc_apply, dim3(nblocks), dim3(256), 0, s, x, dc, n, dv, move, xmin, xmax,
                     (const double*)scal, xnew);
  PYTO_CHECK_HIP(hipGetLastError());
  if (lmid_host || steps_host) {
    int fl[kNumFlags];
    double sc[kNumScalars];
    PYTO_CHECK_HIP(hipMemcpyAsync(fl, flags, sizeof(fl), hipMemcpyDeviceToHost, s));
    PYTO_CHECK_HIP(hipMemcpyAsync(sc, scal, sizeof(sc), hipMemcpyDeviceToHost, s));
    PYTO_CHECK_HIP(hipStreamSynchronize(s));
    if (lmid_host) *lmid_host = sc[S_OC_LAST];
    if (steps_host) *steps_host = fl[F_OC_DONE] ? fl[F_OC_STEPS] : -fl[F_OC_STEPS];
  }
  return PYTO_OK;
}

static int oc_check(const double* x, const double* dc, int64_t n, double dv, double volfrac,
                    double move, double xmin, double xmax, int max_steps, double* xnew) {
  PYTO_REQUIRE(x && dc && xnew, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(n > 0, PYTO_EINVAL, "n must be positive");
  PYTO_REQUIRE(dv > 0 && volfrac > 0 && move >= 0 && xmin <= xmax, PYTO_EINVAL,
               "need dv > 0, volfrac > 0, move >= 0 and xmin <= xmax");
  PYTO_REQUIRE(max_steps > 0 && max_steps <= kOcMaxSteps, PYTO_EINVAL,
               "max_steps %d outside [1, %d]", max_steps, kOcMaxSteps);
  return PYTO_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the same kernels on a file-scope device array.
// Layout (doubles): two parities of kOcMaxBlocks block partials, then
// kNumScalars scalar slots, then kNumFlags ints.  65.9 KB per GPU, whatever
// n is (the kernels stride over the elements with at most kOcMaxBlocks
// workgroups).  k_oc_init rewrites the bisection state and the flags at the
// start of every call, and every partial is written before it is read, so
// nothing needs clearing between calls; calls that overlap on one GPU race
// on this array.
// ---------------------------------------------------------------------------
constexpr int kScratchDoubles = 2 * kOcMaxBlocks + kNumScalars + (kNumFlags + 1) / 2;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of `scratch` per device, looked up once per device with
// that device current (the array's name is a host-side stand-in, not a
// device address).
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptr(double** out) {
  int dev = 0;
  PYTO_CHECK_HIP(hipGetDevice(&dev));
  PYTO_REQUIRE(dev >= 0 && dev < kMaxDevices, PYTO_EINVAL, "device %d beyond %d", dev,
               kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PYTO_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PYTO_REQUIRE(p != nullptr, PYTO_EHIP, "hipGetSymbolAddress returned a null address");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return PYTO_OK;
}

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_oc_update_noalloc(const double* x, const double* dc, int64_t n, double dv,
                           double volfrac, double move, double xmin, double xmax, int max_steps,
                           double* xnew, double* lmid_host, int* steps_host, void* stream) {
  PYTO_TRY(oc_check(x, dc, n, dv, volfrac, move, xmin, xmax, max_steps, xnew));
  double* base = nullptr;
  PYTO_TRY(scratch_ptr(&base));
  double* partials = base;
  double* scal = base + 2 * kOcMaxBlocks;
  int* flags = reinterpret_cast<int*>(scal + kNumScalars);
  return oc_launch(x, dc, n, dv, volfrac, move, xmin, xmax, max_steps, xnew, scal, flags,
                   partials, kOcMaxBlocks, lmid_host, steps_host, as_stream(stream));
}

int pyto_oc_update(pyto_handle* h, const double* x, const double* dc, int64_t n, double dv,
                   double volfrac, double move, double xmin, double xmax, int max_steps,
                   double* xnew, double* lmid_host, int* steps_host, void* stream) {
  PYTO_REQUIRE(h, PYTO_EINVAL, "null handle");
  PYTO_TRY(oc_check(x, dc, n, dv, volfrac, move, xmin, xmax, max_steps, xnew));
  return oc_launch(x, dc, n, dv, volfrac, move, xmin, xmax, max_steps, xnew, h->scal, h->flags,
                   h->partials, kOcMaxBlocks, lmid_host, steps_host, as_stream(stream));
}

}  // extern "C"
