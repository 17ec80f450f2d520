// Connected components of the elements present in a density field
// (HexMesher.find_connected_components, src/hex_mesher.py:1285-1339): the
// elements with density > threshold, joined through the reference's
// 27-entry neighbour array (face, edge and corner neighbours,
// src/hex_mesher.py:126-151).
//
// Label propagation on the device: every present element starts with its
// own index as label; a hook pass lowers each label to the smallest label
// among its present neighbours, and pointer-jumping passes replace a label l
// by the label of element l.  Only thread e writes label e, and every value
// it reads is the label of an element of its own component, so labels only
// decrease and stay inside the component whatever the interleaving.  At the
// fixed point (no pass changes anything) every element carries the smallest
// element index of its component: the result is deterministic, and orders
// the components as the reference's search from each unvisited element in
// element order finds them.
#include "common.hpp"

namespace pyto {

__global__ __launch_bounds__(256) void k_cc_init(const double* __restrict__ x, double thr,
                                                 int64_t n, int32_t* __restrict__ lab) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    lab[e] = x[e] > thr ? (int32_t)e : -1;
}

__global__ __launch_bounds__(256) void k_cc_hook(int nx, int ny, int nz, int32_t* lab,
                                                 int* __restrict__ changed) {
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t l = lab[e];
    if (l < 0) continue;
    const int ix = (int)(e % nx), iy = (int)((e / nx) % ny), iz = (int)(e / ((int64_t)nx * ny));
    int32_t m = l;
    for (int dz = -1; dz <= 1; ++dz) {
      const int jz = iz + dz;
      if (jz < 0 || jz >= nz) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int jy = iy + dy;
        if (jy < 0 || jy >= ny) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int jx = ix + dx;
          if (jx < 0 || jx >= nx) continue;
          const int32_t lj = lab[jx + (int64_t)nx * (jy + (int64_t)ny * jz)];
          if (lj >= 0 && lj < m) m = lj;
        }
      }
    }
    if (m < l) {
      lab[e] = m;
      *changed = 1;
    }
  }
}

__global__ __launch_bounds__(256) void k_cc_jump(int64_t n, int32_t* lab,
                                                 int* __restrict__ changed) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t l = lab[e];
    if (l < 0) continue;
    const int32_t ll = lab[l];
    if (ll < l) {
      lab[e] = ll;
      *changed = 1;
    }
  }
}

__global__ void k_cc_clear(int* changed) { changed[0] = 0; }

}  // namespace pyto

using namespace pyto;

extern "C" {

int pyto_element_components(int nx, int ny, int nz, const double* x, double threshold,
                            int32_t* labels, int* rounds_host, void* stream) {
  PYTO_REQUIRE(x && labels, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, PYTO_EINVAL, "grid must be >= 1x1x1");
  const int64_t n = (int64_t)nx * ny * nz;
  PYTO_REQUIRE(n <= INT32_MAX, PYTO_EINVAL, "more than 2^31 - 1 elements");
  hipStream_t s = as_stream(stream);
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  int* changed = nullptr;
  PYTO_CHECK_HIP(hipMalloc(&changed, sizeof(int)));
  int host_changed = 1, rounds = 0, rc = PYTO_OK;
  hipLaunchKernelGGL(k_cc_init, dim3(blocks), dim3(256), 0, s, x, threshold, n, labels);
  // every round lowers at least one label while it reports a change, so
  // there are at most n rounds with a change before the fixed point
  while (host_changed && rounds <= n) {
    hipLaunchKernelGGL(k_cc_clear, dim3(1), dim3(1), 0, s, changed);
    for (int k = 0; k < 4; ++k) {
      hipLaunchKernelGGL(k_cc_hook, dim3(blocks), dim3(256), 0, s, nx, ny, nz, labels, changed);
      hipLaunchKernelGGL(k_cc_jump, dim3(blocks), dim3(256), 0, s, n, labels, changed);
      hipLaunchKernelGGL(k_cc_jump, dim3(blocks), dim3(256), 0, s, n, labels, changed);
    }
    rounds += 4;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&host_changed, changed, sizeof(int),
                                            hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("connected components: %s", hipGetErrorString(e));
      rc = PYTO_EHIP;
      break;
    }
  }
  hipFree(changed);
  if (rc == PYTO_OK && host_changed) {
    set_error("connected components did not reach a fixed point in %d rounds", rounds);
    rc = PYTO_ENOCONV;
  }
  if (rounds_host) *rounds_host = rounds;
  return rc;
}

}  // extern "C"
