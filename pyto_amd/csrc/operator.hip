// Grid operator: handle lifetime, SIMP densities, Jacobi diagonal and the
// matrix-free K*u of the voxel grid (FP64 and FP32).
//
// K*u kernel layout.  A workgroup of 256 threads is a 16x16 tile of element
// columns; thread (tx,ty) owns element column (15*bx-1+tx, 15*by-1+ty) and the
// node (ex, ey) at its lower-left corner.  The tile marches up a chunk of z:
// per element layer it stages the 17x17 node plane above in LDS, applies the
// element operator in the Hadamard basis (element.hpp), hands the
// contributions for the three neighbouring nodes of each plane to the
// neighbouring threads through LDS, and carries its own top-plane sum in
// registers to the next layer.  Tiles overlap by one element column so every
// owned node (tx,ty >= 1) is complete without atomics; a z chunk starts one
// element layer below its first owned node plane.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "matvec.hpp"

namespace pyto {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

int device_alloc(Handle* h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    return PYTO_ENOMEM;
  }
  h->workspace_bytes += (int64_t)bytes;
  return PYTO_OK;
}

// ---------------------------------------------------------------------------
// Density -> E_e (relative SIMP modulus), src/topopt_material_model.py:88.
// ---------------------------------------------------------------------------
__global__ void k_simp(const double* __restrict__ x, double* __restrict__ E64,
                       float* __restrict__ E32, double* __restrict__ xcopy, int64_t n,
                       double penal, double eps) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xi = x[i];
    const double e = eps + pow(xi, penal) * (1.0 - eps);
    E64[i] = e;
    E32[i] = (float)e;
    xcopy[i] = xi;
  }
}

// Jacobi: d(node,c) = KEdiag_c * sum of E_e over the node's elements;
// |d| < 1e-12 -> 1 (src/linear_solvers.py:84-85); fixed DOFs get 0 so that
// z = M r stays zero there.  nc values per node (3, or 1 for temperature).
// invd32 comes from the FP32 sum of E_e (jacobi_inv32), which esum32 keeps
// when given.
__global__ void k_inv_diag(const double* __restrict__ E64, const uint8_t* __restrict__ mask,
                           GridDims g, int nc, double kd0, double kd1, double kd2,
                           double* __restrict__ invd64, float* __restrict__ invd32,
                           double* __restrict__ diag_out, float* __restrict__ esum32,
                           double* __restrict__ esum64) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nnodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    node_coords(n, g, ix, iy, iz);
    double s = 0.0;
#pragma unroll
    for (int dz = -1; dz <= 0; ++dz)
#pragma unroll
      for (int dy = -1; dy <= 0; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 0; ++dx) {
          const int ex = ix + dx, ey = iy + dy, ez = iz + dz;
          if (ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny && ez >= 0 && ez < g.nz)
            s += E64[ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez)];
        }
    const uint8_t m = mask[n];
    const double kd[3] = {kd0, kd1, kd2};
    const float s32 = (float)s;
    if (esum32) esum32[n] = s32;
    if (esum64) esum64[n] = s;
    for (int c = 0; c < nc; ++c) {
      const bool fixed = (m >> c) & 1;
      double d = kd[c] * s;
      if (diag_out) diag_out[nc * n + c] = fixed ? 0.0 : d;
      if (invd64) {
        if (fabs(d) < 1e-12) d = 1.0;
        invd64[nc * n + c] = fixed ? 0.0 : 1.0 / d;
        float d32;
        invd32[nc * n + c] = jacobi_inv32((float)kd[c], s32, fixed, d32);
      }
    }
  }
}

// At least 4 waves per SIMD for FP32 and 3 for FP64 (128 / 168 VGPRs); 2
// for the FP64 brick element, which needs more than 168 VGPRs.
// TY: element rows of the tile, 16 TY threads per workgroup.
template <typename T, bool CUBE, int NC = 3, bool MASK_IN = true, int TY = kTile,
          int WPE = sizeof(T) == 4 ? 4 : (CUBE || NC == 1 ? 3 : 2)>
__global__ __launch_bounds__(kTile * TY) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_matvec(
    MatvecArgs<T> a) {
  matvec_tile<T, CUBE, NC, MASK_IN, 1, false, TY>(a, xcd_logical_block(blockIdx.x, gridDim.x),
                                                  blockIdx.x);
}

// The operator with the element contraction on the matrix cores: the dense
// 24x24 KE applied by v_mfma_f32_32x32x2_f32 (matvec.hpp, element_mfma).
// Kept for the measurement that compares it with the modal form.
template <int TY>
__global__ __launch_bounds__(kTile * TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_matvec_mfma(
    MatvecArgs<float> a) {
  matvec_tile<float, true, 3, true, 1, true, TY>(a, xcd_logical_block(blockIdx.x, gridDim.x),
                                                 blockIdx.x);
}

// tileflags[tile * sz + kp] = 1 when node plane kp of the tile's staging
// window holds a node with a fixed DOF.  Tile (bxi, byi) stages nodes
// x in [15 bxi - 1, 15 bxi + 15] and y in [oy byi - 1, oy byi + oy] with
// oy = rows - 1.  Racing stores write the same byte value.
__global__ __launch_bounds__(256) void k_tile_flags(const uint8_t* __restrict__ mask, GridDims g,
                                                    int tiles_x, int tiles_y, int oy,
                                                    uint8_t* __restrict__ flags) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= g.nnodes || mask[n] == 0) return;
  const int gx = (int)(n % g.sx), gy = (int)((n / g.sx) % g.sy);
  const int kp = (int)(n / ((int64_t)g.sx * g.sy));
  const int bx0 = max(0, (gx + kOwn - 1) / kOwn - 1), bx1 = min(tiles_x - 1, (gx + 1) / kOwn);
  const int by0 = max(0, (gy + oy - 1) / oy - 1), by1 = min(tiles_y - 1, (gy + 1) / oy);
  for (int by = by0; by <= by1; ++by)
    for (int bx = bx0; bx <= bx1; ++bx) flags[((int64_t)by * tiles_x + bx) * g.sz + kp] = 1;
}

// Fills tileflags for the handle's tile rows (the buffer holds the flags of
// the smallest tile, the largest tile count).
static hipError_t launch_tile_flags(Handle* h, hipStream_t s) {
  int tx = 0;
  const int64_t nt = tiles_for(h->g, h->tile_rows, &tx);
  const int oy = tile_own_y(h->tile_rows);
  hipError_t e = hipMemsetAsync(h->tileflags, 0, nt * h->g.sz, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_tile_flags, dim3((unsigned)((h->g.nnodes + 255) / 256)), dim3(256), 0, s,
                     h->fixmask, h->g, tx, (int)(nt / tx), oy, h->tileflags);
  return hipGetLastError();
}

int set_tile_rows(Handle* h, int rows, hipStream_t s) {
  PYTO_REQUIRE(rows == 8 || rows == 16 || rows == 32, PYTO_EINVAL,
               "operator tile rows must be 8, 16 or 32 (got %d)", rows);
  if (rows == h->tile_rows) return PYTO_OK;
  h->tile_rows = rows;
  PYTO_CHECK_HIP(launch_tile_flags(h, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  return PYTO_OK;
}

// Resident workgroups of a k_matvec variant on the whole device.
template <typename T, bool CUBE, int NC, bool MASK_IN, int TY>
static int matvec_capacity() {
  static int cap = -1;
  if (cap < 0) {
    int nb = 0, dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_matvec<T, CUBE, NC, MASK_IN, TY>,
                                                     kTile * TY, 0) != hipSuccess ||
        hipGetDeviceProperties(&prop, dev) != hipSuccess || nb <= 0)
      return 1024 * 16 / TY;
    cap = nb * prop.multiProcessorCount;
  }
  return cap;
}

template <typename T, int TY>
static int capacity_rows(const Handle* h) {
  const bool cube = h->g.hx == h->g.hy && h->g.hy == h->g.hz;
  return h->ncomp == 1 ? matvec_capacity<T, true, 1, true, TY>()
         : cube        ? matvec_capacity<T, true, 3, true, TY>()
                       : matvec_capacity<T, false, 3, true, TY>();
}

// z-chunk length: minimise (workgroup rounds) x (layers per workgroup,
// including the halo layer) for the occupancy of this kernel variant.
template <typename T>
static void matvec_launch_dims(const Handle* h, dim3& grid, int& zchunk, int& tiles_x,
                               int& ntiles) {
  const GridDims& g = h->g;
  const int64_t tiles = tiles_for(g, h->tile_rows, &tiles_x);
  const int cap = h->tile_rows == 8    ? capacity_rows<T, 8>(h)
                  : h->tile_rows == 32 ? capacity_rows<T, 32>(h)
                                       : capacity_rows<T, 16>(h);
  int best = 16;
  double best_cost = 1e300;
  for (int zc = 4; zc <= 62; ++zc) {   // planes kz0-1 .. kz1 fit one 64-bit mask
    const int64_t blocks = tiles * ((g.sz + zc - 1) / zc);
    if (blocks > kMaxPartials) continue;
    const int64_t rounds = (blocks + cap - 1) / cap;
    const double cost = (double)rounds * (std::min(zc, g.sz) + 1);
    if (cost < best_cost - 1e-9) {
      best_cost = cost;
      best = zc;
    }
  }
  zchunk = best;
  ntiles = (int)tiles;
  grid = dim3((unsigned)(tiles * ((g.sz + zchunk - 1) / zchunk)), 1, 1);
}

// k_matvec for the handle's element kind and tile rows.
template <typename T, bool MASK_IN, int TY>
static void launch_rows(const Handle* h, dim3 grid, hipStream_t s, const MatvecArgs<T>& a) {
  const bool cube = h->g.hx == h->g.hy && h->g.hy == h->g.hz;
  const dim3 block(kTile * TY);
  if (h->ncomp == 1)   // thermal: the modal weights already carry h
    hipLaunchKernelGGL((k_matvec<T, true, 1, MASK_IN, TY>), grid, block, 0, s, a);
  else if (cube)
    hipLaunchKernelGGL((k_matvec<T, true, 3, MASK_IN, TY>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_matvec<T, false, 3, MASK_IN, TY>), grid, block, 0, s, a);
}

template <typename T, bool MASK_IN>
static void launch_kernel(const Handle* h, dim3 grid, hipStream_t s, const MatvecArgs<T>& a) {
  switch (h->tile_rows) {
    case 8: launch_rows<T, MASK_IN, 8>(h, grid, s, a); break;
    case 32: launch_rows<T, MASK_IN, 32>(h, grid, s, a); break;
    default: launch_rows<T, MASK_IN, 16>(h, grid, s, a); break;
  }
}

template <typename T>
int matvec_blocks(const Handle* h) {
  dim3 grid;
  int zc, tx, nt;
  matvec_launch_dims<T>(h, grid, zc, tx, nt);
  return (int)grid.x;
}
template int matvec_blocks<double>(const Handle*);
template int matvec_blocks<float>(const Handle*);

template <typename T>
int launch_matvec(const Handle* h, const T* u, T* y, const T* Ee, const ModalConsts<T>& mc,
                  double* dot_out, const int* done, double* zero_vec, int zero_n,
                  hipStream_t s) {
  MatvecArgs<T> a;
  a.u = u;
  a.y = y;
  a.Ee = Ee;
  a.mask = h->fixmask;
  a.tileflags = h->tileflags;
  a.g = h->g;
  a.mc = mc;
  dim3 grid;
  matvec_launch_dims<T>(h, grid, a.zchunk, a.tiles_x, a.ntiles);
  a.partials = dot_out;
  a.done = done;
  a.zero_vec = zero_vec;
  a.zero_n = zero_n;
  a.ke32 = nullptr;
  launch_kernel<T, true>(h, grid, s, a);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

template int launch_matvec<double>(const Handle*, const double*, double*, const double*,
                                   const ModalConsts<double>&, double*, const int*, double*,
                                   int, hipStream_t);
template int launch_matvec<float>(const Handle*, const float*, float*, const float*,
                                  const ModalConsts<float>&, double*, const int*, double*, int,
                                  hipStream_t);

int ensure_solver_workspace(Handle* h) {
  if (h->r64) return PYTO_OK;
  const size_t nd = (size_t)(h->ncomp * h->g.nnodes);
  PYTO_TRY(device_alloc(h, (void**)&h->r64, nd * 8));
  PYTO_TRY(device_alloc(h, (void**)&h->q64, nd * 8));
  PYTO_TRY(device_alloc(h, (void**)&h->r32, nd * 4));
  PYTO_TRY(device_alloc(h, (void**)&h->q32, nd * 4));
  PYTO_TRY(device_alloc(h, (void**)&h->x32, nd * 4));
  PYTO_TRY(device_alloc(h, (void**)&h->b32, nd * 4));
  PYTO_TRY(device_alloc(h, (void**)&h->pring64, kDirSlots * ring_stride(nd) * 8));
  PYTO_TRY(device_alloc(h, (void**)&h->pring32, kDirSlots32 * ring_stride(nd) * 4));
  PYTO_TRY(device_alloc(h, (void**)&h->alring, kDirSlots * 8));
  return PYTO_OK;
}

int coarse_setup(Handle* h, hipStream_t s);  // coarse.hip
void release_blas(Handle* h);                // coarse.hip

static void host_ke(const Handle* h, double* ke) {
  // Columns of KE from unit vectors through the same modal apply.
  const ModalConsts<double>& mc = h->mc64;
  for (int col = 0; col < 24; ++col) {
    double u[8][3] = {};
    u[col / 3][col % 3] = 1.0;
    double f[8][3];
    element_apply<double>(u, 1.0, mc, f);
    for (int r = 0; r < 24; ++r) ke[r * 24 + col] = f[r / 3][r % 3];
  }
}

// The same for the 8x8 thermal KE_T.
static void host_ke_thermal(const ModalConsts<double>& mc, double* ke) {
  for (int col = 0; col < 8; ++col) {
    double u[8][1] = {};
    u[col][0] = 1.0;
    double f[8][1];
    element_apply_thermal<double>(u, 1.0, mc, f);
    for (int r = 0; r < 8; ++r) ke[r * 8 + col] = f[r][0];
  }
}

// Handle with grid, fixed mask and per-density buffers for nc values per
// node; the physics constants (mc64/mc32, kediag, SIMP law) are set by the
// caller before this runs.
static int create_common(pyto_handle* h, int nelx, int nely, int nelz, double hx, double hy,
                         double hz, int nc, const uint8_t* fixed_mask, hipStream_t s) {
  const int64_t nn = (int64_t)(nelx + 1) * (nely + 1) * (nelz + 1);
  hipGetDevice(&h->device);
  h->ncomp = nc;
  h->g.nx = nelx; h->g.ny = nely; h->g.nz = nelz;
  h->g.sx = nelx + 1; h->g.sy = nely + 1; h->g.sz = nelz + 1;
  h->g.nnodes = nn;
  h->g.nelems = (int64_t)nelx * nely * nelz;
  h->g.hx = hx; h->g.hy = hy; h->g.hz = hz;
  const int64_t ne = h->g.nelems;
  int rc = PYTO_OK;
  const int64_t ntiles = tiles_for(h->g, 8);   // the largest tile count
  if ((rc = device_alloc(h, (void**)&h->fixmask, nn)) ||
      (rc = device_alloc(h, (void**)&h->tileflags, ntiles * h->g.sz)) ||
      (rc = device_alloc(h, (void**)&h->Ee64, ne * 8)) ||
      (rc = device_alloc(h, (void**)&h->Ee32, ne * 4)) ||
      (rc = device_alloc(h, (void**)&h->x64, ne * 8)) ||
      (rc = device_alloc(h, (void**)&h->invd64, nc * nn * 8)) ||
      (rc = device_alloc(h, (void**)&h->invd32, nc * nn * 4)) ||
      (nc == 3 && (rc = device_alloc(h, (void**)&h->esum32, nn * 4))) ||
      (nc == 3 && (rc = device_alloc(h, (void**)&h->esum64, nn * 8))) ||
      (rc = device_alloc(h, (void**)&h->partials, 2 * kMaxPartials * 8)) ||
      (rc = device_alloc(h, (void**)&h->counters, kNumCounters * sizeof(unsigned))) ||
      (rc = device_alloc(h, (void**)&h->scal, kCtlBytes)))
    return rc;
  // scalars and flags in one block (device and host), so that the solver
  // reads both back with one copy
  h->flags = reinterpret_cast<int*>(h->scal + kNumScalars);
  if (hipHostMalloc((void**)&h->host_scal, kCtlBytes, hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    return PYTO_ENOMEM;
  }
  h->host_flags = reinterpret_cast<int*>(h->host_scal + kNumScalars);
  if (hipMemcpyAsync(h->fixmask, fixed_mask, nn, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      launch_tile_flags(h, s) != hipSuccess ||
      hipMemsetAsync(h->counters, 0, kNumCounters * sizeof(unsigned), s) != hipSuccess ||
      hipMemsetAsync(h->scal, 0, kNumScalars * 8, s) != hipSuccess ||
      hipMemsetAsync(h->flags, 0, kNumFlags * sizeof(int), s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("initialising the handle failed: %s", hipGetErrorString(hipGetLastError()));
    return PYTO_EHIP;
  }
  return PYTO_OK;
}

// E_e and the Jacobi inverse diagonal of densities x: the element part of
// pyto_set_density (the coarse matrix follows in coarse.hip).
int density_elements(Handle* h, const double* x, hipStream_t s) {
  const int64_t ne = h->g.nelems;
  int blocks = (int)std::min<int64_t>((ne + 255) / 256, 4096);
  hipLaunchKernelGGL(k_simp, dim3(blocks), dim3(256), 0, s, x, h->Ee64, h->Ee32, h->x64, ne,
                     h->penal, h->eps_void);
  PYTO_CHECK_HIP(hipGetLastError());
  blocks = (int)std::min<int64_t>((h->g.nnodes + 255) / 256, 4096);
  hipLaunchKernelGGL(k_inv_diag, dim3(blocks), dim3(256), 0, s, h->Ee64, h->fixmask, h->g,
                     h->ncomp, h->kediag[0], h->kediag[1], h->kediag[2], h->invd64, h->invd32,
                     (double*)nullptr, h->esum32, h->esum64);
  PYTO_CHECK_HIP(hipGetLastError());
  h->density_set = true;
  return PYTO_OK;
}

}  // namespace pyto

using namespace pyto;

extern "C" {

const char* pyto_last_error(void) { return g_last_error.c_str(); }
int pyto_version(void) { return 1; }

int pyto_ke_hex8(double E, double nu, double hx, double hy, double hz, double* ke_host) {
  PYTO_REQUIRE(ke_host, PYTO_EINVAL, "ke_host is null");
  PYTO_REQUIRE(hx > 0 && hy > 0 && hz > 0, PYTO_EINVAL, "element size must be positive");
  PYTO_REQUIRE(nu > -1.0 && nu < 0.5, PYTO_EINVAL, "Poisson ratio must be in (-1, 0.5)");
  Handle tmp;
  tmp.mc64 = make_modal_consts<double>(E, nu, hx, hy, hz);
  host_ke(&tmp, ke_host);
  return PYTO_OK;
}

int pyto_ke_hex8_thermal(double k0, double hx, double hy, double hz, double* ke_host) {
  PYTO_REQUIRE(ke_host, PYTO_EINVAL, "ke_host is null");
  PYTO_REQUIRE(hx > 0 && hy > 0 && hz > 0, PYTO_EINVAL, "element size must be positive");
  host_ke_thermal(make_thermal_consts<double>(k0, hx, hy, hz), ke_host);
  return PYTO_OK;
}

int pyto_create(pyto_handle** out, int nelx, int nely, int nelz, double hx, double hy,
                double hz, double E0, double nu, double penal, double eps_void,
                const uint8_t* fixed_mask, void* stream) {
  PYTO_REQUIRE(out, PYTO_EINVAL, "out is null");
  *out = nullptr;
  PYTO_REQUIRE(nelx >= 1 && nely >= 1 && nelz >= 1, PYTO_EINVAL, "grid must be >= 1x1x1");
  PYTO_REQUIRE(hx > 0 && hy > 0 && hz > 0, PYTO_EINVAL, "element size must be positive");
  PYTO_REQUIRE(nu > -1.0 && nu < 0.5, PYTO_EINVAL, "Poisson ratio must be in (-1, 0.5)");
  PYTO_REQUIRE(fixed_mask, PYTO_EINVAL, "fixed_mask is null");
  const int64_t nn = (int64_t)(nelx + 1) * (nely + 1) * (nelz + 1);
  // the operator addresses DOF vectors with 32-bit byte offsets (FP64: 24 B per node)
  PYTO_REQUIRE(24 * nn <= (int64_t)UINT32_MAX, PYTO_EINVAL,
               "grid too large: %lld nodes, at most %lld", (long long)nn,
               (long long)(UINT32_MAX / 24));
  pyto_handle* h = new pyto_handle();
  h->E0 = E0; h->nu = nu; h->penal = penal; h->eps_void = eps_void;
  h->mc64 = make_modal_consts<double>(E0, nu, hx, hy, hz);
  h->mc32 = make_modal_consts<float>(E0, nu, hx, hy, hz);
  double ke[576];
  host_ke(h, ke);
  for (int c = 0; c < 3; ++c) h->kediag[c] = ke[c * 24 + c];
  const int rc = create_common(h, nelx, nely, nelz, hx, hy, hz, 3, fixed_mask, as_stream(stream));
  if (rc != PYTO_OK) {
    pyto_destroy(h);
    return rc;
  }
  *out = h;
  return PYTO_OK;
}

int pyto_create_thermal(pyto_handle** out, int nelx, int nely, int nelz, double hx, double hy,
                        double hz, double k0, double penal, double k_void,
                        const uint8_t* fixed_mask, void* stream) {
  PYTO_REQUIRE(out, PYTO_EINVAL, "out is null");
  *out = nullptr;
  PYTO_REQUIRE(nelx >= 1 && nely >= 1 && nelz >= 1, PYTO_EINVAL, "grid must be >= 1x1x1");
  PYTO_REQUIRE(hx > 0 && hy > 0 && hz > 0, PYTO_EINVAL, "element size must be positive");
  PYTO_REQUIRE(k0 > 0, PYTO_EINVAL, "conductivity must be positive");
  PYTO_REQUIRE(fixed_mask, PYTO_EINVAL, "fixed_mask is null");
  const int64_t nn = (int64_t)(nelx + 1) * (nely + 1) * (nelz + 1);
  PYTO_REQUIRE(8 * nn <= (int64_t)UINT32_MAX, PYTO_EINVAL,
               "grid too large: %lld nodes, at most %lld", (long long)nn,
               (long long)(UINT32_MAX / 8));
  pyto_handle* h = new pyto_handle();
  h->E0 = k0; h->nu = 0.0; h->penal = penal; h->eps_void = k_void;
  h->mc64 = make_thermal_consts<double>(k0, hx, hy, hz);
  h->mc32 = make_thermal_consts<float>(k0, hx, hy, hz);
  double ke[64];
  host_ke_thermal(h->mc64, ke);
  h->kediag[0] = ke[0];
  const int rc = create_common(h, nelx, nely, nelz, hx, hy, hz, 1, fixed_mask, as_stream(stream));
  if (rc != PYTO_OK) {
    pyto_destroy(h);
    return rc;
  }
  *out = h;
  return PYTO_OK;
}

int pyto_destroy(pyto_handle* h) {
  if (!h) return PYTO_OK;
  void* ptrs[] = {h->fixmask, h->tileflags, h->Ee64, h->Ee32, h->x64, h->invd64, h->invd32,
                  h->esum32, h->esum64,
                  h->node_group, h->gmask, h->centers, h->sinfo, h->chunks, h->cpart,
                  h->gptr, h->gidx, h->gnodes, h->nchunk,
                  h->nchunk_grp, h->npart, h->nptr, h->nidx, h->Ec, h->Einv, h->cwork, h->ct,
                  h->cmu, h->cres, h->Epk, h->spart, h->r64, h->q64, h->r32, h->q32, h->x32,
                  h->b32, h->partials, h->counters, h->scal /* and flags */, h->elem_buf,
                  h->gal_off, h->gal_part, h->gal_pairs, h->gal_pptr, h->gal_pidx,
                  h->pwave, h->pslot, h->pnodes, h->pcen, h->pE64, h->pE32, h->stencil,
                  h->ke32, h->pring64, h->pring32, h->alring};
  for (void* p : ptrs)
    if (p) hipFree(p);
  release_blas(h);
  for (hipEvent_t e : h->ev_ctl)
    if (e) hipEventDestroy(e);
  if (h->host_ctl) hipHostFree(h->host_ctl);
  if (h->host_scal) hipHostFree(h->host_scal);   // holds host_flags too
  delete h;
  return PYTO_OK;
}

int pyto_info(pyto_handle* h, int64_t* nnodes, int64_t* nelems, int64_t* ndof,
              int64_t* workspace_bytes) {
  PYTO_REQUIRE(h, PYTO_EINVAL, "null handle");
  if (nnodes) *nnodes = h->g.nnodes;
  if (nelems) *nelems = h->g.nelems;
  if (ndof) *ndof = h->ncomp * h->g.nnodes;
  if (workspace_bytes) *workspace_bytes = h->workspace_bytes;
  return PYTO_OK;
}

int pyto_set_tile_rows(pyto_handle* h, int rows, void* stream) {
  PYTO_REQUIRE(h, PYTO_EINVAL, "null handle");
  return set_tile_rows(h, rows, as_stream(stream));
}

int pyto_get_tile_rows(const pyto_handle* h) { return h ? h->tile_rows : PYTO_EINVAL; }

int pyto_set_density(pyto_handle* h, const double* x, void* stream) {
  PYTO_REQUIRE(h && x, PYTO_EINVAL, "null argument");
  hipStream_t s = as_stream(stream);
  PYTO_TRY(density_elements(h, x, s));
  if (h->ngroups > 0) PYTO_TRY(coarse_setup(h, s));
  return PYTO_OK;
}

int pyto_diagonal_f64(pyto_handle* h, double* d, void* stream) {
  PYTO_REQUIRE(h && d, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(h->density_set, PYTO_ESTATE, "pyto_set_density must be called first");
  int blocks = (int)std::min<int64_t>((h->g.nnodes + 255) / 256, 4096);
  hipLaunchKernelGGL(k_inv_diag, dim3(blocks), dim3(256), 0, as_stream(stream), h->Ee64,
                     h->fixmask, h->g, h->ncomp, h->kediag[0], h->kediag[1], h->kediag[2],
                     (double*)nullptr, (float*)nullptr, d, (float*)nullptr, (double*)nullptr);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_matvec_f64(pyto_handle* h, const double* u, double* y, void* stream) {
  PYTO_REQUIRE(h && u && y, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(u != y, PYTO_EINVAL, "u and y must not alias");
  PYTO_REQUIRE(h->density_set, PYTO_ESTATE, "pyto_set_density must be called first");
  return launch_matvec<double>(h, u, y, h->Ee64, h->mc64, nullptr, nullptr, nullptr, 0,
                               as_stream(stream));
}

// y = K_fc g (g's fixed entries kept, y's fixed entries zeroed), FP64.
int pyto_matvec_lift_f64(pyto_handle* h, const double* g, double* y, void* stream) {
  PYTO_REQUIRE(h && g && y, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(g != y, PYTO_EINVAL, "g and y must not alias");
  PYTO_REQUIRE(h->density_set, PYTO_ESTATE, "pyto_set_density must be called first");
  MatvecArgs<double> a;
  a.u = g;
  a.y = y;
  a.Ee = h->Ee64;
  a.mask = h->fixmask;
  a.tileflags = h->tileflags;
  a.g = h->g;
  a.mc = h->mc64;
  dim3 grid;
  matvec_launch_dims<double>(h, grid, a.zchunk, a.tiles_x, a.ntiles);
  a.partials = nullptr;
  a.done = nullptr;
  a.zero_vec = nullptr;
  a.zero_n = 0;
  a.ke32 = nullptr;
  launch_kernel<double, false>(h, grid, as_stream(stream), a);
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

int pyto_matvec_f32(pyto_handle* h, const float* u, float* y, void* stream) {
  PYTO_REQUIRE(h && u && y, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(u != y, PYTO_EINVAL, "u and y must not alias");
  PYTO_REQUIRE(h->density_set, PYTO_ESTATE, "pyto_set_density must be called first");
  return launch_matvec<float>(h, u, y, h->Ee32, h->mc32, nullptr, nullptr, nullptr, 0,
                              as_stream(stream));
}

int pyto_matvec_f32_mfma(pyto_handle* h, const float* u, float* y, void* stream) {
  PYTO_REQUIRE(h && u && y, PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(u != y, PYTO_EINVAL, "u and y must not alias");
  PYTO_REQUIRE(h->ncomp == 3, PYTO_EINVAL, "the MFMA operator variant is structural");
  PYTO_REQUIRE(h->density_set, PYTO_ESTATE, "pyto_set_density must be called first");
  hipStream_t s = as_stream(stream);
  if (!h->ke32) {
    double ke[576];
    host_ke(h, ke);
    float kf[576];
    for (int i = 0; i < 576; ++i) kf[i] = (float)ke[i];
    PYTO_TRY(device_alloc(h, (void**)&h->ke32, sizeof(kf)));
    PYTO_CHECK_HIP(hipMemcpyAsync(h->ke32, kf, sizeof(kf), hipMemcpyHostToDevice, s));
    PYTO_CHECK_HIP(hipStreamSynchronize(s));
  }
  MatvecArgs<float> a;
  a.u = u;
  a.y = y;
  a.Ee = h->Ee32;
  a.mask = h->fixmask;
  a.tileflags = h->tileflags;
  a.g = h->g;
  a.mc = h->mc32;
  dim3 grid;
  matvec_launch_dims<float>(h, grid, a.zchunk, a.tiles_x, a.ntiles);
  a.partials = nullptr;
  a.done = nullptr;
  a.zero_vec = nullptr;
  a.zero_n = 0;
  a.ke32 = h->ke32;
  switch (h->tile_rows) {
    case 8: hipLaunchKernelGGL(k_matvec_mfma<8>, grid, dim3(kTile * 8), 0, s, a); break;
    case 32: hipLaunchKernelGGL(k_matvec_mfma<32>, grid, dim3(kTile * 32), 0, s, a); break;
    default: hipLaunchKernelGGL(k_matvec_mfma<16>, grid, dim3(kTile * 16), 0, s, a); break;
  }
  PYTO_CHECK_HIP(hipGetLastError());
  return PYTO_OK;
}

}  // extern "C"
