// Matrix-free SIMP-scaled hex operator y = K_ff u: the workgroup body shared
// by the single-operator kernel (operator.hip) and the batched kernel
// (batch.hip).
#pragma once

#include "element.hpp"

namespace pyto {

// ---------------------------------------------------------------------------
// Matrix-free y = K_ff u.
// ---------------------------------------------------------------------------
constexpr int kTile = 16;
constexpr int kOwn = 15;
constexpr int kPl = 17;

// Node rows a tile of `rows` element rows owns (x: always kOwn).
constexpr int tile_own_y(int rows) { return rows - 1; }

// Tiles of the operator for tile rows `rows` (and the x count).
inline int64_t tiles_for(const GridDims& g, int rows, int* tiles_x = nullptr) {
  const int tx = (g.sx + kOwn - 1) / kOwn, oy = tile_own_y(rows);
  if (tiles_x) *tiles_x = tx;
  return (int64_t)tx * ((g.sy + oy - 1) / oy);
}

template <typename T>
struct MatvecArgs {
  const T* u;
  T* y;
  const T* Ee;
  const uint8_t* mask;
  GridDims g;
  ModalConsts<T> mc;
  const uint8_t* tileflags;  // per (tile, node plane): window holds a fixed node
  int zchunk;
  int tiles_x;
  int ntiles;
  double* partials;    // per-block partials of u.y (null: no dot)
  const int* done;     // skip when *done != 0 (null: never)
  double* zero_vec;    // zeroed by block 0 (null: none)
  int zero_n;
  const float* ke32;   // dense 24x24 KE (row-major FP32), MFMA variant only
};

// Dense-KE element contraction on the matrix cores (FP32 operator variant):
// f = KE (E u) for the 64 elements of a wave as two v_mfma_f32_32x32x2_f32
// sequences of 12 K-steps (rows: the 24 DOFs of KE padded to 32; columns: 32
// elements).  A holds KE[l % 32][2s + l / 32] (ka, 12 values per lane).  The
// B operand of set 0 (elements of lanes 0-31) wants lane l + 32 to hold DOF
// 2s+1 of element l, and set 1 the mirror, which v_permlane32_swap forms
// from (DOF 2s, DOF 2s+1) of every lane in one instruction; the 32x32
// accumulators hold rows 8q + 4h + m (h = lane / 32) and are swapped back
// the same way, so every lane ends with the 24 forces of its own element.
typedef float mfma_v16f __attribute__((ext_vector_type(16)));
__device__ __forceinline__ void element_mfma(const float u[8][3], float E, const float ka[12],
                                             float f[8][3]) {
  float ue[24];
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) ue[3 * j + c] = E * u[j][c];
  mfma_v16f acc0 = {}, acc1 = {};
#pragma unroll
  for (int s = 0; s < 12; ++s) {
    const auto bw = __builtin_amdgcn_permlane32_swap(__float_as_uint(ue[2 * s]),
                                                     __float_as_uint(ue[2 * s + 1]), false, false);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(ka[s], __uint_as_float(bw[0]), acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(ka[s], __uint_as_float(bw[1]), acc1, 0, 0, 0);
  }
  float ra[16], rb[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const auto w = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc0[r]),
                                                    __float_as_uint(acc1[r]), false, false);
    ra[r] = __uint_as_float(w[0]);   // rows 8q + m   (m < 4) of this lane's element
    rb[r] = __uint_as_float(w[1]);   // rows 8q + 4 + m
  }
#pragma unroll
  for (int d = 0; d < 24; ++d) {
    const int q = d / 8, m = d % 8;
    f[d / 3][d % 3] = m < 4 ? ra[4 * q + m] : rb[4 * q + m - 4];
  }
}

// Lane (tx-1)'s value within the 16-lane row (0 for tx = 0): DPP row_shr:1
// with bound_ctrl, which reads 0 where the row has no source lane, so no
// "old" value has to be materialised and the add can take the DPP operand.
__device__ __forceinline__ float dpp_shr1(float v) {
  const int b = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(b, b, 0x111, 0xF, 0xF, true));
}
__device__ __forceinline__ double dpp_shr1(double v) {
  const int l = __double2loint(v), h = __double2hiint(v);
  const int lo = __builtin_amdgcn_update_dpp(l, l, 0x111, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(h, h, 0x111, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Issue the loads of one node of the 17x17 staging plane without consuming
// them: raw values in a[], the node's fixed bits (pending load) in ml, and
// 7 in mo when the node is outside the grid (the address is then clamped
// into the grid).  The masking happens when the plane is stored to LDS, so
// the loads stay in flight across a whole layer of element work.
template <typename T, int NC>
__device__ __forceinline__ void issue_node(T a[NC], int& mo, int& ml, const T* __restrict__ u,
                                           const uint8_t* __restrict__ mask, uint32_t nxy,
                                           uint32_t nxyb, bool inxy, uint32_t plane, int kp,
                                           int sz, bool any_fixed) {
  // nxy: the node's (clamped) index within its plane, nxyb the same in bytes
  // of a node; inxy: x, y inside the grid.  Only the plane offset changes as
  // the workgroup marches in z.
  // Indices and byte offsets are 32-bit (the handle limits the grid so that
  // every byte offset of a DOF vector fits), so the loads take the vector
  // offset form against the kernel argument's base address; the per-thread
  // part of the byte offset is a constant and the plane's part uniform.
  constexpr uint32_t kNode = NC * sizeof(T);
  const bool inb = inxy && kp >= 0 && kp < sz;
  const uint32_t cz = (uint32_t)min(max(kp, 0), sz - 1);
  uint32_t off = nxyb + (plane * kNode) * cz;
  // an opaque 32-bit offset: keeps the compiler from folding the constant
  // part into a 64-bit per-thread pointer (one 32-bit add per load instead
  // of a 64-bit one)
  asm volatile("" : "+v"(off));
  const T* p = reinterpret_cast<const T*>(reinterpret_cast<const char*>(u) + off);
#pragma unroll
  for (int c = 0; c < NC; ++c) a[c] = p[c];
  mo = inb ? 0 : 7;
  ml = 0;
  if (any_fixed) ml = mask[nxy + plane * cz];
}

// Layer loop, per element layer k (nodes k and k+1):
//   store plane k+1 (prefetched into registers PF layers earlier) to LDS,
//   issue the global loads of plane k+1+PF, one barrier, finish node plane
//   k-1 (owners: its x sum and the y exchange of the row below), then read
//   the corners of plane k+1, apply element (ex, ey, k), finish the x sum of
//   plane k and write its y exchange.
// x-neighbour sums use DPP within the 16-lane rows (tx is the lane's row
// position); only the y-neighbour sums go through LDS.
// One workgroup's share of y = K u: logical block lb of the operator's
// (tile, z-chunk) grid; its u.y partial goes to a.partials[pidx].  Shared by
// the single-operator kernel and the batched kernel.
//
// NC is the number of values per node: 3 (displacement, structural KE) or 1
// (temperature, thermal KE_T); the tile structure is the same.  MASK_IN =
// false keeps the fixed entries of u (only y's fixed entries are zeroed):
// that is K_fc g for lifting non-zero Dirichlet values.
// One node plane's loads in flight: this thread's two staging nodes (raw
// values, outside-grid and fixed bits) and the E of the element layer below.
template <typename T, int NC>
struct PlaneStage {
  T pv0[NC], pv1[NC];
  int mo0, ml0, mo1, ml1;
  T e;
};

// PF: node planes in flight ahead of the one being stored.  One suffices: a
// second plane in flight (11 more VGPRs in FP32) measured no faster.
// TY: element rows of the tile (16 x TY element columns, 16 TY threads; it
// owns 15 x (TY - 1) nodes per plane).
template <typename T, bool CUBE, int NC = 3, bool MASK_IN = true, int PF = 1, bool MFMA = false,
          int TY = kTile>
__device__ __forceinline__ void matvec_tile(const MatvecArgs<T>& a, int lb, int pidx) {
  static_assert(!MFMA || (sizeof(T) == 4 && NC == 3), "the MFMA variant is FP32 structural");
  static_assert(TY >= 4 && TY <= 64 && TY % 4 == 0, "tile rows: whole waves");
  constexpr int kThreads = kTile * TY;
  if (a.done && *a.done) return;
  __shared__ T pl[2][TY + 1][kPl][NC];
  __shared__ T xc[2][TY][kTile][NC];   // odd stride (NC = 3 or 1): no bank conflicts between rows
  __shared__ T es[2][TY][kTile];   // E of the element layer below plane b
  __shared__ uint8_t pm[2][TY + 1][kPl];  // fixed bits of the staged nodes
  __shared__ double red[kThreads / 64];
  const GridDims& g = a.g;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int tile = lb % a.ntiles;
  const int bxi = tile % a.tiles_x, byi = tile / a.tiles_x;
  const int x0 = bxi * kOwn - 1, y0 = byi * (TY - 1) - 1;
  const int ex = x0 + tx, ey = y0 + ty;
  const int kz0 = (lb / a.ntiles) * a.zchunk;
  const int kz1 = min(kz0 + a.zchunk, g.sz);
  // fixed-node flags of the chunk's node planes kz0-1 .. kz1 (<= 64 planes,
  // zchunk <= 62) as one wave-uniform bit mask: no memory access in the loop
  uint64_t pfix;
  {
    const int kp = kz0 - 1 + (int)(threadIdx.x & 63);
    const bool f = kp >= 0 && kp <= kz1 && kp < g.sz &&
                   a.tileflags[(int64_t)tile * g.sz + kp] != 0;
    pfix = __ballot(f);
  }
  auto plane_fixed = [&](int kp) { return ((pfix >> (kp - kz0 + 1)) & 1) != 0; };
  const bool col_valid = ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny;
  const bool owner = tx >= 1 && ty >= 1 && ex < g.sx && ey < g.sy;

  if (a.zero_vec && pidx == 0)
    for (int i = threadIdx.x; i < a.zero_n; i += kThreads) a.zero_vec[i] = 0.0;

  // staging-plane entries of this thread: e0 = tid, e1 = tid + 16 TY (< 17 (TY + 1))
  const int e0 = threadIdx.x, e1 = threadIdx.x + kThreads;
  const bool has1 = e1 < kPl * (TY + 1);
  const int lx0 = e0 % kPl, ly0 = e0 / kPl;
  const int lx1 = has1 ? e1 % kPl : 0, ly1 = has1 ? e1 / kPl : 0;
  PlaneStage<T, NC> st[PF];
#pragma unroll
  for (int q = 0; q < PF; ++q) {
#pragma unroll
    for (int c = 0; c < NC; ++c) st[q].pv1[c] = T(0);
    st[q].mo0 = st[q].ml0 = st[q].ml1 = 0;
    st[q].mo1 = 7;
    st[q].e = T(0);
  }
  // E of element layer kp-1, loaded with node plane kp and staged with it
  uint32_t ecolb = (uint32_t)(min(max(ex, 0), g.nx - 1) + g.nx * min(max(ey, 0), g.ny - 1)) *
                   (uint32_t)sizeof(T);
  asm volatile("" : "+v"(ecolb));
  const uint32_t elayer = (uint32_t)(g.nx * g.ny);
  // fixed bits of this thread's own node, read from the staged plane when
  // it becomes the top plane: mq byte 0 = newest plane
  uint32_t mq = 0;
  // in-plane parts of the two staged nodes' indices (constant over z)
  const uint32_t plane = (uint32_t)(g.sx * g.sy);
  const int gx0 = x0 + lx0, gy0 = y0 + ly0, gx1 = x0 + lx1, gy1 = y0 + ly1;
  const bool inxy0 = gx0 >= 0 && gx0 < g.sx && gy0 >= 0 && gy0 < g.sy;
  const bool inxy1 = gx1 >= 0 && gx1 < g.sx && gy1 >= 0 && gy1 < g.sy;
  const uint32_t nxy0 = (uint32_t)(min(max(gx0, 0), g.sx - 1) + g.sx * min(max(gy0, 0), g.sy - 1));
  const uint32_t nxy1 = (uint32_t)(min(max(gx1, 0), g.sx - 1) + g.sx * min(max(gy1, 0), g.sy - 1));
  // the same in bytes, opaque to the compiler so that it adds the plane's
  // uniform byte offset instead of multiplying the node index per load
  uint32_t nxyb0 = nxy0 * (uint32_t)(NC * sizeof(T)), nxyb1 = nxy1 * (uint32_t)(NC * sizeof(T));
  asm volatile("" : "+v"(nxyb0), "+v"(nxyb1));
  auto fetch = [&](int kp, PlaneStage<T, NC>& s) {
    const bool fx = plane_fixed(kp);
    issue_node<T, NC>(s.pv0, s.mo0, s.ml0, a.u, a.mask, nxy0, nxyb0, inxy0, plane, kp, g.sz, fx);
    if (has1)
      issue_node<T, NC>(s.pv1, s.mo1, s.ml1, a.u, a.mask, nxy1, nxyb1, inxy1, plane, kp, g.sz,
                        fx);
    const int kl = min(max(kp - 1, 0), g.nz - 1);
    uint32_t eoff = ecolb + (elayer * (uint32_t)sizeof(T)) * (uint32_t)kl;
    asm volatile("" : "+v"(eoff));
    s.e = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a.Ee) + eoff);
  };
  // value v with bit c of the fixed bits i set -> 0 (FP32: the sign-extended
  // bit clears the word, two operations instead of a compare and a select)
  auto masked = [](T v, int i, int c) -> T {
    if constexpr (sizeof(T) == 4) {
      int clear = __builtin_amdgcn_sbfe(i, c, 1);   // -1 when the bit is set
      asm volatile("" : "+v"(clear));   // keeps the AND from becoming a compare and select
      return __int_as_float(__float_as_int(v) & ~clear);
    } else {
      return ((i >> c) & 1) ? T(0) : v;
    }
  };
  auto store = [&](int b, const PlaneStage<T, NC>& s) {
    const int m0 = s.mo0 | s.ml0, m1 = s.mo1 | s.ml1;
    const int i0 = MASK_IN ? m0 : s.mo0, i1 = MASK_IN ? m1 : s.mo1;
#pragma unroll
    for (int c = 0; c < NC; ++c) pl[b][ly0][lx0][c] = masked(s.pv0[c], i0, c);
    pm[b][ly0][lx0] = (uint8_t)m0;
    if (has1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) pl[b][ly1][lx1][c] = masked(s.pv1[c], i1, c);
      pm[b][ly1][lx1] = (uint8_t)m1;
    }
    es[b][ty][tx] = s.e;
  };

  // FP32: x/y butterfly stages of the element's bottom face, carried from
  // the layer below (its top face), and u at this thread's node of that face.
  // FP64 keeps the corner values instead: carrying the face stages costs it
  // 8 VGPRs more and one wave per SIMD.
  // (one temperature per node: the faces are small in both precisions)
  constexpr bool kFaces = (sizeof(T) == 4 || NC == 1) && !MFMA;
  // MFMA variant: this lane's A operands (KE entries) and the top-face
  // corner forces of the last element layer
  float ka[12];
  float ctop[4][3];
  if constexpr (MFMA) {
    const int l = threadIdx.x & 63, row = l & 31;
#pragma unroll
    for (int q = 0; q < 12; ++q) ka[q] = row < 24 ? a.ke32[row * 24 + 2 * q + (l >> 5)] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) ctop[j][c] = 0.0f;
  }
  Face<T, NC> bface;
  T bot[4][NC];
  T ucur[NC];
  fetch(kz0 - 1, st[0]);
  store(0, st[0]);
  // planes kz0 .. kz0 + PF - 1 in flight (with them E of layers kz0-1 ..)
#pragma unroll
  for (int q = 0; q < PF; ++q) fetch(min(kz0 + q, kz1), st[q]);
  __syncthreads();
  {
    T c0[NC], c1[NC], c2[NC], c3[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      c0[c] = pl[0][ty][tx][c];
      c1[c] = pl[0][ty][tx + 1][c];
      c2[c] = pl[0][ty + 1][tx + 1][c];
      c3[c] = pl[0][ty + 1][tx][c];
      ucur[c] = c0[c];
      bot[0][c] = c0[c];
      bot[1][c] = c1[c];
      bot[2][c] = c2[c];
      bot[3][c] = c3[c];
    }
    if (kFaces) face_xy<T, NC>(c0, c1, c2, c3, bface);
  }
  mq = pm[0][ty][tx];
  // top face of the previous element layer (z stage of its transposed
  // butterfly), added to this layer's bottom face: each node plane's forces
  // go through the y and x stages once
  Face<T, NC> btop;
  T xs[NC], ys[NC];    // pending sums of the last finished plane (x sum, y exchange)
  T ubot[NC];          // u at this thread's node of that plane
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    btop.s[c] = btop.sy[c] = btop.dx[c] = btop.dxy[c] = T(0);
    xs[c] = ys[c] = ubot[c] = T(0);
  }
  // this thread's share of u.y over its owned nodes, in T: a thread sums at
  // most 3 * zchunk products; the block sum and everything after is FP64
  T dot = T(0);
  // this thread's own node in its plane (owners only write it)
  uint32_t own_xyb = (uint32_t)(ex + g.sx * ey) * (uint32_t)(NC * sizeof(T));
  asm volatile("" : "+v"(own_xyb));
  // one element layer k (see the comment above the kernel); false once the
  // last node plane of the chunk is finished.  Called twice per loop trip so
  // that the values carried from layer to layer alternate between register
  // sets instead of being copied.
  auto layer = [&](int k, PlaneStage<T, NC>& s) -> bool {
    const int b = (k - kz0 + 2) & 1;   // LDS plane buffer of node plane k+1
    // y-exchange buffer of plane k-1 (written in layer k-1).  Parities
    // relative to kz0 are fixed for each of the two calls of a loop trip,
    // so the LDS offsets are constants.
    const int xb_i = (k - kz0) & 1;
    if (k < kz1) store(b, s);
    // unconditional (clamped to plane kz1): a plane past kz1 is never stored,
    // and no load result is carried through a conditional path
    fetch(min(k + 1 + PF, kz1), s);
    __syncthreads();
    // finish node plane k-1: its x sum plus the y exchange of the row below
    if (owner) {
      T sb[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) sb[c] = xs[c] + xc[xb_i][ty - 1][tx][c];
      if (k - 1 >= kz0) {
        const int kp = k - 1;
        const uint32_t m = (mq >> 8) & 0xffu;   // plane k-1 (byte 0: plane k)
        uint32_t yoff = own_xyb + (plane * (uint32_t)(NC * sizeof(T))) * (uint32_t)kp;
        asm volatile("" : "+v"(yoff));
        T* yp = reinterpret_cast<T*>(reinterpret_cast<char*>(a.y) + yoff);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const T v = ((m >> c) & 1u) ? T(0) : sb[c];
          yp[c] = v;
          dot += v * ubot[c];
        }
      }
    }
    if (k >= kz1) return false;
    Face<T, NC> tface;
    T top[4][NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      ubot[c] = ucur[c];
      top[0][c] = pl[b][ty][tx][c];
      top[1][c] = pl[b][ty][tx + 1][c];
      top[2][c] = pl[b][ty + 1][tx + 1][c];
      top[3][c] = pl[b][ty + 1][tx][c];
      ucur[c] = top[0][c];
    }
    if (kFaces) face_xy<T, NC>(top[0], top[1], top[2], top[3], tface);
    mq = (mq << 8) | (uint32_t)pm[b][ty][tx];
    if constexpr (MFMA) {
      float ue[8][3], f[8][3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ue[j][c] = bot[j][c];
          ue[4 + j][c] = top[j][c];
        }
      // the wave-uniform layer test as below; E = 0 outside the grid
      const float E = (k >= 0 && k < g.nz && col_valid) ? (float)es[b][ty][tx] : 0.0f;
      element_mfma(ue, E, ka, f);
      float fc[4][3];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          fc[j][c] = f[j][c] + ctop[j][c];
          ctop[j][c] = f[4 + j][c];
        }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        xs[c] = fc[0][c] + dpp_shr1(fc[1][c]);
        ys[c] = fc[3][c] + dpp_shr1(fc[2][c]);
        asm volatile("" : "+v"(xs[c]));
        xc[(k - kz0 + 1) & 1][ty][tx][c] = ys[c];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < NC; ++c) bot[j][c] = top[j][c];
      return true;
    }
    ModalN<T, NC> F;
    // No branch: an element layer outside the grid (k = -1 or nz at the
    // chunk ends) and a column outside it apply E = 0, and their corner
    // values are 0 as well (masked when staged), so F is 0 there.  A branch
    // on the wave-uniform layer test made the compiler zero the 21 modal
    // forces before it on every layer.
    {
      const T E = (col_valid && k >= 0 && k < g.nz) ? es[b][ty][tx] : T(0);   // E (or k) of layer k, staged with plane k+1
      if constexpr (NC == 1) {
        thermal_modal_faces<T>(bface, tface, E, a.mc, F);
      } else if constexpr (kFaces) {
        element_modal_faces<T, CUBE>(bface, tface, E, a.mc, F);
      } else {
        T ue[8][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            ue[j][c] = bot[j][c];
            ue[4 + j][c] = top[j][c];
          }
        element_modal<T, CUBE>(ue, E, a.mc, F);
      }
    }
    // node plane k from this layer's bottom face and the last layer's top
    // face; x-neighbour sums (DPP) for the nodes (tx,ty) and (tx,ty+1)
    T fc[4][NC];
    face_backward<T, NC>(F, btop, fc);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      xs[c] = fc[0][c] + dpp_shr1(fc[1][c]);
      ys[c] = fc[3][c] + dpp_shr1(fc[2][c]);
      // the x sum is formed here, next to its DPP shift (which then folds
      // into the add), not where the next layer uses it
      asm volatile("" : "+v"(xs[c]));
      // y exchange of plane k for the row above, read after the next
      // layer's barrier; that buffer was last read before this layer's
      // barrier
      xc[(k - kz0 + 1) & 1][ty][tx][c] = ys[c];
    }
    if (kFaces) {
      bface = tface;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < NC; ++c) bot[j][c] = top[j][c];
    }
    return true;
  };
  // plane k+1 is staged in st[(k - kz0 + 1) % PF]: with PF = 2 the two calls
  // of a trip use the two stages, with PF = 1 both use st[0]
  for (int k = kz0 - 1;; k += 2) {
    if (!layer(k, st[0])) break;
    if (!layer(k + 1, st[PF - 1])) break;
  }
  if (a.partials) {
    // block partial of u.y; the consumer kernel sums the partials
    // block partial: wave sums, then wave 0 adds them in a fixed order
    const double ws = wave_reduce_sum((double)dot);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
      double bs = 0.0;
#pragma unroll
      for (int w = 0; w < kThreads / 64; ++w) bs += red[w];
      a.partials[pidx] = bs;
    }
  }
}

}  // namespace pyto
