// The method of moving asymptotes on the device: one call of mmasub (the
// subproblem that src/mmaWrapper.py:90-91 solves through mmapy) and the KKT
// residual of kktcheck (:108-109), for n design variables (one per element,
// millions) and m <= 4 constraints.
//
// mmasub follows Svanberg's published MMA (2007 version): asymptote update,
// move limits alfa/beta, the convex separable approximation p0, q0, P, Q, b,
// then the primal-dual Newton method "subsolv" on barrier levels epsi = 1,
// 0.1, ... > 1e-7, at most 200 Newton steps per level while max|residual| >
// 0.9 epsi, each step scaled to stay 1% inside the bounds and halved (at most
// 50 times) until the residual norm does not grow.
//
// Work split: everything over the n variables is an element pass (HBM
// bound); the m-dimensional part of each Newton step (the (m+1)x(m+1)
// system, the scalar variables y, z, lam, mu, zet, s, the step length and
// the loop decisions) runs in one single-block kernel.  A phase word in
// device memory gates every launch, so the host queues rounds of
//   DIR (new direction), STEP (step length), EVAL (residual and Newton
//   system at a trial point), DECIDE (line search, loop control)
// and reads the phase back once per 16 rounds.  EVAL computes the residual
// norm of the trial point and, in the same pass, the reductions of the Newton
// system at that point (used when the trial is accepted, the common case).
//
// Element state (x, xsi, eta) lives in the caller's output arrays; a trial
// point is x + se*dx, materialised by the next DIR pass.  All reductions are
// per-block partials summed in a fixed order: results are deterministic.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace pyto {

constexpr int kMmaMaxM = 4;
constexpr int kMmaBlocks = 1024;
constexpr int kMmaComp = 2 + 2 * kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2;   // 20
constexpr double kEpsiMin = 1e-7;
constexpr int kRoundsPerCheck = 16;
constexpr int kMaxRounds = 200000;

enum { PH_EVAL = 1, PH_DIR = 2, PH_DONE = 3 };
enum { MODE_LEVEL = 0, MODE_LS = 1 };

struct MmaState {
  int phase, mode, ittt, itto, newton, evals, levels, pad;
  double epsi, se, ss, sp, residunorm, residumax;
  double a0, z, zet, dz, dzet;
  double y[kMmaMaxM], lam[kMmaMaxM], mu[kMmaMaxM], s[kMmaMaxM];
  double dy[kMmaMaxM], dlam[kMmaMaxM], dmu[kMmaMaxM], ds[kMmaMaxM];
  double a[kMmaMaxM], b[kMmaMaxM], c[kMmaMaxM], d[kMmaMaxM];
};

struct MmaScalarsIn {
  double a0;
  double a[kMmaMaxM], c[kMmaMaxM], d[kMmaMaxM], fval[kMmaMaxM];
};
struct MmaVecM {
  double v[kMmaMaxM];
};

// Element data of the subproblem (workspace arrays, length n; P and Q m x n).
struct MmaElem {
  const double *p0, *q0, *P, *Q, *low, *upp, *alfa, *beta;
};
struct MmaDir {
  double *dx, *dxsi, *deta;
};

// Per-block partials: component c of block b at part[c * kMmaBlocks + b].
__device__ __forceinline__ double wave_reduce_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Writes K block sums (component 1 is a max when HAS_MAX) as this block's
// partials.  acc[] is per thread.
template <int K, bool HAS_MAX>
__device__ __forceinline__ void block_partials(const double (&acc)[K], double* part) {
  __shared__ double red[4][K];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double v = (HAS_MAX && k == 1) ? wave_reduce_max(acc[k]) : wave_reduce_sum(acc[k]);
    if (lane == 0) red[w][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    const double r = (HAS_MAX && k == 1) ? fmax(fmax(red[0][k], red[1][k]), fmax(red[2][k], red[3][k]))
                                         : (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
    part[k * kMmaBlocks + blockIdx.x] = r;
  }
}

// In a single-block kernel: out[k] = fixed-order reduction of component k over
// nb partials (a max for k == 1 when HAS_MAX).  One wave per component.
template <int K, bool HAS_MAX>
__device__ __forceinline__ void reduce_partials(const double* part, int nb, double* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = w; k < K; k += 4) {
    const bool is_max = HAS_MAX && k == 1;
    double v = 0.0;   // max components are of non-negative bounds or absolute values
    for (int i = lane; i < nb; i += 64) {
      const double t = part[k * kMmaBlocks + i];
      v = is_max ? fmax(v, t) : v + t;
    }
    v = is_max ? wave_reduce_max(v) : wave_reduce_sum(v);
    if (lane == 0) out[k] = v;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Setup: asymptotes, move limits, approximation, initial interior point.
// ---------------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(256) void k_mma_setup(
    int64_t n, int iter, double move, const double* __restrict__ xval,
    const double* __restrict__ xmin, const double* __restrict__ xmax,
    const double* __restrict__ xold1, const double* __restrict__ xold2,
    const double* __restrict__ df0dx, const double* __restrict__ dfdx, double* __restrict__ low,
    double* __restrict__ upp, double* __restrict__ p0, double* __restrict__ q0,
    double* __restrict__ P, double* __restrict__ Q, double* __restrict__ alfa,
    double* __restrict__ beta, double* __restrict__ x, double* __restrict__ xsi,
    double* __restrict__ eta, MmaDir dir, double* __restrict__ part) {
  double acc[M];
#pragma unroll
  for (int i = 0; i < M; ++i) acc[i] = 0.0;
  for (int64_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const double xv = xval[j], span = xmax[j] - xmin[j];
    double lo, up;
    if (iter <= 2) {
      lo = xv - 0.5 * span;
      up = xv + 0.5 * span;
    } else {
      const double o1 = xold1[j], zzz = (xv - o1) * (o1 - xold2[j]);
      const double f = zzz > 0.0 ? 1.2 : (zzz < 0.0 ? 0.7 : 1.0);
      lo = xv - f * (o1 - low[j]);
      up = xv + f * (upp[j] - o1);
      lo = fmin(fmax(lo, xv - 10.0 * span), xv - 0.01 * span);
      up = fmax(fmin(up, xv + 10.0 * span), xv + 0.01 * span);
    }
    low[j] = lo;
    upp[j] = up;
    const double al = fmax(fmax(lo + 0.1 * (xv - lo), xv - move * span), xmin[j]);
    const double be = fmin(fmin(up - 0.1 * (up - xv), xv + move * span), xmax[j]);
    alfa[j] = al;
    beta[j] = be;
    const double xmamiinv = 1.0 / fmax(span, 1e-5);
    const double ux1 = up - xv, xl1 = xv - lo, ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    const double uxinv = 1.0 / ux1, xlinv = 1.0 / xl1;
    {
      const double g = df0dx[j];
      const double pp = fmax(g, 0.0), qq = fmax(-g, 0.0);
      const double pq = 0.001 * (pp + qq) + 1e-5 * xmamiinv;
      p0[j] = (pp + pq) * ux2;
      q0[j] = (qq + pq) * xl2;
    }
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const double g = dfdx[i * n + j];
      const double pp = fmax(g, 0.0), qq = fmax(-g, 0.0);
      const double pq = 0.001 * (pp + qq) + 1e-5 * xmamiinv;
      const double Pi = (pp + pq) * ux2, Qi = (qq + pq) * xl2;
      P[i * n + j] = Pi;
      Q[i * n + j] = Qi;
      acc[i] += Pi * uxinv + Qi * xlinv;
    }
    const double x0 = 0.5 * (al + be);
    x[j] = x0;
    xsi[j] = fmax(1.0 / (x0 - al), 1.0);
    eta[j] = fmax(1.0 / (be - x0), 1.0);
    dir.dx[j] = 0.0;
    dir.dxsi[j] = 0.0;
    dir.deta[j] = 0.0;
  }
  block_partials<M, false>(acc, part);
}

template <int M>
__global__ __launch_bounds__(256) void k_mma_init(MmaState* st, const double* __restrict__ part,
                                                  int nb, MmaScalarsIn in) {
  __shared__ double r[M];
  reduce_partials<M, false>(part, nb, r);
  if (threadIdx.x != 0) return;
  MmaState s{};
  s.phase = PH_EVAL;
  s.mode = MODE_LEVEL;
  s.epsi = 1.0;
  s.a0 = in.a0;
  s.z = 1.0;
  s.zet = 1.0;
  for (int i = 0; i < M; ++i) {
    s.a[i] = in.a[i];
    s.c[i] = in.c[i];
    s.d[i] = in.d[i];
    s.b[i] = r[i] - in.fval[i];
    s.y[i] = 1.0;
    s.lam[i] = 1.0;
    s.mu[i] = fmax(1.0, 0.5 * in.c[i]);
    s.s[i] = 1.0;
  }
  *st = s;
}

// ---------------------------------------------------------------------------
// Element quantities of the Newton system at one point.
// ---------------------------------------------------------------------------
template <int M>
struct ElemTerms {
  double ux1, xl1, ux2, xl2, plam, qlam, dpsidx, delx, diagx;
  double GG[M];
};

template <int M>
__device__ __forceinline__ ElemTerms<M> elem_terms(const MmaElem& e, int64_t n, int64_t j,
                                                   double x, double xsi, double eta,
                                                   const double (&lam)[M], double epsi,
                                                   double al, double be) {
  ElemTerms<M> t;
  const double lo = e.low[j], up = e.upp[j];
  t.ux1 = up - x;
  t.xl1 = x - lo;
  t.ux2 = t.ux1 * t.ux1;
  t.xl2 = t.xl1 * t.xl1;
  t.plam = e.p0[j];
  t.qlam = e.q0[j];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const double Pi = e.P[i * n + j], Qi = e.Q[i * n + j];
    t.plam += Pi * lam[i];
    t.qlam += Qi * lam[i];
    t.GG[i] = Pi / t.ux2 - Qi / t.xl2;
  }
  t.dpsidx = t.plam / t.ux2 - t.qlam / t.xl2;
  const double xa = x - al, bx = be - x;
  t.delx = t.dpsidx - epsi / xa + epsi / bx;
  t.diagx = 2.0 * (t.plam / (t.ux2 * t.ux1) + t.qlam / (t.xl2 * t.xl1)) + xsi / xa + eta / bx;
  return t;
}

// EVAL: residual norm parts and Newton-system reductions at x + se*dx with
// the scalar variables at old + ss*d.  Components: 0 sum of squares, 1 max
// abs, 2.. gvec[M], then GG*(delx/diagx) [M], then GG diag(1/diagx) GG^T
// (upper triangle, M(M+1)/2).
template <int M>
__global__ __launch_bounds__(256) void k_mma_eval(int64_t n, const MmaState* __restrict__ st,
                                                  MmaElem e, const double* __restrict__ x_,
                                                  const double* __restrict__ xsi_,
                                                  const double* __restrict__ eta_, MmaDir dir,
                                                  double* __restrict__ part) {
  if (st->phase != PH_EVAL) return;
  constexpr int K = 2 + 2 * M + M * (M + 1) / 2;
  const double se = st->se, ss = st->ss, epsi = st->epsi;
  double lam[M];
#pragma unroll
  for (int i = 0; i < M; ++i) lam[i] = st->lam[i] + ss * st->dlam[i];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int64_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const double x = x_[j] + se * dir.dx[j];
    const double xsi = xsi_[j] + se * dir.dxsi[j];
    const double eta = eta_[j] + se * dir.deta[j];
    const double al = e.alfa[j], be = e.beta[j];
    const ElemTerms<M> t = elem_terms<M>(e, n, j, x, xsi, eta, lam, epsi, al, be);
    const double rex = t.dpsidx - xsi + eta;
    const double rexsi = xsi * (x - al) - epsi;
    const double reeta = eta * (be - x) - epsi;
    acc[0] += rex * rex + rexsi * rexsi + reeta * reeta;
    acc[1] = fmax(acc[1], fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
    const double uxinv = 1.0 / t.ux1, xlinv = 1.0 / t.xl1, dinv = 1.0 / t.diagx;
    const double w = t.delx * dinv;
    int k = 2 + 2 * M;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      acc[2 + i] += e.P[i * n + j] * uxinv + e.Q[i * n + j] * xlinv;
      acc[2 + M + i] += t.GG[i] * w;
#pragma unroll
      for (int l = i; l < M; ++l) acc[k++] += t.GG[i] * t.GG[l] * dinv;
    }
  }
  block_partials<K, true>(acc, part);
}

// Dense (m+1)x(m+1) solve with partial pivoting (one thread).
__device__ __forceinline__ void small_solve(double (*A)[kMmaMaxM + 1], double* b, int n) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
    if (piv != c) {
      for (int k = 0; k < n; ++k) {
        const double t = A[c][k];
        A[c][k] = A[piv][k];
        A[piv][k] = t;
      }
      const double t = b[c];
      b[c] = b[piv];
      b[piv] = t;
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int k = c; k < n; ++k) A[r][k] -= f * A[c][k];
      b[r] -= f * b[c];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    double v = b[c];
    for (int k = c + 1; k < n; ++k) v -= A[c][k] * b[k];
    b[c] = v / A[c][c];
  }
}

// DECIDE: the line search and loop control of subsolv, and the scalar part of
// a new Newton step.
template <int M>
__global__ __launch_bounds__(256) void k_mma_decide(MmaState* st, const double* __restrict__ part,
                                                    int nb) {
  if (st->phase != PH_EVAL) return;
  constexpr int K = 2 + 2 * M + M * (M + 1) / 2;
  __shared__ double r[K];
  reduce_partials<K, true>(part, nb, r);
  if (threadIdx.x != 0) return;
  MmaState s = *st;
  s.evals += 1;
  const double epsi = s.epsi, ss = s.ss;
  double y[M], lam[M], mu[M], sv[M];
  for (int i = 0; i < M; ++i) {
    y[i] = s.y[i] + ss * s.dy[i];
    lam[i] = s.lam[i] + ss * s.dlam[i];
    mu[i] = s.mu[i] + ss * s.dmu[i];
    sv[i] = s.s[i] + ss * s.ds[i];
  }
  const double z = s.z + ss * s.dz, zet = s.zet + ss * s.dzet;
  // scalar residual parts (rey, rez, relam, remu, rezet, res)
  double sq = r[0], mx = r[1];
  double alam = 0.0;
  for (int i = 0; i < M; ++i) alam += s.a[i] * lam[i];
  auto add = [&](double v) {
    sq += v * v;
    mx = fmax(mx, fabs(v));
  };
  for (int i = 0; i < M; ++i) add(s.c[i] + s.d[i] * y[i] - mu[i] - lam[i]);
  add(s.a0 - zet - alam);
  for (int i = 0; i < M; ++i) add(r[2 + i] - s.a[i] * z - y[i] + sv[i] - s.b[i]);
  for (int i = 0; i < M; ++i) add(mu[i] * y[i] - epsi);
  add(zet * z - epsi);
  for (int i = 0; i < M; ++i) add(lam[i] * sv[i] - epsi);
  const double resinew = sqrt(sq);

  if (s.mode == MODE_LS) {
    s.itto += 1;
    if (resinew > s.residunorm && s.itto < 50) {
      s.se *= 0.5;
      s.ss *= 0.5;
      *st = s;
      return;   // evaluate the halved step next
    }
    for (int i = 0; i < M; ++i) {
      s.y[i] = y[i];
      s.lam[i] = lam[i];
      s.mu[i] = mu[i];
      s.s[i] = sv[i];
    }
    s.z = z;
    s.zet = zet;
    s.sp = s.se;
    s.ss = 0.0;
  } else {
    s.ittt = 0;
  }
  s.residunorm = resinew;
  s.residumax = mx;

  if (s.residumax > 0.9 * epsi && s.ittt < 200) {
    s.ittt += 1;
    s.newton += 1;
    // Newton system, case m < n (src: subsolv); r[] holds the sums at this
    // (accepted) point
    double dely[M], diagy[M], blam[M];
    double A[kMmaMaxM + 1][kMmaMaxM + 1] = {};
    double rhs[kMmaMaxM + 1];
    int k = 2 + 2 * M;
    for (int i = 0; i < M; ++i)
      for (int l = i; l < M; ++l) {
        A[i][l] = r[k];
        A[l][i] = r[k];
        ++k;
      }
    for (int i = 0; i < M; ++i) {
      dely[i] = s.c[i] + s.d[i] * s.y[i] - s.lam[i] - epsi / s.y[i];
      diagy[i] = s.d[i] + s.mu[i] / s.y[i];
      const double dellam = r[2 + i] - s.a[i] * s.z - s.y[i] - s.b[i] + epsi / s.lam[i];
      blam[i] = dellam + dely[i] / diagy[i] - r[2 + M + i];
      A[i][i] += s.s[i] / s.lam[i] + 1.0 / diagy[i];
      A[i][M] = s.a[i];
      A[M][i] = s.a[i];
      rhs[i] = blam[i];
    }
    A[M][M] = -s.zet / s.z;
    double alam2 = 0.0;
    for (int i = 0; i < M; ++i) alam2 += s.a[i] * s.lam[i];
    rhs[M] = s.a0 - alam2 - epsi / s.z;
    small_solve(A, rhs, M + 1);
    s.dz = rhs[M];
    for (int i = 0; i < M; ++i) {
      s.dlam[i] = rhs[i];
      s.dy[i] = -dely[i] / diagy[i] + rhs[i] / diagy[i];
      s.dmu[i] = -s.mu[i] + epsi / s.y[i] - (s.mu[i] * s.dy[i]) / s.y[i];
      s.ds[i] = -s.s[i] + epsi / s.lam[i] - (s.s[i] * s.dlam[i]) / s.lam[i];
    }
    s.dzet = -s.zet + epsi / s.z - s.zet * s.dz / s.z;
    s.phase = PH_DIR;
  } else {
    s.epsi = 0.1 * epsi;
    s.levels += 1;
    if (s.epsi > kEpsiMin) {
      s.mode = MODE_LEVEL;
      s.se = s.sp;
      s.ss = 0.0;
      s.phase = PH_EVAL;
    } else {
      s.phase = PH_DONE;
    }
  }
  *st = s;
}

// DIR: materialise the accepted point, new direction dx, dxsi, deta, and the
// step-length bound max(-1.01 dxsi/xsi, -1.01 deta/eta, -1.01 dx/(x-alfa),
// 1.01 dx/(beta-x)) as partial 1 (the max component).
template <int M>
__global__ __launch_bounds__(256) void k_mma_dir(int64_t n, const MmaState* __restrict__ st,
                                                 MmaElem e, double* __restrict__ x_,
                                                 double* __restrict__ xsi_,
                                                 double* __restrict__ eta_, MmaDir dir,
                                                 double* __restrict__ part) {
  if (st->phase != PH_DIR) return;
  const double sp = st->sp, epsi = st->epsi;
  double lam[M], dlam[M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    lam[i] = st->lam[i];
    dlam[i] = st->dlam[i];
  }
  double acc[2] = {0.0, 0.0};
  for (int64_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const double x = x_[j] + sp * dir.dx[j];
    const double xsi = xsi_[j] + sp * dir.dxsi[j];
    const double eta = eta_[j] + sp * dir.deta[j];
    const double al = e.alfa[j], be = e.beta[j];
    const ElemTerms<M> t = elem_terms<M>(e, n, j, x, xsi, eta, lam, epsi, al, be);
    double gl = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) gl += t.GG[i] * dlam[i];
    const double xa = x - al, bx = be - x;
    const double dx = -t.delx / t.diagx - gl / t.diagx;
    const double dxsi = -xsi + epsi / xa - (xsi * dx) / xa;
    const double deta = -eta + epsi / bx + (eta * dx) / bx;
    x_[j] = x;
    xsi_[j] = xsi;
    eta_[j] = eta;
    dir.dx[j] = dx;
    dir.dxsi[j] = dxsi;
    dir.deta[j] = deta;
    acc[1] = fmax(acc[1], fmax(fmax(-1.01 * dxsi / xsi, -1.01 * deta / eta),
                               fmax(-1.01 * dx / xa, 1.01 * dx / bx)));
  }
  block_partials<2, true>(acc, part);
}

// STEP: steg = 1 / max(1, element bound, scalar bounds); start the line search.
template <int M>
__global__ __launch_bounds__(256) void k_mma_step(MmaState* st, const double* __restrict__ part,
                                                  int nb) {
  if (st->phase != PH_DIR) return;
  __shared__ double r[2];
  reduce_partials<2, true>(part, nb, r);
  if (threadIdx.x != 0) return;
  MmaState s = *st;
  double stm = fmax(r[1], 1.0);
  auto bound = [&](double v, double dv) { stm = fmax(stm, -1.01 * dv / v); };
  for (int i = 0; i < M; ++i) {
    bound(s.y[i], s.dy[i]);
    bound(s.lam[i], s.dlam[i]);
    bound(s.mu[i], s.dmu[i]);
    bound(s.s[i], s.ds[i]);
  }
  bound(s.z, s.dz);
  bound(s.zet, s.dzet);
  const double steg = 1.0 / stm;
  s.se = steg;
  s.ss = steg;
  s.sp = 0.0;
  s.itto = 0;
  s.mode = MODE_LS;
  s.phase = PH_EVAL;
  *st = s;
}

// Final point: x + sp*dx (and xsi, eta) in place.
__global__ __launch_bounds__(256) void k_mma_final(int64_t n, const MmaState* __restrict__ st,
                                                   double* __restrict__ x_,
                                                   double* __restrict__ xsi_,
                                                   double* __restrict__ eta_, MmaDir dir) {
  const double sp = st->sp;
  for (int64_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    x_[j] += sp * dir.dx[j];
    xsi_[j] += sp * dir.dxsi[j];
    eta_[j] += sp * dir.deta[j];
  }
}

// KKT residual rows over the variables (kktcheck): rex = df0dx + dfdx^T lam -
// xsi + eta, rexsi = xsi (x - xmin), reeta = eta (xmax - x); sum of squares
// and max abs as partials 0 and 1.
template <int M>
__global__ __launch_bounds__(256) void k_mma_kkt(int64_t n, const double* __restrict__ x,
                                                 const double* __restrict__ xsi,
                                                 const double* __restrict__ eta,
                                                 const double* __restrict__ xmin,
                                                 const double* __restrict__ xmax,
                                                 const double* __restrict__ df0dx,
                                                 const double* __restrict__ dfdx, MmaVecM lam,
                                                 double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int64_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    double rex = df0dx[j];
#pragma unroll
    for (int i = 0; i < M; ++i) rex += dfdx[i * n + j] * lam.v[i];
    rex += -xsi[j] + eta[j];
    const double r1 = xsi[j] * (x[j] - xmin[j]), r2 = eta[j] * (xmax[j] - x[j]);
    acc[0] += rex * rex + r1 * r1 + r2 * r2;
    acc[1] = fmax(acc[1], fmax(fabs(rex), fmax(fabs(r1), fabs(r2))));
  }
  block_partials<2, true>(acc, part);
}

__global__ void k_mma_reduce2(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double r[2];
  reduce_partials<2, true>(part, nb, r);
  if (threadIdx.x == 0) {
    out[0] = r[0];
    out[1] = r[1];
  }
}

}  // namespace pyto

using namespace pyto;

struct pyto_mma {
  int device = 0;
  int64_t n = 0;
  int m = 0;
  double* buf = nullptr;     // p0 q0 alfa beta dx dxsi deta (7n) + P Q (2mn)
  double* part = nullptr;    // kMmaComp * kMmaBlocks
  MmaState* st = nullptr;
  MmaState* host_st = nullptr;   // pinned
};

extern "C" {

int pyto_mma_destroy(pyto_mma* w) {
  if (!w) return PYTO_OK;
  if (w->buf) hipFree(w->buf);
  if (w->part) hipFree(w->part);
  if (w->st) hipFree(w->st);
  if (w->host_st) hipHostFree(w->host_st);
  delete w;
  return PYTO_OK;
}

int pyto_mma_create(pyto_mma** out, int64_t n, int m) {
  PYTO_REQUIRE(out, PYTO_EINVAL, "null argument");
  *out = nullptr;
  PYTO_REQUIRE(n > 0 && m >= 1 && m <= kMmaMaxM, PYTO_EINVAL,
               "MMA workspace needs n > 0 and 1 <= m <= %d (got n=%lld, m=%d)", kMmaMaxM,
               (long long)n, m);
  PYTO_REQUIRE(m < n, PYTO_EINVAL, "the device subproblem solver needs m < n");
  pyto_mma* w = new pyto_mma();
  hipGetDevice(&w->device);
  w->n = n;
  w->m = m;
  const size_t bytes = sizeof(double) * (size_t)n * (7 + 2 * (size_t)m);
  if (hipMalloc(&w->buf, bytes) != hipSuccess ||
      hipMalloc(&w->part, sizeof(double) * kMmaComp * kMmaBlocks) != hipSuccess ||
      hipMalloc(&w->st, sizeof(MmaState)) != hipSuccess ||
      hipHostMalloc(&w->host_st, sizeof(MmaState)) != hipSuccess) {
    set_error("MMA workspace allocation of %zu bytes failed", bytes);
    pyto_mma_destroy(w);
    return PYTO_ENOMEM;
  }
  *out = w;
  return PYTO_OK;
}

}  // extern "C"

namespace {

template <int M>
int mma_sub_impl(pyto_mma* w, int iter, const double* xval, const double* xmin,
                 const double* xmax, const double* xold1, const double* xold2,
                 const double* df0dx, const double* dfdx, const MmaScalarsIn& in, double move,
                 double* low, double* upp, double* xmma, double* xsi, double* eta,
                 double* scal_host, pyto_mma_stats* stats, hipStream_t s) {
  const int64_t n = w->n;
  const int nb = (int)std::min<int64_t>((n + 255) / 256, kMmaBlocks);
  double* b = w->buf;
  double *p0 = b, *q0 = b + n, *alfa = b + 2 * n, *beta = b + 3 * n;
  MmaDir dir{b + 4 * n, b + 5 * n, b + 6 * n};
  double *P = b + 7 * n, *Q = b + (7 + M) * n;
  MmaElem e{p0, q0, P, Q, low, upp, alfa, beta};
  hipLaunchKernelGGL(k_mma_setup<M>, dim3(nb), dim3(256), 0, s, n, iter, move, xval, xmin, xmax,
                     xold1, xold2, df0dx, dfdx, low, upp, p0, q0, P, Q, alfa, beta, xmma, xsi,
                     eta, dir, w->part);
  hipLaunchKernelGGL(k_mma_init<M>, dim3(1), dim3(256), 0, s, w->st, w->part, nb, in);
  PYTO_CHECK_HIP(hipGetLastError());
  int rounds = 0;
  for (;;) {
    for (int k = 0; k < kRoundsPerCheck; ++k) {
      hipLaunchKernelGGL(k_mma_dir<M>, dim3(nb), dim3(256), 0, s, n, w->st, e, xmma, xsi, eta,
                         dir, w->part);
      hipLaunchKernelGGL(k_mma_step<M>, dim3(1), dim3(256), 0, s, w->st, w->part, nb);
      hipLaunchKernelGGL(k_mma_eval<M>, dim3(nb), dim3(256), 0, s, n, w->st, e, xmma, xsi, eta,
                         dir, w->part);
      hipLaunchKernelGGL(k_mma_decide<M>, dim3(1), dim3(256), 0, s, w->st, w->part, nb);
    }
    PYTO_CHECK_HIP(hipGetLastError());
    rounds += kRoundsPerCheck;
    PYTO_CHECK_HIP(hipMemcpyAsync(w->host_st, w->st, sizeof(MmaState), hipMemcpyDeviceToHost, s));
    PYTO_CHECK_HIP(hipStreamSynchronize(s));
    if (w->host_st->phase == PH_DONE) break;
    PYTO_REQUIRE(rounds < kMaxRounds, PYTO_ENOCONV,
                 "MMA subproblem did not finish in %d rounds", kMaxRounds);
  }
  hipLaunchKernelGGL(k_mma_final, dim3(nb), dim3(256), 0, s, n, w->st, xmma, xsi, eta, dir);
  PYTO_CHECK_HIP(hipGetLastError());
  const MmaState& hs = *w->host_st;
  // y[m], z, lam[m], mu[m], zet, s[m]
  if (scal_host) {
    double* o = scal_host;
    for (int i = 0; i < M; ++i) *o++ = hs.y[i];
    *o++ = hs.z;
    for (int i = 0; i < M; ++i) *o++ = hs.lam[i];
    for (int i = 0; i < M; ++i) *o++ = hs.mu[i];
    *o++ = hs.zet;
    for (int i = 0; i < M; ++i) *o++ = hs.s[i];
  }
  if (stats) {
    stats->newton_steps = hs.newton;
    stats->evaluations = hs.evals;
    stats->levels = hs.levels;
    stats->rounds = rounds;
    stats->residual_norm = hs.residunorm;
  }
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  return PYTO_OK;
}

template <int M>
int mma_kkt_impl(pyto_mma* w, const double* x, const double* xsi, const double* eta,
                 const double* xmin, const double* xmax, const double* df0dx, const double* dfdx,
                 const MmaVecM& lam, double* out_host, hipStream_t s) {
  const int64_t n = w->n;
  const int nb = (int)std::min<int64_t>((n + 255) / 256, kMmaBlocks);
  hipLaunchKernelGGL(k_mma_kkt<M>, dim3(nb), dim3(256), 0, s, n, x, xsi, eta, xmin, xmax, df0dx,
                     dfdx, lam, w->part);
  double* dev_out = w->part + 2 * kMmaBlocks;   // partial slots the KKT pass does not use
  hipLaunchKernelGGL(k_mma_reduce2, dim3(1), dim3(256), 0, s, w->part, nb, dev_out);
  PYTO_CHECK_HIP(hipGetLastError());
  PYTO_CHECK_HIP(hipMemcpyAsync(w->host_st, dev_out, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  PYTO_CHECK_HIP(hipStreamSynchronize(s));
  const double* r = reinterpret_cast<const double*>(w->host_st);
  out_host[0] = r[0];
  out_host[1] = r[1];
  return PYTO_OK;
}

}  // namespace

extern "C" {

int pyto_mma_sub(pyto_mma* w, int iter, const double* xval, const double* xmin,
                 const double* xmax, const double* xold1, const double* xold2,
                 const double* df0dx, const double* dfdx, const double* fval_host, double a0,
                 const double* a_host, const double* c_host, const double* d_host, double move,
                 double* low, double* upp, double* xmma, double* xsi, double* eta,
                 double* scal_host, pyto_mma_stats* stats, void* stream) {
  PYTO_REQUIRE(w && xval && xmin && xmax && xold1 && xold2 && df0dx && dfdx && fval_host &&
                   a_host && c_host && d_host && low && upp && xmma && xsi && eta,
               PYTO_EINVAL, "null argument");
  PYTO_REQUIRE(iter >= 1 && move > 0.0, PYTO_EINVAL, "need iter >= 1 and move > 0");
  int dev = 0;
  hipGetDevice(&dev);
  PYTO_REQUIRE(dev == w->device, PYTO_ESTATE, "MMA workspace belongs to device %d, current %d",
               w->device, dev);
  MmaScalarsIn in{};
  in.a0 = a0;
  for (int i = 0; i < w->m; ++i) {
    PYTO_REQUIRE(c_host[i] >= 0.0 && d_host[i] > 0.0 && a_host[i] >= 0.0, PYTO_EINVAL,
                 "need a >= 0, c >= 0, d > 0");
    in.a[i] = a_host[i];
    in.c[i] = c_host[i];
    in.d[i] = d_host[i];
    in.fval[i] = fval_host[i];
  }
  hipStream_t s = as_stream(stream);
  switch (w->m) {
    case 1: return mma_sub_impl<1>(w, iter, xval, xmin, xmax, xold1, xold2, df0dx, dfdx, in, move, low, upp, xmma, xsi, eta, scal_host, stats, s);
    case 2: return mma_sub_impl<2>(w, iter, xval, xmin, xmax, xold1, xold2, df0dx, dfdx, in, move, low, upp, xmma, xsi, eta, scal_host, stats, s);
    case 3: return mma_sub_impl<3>(w, iter, xval, xmin, xmax, xold1, xold2, df0dx, dfdx, in, move, low, upp, xmma, xsi, eta, scal_host, stats, s);
    default: return mma_sub_impl<4>(w, iter, xval, xmin, xmax, xold1, xold2, df0dx, dfdx, in, move, low, upp, xmma, xsi, eta, scal_host, stats, s);
  }
}

int pyto_mma_kkt(pyto_mma* w, const double* x, const double* xsi, const double* eta,
                 const double* xmin, const double* xmax, const double* df0dx, const double* dfdx,
                 const double* lam_host, double* out_host, void* stream) {
  PYTO_REQUIRE(w && x && xsi && eta && xmin && xmax && df0dx && dfdx && lam_host && out_host,
               PYTO_EINVAL, "null argument");
  MmaVecM lam{};
  for (int i = 0; i < w->m; ++i) lam.v[i] = lam_host[i];
  hipStream_t s = as_stream(stream);
  switch (w->m) {
    case 1: return mma_kkt_impl<1>(w, x, xsi, eta, xmin, xmax, df0dx, dfdx, lam, out_host, s);
    case 2: return mma_kkt_impl<2>(w, x, xsi, eta, xmin, xmax, df0dx, dfdx, lam, out_host, s);
    case 3: return mma_kkt_impl<3>(w, x, xsi, eta, xmin, xmax, df0dx, dfdx, lam, out_host, s);
    default: return mma_kkt_impl<4>(w, x, xsi, eta, xmin, xmax, df0dx, dfdx, lam, out_host, s);
  }
}

}  // extern "C"
