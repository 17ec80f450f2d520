/*
 * pyto_hip.h -- C ABI of libpytohip.so, the MI355X (gfx950) implementation of
 * PyTO's voxel-hex structural FEA path: the SIMP-scaled 8-node-hex operator,
 * the Jacobi-preconditioned deflated PCG, the cone filter and the element
 * kernels of the compliance and p-norm stress sensitivities.
 *
 * The reference (UW-ERSL/PyTO) is pure Python with no FFI layer; each entry
 * point below names the reference function or expression it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative PYTO_E* code; the
 *     message of the last error on the calling thread is pyto_last_error().
 *     No C++ exception crosses this boundary.
 *   - All array arguments are caller-owned DEVICE pointers (e.g. torch tensor
 *     data_ptr()), except where a name ends in _host.  Vectors over DOFs use
 *     the reference's AoS order: dof = 3*node + component, node index
 *     ix + (nx+1)*iy + (nx+1)*(ny+1)*iz (src/hex_mesher.py:104-109,1085-1093).
 *     Element arrays use element index ix + nx*iy + nx*ny*iz
 *     (src/hex_mesher.py:117-124).
 *   - `stream` is a hipStream_t passed as void* (0 = null stream).  Calls are
 *     asynchronous except the drivers that return scalars to the host.
 *   - A handle owns only its constants and scratch workspace, which
 *     pyto_destroy frees.  One thread at a time may use a handle, and the
 *     calls on one handle must be ordered by the caller (one stream, or
 *     streams the caller synchronises).  Use of separate handles from
 *     several host threads at once is not verified: density updates of
 *     separate handles from concurrent threads once gave a non-positive
 *     coarse pivot that sequential updates of the same handles do not, with
 *     the cause unexplained.  Drive the handles of one process from one
 *     thread.
 */
#ifndef PYTO_HIP_H
#define PYTO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYTO_OK 0
#define PYTO_EINVAL -1     /* bad argument (shape, range, null pointer) */
#define PYTO_EHIP -2       /* HIP runtime error */
#define PYTO_ENOMEM -3     /* device allocation failed */
#define PYTO_ESTATE -4     /* call out of order (e.g. solve before density) */
#define PYTO_ENOCONV -5    /* solver stopped at maxIters or stagnated (result still written) */
#define PYTO_ENUMERIC -6   /* coarse matrix W^T K W not positive definite */

typedef struct pyto_handle pyto_handle;

/* Solver statistics written by the solve drivers. */
typedef struct pyto_solve_stats {
  int iterations;        /* inner PCG iterations (sum over refinement steps) */
  int outer_iterations;  /* refinement steps (1 for the plain FP64 solve) */
  int converged;         /* 1 if the stopping rule was met */
  int matvecs64;         /* FP64 operator applications */
  int matvecs32;         /* FP32 operator applications */
  double rel;            /* final sqrt(r^T M r / r0^T M r0) */
  double rel_true;       /* ||f - K u|| / ||f|| recomputed in FP64 at exit */
  double setup_ms;       /* density set-up time inside the call, if any */
} pyto_solve_stats;

/* No reference counterpart: the calling thread's last error message (the
 * reference raises Python exceptions instead) and the ABI version. */
const char* pyto_last_error(void);
int pyto_version(void);

/* Element stiffness KE (24x24, row-major FP64) of one brick element.
 * Replaces hex8_stiffness_matrix_structural, src/hex_element_stiffness.py:231-286. */
int pyto_ke_hex8(double E, double nu, double hx, double hy, double hz, double* ke_host);

/* Thermal element matrix KE_T (8x8, row-major FP64) of one brick of
 * conductivity k0.  Replaces hex8_stiffness_matrix_thermal,
 * src/hex_element_stiffness.py:339-385. */
int pyto_ke_hex8_thermal(double k0, double hx, double hy, double hz, double* ke_host);

/* Create the grid operator on the current device.
 * nelx/nely/nelz, element size and material replace HexMesher.grid_mesh +
 * HexStructuralFEA.__init__ (src/hex_mesher.py:72-169,
 * src/hex_structural_fea.py:21-52).  fixed_mask: uint8 per node, bit c set if
 * DOF 3*node+c is fixed (BC.fixed_dofs, src/bound_cond.py:8-32); only zero
 * Dirichlet values are supported.  penal/eps_void parametrise the SIMP law
 * E_e = E0*(eps + x^penal (1-eps)) (src/topopt_material_model.py:88). */
int pyto_create(pyto_handle** out, int nelx, int nely, int nelz,
                double hx, double hy, double hz, double E0, double nu,
                double penal, double eps_void, const uint8_t* fixed_mask,
                void* stream);

/* Create a thermal grid operator (one temperature per node) on the current
 * device: the 8x8 KE_T of conductivity k0 applied matrix-free, replacing the
 * COO assembly of HexThermalFEA.solve (src/hex_thermal_fea.py:74-117).
 * fixed_mask: uint8 per node, bit 0 set if the node's temperature is fixed
 * (zero Dirichlet value).  penal/k_void parametrise the conductivity law
 * k_e = k0*(k_void + x^penal (1-k_void)) (src/topopt_material_model.py:98-109).
 * The same matvec, diagonal, set_density, solve and compliance entry points
 * then act on vectors of one value per node.  Deflation groups are not
 * supported on a thermal handle (the solves are Jacobi PCG). */
int pyto_create_thermal(pyto_handle** out, int nelx, int nely, int nelz, double hx, double hy,
                        double hz, double k0, double penal, double k_void,
                        const uint8_t* fixed_mask, void* stream);
int pyto_destroy(pyto_handle* h);
int pyto_info(pyto_handle* h, int64_t* nnodes, int64_t* nelems, int64_t* ndof,
              int64_t* workspace_bytes);

/* Densities x (FP64, per element) -> E_e (FP64 and FP32 copies) and the
 * Jacobi inverse diagonal; if deflation groups are set, also assembles the
 * coarse Galerkin matrix sym(W^T K W) + shift (default 1e-10 I) and inverts
 * it in FP64 (Cholesky; PYTO_ENUMERIC if it is not positive definite).
 * Replaces the per-solve part of HexStructuralFEA.solve
 * (src/hex_structural_fea.py:87-102), _jacobi_preconditioner
 * (src/linear_solvers.py:71-88) and the set-up of deflatedPCG_CPU
 * (src/deflation.py:514-521). */
int pyto_set_density(pyto_handle* h, const double* x, void* stream);

/* pyto_set_density for n distinct handles at once (the designs of a batch):
 * every handle's E_e, Jacobi diagonal and Galerkin matrix, then one batched
 * FP64 Cholesky factorisation and inverse of all their coarse matrices
 * (rocSOLVER's batched potrf/potri on copies padded with identity to the
 * largest order) and one synchronisation of `stream` for the pivot check.
 * x[i] holds handle i's densities.  No reference counterpart (the reference
 * updates one design per solve). */
int pyto_set_density_batch(int n, pyto_handle* const* handles, const double* const* x,
                           void* stream);

/* Deflation groups: node_group (int32 per node, 0..ngroups-1) and group
 * centres (FP64, ngroups x 3).  W is implicit (rigid-body modes about the
 * centres, column map u:{0,4,5} v:{1,3,5} w:{2,3,4}); replaces
 * DeflationSolver.create_deflation_matrix (src/deflation.py:386-458).
 * On a thermal handle (pyto_create_thermal) W holds one constant per group
 * instead (ngroups coarse unknowns; the centres are not used).
 * node_group and centers are device pointers; the handle keeps copies. */
int pyto_set_groups(pyto_handle* h, const int32_t* node_group, int ngroups,
                    const double* centers, void* stream);

/* Operator tile shape: 16 x rows element columns per workgroup (rows = 8,
 * 16 or 32; 16 rows by default), re-labels the tiles' fixed-node planes.
 * Results are the same for every shape (each node is summed by one tile in
 * a fixed order); the dot-product partials change their grouping.  No
 * reference counterpart: a launch-shape setting of this operator, chosen by
 * measurement (GridOperator.autotune_tile_rows).  A batch sets its
 * systems' handles to its own rows (pyto_batch_set_tile_rows). */
int pyto_set_tile_rows(pyto_handle* h, int rows, void* stream);
int pyto_get_tile_rows(const pyto_handle* h);

/* y = K_ff u (fixed DOFs of u ignored, of y zeroed): the SIMP-scaled
 * matrix-free product replacing K@p with the assembled CSR K
 * (src/deflation.py:540,558,573; src/hex_structural_fea.py:101). */
int pyto_matvec_f64(pyto_handle* h, const double* u, double* y, void* stream);
int pyto_matvec_f32(pyto_handle* h, const float* u, float* y, void* stream);
/* The same FP32 product with the element contraction on the matrix cores
 * (the dense 24x24 KE by v_mfma_f32_32x32x2_f32 instead of the modal
 * element).  No reference counterpart: a measurement variant, slower than
 * pyto_matvec_f32 (the dense contraction does about 10x the flops of the
 * modal element). */
int pyto_matvec_f32_mfma(pyto_handle* h, const float* u, float* y, void* stream);

/* y = K_fc g: the operator with g's fixed entries kept and y's fixed entries
 * zeroed (FP64).  f_f - K_fc g is the right-hand side with the non-zero
 * Dirichlet values g eliminated: b' = b - A[:, fixed] g of impose_dirichlet_bc
 * (src/bound_cond.py:57-66). */
int pyto_matvec_lift_f64(pyto_handle* h, const double* g, double* y, void* stream);

/* diag(K_ff) (fixed DOFs 0); replaces A.diagonal() in
 * _jacobi_preconditioner (src/linear_solvers.py:84). */
int pyto_diagonal_f64(pyto_handle* h, double* d, void* stream);

/* Coarse operators (need groups + density).  restrict: t = W^T v (6G, or G
 * on a thermal handle);
 * prolong: v = W mu (fixed DOFs 0); coarse_solve: mu = E^{-1} t with the
 * FP64 factor; replaces WT@v, W@mu, spsolve(WKW, .) (src/deflation.py:533-551). */
int pyto_restrict_f64(pyto_handle* h, const double* v, double* t, void* stream);
int pyto_prolong_f64(pyto_handle* h, const double* mu, double* v, void* stream);
int pyto_coarse_solve_f64(pyto_handle* h, const double* t, double* mu, void* stream);
/* Diagonal shift of the coarse matrix: E = sym(W^T K W) + (abs + rel*|E_ii|) on
 * the diagonal.  Default abs = 1e-10, rel = 0: the reference's
 * `WKW += eye * 1e-10` (src/deflation.py:521).  That absolute shift makes the
 * deflation inexact (the iteration stalls below rtol ~1e-9, and earlier when
 * void elements make coarse entries small); abs = rel = 0 gives the exact
 * Galerkin matrix.  Takes effect at the next pyto_set_density. */
int pyto_set_coarse_shift(pyto_handle* h, double abs_shift, double rel_shift);
/* Copy the coarse Galerkin matrix (6G x 6G FP64, G x G on a thermal handle,
 * device) for inspection. */
int pyto_coarse_matrix_f64(pyto_handle* h, double* E, void* stream);

/* Plain FP64 deflated PCG: the reference algorithm of deflatedPCG_CPU
 * (src/deflation.py:504-584) with the stopping rule sqrt(rz/rz0) <= rtol and
 * maxIters; f and u are full-length DOF vectors (u[fixed] = 0), i.e.
 * linear_solvers.solve(..., Solvers.DPCG, ...) (src/linear_solvers.py:177-202).
 * Synchronises `stream`; returns PYTO_ENOCONV (u still written) at maxIters,
 * or when the iteration stagnates past its attainable accuracy (once the
 * best relative residual is below 1e-11: it grows 1e3x above that best, or
 * gains < 1% for max(200, k/2) iterations), which the reference loop does
 * not detect. */
int pyto_solve_dpcg_f64(pyto_handle* h, const double* f, double* u, double rtol,
                        int max_iters, pyto_solve_stats* stats, void* stream);

/* FP32 deflated PCG inside FP64 iterative refinement: the outer residual
 * f - K u is recomputed in FP64, inner products accumulate in FP64, the coarse
 * factor is FP64; stops when the FP64 sqrt(r^T M r / r0^T M r0) <= rtol
 * (the reference's rule on the true residual).  Each refinement step runs
 * the FP32 iteration until its residual has dropped by max(inner_rtol,
 * 0.3 rtol/rel) (at most 0.1), then replaces it by the FP64 residual and
 * continues the same search direction (reliable update) unless the two
 * residuals differ by more than 1.25x in norm, where it restarts.  inner_rtol
 * around 0.1 keeps the FP32 drift below the residual.  No reference
 * counterpart: this is the precision form of the same solve. */
int pyto_solve_dpcg_mixed(pyto_handle* h, const double* f, double* u, double rtol,
                          double inner_rtol, int max_iters, int max_outer,
                          pyto_solve_stats* stats, void* stream);

/* Compliance: ce_e = u_e^T KE u_e, C = sum E(x_e) ce_e (FP64 reduction into
 * *C, a device scalar), grad_e = -E'(x_e) ce_e; replaces
 * compute_compliance_and_gradient (src/topopt_obj_cons_sensitivities.py:148-179).
 * ce and grad may be null. */
int pyto_compliance(pyto_handle* h, const double* u, const double* x, double* ce,
                    double* grad, double* C, void* stream);

/* Centre temperature gradient in reference coordinates of a thermal handle:
 * grad (3 x N_el, FP64) = gradN @ T[edofMatThermal].T with gradN the +-1/8
 * table of HexThermalFEA.postprocess (src/hex_thermal_fea.py:250-270). */
int pyto_thermal_gradient(pyto_handle* h, const double* T, double* grad, void* stream);

/* Bilinear element term be_e = a_e^T KE b_e (FP64); with a = adjoint and
 * b = u it is the T2 factor of src/topopt_obj_cons_sensitivities.py:279-282. */
int pyto_element_bilinear(pyto_handle* h, const double* a, const double* b,
                          double* be, void* stream);

/* Centre-point engineering strains and stresses (N_el x 6 each, order
 * xx,yy,zz,xy,xz,yz; stresses relaxed by x^q), von Mises (N_el), element
 * strain energy 0.5 eps.sigma (N_el) and sum vm^p (device scalar); replaces
 * HexStructuralFEA.postprocess (src/hex_structural_fea.py:152-232).
 * strain, stress and energy may be null. */
int pyto_stress(pyto_handle* h, const double* u, const double* x, double q,
                double pexp, double* strain, double* stress, double* vm,
                double* energy, double* sum_vm_p, void* stream);

/* p-norm adjoint right-hand side g (N_dof) and direct term T1 (N_el), given
 * vm and sum_vm_p from pyto_stress; the intended math of
 * src/topopt_obj_cons_sensitivities.py:229-266 (g_e as a flat 24-vector). */
int pyto_pnorm_adjoint_rhs(pyto_handle* h, const double* u, const double* x,
                           const double* vm, const double* sum_vm_p, double q,
                           double pexp, double* g, double* T1, void* stream);

/* Cone filter of radius rel_radius*hx (src/topopt_filters.py:8-39) as a fixed
 * stencil.  mode 0: out = H v; 1: out = (H v)/Hs (density filter,
 * src/topopt_mma.py:255); 2: out = (H (x v))/Hs/x with x = aux (sensitivity
 * filter, src/topopt_ocm.py:239); 3: out = H (v/Hs) (density-filter chain
 * rule, src/topopt_mma.py:342-345); 4: out = Hs (v ignored). */
int pyto_filter_apply(pyto_handle* h, double rel_radius, int mode, const double* v,
                      const double* aux, double* out, void* stream);
/* Stencil of the filter (host arrays): *count offsets (dx,dy,dz) and weights;
 * capacity is the size of the output arrays. */
int pyto_filter_stencil(pyto_handle* h, double rel_radius, int capacity, int* count,
                        int32_t* offsets_host, double* weights_host);

/* ---- optimality-criteria update ----------------------------------------- */
/* OC density update with the volume bisection of src/topopt_ocm.py:247-269,
 * run on the device: l1 = 0, l2 = 1e12; while (l2-l1)/((l1+l2)/2+1e-10) > 1e-7:
 * lmid = (l1+l2)/2, D = clip(-dc/(dv*lmid), 0.1, 10),
 * xnew = max(xmin, max(x-move, min(xmax, min(x+move, x*sqrt(D))))),
 * l1 = lmid if mean(xnew)/volfrac - 1 > 0 else l2 = lmid.  Writes the xnew of
 * the last lmid evaluated.  x, dc, xnew: device arrays of n values; dv is the
 * (uniform) volume-constraint gradient 1/(volfrac*n).  At most max_steps
 * bisection steps are enqueued; nothing is read back unless lmid_host or
 * steps_host is given (then the stream is synchronised; *steps_host < 0 means
 * max_steps ran out before the stopping rule held). */
int pyto_oc_update(pyto_handle* h, const double* x, const double* dc, int64_t n, double dv,
                   double volfrac, double move, double xmin, double xmax, int max_steps,
                   double* xnew, double* lmid_host, int* steps_host, void* stream);
/* No reference counterpart: pyto_oc_update for callers that cannot allocate
 * device memory (code under HIP graph capture, tight optimisation or
 * serving loops, harnesses that forbid hipMalloc).  It needs no handle and
 * calls no hipMalloc/hipFree or allocating library routine: its scratch (the
 * bisection's block partials, state and flags) is one static device array of
 * 65.9 KB per GPU, held until the process ends, whose size does not depend
 * on n.  Same kernels, same launch shape and so the same results as
 * pyto_oc_update, bit for bit.  Every call on a GPU shares that GPU's array
 * (rewritten at the start of each call): calls that overlap on one GPU, on
 * different streams or threads, race.  lmid_host/steps_host as in
 * pyto_oc_update (they synchronise the stream; pass null under capture). */
int pyto_oc_update_noalloc(const double* x, const double* dc, int64_t n, double dv,
                           double volfrac, double move, double xmin, double xmax, int max_steps,
                           double* xnew, double* lmid_host, int* steps_host, void* stream);

/* ---- batches of independent systems ------------------------------------- */
/* A batch solves many independent systems K_i u_i = f_i at once: several
 * designs (handles, each with its own grid, densities and BCs) and several
 * load cases per design (the same handle listed once per load case).  Every
 * kernel of the iteration is one launch for the whole batch; each system has
 * its own scalars and stopping test.  Per system the algorithm is PCG in FP32
 * inside FP64 iterative refinement, as pyto_solve_dpcg_mixed: the deflated
 * PCG of src/deflation.py:504-584 when the handle has deflation groups
 * (pyto_set_groups before the densities), otherwise Jacobi PCG (the
 * reference's PCG branch, src/linear_solvers.py:157-176).  No reference
 * counterpart: the reference solves one system per call.
 * All handles must be structural (pyto_create; a thermal handle is
 * refused with PYTO_EINVAL) and use cube elements, or all bricks.  New densities keep a
 * batch valid; new deflation groups on a handle need a new batch
 * (pyto_batch_solve returns PYTO_ESTATE).  The batch sets its handles'
 * operator tile rows to its own (16 at creation). */
typedef struct pyto_batch pyto_batch;
int pyto_batch_create(pyto_batch** out, int nsys, pyto_handle* const* handles, void* stream);
int pyto_batch_destroy(pyto_batch* b);
/* Node planes per workgroup of the batched operator (the tile depth; the
 * autotuner times candidates with pyto_batch_time_matvec). */
int pyto_batch_set_zchunk(pyto_batch* b, int zchunk, void* stream);
/* Tile rows of the batched operator (8, 16 or 32; 16 at creation).  Sets
 * every system's handle to the same rows (pyto_set_tile_rows).  No reference
 * counterpart: the batch autotuner's second launch-shape axis. */
int pyto_batch_set_tile_rows(pyto_batch* b, int rows, void* stream);
int pyto_batch_get_tile_rows(const pyto_batch* b);
int pyto_batch_info(pyto_batch* b, int* zchunk, int* matvec_blocks, int* vector_blocks);
/* Average duration (ms) of `reps` launches of the batched FP32 operator. */
int pyto_batch_time_matvec(pyto_batch* b, int reps, float* ms_per_launch, void* stream);
/* Microseconds per grid-wide barrier of a persistent kernel (one workgroup
 * of 256 threads on each of blocks_per_cu x CU count slots, cooperative
 * launch, counter barrier with a release fence and an acquiring poll; the
 * launch itself subtracted).  No reference counterpart: the measurement
 * behind replaying captured iteration graphs instead of one persistent
 * launch (each batched iteration has one dependent phase per kernel).
 * PYTO_ENOCONV if a wait exceeded its poll limit (all workgroups then
 * leave). */
int pyto_measure_grid_barrier(int blocks_per_cu, int rounds, float* us_per_barrier,
                              void* stream);
/* Solve all systems: f[i], u[i] are device FP64 vectors of system i's DOFs
 * (u overwritten).  Densities must be set on the handles.  stats (nsys
 * entries, may be null) per system.  Synchronises `stream`; PYTO_ENOCONV if
 * any system misses rtol within max_iters inner iterations / max_outer
 * refinement steps. */
int pyto_batch_solve(pyto_batch* b, const double* const* f, double* const* u, double rtol,
                     double inner_rtol, int max_iters, int max_outer, pyto_solve_stats* stats,
                     void* stream);

/* Thermo-elastic coupling.  H is a 24 x 8 row-major FP64 host array whose row
 * 3 j + c belongs to element node j, displacement component c (how the
 * reference indexes its getHMatrix, src/hex_thermal_fea.py:119-237); the
 * law E(x) = eps + x^penal (1 - eps) is the structural one.  Any handle of
 * the grid (structural or thermal) may be passed; T is one value per node,
 * d and f three.
 *
 * pyto_thermoelastic_force: f = sum_e e_alpha E(x_e) H (T_e - t_ref) over all
 * nodes (get_thermoelastic_force, src/hex_thermal_fea.py:198-237).
 * pyto_thermoelastic_adjoint_rhs: r = sum_e coef E(x_e) H^T d_e (the
 * right-hand side of solve_thermal_adjoint with coef = -2 E alpha,
 * src/topopt_obj_cons_sensitivities.py:352-366).
 * pyto_thermoelastic_bilinear: out_e = d_e^T H (T_e - t_ref) (the direct term
 * of compute_thermoelastic_compliance_and_gradient,
 * src/topopt_obj_cons_sensitivities.py:455-458). */
int pyto_thermoelastic_force(pyto_handle* h, const double* T, const double* x, double penal,
                             double eps, double e_alpha, double t_ref, const double* H_host,
                             double* f, void* stream);
int pyto_thermoelastic_adjoint_rhs(pyto_handle* h, const double* d, const double* x,
                                   double penal, double eps, double coef, const double* H_host,
                                   double* r, void* stream);
int pyto_thermoelastic_bilinear(pyto_handle* h, const double* d, const double* T, double t_ref,
                                const double* H_host, double* out, void* stream);

/* ---- method of moving asymptotes ------------------------------------------ */
/* One MMA step (mmasub) and its KKT residual (kktcheck) for n variables and
 * 1 <= m <= 4 constraints, m < n.  These replace the mmapy calls of runMMA,
 * src/mmaWrapper.py:90-91 and :108-109 (mmapy implements K. Svanberg's
 * published MMA; same constants: asymptote factors 0.5/1.2/0.7, bounds 0.01
 * and 10 times xmax - xmin, raa0 1e-5, albefa 0.1, epsimin 1e-7).
 *
 * pyto_mma_create: workspace for n variables and m constraints
 * (8 (7 + 2m) n bytes).
 * pyto_mma_sub: xval, xmin, xmax, xold1, xold2, df0dx (n) and dfdx (m x n,
 * row-major) are device inputs; low and upp (n) are read (iteration > 2) and
 * overwritten with the new asymptotes; xmma, xsi, eta (n) receive the
 * subproblem solution.  fval_host, a_host, c_host, d_host hold m values.
 * scal_host (may be null) receives y[m], z, lam[m], mu[m], zet, s[m]
 * (4m + 2 values).  Synchronous: returns when the result is written.
 * pyto_mma_kkt: the rows of the KKT residual over the variables,
 * df0dx + dfdx^T lam - xsi + eta, xsi (x - xmin), eta (xmax - x); writes
 * their sum of squares and max abs to out_host[0..1] (the m + 1 scalar rows
 * are the caller's). */
typedef struct pyto_mma pyto_mma;
typedef struct pyto_mma_stats {
  int newton_steps;      /* Newton steps over all barrier levels */
  int evaluations;       /* residual evaluations (line-search trials + level starts) */
  int levels;            /* barrier levels */
  int rounds;            /* launch rounds queued (a multiple of 16) */
  double residual_norm;  /* final residual norm of the subproblem */
} pyto_mma_stats;
int pyto_mma_create(pyto_mma** out, int64_t n, int m);
int pyto_mma_destroy(pyto_mma* w);
int pyto_mma_sub(pyto_mma* w, int iter, const double* xval, const double* xmin,
                 const double* xmax, const double* xold1, const double* xold2,
                 const double* df0dx, const double* dfdx, const double* fval_host, double a0,
                 const double* a_host, const double* c_host, const double* d_host, double move,
                 double* low, double* upp, double* xmma, double* xsi, double* eta,
                 double* scal_host, pyto_mma_stats* stats, void* stream);
int pyto_mma_kkt(pyto_mma* w, const double* x, const double* xsi, const double* eta,
                 const double* xmin, const double* xmax, const double* df0dx, const double* dfdx,
                 const double* lam_host, double* out_host, void* stream);

/* ---- problem builders on the device --------------------------------------- */
/* The per-node and per-element set-up of a voxel problem, bit-identical to the
 * reference's arithmetic; all arrays are device memory.
 * pyto_grid_arrays: node coordinates (3 per node, i * h), element corner
 * nodes (8 per element, int32) and structural element DOFs (24 per element,
 * int64); any output may be null.  Replaces grid_mesh's node and element
 * loops and createEdofMatStructural (src/hex_mesher.py:104-152,1080-1095).
 * pyto_plane_mask: ORs comp_bits (bit c = DOF c) into mask[n] for the nodes of
 * the plane index[axis] == (at_max ? n_axis : 0); the fixed-DOF mask of
 * pyto_create from getNodesOnBoundingBoxPlane (src/hex_mesher.py:732-760).
 * pyto_face_load: f[dofs_per_node n + comp] = -w (total / sum w) on that plane,
 * w = 4 inside, 2 on an edge, 1 at a corner (src/hex_structural_examples.py:
 * 360-395); other entries are left as they are.
 * pyto_node_groups: node_group[n] = box_to_group[tx[ix] + nbx (ty[iy] + nby
 * tz[iz])] with per-axis box tables of nx+1, ny+1, nz+1 entries (-1 for a box
 * index outside the table); the assignment step of create_deflation_groups
 * (src/deflation.py:140-160,205-207).
 * pyto_group_sums: per group g and coordinate c, the sum of the coordinates of
 * the group's nodes in node order inside the half-open node box
 * bbox[6g .. 6g+5] = (x0, x1, y0, y1, z0, z1), sequential like numpy's add.at
 * (src/deflation.py:163-165,212-214); node_group null = every node of the box. */
int pyto_grid_arrays(int nx, int ny, int nz, double hx, double hy, double hz, double* node_xyz,
                     int32_t* elem_nodes, int64_t* edof, void* stream);
int pyto_plane_mask(int nx, int ny, int nz, int axis, int at_max, int comp_bits, uint8_t* mask,
                    void* stream);
int pyto_face_load(int nx, int ny, int nz, int axis, int at_max, int comp, int dofs_per_node,
                   double total, double* f, void* stream);
int pyto_node_groups(int nx, int ny, int nz, const int32_t* tx, const int32_t* ty,
                     const int32_t* tz, int nbx, int nby, int nbz, const int32_t* box_to_group,
                     int32_t* node_group, void* stream);
int pyto_group_sums(int nx, int ny, int nz, double hx, double hy, double hz,
                    const int32_t* node_group, int ngroups, const int32_t* bbox, double* sums,
                    void* stream);

/* ---- design clean-up ----------------------------------------------------- */
/* Connected components of the elements with x > threshold under the
 * reference's 27-entry element neighbour array (face, edge and corner
 * neighbours, src/hex_mesher.py:126-151): labels (int32 per element, device)
 * receives -1 for an absent element and otherwise the smallest element index
 * of its component, so the components come ordered as
 * HexMesher.find_connected_components (src/hex_mesher.py:1285-1339) lists
 * them.  Synchronises `stream`; *rounds_host (may be null) = propagation
 * rounds run. */
int pyto_element_components(int nx, int ny, int nz, const double* x, double threshold,
                            int32_t* labels, int* rounds_host, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYTO_HIP_H */
