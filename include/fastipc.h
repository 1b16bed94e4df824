/*
 * fastipc.h -- C ABI of libfastipc.so, the MI355X (gfx950) implementation of
 * the FastIPC projected-Newton IPC time step.
 *
 * Two groups of entry points:
 *
 *  1. Scalar drop-ins with the names, argument order and by-value / double*
 *     convention of the reference's common/math/wrapper/wrapper.cpp, so that
 *     every ti.external_func_call site of the reference can bind them
 *     unchanged (FP64, reentrant, host-side).
 *
 *  2. Batched device entry points on a context (fipc_*), one per step
 *     function of projects/FastIPC/NEWTON_IPC.py.  All return 0 on success or
 *     a FIPC_ERR_* status.  A context owns its device buffers; calls on one
 *     context must come from one host thread and are ordered on the
 *     context's HIP stream (fipc_assemble and fipc_ccd_alpha also run part
 *     of their work -- the contact rows, the edge-edge CCD half -- on a
 *     second stream of the context, which the first waits for before the
 *     call's later work).  Host arrays are copied in and out.
 *
 * Vertex arrays are float64 (nv, 3) row-major; index arrays are int32.
 * Constraint types are numbered as the reference's row sets:
 *   0 PP (2 ints/row), 1 PE (3), 2 PT (4), 3 EE (4), 4 EEM (4), 5 PPM (4), 6 PEM (4)
 */
#ifndef FASTIPC_H
#define FASTIPC_H
#ifdef __cplusplus
extern "C" {
#endif

#define FIPC_OK 0
#define FIPC_ERR_ARG 1          /* invalid argument */
#define FIPC_ERR_HIP 2          /* HIP runtime error */
#define FIPC_ERR_CAPACITY 3     /* a fixed limit exceeded: the packed-key limit (2^21 vertices), a batch maximum */
#define FIPC_ERR_NONFINITE 4    /* NaN/Inf in a result */
#define FIPC_ERR_NOT_CONVERGED 5 /* PCG hit its iteration limit */
#define FIPC_ERR_STATE 6        /* call out of order (e.g. no mesh) */

#define FIPC_TYPES 7

/* ---------------------------------------------------------------------- */
/* 1. scalar drop-ins                                                       */
/* ---------------------------------------------------------------------- */
/* replaces wrapper.cpp:50-56 project_pd_3 (JGSL::makePD, UTILS.h:7-25) */
void project_pd_3(double in_0, double in_1, double in_2, double in_3, double in_4, double in_5,
                  double in_6, double in_7, double in_8,
                  double* out_0, double* out_1, double* out_2, double* out_3, double* out_4,
                  double* out_5, double* out_6, double* out_7, double* out_8);
/* replaces wrapper.cpp:66-88 project_pd_6 / project_pd_9 / project_pd_12: n*n
 * doubles by value (row-major) followed by n*n double* outputs */
void project_pd_6(double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double,
                   double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*);
void project_pd_9(double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double,
                   double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*);
void project_pd_12(double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double,
                   double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*);
/* replaces wrapper.cpp:42-48 project_pd_2 and :58-64 project_pd_4 (JGSL::makePD) */
void project_pd_2(double, double, double, double,
                  double*, double*, double*, double*);
void project_pd_4(double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double,
                  double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*, double*);
/* replaces wrapper.cpp:90-198 solve_2 / solve_3 / solve_4 / solve_6 / solve_9:
 * n*n matrix entries and n right-hand-side entries by value, n outputs.  LDL^T
 * with diagonal pivoting, D clamped from below at 1e-8, then (L D L^T) x = rhs
 * in the pivoted order, as the reference computes it (which leaves the
 * permutation in place: x solves the clamped P F P^T) */
void solve_2(double, double, double, double, double, double,
             double*, double*);
void solve_3(double, double, double, double, double, double, double, double, double, double, double, double,
             double*, double*, double*);
void solve_4(double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double,
             double*, double*, double*, double*);
void solve_6(double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double,
             double*, double*, double*, double*, double*, double*);
void solve_9(double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double, double,
             double*, double*, double*, double*, double*, double*, double*, double*, double*);
/* replaces wrapper.cpp:12-25 svd_2 (SINGULAR_VALUE_DECOMPOSITION.h:25-105):
 * A = U diag(s00, s11) V^T, U and V rotations, s00 >= s11 (s11 may be negative) */
void svd_2(double F00, double F01, double F10, double F11, double* U00, double* U01, double* U10, double* U11,
           double* s00, double* s01, double* s10, double* s11, double* V00, double* V01, double* V10, double* V11);
/* array form of the projection for n = 3, 4, 6, 9, 12 (JGSL::makePD); n = 2 is
 * the neo-Hookean 2x2 closed form of common/math/math_tools.py:39-64 */
void fipc_project_pd(int n, const double* in, double* out);
/* replaces wrapper.cpp:27-40 svd_3 (rotation-variant SVD, det U = det V = 1) */
void svd_3(double F00, double F01, double F02, double F10, double F11, double F12, double F20, double F21, double F22,
           double* U00, double* U01, double* U02, double* U10, double* U11, double* U12, double* U20, double* U21, double* U22,
           double* s00, double* s01, double* s02, double* s10, double* s11, double* s12, double* s20, double* s21, double* s22,
           double* V00, double* V01, double* V02, double* V10, double* V11, double* V12, double* V20, double* V21, double* V22);
/* replaces wrapper.cpp:200-218 */
void get_smallest_positive_real_quad_root(double a, double b, double c, double tol, double* ret);
/* replaces wrapper.cpp:220-252 */
void get_smallest_positive_real_cubic_root(double a, double b, double c, double d, double tol, double* ret);
/* replaces wrapper.cpp:254-290 (CTCD::vertexFaceCTCD, EVCTCD/CTCD.cpp:383-472) */
void point_triangle_ccd(double p0, double p1, double p2, double t00, double t01, double t02,
                        double t10, double t11, double t12, double t20, double t21, double t22,
                        double dp0, double dp1, double dp2, double dt00, double dt01, double dt02,
                        double dt10, double dt11, double dt12, double dt20, double dt21, double dt22,
                        double eta, double dist2_cur, double* ret);
/* replaces wrapper.cpp:292-328 (CTCD::edgeEdgeCTCD, EVCTCD/CTCD.cpp:245-381) */
void edge_edge_ccd(double a00, double a01, double a02, double a10, double a11, double a12,
                   double b00, double b01, double b02, double b10, double b11, double b12,
                   double da00, double da01, double da02, double da10, double da11, double da12,
                   double db00, double db01, double db02, double db10, double db11, double db12,
                   double eta, double dist2_cur, double* ret);

/* ---------------------------------------------------------------------- */
/* 2a. batched per-constraint / per-element functions on the device         */
/* ---------------------------------------------------------------------- */
/* {PP,PE,PT,EE,EEM,PPM,PEM}_g_and_H (common/math/ipc.py:66-329) for n
 * constraints of one type.  x: n*k*3 points (k = 2,3,4 by type; mollified
 * types also read x0: n*4*3 rest points).  Outputs E[n], g[n*3k],
 * H[n*3k*3k] (row-major, SPD-projected when project != 0).  Runs the same
 * kernels as the assembly. */
int fipc_batch_constraint(int type, int n, const double* x, const double* x0, double dHat2, double kappa,
                          int project, double* E, double* g, double* H);
/* neo-Hookean tet (common/physics/neo_hookean.py + NEWTON_IPC.py:568-691):
 * F[n*9] row-major -> Psi[n], P[n*9], and the 12x12 Hessian of the tet
 * energy with respect to its vertices for a rest shape with D_m^{-1} = IB[n*9],
 * scaled by scale[n] (dt^2 vol0).  Hout may be NULL. */
int fipc_batch_tet(int n, const double* F, const double* IB, const double* scale, const double* la,
                   const double* mu, double* psi, double* P, double* Hout);
/* CCD on the device: type 0 point-triangle, 1 edge-edge.  pts[n*12], disp[n*12],
 * dist2[n] -> toc[n] (wrapper.cpp:254-328 semantics) */
int fipc_batch_ccd(int type, int n, const double* pts, const double* disp, double eta, const double* dist2, double* toc);
/* fipc_batch_ccd without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * device copies of the inputs and of toc live in a fixed device array of
 * FIPC_CCD_NOALLOC_MAX pairs per GPU (3.25 MiB, held until the process
 * ends) instead of buffers allocated per call; larger n returns
 * FIPC_ERR_CAPACITY.  Calls that overlap on one GPU share that array and
 * race.  Same arguments and results as fipc_batch_ccd. */
#define FIPC_CCD_NOALLOC_MAX 16384
int fipc_batch_ccd_noalloc(int type, int n, const double* pts, const double* disp, double eta, const double* dist2,
                           double* toc);

/* host-side evaluation of the same per-constraint / per-tet math that the
 * kernels run (for CPU tests and debugging; no GPU needed).  Arguments as for
 * the batched versions with n = 1; fipc_host_tet takes scale = dt^2 vol0. */
int fipc_host_constraint(int type, const double* x, const double* x0, double dHat2, double kappa, int project,
                         double* E, double* g, double* H);
int fipc_host_tet(const double* F, const double* IB, double scale, double la, double mu, double* psi, double* P,
                  double* H);

/* ---------------------------------------------------------------------- */
/* 2b. the Newton step on a context                                        */
/* ---------------------------------------------------------------------- */
typedef struct fipc_ctx fipc_ctx;

int fipc_create(fipc_ctx** out);
int fipc_destroy(fipc_ctx* c);
/* mesh and rest state.  IB = D_m^{-1} per tet (row-major 3x3), vol = rest
 * volume, la/mu Lame parameters, m lumped vertex mass, x0 rest positions
 * (used by the mollifier), boundary points/edges/triangles
 * (common/math/graph_tools.py:26-45), dirichlet: nv bytes, nonzero = fixed
 * (NEWTON_IPC.py:505-511, :750-772). */
int fipc_upload_mesh(fipc_ctx* c, int nv, const double* x0, const double* m, int nt, const int* tets,
                     const double* IB, const double* vol, const double* la, const double* mu,
                     int nbp, const int* bp, int nbe, const int* be, int nbt, const int* bt,
                     const unsigned char* dirichlet);
/* dt, dHat (not squared), kappa, 1/cell size of the spatial hash, gravity on axis 1 */
int fipc_set_params(fipc_ctx* c, double dt, double dHat, double kappa, double inv_dx, double gravity);
int fipc_set_state(fipc_ctx* c, const double* x, const double* v);
int fipc_get_state(fipc_ctx* c, double* x, double* v);
/* compute_xn_and_xTilde + move_nodes (NEWTON_IPC.py:497-511): xn = x,
 * xTilde = x + dt v + dt^2 g e_1; Dirichlet vertices get xTilde = target
 * (target: nv*3, only Dirichlet rows are read).  x of Dirichlet vertices is
 * left where it is; the solve prescribes their increment xTilde - x. */
int fipc_begin_step(fipc_ctx* c, const double* target);
/* the same with the prescribed positions of the listed rows only (the
 * Dirichlet vertices; rows[n], target[n][3]); the other rows are not used */
int fipc_begin_step_rows(fipc_ctx* c, int n, const int* rows, const double* target);
/* compute_v (NEWTON_IPC.py:799-802) */
int fipc_end_step(fipc_ctx* c);
/* find_constraints (NEWTON_IPC.py:337-357): counts[7] after deduplication */
int fipc_find_constraints(fipc_ctx* c, int* counts);
/* broad-phase candidates of the last fipc_find_constraints: cand[0] the
 * (boundary point, boundary triangle) pairs, cand[1] the (edge, edge) pairs
 * i < j without a shared vertex, whose boxes overlap within dHat
 * (point_triangle_ccd_broadphase / edge_edge_ccd_broadphase, ipc.py:332-354)
 * and so go on to the closest-feature distance test (NEWTON_IPC.py:181-312).
 * Counted only while fipc_search_stats(c, 1) is on (the counting costs the
 * queries time); -1 otherwise.  No counterpart in the reference. */
int fipc_search_stats(fipc_ctx* c, int on);
int fipc_search_counts(fipc_ctx* c, int* cand);
/* rows of one type, in lexicographic order (np.unique order,
 * NEWTON_IPC.py:348-356): PP/PE are radix-sorted and deduplicated on the
 * device, the four-index types (unique by construction) merge-sorted;
 * rows_out: counts[type]*width */
int fipc_get_constraints(fipc_ctx* c, int type, int* rows_out);
/* output_current_minimal_distance (NEWTON_IPC.py:815-839) */
int fipc_min_distance(fipc_ctx* c, double* d2min);
/* compute_hessian_and_gradient (NEWTON_IPC.py:736-747): precision 0 = FP64
 * (every block and gradient entry sums its tets in index order),
 * 1 = FP32 Hessian storage (FP32 tet blocks summed per chunk of up to 64
 * consecutive tets in FP32, the chunk partials in FP64, stored in FP32; the
 * gradient stays FP64, summed per chunk, then over chunks; contact rows in
 * FP32 too except those closer than 1e-2 dHat).  The contact rows' gradient
 * and diagonal blocks are summed per vertex in a fixed order; no sum uses
 * atomics, so the result is the same bits on every run. */
int fipc_assemble(fipc_ctx* c, int precision);
/* solve_system (NEWTON_IPC.py:750-777) with block-Jacobi PCG instead of
 * CHOLMOD: relative residual tolerance tol, at most maxit iterations */
int fipc_solve(fipc_ctx* c, double tol, int maxit, int* iters, double* relres);
/* continues the last fipc_solve of the current system (no fipc_assemble,
 * fipc_upload_mesh or fipc_set_vector of anything but the solution since)
 * from its solution to the relative tolerance tol, against the same initial
 * residual norm; iters counts the added iterations.  Lets a caller run the
 * search directions of an inexact Newton method at a loose tolerance and
 * tighten only the solve whose direction decides convergence.  No
 * counterpart in the reference, whose direct solve is exact. */
int fipc_solve_refine(fipc_ctx* c, double tol, int maxit, int* iters, double* relres);
/* fipc_solve_refine that ends early when the exact direction p* provably
 * passes the Newton test max|p*| < bound: since A >= M (lumped mass; the
 * elastic and barrier Hessians are projected to PSD), max|p*| <= max|p| +
 * |b - A p|_2 / (m_min / 2) with m_min the least lumped mass of a free
 * vertex (halved for the FP32 storage of A).  Checked with the FP64
 * residual of a restart; *certified = 1 when it held.  FP32 mode only (in
 * FP64 mode it is fipc_solve_refine).  No counterpart in the reference,
 * whose direct solve is exact. */
int fipc_solve_certify(fipc_ctx* c, double tol, int maxit, double bound, int* iters, double* relres,
                       int* certified);
/* output_residual (NEWTON_IPC.py:805-812): max |p_i| */
int fipc_residual(fipc_ctx* c, double* res);
/* compute_energy (NEWTON_IPC.py:514-542); parts[3] = inertia, elastic, barrier (may be NULL) */
int fipc_energy(fipc_ctx* c, double* E, double* parts);
/* compute_intersection_free_step_size (NEWTON_IPC.py:456-468): PT and EE CCD
 * filters and (inversion != 0) the per-tet inversion filter; parts[3] may be NULL */
int fipc_ccd_alpha(fipc_ctx* c, double eta, int inversion, double* alpha, double* parts);
/* candidate pairs (point-triangle, edge-edge) that passed the lower-bound
 * filter in the last fipc_ccd_alpha and went to the narrow phase */
int fipc_ccd_counts(fipc_ctx* c, int* pairs);
/* save_xPrev (NEWTON_IPC.py:786-789) */
int fipc_save_prev(fipc_ctx* c);
/* apply_sol (NEWTON_IPC.py:792-796): x = xPrev + alpha p */
int fipc_apply_sol(fipc_ctx* c, double alpha);
/* check_collision (NEWTON_IPC.py:855-886) */
int fipc_check_collision(fipc_ctx* c, int* hit);
/* accessors for tests and drivers */
int fipc_get_vector(fipc_ctx* c, int which, double* out);   /* 0 x, 1 p (solution), 2 rhs, 3 xTilde */
int fipc_set_vector(fipc_ctx* c, int which, const double* in);
/* BSR of the elastic + inertia Hessian, both triangles: rowptr[nv+1],
 * col[nnzb], val[nnzb*9] (block row-major, columns ascending per row).  The
 * device stores the upper triangle only; this expands it.  With rowptr, col
 * and val all null only *nnzb is written. */
int fipc_get_bsr(fipc_ctx* c, int* nnzb, int* rowptr, int* col, double* val);
/* blocks the device stores (upper triangle incl. diagonal) and the lower
 * blocks the SpMV reads transposed through (block, column) index pairs */
int fipc_bsr_storage(fipc_ctx* c, int* nupper, int* nlower);
/* y = A x with the assembled Hessian (elastic BSR + contact rows), Dirichlet rows zeroed.
 * FIPC_ERR_STATE if constraints were found again since the last fipc_assemble
 * (the product would mix old Hessians with new rows); the same holds for
 * fipc_solve and fipc_solve_refine. */
int fipc_spmv(fipc_ctx* c, const double* x, double* y);
/* timing of the hot kernels with HIP events on the context stream, recorded
 * without synchronisation while enabled and folded when read.  Classes:
 * 0 SpMV inside the PCG iterations (contact + BSR rows), 1 per-tet elastic
 * pass, 2 CCD query, 3 constraint / collision query, 4 energy, 5 contact
 * assembly, 6 PCG vector updates, 7 grid build (phase time on the stream),
 * 8 assembly gathers, 9 residual SpMVs (initial guess, FP64 refinement),
 * 10 CCD narrow phase (the root finding over the candidate pairs; class 2 is
 * then the candidate queries).  Writes the first n (<= 11) totals (ms) and
 * launch counts. */
int fipc_enable_timing(fipc_ctx* c, int on);
int fipc_kernel_times(fipc_ctx* c, int n, double* ms, long* launches);
int fipc_sync(fipc_ctx* c);
/* plain device-to-device copy of nbytes (multiple of 16) between two
 * temporary buffers on the context stream; *ms = its HIP-event time.  A
 * bandwidth reference and the calibration kernel of hardware byte counters
 * (no counterpart in the reference). */
int fipc_stream_copy(fipc_ctx* c, long nbytes, double* ms);

/* ---------------------------------------------------------------------- */
/* 3. learned Newton warm start.  No counterpart in the reference: its only */
/*    contact point is the initial guess x = x_n of NEWTON_IPC.py:958-960,  */
/*    which stays the fallback.                                             */
/* ---------------------------------------------------------------------- */
typedef struct fipc_mlp fipc_mlp;
/* per-vertex network n_in -> depth x width (SiLU) -> n_out (linear); products
 * on the BF16 matrix cores with FP32 accumulation, FP32 master parameters
 * packed per layer as W [fan_in][fan_out] then b [fan_out]; initialised from
 * seed (variance 1/fan_in, output layer 10x smaller) */
int fipc_mlp_create(fipc_mlp** out, int n_in, int width, int depth, int n_out, unsigned long long seed);
int fipc_mlp_destroy(fipc_mlp* m);
int fipc_mlp_nparams(fipc_mlp* m, long* n);
int fipc_mlp_get_params(fipc_mlp* m, float* w);
/* sets the parameters and resets the Adam state */
int fipc_mlp_set_params(fipc_mlp* m, const float* w);
/* Y[n][n_out] = network(X[n][n_in]) (host arrays) */
int fipc_mlp_forward(fipc_mlp* m, int n, const float* X, float* Y);
/* mean squared error over n rows and its gradient (nparams, packed like the
 * parameters; grad may be NULL) */
int fipc_mlp_grad(fipc_mlp* m, int n, const float* X, const float* T, float* loss, float* grad);
/* device-resident training set of (feature, target) rows */
int fipc_mlp_dataset_clear(fipc_mlp* m);
int fipc_mlp_dataset_size(fipc_mlp* m, long* n);
int fipc_mlp_dataset_add(fipc_mlp* m, int n, const float* X, const float* T);
int fipc_mlp_dataset_get(fipc_mlp* m, long first, long n, float* X, float* T);
/* iters Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on batches of `batch`
 * rows drawn from the training set; losses[iters] (may be NULL) = mean
 * squared error of each batch before its update */
int fipc_mlp_train(fipc_mlp* m, int iters, int batch, float lr, unsigned long long seed, float* losses);

/* Warm start on a context, after fipc_begin_step and fipc_find_constraints
 * at x_n.  Features per vertex (16, scaled by `scale`, a length):
 *   0-2 v dt, 3-5 the previous step's correction x_n - xTilde_{n-1},
 *   6-8 the gravity increment dt^2 g, 9 nearest contact distance / dHat
 *   (1 without contact), 10-12 unit normal away from that contact (0 without),
 *   13 Dirichlet flag, 14-15 the contact partner's inertial (or prescribed)
 *   increment relative to the vertex's own: approach along the normal and
 *   tangential size (0 without contact).
 * fipc_ws_features copies them to feat[nv][16] (may be NULL).
 * fipc_ws_predict writes the increment
 *   p = xTilde - x_n + c_prev + scale * network(f),
 * c_prev = the previous step's correction (features 3-5 times scale), i.e.
 * the network predicts the change of the correction; Dirichlet vertices get
 * their prescribed increment.  p is the solution vector, to be filtered by
 * fipc_ccd_alpha and applied with fipc_apply_sol.
 * fipc_ws_capture_begin keeps the features of nsample free vertices drawn
 * from seed, a share near_frac of them (when there are any) among the
 * vertices within d-hat of a contact primitive (feature 9 below 1);
 * fipc_ws_capture_end (after the step, before fipc_end_step)
 * appends them to the training set with targets
 * ((x_{n+1} - xTilde) - c_prev) / scale. */
int fipc_ws_features(fipc_ctx* c, double scale, float* feat);
int fipc_ws_predict(fipc_ctx* c, fipc_mlp* m, double scale);
int fipc_ws_capture_begin(fipc_ctx* c, fipc_mlp* m, double scale, int nsample, unsigned long long seed,
                          double near_frac);
int fipc_ws_capture_end(fipc_ctx* c, fipc_mlp* m, double scale);

#ifdef __cplusplus
}
#endif
#endif
