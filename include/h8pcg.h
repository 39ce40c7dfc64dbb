/* h8pcg.h -- C ABI of the MI355X (gfx950) H8 linear-elastic solver library.
 *
 * One shared library, libh8pcg.so, built from computation-suite_amd/csrc/.
 * Every buffer argument is a DEVICE pointer (for example torch.Tensor.data_ptr()
 * of a tensor on the GPU); `stream` is a hipStream_t passed as void* (0 = the
 * null stream).  Functions return an int status (H8_OK on success) and never
 * throw.  h8_last_error() returns a thread-local message for the last failure.
 *
 * Reference citations are file:line in abhi2k16/computation-suite, with
 * fe/ = fea_engine/src/fea_engine/.
 *
 * Structured grids.  A `h8_grid` is the box mesh of fe/mesh.py:1288-1289:
 * node(i,j,k) = k(nx+1)(ny+1) + i(ny+1) + j (y fastest), element
 * e(i,j,k) = k*nx*ny + i*ny + j, corner order [n(i,j,k), n(i+1,j,k),
 * n(i+1,j+1,k), n(i,j+1,k)] then the same at k+1, DOF 3*node + c.  Every
 * element shares one 24x24 matrix K0 (uniform spacing) scaled by a per-element
 * factor s_e (per-element Young's modulus, or the SIMP scale).  Full-length
 * vectors have 3*(nx+1)(ny+1)(nz+1) entries; fixed DOFs are marked by a zero in
 * `dinv` (the inverse Jacobi diagonal), which keeps every Krylov vector zero
 * there -- the same iterates as the reduced system K_ff u_f = F_f.
 *
 * A `K0` argument is H8_K0_WORDS values: the 24x24 element matrix (row-major,
 * entry 24*i + j) followed by its 72-value reflection-sector table, which
 * h8_sector_table_f64 fills from the matrix.  An FP32 K0 (h8_spmv_f32,
 * h8_refine_solve, h8_time_kernels at precision 32) is H8_K0_WORDS_F32 words:
 * the cast of the FP64 one followed by the FP64 table itself.  Only the sector
 * and resident kernels read the tables.
 */
#ifndef H8PCG_H
#define H8PCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped to the reference's exception types by the host) */
#define H8_OK 0
#define H8_ERR_ARG 1        /* ValueError: bad shape/size/pointer            */
#define H8_ERR_HIP 2        /* RuntimeError: a HIP call failed               */
#define H8_NOT_SPD 3        /* LinAlgError: p^T A p <= 0                     */
#define H8_MAXITER 4        /* RuntimeError: not converged in maxiter        */
#define H8_ZERO_DIAG 5      /* ValueError: zero Jacobi diagonal at a free DOF */
#define H8_ERR_WORKSPACE 6  /* ValueError: workspace too small               */
#define H8_MG_DECLINED 7    /* the multigrid preconditioner does not apply (singular
                               coarse operator): solve with the Jacobi form instead */

#define H8_VARIANT_AUTO 0   /* library default for the precision             */
#define H8_VARIANT_NODE 1   /* node-centric gather (VALU), deterministic     */
#define H8_VARIANT_MFMA 2   /* element-batched 24x24 contraction on MFMA     */
#define H8_VARIANT_NODE_PERSIST 3 /* node-centric, persistent XCD-local brick runs */
#define H8_VARIANT_NODE_SPLIT 4   /* node-centric, two threads per node (elements above / below) */
#define H8_VARIANT_SECTOR 5       /* element products through the reflection sectors, marching in z */
#define H8_VARIANT_SECTOR_NARROW 6 /* the sector kernel with 8 x 16 element columns (half the LDS) */
#define H8_VARIANT_RESIDENT 7     /* PCG loop in one cooperative launch, vectors resident in LDS
                                     (grids that fit one block per CU; other products: the default) */

#define H8_K0_WORDS 648           /* 576 matrix entries + 72 sector-table entries */
#define H8_K0_WORDS_F32 792       /* an FP32 K0: the cast of the FP64 K0 (648 words), then the
                                     FP64 sector table bit-copied (72 doubles = 144 words, 8-byte
                                     aligned) for FP64 element products of FP32 data */

typedef struct h8_grid {
    int32_t nx, ny, nz;     /* elements per axis (>= 1)                      */
    int32_t variant;        /* H8_VARIANT_* for the K*u kernels              */
    int32_t layers;         /* node layers per block of H8_VARIANT_SECTOR (0 = library choice) */
    int32_t arith;          /* element-product arithmetic of the FP32 sector and resident kernels:
                               0 = library choice (32), 32 = FP32 products, 64 = FP64 products
                               of the FP32 data (outputs rounded to FP32).  The node-centric
                               and MFMA kernels compute in their storage precision. */
    int32_t mg_degree;      /* Chebyshev steps per smoothing pass on the finest level of the
                               multigrid preconditioner (h8_refine_solve_mg, h8_mg_vcycle_f32):
                               0 = library choice (4 up to 2^20 DOFs, 3 above), else 2..6;
                               coarser levels take 2 */
} h8_grid;

typedef struct h8_pcg_stats {
    int64_t iterations;     /* k at which ||r_k||/||b|| < tol (0: early out) */
    int32_t status;         /* H8_OK, H8_NOT_SPD or H8_MAXITER               */
    int32_t outer;          /* FP64 residual replacements (refined solve)    */
    double rel_residual;    /* recurrence ||r_k||/||b|| at exit              */
    double true_rel_residual; /* FP64 ||f - K u||/||f|| (refined), else -1   */
    double bnorm;           /* ||b|| (1 when b = 0)                          */
    double pAp;             /* the offending p^T A p when status = NOT_SPD   */
    double elapsed_ms;      /* device time of the solve                      */
    int64_t inner_iterations; /* FP32 iterations (refined solve)             */
    int64_t launches;       /* graph launches (host polls)                   */
} h8_pcg_stats;

/* ---- diagnostics ------------------------------------------------------- */
const char* h8_last_error(void);
int h8_version(void);
/* Writes the device name / CU count / arch of the current device. */
int h8_device_info(char* buf, int buflen);
/* Destroys what the library keeps between calls: its cached HIP graphs (the
 * captured PCG iteration chunks) and its per-device capture streams.  Call it
 * before the HIP runtime is torn down (the Python binding registers it with
 * atexit); the next solve re-creates what it needs.  No counterpart in the
 * reference (it keeps no device state). */
int h8_release(void);

/* Fills K0[576 .. 647] with the reflection-sector table of the 24x24 matrix in
 * K0[0 .. 575] (device pointer): Kh[psi][c][c'] = (1/8) sum_x
 * (-1)^popcount((psi ^ 2^c') & x) K0[c][3x + c'] for the corner bit masks x.
 * No counterpart in the reference (a layout helper of this library). */
int h8_sector_table_f64(double* K0, void* stream);

/* ---- element stiffness ------------------------------------------------- */
/* ke[e] = sum over the 2x2x2 Gauss rule of B^T D B detJ w, general geometry.
 * X: (m,8,3) nodal coordinates; D: (6,6) when d_stride == 0, else (m,36) with
 * d_stride = 36; ke: (m,24,24).  Replaces Element.stiffness +
 * Hex8Solid3D.B_matrix (fe/elements/base.py:494-514, fe/elements/solids.py:
 * 144-158) and tensorized_element_stiffness (fe/vectorized_assembly.py:165-214). */
int h8_ke_batched_f64(const double* X, const double* D, int64_t d_stride, int64_t m,
                      double* ke, void* stream);

/* The B-bar (mean-dilatation) Hex8 stiffness, same arguments and layout:
 * B(p) with its volumetric part replaced by the element volume average at
 * every Gauss point (Hex8SolidBbar._bbar_matrices / stiffness,
 * fe/elements/solids.py:225-262). */
int h8_ke_bbar_batched_f64(const double* X, const double* D, int64_t d_stride, int64_t m,
                           double* ke, void* stream);

/* ---- structured matrix-free operator ------------------------------------ */
/* dinv[i] = 1/K_ii at free DOFs (free_mask[i] != 0), 0 at fixed DOFs, with
 * K_ii = sum_e s_e K0_aa.  Returns H8_ZERO_DIAG if a free diagonal is 0.
 * Replaces jacobi_preconditioner / jacobi_preconditioner_torch
 * (fe/iterative_solvers.py:193-202,305-321). */
int h8_diag_f64(const h8_grid* g, const double* s_e, const double* K0,
                const uint8_t* free_mask, double* dinv, int32_t* d_flag, void* stream);
/* h8_diag_f64 without the host wait, for many models in a row (the per-model
 * jacobi_preconditioner of solve_static_batched / batch.map, fe/batched_solve.py:
 * 105-164, fe/batch.py:60-128): ORs 1 into *d_flag where a free diagonal is 0.
 * The caller zeroes *d_flag before the first call and checks it once after the
 * last (H8_ZERO_DIAG semantics). */
int h8_diag_async_f64(const h8_grid* g, const double* s_e, const double* K0,
                      const uint8_t* free_mask, double* dinv, int32_t* d_flag, void* stream);

/* y = K u over the whole grid (every DOF), K = assembled sum_e s_e K0.
 * Replaces the CSR product A @ p of preconditioned_cg and _cg_torch
 * (fe/iterative_solvers.py:439-441, fe/torch_sparse_solver.py:183) and the
 * matrix-free _Ku_free of fe/topopt.py:214-220. */
int h8_spmv_f64(const h8_grid* g, const double* s_e, const double* K0,
                const double* u, double* y, void* stream);
int h8_spmv_f32(const h8_grid* g, const float* s_e, const float* K0,
                const float* u, float* y, void* stream);

/* Workspace for h8_pcg_* / h8_refine_solve (bytes), precision 64 or 32 (32 =
 * the refined solve, which needs both FP32 and FP64 vectors); 0 (and
 * h8_last_error) for an invalid grid or precision. */
size_t h8_pcg_workspace_bytes(const h8_grid* g, int precision);

/* Jacobi-PCG, Hestenes-Stiefel recurrence of preconditioned_cg
 * (fe/iterative_solvers.py:434-469) on the masked full-length system:
 * x0 = x if use_x0 else 0, bnorm = ||b|| (1 if 0), early return when
 * ||r0||/bnorm < tol, H8_NOT_SPD when p^T A p <= 0, stop when
 * ||r_k||/bnorm < tol, H8_MAXITER after maxiter iterations.  b must be zero at
 * fixed DOFs.  Replaces FESystem._solve_static_iterative's "pcg" branch
 * (fe/solver.py:1922-2029) and fesystem_solve_static_torch's CG
 * (fe/torch_sparse_solver.py:268-314). */
int h8_pcg_f64(const h8_grid* g, const double* s_e, const double* K0, const double* dinv,
               const double* b, double* x, int use_x0, double tol, int64_t maxiter,
               void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream);

/* h8_pcg_f64 without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops): the
 * workspace is a file-scope device array of H8_SCRATCH_BYTES bytes that each
 * GPU holds until the process ends, and the call creates no events, captures
 * no graphs and allocates nothing (elapsed_ms = -1).  Grids whose
 * h8_pcg_workspace_bytes(g, 64) exceeds H8_SCRATCH_BYTES are refused with
 * H8_ERR_WORKSPACE (about 590k DOFs).  Same recurrence, kernels and results
 * as h8_pcg_f64.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (two host threads, or two streams) race.  No
 * counterpart in the reference (its solvers allocate per call); it exists for
 * callers that cannot allocate. */
#define H8_SCRATCH_BYTES (32u << 20)
int h8_pcg_f64_noalloc(const h8_grid* g, const double* s_e, const double* K0, const double* dinv,
                       const double* b, double* x, int use_x0, double tol, int64_t maxiter,
                       h8_pcg_stats* stats, void* stream);

/* Mixed precision: FP32 Jacobi-PCG iterations (s_e32, K032, FP32 vectors,
 * FP64 dot products) whose residual is replaced by the FP64 true residual
 * r = f - K u (s_e64, K064) and whose iterate is folded into the FP64 u
 * whenever the FP32 residual has fallen by `delta` (reliable updates; the
 * search direction is kept).  Stops when the FP64 ||f - K u||/||f|| <= tol.
 * No counterpart in the reference (it has one precision); it returns the same
 * u as h8_pcg_f64 to within tol.  f zero at fixed DOFs; dinv64 as above. */
int h8_refine_solve(const h8_grid* g, const double* s_e64, const double* K064,
                    const float* s_e32, const float* K032, const double* dinv64,
                    const double* f, double* u, double tol, int64_t maxiter, double delta,
                    void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream);
/* h8_refine_solve started from the guess in u (its fixed-DOF entries are
 * zeroed first): the first FP64 true residual is f - K u, so a guess that
 * already meets tol returns after it with no FP32 iteration.  The role of the
 * x0 argument of preconditioned_cg (fe/iterative_solvers.py:362-469) and of the
 * warm-started forward solves of the SIMP loop (fe/topopt.py:302-377). */
int h8_refine_solve_x0(const h8_grid* g, const double* s_e64, const double* K064,
                       const float* s_e32, const float* K032, const double* dinv64,
                       const double* f, double* u, double tol, int64_t maxiter, double delta,
                       void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream);

/* Geometric multigrid preconditioner of the FP32 iterations.  The hierarchy
 * halves the even axes of the grid until a level has at most 128 DOFs; every
 * coarse operator is the Galerkin product P^T A P with trilinear P restricted
 * to the free DOFs (27-point block stencils, FP32), the coarsest is inverted
 * densely, smoothing is degree-2 Chebyshev-Jacobi on Gershgorin bounds.
 * h8_mg_workspace_bytes: the hierarchy's workspace (bytes), 0 when the grid
 * does not coarsen to at most 512 DOFs that way (then only the Jacobi form
 * applies).  No counterpart in the reference (its PCG preconditioners are
 * Jacobi, SSOR and IC(0), fe/iterative_solvers.py:193-360). */
size_t h8_mg_workspace_bytes(const h8_grid* g);

/* h8_refine_solve (use_x0 = 0) or h8_refine_solve_x0 (use_x0 = 1) with the
 * V-cycle as the preconditioner of the FP32 iterations (z = V(r) instead of
 * dinv r); same arguments, stopping rule, statistics and errors, plus the
 * hierarchy's workspace mg_ws.  The hierarchy is built from s_e64, K064 and
 * the free mask (dinv64 != 0) on every call.  Returns H8_MG_DECLINED when the
 * coarsest operator is singular (supports that leave rigid-body motions free):
 * nothing solved, u holds the initial guess with its fixed entries zeroed (0
 * without use_x0), as the Jacobi form starts from.  Returns the u of
 * h8_refine_solve to within tol. */
int h8_refine_solve_mg(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                       const float* K032, const double* dinv64, const double* f, double* u, double tol,
                       int64_t maxiter, double delta, int use_x0, void* ws, size_t ws_bytes, void* mg_ws,
                       size_t mg_ws_bytes, h8_pcg_stats* stats, void* stream);

/* One V-cycle z = V(r) (FP32 vectors, r zero at fixed DOFs) after building the
 * hierarchy: the preconditioner of h8_refine_solve_mg as an operator, for
 * checking it.  ws: a workspace of h8_pcg_workspace_bytes(g, 32).
 * Diagnostics helper with no counterpart in the reference. */
int h8_mg_vcycle_f32(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                     const float* K032, const double* dinv64, const float* r, float* z, void* ws,
                     size_t ws_bytes, void* mg_ws, size_t mg_ws_bytes, void* stream);

/* Average device time (ms) of the multigrid kernels, each graph-launched
 * `reps` times alone between HIP events on `stream` (after one whole cycle
 * has filled every level's vectors); avg_ms has H8_MG_TIMES slots:
 *   0 the level-0 FP32 product y = K x (sector kernel, plain mode),
 *   1 one Chebyshev step on level 1 (stencil product fused with the update),
 *   2 one whole V-cycle,
 *   3 one level-0 Chebyshev step (sector kernel),
 *   4 the level-0 first step pair from zero (sector kernel),
 *   5 the level-0 masked residual (sector kernel),
 *   6 the restriction 0 -> 1 with level 1's first step,
 *   7 the prolongation 1 -> 0,
 *   8 the single-workgroup kernel of the small levels,
 *   9 the level-1 masked residual (stencil product);
 * a slot that does not apply to the grid's hierarchy is 0.  info[0..3] (may
 * be NULL) = levels, DOFs and nodes of level 1, and the Chebyshev steps per
 * level-0 smoothing pass (2 deg - 2 level-0 steps of slot 3 per cycle).  ws:
 * h8_pcg_workspace_bytes(g, 32); mg_ws: h8_mg_workspace_bytes(g).
 * Measurement helper with no counterpart in the reference; scribbles on both
 * workspaces. */
#define H8_MG_TIMES 10
int h8_mg_time_kernels(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                       const float* K032, const double* dinv64, void* ws, size_t ws_bytes, void* mg_ws,
                       size_t mg_ws_bytes, int reps, double* avg_ms, int64_t* info, void* stream);

/* ---- element bilinear forms (SIMP sensitivities) ----------------------- */
/* out[e] = w[e] * u_e^T K0 v_e for every element (u_e = u at the element's 24
 * DOFs).  With w = -s'(rho), u = v this is compliance_sensitivity_closed_form
 * (fe/topopt.py:161-177); with w = s'(rho), u = lambda, v = displacement it is
 * the vector-Jacobian product of compliance_sensitivity_via_adjoint
 * (fe/topopt.py:180-229, fe/differentiable.py:478-512). */
int h8_elem_bilinear_f64(const h8_grid* g, const double* K0, const double* u, const double* v,
                         const double* w, double* out, void* stream);

/* s[e] = rho_min + (1-rho_min) rho^p and ds[e] = p (1-rho_min) rho^(p-1)
 * (simp_scale / simp_scale_grad, fe/topopt.py:74-90).  ds may be NULL. */
int h8_simp_scale_f64(const double* rho, double p, double rho_min, int64_t m,
                      double* s, double* ds, void* stream);

/* ---- batched: many structured instances x many load cases --------------- */
/* One instance of a batched solve (HOST array).  Instance i has
 * ndof_i = 3(nx+1)(ny+1)(nz+1) DOFs and nrhs load cases; its element scales
 * are s_e[elem_off .. elem_off + nx*ny*nz), its inverse Jacobi diagonal is
 * dinv[3*node_off .. 3*node_off + ndof_i) (zero at fixed DOFs, shared by its
 * cases), and case c occupies b/x[vec_off + c*ndof_i .. + ndof_i). */
typedef struct h8_instance {
    int32_t nx, ny, nz, nrhs;
    int64_t node_off, elem_off, vec_off;
} h8_instance;

/* Workspace (bytes) of h8_pcg_batched_f64 for this instance list. */
size_t h8_batched_workspace_bytes(const h8_instance* inst, int32_t n_inst, int64_t total_vec);

/* Jacobi-PCG of every (instance, case) system with the recurrence and
 * stopping rule of h8_pcg_f64, all systems advanced together: each launch
 * covers every system still running (per-system convergence masks), and the
 * iterations are captured in HIP graphs; one host poll per graph launch reads
 * the number of systems still running.  K0 is one 24x24 matrix
 * (k0_per_instance = 0) or one per instance (= 1).  x receives the solutions
 * (x0 = 0).  status_out/iters_out/rel_out are HOST arrays of length
 * sum(nrhs) (system order: instance by instance, case by case).  Replaces the
 * per-model loop of batch.map (fe/batch.py:60-128) and, for one instance, the
 * multi-RHS solve_static_batched (fe/batched_solve.py:105-164).  Returns H8_OK
 * when every system converged, else the first failing status. */
int h8_pcg_batched_f64(const h8_instance* inst, int32_t n_inst, int64_t total_nodes, int64_t total_elems,
                       int64_t total_vec, const double* s_e, const double* K0, int32_t k0_per_instance,
                       const double* dinv, const double* b, double* x, double tol, int64_t maxiter,
                       void* ws, size_t ws_bytes, int32_t* status_out, int64_t* iters_out,
                       double* rel_out, void* stream);

/* h8_pcg_batched_f64 in a chosen form (same arguments and results):
 *   H8_BATCHED_LAUNCHED (0): h8_pcg_batched_f64 itself -- every system
 *     advanced together, two launches per iteration in captured graphs;
 *   H8_BATCHED_PERSISTENT_8X16 (1), _16X16 (2): one persistent launch; each
 *     workgroup takes whole systems from a queue (largest first) and runs each
 *     one's PCG to the end with every reduction inside the workgroup, the
 *     direction kernel's sector march on 8 x 16 or 16 x 16 element columns.
 * The iterations of every form are those of h8_pcg_f64 up to the summation
 * order of the dot products.  The choice between them is the project's
 * batched autotuner's (h8pcg.batched).  The workspace is
 * h8_batched_workspace_bytes for every form.  Replaces, like
 * h8_pcg_batched_f64, the per-model loop of batch.map (fe/batch.py:60-128)
 * and the multi-RHS solve_static_batched (fe/batched_solve.py:105-164). */
#define H8_BATCHED_LAUNCHED 0
#define H8_BATCHED_PERSISTENT_8X16 1
#define H8_BATCHED_PERSISTENT_16X16 2
int h8_pcg_batched_form_f64(const h8_instance* inst, int32_t n_inst, int64_t total_nodes, int64_t total_elems,
                            int64_t total_vec, const double* s_e, const double* K0, int32_t k0_per_instance,
                            const double* dinv, const double* b, double* x, double tol, int64_t maxiter,
                            void* ws, size_t ws_bytes, int32_t* status_out, int64_t* iters_out,
                            double* rel_out, int32_t form, void* stream);

/* Device time (ms) and graph launches of this thread's last batched solve. */
double h8_batched_last_ms(void);
int64_t h8_batched_last_launches(void);

/* ---- stress / strain recovery ----------------------------------------- */
#define H8_RECOVERY_WORDS 252     /* 216 projected shape-function gradients + D (36) */

/* The recovery table of the brick element with spacing (hx, hy, hz):
 * A[c][d][b] = sum_p N_c(p) dN_b/dx_d(p) over the 2x2x2 Gauss points for the
 * 8 corners c, the volume average (1/8) sum_p dN_b/dx_d(p) as c = 8, then
 * the 6x6 D (device pointer; NULL = identity, i.e. strain).  Layout helper of
 * h8_recover_f64 (fe/recovery.py:193-231 evaluates B(p) per Gauss point). */
int h8_recovery_table_f64(double hx, double hy, double hz, const double* D, double* tab, void* stream);

/* Stress sigma = s_e D B u (strain with the identity table) recovered as the
 * reference does (fe/recovery.py:234-330, _block_loop / _block_batched /
 * _recover): Gauss-point values projected to the corners with the linear
 * shape functions, sum N_a val |J| w / sum N_a |J| w over the elements sharing
 * a node -> nodal (n_nodes x 6, Voigt xx, yy, zz, xy, yz, xz); volume averages
 * -> elem (m x 6); their von Mises values (fe/recovery.py:364-375) -> vm_nodal
 * (n_nodes), vm_elem (m).  Any output may be NULL; s_e NULL = 1 for every
 * element.  Replaces stress / strain / von_mises (fe/recovery.py:378-420). */
int h8_recover_f64(const h8_grid* g, const double* tab, const double* s_e, const double* u,
                   double* nodal, double* vm_nodal, double* elem, double* vm_elem, void* stream);

/* ---- linear multi-point constraints --------------------------------------- */
/* The reduced form of the constraints sum_i c_i u_i = value (add_constraint /
 * tie, fe/solver.py:1341-1397), as the host's reduction (fe/boundary.py:98-166)
 * leaves them: slave DOFs u_s = g_s + sum_m T_sm u_m over kept master DOFs.
 * Device arrays: for slave t, its masters s_master[s_ptr[t] .. s_ptr[t+1]) with
 * coefficients s_coef; the transpose by master likewise (m_ptr, m_slave,
 * m_coef).  Masters are never slaves. */
typedef struct h8_cmap {
    int64_t n_slave;
    const int64_t* slave;
    const int32_t* s_ptr;
    const int64_t* s_master;
    const double* s_coef;
    int64_t n_master;
    const int64_t* master;
    const int32_t* m_ptr;
    const int64_t* m_slave;
    const double* m_coef;
} h8_cmap;

/* v_s = sum_m T_sm v_m at every slave (the other entries unchanged). */
int h8_constraint_expand_f64(const h8_cmap* c, double* v, void* stream);
/* v_m += sum_s T_sm v_s at every master (T^T applied; the slave entries are
 * left for the caller to mask). */
int h8_constraint_contract_f64(const h8_cmap* c, double* v, void* stream);
/* dinv_m = 1 / diag(T^T K T)_m at every master, the K entries evaluated from
 * K0 and s_e (0 where that diagonal is not positive); the other entries of
 * dinv (1/K_ii at kept DOFs, 0 at fixed and slave DOFs, from h8_diag_f64 with
 * the slaves masked) are left as they are. */
int h8_constraint_diag_f64(const h8_grid* g, const double* s_e, const double* K0, const h8_cmap* c,
                           double* dinv, void* stream);
/* Jacobi-PCG (the recurrence and stopping rule of h8_pcg_f64, x0 = 0) on the
 * eliminated system T^T K T x = b over the kept DOFs, on full-length vectors
 * zero at fixed and slave DOFs (b contracted and masked by the caller; dinv
 * from h8_diag_f64 + h8_constraint_diag_f64).  The caller recovers
 * u = expand(x) + w.  Replaces the constrained solve of form_linear_system +
 * ReducedSystem.solve (fe/solver.py:1403-1479, fe/linear_system.py:101-140);
 * workspace of h8_pcg_workspace_bytes(g, 64). */
int h8_pcg_constrained_f64(const h8_grid* g, const double* s_e, const double* K0, const double* dinv,
                           const h8_cmap* c, const double* b, double* x, double tol, int64_t maxiter,
                           void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream);

/* ---- SIMP loop: sensitivity filter, optimality-criteria update ---------- */
/* Sigmund's sensitivity filter on the grid's elements (fe/topopt.py:235-252):
 * out_e = sum_j H_ej rho_j dc_j / max(max(rho_e, eps) sum_j H_ej, eps) with
 * H_ej = max(0, radius - |c_e - c_j|) between the element centroids (brick
 * spacing hx, hy, hz); the dense m x m H of build_filter_weights is a stencil
 * here (radius at most 64 element lengths).  rho, dc, out: m values. */
int h8_sens_filter_f64(const h8_grid* g, double hx, double hy, double hz, double radius,
                       const double* rho, const double* dc, double eps, double* out, void* stream);

/* Workspace (bytes) of h8_oc_update_f64 for m elements. */
size_t h8_oc_workspace_bytes(int64_t m);

/* Optimality-criteria density update (fe/topopt.py:258-296): dc clipped to
 * <= -1e-12; rho_new = clip(clip(rho sqrt(-dc / l), rho - move, rho + move),
 * rho_min_bound, rho_max_bound) with the volume multiplier l found as the
 * reference does -- the bracket [1e-12, 1] doubled at its upper end (at most
 * 60 times) until mean(rho_new) <= volume_fraction, then bisected until
 * l2 - l1 < bisection_tol max(l2, 1) or max_bisection_iter steps; rho_new is
 * the last bisection step's.  *change (HOST pointer, may be NULL) receives
 * max |rho_new - rho|.  Synchronizes `stream` between chunks of passes. */
int h8_oc_update_f64(const double* rho, const double* dc, int64_t m, double volume_fraction, double move,
                     double rho_min_bound, double rho_max_bound, double bisection_tol,
                     int32_t max_bisection_iter, double* rho_new, double* change, void* ws,
                     size_t ws_bytes, void* stream);

/* ---- Hex20: the 20-node serendipity hexahedron ------------------------- */
/* ke[e] = sum over the 3x3x3 Gauss rule of B^T D B detJ w for the 20-node
 * element, general geometry.  X: (m,20,3) nodal coordinates in the element's
 * node order (corners 0-7 as Hex8, mid-edge nodes 8-11 on the bottom edges
 * (0,1) (1,2) (2,3) (3,0), 12-15 on the top edges (4,5) (5,6) (6,7) (7,4),
 * 16-19 on the vertical edges (0,4) .. (3,7)); D: (6,6) when d_stride == 0,
 * else (m,36) with d_stride = 36; ke: (m,60,60).  Replaces Element.stiffness
 * + Hex20Solid3D.shape_functions / B_matrix (fe/elements/base.py:494-514,
 * fe/elements/solids.py:331-406). */
int h8_ke20_batched_f64(const double* X, const double* D, int64_t d_stride, int64_t m,
                        double* ke, void* stream);

/* The Hex20 box of nx x ny x nz congruent bricks: nodes numbered corners
 * (i,j,k) -> (k NI + i) NJ + j, then x-edge midpoints Nc + (k nx + i) NJ + j,
 * y-edge midpoints Nc + Nx + (k NI + i) ny + j, z-edge midpoints
 * Nc + Nx + Ny + (k NI + i) NJ + j (NI = nx+1, NJ = ny+1, NK = nz+1,
 * Nc = NI NJ NK, Nx = nx NJ NK, Ny = NI ny NK); DOF 3 node + c; element
 * (i,j,k) -> (k nx + i) ny + j.  K0: the (60,60) brick stiffness
 * (h8_ke20_batched_f64), s_e: per-element scale.  y = K u over every DOF,
 * in 8 parity-colour passes (deterministic).  Replaces the CSR product A @ p
 * of preconditioned_cg on a Hex20 mesh (fe/iterative_solvers.py:439-441). */
int h20_spmv_f64(int32_t nx, int32_t ny, int32_t nz, const double* s_e, const double* K0,
                 const double* u, double* y, void* stream);

/* dinv for the Hex20 box as h8_diag_f64 (1/K_ii at free DOFs, 0 at fixed;
 * H8_ZERO_DIAG on a zero free diagonal).  Replaces jacobi_preconditioner
 * (fe/iterative_solvers.py:193-202) on a Hex20 mesh. */
int h20_diag_f64(int32_t nx, int32_t ny, int32_t nz, const double* s_e, const double* K0,
                 const uint8_t* free_mask, double* dinv, int32_t* d_flag, void* stream);

/* Workspace (bytes) of h20_pcg_f64; 0 for an invalid box. */
size_t h20_pcg_workspace_bytes(int32_t nx, int32_t ny, int32_t nz);

/* Jacobi-PCG on the Hex20 box, FP64, x0 = 0: the same iteration, stopping
 * test and errors as h8_pcg_f64 (preconditioned_cg, fe/iterative_solvers.py:
 * 434-469): b is zero at fixed DOFs, dinv from h20_diag_f64. */
int h20_pcg_f64(int32_t nx, int32_t ny, int32_t nz, const double* s_e, const double* K0,
                const double* dinv, const double* b, double* x, double tol, int64_t maxiter,
                void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream);

/* ---- measurement ------------------------------------------------------- */
/* Average device time (ms) of the two kernels of one PCG iteration, each
 * launched `reps` times alone between HIP events on `stream`: avg_ms[0] = the
 * fused direction + K*p kernel, avg_ms[1] = the update kernel.  precision 64
 * (workspace of h8_pcg_f64, s_e/K0 double) or 32 (workspace of
 * h8_refine_solve, s_e/K0 float); dinv64 as for the solves.  Measurement
 * helper with no counterpart in the reference; scribbles on the workspace. */
int h8_time_kernels(const h8_grid* g, int precision, const void* s_e, const void* K0,
                    const double* dinv64, void* ws, size_t ws_bytes, int reps, double* avg_ms,
                    void* stream);

/* ---- small vector helpers used by the host around the solves ----------- */
/* y = float(x) */
int h8_cast_f64_f32(const double* x, float* y, int64_t n, void* stream);
/* out = mask ? (a - b) : 0 with mask = (dinv != 0)  (b may be NULL -> a) */
int h8_masked_sub_f64(const double* a, const double* b, const double* dinv, double* out,
                      int64_t n, void* stream);
/* The same for ncols columns of length n stored one after another in a and out,
 * b (length n, may be NULL) shared: the Dirichlet-corrected right-hand sides
 * F[:, c] - K u_D of a batch (fe/batched_solve.py:105-164), one launch per model. */
int h8_masked_sub_cols_f64(const double* a, const double* b, const double* dinv, double* out,
                           int64_t n, int64_t ncols, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* H8PCG_H */
