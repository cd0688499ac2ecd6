/* eqhip.h — C-ABI of the MI355X-native Equelle runtime (libeqhip.so).
 *
 * The reference exposes this path only as the C++ class
 * equelleCUDA::EquelleRuntimeCUDA (backends/cuda/include/EquelleRuntimeCUDA.hpp:89-331);
 * the same surface is available to C++ callers through include/EquelleRuntimeCUDA.hpp.
 * This C-ABI is the binding a foreign caller (ctypes, cgo, JNI, N-API) uses.  Each
 * entry point cites the reference member it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success and nonzero on error; the message of the
 *     last error on the calling thread is eqhip_last_error().  Errors are where the
 *     reference throws std::runtime_error (size mismatches, bad input sets, ...).
 *   - A runtime (grid + HelperOps + solver settings) is an opaque pointer created from
 *     a "key=value key=value" parameter string (Opm::ParameterGroup keys).
 *   - Values are integer handles owned by the runtime: collections of scalars (with an
 *     optional CSR Jacobian), index sets (CollOfCell / CollOfFace) and boolean
 *     collections.  They live in device memory (HBM); release them with eqhip_release.
 *   - All work is ordered on one HIP stream; calls that return host data synchronise.
 *   - One controller thread per runtime; the library is not re-entrant per runtime.
 */
#ifndef EQHIP_H_INCLUDED
#define EQHIP_H_INCLUDED

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
const char* eqhip_last_error(void);

/* ------------------------------------------------------------------ runtime */
/* EquelleRuntimeCUDA::EquelleRuntimeCUDA(const Opm::ParameterGroup&) — EquelleRuntimeCUDA.cpp:65-82.
 * Keys: grid_dim nx ny nz dx dy dz; solver (BiCGStab|CG|GMRes) preconditioner (diagonal|ilu0|none)
 * solver_max_iter solver_tol; max_iter abs_res_tol verbose output_to_file; extensions:
 * rel_res_tol precision (fp64|mixed) gmres_restart inner_tol max_refine. */
int eqhip_runtime_create(const char* params, void** rt);
/* ~EquelleRuntimeCUDA() — EquelleRuntimeCUDA.cpp:85-88 */
int eqhip_runtime_destroy(void* rt);
/* info[5] = {dim, cells, faces, interior faces, boundary faces}; UnstructuredGrid fields
 * read by EquelleRuntimeCUDA::getGrid() — EquelleRuntimeCUDA.hpp:297 */
int eqhip_grid_info(void* rt, int* info);
/* which: 0 face_cells (2*faces), 1 cell_facepos (cells+1), 2 cell_faces — getGrid() */
int eqhip_grid_int_array(void* rt, int which, int* out);
/* which: 0 cell_centroids, 1 cell_volumes, 2 face_centroids, 3 face_areas, 4 face_normals — getGrid() */
int eqhip_grid_double_array(void* rt, int which, double* out);
/* HelperOps matrices, which: 0 grad, 1 div, 2 fulldiv — getGradMatrix/getDivMatrix/getFulldivMatrix,
 * EquelleRuntimeCUDA.hpp:299-303.  info[3] = {rows, cols, nnz}. */
int eqhip_opmat_info(void* rt, int which, int* info);
int eqhip_opmat_copy(void* rt, int which, int* row_ptr, int* col_ind, double* vals);

/* ------------------------------------------------------------------ index sets */
/* which: 0 allCells 1 boundaryCells 2 interiorCells 3 allFaces 4 boundaryFaces 5 interiorFaces
 * — EquelleRuntimeCUDA.hpp:100-110 */
int eqhip_set_builtin(void* rt, int which, int* set);
/* CollOfIndices<codim>(indices) — CollOfIndices.hpp:57 */
int eqhip_set_from_host(void* rt, int codim, const int* idx, int n, int* set);
int eqhip_set_size(void* rt, int set, int* n);                 /* CollOfIndices::size() */
int eqhip_set_copy(void* rt, int set, int* out);               /* CollOfIndices::stdToHost() */
int eqhip_first_cell(void* rt, int faces, int* set);           /* firstCell — EquelleRuntimeCUDA.hpp:116 */
int eqhip_second_cell(void* rt, int faces, int* set);          /* secondCell — EquelleRuntimeCUDA.hpp:122 */
int eqhip_set_on(void* rt, int data, int from, int to, int* set); /* operatorOn<cd,cs> — .hpp:207-210 */
int eqhip_set_trinary(void* rt, int pred, int t, int f, int* set); /* trinaryIf<codim> — .hpp:223-226 */
int eqhip_is_empty(void* rt, int set, int* pred);              /* isEmpty<codim> — .hpp:176-177 */
/* superset.contains(subset) check of inputDomainSubsetOf — EquelleRuntimeCUDA_impl.hpp:98 */
int eqhip_set_contains(void* rt, int superset, int subset);

/* ------------------------------------------------------------------ booleans (CollOfBool) */
int eqhip_bool_size(void* rt, int pred, int* n);
int eqhip_bool_copy(void* rt, int pred, unsigned char* out);
/* op: 0 >, 1 >=, 2 <, 3 <=, 4 ==, 5 != — CollOfScalar comparisons, CollOfScalar.hpp:282-407 */
int eqhip_compare(void* rt, int op, int a, int b, int* pred);
/* scalar_left: 0 => a OP s, 1 => s OP a */
int eqhip_compare_scalar(void* rt, int op, int a, double s, int scalar_left, int* pred);

/* ------------------------------------------------------------------ collections of scalars */
int eqhip_scalar_from_host(void* rt, const double* v, int n, int* h);  /* CollOfScalar(std::vector) */
int eqhip_scalar_uniform(void* rt, int n, double s, int* h);          /* CollOfScalar(size, value) */
/* CollOfScalar(CudaArray, CudaMatrix): values + CSR Jacobian (rows = n) */
int eqhip_scalar_with_jacobian(void* rt, const double* v, int n, const int* row_ptr, const int* col_ind,
                               const double* jac, int cols, int* h);
/* CollOfScalar(const CollOfScalar&, bool primaryVariable) — collOfScalar.cu:68-78 */
int eqhip_primary(void* rt, int a, int* h);
/* info[5] = {size, useAutoDiff, jac rows, jac cols, jac nnz}; empty Jacobian => zeros */
int eqhip_scalar_info(void* rt, int h, int* info);
int eqhip_scalar_copy(void* rt, int h, double* out);                              /* copyToHost() */
int eqhip_jacobian_copy(void* rt, int h, int* row_ptr, int* col_ind, double* vals); /* matrixToHost() */
/* kind: 0 scalar collection, 1 index set, 2 boolean collection */
int eqhip_release(void* rt, int kind, int h);
/* op: 0 +, 1 -, 2 *, 3 / — CollOfScalar operators, collOfScalar.cu:273-341 */
int eqhip_binop(void* rt, int op, int a, int b, int* h);
/* op 2: s*a; op 3: a/s (scalar_left 0) or s/a (scalar_left 1) — collOfScalar.cu:367-399 */
int eqhip_scalar_binop(void* rt, int op, double s, int a, int scalar_left, int* h);
int eqhip_negate(void* rt, int a, int* h);                      /* unary minus, collOfScalar.cu:401-403 */
int eqhip_sqrt(void* rt, int a, int* h);                        /* sqrt — EquelleRuntimeCUDA.cpp:324-327 */
int eqhip_trinary(void* rt, int pred, int t, int f, int* h);    /* trinaryIf — EquelleRuntimeCUDA.cpp:239-248 */
int eqhip_gradient(void* rt, int a, int* h);                    /* gradient — EquelleRuntimeCUDA.cpp:251-264 */
int eqhip_divergence(void* rt, int a, int* h);                  /* divergence — EquelleRuntimeCUDA.cpp:273-291 */
/* negGradient: ngrad * a on AllCells() (u[first] - u[second] per interior face) and
 * interiorDivergence: div * a on InteriorFaces() — EquelleRuntimeCPU.cpp:345-365 (the CUDA
 * backend throws on both, EquelleRuntimeCUDA.cpp:311-319) */
int eqhip_neg_gradient(void* rt, int a, int* h);
int eqhip_interior_divergence(void* rt, int a, int* h);
/* norm(CollOfScalar): |a| with Jacobian diag(sign) a' — EquelleRuntimeCPU.cpp:241-254 */
int eqhip_norm_scalar(void* rt, int a, int* h);
int eqhip_extend(void* rt, int a, int from, int to, int* h);    /* operatorExtend — _impl.hpp:112-125 */
int eqhip_extend_scalar(void* rt, double s, int to, int* h);    /* operatorExtend(Scalar, set) — _impl.hpp:127-131 */
int eqhip_on(void* rt, int a, int from, int to, int* h);        /* operatorOn — _impl.hpp:134-147 */
/* kind: 0 sum, 1 product, 2 max, 3 min, 4 twoNorm — EquelleRuntimeCUDA.cpp:148-151,332-346 */
int eqhip_reduce(void* rt, int kind, int a, double* out);
/* The same reduction, allocating no device memory (no hipMalloc/hipFree, no pool block).
 * No counterpart in the reference: it is for callers that cannot allocate (code under
 * graph capture, tight serving loops, harnesses that forbid allocation).  Its partials live
 * in a fixed 16 KiB device array per GPU (2048 doubles; at most 1025 are used, so every
 * collection size fits); calls that overlap on one GPU share that array and race.  Results
 * are bit-identical to eqhip_reduce.  kind outside 0..4 is an error. */
int eqhip_reduce_noalloc(void* rt, int kind, int a, double* out);
int eqhip_norm_set(void* rt, int set, int* h);                  /* norm<codim>(set) — _impl.hpp:171-182 */
int eqhip_centroid(void* rt, int set, double* out);             /* centroid — _impl.hpp:203-210 (n*dim) */
int eqhip_normal(void* rt, int faces, double* out);             /* normal — EquelleRuntimeCUDA.cpp:137-140 (unit) */
/* The linear solve of one Newton step: J du = r with r = residual value, J its Jacobian
 * — LinearSolver::solve, linearSolver.cu:107-201.  *iters = Krylov iterations. */
int eqhip_solve_for_update(void* rt, int residual, int* du, int* iters);

/* ------------------------------------------------------------------ CSR algebra (CudaMatrix) */
/* op: 0 A+B, 1 A-B, 2 A*B on host CSR inputs, computed on the device — cudaMatrix.cu:640-694.
 * An absent operand (null row_ptr) is the empty (zero) matrix.  out_info = {rows, cols, nnz}. */
int eqhip_csr_binop(int op, int ar, int ac, const int* arp, const int* aci, const double* av, int br, int bc,
                    const int* brp, const int* bci, const double* bv, int* out_info, int* orp, int* oci,
                    double* ov, int cap);
/* y = op(A) x — operator*(CudaMatrix, CudaArray), cudaMatrix.cu:698-733 */
int eqhip_csr_spmv(int rows, int cols, const int* rp, const int* ci, const double* v, const double* x, double* y,
                   int transpose);
/* y = J x through the Krylov solver's SELL-64 kernels (fp32 != 0: FP32 kernel) for the
 * Jacobian of collection h — the SpMV inside LinearSolver::solve. */
int eqhip_sell_spmv(void* rt, int h, const double* x, double* y, int fp32);

/* ------------------------------------------------------------------ workloads (heat) */
/* Runs the heat program of examples/app/out_heateq.cpp:28-73 (k_model=const) or its
 * nonlinear-conductivity form (k_model=quadratic) through the device AD engine:
 * one newtonSolve per time step.  newton_iters / final_norms: per step (may be null). */
int eqhip_heat_run(const char* params, const double* u_initial, int n, const int* dir_faces,
                   const double* dir_vals, int ndir, const double* dts, int nsteps, double* u_out,
                   int* newton_iters, double* final_norms);
/* One residual + Jacobian evaluation of the same program; info = {n, nnz}; null outputs => query sizes. */
int eqhip_heat_residual(const char* params, const double* u, const double* u0, int n, const int* dir_faces,
                        const double* dir_vals, int ndir, double dt, int* info, double* F, int* row_ptr,
                        int* col_ind, double* jac);

/* A heat simulation session: the program above on one runtime with the state kept in
 * device memory between steps (the time loop of out_heateq.cpp:64-73).
 * stats = {Newton iterations, Krylov iterations, stop reason (1 abs, 2 rel, 0 max_iter)},
 * norms = {initial, final residual 2-norm}; either may be null. */
int eqhip_heat_create(const char* params, const double* u_initial, int n, const int* dir_faces,
                      const double* dir_vals, int ndir, void** sim);
int eqhip_heat_step(void* sim, double dt, int* stats, double* norms);
int eqhip_heat_state(void* sim, double* u);
/* replaces the session's state (the field the next step starts from) */
int eqhip_heat_set_state(void* sim, const double* u);
/* Time of one residual + Jacobian evaluation of the program at the session's state (the
 * work of one Newton iteration before its linear solve, EquelleRuntimeCUDA_impl.hpp:216-300
 * evaluated by the device AD engine), averaged over reps evaluations after one untimed one;
 * info = {n, Jacobian nnz, 1 when the recorded residual program ran as one fused kernel
 * (fuse_residual=1, the default) or 0 when it ran op by op}.  Measurement only: the state
 * is not changed. */
int eqhip_heat_eval_time(void* sim, double dt, int reps, double* ms_per_eval, int* info);
int eqhip_heat_destroy(void* sim);

/* ------------------------------------------------------------------ recorded residuals */
/* No counterpart in the reference (which evaluates the residual functor op by op,
 * EquelleRuntimeCUDA_impl.hpp:216-300).  Counters of the recorded residual programs of
 * this process: out = {recording evaluations, symbolic recordings that found their kernel,
 * kernels built, fused evaluations, recordings stopped by an unsupported operation,
 * kernels rejected by the check against the recorded evaluation}; seconds = total kernel
 * build time. */
int eqhip_residual_stats(long* out, double* seconds);
/* Builds the fused kernel of a small heat-like residual program with the run-time compiler
 * for gfx950 (no device needed); 0 on success, the compiler log (truncated) in log. */
int eqhip_residual_codegen_selftest(char* log, int len);

/* ------------------------------------------------------------------ two-phase sessions */
/* The two-phase program of examples/dsl/twophase.equelle (examples/app/out_twophase.cpp)
 * on the device AD engine: per step a pressure newtonSolve then a saturation newtonSolve
 * (twophase.equelle:99-113).  params: grid (grid_dim=3 nx ny nz dx dy dz), watervisc,
 * oilvisc, min_poro, corey (1|2), and the solver/Newton parameters of runtime_create.
 * fully_implicit=1 runs examples/dsl/twophase_fully_implicit.equelle instead: one
 * newtonSolveSystem on (p, sw) per step (EquelleRuntimeCPU_impl.hpp:282-339).
 * perm, poro, sw_initial: one value per cell; sources: ascending cells with rates. */
int eqhip_twophase_create(const char* params, const double* perm, const double* poro, const double* sw_initial,
                          int n, const int* src_cells, const double* src_vals, int nsrc, void** sim);
/* stats = {pressure Newton, Krylov, stop; transport Newton, Krylov, stop};
 * norms = {pressure ||F0||, ||Fn||, transport ||F0||, ||Fn||}; fully implicit: the system solve's
 * Newton, Krylov, stop in stats[0..2], stats[3..4] = 0, stats[5] = stop, norms repeated */
int eqhip_twophase_step(void* sim, double dt, int* stats, double* norms);
int eqhip_twophase_state(void* sim, double* p, double* sw);
/* replaces the session's pressure and/or saturation (null leaves that field) */
int eqhip_twophase_set_state(void* sim, const double* p, const double* sw);
/* residual + Jacobian (CSR): which = 0 pressure residual at p (total mobility of sw0),
 * which = 1 transport residual at sw (flux from p, sw0; step dt), which = 2 the fully
 * implicit system at (p, sw) (2n rows [pressure; transport], columns [p; sw], previous
 * saturation sw0); info = {rows, nnz}; pass null F to query sizes. */
int eqhip_twophase_residual(void* sim, int which, const double* p, const double* sw0, const double* sw, double dt,
                            int* info, double* F, int* row_ptr, int* col_ind, double* jac);
int eqhip_twophase_destroy(void* sim);

/* ------------------------------------------------------------------ fused heat batches */
/* Many independent nonlinear-heat instances (the program above, k_model=quadratic,
 * zero-flux boundary, unit or given cells) advanced together by fused kernels with
 * per-instance Newton and Krylov convergence.  No reference counterpart: the reference
 * runs one newtonSolve at a time (EquelleRuntimeCUDA_impl.hpp:216-300).
 * params: precision (fp64|mixed) preconditioner (diagonal|ilu0|none) rel_res_tol abs_res_tol
 * max_iter solver_tol solver_max_iter inner_tol (heat, default 1e-6) inner_tol_twophase (default 1e-4)
 * max_refine graph (0|1) cpt xcd tune krylov_form_heat (reduced|full, default reduced)
 * krylov_form_twophase (reduced|full|auto, default auto: full when the pressure stage runs
 * multigrid CG, else reduced in fp64 and full in mixed): with preconditioner ilu0, the Krylov
 * solve on the red-black reduced (Schur complement) system or on the full system with
 * red-black ILU(0).  twophase_pressure_cg (0|1, default 1): CG for the symmetric pressure
 * stage of two-phase instances; twophase_pressure_mg (0|1, default 1): with the aggregation
 * multigrid V-cycle as its preconditioner (full form).  pair_vectors (0|1, default 1): four
 * pair ranks per lane in the Krylov kernels of 2D heat instances with nx % 8 == 0;
 * operator_tiles (0|1, default 1): of those, instances with nx % 512 == 0 apply the reduced
 * operator in tiles of four rows x 256 pair ranks per workgroup. */
int eqhip_batch_create(const char* params, void** batch);
int eqhip_batch_destroy(void* batch);
int eqhip_batch_add(void* batch, int nx, int ny, double k, double cv, const double* u, int* slot);
/* Adds a two-phase instance (examples/dsl/twophase.equelle; generated form
 * examples/app/out_twophase.cpp:26-113) to the batch: Cartesian nx x ny x nz cells
 * (nz = 1: 2D), per-cell perm, poro, initial sw, rate sources on ascending cells.  params:
 * dx dy dz watervisc oilvisc min_poro corey (1|2).  A step is a pressure Newton solve then
 * a saturation Newton solve (twophase.equelle:99-113); pressure starts at 0.  Two-phase
 * and heat instances share the batch's launches. */
int eqhip_batch_add_twophase(void* batch, const char* params, int nx, int ny, int nz, const double* perm,
                             const double* poro, const double* sw, int nsrc, const int* src_cells,
                             const double* src_vals, int* slot);
int eqhip_batch_twophase_state(void* batch, int slot, double* p, double* sw);
/* stats = {pressure Newton iterations, pressure Krylov iterations, pressure stop reason};
 * eqhip_batch_stats gives the totals of the step and the saturation solve's stop reason */
int eqhip_batch_twophase_stats(void* batch, int slot, int* stats);
/* pressure (stage 0) or transport (stage 1) residual and Jacobian (CSR, ascending columns)
 * at (p, sw0, sw, dt); overwrites the slot's state; caller provides n, n+1, 7n, 7n entries */
int eqhip_batch_twophase_residual(void* batch, int slot, int stage, const double* p, const double* sw0,
                                  const double* sw, double dt, int* info, double* F, int* row_ptr, int* col_ind,
                                  double* jac);
int eqhip_batch_remove(void* batch, int slot);
/* starts one implicit step of length dt for the slot (u0 := u) */
int eqhip_batch_begin_step(void* batch, int slot, double dt);
/* one Newton iteration for every slot with a step in flight (mixed precision: or one more
 * refinement round of its linear solve); returns the slots that finished */
int eqhip_batch_iterate(void* batch, int* finished, int cap, int* nfinished);
int eqhip_batch_in_flight(void* batch, int slot, int* flag);
/* whole steps for the listed slots */
int eqhip_batch_step(void* batch, const int* slots, int nslots, double dt);
/* stats = {Newton, Krylov, refinements, stop reason, failed}, norms = {||F(u_0)||, ||F(u_k)||} */
int eqhip_batch_stats(void* batch, int slot, int* stats, double* norms);
/* the same stats for nslots slots in one call: stats[5 q .. 5 q + 4] of slots[q] (host-side
 * counters, no device access) */
int eqhip_batch_stats_many(void* batch, const int* slots, int nslots, int* stats);
int eqhip_batch_state(void* batch, int slot, double* u);
int eqhip_batch_set_state(void* batch, int slot, const double* u);
/* two-phase slot: pressure and/or saturation (null leaves that field) */
int eqhip_batch_set_twophase_state(void* batch, int slot, const double* p, const double* sw);
/* device-memory footprint of the slot in bytes (admission control) */
int eqhip_batch_bytes(void* batch, int slot, double* bytes);
/* launch shape of the most recent launch: info = {cells per thread of the SpMV, XCD-contiguous
 * block order, membership signatures tuned so far}.  Shapes are chosen per membership by
 * timing each kernel class (params cpt=1|2|4|8 and xcd=0|1 fix them for every class;
 * tune=0 uses cpt=4, xcd=1). */
int eqhip_batch_shape(void* batch, int* info);
/* cells per thread of each kernel class of the most recent launch: cpt = {assembly, defect,
 * SpMV, preconditioner, vector updates} (5 entries) */
int eqhip_batch_shape_classes(void* batch, int* cpt);
/* average execution time in ms of one launch over all slots (every slot active, tuned
 * shape): each of `reps` launches is bracketed by its own HIP start/stop events
 * (hipExtLaunchKernel) and waited for; event_ms (optional) gets the average of `reps`
 * further launches queued back to back between one event pair (dispatch gaps included).
 * kernel = "stencil_spmv" | "assemble" | "ilu_black" | "ilu_red" | "defect" */
int eqhip_batch_kernel_time(void* batch, const char* kernel, int reps, double* avg_ms, double* event_ms);
/* F and the Jacobian (CSR, ascending columns) at (u, u0, dt); info = {n, nnz};
 * caller provides n, n+1, 5n, 5n entries. */
int eqhip_batch_residual(void* batch, int slot, const double* u, const double* u0, double dt, int* info,
                         double* F, int* row_ptr, int* col_ind, double* jac);

/* ------------------------------------------------------------------ request loop */
/* Continuous batching of incoming solves.  A request = one nonlinear-heat or two-phase
 * instance advanced by `steps` implicit steps of length dt.  Requests queue FIFO until their HBM
 * footprint fits `hbm_budget` (bytes); admitted requests share launches, one Newton
 * iteration per progress() tick, and keep their state device-resident between steps.
 * No reference counterpart (the reference runs one program per process,
 * compiler/PrintCUDABackendASTVisitor.cpp:18-77); this is the request loop of the
 * serving workload.  params: the batch parameters above plus hbm_budget, max_in_flight,
 * max_in_flight_twophase (two-phase requests in flight at once; default unlimited) and
 * twophase_precision (fp64|mixed, default fp64: Krylov precision of the two-phase requests).
 * Admission is FIFO per request kind (heat, two-phase), the older queue head first. */
int eqhip_server_create(const char* params, void** server);
int eqhip_server_destroy(void* server);
int eqhip_server_submit(void* server, int nx, int ny, double k, double cv, const double* u0, int steps, double dt,
                        int* request);
/* a two-phase request (the instance of eqhip_batch_add_twophase, same params) */
int eqhip_server_submit_twophase(void* server, const char* params, int nx, int ny, int nz, const double* perm,
                                 const double* poro, const double* sw, int nsrc, const int* src_cells,
                                 const double* src_vals, int steps, double dt, int* request);
/* one scheduling tick; returns the requests that finished in it */
int eqhip_server_progress(void* server, int* finished, int cap, int* nfinished);
/* copies a finished request's field (heat: u; two-phase: sw into u and, if p is not null,
 * the pressure; one value per cell) and releases it;
 * stats = {steps, Newton, Krylov, failed}, times = {queued seconds, latency seconds} */
int eqhip_server_result(void* server, int request, double* u, double* p, int* stats, double* times);
/* Newton iterations per solve allowed for one request, queued or in flight (at most the
 * server's max_iter; 0 restores it).  A step that reaches the limit unconverged ends the
 * request with failed = 1; no other request is affected. */
int eqhip_server_set_max_newton(void* server, int request, int max_newton);
/* counts = {queued, in flight}, bytes = {admitted footprint, budget} */
int eqhip_server_status(void* server, int* counts, double* bytes);

/* ------------------------------------------------------------------ measurement */
/* Live kernel timing with HIP events on the runtime stream: mode 1 on, 0 off, 2 reset.
 * Timed launches: "spmv_fp64", "spmv_fp32" (CSR/SELL Krylov SpMV of the AD engine),
 * "stencil_spmv_fp64", "stencil_spmv_fp32", "assemble", "defect_fp64" (fused batches). No reference counterpart
 * (the reference only prints wall times, EquelleRuntimeCUDA_impl.hpp:295-297). */
int eqhip_timing(int mode);
int eqhip_timing_query(const char* name, int* count, double* total_ms);
/* Waits for all work queued by the runtime. */
int eqhip_device_sync(void);

#ifdef __cplusplus
}
#endif
#endif /* EQHIP_H_INCLUDED */
