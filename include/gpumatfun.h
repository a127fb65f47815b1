/*
 * gpumatfun.h -- C ABI of the MI355X (gfx950) f(A)b engine.
 *
 * The entry points mirror the reference's nine MATLAB MEX gateways one for
 * one; each declaration cites the gateway (and the solver method behind it)
 * that it replaces.  A MEX shim, a ctypes binding or any FFI binds these
 * symbols directly: plain pointers and sizes, no C++ or torch types.
 *
 * Conventions (all entry points):
 *   - every function returns gm_status and never aborts; the text of the
 *     last error of the calling thread is available from gm_last_error();
 *   - dense arrays are column-major double precision, as MATLAB passes them;
 *   - all host arrays given to gm_lanczos_create are deep-copied;
 *   - a solver owns one HIP stream; the stepping entry points only enqueue
 *     work on it, the getters synchronise it; one thread at a time per solver;
 *   - handles are checked against the set of live solvers, so a deleted or
 *     unknown handle gives GM_EBADHANDLE instead of undefined behaviour.
 */
#ifndef GPUMATFUN_H
#define GPUMATFUN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gm_status {
    GM_OK = 0,
    GM_EINVAL = 1,       /* invalid argument (shape, type, value)            */
    GM_ENOMEM = 2,       /* device or host allocation failed                 */
    GM_EDEVICE = 3,      /* HIP runtime error or no usable gfx950 device     */
    GM_EBREAKDOWN = 4,   /* Lanczos breakdown (beta = 0 or non-finite)       */
    GM_EBADHANDLE = 5,   /* null, deleted or unknown solver handle           */
    GM_ESTATE = 6        /* call out of order (e.g. apply_V before a build)  */
} gm_status;

typedef struct gm_solver gm_solver;

/* Sparse input format of the matrix A. */
enum {
    GM_SPARSE_CSC = 0,   /* MATLAB layout: ptr = jc (n+1), ind = ir (nnz)    */
    GM_SPARSE_CSR = 1
};

/* Flags of gm_create_desc.flags. */
enum {
    GM_FLAG_NO_CLAMP = 1u << 0,  /* update_U without the u >= 0 clamp of
                                    update_U.cu:9 (the CPU path's semantics) */
    GM_FLAG_AUTOTUNE = 1u << 1,  /* time the SpMM row-to-lane schedules on A at
                                    create and keep the fastest              */
    GM_FLAG_ROW_GROUPS = 1u << 2, /* without AUTOTUNE: the row-group SpMM
                                    (lanes per row from the mean row length) */
    GM_FLAG_ROW_BLOCKS = 1u << 3, /* without AUTOTUNE: the row-block SpMM with
                                    1024-nonzero LDS tiles.  Neither flag: row
                                    blocks when rows average <= 16 nonzeros */
    GM_FLAG_INDEX32 = 1u << 4,    /* never store the row-block SpMM's column
                                    indices as 16-bit offsets from the
                                    block's first row (the autotuner times
                                    both widths where the offsets fit)      */
    GM_FLAG_INDEX16 = 1u << 5,    /* without AUTOTUNE: use the 16-bit offsets
                                    where they fit (default: 32-bit)        */
    GM_FLAG_EARLY_LOADS = 1u << 6, /* without AUTOTUNE: the row-block SpMM loads
                                    each row's own operands before its
                                    gathers (default: after)                */
    GM_FLAG_WAVE_SOLVES = 1u << 7, /* without AUTOTUNE: the shifted solves run
                                    one wavefront per system (cyclic
                                    reduction; m <= 128) instead of one lane
                                    per system.  AUTOTUNE times the forms   */
    GM_FLAG_COLUMN_SLABS = 1u << 8, /* gm_batch: chunks of >= 32 columns run
                                    their SpMM in slabs of 32 columns (at 256
                                    columns one slab per XCD) instead of
                                    whole rows.  AUTOTUNE times both forms  */
    GM_FLAG_LOCALITY_ORDER = 1u << 9 /* gm_batch: renumber the nodes inside the
                                    engine by gm_locality_order (leaves of
                                    512 nodes) so that rows gathered together
                                    lie close; inputs and results stay in the
                                    caller's order                          */
};

/*
 * The 26 positional inputs of lanczos_create (lanczos_create.cpp:7-32), in
 * the same order, plus the sizes MATLAB carries inside its mxArrays.
 */
typedef struct gm_create_desc {
    int32_t n;                     /* rows of A (= rows of U)               */
    const double* U;               /* [0]  n x 2: [u, g(u)]                 */
    int32_t sparse_format;         /* [1]  A: GM_SPARSE_CSC or _CSR         */
    int64_t nnz;
    const double* A_values;        /*      nnz                              */
    const int32_t* A_ptr;          /*      n + 1                            */
    const int32_t* A_ind;          /*      nnz                              */
    int32_t num_eigs;              /*      K (columns of Q), may be 0       */
    const double* Q;               /* [2]  n x K                            */
    const double* lambda;          /* [3]  K (unused, as in the reference)  */
    int32_t subspace_size;         /* [4]  m, the "N" argument              */
    int32_t gamma_size;            /*      d + 1                            */
    const double* gamma;           /* [5]  d + 1, highest power first       */
    int32_t contour_size;          /*      n_s                              */
    const double* gamma_s_real;    /* [6]  (d+1) x n_s                      */
    const double* gamma_s_imag;    /* [7]  (d+1) x n_s                      */
    const double* exp_lambda;      /* [8]  K                                */
    const double* phi_lambda;      /* [9]  K                                */
    double dt;                     /* [10]                                  */
    const double* ode;             /* [11] n (may be NULL: zeros)           */
    double aode, bode, code;       /* [12..14]                              */
    double vrest, vamp, vth;       /* [15..17] (vth unused)                 */
    double vpeak, c1ion, c2ion;    /* [18..20] (c2ion unused)               */
    const double* mass;            /* [21] n (may be NULL: ones)            */
    const double* wts1;            /* [22] n_s x 2 [Re Im], includes s      */
    const double* wts2;            /* [23] n_s x 2                          */
    const double* shifts;          /* [24] 4 * m * n_s, packed as in
                                            problem_startup.m:140-144       */
    const double* valvec;          /* [25] must be [-1, 0, 1] (may be NULL) */
    uint32_t flags;                /* GM_FLAG_*                             */
    int32_t device;                /* HIP device ordinal                    */
} gm_create_desc;

/* Text of the last error raised on the calling thread ("" if none). */
const char* gm_last_error(void);

/* replaces lanczos_create.cpp:35-768 (mexFunction) and the constructor
 * lanczos_impl.cu:5-91: upload, U(:,1) *= dt, deflation Z = Q'U,
 * U -= QZ, update = Q diag(e^L, phi(L)) Z.                                 */
gm_status gm_lanczos_create(const gm_create_desc* desc, gm_solver** out);

/* replaces lanczos_delete.cpp:6-35.                                        */
gm_status gm_lanczos_delete(gm_solver* s);

/* replaces build_subspace.cpp:6-35 -> build_subspace.cu:5-91.            */
gm_status gm_build_subspace(gm_solver* s);

/* replaces lin_sys_solves.cpp:6-35 -> lin_sys_solves.cu:5-21
 * (copyforsolve/linsyssolve/copyafter, lanczos_impl.cu:292-406).          */
gm_status gm_lin_sys_solves(gm_solver* s);

/* replaces apply_V.cpp:6-35 -> apply_V.cu:5-95.                           */
gm_status gm_apply_V(gm_solver* s);

/* replaces update_U.cpp:6-35 -> update_U.cu:5-32.                         */
gm_status gm_update_U(gm_solver* s);

/* replaces update_U_monodomain.cpp:6-35 -> update_U_monodomain.cu:5-43.   */
gm_status gm_update_U_monodomain(gm_solver* s);

/* Mitchell-Schaeffer reaction (no reference counterpart: the reference's
 * monodomain reaction is FitzHugh-Nagumo, update_U_monodomain.cu:11-22).
 * The gate h is kept in the ode vector; one update does
 *   u  = X(:,1) + X(:,2) + update(:,1) + update(:,2)
 *   h += dt*((1-h)/tau_open)  if u < v_gate,  h += dt*(-h/tau_close) otherwise
 *   g  = h u^2 (1-u)/tau_in - u/tau_out,  U(:,2) = dt*g
 * then the deflation split as in update_U.  Defaults: 0.3, 6, 120, 150, 0.13. */
typedef struct gm_ms_params {
    double tau_in, tau_out, tau_open, tau_close, v_gate;
} gm_ms_params;
gm_status gm_set_ms_params(gm_solver* s, const gm_ms_params* p);
gm_status gm_update_U_ms(gm_solver* s);

/* replaces get_U.cpp:6-38 -> lanczos_impl.cu:502-506: dst is n x 2
 * column-major host memory, receives Usol = [u, dt*g(u)].                  */
gm_status gm_get_U(gm_solver* s, double* dst);

/* replaces get_xsol.cpp:6-38 -> lanczos_impl.cu:508-512: dst is n x 2,
 * receives X = [f1(A) ubar, f2(A) dt*gbar] (without the deflation update). */
gm_status gm_get_xsol(gm_solver* s, double* dst);

/* replaces get_aabb_mex, called at lanczosEngine.m:25-27 but absent from
 * the reference tree: dst is (m+1) x 4 [alpha1 alpha2 beta1 beta2].       */
gm_status gm_get_aabb(gm_solver* s, double* dst);

/* replaces set_y_mex, called at lanczosEngine.m:29-31 but absent from the
 * reference tree: overwrites Yreal/Yimag (each m x 2n_s).                  */
gm_status gm_set_y(gm_solver* s, const double* yreal, const double* yimag);

/* Reads Yreal/Yimag (m x 2n_s each), the output of lin_sys_solves.        */
gm_status gm_get_y(gm_solver* s, double* yreal, double* yimag);

/* State get/set for checkpoint/resume (no reference counterpart: the
 * reference's get_U returns [u, dt*g] and create multiplies by dt again,
 * lanczos_impl.cu:79, so it cannot round-trip).  u, g (without dt) and ode
 * are n-vectors; any pointer may be NULL to skip it.  set_state recomputes
 * the deflation split exactly as update_U does.                           */
gm_status gm_get_state(gm_solver* s, double* u, double* g, double* ode);
gm_status gm_set_state(gm_solver* s, const double* u, const double* g, const double* ode);

/* One or more full time steps: build_subspace, lin_sys_solves, apply_V,
 * then update_U (monodomain = 0), update_U_monodomain (1) or update_U_ms (2),
 * the sequence of
 * nicksmethodgpu_problem.m:7-16 / nicksmethodgpu_problem_monodomain.m:7-16.
 * use_graph != 0 captures one step into a HIP graph on first use and
 * replays it.  Asynchronous: returns once the work is enqueued.            */
gm_status gm_step(gm_solver* s, int32_t nsteps, int32_t monodomain, int32_t use_graph);

/* Waits for the solver's stream; returns GM_EBREAKDOWN if any Lanczos run
 * since create met beta = 0 or a non-finite beta (the affected basis
 * vectors are set to zero instead of propagating NaN).                     */
gm_status gm_synchronize(gm_solver* s);

/* Introspection: sizes and the chosen SpMM schedule: spmm_lanes > 0 is the
 * row-group kernel with that many lanes per row, spmm_lanes < 0 the
 * row-block kernel with -spmm_lanes nonzeros per LDS tile.                 */
gm_status gm_get_info(gm_solver* s, int64_t* n, int64_t* nnz, int32_t* m,
                      int32_t* gamma_size, int32_t* contour_size,
                      int32_t* num_eigs, int32_t* spmm_lanes);
/* The SpMM schedule in use: kind 0 = row groups (param = lanes per row),
 * 1 = row blocks (param = nonzeros per LDS tile), 2 = row blocks with the
 * epilogue operands loaded before the gathers; index_bytes = bytes per
 * stored column index (4, or 2 for 16-bit block offsets).                  */
gm_status gm_get_schedule(gm_solver* s, int32_t* kind, int32_t* param, int32_t* index_bytes);
/* The shifted solves' form (lin_sys_solves.cu:5-21): 0 = one lane per
 * system, factors in global memory; 1 = one lane per system, factors in LDS;
 * 2 = one wavefront per system (cyclic reduction).  systems_per_block may be
 * NULL.                                                                    */
gm_status gm_get_solve_form(gm_solver* s, int32_t* form, int32_t* systems_per_block);

/* The solver's HIP stream (as void*), for callers that time it with
 * events of their own.                                                      */
gm_status gm_get_stream(gm_solver* s, void** stream);

/* Measurement helpers (no reference counterpart).  gm_profile_spmm launches
 * one SpMM variant of the step `reps` times on the solver's stream between
 * two HIP events and returns the average duration of one launch in ms:
 * mode 0 = the Horner inner step of build_subspace (y = A x + g v_k, the
 * most frequent kernel of a step with a polynomial preconditioner),
 * mode 1 = the Lanczos step without preconditioner (y = A v - b v_{k-1}
 * fused with the alpha dot product).  It overwrites only scratch buffers;
 * call it between steps.  gm_profile_copy times a plain device-to-device
 * copy kernel over `bytes` bytes (read once, written once), the reference
 * point for the HBM counters.                                               */
gm_status gm_profile_spmm(gm_solver* s, int32_t mode, int32_t reps, double* avg_ms);
/* As gm_profile_spmm, but every timed launch follows a streaming copy of
 * flush_bytes (half read, half written) and is timed alone between its own two
 * events, so L2 and the Infinity Cache hold nothing of the SpMM's working set
 * when it starts (flush_bytes well above 256 MB); the average includes one
 * launch latency per call.                                                  */
gm_status gm_profile_spmm_flushed(gm_solver* s, int32_t mode, int32_t reps, int64_t flush_bytes,
                                  double* avg_ms);
gm_status gm_profile_copy(gm_solver* s, int64_t bytes, int32_t reps, double* avg_ms);
/* What a persistent kernel would pay instead of a kernel boundary: the average
 * cost (us) of one cooperative grid barrier over `blocks` resident workgroups
 * (each thread stores one value before it), and of launching the same store as
 * back-to-back kernels of `blocks` workgroups.  No solver needed.           */
gm_status gm_profile_barrier(int32_t device, int32_t blocks, int32_t iters, double* grid_sync_us,
                             double* launch_us);

/* ---------------------------------------------------------------------------
 * Batched form (no reference counterpart): P independent problems that share
 * A, the contour nodes/weights and the polynomial preconditioner, stepped
 * together.  Problem p owns the column pair [u_p, g_p] (the plain path's exp
 * and phi columns), so one step is 2P f(A)b solves; the engine processes
 * columns in chunks whose Lanczos bases fit in device memory.  The step is
 * the plain path's step per pair: deflation by the shared (Q, Lambda) when
 * num_eigs > 0 (lanczos_impl.cu:84-89, update_U.cu:18-27), and the Fisher-KPP
 * (update_U.cu:8-16) or Mitchell-Schaeffer reaction.
 * ------------------------------------------------------------------------- */
typedef struct gm_batch gm_batch;

enum { GM_REACTION_FISHER = 0, GM_REACTION_MS = 2 };

typedef struct gm_batch_desc {
    int32_t n;
    int32_t sparse_format;         /* GM_SPARSE_CSC or GM_SPARSE_CSR         */
    int64_t nnz;
    const double* A_values;
    const int32_t* A_ptr;
    const int32_t* A_ind;
    int32_t num_problems;          /* P                                      */
    const double* U;               /* n x 2P column-major [u_1 g_1 u_2 g_2 ...] */
    int32_t subspace_size;         /* m                                      */
    int32_t gamma_size;
    const double* gamma;
    int32_t contour_size;          /* n_s                                    */
    const double* gamma_s_real;    /* (d+1) x n_s                            */
    const double* gamma_s_imag;
    const double* wts1;            /* n_s x 2 [Re Im], includes s            */
    const double* wts2;
    const double* etaxi;           /* n_s complex shifts, [re im] interleaved */
    double dt;
    const double* mass;            /* n (may be NULL: ones)                  */
    int32_t chunk_columns;         /* columns per pass, power of two dividing
                                      2P (0: largest that fits, <= 256)      */
    uint32_t flags;                /* GM_FLAG_NO_CLAMP; GM_FLAG_AUTOTUNE times
                                      each chunk width <= the allocated one and
                                      keeps the cheapest per column          */
    int32_t device;
    int32_t num_eigs;              /* K <= 64 shared deflation vectors (0: none) */
    const double* Q;               /* n x K column-major                     */
    const double* exp_lambda;      /* K                                      */
    const double* phi_lambda;      /* K                                      */
    int32_t reaction;              /* GM_REACTION_FISHER or GM_REACTION_MS   */
    const double* ode;             /* MS: n x P column-major gate h (NULL: ones) */
    gm_ms_params ms;               /* MS parameters                          */
} gm_batch_desc;

const char* gm_batch_last_error(void);
gm_status gm_batch_create(const gm_batch_desc* desc, gm_batch** out);
gm_status gm_batch_delete(gm_batch* b);
/* nsteps time steps of all problems (asynchronous; HIP graph when use_graph). */
gm_status gm_batch_step(gm_batch* b, int32_t nsteps, int32_t use_graph);
/* dst: n x 2P column-major [u_p, dt g_p].                                  */
gm_status gm_batch_get_U(gm_batch* b, double* dst);
gm_status gm_batch_synchronize(gm_batch* b);
gm_status gm_batch_info(gm_batch* b, int32_t* num_problems, int32_t* chunk_columns);
/* on = 1 when the chunks run their SpMM in column slabs.                  */
gm_status gm_batch_get_column_slabs(gm_batch* b, int32_t* on);
gm_status gm_batch_get_stream(gm_batch* b, void** stream);
/* dst: n x P column-major MS gate h of every problem.                      */
gm_status gm_batch_get_ode(gm_batch* b, double* dst);

/* ---------------------------------------------------------------------------
 * Differentiation (no reference counterpart): u = f(A) b for one f given by
 * its contour weights c_j = s f(xi_j) dxi/dt_j (the quadrature of
 * problem_startup.m:95-103, no preconditioner, no deflation), and the
 * vector-Jacobian product for an upstream gradient g = dL/du:
 *   dL/db = f(A) g,  dL/dkappa (variable-coefficient model A(kappa) with
 *   off-diagonals -(kappa_i + kappa_k)/(2h^2) and boundary terms linear in
 *   kappa_i),  dL/dalpha = Im sum_j (dc_j/dalpha) g^T (xi_j - A)^{-1} b.
 * ------------------------------------------------------------------------- */
typedef struct gm_fab gm_fab;

enum { GM_FAB_RECOMPUTE = 1u << 0 };   /* rebuild the forward basis in backward
                                          instead of keeping it                */

typedef struct gm_fab_desc {
    int32_t n;
    int32_t sparse_format;         /* GM_SPARSE_CSR, diagonal stored         */
    int64_t nnz;
    const double* A_values;
    const int32_t* A_ptr;
    const int32_t* A_ind;
    int32_t subspace_size;         /* m                                      */
    int32_t contour_size;          /* n_s                                    */
    const double* xi;              /* n_s complex nodes, [re im] interleaved */
    const double* wts;             /* n_s complex c_j, interleaved           */
    const double* dwts;            /* n_s complex dc_j/dalpha (may be NULL)  */
    uint32_t flags;                /* GM_FAB_*                               */
    int32_t device;
    /* Deflation (as lanczos_create takes it): f(A) b = Q f(Lambda) Q^T b +
     * the Krylov/contour f of (I - Q Q^T) b.  The contour then only needs to
     * cover the rest of the spectrum.                                       */
    int32_t num_eigs;              /* K (0: none), K <= 64                   */
    const double* Q;               /* n x K column-major, orthonormal        */
    const double* lambda;          /* K eigenvalues of A                     */
    const double* f_lambda;        /* K values f(lambda_k)                   */
    const double* f_divdiff;       /* K x K column-major divided differences
                                      f[lambda_k, lambda_l], diagonal
                                      f'(lambda_k); used by
                                      gm_fab_backward_param (may be NULL)   */
} gm_fab_desc;

const char* gm_fab_last_error(void);
gm_status gm_fab_create(const gm_fab_desc* desc, gm_fab** out);
gm_status gm_fab_delete(gm_fab* f);
/* new values of A (same sparsity), e.g. after kappa changes               */
gm_status gm_fab_set_matrix(gm_fab* f, const double* A_values);
/* new weights (and their alpha derivatives), e.g. after alpha changes      */
gm_status gm_fab_set_weights(gm_fab* f, const double* wts, const double* dwts);
/* u = f(A) b, vectors of length n; u may be NULL to (re)build only the
 * forward state a backward pass needs.  In the gm_fab_* calls the vectors
 * b, u, g, grad_b, kappa and grad_kappa may be host memory or device memory
 * of the engine's GPU (copies by unified addressing); the calls return with
 * the engine's stream synchronised.                                        */
gm_status gm_fab_forward(gm_fab* f, const double* b, double* u);
/* after a forward: any output pointer may be NULL to skip that gradient;
 * kappa (n) is needed for grad_kappa.  With deflation (num_eigs > 0) only
 * grad_b is available here (GM_EINVAL for grad_kappa / grad_alpha).        */
gm_status gm_fab_backward(gm_fab* f, const double* g, const double* kappa, double* grad_b,
                          double* grad_kappa, double* grad_alpha);
/* after a forward: gradients with respect to nparams scalar parameters theta_q
 * of A, given each dA/dtheta_q on A's sparsity pattern (nnz values,
 * symmetric):
 *   grad_theta[q] = Im sum_j c_j z_j^T (dA/dtheta_q) x_j
 * (e.g. theta = the conductivity ratio of an anisotropic operator, which A
 * depends on linearly); grad_b (may be NULL) = f(A) g as in gm_fab_backward.
 * With deflation the Frechet derivative of f splits over the eigenpairs Q
 * and the rest: the Q-Q block by f_divdiff, the cross blocks by real-shift
 * solves (lambda_k - T)^{-1} in the Krylov bases, the rest by the contour. */
gm_status gm_fab_backward_param(gm_fab* f, const double* g, int32_t nparams,
                                const double* const* dA_values, double* grad_b,
                                double* grad_theta);
/* Keeps nparams dA/dtheta_q (A's pattern) on the device, with their products
 * with Q under deflation; gm_fab_backward_param with dA_values = NULL then
 * uses them instead of uploading the derivatives on every call (a time
 * adjoint calls it twice per step with the same ones).                     */
gm_status gm_fab_set_param_derivs(gm_fab* f, int32_t nparams, const double* const* dA_values);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): u = f(A) b as gm_fab_forward computes it without deflation, the
 * same kernels and the same results, with no handle and no allocation.  All
 * arrays are device memory of `device`: A in CSR (n + 1, nnz, nnz), the n_s
 * complex nodes xi and weights wts ([re im] interleaved), b and u (n).  The
 * Krylov basis and the other work buffers live in one device array of 64 MiB
 * that the library keeps on each GPU it runs on until the process ends, which
 * limits n (m + 2) + 10 m n_s to about 8.3 million doubles (GM_EINVAL beyond:
 * e.g. n = 100,000 at m = 80); calls that overlap on one GPU race for it.
 * Work is enqueued on `stream` (NULL: the null stream); with status non-NULL
 * the call waits for it and stores the status bits (non-zero bit 1 or 2:
 * breakdown, returned as GM_EBREAKDOWN), otherwise it returns once enqueued.
 * Errors: gm_fab_last_error().                                              */
gm_status gm_fab_forward_noalloc(int32_t n, int64_t nnz, const int32_t* A_ptr,
                                 const int32_t* A_ind, const double* A_values,
                                 int32_t subspace_size, int32_t contour_size, const double* xi,
                                 const double* wts, const double* b, double* u, void* stream,
                                 int32_t device, uint32_t* status);

/* ---------------------------------------------------------------------------
 * Uneven batch (no reference counterpart): P independent problems of
 * different sizes -- each its own A_p, [u_p, g_p], contour weights, shifted
 * gammas and shifts, as lanczos_create takes them for one problem -- stepped
 * together, every stage of a step one launch over all problems.  m, gamma,
 * n_s and dt are shared; no deflation; A_p in CSR.
 * ------------------------------------------------------------------------- */
typedef struct gm_ubatch gm_ubatch;

typedef struct gm_ubatch_desc {
    int32_t num_problems;               /* P                                  */
    const int32_t* n;                   /* P row counts                       */
    const int64_t* nnz;                 /* P nonzero counts                   */
    int32_t sparse_format;              /* GM_SPARSE_CSR                      */
    const double* const* A_values;      /* P arrays (CSR of each A_p)         */
    const int32_t* const* A_ptr;
    const int32_t* const* A_ind;
    const double* const* U;             /* P arrays, n_p x 2 column-major
                                           [u, g(u)] (column 1 scaled by dt) */
    const double* const* mass;          /* P arrays of n_p, or NULL (ones)    */
    int32_t subspace_size;              /* m                                  */
    int32_t gamma_size;                 /* d + 1                              */
    const double* gamma;                /* shared polynomial, highest first   */
    int32_t contour_size;               /* n_s                                */
    const double* gamma_s_real;         /* P blocks of (d+1) x n_s            */
    const double* gamma_s_imag;
    const double* wts1;                 /* P blocks of n_s x 2 [Re Im]        */
    const double* wts2;
    const double* shifts;               /* P blocks of 4 m n_s                */
    double dt;
    uint32_t flags;                     /* GM_FLAG_NO_CLAMP; GM_FLAG_AUTOTUNE
                                           times whole steps at create under each
                                           SpMM schedule (512/1024/2048-nonzero
                                           row blocks, 16- or 32-bit column
                                           offsets, early or late epilogue
                                           loads) and keeps the fastest       */
    int32_t device;
} gm_ubatch_desc;

const char* gm_ubatch_last_error(void);
gm_status gm_ubatch_create(const gm_ubatch_desc* desc, gm_ubatch** out);
gm_status gm_ubatch_delete(gm_ubatch* u);
/* nsteps time steps of every problem (asynchronous; HIP graph when use_graph). */
gm_status gm_ubatch_step(gm_ubatch* u, int32_t nsteps, int32_t use_graph);
/* dst: n_p x 2 column-major [u_p, dt g_p] of problem p.                    */
gm_status gm_ubatch_get_U(gm_ubatch* u, int32_t problem, double* dst);
gm_status gm_ubatch_synchronize(gm_ubatch* u);
gm_status gm_ubatch_info(gm_ubatch* u, int32_t* num_problems, int64_t* total_rows,
                         int32_t* row_blocks);
/* The SpMM schedule in use: nonzeros per row block, bytes per stored column
 * index (2 or 4), and whether the epilogue operands are loaded early.      */
gm_status gm_ubatch_get_schedule(gm_ubatch* u, int32_t* tile, int32_t* index_bytes,
                                 int32_t* early_loads);
/* The shifted solves' form, as gm_get_solve_form: 0 / 1 one lane per system
 * (factors in global memory / LDS), 2 one wavefront per system.            */
gm_status gm_ubatch_get_solve_form(gm_ubatch* u, int32_t* form);

/* ---------------------------------------------------------------------------
 * Host set-up of the create inputs (host code only; usable without a GPU).
 * What problem_startup.m:71-144 computes in MATLAB before lanczos_create.
 * Complex arrays are [re im] interleaved unless stated; wts1/wts2 are n x 2
 * column-major [Re Im] as lanczos_create takes them.
 * ------------------------------------------------------------------------- */
/* gamma (degree+1), highest power first: polynomial_coeffs.m:11-28.        */
gm_status gm_polynomial_coeffs(int32_t degree, double prob_dim, double* gamma);
/* HHT nodes xi (n complex), dxi/dt (n complex, may be NULL), weights
 * s f1(xi) dxi/dt and s f2(xi) dxi/dt (n x 2 each), s (may be NULL), for
 * f1 = exp(-eta x^(alpha/2)), f2 = phi(-eta x^(alpha/2)):
 * problem_startup.m:72-74, 80-103 (n is an input here; the reference picks
 * it with fminbnd at problem_startup.m:92).                                */
gm_status gm_contour_setup(double l1, double lN, int32_t n, double eta, double alpha,
                           double* xi, double* dxidt, double* wts1, double* wts2,
                           double* s);
/* gamma_s_real/imag ((d+1) x n column-major), etaxi (n complex, may be
 * NULL) and the packed shifts vector (4 m n doubles, may be NULL):
 * problem_startup.m:106-123, 140-144.                                      */
gm_status gm_shift_setup(const double* gamma, int32_t gamma_size, const double* xi,
                         int32_t n, int32_t m, double* gamma_s_real, double* gamma_s_imag,
                         double* etaxi, double* shifts);
/* exp_lambda, phi_lambda (K) with phi_lambda[0] = 1: problem_startup.m:72-77. */
gm_status gm_eig_weights(const double* lambda, int32_t K, double eta, double alpha,
                         double* exp_lambda, double* phi_lambda);
/* Subspace size m = ceil(log(tol/2) / log(k)), k from the extreme values of
 * x p(x) on [l1, lN]: problem_startup_nonuniform.m:101-113 (real critical
 * points only).                                                            */
gm_status gm_subspace_size(const double* gamma, int32_t gamma_size, double l1, double lN,
                           double tol, int32_t* m);
/* Reverse Cuthill-McKee order of the symmetric pattern (ptr n + 1, ind) of an
 * n x n matrix, in place of symrcm (process_nonuniform.m:15): perm[q] is the
 * old index of the node placed at q (A(perm, perm) is banded).  Host only.  */
gm_status gm_rcm_order(int32_t n, const int32_t* ptr, const int32_t* ind, int32_t* perm);
/* Cache-locality order (no reference counterpart): recursive breadth-first
 * bisection of the symmetric pattern down to parts of at most `leaf` nodes,
 * the parts depth-first and each part's nodes in increasing index; perm as
 * for gm_rcm_order.  Used by gm_batch with GM_FLAG_LOCALITY_ORDER.  Host only. */
gm_status gm_locality_order(int32_t n, const int32_t* ptr, const int32_t* ind, int32_t leaf,
                            int32_t* perm);

/* ---------------------------------------------------------------------------
 * f(A) on R columns at once with the parameter adjoint (no reference
 * counterpart): the multi-column form of gm_fab_* for problems that share A,
 * the contour and the deflation pairs (e.g. the stimuli of a monodomain run).
 * Columns pair like the batched engine's: even columns use wts1 and
 * f1_lambda, odd columns wts2 and f2_lambda.  Vectors are N x R row-major
 * (column c of row i at i*R + c), in host or device memory.
 * ------------------------------------------------------------------------- */
typedef struct gm_bfab gm_bfab;

typedef struct gm_bfab_desc {
    int32_t n;
    int32_t sparse_format;         /* GM_SPARSE_CSR, A symmetric             */
    int64_t nnz;
    const double* A_values;
    const int32_t* A_ptr;
    const int32_t* A_ind;
    int32_t columns;               /* R: 2, 4, 8, 16, 32 or 64               */
    int32_t subspace_size;         /* m                                      */
    int32_t contour_size;          /* n_s                                    */
    const double* xi;              /* n_s complex nodes, interleaved         */
    const double* wts1;            /* n_s complex weights, even columns      */
    const double* wts2;            /* n_s complex weights, odd columns       */
    int32_t num_eigs;              /* K (0: none), K <= 64                   */
    const double* Q;               /* n x K column-major, orthonormal        */
    const double* lambda;          /* K                                      */
    const double* f1_lambda;       /* K: f1(lambda_k) (even columns)         */
    const double* f2_lambda;       /* K: f2(lambda_k) (odd columns)          */
    const double* f1_divdiff;      /* K x K column-major divided differences */
    const double* f2_divdiff;      /*   of f1, f2 (parameter gradient only)  */
    int32_t device;
} gm_bfab_desc;

const char* gm_bfab_last_error(void);
gm_status gm_bfab_create(const gm_bfab_desc* desc, gm_bfab** out);
gm_status gm_bfab_delete(gm_bfab* f);
/* U = f(A) B column by column; U may be NULL to build only the forward state.
 * With U NULL the call returns once B has been copied and leaves the run in
 * flight on the engine's second stream; the next call on the engine waits for
 * it (a gm_bfab_backward_param overlaps it with the adjoint's own run) and
 * reports its breakdown flags.                                              */
gm_status gm_bfab_forward(gm_bfab* f, const double* B, double* U);
/* dA/dtheta_q (A's pattern, symmetric) kept on the device for the backward.  */
gm_status gm_bfab_set_param_derivs(gm_bfab* f, int32_t nparams, const double* const* dA_values);
/* After a forward: grad_B = f(A) G (may be NULL) and, per stored parameter q
 * and column c, grad_theta[q*R + c] = d(g_c^T f(A) b_c)/dtheta_q (host
 * memory).  With m > 48 or fewer than 16 columns, the first call with
 * stored parameters allocates one N x R x m work block, kept until
 * gm_bfab_delete.                                                           */
gm_status gm_bfab_backward_param(gm_bfab* f, const double* G, double* grad_B, double* grad_theta);
gm_status gm_bfab_info(gm_bfab* f, int32_t* columns, int32_t* num_params);

/* ---------------------------------------------------------------------------
 * Deflation eigenpairs on the GPU.  Replaces the host eigensolve that
 * produces Q and Lambda before lanczos_create -- eigs(K, M, 'sm') in
 * process_nonuniform.m:106-110 (MATLAB, sparse factorisation), ARPACK
 * (scipy eigsh) in the Python drivers.  The K smallest eigenpairs of the
 * symmetric A by Chebyshev-filtered subspace iteration on a block of
 * block_size columns; converges quickly when lambda_max / lambda_K is
 * moderate (grid operators), slowly when it is huge (sliver elements).
 * ------------------------------------------------------------------------- */
typedef struct gm_eigs_desc {
    int32_t n;
    int64_t nnz;
    int32_t sparse_format;         /* CSR or CSC: equal arrays, A symmetric */
    const double* A_values;
    const int32_t* A_ptr;
    const int32_t* A_ind;
    int32_t num_eigs;              /* K >= 1                                */
    int32_t block_size;            /* 16, 32 or 64; > K, <= n               */
    int32_t degree;                /* Chebyshev filter degree per pass      */
    int32_t max_passes;            /* filter passes before giving up        */
    double tol;                    /* max_i ||A q_i - l_i q_i|| / |l_i|     */
    double upper_bound;            /* >= lambda_max, or <= 0: Gershgorin    */
    uint64_t seed;                 /* start block: uniform [-1, 1) deviates */
    int32_t device;
} gm_eigs_desc;

const char* gm_eigs_last_error(void);
/* lambda (K, ascending) and, when Q is not NULL, Q (n x K column-major,
 * orthonormal columns).  passes / max_residual (either may be NULL) report
 * the filter passes taken and the final max relative residual.  Returns
 * GM_EBREAKDOWN, with the outputs filled, when tol is not met in
 * max_passes.                                                             */
gm_status gm_eigs_smallest(const gm_eigs_desc* desc, double* lambda, double* Q,
                           int32_t* passes, double* max_residual);
/* Host helper (no GPU): A = V diag(w) V^T for a symmetric n x n row-major A
 * by cyclic Jacobi; w ascending, V (may be NULL) row-major with the
 * eigenvectors as columns.                                                */
gm_status gm_sym_eig(int32_t n, const double* A, double* w, double* V);

/* ---------------------------------------------------------------------------
 * Deflation eigenpairs through a sparse factorisation (host code, threads;
 * no GPU needed).  Replaces eigs(K, M, k, 'sm') of process_nonuniform.m:38
 * and :106-110 -- MATLAB factors K - sigma M and runs shift-invert Lanczos --
 * for operators whose lambda_max / lambda_K is too large for the Chebyshev
 * filter of gm_eigs_smallest (sliver elements of an unstructured mesh), and
 * eigs(K, M, 1, 'lm') of process_nonuniform.m:40 (gm_lambda_max).  A is the
 * symmetric n x n CSR matrix with both triangles (int32 indices); with the
 * lumped mass, A = M^(-1/2) K M^(-1/2) and its orthonormal eigenvectors are
 * M^(1/2) times the M-orthonormal ones of (K, M) (process_nonuniform.m:107).
 * Results do not depend on the thread count.
 * ------------------------------------------------------------------------- */
typedef struct gm_chol gm_chol;

const char* gm_chol_last_error(void);
/* Nested-dissection order of the symmetric pattern: perm[q] is the old index
 * of the node placed at q.                                                  */
gm_status gm_nd_order(int32_t n, const int32_t* ptr, const int32_t* ind, int32_t* perm);
/* Cholesky factor of P (A - sigma I) P^T (multifrontal, supernodal) with P
 * from gm_nd_order refined to a postorder of the elimination tree.
 * threads <= 0: all hardware threads.  GM_EBREAKDOWN if A - sigma I is not
 * positive definite.                                                        */
gm_status gm_chol_factor(int32_t n, const int32_t* ptr, const int32_t* ind, const double* val,
                         double sigma, int32_t threads, gm_chol** out);
/* X <- (A - sigma I)^(-1) X for nrhs columns (n x nrhs column-major, host). */
gm_status gm_chol_solve(gm_chol* f, int32_t nrhs, double* X);
/* nnz(L) (stored, incl. the diagonal), factorisation flops, supernodes, and
 * seconds[3] = ordering, symbolic, numeric (any may be NULL).               */
gm_status gm_chol_info(gm_chol* f, int64_t* nnz_l, double* flops, int32_t* nsuper,
                       double* seconds);
/* The factor's final order (n): perm[q] = old index at q.                   */
gm_status gm_chol_perm(gm_chol* f, int32_t* perm);
gm_status gm_chol_delete(gm_chol* f);

typedef struct gm_sieigs_desc {
    int32_t n;
    int64_t nnz;
    const double* A_values;        /* CSR, both triangles, symmetric         */
    const int32_t* A_ptr;
    const int32_t* A_ind;
    int32_t num_eigs;              /* K >= 1 smallest                        */
    int32_t block_size;            /* K <= block_size <= min(512, n / 2)     */
    int32_t max_blocks;            /* Lanczos blocks per restart (0: 12)     */
    int32_t max_restarts;          /* < 0: 4                                 */
    double tol;                    /* ||A q - l q||_2 / ||A||_inf per pair   */
                                   /*  (<= 0: 1e-14)                         */
    double sigma;                  /* shift below lambda_1 (A - sigma I SPD) */
    uint64_t seed;                 /* start block: uniform [-1, 1) deviates  */
    int32_t threads;               /* <= 0: all hardware threads             */
} gm_sieigs_desc;

/* lambda (K, ascending) and, when Q is not NULL, Q (n x K column-major,
 * orthonormal columns) by block Lanczos on (A - sigma I)^(-1) with full
 * reorthogonalisation; a restart cycle ends when the recurrence's residual
 * estimate passes tol, and the pairs are accepted on their explicit
 * residuals.  blocks (may be NULL) = block solves taken, max_residual (may
 * be NULL) = max ||A q - l q||_2 / ||A||_inf.  GM_EBREAKDOWN, with the
 * outputs filled, if tol is not met in max_restarts restarts.              */
gm_status gm_eigs_shift_invert(const gm_sieigs_desc* desc, double* lambda, double* Q,
                               int32_t* blocks, double* max_residual);
/* lambda_max by Lanczos from a seeded start vector (the largest Ritz value,
 * stopped when it changes by <= tol relative three times running; tol <= 0:
 * 1e-12, max_iter <= 0: 300).  iters may be NULL.                          */
gm_status gm_lambda_max(int32_t n, const int32_t* ptr, const int32_t* ind, const double* val,
                        double tol, int32_t max_iter, uint64_t seed, int32_t threads,
                        double* lmax, int32_t* iters);

#ifdef __cplusplus
}
#endif

#endif /* GPUMATFUN_H */
