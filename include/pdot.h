/*
 * pdot.h -- C ABI of libpdot_hip.so, the MI355X (gfx950) hot path of PDOT:
 * restarted adaptive PDHG for discrete optimal transport
 *     min <C,X>  s.t.  X 1 = a,  X^T 1 = b,  X >= 0.
 *
 * The library owns all device memory and one HIP stream per context.  Host
 * pointers are borrowed for the duration of a call.  Matrices cross the
 * boundary column-major with leading dimension n (element (i,j) at i + j*n),
 * as Julia's Matrix{Float64} lays them out.  Every call returns 0 on success
 * or a negative PDOT_E_* code; pdot_last_error() then describes the failure.
 * A context is not thread-safe; distinct contexts may be used concurrently.
 *
 * The host control logic (restart scheme, termination, primal-weight update,
 * logs) stays with the caller; the library performs the n*m work and the
 * per-iteration step-size recurrence on the device.  Each entry point names
 * the reference code it replaces (paths relative to the PDOT.jl tree).
 */
#ifndef PDOT_H
#define PDOT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdot_ctx pdot_ctx;

enum {
    PDOT_OK = 0,
    PDOT_E_ARG = -1,   /* invalid argument (sizes, pointers, slot state) */
    PDOT_E_HIP = -2,   /* a HIP runtime call failed */
    PDOT_E_STATE = -3, /* call not valid in the current solver state */
    PDOT_E_NOMEM = -4  /* device allocation failed */
};

enum { PDOT_STORAGE_F64 = 0, PDOT_STORAGE_F32 = 1 };
enum { PDOT_POLICY_ADAPTIVE = 0, PDOT_POLICY_CONSTANT = 1 };
enum { PDOT_POINT_CURRENT = 0, PDOT_POINT_AVERAGE = 1 };
enum { PDOT_COST_SQEUCLID = 1, PDOT_COST_L1 = 2 };

/* Device-resident solver scalars: the scalar fields of CuPdhgSolverState
 * (src/solver_gpu.jl:23-36) and of CuSolutionWeightedAverage
 * (src/saddle_point_gpu.jl:52-65), plus the outputs of the last pass.
 * cumulative_kkt_passes counts step passes only; the caller adds the 2.0 per
 * evaluation (src/solver_gpu.jl:484, 534). */
typedef struct pdot_solver_state {
    double step_size;             /* solver_state.step_size (eta) */
    double primal_weight;         /* solver_state.primal_weight (omega) */
    double cumulative_kkt_passes; /* +1 per PDHG pass (solver_gpu.jl:336) */
    double sum_weights;           /* sum_*_solution_weights (all three are equal) */
    int64_t total_number_iterations;
    int64_t avg_count;            /* *_solutions_count (all three are equal) */
    int64_t take_step_calls;      /* take_step! invocations so far */
    int64_t error_call;           /* take_step! call that set numerical_error, else 0 */
    double last_interaction;      /* compute_interaction_and_movement (solver_gpu.jl:267-288) */
    double last_movement;
    double last_step_limit;
    double last_dx2;              /* ||Delta X||_F^2 of the last pass */
    double last_dot;              /* dr'du + dc'dv of the last pass */
    double last_dy2;              /* ||[du; dv]||^2 of the last pass */
    double reduction_exponent;    /* AdaptiveStepsizeParams (solver_gpu.jl:2-5) */
    double growth_exponent;
    int32_t step_policy;          /* PDOT_POLICY_* */
    int32_t numerical_error;      /* solver_state.numerical_error; halts further passes */
    int32_t gzero;                /* the primal gradient is still the initial zero matrix */
    int32_t cur_slot;             /* internal: slot holding the current point */
    int32_t avg_slot;             /* internal: slot holding the average (valid if avg_count > 0) */
    int32_t pending;              /* internal: a speculative pass awaits commit */
    int32_t spare0, spare1;       /* internal: slots written by the pending pass */
    int64_t step_budget;          /* internal (batches): passes left in the current run */
    int32_t defer_ok;             /* internal: passes may defer the average's update */
    int32_t avg_pending;          /* internal: the average lacks the current point's contribution */
    double pend_nw, pend_aw;      /* internal: the averaging weights of that contribution */
} pdot_solver_state;

/* Reductions of one point (X, u, v, r, c, G) with G = (C - u 1^T) - 1 v^T,
 * from which the caller forms ConvergenceInformation (src/utils_gpu.jl:68-129)
 * and the weighted KKT residual (src/saddle_point_gpu.jl:299-330). */
typedef struct pdot_point_stats {
    double primal_objective;  /* <C, X>                       utils_gpu.jl:84 */
    double x_inf;             /* ||X||_inf                    utils_gpu.jl:102 */
    double x_sq;              /* ||X||_F^2                    utils_gpu.jl:103 */
    double rc_inf;            /* ||min(G,0)||_inf             utils_gpu.jl:52 */
    double rc_sq;             /* ||min(G,0)||_F^2             utils_gpu.jl:108 */
    double p_inf;             /* ||[a-r; b-c]||_inf           utils_gpu.jl:94 */
    double p_sq;              /* ||[a-r; b-c]||_2^2           utils_gpu.jl:95 */
    double dual_objective;    /* a'u + b'v                    utils_gpu.jl:54 */
    double u_inf;             /* ||u||_inf  (= ||[u;u]||_inf, utils_gpu.jl:115) */
    double u_sq;              /* ||u||_2^2                    utils_gpu.jl:116 */
    double y_sq;              /* ||[u; v]||_2^2               saddle_point_gpu.jl:765 */
} pdot_point_stats;

/* Create a context for an n x m problem on HIP device `device`, storing C, X
 * and the running average X_avg in FP64 (PDOT_STORAGE_F64) or FP32
 * (PDOT_STORAGE_F32, arithmetic and duals stay FP64).  Replaces the
 * CUDA.zeros allocations of src/solver_gpu.jl:428-471. */
int pdot_create(pdot_ctx** out, int64_t n, int64_t m, int32_t storage, int32_t device);
void pdot_destroy(pdot_ctx* ctx);
const char* pdot_last_error(const pdot_ctx* ctx);

/* Host-to-device copy of the problem, ot_cpu_to_gpu (src/io.jl:18-36).
 * C: n*m column-major FP64; a: n; b: m.  C streams through a pinned
 * two-buffer staging ring (host copy of one chunk overlapping the DMA of the
 * previous one); FP32 storage is narrowed on the device; the same pass
 * computes the norms of cached_ot_info. */
int pdot_upload(pdot_ctx* ctx, const double* C, const double* a, const double* b);

/* Point-cloud form of the problem, an extension beyond the reference's dense
 * Matrix{Float64} contract: the cost is never stored but formed in registers
 * each pass, C_ij = (sum_d f(y_jd - x_id)) / s with f(t) = t*t
 * (PDOT_COST_SQEUCLID) or |t| (PDOT_COST_L1), summed from 0 in d order, and
 * s = max_ij of the raw sum if `normalize`, else 1 -- bit for bit the matrix a
 * dense caller builds that way, so results equal those of pdot_upload of
 * that matrix (FP64 storage).  With FP32 storage the division by s is a
 * multiply by 1/s, within about one FP64 ulp of the quotient (the dense FP32
 * form stores C rounded to FP32, 2^29 times coarser); a normalising maximum
 * that is finite and outside [2^-1022, 2^1022] is then rejected (PDOT_E_ARG).
 * x: n x dim, y: m x dim (row-major), dim 1..3.  One n*m stream per pass
 * fewer than the dense form; no n*m allocation for C. */
int pdot_upload_points(pdot_ctx* ctx, int32_t cost_kind, int32_t dim, const double* x, const double* y,
                       int32_t normalize, const double* a, const double* b);

/* cached_ot_info (src/termination.jl:94-101) of the uploaded problem, in the
 * CachedOTInfo field order: ||C||_inf, ||[a;b]||_inf, ||C||_F, ||[a;b]||_2.
 * The matrix norms come from the upload pass on the device (from the FP64
 * input, for either storage); no extra host pass over C. */
int pdot_cached_ot_info(pdot_ctx* ctx, double* out4);

/* Reset the iterate: X = 0, r = c = 0, empty average, counters 0
 * (src/solver_gpu.jl:428-441, 476-496).  u0/v0 NULL gives the reference's zero
 * start with G = 0; otherwise a warm start (u0, v0) with G0 = (C - u0) - v0.
 * step_policy/exponents select AdaptiveStepsizeParams or ConstantStepsizeParams;
 * step_size and primal_weight are the initial eta and omega. */
int pdot_set_start(pdot_ctx* ctx, const double* u0, const double* v0, int32_t step_policy,
                   double reduction_exponent, double growth_exponent, double step_size,
                   double primal_weight);

/* Run `count` take_step! calls (src/solver_gpu.jl:290-367 adaptive, :370-410
 * constant) with the step-size recurrence on the device and no host round
 * trip; a pass pending from pdot_eval is committed first.  Returns once the
 * state is committed; `out` (may be NULL) receives it.  A numerical error halts
 * the remaining passes and is reported in out->numerical_error/error_call. */
int pdot_run_steps(pdot_ctx* ctx, int64_t count, pdot_solver_state* out);

/* Evaluate the committed point and its running average in one streaming pass
 * (compute_convergence_information, src/utils_gpu.jl:68-129, and the
 * n*m parts of compute_weight_kkt_residual, src/saddle_point_gpu.jl:299-330),
 * while speculatively performing the next take_step! into spare buffers.
 * The pass is committed by the next pdot_run_steps/pdot_eval and discarded by
 * pdot_restart or pdot_discard.  `average` is zeroed when the average is empty.
 * `state` (may be NULL) receives the pre-pass state. */
int pdot_eval(pdot_ctx* ctx, pdot_point_stats* current, pdot_point_stats* average,
              pdot_solver_state* state);

/* Discard a pending speculative pass (no-op if none). */
int pdot_discard(pdot_ctx* ctx);

/* Restart, src/saddle_point_gpu.jl:639-696: discard any pending pass; if
 * to_average, the average becomes the current point (X, u, v, r, c and,
 * implicitly, G); the average is reset (:667); omega is set to
 * new_primal_weight (compute_new_primal_weight, :701-728, runs on the host). */
int pdot_restart(pdot_ctx* ctx, int32_t to_average, double new_primal_weight);

/* Read or overwrite the device solver scalars. */
int pdot_get_state(pdot_ctx* ctx, pdot_solver_state* out);
int pdot_set_state(pdot_ctx* ctx, const pdot_solver_state* in);

/* Device-to-host copy of the current (PDOT_POINT_CURRENT) or average point,
 * gpu_to_cpu! (src/io.jl:41-53).  X: n*m column-major FP64 (may be NULL),
 * u: n, v: m (may be NULL).  FP32 storage is widened on the device. */
int pdot_download(pdot_ctx* ctx, int32_t which, double* X, double* u, double* v);

/* Feasibility rounding of the current or average plan onto the transport
 * polytope of (a, b), rounding(problem, primal_solution) of
 * src/utils_gpu.jl:296-318, computed on the device from the resident point
 * (FP32 storage is widened to FP64 first).  X_rounded: n*m column-major
 * FP64.  The point itself is left unchanged. */
int pdot_round(pdot_ctx* ctx, int32_t which, double* X_rounded);

/* The same rounding of a host plan X (n*m column-major FP64), in place, as
 * the reference's rounding() overwrites its primal_solution argument
 * (src/utils_gpu.jl:300-318).  Needs no context; errbuf (errlen bytes, may be
 * NULL) receives the error text. */
int pdot_round_matrix(int32_t device, int64_t n, int64_t m, const double* a, const double* b, double* X,
                      char* errbuf, int64_t errlen);

/* No counterpart in the reference: the rounding of pdot_round_matrix for a
 * plan already in device memory, for callers that cannot allocate device
 * memory (tight serving loops, code that must not call hipMalloc).  It calls
 * no hipMalloc, hipFree or other allocating routine: its workspace (row and
 * column partial sums, scale and error vectors) is a static device array of
 * 64 MiB per GPU, held until the process ends, which covers plans up to about
 * 30000 x 30000 on a 256-CU device; larger plans are rejected with
 * PDOT_E_ARG.  X (ldx >= n), Y (ldy >= n), a and b are FP64 device pointers
 * of `device` whose allocations hold the given sizes (checked); Y may be X
 * (in place, ldy == ldx) but must not otherwise overlap it.  The kernels run
 * on `stream` (a hipStream_t of `device`, NULL for the null stream), and the
 * call returns once it has synchronised that stream.  Calls on one GPU share
 * the workspace: calls that overlap in time on the same GPU race and must be
 * serialised by the caller. */
int pdot_round_device_noalloc(int32_t device, void* stream, int64_t n, int64_t m, const double* a, const double* b,
                              const double* X, int64_t ldx, double* Y, int64_t ldy, char* errbuf, int64_t errlen);

/* Host-side text output, write_vector_to_file of scripts/solve.jl:5-11: one
 * element per line (a matrix in column-major order), each as Julia's println
 * prints a Float64 (shortest round-trip digits; "0.0001", "123456.0",
 * "1.0e-5", "1.234567e6", "NaN", "Inf").  Formatting runs on host threads. */
int pdot_write_julia_vector(const char* path, const double* x, int64_t count, char* errbuf, int64_t errlen);

/* The Julia text of one Float64 into out (outlen >= 33); returns its length
 * or a negative PDOT_E_* code. */
int pdot_format_julia_float(double x, char* out, int64_t outlen);

/* Timing hook: run `count` take_step! calls like pdot_run_steps while
 * recording HIP events on the context's stream around every launch of the
 * streaming step kernel.  step_kernel_ms = summed step-kernel time,
 * total_ms = elapsed time of the whole sequence. */
int pdot_time_steps(pdot_ctx* ctx, int64_t count, double* step_kernel_ms, double* total_ms);

/* Use a captured hipGraph for pdot_run_steps sequences (1) or plain stream
 * launches (0, the default). */
int pdot_use_graphs(pdot_ctx* ctx, int32_t enable);

/* Deferred averaging (on by default): an ordinary pass may leave the running
 * average's update to the next pass, which applies it in registers, so that
 * every second pass streams three n x m arrays instead of five.  Results are
 * bit for bit those of updating the average every pass.  enable = 0 turns it
 * off (for comparison); a lagging contribution is written out first.  No
 * counterpart in the reference (a data-movement choice of this library). */
int pdot_set_defer(pdot_ctx* ctx, int32_t enable);

/* Passes that leave the average alone (every second ordinary pass with
 * deferred averaging) run a kernel compiled for that kind, without the
 * average's loads and registers, where the form has one (stored costs in FP64
 * and FP32, point clouds with FP32 storage); enable = 0 runs the kernel that
 * decides at run time for every pass, for comparison.  Results are identical either way.  No
 * counterpart in the reference (a kernel-selection choice of this library). */
int pdot_set_pass_kernels(pdot_ctx* ctx, int32_t enable);

/* Where the controller (take_step!'s acceptance test, step-size recurrence
 * and commit) runs: mode 1 in the last block of the reduce kernel to finish
 * (its blocks publish their scalars without a device-scope fence), 0 in a
 * one-wave kernel after it, -1 (the default) by size (1 for problems of up
 * to 256 reduce blocks, n + m <= 16384).  Results are identical either way.
 * No counterpart in the reference (a kernel-selection choice of this
 * library). */
int pdot_set_controller_fold(pdot_ctx* ctx, int32_t mode);

/* How the step kernel streams the matrices: mode 0 with the nontemporal hint,
 * 1 without it (the lines may stay in the 256 MB Infinity Cache for the next
 * pass), -1 (the default) by size (1 while five n x m arrays of the storage
 * type take at most 352 MiB).  Results are identical either way.  No
 * counterpart in the reference (a data-movement choice of this library). */
int pdot_set_stream_policy(pdot_ctx* ctx, int32_t mode);

/* Zero maps (on by default): every iterate slot keeps one bit per 16 bytes of
 * each column saying whether those entries are all +0.0; a pass neither
 * reads such entries nor writes a new point's entries where a group of 64
 * bytes holds only +0.0, so the iterates' zeros (most of an optimal
 * transport plan) cost no matrix traffic.  Results are bit for bit those of
 * dense passes.  enable = 0 streams every entry (for comparison); the slots
 * are made dense first.  No counterpart in the reference (a data-movement
 * choice of this library). */
int pdot_set_sparse(pdot_ctx* ctx, int32_t enable);

/* Share of the 16-byte column units of the slot holding the point (`which`:
 * PDOT_POINT_CURRENT or PDOT_POINT_AVERAGE; the average's slot as the next
 * pass reads it, without a lagging contribution) that its zero map records
 * as +0.0, i.e. that passes do not read; 0 when zero maps are off.  A unit
 * is recorded as +0.0 only when it is; the share is exact for a slot written
 * by a pass with zero maps on.  Changes no state.  No counterpart in the
 * reference (a diagnostic of this library). */
int pdot_zero_share(pdot_ctx* ctx, int32_t which, double* share);

/* Launch geometry of the streaming step kernel: rows per block, column
 * blocks, columns per block, total blocks.  col_blocks_hint > 0 overrides the
 * default choice (used by the launch-shape tuner); 0 keeps the default.
 * interleave = 1 (the default) makes column block cb take columns cb,
 * cb + Cb, ... instead of a contiguous range. */
int pdot_set_launch(pdot_ctx* ctx, int32_t col_blocks_hint, int32_t interleave);
int pdot_get_launch(pdot_ctx* ctx, int32_t* rows_per_block, int32_t* col_blocks, int32_t* cols_per_block,
                    int32_t* blocks);

/* Bytes of device memory held by the context. */
int64_t pdot_device_bytes(const pdot_ctx* ctx);

/* Measurement hook with no counterpart in the reference: copies `bytes`
 * (a multiple of 16) between two scratch buffers with a plain 16-byte copy
 * kernel, `reps` times after one warm-up, and returns the mean time per copy.
 * Used to calibrate HBM byte counters and bandwidth against a kernel of known
 * traffic (2 * bytes per copy). */
int pdot_copy_probe(pdot_ctx* ctx, int64_t bytes, int32_t reps, double* ms_per_copy);

/* ------------------------------------------------------------------------
 * Batches: many independent problems of uneven n x m advanced together.
 * No counterpart in the reference, which solves one problem per optimize
 * call (src/solver_gpu.jl:413); every instance follows exactly the
 * single-problem semantics above, with its own step size, primal weight,
 * averaging weights, restart and termination state.  Per-instance arrays
 * have `count` entries.
 * ------------------------------------------------------------------------ */
typedef struct pdot_batch pdot_batch;

int pdot_batch_create(pdot_batch** out, int32_t count, const int64_t* n, const int64_t* m, int32_t storage,
                      int32_t device);
void pdot_batch_destroy(pdot_batch* b);
const char* pdot_batch_last_error(const pdot_batch* b);
/* ot_cpu_to_gpu for instance k (src/io.jl:18-36). */
int pdot_batch_upload(pdot_batch* b, int32_t k, const double* C, const double* a, const double* bvec);
/* Zero start for every instance (src/solver_gpu.jl:428-496); step_size and
 * primal_weight are per instance.  All instances become active. */
int pdot_batch_set_start(pdot_batch* b, int32_t step_policy, double reduction_exponent, double growth_exponent,
                         const double* step_size, const double* primal_weight);
/* Optional warm start (u0, v0) of instance k, after pdot_batch_set_start and
 * before any pass: X0 = 0, G0 = (C - u0) - v0. */
int pdot_batch_set_warm_start(pdot_batch* b, int32_t k, const double* u0, const double* v0);
/* Active instances take part in passes and evaluations; inactive ones (for
 * example terminated) are skipped.  Rebuilds the work lists. */
int pdot_batch_set_active(pdot_batch* b, const int32_t* active);
/* counts[k] take_step! calls for each active instance (the pending pass of an
 * evaluation is committed first); instances run concurrently in one launch
 * sequence of max(counts) passes. */
int pdot_batch_run_steps(pdot_batch* b, const int64_t* counts, pdot_solver_state* out);
/* pdot_eval for every active instance. */
int pdot_batch_eval(pdot_batch* b, pdot_point_stats* current, pdot_point_stats* average, pdot_solver_state* states);
/* Per instance: action 0 none, 1 restart to the current point, 2 restart to
 * the average (src/saddle_point_gpu.jl:639-696); new_primal_weight applies
 * where action != 0. */
int pdot_batch_restart(pdot_batch* b, const int32_t* action, const double* new_primal_weight);
int pdot_batch_discard(pdot_batch* b, const int32_t* which);
int pdot_batch_get_states(pdot_batch* b, pdot_solver_state* out);
int pdot_batch_set_states(pdot_batch* b, const pdot_solver_state* in);
int pdot_batch_download(pdot_batch* b, int32_t k, int32_t which, double* X, double* u, double* v);
int pdot_batch_time_steps(pdot_batch* b, int64_t count, double* step_kernel_ms, double* total_ms);
int pdot_batch_use_graphs(pdot_batch* b, int32_t enable);
/* Deferred averaging for every instance of the batch, as pdot_set_defer does
 * for a single problem (on by default; bit-identical results).  No
 * counterpart in the reference. */
int pdot_batch_set_defer(pdot_batch* b, int32_t enable);

/* Zero maps in the batch (on by default), as pdot_set_sparse for every
 * instance: iterate entries recorded as +0.0 are neither read nor written.
 * Results are bit for bit those of dense passes; enable = 0 streams every
 * entry (the slots are expanded first).  No counterpart in the reference. */
int pdot_batch_set_sparse(pdot_batch* b, int32_t enable);
/* Launch shape of the persistent step kernel (tuned by pdot_amd.batch.autotune):
 * columns per tile and the grid as a multiple of the resident block count. */
int pdot_batch_set_launch(pdot_batch* b, int32_t cols_per_tile, int32_t grid_mult);
int pdot_batch_get_launch(pdot_batch* b, int32_t* rows_per_tile, int32_t* cols_per_tile, int32_t* tiles,
                          int32_t* step_grid);
int64_t pdot_batch_device_bytes(const pdot_batch* b);

/* ------------------------------------------------------------------------
 * Learned warm start: a point-wise network with mean-pooled set context that
 * predicts dual potentials (u, v) from point coordinates and masses.  No
 * counterpart in the reference; it only chooses the starting iterate handed
 * to pdot_set_start / pdot_batch_set_warm_start.  All matrix products run on
 * bf16 MFMA with FP32 accumulation.  Features are N x 4 row-major FP32 rows
 * (x, y, scaled mass, side), the nsrc source rows first; the parameter
 * vector layout is documented in pdot_mlp.hip.
 * ------------------------------------------------------------------------ */
typedef struct pdot_mlp pdot_mlp;

int pdot_mlp_create(pdot_mlp** out, int32_t device);
void pdot_mlp_destroy(pdot_mlp* m);
const char* pdot_mlp_last_error(const pdot_mlp* m);
int64_t pdot_mlp_num_params(const pdot_mlp* m);
int pdot_mlp_set_params(pdot_mlp* m, const float* params);  /* also resets the Adam moments */
int pdot_mlp_get_params(pdot_mlp* m, float* params);
int pdot_mlp_forward(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, float* out);
/* Mean squared error against target (N) and its gradient (num_params, may
 * be NULL); the gradient stays on the device for pdot_mlp_adam_step. */
int pdot_mlp_loss_grad(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, const float* target,
                       float* loss, float* grad);
int pdot_mlp_adam_step(pdot_mlp* m, float lr, float beta1, float beta2, float eps);
/* The weight head: the natural log of a per-instance initial primal weight
 * (the reference's primal_importance, src/solver_gpu.jl:486-496) predicted
 * from the mean-pooled context of the first layer and the instance size.
 * nsrc, ntgt >= 1. */
int pdot_mlp_predict_weight(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, float* log_weight);
/* Squared error of the weight head against target (a log weight) and its
 * gradient (num_params, may be NULL; zero outside the head, and outside the
 * head and the first layer when train_trunk), kept on the device for
 * pdot_mlp_adam_step; log_weight (may be NULL) receives the prediction. */
int pdot_mlp_weight_loss_grad(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, float target,
                              int32_t train_trunk, float* log_weight, float* loss, float* grad);
/* Test hook for the MFMA matrix product: C (M x N) = op(A) op(B), operands
 * rounded to bf16, FP32 accumulation; row-major, op = transpose if set. */
int pdot_mlp_gemm_probe(pdot_mlp* m, const float* A, const float* B, float* C, int64_t M, int64_t N, int64_t K,
                        int32_t trans_a, int32_t trans_b);

#ifdef __cplusplus
}
#endif

#endif /* PDOT_H */
