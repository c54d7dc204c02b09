/*
 * pdot_oracle.c -- CPU restatement of PDOT.jl's restarted adaptive PDHG for
 * discrete optimal transport, in FP64 on host arrays (OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library and
 * the CPU baseline that bench.py times beside it.  Product code never links,
 * loads or calls it; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do, and only as the thing results are compared against or
 * timed beside.
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the PDOT.jl source tree).  Two modes share one code path:
 *
 *   ORACLE_LITERAL (0)  stores the primal gradient G, its running average and
 *                       the last-restart copies exactly like the reference
 *                       (8 n*m arrays) and replays a rejected step by
 *                       recomputing the pass (solver_gpu.jl:299-365).
 *   ORACLE_FUSED   (1)  recomputes G = (C - u) - v^T and its average from
 *                       (C, u_avg, v_avg) instead of storing them, replays a
 *                       rejected step as a scalar recurrence only, and keeps
 *                       the last restart point as cached scalars.  This is the
 *                       arithmetic the fused GPU kernels perform.
 *
 * Reductions are deterministic: sums over a matrix are formed per column in
 * row order and then summed over columns in order, independent of the thread
 * count.  Floating-point expressions keep the reference's evaluation order and
 * are compiled with -ffp-contract=off so that no multiply-add is fused.
 *
 * Parity pins (see tests/): closed-form step-1/step-2 known answers, LP optima
 * from scipy HiGHS and linear_sum_assignment committed as fixtures, separable
 * and 1-D closed forms, and literal-vs-fused agreement.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* OpenMP only pays off above this many matrix elements */
#define PAR_MIN 131072


#define ORACLE_LITERAL 0
#define ORACLE_FUSED 1

/* enums, values as declared in the reference */
/* RestartChoice solve_log.jl:15-20 */
enum { RESTART_CHOICE_UNSPECIFIED = 0, RESTART_CHOICE_NO_RESTART = 1,
       RESTART_CHOICE_WEIGHTED_AVERAGE_RESET = 2, RESTART_CHOICE_RESTART_TO_AVERAGE = 3 };
/* PointType solve_log.jl:35-41 */
enum { POINT_TYPE_UNSPECIFIED = 0, POINT_TYPE_CURRENT_ITERATE = 1,
       POINT_TYPE_ITERATE_DIFFERENCE = 2, POINT_TYPE_AVERAGE_ITERATE = 3, POINT_TYPE_NONE = 4 };
/* TerminationReason solve_log.jl:245-254; 0 doubles as "do not terminate" */
enum { TERM_NONE = -1, TERMINATION_REASON_UNSPECIFIED = 0, TERMINATION_REASON_OPTIMAL = 1,
       TERMINATION_REASON_TIME_LIMIT = 2, TERMINATION_REASON_ITERATION_LIMIT = 3,
       TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT = 4, TERMINATION_REASON_NUMERICAL_ERROR = 5,
       TERMINATION_REASON_INVALID_PROBLEM = 6, TERMINATION_REASON_OTHER = 7 };
/* RestartScheme saddle_point_gpu.jl:401, RestartToCurrentMetric :408, OptimalityNorm termination.jl:2 */
enum { NO_RESTARTS = 0, FIXED_FREQUENCY = 1, ADAPTIVE_KKT = 2 };
enum { NO_RESTART_TO_CURRENT = 0, KKT_GREEDY = 1 };
enum { L_INF = 0, L_2 = 1 };

/* PdhgParameters solver_gpu.jl:9-21 + TerminationCriteria termination.jl:5-51 +
 * RestartParameters saddle_point_gpu.jl:411-441 + step-size policy :2-7, flattened. */
typedef struct {
    double primal_importance;
    int32_t scale_invariant_initial_primal_weight;
    int32_t verbosity;
    int32_t record_iteration_stats;
    int32_t termination_evaluation_frequency;
    int32_t optimality_norm;
    int32_t iteration_limit;
    double eps_optimal_absolute;
    double eps_optimal_relative;
    double time_sec_limit;
    double kkt_matrix_pass_limit;
    int32_t restart_scheme;
    int32_t restart_to_current_metric;
    int64_t restart_frequency_if_fixed;
    double artificial_restart_threshold;
    double sufficient_reduction_for_restart;
    double necessary_reduction_for_restart;
    double primal_weight_update_smoothing;
    int32_t step_policy; /* 0 AdaptiveStepsizeParams, 1 ConstantStepsizeParams */
    int32_t pad0;
    double reduction_exponent;
    double growth_exponent;
} oracle_params;

/* IterationStats solve_log.jl:154-214 with its single ConvergenceInformation
 * (solve_log.jl:47-147) flattened in. */
typedef struct {
    int32_t iteration_number;
    int32_t restart_used;
    int32_t candidate_type;
    int32_t cumulative_rejected_steps;
    double cumulative_kkt_matrix_passes;
    double cumulative_time_sec;
    double step_size;
    double primal_weight;
    double time_spent_doing_basic_algorithm;
    double primal_objective;
    double dual_objective;
    double corrected_dual_objective;
    double l_inf_primal_residual;
    double l2_primal_residual;
    double l_inf_dual_residual;
    double l2_dual_residual;
    double relative_l_inf_primal_residual;
    double relative_l2_primal_residual;
    double relative_l_inf_dual_residual;
    double relative_l2_dual_residual;
    double relative_optimality_gap;
    double l_inf_primal_variable;
    double l2_primal_variable;
    double l_inf_dual_variable;
    double l2_dual_variable;
} oracle_stats;

/* ---------------------------------------------------------------- helpers */

static double wall_time(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Julia's max(x, 0.0) / min(x, y) propagate NaN (C's fmax/fmin do not). */
static inline double jl_max0(double x) { return x != x ? x : (x > 0.0 ? x : 0.0); }
static inline double jl_min0(double x) { return x != x ? x : (x < 0.0 ? x : 0.0); }
static inline double jl_min(double x, double y) {
    if (x != x || y != y) return NAN;
    return x < y ? x : y;
}

/* (k+1)^e as Julia evaluates Int^Float64: an integral exponent goes through
 * integer powering, exact for the integer bases < 2^26 that occur here. */
static double jl_pow_int_base(int64_t base, double e) {
    double x = (double)base;
    if (e == floor(e) && fabs(e) <= 64.0) {
        int64_t k = (int64_t)fabs(e);
        double r = 1.0;
        for (int64_t t = 0; t < k; ++t) r *= x;
        return e < 0 ? 1.0 / r : r;
    }
    return pow(x, e);
}

typedef struct {
    int64_t n, m;
    const double *C, *a, *b;
} prob_t;

/* G_ij of a point: stored matrix, or zero (initial state, solver_gpu.jl:434),
 * or recomputed (C_ij - u_i) - v_j in the order of solver_gpu.jl:249-250. */
typedef struct {
    const double* G; /* stored gradient or NULL */
    int gzero;       /* G is the initial zero matrix */
    const double *u, *v;
} grad_t;

static inline double grad_at(const prob_t* P, const grad_t* g, int64_t i, int64_t j) {
    if (g->G) return g->G[i + j * P->n];
    if (g->gzero) return 0.0;
    return (P->C[i + j * P->n] - g->u[i]) - g->v[j];
}

/* deterministic sum over the columns of per-column partials */
static double sum_cols(const double* t, int64_t m) {
    double s = 0.0;
    for (int64_t j = 0; j < m; ++j) s += t[j];
    return s;
}
static double max_cols(const double* t, int64_t m) {
    double s = 0.0;
    for (int64_t j = 0; j < m; ++j) {
        double x = t[j];
        if (x != x) return x;
        if (x > s) s = x;
    }
    return s;
}
static double vec_dot(const double* x, const double* y, int64_t k) {
    double s = 0.0;
    for (int64_t i = 0; i < k; ++i) s += x[i] * y[i];
    return s;
}
static double vec_sumsq(const double* x, int64_t k) {
    double s = 0.0;
    for (int64_t i = 0; i < k; ++i) s += x[i] * x[i];
    return s;
}
static double vec_maxabs(const double* x, int64_t k) {
    double s = 0.0;
    for (int64_t i = 0; i < k; ++i) {
        double t = fabs(x[i]);
        if (t != t) return t;
        if (t > s) s = t;
    }
    return s;
}

/* row sums: out[i] = sum_j M[i,j] + add[i] (add may be NULL); row order over j */
static void row_sums(int64_t n, int64_t m, const double* M, const double* add, double* out) {
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t i0 = 0; i0 < n; i0 += 256) {
        int64_t i1 = i0 + 256 < n ? i0 + 256 : n;
        double acc[256];
        for (int64_t i = i0; i < i1; ++i) acc[i - i0] = 0.0;
        for (int64_t j = 0; j < m; ++j) {
            const double* col = M + j * n;
            for (int64_t i = i0; i < i1; ++i) acc[i - i0] += col[i];
        }
        for (int64_t i = i0; i < i1; ++i) out[i] = add ? acc[i - i0] + add[i] : acc[i - i0];
    }
}
/* column sums: out[j] = sum_i M[i,j] + add[j] */
static void col_sums(int64_t n, int64_t m, const double* M, const double* add, double* out) {
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j) {
        const double* col = M + j * n;
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += col[i];
        out[j] = add ? s + add[j] : s;
    }
}

/* ------------------------------------------------------- PDHG step pieces */

/* compute_next_primal_solution_kernel! + compute_next_primal_solution!
 * solver_gpu.jl:125-174: D = max(X - (eta/omega) G, 0) - X, row/col sums of D
 * added to the tracked sums. */
static void primal_pass(const prob_t* P, const double* X, const grad_t* g, double eta, double omega,
                        double* D, const double* r, const double* c, double* rn, double* cn) {
    const int64_t n = P->n, m = P->m;
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j) {
        for (int64_t i = 0; i < n; ++i) {
            int64_t idx = i + j * n;
            double x = X[idx];
            double d = x - (eta / omega) * grad_at(P, g, i, j);
            d = jl_max0(d);
            d -= x;
            D[idx] = d;
        }
    }
    row_sums(n, m, D, r, rn);
    col_sums(n, m, D, c, cn);
}

/* compute_next_dual_solution_kernel! solver_gpu.jl:177-195:
 * y+ = y + eta*omega*(q - 2 s+ + s) */
static void dual_pass(int64_t k, const double* q, const double* y, const double* s, const double* sn,
                      double eta, double omega, double* yn) {
    for (int64_t i = 0; i < k; ++i) yn[i] = y[i] + eta * omega * (q[i] - 2 * sn[i] + s[i]);
}

/* compute_interaction_and_movement solver_gpu.jl:267-288 */
static void interaction_movement(const prob_t* P, const double* D, const double* r, const double* rn,
                                 const double* c, const double* cn, const double* u, const double* un,
                                 const double* v, const double* vn, double omega, double* colbuf,
                                 double* interaction, double* movement) {
    const int64_t n = P->n, m = P->m;
    double dot_r = 0.0, dot_c = 0.0, yy = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double dr = rn[i] - r[i], du = un[i] - u[i];
        dot_r += dr * du;
        yy += du * du;
    }
    for (int64_t j = 0; j < m; ++j) {
        double dc = cn[j] - c[j], dv = vn[j] - v[j];
        dot_c += dc * dv;
        yy += dv * dv;
    }
    double pdi = dot_r;
    pdi += dot_c;
    *interaction = fabs(pdi);
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j) {
        const double* col = D + j * n;
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += col[i] * col[i];
        colbuf[j] = s;
    }
    double nd = sqrt(sum_cols(colbuf, m));
    double ny = sqrt(yy);
    *movement = 0.5 * omega * (nd * nd) + (0.5 / omega) * (ny * ny);
}

/* ---------------------------------------------------- weighted average */

/* CuSolutionWeightedAverage saddle_point_gpu.jl:52-65 */
typedef struct {
    double *X, *u, *v, *r, *c, *G; /* G only in literal mode */
    int64_t primal_count, dual_source_count, dual_target_count;
    double sum_primal, sum_dual_source, sum_dual_target;
} wavg_t;

/* new/avg weights of saddle_point_gpu.jl:122-127 */
static inline void avg_weights(double S, double w, double* nw, double* aw) {
    if (S > 0.0) {
        *aw = S / (S + w);
        *nw = w / S;
    } else {
        *nw = 1.0;
        *aw = 1.0;
    }
}
/* avg .+= x * nw; avg *= aw  (saddle_point_gpu.jl:129-130 and siblings) */
static void avg_vec(double* avg, const double* x, int64_t k, double nw, double aw) {
    for (int64_t i = 0; i < k; ++i) {
        double t = avg[i] + x[i] * nw;
        avg[i] = t * aw;
    }
}
static void avg_mat(double* avg, const double* x, int64_t nm, double nw, double aw) {
#pragma omp parallel for schedule(static) if(nm >= PAR_MIN)
    for (int64_t i = 0; i < nm; ++i) {
        double t = avg[i] + x[i] * nw;
        avg[i] = t * aw;
    }
}

/* add_to_solution_weighted_average! saddle_point_gpu.jl:244-289 and the six
 * add_to_* helpers :115-241; every helper uses the pre-update weight sums. */
static void add_to_average(const prob_t* P, wavg_t* A, const double* X, const double* u, const double* v,
                           double w, const double* r, const double* c, const double* G) {
    const int64_t n = P->n, m = P->m;
    double nw, aw;
    avg_weights(A->sum_primal, w, &nw, &aw);
    avg_vec(A->r, r, n, nw, aw);                  /* :185-202 */
    avg_vec(A->c, c, m, nw, aw);                  /* :204-221 */
    if (A->G) {                                   /* :223-241 */
        avg_weights(A->sum_dual_target, w, &nw, &aw);
        avg_mat(A->G, G, n * m, nw, aw);
    }
    avg_weights(A->sum_primal, w, &nw, &aw);      /* :115-135 */
    avg_mat(A->X, X, n * m, nw, aw);
    A->primal_count += 1;
    A->sum_primal += w;
    avg_weights(A->sum_dual_source, w, &nw, &aw); /* :140-160 */
    avg_vec(A->u, u, n, nw, aw);
    A->dual_source_count += 1;
    A->sum_dual_source += w;
    avg_weights(A->sum_dual_target, w, &nw, &aw); /* :162-182 */
    avg_vec(A->v, v, m, nw, aw);
    A->dual_target_count += 1;
    A->sum_dual_target += w;
}

/* reset_solution_weighted_average! saddle_point_gpu.jl:93-110 */
static void reset_average(const prob_t* P, wavg_t* A) {
    const int64_t n = P->n, m = P->m;
    memset(A->X, 0, sizeof(double) * n * m);
    memset(A->u, 0, sizeof(double) * n);
    memset(A->v, 0, sizeof(double) * m);
    A->primal_count = A->dual_source_count = A->dual_target_count = 0;
    A->sum_primal = A->sum_dual_source = A->sum_dual_target = 0.0;
    memset(A->r, 0, sizeof(double) * n);
    memset(A->c, 0, sizeof(double) * m);
    if (A->G) memset(A->G, 0, sizeof(double) * n * m);
}

/* ------------------------------------------------------------ evaluation */

/* CachedOTInfo termination.jl:87-101 */
typedef struct {
    double l_inf_norm_primal_linear_objective, l_inf_norm_primal_right_hand_side;
    double l2_norm_primal_linear_objective, l2_norm_primal_right_hand_side;
} ot_cache_t;

void oracle_cached_ot_info(int64_t n, int64_t m, const double* C, const double* a, const double* b,
                           double* out4) {
    double* colbuf = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* colmax = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j) {
        double s = 0.0, mx = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double x = C[i + j * n];
            s += x * x;
            double t = fabs(x);
            if (t != t || t > mx) mx = t;
        }
        colbuf[j] = s;
        colmax[j] = mx;
    }
    double ab_inf = vec_maxabs(a, n), t = vec_maxabs(b, m);
    if (t > ab_inf || t != t) ab_inf = t;
    out4[0] = max_cols(colmax, m);                          /* norm(C, Inf) */
    out4[1] = ab_inf;                                       /* norm([a;b], Inf) */
    out4[2] = sqrt(sum_cols(colbuf, m));                    /* norm(C, 2): Frobenius */
    out4[3] = sqrt(vec_sumsq(a, n) + vec_sumsq(b, m));      /* norm([a;b], 2) */
    free(colbuf);
    free(colmax);
}

/* The n*m reductions of one point: <C,X>, ||X||_inf, ||X||_2^2, and of
 * min(G,0): ||.||_inf, ||.||_2^2 (utils_gpu.jl:84, 102-103, 39, 52, 108). */
typedef struct {
    double pobj, x_inf, x_sq, rc_inf, rc_sq;
} matstats_t;

static void point_matstats(const prob_t* P, const double* X, const grad_t* g, double* scratch, matstats_t* o) {
    const int64_t n = P->n, m = P->m;
    double* t_pobj = scratch;
    double* t_xinf = scratch + m;
    double* t_xsq = scratch + 2 * m;
    double* t_rinf = scratch + 3 * m;
    double* t_rsq = scratch + 4 * m;
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j) {
        double sp = 0.0, xi = 0.0, xs = 0.0, ri = 0.0, rs = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            int64_t idx = i + j * n;
            double x = X[idx];
            sp += P->C[idx] * x;
            double ax = fabs(x);
            if (ax != ax || ax > xi) xi = ax;
            xs += x * x;
            double rc = jl_min0(grad_at(P, g, i, j));
            double ar = fabs(rc);
            if (ar != ar || ar > ri) ri = ar;
            rs += rc * rc;
        }
        t_pobj[j] = sp;
        t_xinf[j] = xi;
        t_xsq[j] = xs;
        t_rinf[j] = ri;
        t_rsq[j] = rs;
    }
    o->pobj = sum_cols(t_pobj, m);
    o->x_inf = max_cols(t_xinf, m);
    o->x_sq = sum_cols(t_xsq, m);
    o->rc_inf = max_cols(t_rinf, m);
    o->rc_sq = sum_cols(t_rsq, m);
}

/* The vector reductions of one point (utils_gpu.jl:26-27, 54, 94-95, 115-116). */
typedef struct {
    double p_inf, p_sq, dobj, u_inf, u_sq, y_sq;
} vecstats_t;

static void point_vecstats(const prob_t* P, const double* u, const double* v, const double* r,
                           const double* c, vecstats_t* o) {
    const int64_t n = P->n, m = P->m;
    double pinf = 0.0, psq = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double t = P->a[i] - r[i];
        double at = fabs(t);
        if (at != at || at > pinf) pinf = at;
        psq += t * t;
    }
    for (int64_t j = 0; j < m; ++j) {
        double t = P->b[j] - c[j];
        double at = fabs(t);
        if (at != at || at > pinf) pinf = at;
        psq += t * t;
    }
    o->p_inf = pinf;
    o->p_sq = psq;
    o->dobj = vec_dot(P->a, u, n) + vec_dot(P->b, v, m);
    o->u_inf = vec_maxabs(u, n);
    o->u_sq = vec_sumsq(u, n);
    o->y_sq = o->u_sq + vec_sumsq(v, m);
}

/* compute_convergence_information utils_gpu.jl:68-129 from the reductions */
static void convergence_info(const ot_cache_t* oc, double eps_ratio, const matstats_t* ms,
                             const vecstats_t* vs, oracle_stats* s) {
    s->primal_objective = ms->pobj;
    s->l_inf_primal_residual = vs->p_inf;
    s->l2_primal_residual = sqrt(vs->p_sq);
    s->relative_l_inf_primal_residual = s->l_inf_primal_residual / (eps_ratio + oc->l_inf_norm_primal_right_hand_side);
    s->relative_l2_primal_residual = s->l2_primal_residual / (eps_ratio + oc->l2_norm_primal_right_hand_side);
    s->l_inf_primal_variable = ms->x_inf;
    s->l2_primal_variable = sqrt(ms->x_sq);
    s->dual_objective = vs->dobj;
    s->l_inf_dual_residual = ms->rc_inf;
    s->l2_dual_residual = sqrt(ms->rc_sq);
    s->relative_l_inf_dual_residual = s->l_inf_dual_residual / (eps_ratio + oc->l_inf_norm_primal_linear_objective);
    s->relative_l2_dual_residual = s->l2_dual_residual / (eps_ratio + oc->l2_norm_primal_linear_objective);
    /* [u; u], not [u; v] (utils_gpu.jl:115-116) */
    s->l_inf_dual_variable = vs->u_inf;
    s->l2_dual_variable = sqrt(vs->u_sq + vs->u_sq);
    s->corrected_dual_objective = (ms->rc_inf == 0.0) ? vs->dobj : -INFINITY; /* :57-63 */
    double gap = fabs(s->primal_objective - s->dual_objective);
    double abs_obj = fabs(s->primal_objective) + fabs(s->dual_objective);
    s->relative_optimality_gap = gap / (eps_ratio + abs_obj);
    s->candidate_type = POINT_TYPE_AVERAGE_ITERATE; /* solver_gpu.jl:551, 571 */
}

/* compute_weight_kkt_residual saddle_point_gpu.jl:299-330, from the point's
 * L2 primal residual p2, L2 dual residual d2 and objectives. */
double oracle_weighted_kkt(double omega, double p2, double d2, double pobj, double dobj) {
    double g = fabs(pobj - dobj);
    return sqrt(omega * (p2 * p2) + 1 / omega * (d2 * d2) + g * g);
}

/* ------------------------------------------------- restart decisions */

/* One record per evaluation that reaches run_restart_scheme (every
 * non-terminating evaluation): what the restart scheme read and decided. */
typedef struct {
    int32_t iteration;        /* the evaluation iteration */
    int32_t restart_used;     /* RestartChoice */
    int32_t artificial;       /* restart_length >= threshold * iterations_completed (:561-564) */
    int32_t adaptive_checked; /* should_do_adaptive_restart_kkt ran */
    int32_t reset_to_average; /* should_reset_to_average (:477-487) */
    int32_t pad;
    int64_t avg_count;        /* restart_length (:558) */
    int64_t iterations_completed;
    double step_size, primal_weight_before, primal_weight_after;
    double kkt_current, kkt_average, kkt_last; /* weighted KKT at primal_weight_before (NaN: not formed) */
    double ratio;                              /* candidate / last (NaN unless checked) */
    double last_trial_before, last_trial_after;
    double last_x_norm, last_y_norm;           /* norms of the previous restart point (restarts only) */
    double primal_distance, dual_distance;     /* update_last_restart_info! (restarts only) */
    /* (p2, d2, pobj, dobj, ||X||, ||[u;v]||) of the current point, the average and the last restart point */
    double pieces[3][6];
} oracle_restart_rec;

/* should_do_adaptive_restart_kkt saddle_point_gpu.jl:492-533, given the
 * weighted KKT of the candidate and of the last restart point, both at the
 * current primal weight.  The trial's ratio is stored in *last_trial whatever
 * the outcome (:530). */
int oracle_should_do_adaptive_restart_kkt(double candidate_kkt, double last_kkt, double sufficient,
                                          double necessary, double* last_trial) {
    int do_restart = 0;
    double ratio = candidate_kkt / last_kkt;
    if (ratio < necessary) {
        if (ratio < sufficient) do_restart = 1;
        else if (ratio > *last_trial) do_restart = 1;
    }
    *last_trial = ratio;
    return do_restart;
}

/* The decision of run_restart_scheme (saddle_point_gpu.jl:539-633) from the
 * weighted KKT residuals at the current primal weight: kc of the current
 * point, ka of the average, kl of the last restart point (kc and ka are not
 * read under NO_RESTARTS, kl only where the adaptive test runs).  Returns the
 * RestartChoice; *last_trial is the CuRestartInfo field; rec (may be NULL)
 * receives the flags and the ratio. */
int oracle_restart_decision(const oracle_params* p, int64_t avg_count, int64_t iterations_completed, double kc,
                            double ka, double kl, double* last_trial, oracle_restart_rec* rec) {
    if (rec) {
        rec->artificial = rec->adaptive_checked = rec->reset_to_average = 0;
        rec->ratio = NAN;
        rec->last_trial_before = *last_trial;
    }
    if (!(avg_count > 0)) { /* :558-562 */
        if (rec) rec->last_trial_after = *last_trial;
        return RESTART_CHOICE_NO_RESTART;
    }
    int do_restart = 0, reset_to_average = 0;
    double cand = NAN;
    if ((double)avg_count >= p->artificial_restart_threshold * (double)iterations_completed) { /* :568-571 */
        do_restart = 1;
        if (rec) rec->artificial = 1;
    }
    if (p->restart_scheme != NO_RESTARTS) { /* :573-615 */
        reset_to_average = (p->restart_to_current_metric == KKT_GREEDY) ? (kc >= ka) : 1;
        cand = reset_to_average ? ka : kc;
    }
    if (!do_restart) { /* :617-633 */
        if (p->restart_scheme == ADAPTIVE_KKT) {
            do_restart = oracle_should_do_adaptive_restart_kkt(cand, kl, p->sufficient_reduction_for_restart,
                                                               p->necessary_reduction_for_restart, last_trial);
            if (rec) {
                rec->adaptive_checked = 1;
                rec->ratio = *last_trial;
            }
        } else if (p->restart_scheme == FIXED_FREQUENCY && p->restart_frequency_if_fixed <= avg_count) {
            do_restart = 1;
        }
    }
    if (rec) {
        rec->reset_to_average = reset_to_average;
        rec->last_trial_after = *last_trial;
    }
    if (!do_restart) return RESTART_CHOICE_NO_RESTART;
    return reset_to_average ? RESTART_CHOICE_RESTART_TO_AVERAGE : RESTART_CHOICE_WEIGHTED_AVERAGE_RESET;
}

/* update_last_restart_info! saddle_point_gpu.jl:753-767 after the reset of
 * the average (:667): the distances moved are the weighted norms
 * (primal_weighted_norm / dual_weighted_norm, :36-50) of 0 minus the previous
 * restart point, whose norms are last_x_norm = ||X|| and last_y_norm =
 * ||[u; v]||. */
void oracle_restart_distances(double last_x_norm, double last_y_norm, double primal_norm_params,
                              double dual_norm_params, double primal_weight, double* primal_distance,
                              double* dual_distance) {
    *primal_distance = sqrt(primal_norm_params) * last_x_norm / sqrt(primal_weight);
    *dual_distance = sqrt(dual_norm_params) * last_y_norm * sqrt(primal_weight);
}

/* compute_new_primal_weight saddle_point_gpu.jl:701-728 */
double oracle_new_primal_weight(double primal_distance, double dual_distance, double primal_weight,
                                double smoothing) {
    const double eps = 2.220446049250313e-16; /* eps(Float64) */
    if (primal_distance > eps && dual_distance > eps) {
        double est = dual_distance / primal_distance;
        double lw = smoothing * log(est) + (1 - smoothing) * log(primal_weight);
        return exp(lw);
    }
    return primal_weight;
}

/* optimality_criteria_met termination.jl:106-134 */
static int optimality_met(int norm, double abs_tol, double rel_tol, const oracle_stats* ci,
                          const ot_cache_t* oc) {
    double abs_obj = fabs(ci->primal_objective) + fabs(ci->dual_objective);
    double gap = fabs(ci->primal_objective - ci->dual_objective);
    double pe, pb, de, db;
    if (norm == L_INF) {
        pe = ci->l_inf_primal_residual;
        pb = oc->l_inf_norm_primal_right_hand_side;
        de = ci->l_inf_dual_residual;
        db = oc->l_inf_norm_primal_linear_objective;
    } else {
        pe = ci->l2_primal_residual;
        pb = oc->l2_norm_primal_right_hand_side;
        de = ci->l2_dual_residual;
        db = oc->l2_norm_primal_linear_objective;
    }
    return de < abs_tol + rel_tol * db && pe < abs_tol + rel_tol * pb && gap < abs_tol + rel_tol * abs_obj;
}

/* check_termination_criteria termination.jl:140-166 */
static int check_termination(const oracle_params* p, const ot_cache_t* oc, const oracle_stats* s) {
    if (optimality_met(p->optimality_norm, p->eps_optimal_absolute, p->eps_optimal_relative, s, oc))
        return TERMINATION_REASON_OPTIMAL;
    if (s->iteration_number >= p->iteration_limit) return TERMINATION_REASON_ITERATION_LIMIT;
    if (s->cumulative_kkt_matrix_passes >= p->kkt_matrix_pass_limit) return TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT;
    if (s->cumulative_time_sec >= p->time_sec_limit) return TERMINATION_REASON_TIME_LIMIT;
    return TERM_NONE;
}

/* ------------------------------------------------------------- optimize */

/* Last restart point, CuRestartInfo saddle_point_gpu.jl:332-393.  Literal mode
 * stores the copies; fused mode keeps ||X||, ||[u;v]|| and the KKT pieces. */
typedef struct {
    double *X, *u, *v, *r, *c, *G;
    double x_norm, y_norm, p2, d2, pobj, dobj;
    int64_t last_restart_length;
    double primal_distance, dual_distance, kkt_reduction_ratio_last_trial;
} last_t;

typedef struct {
    double* X;
    double *u, *v, *r, *c;
    double* G; /* literal only */
    int gzero;
} point_t;

static grad_t point_grad(const point_t* pt) {
    grad_t g = {pt->G, pt->gzero, pt->u, pt->v};
    return g;
}

/* optimize (solver_gpu.jl:413-720).  stats receives the recorded
 * IterationStats, eta_trace the step size after every take_step! call and
 * rtrace (may be NULL) one oracle_restart_rec per non-terminating evaluation. */
int oracle_optimize_ex(const oracle_params* p, int64_t n, int64_t m, const double* C, const double* a,
                       const double* b, int32_t mode, const double* u0, const double* v0, const double* X0,
                       double* X_out, double* u_out, double* v_out, int32_t* termination_reason,
                       int32_t* iteration_count, oracle_stats* stats, int32_t stats_cap, int32_t* n_stats,
                       double* eta_trace, int64_t trace_cap, int64_t* trace_len, oracle_restart_rec* rtrace,
                       int32_t rtrace_cap, int32_t* n_rtrace) {
    const int literal = (mode == ORACLE_LITERAL);
    if (n <= 0 || m <= 0 || p->termination_evaluation_frequency <= 0) return -4;
    /* solver_gpu.jl:420-422 */
    if (p->primal_importance <= 0 || !isfinite(p->primal_importance)) return -1;
    /* construct_restart_parameters asserts saddle_point_gpu.jl:455-461 */
    if (!(p->restart_frequency_if_fixed > 1) || !(0.0 < p->artificial_restart_threshold && p->artificial_restart_threshold <= 1.0) ||
        !(0.0 < p->sufficient_reduction_for_restart && p->sufficient_reduction_for_restart <= p->necessary_reduction_for_restart &&
          p->necessary_reduction_for_restart <= 1.0) ||
        !(0.0 <= p->primal_weight_update_smoothing && p->primal_weight_update_smoothing <= 1.0))
        return -3;

    prob_t P = {n, m, C, a, b};
    const int64_t nm = n * m;
    double oc4[4];
    oracle_cached_ot_info(n, m, C, a, b, oc4); /* termination.jl:94-101, before start_time */
    ot_cache_t oc = {oc4[0], oc4[1], oc4[2], oc4[3]};

    /* allocations, solver_gpu.jl:428-472 */
    int nmat = literal ? 8 : 3;
    double* mats = (double*)calloc((size_t)nmat * (size_t)nm, sizeof(double));
    double* vecs = (double*)calloc((size_t)(12 * (n + m)), sizeof(double));
    double* scratch = (double*)calloc((size_t)(6 * m + 16), sizeof(double));
    if (!mats || !vecs || !scratch) {
        free(mats);
        free(vecs);
        free(scratch);
        return -2;
    }
    point_t cur;
    wavg_t avg;
    last_t last;
    memset(&avg, 0, sizeof(avg));
    memset(&last, 0, sizeof(last));
    cur.X = mats;
    avg.X = mats + nm;
    double* D = mats + 2 * nm;
    cur.G = literal ? mats + 3 * nm : NULL;
    avg.G = literal ? mats + 4 * nm : NULL;
    last.X = literal ? mats + 5 * nm : NULL;
    last.G = literal ? mats + 6 * nm : NULL;
    double* vp = vecs;
    cur.u = vp; vp += n; cur.v = vp; vp += m; cur.r = vp; vp += n; cur.c = vp; vp += m;
    avg.u = vp; vp += n; avg.v = vp; vp += m; avg.r = vp; vp += n; avg.c = vp; vp += m;
    double* un = vp; vp += n; double* vn = vp; vp += m; double* rn = vp; vp += n; double* cn = vp; vp += m;
    last.u = vp; vp += n; last.v = vp; vp += m; last.r = vp; vp += n; last.c = vp; vp += m;

    /* initial state: X = 0, duals 0 (or a warm start), G = 0 (solver_gpu.jl:429-434);
     * a warm start defines G0 = (C - u0) - v0 and may also set X0 (then r, c
     * are its row and column sums). */
    if (X0) {
        memcpy(cur.X, X0, sizeof(double) * (size_t)nm);
        row_sums(n, m, cur.X, NULL, cur.r);
        col_sums(n, m, cur.X, NULL, cur.c);
    }
    cur.gzero = 1;
    if (u0 && v0) {
        memcpy(cur.u, u0, sizeof(double) * n);
        memcpy(cur.v, v0, sizeof(double) * m);
        cur.gzero = 0;
        if (literal) {
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
            for (int64_t j = 0; j < m; ++j)
                for (int64_t i = 0; i < n; ++i) cur.G[i + j * n] = (C[i + j * n] - u0[i]) - v0[j];
        }
    }

    double step_size;
    if (p->step_policy == 0) step_size = 1.0; /* solver_gpu.jl:476-477 */
    else step_size = (1 - 0.2) / sqrt((double)(n + m)); /* :479-481 */
    const double KKT_PASSES_PER_TERMINATION_EVALUATION = 2.0;
    double primal_weight;
    if (p->scale_invariant_initial_primal_weight) { /* select_initial_primal_weight saddle_point_gpu.jl:848-866 */
        double rhs = sqrt(1.0) * oc.l2_norm_primal_right_hand_side;
        double obj = sqrt(1.0) * oc.l2_norm_primal_linear_objective;
        primal_weight = (obj > 0.0 && rhs > 0.0) ? p->primal_importance * (obj / rhs) : p->primal_importance;
    } else {
        primal_weight = p->primal_importance;
    }
    const double smoothing = p->primal_weight_update_smoothing;
    int numerical_error = 0;
    double cumulative_kkt_passes = 0.0;
    int64_t total_number_iterations = 0;

    int32_t nst = 0;
    int32_t nrt = 0;
    int64_t ntrace = 0;
    double start_time = wall_time();
    double time_basic = 0.0;

    /* create_last_restart_info saddle_point_gpu.jl:371-393: copies of the start */
    last.last_restart_length = 1;
    last.primal_distance = 0.0;
    last.dual_distance = 0.0;
    last.kkt_reduction_ratio_last_trial = 1.0;
    memcpy(last.u, cur.u, sizeof(double) * n);
    memcpy(last.v, cur.v, sizeof(double) * m);
    if (literal) {
        memset(last.X, 0, sizeof(double) * nm);
        memcpy(last.G, cur.G, sizeof(double) * nm);
    } else {
        /* scalar cache of the start point's KKT pieces */
        matstats_t ms;
        vecstats_t vs;
        grad_t g = point_grad(&cur);
        point_matstats(&P, cur.X, &g, scratch, &ms);
        point_vecstats(&P, cur.u, cur.v, cur.r, cur.c, &vs);
        last.x_norm = 0.0;
        last.y_norm = sqrt(vs.y_sq);
        last.p2 = sqrt(vs.p_sq);
        last.d2 = sqrt(ms.rc_sq);
        last.pobj = ms.pobj;
        last.dobj = vs.dobj;
    }

    const int64_t iteration_limit = p->iteration_limit;
    const int64_t freq = p->termination_evaluation_frequency;
    int64_t iteration = 0;
    int rc = 0;
    for (;;) {
        iteration += 1;
        if ((iteration - 1) % freq == 0 || iteration == iteration_limit + 1 || iteration <= 10 || numerical_error) {
            cumulative_kkt_passes += KKT_PASSES_PER_TERMINATION_EVALUATION;
            /* solver_gpu.jl:537: current point on numerical error or an empty average (Q6) */
            int use_current = numerical_error || avg.primal_count == 0;
            oracle_stats st;
            memset(&st, 0, sizeof(st));
            matstats_t ms_eval;
            vecstats_t vs_eval;
            if (use_current) {
                grad_t g = point_grad(&cur);
                point_matstats(&P, cur.X, &g, scratch, &ms_eval);
                point_vecstats(&P, cur.u, cur.v, cur.r, cur.c, &vs_eval);
            } else {
                grad_t g = {avg.G, 0, avg.u, avg.v};
                point_matstats(&P, avg.X, &g, scratch, &ms_eval);
                point_vecstats(&P, avg.u, avg.v, avg.r, avg.c, &vs_eval);
            }
            double eps_ratio = p->eps_optimal_absolute / p->eps_optimal_relative;
            convergence_info(&oc, eps_ratio, &ms_eval, &vs_eval, &st);
            st.iteration_number = (int32_t)iteration;
            st.cumulative_kkt_matrix_passes = cumulative_kkt_passes;
            st.cumulative_time_sec = wall_time() - start_time;
            st.step_size = step_size;
            st.primal_weight = primal_weight;
            st.time_spent_doing_basic_algorithm = time_basic;
            st.restart_used = RESTART_CHOICE_UNSPECIFIED;

            /* define_norms solver_gpu.jl:51-56 */
            double primal_norm_params = 1 / step_size * primal_weight;
            double dual_norm_params = 1 / step_size / primal_weight;

            int term = check_termination(p, &oc, &st);
            if (numerical_error && term == TERM_NONE) term = TERMINATION_REASON_NUMERICAL_ERROR;
            int32_t rec_idx = -1;
            if (p->record_iteration_stats || term != TERM_NONE) {
                if (nst < stats_cap) {
                    stats[nst] = st;
                    rec_idx = nst;
                }
                nst++;
            }
            if (term != TERM_NONE) {
                /* solver_gpu.jl:614-663: return the average, or the current point */
                if (avg.primal_count == 0 || avg.dual_source_count == 0 || avg.dual_target_count == 0) {
                    memcpy(X_out, cur.X, sizeof(double) * nm);
                    memcpy(u_out, cur.u, sizeof(double) * n);
                    memcpy(v_out, cur.v, sizeof(double) * m);
                } else {
                    memcpy(X_out, avg.X, sizeof(double) * nm);
                    memcpy(u_out, avg.u, sizeof(double) * n);
                    memcpy(v_out, avg.v, sizeof(double) * m);
                }
                *termination_reason = term;
                *iteration_count = (int32_t)(iteration - 1);
                break;
            }

            /* run_restart_scheme saddle_point_gpu.jl:539-696 */
            int restart_used = RESTART_CHOICE_NO_RESTART;
            oracle_restart_rec rr;
            memset(&rr, 0, sizeof(rr));
            rr.iteration = (int32_t)iteration;
            rr.avg_count = avg.primal_count;
            rr.iterations_completed = iteration - 1;
            rr.step_size = step_size;
            rr.primal_weight_before = primal_weight;
            rr.kkt_current = rr.kkt_average = rr.kkt_last = rr.ratio = NAN;
            rr.last_x_norm = rr.last_y_norm = rr.primal_distance = rr.dual_distance = NAN;
            for (int q = 0; q < 3; ++q)
                for (int f = 0; f < 6; ++f) rr.pieces[q][f] = NAN;
            if (avg.primal_count > 0 && avg.dual_source_count > 0 && avg.dual_target_count > 0) {
                /* the weighted KKT residuals the scheme reads (:575-612, 505-515) */
                matstats_t msc, msa;
                vecstats_t vsc, vsa;
                grad_t gc = point_grad(&cur);
                point_matstats(&P, cur.X, &gc, scratch, &msc);
                point_vecstats(&P, cur.u, cur.v, cur.r, cur.c, &vsc);
                grad_t ga = {avg.G, 0, avg.u, avg.v};
                point_matstats(&P, avg.X, &ga, scratch, &msa);
                point_vecstats(&P, avg.u, avg.v, avg.r, avg.c, &vsa);
                double pc[6] = {sqrt(vsc.p_sq), sqrt(msc.rc_sq), msc.pobj, vsc.dobj, sqrt(msc.x_sq), sqrt(vsc.y_sq)};
                double pa[6] = {sqrt(vsa.p_sq), sqrt(msa.rc_sq), msa.pobj, vsa.dobj, sqrt(msa.x_sq), sqrt(vsa.y_sq)};
                double pl[6];
                if (literal) {
                    matstats_t msl;
                    vecstats_t vsl;
                    grad_t gl = {last.G, 0, last.u, last.v};
                    point_matstats(&P, last.X, &gl, scratch, &msl);
                    point_vecstats(&P, last.u, last.v, last.r, last.c, &vsl);
                    double t[6] = {sqrt(vsl.p_sq), sqrt(msl.rc_sq), msl.pobj, vsl.dobj, sqrt(msl.x_sq), sqrt(vsl.y_sq)};
                    memcpy(pl, t, sizeof(pl));
                } else {
                    double t[6] = {last.p2, last.d2, last.pobj, last.dobj, last.x_norm, last.y_norm};
                    memcpy(pl, t, sizeof(pl));
                }
                memcpy(rr.pieces[0], pc, sizeof(pc));
                memcpy(rr.pieces[1], pa, sizeof(pa));
                memcpy(rr.pieces[2], pl, sizeof(pl));
                if (p->restart_scheme != NO_RESTARTS) {
                    rr.kkt_current = oracle_weighted_kkt(primal_weight, pc[0], pc[1], pc[2], pc[3]);
                    rr.kkt_average = oracle_weighted_kkt(primal_weight, pa[0], pa[1], pa[2], pa[3]);
                }
                rr.kkt_last = oracle_weighted_kkt(primal_weight, pl[0], pl[1], pl[2], pl[3]);
                restart_used = oracle_restart_decision(p, avg.primal_count, iteration - 1, rr.kkt_current,
                                                       rr.kkt_average, rr.kkt_last,
                                                       &last.kkt_reduction_ratio_last_trial, &rr);
                if (restart_used != RESTART_CHOICE_NO_RESTART) {
                    const int to_average = (restart_used == RESTART_CHOICE_RESTART_TO_AVERAGE);
                    /* the point the scheme restarts from: under NO_RESTARTS (no
                     * candidate) the current point (:573-576, 654-657) */
                    const double* pn = to_average ? pa : pc;
                    if (to_average) { /* :646-651 */
                        memcpy(cur.X, avg.X, sizeof(double) * nm);
                        memcpy(cur.u, avg.u, sizeof(double) * n);
                        memcpy(cur.v, avg.v, sizeof(double) * m);
                        memcpy(cur.r, avg.r, sizeof(double) * n);
                        memcpy(cur.c, avg.c, sizeof(double) * m);
                        if (literal) memcpy(cur.G, avg.G, sizeof(double) * nm);
                        cur.gzero = 0;
                    }
                    reset_average(&P, &avg); /* :667 */
                    /* update_last_restart_info! :733-780; the average was just
                     * reset, so avg - last = -last (Q2). */
                    double last_x_norm, last_y_norm;
                    if (literal) {
                        double* colbuf = scratch;
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
                        for (int64_t j = 0; j < m; ++j) {
                            double s = 0.0;
                            for (int64_t i = 0; i < n; ++i) {
                                double d = avg.X[i + j * n] - last.X[i + j * n];
                                s += d * d;
                            }
                            colbuf[j] = s;
                        }
                        last_x_norm = sqrt(sum_cols(colbuf, m));
                        double yy = 0.0;
                        for (int64_t i = 0; i < n; ++i) { double d = avg.u[i] - last.u[i]; yy += d * d; }
                        for (int64_t j = 0; j < m; ++j) { double d = avg.v[j] - last.v[j]; yy += d * d; }
                        last_y_norm = sqrt(yy);
                    } else {
                        last_x_norm = last.x_norm;
                        last_y_norm = last.y_norm;
                    }
                    oracle_restart_distances(last_x_norm, last_y_norm, primal_norm_params, dual_norm_params,
                                             primal_weight, &last.primal_distance, &last.dual_distance);
                    rr.last_x_norm = last_x_norm;
                    rr.last_y_norm = last_y_norm;
                    rr.primal_distance = last.primal_distance;
                    rr.dual_distance = last.dual_distance;
                    if (literal) {
                        memcpy(last.X, cur.X, sizeof(double) * nm);
                        memcpy(last.u, cur.u, sizeof(double) * n);
                        memcpy(last.v, cur.v, sizeof(double) * m);
                        memcpy(last.r, cur.r, sizeof(double) * n);
                        memcpy(last.c, cur.c, sizeof(double) * m);
                        memcpy(last.G, cur.G, sizeof(double) * nm);
                    } else {
                        last.p2 = pn[0];
                        last.d2 = pn[1];
                        last.pobj = pn[2];
                        last.dobj = pn[3];
                        last.x_norm = pn[4];
                        last.y_norm = pn[5];
                    }
                    last.last_restart_length = rr.avg_count;
                }
            } else {
                rr.last_trial_before = rr.last_trial_after = last.kkt_reduction_ratio_last_trial;
            }
            if (rec_idx >= 0) stats[rec_idx].restart_used = restart_used;
            if (restart_used != RESTART_CHOICE_NO_RESTART) {
                /* compute_new_primal_weight saddle_point_gpu.jl:701-728 */
                primal_weight = oracle_new_primal_weight(last.primal_distance, last.dual_distance, primal_weight,
                                                         smoothing);
            }
            rr.restart_used = restart_used;
            rr.primal_weight_after = primal_weight;
            if (rtrace && nrt < rtrace_cap) rtrace[nrt] = rr;
            nrt++;
        }

        double t_basic0 = wall_time();
        /* take_step! */
        if (p->step_policy == 0) {
            /* AdaptiveStepsizeParams solver_gpu.jl:290-367 */
            double eta_loc = step_size;
            int done = 0;
            int replay = 0;
            double L = INFINITY;
            while (!done) {
                total_number_iterations += 1;
                if (literal || !replay) {
                    grad_t g = point_grad(&cur);
                    primal_pass(&P, cur.X, &g, step_size, primal_weight, D, cur.r, cur.c, rn, cn);
                    dual_pass(n, a, cur.u, cur.r, rn, step_size, primal_weight, un);
                    dual_pass(m, b, cur.v, cur.c, cn, step_size, primal_weight, vn);
                    double interaction, movement;
                    interaction_movement(&P, D, cur.r, rn, cur.c, cn, cur.u, un, cur.v, vn, primal_weight,
                                         scratch, &interaction, &movement);
                    cumulative_kkt_passes += 1;
                    if (interaction > 0) {
                        L = movement / interaction;
                        if (movement == 0.0) {
                            numerical_error = 1;
                            break;
                        }
                    } else {
                        L = INFINITY;
                    }
                } else {
                    /* fused mode: the replay recomputes the identical pass (Q1) */
                    cumulative_kkt_passes += 1;
                }
                if (eta_loc <= L) {
                    /* update_solution_in_solver_state! :237-264 */
#pragma omp parallel for schedule(static) if(nm >= PAR_MIN)
                    for (int64_t t = 0; t < nm; ++t) cur.X[t] += D[t];
                    memcpy(cur.u, un, sizeof(double) * n);
                    memcpy(cur.v, vn, sizeof(double) * m);
                    memcpy(cur.r, rn, sizeof(double) * n);
                    memcpy(cur.c, cn, sizeof(double) * m);
                    if (literal) {
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
                        for (int64_t j = 0; j < m; ++j)
                            for (int64_t i = 0; i < n; ++i) {
                                double t = C[i + j * n] - cur.u[i];
                                cur.G[i + j * n] = t - cur.v[j];
                            }
                    }
                    cur.gzero = 0;
                    add_to_average(&P, &avg, cur.X, cur.u, cur.v, step_size, cur.r, cur.c, cur.G);
                    done = 1;
                }
                double first = (1 - 1 / jl_pow_int_base(total_number_iterations + 1, p->reduction_exponent)) * L;
                double second = (1 + 1 / jl_pow_int_base(total_number_iterations + 1, p->growth_exponent)) * eta_loc;
                eta_loc = jl_min(first, second);
                replay = 1;
                if (!done && eta_loc != eta_loc) {
                    /* the reference loops forever on a NaN limit; stop and report */
                    numerical_error = 1;
                    break;
                }
            }
            step_size = eta_loc;
        } else {
            /* ConstantStepsizeParams solver_gpu.jl:370-410 */
            grad_t g = point_grad(&cur);
            primal_pass(&P, cur.X, &g, step_size, primal_weight, D, cur.r, cur.c, rn, cn);
            dual_pass(n, a, cur.u, cur.r, rn, step_size, primal_weight, un);
            dual_pass(m, b, cur.v, cur.c, cn, step_size, primal_weight, vn);
            cumulative_kkt_passes += 1;
#pragma omp parallel for schedule(static) if(nm >= PAR_MIN)
            for (int64_t t = 0; t < nm; ++t) cur.X[t] += D[t];
            memcpy(cur.u, un, sizeof(double) * n);
            memcpy(cur.v, vn, sizeof(double) * m);
            memcpy(cur.r, rn, sizeof(double) * n);
            memcpy(cur.c, cn, sizeof(double) * m);
            if (literal) {
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
                for (int64_t j = 0; j < m; ++j)
                    for (int64_t i = 0; i < n; ++i) {
                        double t = C[i + j * n] - cur.u[i];
                        cur.G[i + j * n] = t - cur.v[j];
                    }
            }
            cur.gzero = 0;
            add_to_average(&P, &avg, cur.X, cur.u, cur.v, step_size, cur.r, cur.c, cur.G);
        }
        if (eta_trace && ntrace < trace_cap) eta_trace[ntrace] = step_size;
        ntrace++;
        time_basic += wall_time() - t_basic0;
    }
    *n_stats = nst;
    if (trace_len) *trace_len = ntrace;
    if (n_rtrace) *n_rtrace = nrt;
    free(mats);
    free(vecs);
    free(scratch);
    return rc;
}

int oracle_optimize(const oracle_params* p, int64_t n, int64_t m, const double* C, const double* a,
                    const double* b, int32_t mode, const double* u0, const double* v0, const double* X0, double* X_out,
                    double* u_out, double* v_out, int32_t* termination_reason, int32_t* iteration_count,
                    oracle_stats* stats, int32_t stats_cap, int32_t* n_stats, double* eta_trace,
                    int64_t trace_cap, int64_t* trace_len) {
    return oracle_optimize_ex(p, n, m, C, a, b, mode, u0, v0, X0, X_out, u_out, v_out, termination_reason,
                              iteration_count, stats, stats_cap, n_stats, eta_trace, trace_cap, trace_len, NULL, 0,
                              NULL);
}

/* --------------------------------------------- single-pass primitives */

/* One committed PDHG pass in the fused arithmetic, for kernel-level parity:
 * D = max(X - tau G, 0) - X with G = 0 (gzero) or (C - u) - v; X += D;
 * r+ = rowsum(D) + r, c+ = colsum(D) + c; dual update; weighted-average add
 * with weight eta and pre-update sum S (X_avg etc. updated in place; an
 * average with S == 0 is overwritten).  out6 = {interaction, movement,
 * ||D||^2, dot(dr,du)+dot(dc,dv), ||dy||^2, 0}. */
void oracle_fused_pass(int64_t n, int64_t m, const double* C, const double* a, const double* b, double* X,
                       double* u, double* v, double* r, double* c, int32_t gzero, double eta, double omega,
                       double* Xavg, double* uavg, double* vavg, double* ravg, double* cavg, double S,
                       double* out6) {
    prob_t P = {n, m, C, a, b};
    double* D = (double*)malloc(sizeof(double) * (size_t)(n * m));
    double* buf = (double*)malloc(sizeof(double) * (size_t)(2 * (n + m) + m));
    double *un = buf, *vn = buf + n, *rn = buf + n + m, *cn = buf + 2 * n + m, *colbuf = buf + 2 * (n + m);
    grad_t g = {NULL, gzero, u, v};
    primal_pass(&P, X, &g, eta, omega, D, r, c, rn, cn);
    dual_pass(n, a, u, r, rn, eta, omega, un);
    dual_pass(m, b, v, c, cn, eta, omega, vn);
    double inter, mov;
    interaction_movement(&P, D, r, rn, c, cn, u, un, v, vn, omega, colbuf, &inter, &mov);
    double dd = 0.0;
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += D[i + j * n] * D[i + j * n];
        colbuf[j] = s;
    }
    dd = sum_cols(colbuf, m);
    double dot = 0.0, yy = 0.0;
    for (int64_t i = 0; i < n; ++i) { dot += (rn[i] - r[i]) * (un[i] - u[i]); yy += (un[i] - u[i]) * (un[i] - u[i]); }
    for (int64_t j = 0; j < m; ++j) { dot += (cn[j] - c[j]) * (vn[j] - v[j]); yy += (vn[j] - v[j]) * (vn[j] - v[j]); }
    for (int64_t t = 0; t < n * m; ++t) X[t] += D[t];
    memcpy(u, un, sizeof(double) * n);
    memcpy(v, vn, sizeof(double) * m);
    memcpy(r, rn, sizeof(double) * n);
    memcpy(c, cn, sizeof(double) * m);
    double nw, aw;
    avg_weights(S, eta, &nw, &aw);
    avg_vec(ravg, r, n, nw, aw);
    avg_vec(cavg, c, m, nw, aw);
    avg_mat(Xavg, X, n * m, nw, aw);
    avg_vec(uavg, u, n, nw, aw);
    avg_vec(vavg, v, m, nw, aw);
    out6[0] = inter;
    out6[1] = mov;
    out6[2] = dd;
    out6[3] = dot;
    out6[4] = yy;
    out6[5] = 0.0;
    free(D);
    free(buf);
}

/* The reductions of an evaluation of one point in the fused arithmetic:
 * out11 = {pobj, ||X||_inf, ||X||^2, ||min(G,0)||_inf, ||min(G,0)||^2,
 *          ||a-r;b-c||_inf, ||a-r;b-c||^2, a'u+b'v, ||u||_inf, ||u||^2, ||[u;v]||^2}. */
void oracle_point_reductions(int64_t n, int64_t m, const double* C, const double* a, const double* b,
                             const double* X, const double* u, const double* v, const double* r,
                             const double* c, int32_t gzero, double* out11) {
    prob_t P = {n, m, C, a, b};
    double* scratch = (double*)malloc(sizeof(double) * (size_t)(6 * m + 16));
    grad_t g = {NULL, gzero, u, v};
    matstats_t ms;
    vecstats_t vs;
    point_matstats(&P, X, &g, scratch, &ms);
    point_vecstats(&P, u, v, r, c, &vs);
    out11[0] = ms.pobj; out11[1] = ms.x_inf; out11[2] = ms.x_sq; out11[3] = ms.rc_inf; out11[4] = ms.rc_sq;
    out11[5] = vs.p_inf; out11[6] = vs.p_sq; out11[7] = vs.dobj; out11[8] = vs.u_inf; out11[9] = vs.u_sq;
    out11[10] = vs.y_sq;
    free(scratch);
}

/* rounding utils_gpu.jl:296-318 (Altschuler-Weed-Rigollet), in place on X */
void oracle_rounding(int64_t n, int64_t m, const double* a, const double* b, double* X) {
    double* rs = (double*)malloc(sizeof(double) * (size_t)n);
    double* cs = (double*)malloc(sizeof(double) * (size_t)m);
    double* er = (double*)malloc(sizeof(double) * (size_t)n);
    double* ec = (double*)malloc(sizeof(double) * (size_t)m);
    row_sums(n, m, X, NULL, rs);
    for (int64_t i = 0; i < n; ++i) {
        double t = a[i] / rs[i];
        if (t != t) t = 0.0;
        rs[i] = t < 1.0 ? t : 1.0; /* min.(tmp, 1.0) */
    }
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j)
        for (int64_t i = 0; i < n; ++i) X[i + j * n] = rs[i] * X[i + j * n];
    col_sums(n, m, X, NULL, cs);
    for (int64_t j = 0; j < m; ++j) {
        double t = b[j] / cs[j];
        if (t != t) t = 0.0;
        cs[j] = t < 1.0 ? t : 1.0;
    }
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j)
        for (int64_t i = 0; i < n; ++i) X[i + j * n] = X[i + j * n] * cs[j];
    row_sums(n, m, X, NULL, er);
    double l1 = 0.0;
    for (int64_t i = 0; i < n; ++i) { er[i] = a[i] - er[i]; l1 += fabs(er[i]); }
    if (l1 != 0.0)
        for (int64_t i = 0; i < n; ++i) er[i] /= l1;
    col_sums(n, m, X, NULL, ec);
    for (int64_t j = 0; j < m; ++j) ec[j] = b[j] - ec[j];
#pragma omp parallel for schedule(static) if(n * m >= PAR_MIN)
    for (int64_t j = 0; j < m; ++j)
        for (int64_t i = 0; i < n; ++i) X[i + j * n] += er[i] * ec[j];
    free(rs); free(cs); free(er); free(ec);
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
