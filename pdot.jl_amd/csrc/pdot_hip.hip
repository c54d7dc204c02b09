// pdot_hip.hip -- gfx950 (MI355X, CDNA4) kernels and the C ABI of
// libpdot_hip.so: the PDHG hot path of PDOT.jl re-designed as HBM-streaming
// kernels.  See include/pdot.h for the interface contract.
//
// Data layout in HBM (one context = one n x m problem):
//   C, and four "slots" each holding a point (X, u, v, r, c).  Matrices are
//   column-major with a padded leading dimension ld = roundup(n, 64), FP64 or
//   FP32.  The current point and the running average each live in a slot; a
//   PDHG pass reads the current slot (and the average slot) and writes the new
//   point and the new average into two spare slots, so a pass can be
//   committed, or discarded (restart, termination, numerical error), by
//   changing slot indices only.  G = (C - u 1^T) - 1 v^T is never stored: it
//   is recomputed in registers from C, u and v, exactly as
//   src/solver_gpu.jl:249-250 forms it.
//
// Per PDHG pass (one take_step!, src/solver_gpu.jl:290-367):
//   step_kernel    streams C, X, X_avg -> X+, X_avg+ and emits per-block
//                  partial row sums, column sums and ||Delta||^2 of
//                  Delta = max(X - tau G, 0) - X (5 n*m passes in total).
//   reduce_kernel  finishes the row/column sums (fixed order), applies the
//                  dual update, the averaging of u, v, r, c and the vector
//                  parts of interaction and movement.  Then the controller
//                  (the adaptive step-size recurrence, the replay of a
//                  rejected step and the commit) runs over the blocks'
//                  scalars in one wave: for problems of up to
//                  kFoldMaxBlocks reduce blocks in the reduce kernel's last
//                  block to finish (the blocks publish their scalars with
//                  agent-scope stores and an agent-scope completion count,
//                  with no device-scope fence, an L2 write-back and
//                  invalidate on this GPU), for larger ones in
//                  controller_kernel after the kernel boundary (the
//                  controller's registers would halve the reduce kernel's
//                  resident blocks).
//   commit_kernel  one wave: commits a pass left pending by an evaluation.
// Evaluation of the point and its average (src/utils_gpu.jl:68-129) is fused
// into the step kernel as extra reductions of the values it already loads.
//
// All reductions have a fixed order, so results are bitwise reproducible.
// Compiled with -ffp-contract=off: every expression rounds as the reference's
// broadcasts do.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "pdot.h"
#include "pdot_common.h"
#include "pdot_round.h"

namespace {

struct SlotPtrs {
    void* X[4];
    double* u[4];
    double* v[4];
    double* r[4];
    double* c[4];
};

struct StepParams {
    const void* C;
    const double* px;  // point costs (cost_kind != kCostDense)
    const double* py;
    double cdiv;
    int32_t dim;
    SlotPtrs s;
    const pdot_solver_state* st;
    double* rowpart;
    double* colpart;
    double* sqpart;
    double* evalpart;
    int64_t n, m, ld;
    int32_t TC, Cb, R;
    int32_t TRB;         // rows per row block: n split evenly over the R blocks (<= the block's rows)
    int32_t interleave;  // 1: block cb takes columns cb, cb+Cb, ...; 0: a contiguous range
    int32_t* kind_err;   // set by a leave-alone pass whose state says it must update the average
    uint64_t* zmap[4];   // zero map of each X slot (see ZGeom)
    int64_t ztc;         // map words per wave and tile: the most columns a tile has
    int32_t sparse;      // passes read and write through the zero maps
};

struct ReduceParams {
    SlotPtrs s;
    const double* a;
    const double* b;
    pdot_solver_state* st;
    const double* rowpart;
    const double* colpart;
    const double* sqpart;
    double* redpart;
    double* evalvec;
    uint32_t* done;           // blocks finished in this launch (0 between launches)
    const int32_t* kind_err;  // see step_kernel
    int64_t n, m, ld;
    int32_t Cb, R, NBr, NBc, nred, nstep, sqchunk;
};

// ------------------------------------------------------------- step kernel

// One PDHG pass over the whole problem: block (rb, cb) owns rows
// [rb*TRB, min(n, (rb+1)*TRB)) of the columns cb, cb+Cb, ... (interleaved) or of the
// contiguous range [cb*TC, (cb+1)*TC); step_tile does the work.
// KIND = kKindLeave: a pass the host knows to leave the average alone (the
// passes of a launch sequence alternate deterministically); should the state
// say otherwise, the pass does nothing and flags kind_err, and the controller
// then stops the solve instead of committing.  NT: stream with the
// nontemporal hint (see kCacheableMaxBytes).
template <typename T, bool EVAL, int COST, int KIND = kKindAny, bool NT = true>
__global__ __launch_bounds__((StepShape<T, COST != kCostDense>::kWaves * kWave)) void step_kernel(StepParams p) {
    using Shape = StepShape<T, COST != kCostDense>;
    using LShape = LeaveShape<T, COST != kCostDense>;
    constexpr int WAVES = Shape::kWaves;
    constexpr int UNROLL = KIND == kKindLeave ? LShape::kUnroll : Shape::kUnroll;
    constexpr int PIPE = KIND == kKindLeave ? LShape::kPipe : Shape::kPipe;
    constexpr int TR = WAVES * kWave * VecT<T>::N;
    __shared__ pdot_solver_state sst;
    __shared__ double red[WAVES * kWave];
    __shared__ double colbuf[WAVES * col_rows<T>()][kChunk];
    const int tid = threadIdx.x;
    if (tid == 0) sst = *p.st;
    __syncthreads();
    const bool do_step = !sst.numerical_error;
    if (!do_step && !EVAL) return;
    if (KIND == kKindLeave && !pass_defers(sst, false)) {
        if (tid == 0 && blockIdx.x == 0) *p.kind_err = 1;
        return;
    }

    const int rb = blockIdx.x % p.R;
    const int cb = blockIdx.x / p.R;
    const bool ilv = p.interleave != 0;
    TileArgs<T> a;
    a.n = p.n;
    a.ld = p.ld;
    a.row0 = (int64_t)rb * p.TRB;
    a.row_end = (a.row0 + p.TRB < p.n) ? a.row0 + p.TRB : p.n;
    a.j0 = ilv ? cb : (int64_t)cb * p.TC;
    a.jstep = ilv ? p.Cb : 1;
    if (ilv) a.ncols = (p.m - cb + p.Cb - 1) / p.Cb;
    else a.ncols = ((a.j0 + p.TC < p.m) ? a.j0 + p.TC : p.m) - a.j0;
    a.avg_empty = sst.avg_count == 0;
    const int cs = sst.cur_slot;
    const int as = a.avg_empty ? cs : sst.avg_slot;
    a.C = static_cast<const T*>(p.C);
    a.px = p.px;
    a.py = p.py;
    a.cdiv = p.cdiv;
    a.dim = p.dim;
    a.Xc = static_cast<const T*>(p.s.X[cs]);
    a.Xa = static_cast<const T*>(p.s.X[as]);
    a.Xn = static_cast<T*>(p.s.X[sst.spare0]);
    a.Xan = static_cast<T*>(p.s.X[sst.spare1]);
    a.uc = p.s.u[cs];
    a.ua = p.s.u[as];
    a.vc = p.s.v[cs];
    a.va = p.s.v[as];
    a.rowpart = p.rowpart + (int64_t)cb * p.ld;
    a.colpart = p.colpart + (int64_t)rb * p.m;
    a.eta = sst.step_size;
    a.omega = sst.primal_weight;
    const double S = sst.sum_weights;
    a.nw = 1.0;
    a.aw = 1.0;
    if (S > 0.0) {
        a.aw = S / (S + a.eta);
        a.nw = a.eta / S;
    }
    a.gzero = sst.gzero != 0;
    a.do_step = do_step;
    a.defer = pass_defers(sst, EVAL);
    a.pend = sst.avg_pending != 0;
    a.pnw = sst.pend_nw;
    a.paw = sst.pend_aw;
    a.sparse = p.sparse != 0;
    const int64_t zbase = ((int64_t)rb * p.Cb + cb) * WAVES * p.ztc;
    a.zc = p.zmap[cs] + zbase;
    a.za = p.zmap[as] + zbase;
    a.zn = p.zmap[sst.spare0] + zbase;
    a.zan = p.zmap[sst.spare1] + zbase;
    a.zstride = p.ztc;
    double dd = 0.0, ev[kMatStats];
    step_tile<T, EVAL, UNROLL, WAVES, PIPE, COST, KIND, NT>(a, colbuf, red, &dd, ev);
    if (tid == 0) {
        if (do_step) p.sqpart[blockIdx.x] = dd;
        if (EVAL) {
            for (int k = 0; k < kMatStats; ++k) p.evalpart[(int64_t)blockIdx.x * kMatStats + k] = ev[k];
        }
    }
}

// ----------------------------------------------------------- reduce kernel

// Reduce blocks up to which the controller runs in the reduce kernel's last
// block (n + m <= 16384).  Measured per iteration (host clock, one process):
// 34.0 -> 20.7 us at 1000 x 1000, 66.1 -> 60.0 us at 4096 x 4096 (FP32
// storage), 580 -> 584 us at 10000 x 10000 (FP64, 314 blocks), and the
// reduce kernel 6.0 + 6.9 -> 18.8 us at 30000 x 30000 (938 blocks, which no
// longer fit in one round at the folded kernel's 90 registers).
constexpr int32_t kFoldMaxBlocks = 256;

// Block b < NBr owns rows [64b, 64b+64), the next NBc blocks the columns;
// reduce_chunk does the work; each block leaves its three scalars in
// redpart.  FOLD: the blocks publish them (agent-scope stores, then an
// agent-scope completion count, no device-scope fence), and the last block
// to finish runs the controller over them in one wave: it commits the pass,
// or leaves an evaluation's pass pending for the host's decision.  Every
// block reads the state before it counts itself done, so the last block may
// update it in place.  A halted solve (numerical error) counts nothing and
// commits nothing.  Without FOLD, controller_kernel follows.
template <bool EVAL, bool FOLD>
__global__ __launch_bounds__(kRedBlock) void reduce_kernel(ReduceParams p) {
    __shared__ pdot_solver_state sst;
    __shared__ double psum[kRedWaves][kWave];
    __shared__ int32_t kind_err;
    const int tid = threadIdx.x;
    if (tid == 0) {
        sst = *p.st;
        kind_err = *p.kind_err;  // written by the step kernel before this one
    }
    __syncthreads();
    const bool do_step = !sst.numerical_error;
    if (!do_step && !EVAL) return;
    const bool rows = (int)blockIdx.x < p.NBr;
    const bool avg_empty = sst.avg_count == 0;
    const int cs = sst.cur_slot;
    const int as = avg_empty ? cs : sst.avg_slot;
    ChunkArgs c;
    c.rows = rows;
    c.k0 = (int64_t)(rows ? blockIdx.x : blockIdx.x - p.NBr) * kWave;
    c.len = rows ? p.n : p.m;
    c.part = rows ? p.rowpart : p.colpart;
    c.pstride = rows ? p.ld : p.m;
    c.np = rows ? p.Cb : p.R;
    c.q = rows ? p.a : p.b;
    double* const* ys = rows ? p.s.u : p.s.v;
    double* const* ss = rows ? p.s.r : p.s.c;
    c.y_cur = ys[cs];
    c.s_cur = ss[cs];
    c.y_avg = ys[as];
    c.s_avg = ss[as];
    c.y_new = ys[sst.spare0];
    c.s_new = ss[sst.spare0];
    c.y_anew = ys[sst.spare1];
    c.s_anew = ss[sst.spare1];
    c.sqpart = p.sqpart;
    c.sq_lo = (int64_t)blockIdx.x * p.sqchunk;
    c.sq_hi = c.sq_lo + p.sqchunk < p.nstep ? c.sq_lo + p.sqchunk : p.nstep;
    c.eta = sst.step_size;
    c.omega = sst.primal_weight;
    const double S = sst.sum_weights;
    c.nw = 1.0;
    c.aw = 1.0;
    if (S > 0.0) {
        c.aw = S / (S + c.eta);
        c.nw = c.eta / S;
    }
    c.avg_empty = avg_empty;
    c.do_step = do_step;
    c.defer = pass_defers(sst, EVAL);
    c.pend = sst.avg_pending != 0;
    c.pnw = sst.pend_nw;
    c.paw = sst.pend_aw;
    double out3[3], ev[kVecStats];
    reduce_chunk<EVAL, kRedWaves>(c, psum, out3, ev);
    __shared__ int last;
    if (tid == 0) {
        if (EVAL) {
            for (int t = 0; t < kVecStats; ++t) p.evalvec[(int64_t)blockIdx.x * kVecStats + t] = ev[t];
        }
        last = 0;
        if (do_step && !FOLD) {
            p.redpart[3 * blockIdx.x + 0] = out3[0];
            p.redpart[3 * blockIdx.x + 1] = out3[1];
            p.redpart[3 * blockIdx.x + 2] = out3[2];
        }
        if (do_step && FOLD) {
            publish_f64(&p.redpart[3 * blockIdx.x + 0], out3[0]);
            publish_f64(&p.redpart[3 * blockIdx.x + 1], out3[1]);
            publish_f64(&p.redpart[3 * blockIdx.x + 2], out3[2]);
            wait_memory_ops();
            last = count_done(p.done, gridDim.x) ? 1 : 0;
        }
    }
    if (!FOLD) return;
    __syncthreads();
    if (!last || tid >= kWave) return;
    // the controller, in wave 0 of the last block
    if (EVAL) {  // an evaluation's pass waits for the host's decision
        if (tid == 0) p.st->pending = 1;
        return;
    }
    if (kind_err) {  // the step kernel did not perform the pass (see step_kernel)
        if (tid == 0) {
            p.st->numerical_error = 1;
            p.st->error_call = sst.take_step_calls + 1;
        }
        return;
    }
    reduce_and_commit<true, 8>(&sst, p.redpart, p.nred);
    if (tid == 0) *p.st = sst;
}

// Final fixed-order reduction of the evaluation partials into 22 values.
__global__ __launch_bounds__(kBlock) void evalfin_kernel(const double* evalpart, int64_t nstep, const double* evalvec,
                                                         int64_t nred, double* out) {
    __shared__ double red[kBlock];
    const int tid = threadIdx.x;
    for (int k = 0; k < kMatStats; ++k) {
        const bool is_max = (k == 1 || k == 3 || k == 6 || k == 8);
        double acc = 0.0;
        for (int64_t b = tid; b < nstep; b += kBlock) {
            const double v = evalpart[b * kMatStats + k];
            acc = is_max ? nanmax(acc, v) : acc + v;
        }
        acc = is_max ? block_max(acc, red) : block_sum(acc, red);
        if (tid == 0) out[k] = acc;
    }
    for (int k = 0; k < kVecStats; ++k) {
        const bool is_max = (k == 0 || k == 3 || k == 6 || k == 9);
        double acc = 0.0;
        for (int64_t b = tid; b < nred; b += kBlock) {
            const double v = evalvec[b * kVecStats + k];
            acc = is_max ? nanmax(acc, v) : acc + v;
        }
        acc = is_max ? block_max(acc, red) : block_sum(acc, red);
        if (tid == 0) out[kMatStats + k] = acc;
    }
}

// Commits the pass the reduce kernel (without FOLD) has reduced; an
// evaluation's pass is left pending for the host's decision instead (one
// wave).  A halted solve (numerical error) commits nothing.
template <bool EVAL>
__global__ __launch_bounds__(kWave) void controller_kernel(pdot_solver_state* st, const double* redpart, int32_t nred,
                                                           const int32_t* kind_err) {
    __shared__ pdot_solver_state sst;
    if (threadIdx.x == 0) sst = *st;
    __syncthreads();
    if (sst.numerical_error) return;
    if (EVAL) {
        if (threadIdx.x == 0) st->pending = 1;
        return;
    }
    if (*kind_err) {  // the step kernel did not perform the pass (see step_kernel)
        if (threadIdx.x == 0) {
            st->numerical_error = 1;
            st->error_call = sst.take_step_calls + 1;
        }
        return;
    }
    reduce_and_commit(&sst, redpart, nred);
    if (threadIdx.x == 0) *st = sst;
}

// Commits a pending pass (one wave).
__global__ __launch_bounds__(kWave) void commit_kernel(pdot_solver_state* st, const double* redpart, int32_t nred) {
    __shared__ pdot_solver_state sst;
    if (threadIdx.x == 0) sst = *st;
    __syncthreads();
    if (!sst.pending) return;
    reduce_and_commit(&sst, redpart, nred);
    if (threadIdx.x == 0) *st = sst;
}

constexpr int kIngestBlocks = 1024;  // blocks per ingest chunk

// Ingest of one column block of the FP64 cost matrix (n x ncols, source
// pitch lds): per-block max |C| (NaN propagates, as Julia's maximum does) and
// sum of C^2, the inputs of cached_ot_info (src/termination.jl:94-101); with
// NARROW the block is also stored as FP32 (pitch ld).  Block b takes columns
// b, b + gridDim.x, ...; threads stride the rows.
template <bool NARROW>
__global__ __launch_bounds__(256) void ingest_kernel(const double* __restrict__ src, int64_t lds, float* __restrict__ dst,
                                                     int64_t ld, int64_t n, int64_t ncols, double* __restrict__ part) {
    __shared__ double red[256];
    double mx = 0.0, ss = 0.0;
    for (int64_t j = blockIdx.x; j < ncols; j += gridDim.x) {
        const double* col = src + j * lds;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            const double x = col[i];
            if (NARROW) dst[j * ld + i] = (float)x;
            mx = nanmax(mx, fabs(x));
            ss += x * x;
        }
    }
    mx = block_max(mx, red);
    ss = block_sum(ss, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mx;
        part[2 * blockIdx.x + 1] = ss;
    }
}

// cached_ot_info inputs for point costs: the same partials, in the same
// order, as ingest_kernel over the matrix the points define (one launch per
// upload chunk of ncols columns from column j0), formed in registers.
template <int COST>
__global__ __launch_bounds__(256) void point_stats_kernel(const double* __restrict__ px, const double* __restrict__ py,
                                                          int64_t n, int dim, double cdiv, int64_t j0, int64_t ncols,
                                                          double* __restrict__ part) {
    __shared__ double red[256];
    double mx = 0.0, ss = 0.0;
    const double crcp = 1.0 / cdiv;
    for (int64_t j = blockIdx.x; j < ncols; j += gridDim.x) {
        double yj[kMaxDim];
        for (int d = 0; d < kMaxDim; ++d) yj[d] = d < dim ? py[(j0 + j) * dim + d] : 0.0;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            double xi[kMaxDim];
            for (int d = 0; d < kMaxDim; ++d) xi[d] = d < dim ? px[(int64_t)d * n + i] : 0.0;
            const double x = point_cost<COST, true>(xi, yj, dim, cdiv, crcp);
            mx = nanmax(mx, fabs(x));
            ss += x * x;
        }
    }
    mx = block_max(mx, red);
    ss = block_sum(ss, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mx;
        part[2 * blockIdx.x + 1] = ss;
    }
}

// Plain 16-byte copy (counter and bandwidth calibration).
__global__ __launch_bounds__(256) void copy_kernel(const dbl2* __restrict__ src, dbl2* __restrict__ dst, int64_t nv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) dst[i] = src[i];
}

// FP64 host layout (pitch n) -> device storage (pitch ld), for ncols columns.
template <typename T>
__global__ void narrow_kernel(const double* __restrict__ src, T* __restrict__ dst, int64_t n, int64_t ld, int64_t ncols) {
    const int64_t total = n * ncols;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        dst[j * ld + i] = (T)src[t];
    }
}
template <typename T>
__global__ void widen_kernel(const T* __restrict__ src, double* __restrict__ dst, int64_t n, int64_t ld, int64_t ncols) {
    const int64_t total = n * ncols;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        dst[t] = (double)src[j * ld + i];
    }
}

// Lanes marked as holding a value (set bits of the lanes below each row
// block's end), summed into *out.
__global__ __launch_bounds__(256) void zcount_kernel(const uint64_t* __restrict__ z, ZGeom g,
                                                     unsigned long long* __restrict__ out) {
    unsigned long long live = 0;
    for (int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; wi < g.words; wi += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = wi % g.tc;
        int64_t r = wi / g.tc;
        const int64_t w = r % g.waves;
        r /= g.waves;
        const int64_t cb = r % g.Cb, rb = r / g.Cb;
        const int64_t ncols = g.interleave ? (g.m - cb + g.Cb - 1) / g.Cb
                                           : ((cb * g.TC + g.TC < g.m) ? cb * g.TC + g.TC : g.m) - cb * g.TC;
        const int64_t row0 = rb * g.TRB + w * kWave * g.vec;
        const int64_t row_end = (rb * g.TRB + g.TRB < g.n) ? rb * g.TRB + g.TRB : g.n;
        if (t >= ncols || row0 >= row_end) continue;
        const int64_t lanes = (row_end - row0 + g.vec - 1) / g.vec;
        const uint64_t mask = lanes >= 64 ? ~0ull : ((1ull << lanes) - 1);
        live += (unsigned long long)__popcll(z[wi] & mask);
    }
    for (int off = 32; off > 0; off >>= 1) live += __shfl_xor(live, off);
    if ((threadIdx.x & (kWave - 1)) == 0 && live) atomicAdd(out, live);
}

}  // namespace

// ====================================================================== host

struct pdot_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0, m = 0, ld = 0;
    int32_t storage = PDOT_STORAGE_F64;
    size_t esz = 8;
    void* C = nullptr;
    SlotPtrs s{};
    double* a = nullptr;
    double* b = nullptr;
    double* rowpart = nullptr;
    double* colpart = nullptr;
    double* sqpart = nullptr;
    double* evalpart = nullptr;
    double* redpart = nullptr;
    double* evalvec = nullptr;
    double* evalout = nullptr;
    double* stage = nullptr;
    int64_t stage_cols = 0;
    pdot_solver_state* stA = nullptr;
    int32_t resident_blocks = 0;          // step-kernel blocks resident at once
    pdot_solver_state* h_state = nullptr;  // pinned
    double* h_eval = nullptr;              // pinned
    double* h_pin[2] = {nullptr, nullptr};  // pinned upload staging ring
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    int64_t pin_cols = 0;                   // columns per upload chunk
    double* ot_part = nullptr;              // per-chunk, per-block (max |C|, sum C^2)
    int64_t ot_part_bytes = 0;
    double ot_info[4] = {0, 0, 0, 0};
    bool ot_ready = false;
    int32_t cost_kind = kCostDense;         // kCostDense: C stored; else formed from px, py
    int32_t dim = 0;
    double cdiv = 1.0;
    double* px = nullptr;                   // dim x n source coordinates
    double* py = nullptr;                   // m x dim target coordinates
    int32_t R = 0, Cb = 0, TC = 0, TRB = 0, nstep = 0, NBr = 0, NBc = 0, nred = 0, sqchunk = 0;
    int32_t cb_hint = 0;
    int32_t interleave = 1;
    bool uploaded = false, started = false;
    bool use_graph = false;
    bool defer = true;  // ordinary passes alternate deferring / applying the average's update
    bool pass_kernels = true;     // the passes that leave the average alone use their own kernel
    int32_t fold_mode = -1;       // controller in the reduce kernel: -1 by size, 0 never, 1 always
    int32_t stream_policy = -1;   // step kernel streams: -1 by size, 0 nontemporal, 1 cacheable
    bool sparse = true;           // passes skip the zero rows recorded in the zero maps
    uint64_t* zmap[4] = {nullptr, nullptr, nullptr, nullptr};  // per X slot (geometry zg)
    ZGeom zg;
    int32_t* kind_err = nullptr;  // device flag, see step_kernel
    uint32_t* done = nullptr;     // reduce_kernel's completion count, see there
    std::map<int64_t, hipGraphExec_t> graphs;
    std::vector<hipEvent_t> events;
    int64_t bytes = 0;
    std::string err;
};

namespace {

int fail(pdot_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            return fail((ctx), PDOT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));         \
        }                                                                                              \
    } while (0)

int step_waves(int32_t storage, int cost) {
    const bool pts = cost != kCostDense;
    if (storage == PDOT_STORAGE_F64) return pts ? StepShape<double, true>::kWaves : StepShape<double, false>::kWaves;
    return pts ? StepShape<float, true>::kWaves : StepShape<float, false>::kWaves;
}
int rows_per_block(int32_t storage, int cost) {
    return step_waves(storage, cost) * kWave * (storage == PDOT_STORAGE_F64 ? 2 : 4);
}
static_assert(StepShape<double, false>::kWaves >= 1 && StepShape<double, false>::kWaves <= 16, "step kernel waves");
static_assert(StepShape<float, false>::kWaves >= 1 && StepShape<float, false>::kWaves <= 16, "step kernel waves");
static_assert(StepShape<double, true>::kWaves >= 1 && StepShape<double, true>::kWaves <= 16, "step kernel waves");
static_assert(StepShape<float, true>::kWaves >= 1 && StepShape<float, true>::kWaves <= 16, "step kernel waves");

int dev_alloc(pdot_ctx* ctx, void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ctx, PDOT_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    ctx->bytes += (int64_t)bytes;
    return PDOT_OK;
}

void free_graphs(pdot_ctx* ctx) {
    for (auto& kv : ctx->graphs) hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
}

void free_partials(pdot_ctx* ctx) {
    double** ps[] = {&ctx->rowpart, &ctx->colpart, &ctx->sqpart, &ctx->evalpart, &ctx->redpart, &ctx->evalvec};
    for (double** q : ps) {
        if (*q) hipFree(*q);
        *q = nullptr;
    }
}

int fill_map(pdot_ctx* ctx, int slot, uint64_t value);

// Zero maps.  Makes slot `slot`'s memory hold every value in the dense
// layout, and its map all ones (which describes that layout).
int densify_slot(pdot_ctx* ctx, int slot) {
    if (!ctx->zmap[slot] || ctx->zg.words == 0) return PDOT_OK;
    const ZGeom& g = ctx->zg;
    const int64_t blocks = std::min<int64_t>((g.words + 3) / 4, 8192);
    if (ctx->storage == PDOT_STORAGE_F64)
        hipLaunchKernelGGL(densify_kernel<double>, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<double*>(ctx->s.X[slot]), ctx->zmap[slot], g);
    else
        hipLaunchKernelGGL(densify_kernel<float>, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<float*>(ctx->s.X[slot]), ctx->zmap[slot], g);
    HIPCHK(ctx, hipGetLastError());
    return fill_map(ctx, slot, ~0ull);
}

int fill_map(pdot_ctx* ctx, int slot, uint64_t value) {
    if (!ctx->zmap[slot] || ctx->zg.words == 0) return PDOT_OK;
    const int64_t blocks = std::min<int64_t>((ctx->zg.words + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks), dim3(256), 0, ctx->stream, ctx->zmap[slot], ctx->zg.words, value);
    HIPCHK(ctx, hipGetLastError());
    return PDOT_OK;
}

// Before the maps are dropped or no longer kept up to date: every slot's
// memory holds every value.
int densify_all(pdot_ctx* ctx) {
    int rc;
    for (int s = 0; s < 4; ++s)
        if ((rc = densify_slot(ctx, s))) return rc;
    return PDOT_OK;
}

void free_maps(pdot_ctx* ctx) {
    for (int s = 0; s < 4; ++s) {
        if (ctx->zmap[s]) {
            hipFree(ctx->zmap[s]);
            ctx->bytes -= (int64_t)(sizeof(uint64_t) * (size_t)ctx->zg.words);
        }
        ctx->zmap[s] = nullptr;
    }
    ctx->zg = ZGeom();
}

// Maps for the present geometry, all bits set (the slots' memory holds every value).
int alloc_maps(pdot_ctx* ctx) {
    ZGeom g;
    g.n = ctx->n;
    g.m = ctx->m;
    g.ld = ctx->ld;
    g.R = ctx->R;
    g.Cb = ctx->Cb;
    g.TRB = ctx->TRB;
    g.TC = ctx->TC;
    g.interleave = ctx->interleave;
    g.waves = step_waves(ctx->storage, ctx->cost_kind);
    g.vec = (int32_t)(16 / ctx->esz);
    g.tc = ctx->interleave ? (ctx->m + ctx->Cb - 1) / ctx->Cb : ctx->TC;
    g.words = (int64_t)g.R * g.Cb * g.waves * g.tc;
    int rc;
    for (int s = 0; s < 4; ++s)
        if ((rc = dev_alloc(ctx, (void**)&ctx->zmap[s], sizeof(uint64_t) * (size_t)g.words))) return rc;
    ctx->zg = g;
    for (int s = 0; s < 4; ++s)
        if ((rc = fill_map(ctx, s, ~0ull))) return rc;
    return PDOT_OK;
}

// Launch geometry: R row blocks of rows_per_block rows; Cb column blocks.
int setup_geometry(pdot_ctx* ctx) {
    free_graphs(ctx);
    free_partials(ctx);  // evalvec, allocated last, stays null if an allocation below fails
    // the maps describe the slots in the geometry being replaced
    int rc0 = ctx->sparse ? densify_all(ctx) : PDOT_OK;
    if (rc0) return rc0;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    free_maps(ctx);
    const int TR = rows_per_block(ctx->storage, ctx->cost_kind);
    const int64_t R = (ctx->n + TR - 1) / TR;
    // One full wave of resident blocks: every block streams the same number
    // of columns and no straggler round follows.  Both kinds of pass use this
    // geometry (the generic kernel's occupancy): the leave-alone kernel's
    // lower register count would fit more column blocks (2% faster for FP32
    // storage), but the partial row sums would then be grouped differently
    // in every second pass and the results would no longer equal those of
    // the eager average bit for bit.
    int64_t Cb;
    if (ctx->cb_hint > 0) {
        Cb = ctx->cb_hint;
    } else {
        Cb = ctx->resident_blocks / R;
    }
    if (Cb < 1) Cb = 1;
    if (Cb > ctx->m) Cb = ctx->m;
    int64_t TC = (ctx->m + Cb - 1) / Cb;
    Cb = (ctx->m + TC - 1) / TC;
    if (R * Cb > 0x7fffffffLL) return fail(ctx, PDOT_E_ARG, "grid too large");
    ctx->R = (int32_t)R;
#ifndef PDOT_ROW_SPLIT_ALIGN
#define PDOT_ROW_SPLIT_ALIGN 128  // bytes (measured best of 64, 128, 256 and no split); 0: full-height row blocks
#endif
    if (PDOT_ROW_SPLIT_ALIGN > 0) {
        // even row split with aligned row offsets, so that no row block is short
        const int64_t align = PDOT_ROW_SPLIT_ALIGN / (int64_t)ctx->esz;
        const int64_t per = (ctx->n + R - 1) / R;
        ctx->TRB = (int32_t)std::min<int64_t>(TR, (per + align - 1) / align * align);
    } else {
        ctx->TRB = TR;
    }
    ctx->Cb = (int32_t)Cb;
    ctx->TC = (int32_t)TC;
    ctx->nstep = (int32_t)(R * Cb);
    ctx->NBr = (int32_t)((ctx->n + kWave - 1) / kWave);
    ctx->NBc = (int32_t)((ctx->m + kWave - 1) / kWave);
    ctx->nred = ctx->NBr + ctx->NBc;
    ctx->sqchunk = (ctx->nstep + ctx->nred - 1) / ctx->nred;
    int rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->rowpart, sizeof(double) * (size_t)(Cb * ctx->ld)))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->colpart, sizeof(double) * (size_t)(R * ctx->m)))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->sqpart, sizeof(double) * (size_t)ctx->nstep))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->evalpart, sizeof(double) * (size_t)ctx->nstep * kMatStats))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->redpart, sizeof(double) * (size_t)ctx->nred * 3))) return rc;
    if ((rc = alloc_maps(ctx))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->evalvec, sizeof(double) * (size_t)ctx->nred * kVecStats))) return rc;
    return PDOT_OK;
}

StepParams step_params(pdot_ctx* ctx) {
    StepParams p;
    p.C = ctx->C;
    p.px = ctx->px;
    p.py = ctx->py;
    p.cdiv = ctx->cdiv;
    p.dim = ctx->dim;
    p.s = ctx->s;
    p.st = ctx->stA;
    p.rowpart = ctx->rowpart;
    p.colpart = ctx->colpart;
    p.sqpart = ctx->sqpart;
    p.evalpart = ctx->evalpart;
    p.n = ctx->n;
    p.m = ctx->m;
    p.ld = ctx->ld;
    p.TC = ctx->TC;
    p.TRB = ctx->TRB;
    p.Cb = ctx->Cb;
    p.R = ctx->R;
    p.interleave = ctx->interleave;
    p.kind_err = ctx->kind_err;
    for (int k = 0; k < 4; ++k) p.zmap[k] = ctx->zmap[k];
    p.ztc = ctx->zg.tc;
    p.sparse = (ctx->sparse && ctx->zg.words > 0) ? 1 : 0;
    return p;
}

ReduceParams reduce_params(pdot_ctx* ctx) {
    ReduceParams p;
    p.s = ctx->s;
    p.a = ctx->a;
    p.b = ctx->b;
    p.st = ctx->stA;
    p.rowpart = ctx->rowpart;
    p.colpart = ctx->colpart;
    p.sqpart = ctx->sqpart;
    p.redpart = ctx->redpart;
    p.evalvec = ctx->evalvec;
    p.done = ctx->done;
    p.kind_err = ctx->kind_err;
    p.n = ctx->n;
    p.m = ctx->m;
    p.ld = ctx->ld;
    p.Cb = ctx->Cb;
    p.R = ctx->R;
    p.NBr = ctx->NBr;
    p.NBc = ctx->NBc;
    p.nred = ctx->nred;
    p.nstep = ctx->nstep;
    p.sqchunk = ctx->sqchunk;
    return p;
}

template <typename T, int COST, bool NT>
void launch_step_c(bool eval, bool leave, dim3 grid, dim3 block, hipStream_t stream, const StepParams& p) {
    if (eval) hipLaunchKernelGGL((step_kernel<T, true, COST, kKindAny, NT>), grid, block, 0, stream, p);
    else if constexpr (LeaveShape<T, COST != kCostDense>::kOwnKernel) {
        if (leave) hipLaunchKernelGGL((step_kernel<T, false, COST, kKindLeave, NT>), grid, block, 0, stream, p);
        else hipLaunchKernelGGL((step_kernel<T, false, COST, kKindAny, NT>), grid, block, 0, stream, p);
    } else {
        hipLaunchKernelGGL((step_kernel<T, false, COST, kKindAny, NT>), grid, block, 0, stream, p);
    }
}

template <typename T, bool NT>
void launch_step_n(int cost, bool eval, bool leave, dim3 grid, dim3 block, hipStream_t stream, const StepParams& p) {
    if (cost == kCostSqEuclid) launch_step_c<T, kCostSqEuclid, NT>(eval, leave, grid, block, stream, p);
    else if (cost == kCostL1) launch_step_c<T, kCostL1, NT>(eval, leave, grid, block, stream, p);
    else launch_step_c<T, kCostDense, NT>(eval, leave, grid, block, stream, p);
}

template <typename T>
void launch_step_t(int cost, bool nt, bool eval, bool leave, dim3 grid, dim3 block, hipStream_t stream,
                   const StepParams& p) {
    if (nt) launch_step_n<T, true>(cost, eval, leave, grid, block, stream, p);
    else launch_step_n<T, false>(cost, eval, leave, grid, block, stream, p);
}

// leave: the pass is known to leave the average alone (kKindLeave kernel
// where the form has one and the per-kind kernels are enabled).
bool uses_leave_kernel(pdot_ctx* ctx, bool eval, bool leave) {
    if (!leave || eval || !ctx->pass_kernels) return false;
    const bool points = ctx->cost_kind != kCostDense;
    if (ctx->storage == PDOT_STORAGE_F64) return points ? LeaveShape<double, true>::kOwnKernel : LeaveShape<double, false>::kOwnKernel;
    return points ? LeaveShape<float, true>::kOwnKernel : LeaveShape<float, false>::kOwnKernel;
}

// Largest footprint of a pass's matrices (five n x m arrays at the storage
// type, whether or not C is stored, so that the point-cloud and dense forms
// choose alike) streamed without the nontemporal hint.  Below it the lines a
// pass leaves in the 256 MB Infinity Cache serve part of the next pass.
// Measured per iteration without the hint against with it (one process,
// gaussian-shift instances): 160 MiB (2048^2 FP64) -14%, 172 MiB (3000^2
// FP32) -15%, 209 MiB -16%, 250 MiB -9%, 320 MiB (4096^2 FP32) -4%, 344 MiB
// (3000^2 FP64) -7%; 390 MiB (4500^2 FP32) +4%, 419 MiB +8%, 482-964 MiB
// +8-11%; at 20-39 MiB (1000^2) within 2% either way.
constexpr int64_t kCacheableMaxBytes = int64_t(352) << 20;

bool streams_nontemporal(const pdot_ctx* ctx) {
    if (ctx->stream_policy == 0) return true;
    if (ctx->stream_policy == 1) return false;
    return 5 * ctx->ld * ctx->m * (int64_t)ctx->esz > kCacheableMaxBytes;
}

hipError_t launch_step(pdot_ctx* ctx, bool eval, bool leave = false) {
    leave = uses_leave_kernel(ctx, eval, leave);
    StepParams p = step_params(ctx);
    dim3 grid(ctx->nstep), block(step_waves(ctx->storage, ctx->cost_kind) * kWave);
    const bool nt = streams_nontemporal(ctx);
    if (ctx->storage == PDOT_STORAGE_F64) launch_step_t<double>(ctx->cost_kind, nt, eval, leave, grid, block, ctx->stream, p);
    else launch_step_t<float>(ctx->cost_kind, nt, eval, leave, grid, block, ctx->stream, p);
    return hipGetLastError();
}

hipError_t launch_reduce(pdot_ctx* ctx, bool eval) {
    ReduceParams p = reduce_params(ctx);
    dim3 grid(ctx->nred), block(kRedBlock);
    const bool fold = ctx->fold_mode == 1 || (ctx->fold_mode < 0 && ctx->nred <= kFoldMaxBlocks);
    if (fold) {
        if (eval) hipLaunchKernelGGL((reduce_kernel<true, true>), grid, block, 0, ctx->stream, p);
        else hipLaunchKernelGGL((reduce_kernel<false, true>), grid, block, 0, ctx->stream, p);
        return hipGetLastError();
    }
    if (eval) hipLaunchKernelGGL((reduce_kernel<true, false>), grid, block, 0, ctx->stream, p);
    else hipLaunchKernelGGL((reduce_kernel<false, false>), grid, block, 0, ctx->stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (eval) hipLaunchKernelGGL((controller_kernel<true>), dim3(1), dim3(kWave), 0, ctx->stream, ctx->stA, ctx->redpart, ctx->nred,
                                 ctx->kind_err);
    else hipLaunchKernelGGL((controller_kernel<false>), dim3(1), dim3(kWave), 0, ctx->stream, ctx->stA, ctx->redpart, ctx->nred,
                            ctx->kind_err);
    return hipGetLastError();
}

hipError_t launch_commit(pdot_ctx* ctx) {
    hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(kWave), 0, ctx->stream, ctx->stA, ctx->redpart, ctx->nred);
    return hipGetLastError();
}

// Whether pass k (from 0) of a launch sequence leaves the average alone.
// With deferral, a committed pass that defers leaves the update pending and
// the next applies it, so the kinds alternate from the first: the first
// defers if the average is non-empty and holds no lagging contribution once a
// pending evaluation's pass (which always updates the average) is committed.
// A numerical error halts the remaining passes, whatever their kind.
bool first_pass_leaves(const pdot_solver_state& st) {
    if (!st.defer_ok) return false;
    if (st.pending) return true;
    return st.avg_pending == 0 && st.avg_count > 0;
}
bool pass_leaves(bool first_leaves, bool defer_ok, int64_t k) {
    return defer_ok && ((k % 2 == 0) == first_leaves);
}

// A pass left pending by an evaluation is committed first; each later
// pass commits itself in its reduce kernel.
int enqueue_steps(pdot_ctx* ctx, int64_t count, bool first_leaves, bool defer_ok) {
    HIPCHK(ctx, launch_commit(ctx));
    for (int64_t k = 0; k < count; ++k) {
        const bool leave = pass_leaves(first_leaves, defer_ok, k);
        HIPCHK(ctx, launch_step(ctx, false, leave));
        HIPCHK(ctx, launch_reduce(ctx, false));
    }
    return PDOT_OK;
}

// After a launch sequence: a leave-alone pass whose state said otherwise
// (never expected) stopped the solve; report it.
int check_pass_kinds(pdot_ctx* ctx) {
    int32_t err = 0;
    HIPCHK(ctx, hipMemcpyAsync(&err, ctx->kind_err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (err) {
        HIPCHK(ctx, hipMemsetAsync(ctx->kind_err, 0, sizeof(int32_t), ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return fail(ctx, PDOT_E_STATE, "internal: a pass launched as leaving the average alone had to update it; "
                                       "the solve was stopped (numerical_error)");
    }
    return PDOT_OK;
}

// Host copy of choose_spares.
void host_choose_spares(pdot_solver_state& st) {
    int used0 = st.cur_slot, used1 = st.avg_count > 0 ? st.avg_slot : st.cur_slot;
    int k = 0;
    for (int s = 0; s < 4; ++s) {
        if (s == used0 || s == used1) continue;
        if (k == 0) st.spare0 = s;
        else if (k == 1) st.spare1 = s;
        ++k;
    }
}

int read_state(pdot_ctx* ctx, pdot_solver_state* out) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_state, ctx->stA, sizeof(pdot_solver_state), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (out) *out = *ctx->h_state;
    return PDOT_OK;
}

int write_state(pdot_ctx* ctx, const pdot_solver_state* in) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *ctx->h_state = *in;
    HIPCHK(ctx, hipMemcpyAsync(ctx->stA, ctx->h_state, sizeof(pdot_solver_state), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PDOT_OK;
}

// Writes the average's lagging contribution (if any) into the average's slot,
// so that the slot holds the average itself; st is updated and written back.
int materialize_average(pdot_ctx* ctx, pdot_solver_state& st) {
    if (!st.avg_pending) return PDOT_OK;
    const int64_t total = ctx->ld * ctx->m;
    int rc;
    if (ctx->sparse) {  // the kernel below reads and writes every value
        if ((rc = densify_slot(ctx, st.avg_slot))) return rc;
        if ((rc = densify_slot(ctx, st.cur_slot))) return rc;
    }
    if (ctx->storage == PDOT_STORAGE_F64)
        hipLaunchKernelGGL(materialize_kernel<double>, dim3(4096), dim3(256), 0, ctx->stream,
                           static_cast<double*>(ctx->s.X[st.avg_slot]), static_cast<const double*>(ctx->s.X[st.cur_slot]),
                           total, st.pend_nw, st.pend_aw);
    else
        hipLaunchKernelGGL(materialize_kernel<float>, dim3(4096), dim3(256), 0, ctx->stream,
                           static_cast<float*>(ctx->s.X[st.avg_slot]), static_cast<const float*>(ctx->s.X[st.cur_slot]),
                           total, st.pend_nw, st.pend_aw);
    HIPCHK(ctx, hipGetLastError());
    VecAverages va;
    double* const* ys[4] = {ctx->s.u, ctx->s.r, ctx->s.v, ctx->s.c};
    for (int w = 0; w < 4; ++w) {
        va.avg[w] = ys[w][st.avg_slot];
        va.cur[w] = ys[w][st.cur_slot];
        va.len[w] = w < 2 ? ctx->n : ctx->m;
    }
    hipLaunchKernelGGL(materialize_vectors_kernel, dim3(64, 4), dim3(256), 0, ctx->stream, va, st.pend_nw, st.pend_aw);
    HIPCHK(ctx, hipGetLastError());
    st.avg_pending = 0;
    return write_state(ctx, &st);
}

int ensure_stage(pdot_ctx* ctx) {
    if (ctx->stage) return PDOT_OK;
    const int64_t max_bytes = (int64_t)256 << 20;
    int64_t cols = max_bytes / (8 * (ctx->n > 0 ? ctx->n : 1));
    if (cols < 1) cols = 1;
    if (cols > ctx->m) cols = ctx->m;
    ctx->stage_cols = cols;
    return dev_alloc(ctx, (void**)&ctx->stage, sizeof(double) * (size_t)(cols * ctx->n));
}

// Pinned staging ring for uploads: two buffers of up to 64 MiB of whole
// columns.
int ensure_pinned(pdot_ctx* ctx) {
    if (ctx->h_pin[0]) return PDOT_OK;
    const int64_t max_bytes = (int64_t)64 << 20;
    int64_t cols = max_bytes / (8 * ctx->n);
    if (cols < 1) cols = 1;
    if (cols > ctx->m) cols = ctx->m;
    ctx->pin_cols = cols;
    for (int k = 0; k < 2; ++k) {
        HIPCHK(ctx, hipHostMalloc((void**)&ctx->h_pin[k], sizeof(double) * (size_t)(cols * ctx->n), hipHostMallocDefault));
        HIPCHK(ctx, hipEventCreateWithFlags(&ctx->pin_ev[k], hipEventDisableTiming));
    }
    return PDOT_OK;
}

// memcpy split over host threads (pageable source into pinned staging).
void parallel_copy(void* dst, const void* src, size_t bytes) {
    const size_t min_piece = (size_t)4 << 20;
    unsigned hw = std::thread::hardware_concurrency();
    size_t nt = hw == 0 ? 1 : (hw < 8 ? hw : 8);
    if (bytes / min_piece < nt) nt = bytes / min_piece > 0 ? bytes / min_piece : 1;
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t piece = (bytes / nt + 63) / 64 * 64;
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nt; ++t) {
        const size_t lo = t * piece;
        if (lo >= bytes) break;
        const size_t len = (lo + piece < bytes) ? piece : bytes - lo;
        pool.emplace_back([=]() { std::memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, len); });
    }
    for (auto& th : pool) th.join();
}

double host_nanmax(double acc, double t) { return (acc != acc || t != t || t > acc) ? (acc != acc ? acc : t) : acc; }

// cached_ot_info, src/termination.jl:94-101, from the upload pass's
// per-block (max |C|, sum C^2) partials, in a fixed order.
void finish_ot_info(pdot_ctx* ctx, const std::vector<double>& parts, const double* a, const double* b) {
    double cmax = 0.0, csq = 0.0;
    for (size_t t = 0; t < parts.size(); t += 2) {
        cmax = host_nanmax(cmax, parts[t]);
        csq += parts[t + 1];
    }
    double qmax = 0.0, qsq = 0.0;
    for (int64_t i = 0; i < ctx->n; ++i) {
        qmax = host_nanmax(qmax, std::fabs(a[i]));
        qsq += a[i] * a[i];
    }
    for (int64_t j = 0; j < ctx->m; ++j) {
        qmax = host_nanmax(qmax, std::fabs(b[j]));
        qsq += b[j] * b[j];
    }
    ctx->ot_info[0] = cmax;
    ctx->ot_info[1] = qmax;
    ctx->ot_info[2] = std::sqrt(csq);
    ctx->ot_info[3] = std::sqrt(qsq);
    ctx->ot_ready = true;
}

template <typename T, int COST>
hipError_t step_occupancy(int* per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, step_kernel<T, false, COST>,
                                                        StepShape<T, COST != kCostDense>::kWaves * kWave, 0);
}

// Resident step blocks for the kernel the context will run (storage and cost
// form) and the launch geometry derived from it.
int configure_kernel(pdot_ctx* ctx) {
    int ncu = 0, per_cu = 0;
    HIPCHK(ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const bool f64 = ctx->storage == PDOT_STORAGE_F64;
    hipError_t e;
    if (ctx->cost_kind == kCostSqEuclid) e = f64 ? step_occupancy<double, kCostSqEuclid>(&per_cu) : step_occupancy<float, kCostSqEuclid>(&per_cu);
    else if (ctx->cost_kind == kCostL1) e = f64 ? step_occupancy<double, kCostL1>(&per_cu) : step_occupancy<float, kCostL1>(&per_cu);
    else e = f64 ? step_occupancy<double, kCostDense>(&per_cu) : step_occupancy<float, kCostDense>(&per_cu);
    HIPCHK(ctx, e);
    ctx->resident_blocks = (per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 1);
    return setup_geometry(ctx);
}

}  // namespace

extern "C" {

int pdot_create(pdot_ctx** out, int64_t n, int64_t m, int32_t storage, int32_t device) {
    if (!out) return PDOT_E_ARG;
    *out = nullptr;
    if (n <= 0 || m <= 0 || (storage != PDOT_STORAGE_F64 && storage != PDOT_STORAGE_F32)) return PDOT_E_ARG;
    pdot_ctx* ctx = new pdot_ctx();
    ctx->n = n;
    ctx->m = m;
    ctx->storage = storage;
    ctx->esz = storage == PDOT_STORAGE_F64 ? 8 : 4;
    ctx->ld = (n + kLdAlign - 1) / kLdAlign * kLdAlign;
    ctx->device = device;
    *out = ctx;
    HIPCHK(ctx, hipSetDevice(device));
    HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const size_t mat = ctx->esz * (size_t)ctx->ld * (size_t)m;
    int rc;
    for (int s = 0; s < 4; ++s) {
        if ((rc = dev_alloc(ctx, &ctx->s.X[s], mat))) return rc;
        if ((rc = dev_alloc(ctx, (void**)&ctx->s.u[s], sizeof(double) * (size_t)ctx->ld))) return rc;
        if ((rc = dev_alloc(ctx, (void**)&ctx->s.r[s], sizeof(double) * (size_t)ctx->ld))) return rc;
        if ((rc = dev_alloc(ctx, (void**)&ctx->s.v[s], sizeof(double) * (size_t)m))) return rc;
        if ((rc = dev_alloc(ctx, (void**)&ctx->s.c[s], sizeof(double) * (size_t)m))) return rc;
        HIPCHK(ctx, hipMemsetAsync(ctx->s.X[s], 0, mat, ctx->stream));
    }
    if ((rc = dev_alloc(ctx, (void**)&ctx->a, sizeof(double) * (size_t)ctx->ld))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->b, sizeof(double) * (size_t)m))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->evalout, sizeof(double) * kEvalOut))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->stA, sizeof(pdot_solver_state)))) return rc;
    if ((rc = dev_alloc(ctx, (void**)&ctx->kind_err, sizeof(int32_t)))) return rc;
    HIPCHK(ctx, hipMemsetAsync(ctx->kind_err, 0, sizeof(int32_t), ctx->stream));
    if ((rc = dev_alloc(ctx, (void**)&ctx->done, sizeof(uint32_t)))) return rc;
    HIPCHK(ctx, hipMemsetAsync(ctx->done, 0, sizeof(uint32_t), ctx->stream));
    HIPCHK(ctx, hipHostMalloc((void**)&ctx->h_state, sizeof(pdot_solver_state), hipHostMallocDefault));
    HIPCHK(ctx, hipHostMalloc((void**)&ctx->h_eval, sizeof(double) * kEvalOut, hipHostMallocDefault));
    if ((rc = configure_kernel(ctx))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PDOT_OK;
}

void pdot_destroy(pdot_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    free_graphs(ctx);
    for (hipEvent_t e : ctx->events) hipEventDestroy(e);
    free_partials(ctx);
    free_maps(ctx);
    void* ps[] = {ctx->C, ctx->a, ctx->b, ctx->evalout, ctx->stA, ctx->kind_err, ctx->done, ctx->stage};
    for (void* q : ps)
        if (q) hipFree(q);
    for (int s = 0; s < 4; ++s) {
        void* qs[] = {ctx->s.X[s], ctx->s.u[s], ctx->s.v[s], ctx->s.r[s], ctx->s.c[s]};
        for (void* q : qs)
            if (q) hipFree(q);
    }
    if (ctx->h_state) hipHostFree(ctx->h_state);
    if (ctx->h_eval) hipHostFree(ctx->h_eval);
    for (int k = 0; k < 2; ++k) {
        if (ctx->h_pin[k]) hipHostFree(ctx->h_pin[k]);
        if (ctx->pin_ev[k]) hipEventDestroy(ctx->pin_ev[k]);
    }
    if (ctx->ot_part) hipFree(ctx->ot_part);
    if (ctx->px) hipFree(ctx->px);
    if (ctx->py) hipFree(ctx->py);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* pdot_last_error(const pdot_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pdot_upload(pdot_ctx* ctx, const double* C, const double* a, const double* b) {
    if (!ctx) return PDOT_E_ARG;
    if (!C || !a || !b) return fail(ctx, PDOT_E_ARG, "pdot_upload: null pointer");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int64_t n = ctx->n, m = ctx->m, ld = ctx->ld;
    const bool f32 = ctx->storage == PDOT_STORAGE_F32;
    ctx->ot_ready = false;
    int rc;
    if (!ctx->C) {  // the cost matrix is allocated by the first dense upload
        const size_t mat = ctx->esz * (size_t)ld * (size_t)m;
        if ((rc = dev_alloc(ctx, &ctx->C, mat))) return rc;
        HIPCHK(ctx, hipMemsetAsync(ctx->C, 0, mat, ctx->stream));
    }
    if (ctx->cost_kind != kCostDense) {
        ctx->cost_kind = kCostDense;
        if ((rc = configure_kernel(ctx))) return rc;
    }
    if ((rc = ensure_pinned(ctx))) return rc;
    if (f32 && (rc = ensure_stage(ctx))) return rc;
    const int64_t cols = f32 ? (ctx->pin_cols < ctx->stage_cols ? ctx->pin_cols : ctx->stage_cols) : ctx->pin_cols;
    const int64_t nchunks = (m + cols - 1) / cols;
    const int64_t G = cols < kIngestBlocks ? cols : kIngestBlocks;
    if (ctx->ot_part) {
        hipFree(ctx->ot_part);
        ctx->bytes -= ctx->ot_part_bytes;
        ctx->ot_part = nullptr;
    }
    ctx->ot_part_bytes = (int64_t)(sizeof(double) * 2 * G * nchunks);
    if ((rc = dev_alloc(ctx, (void**)&ctx->ot_part, (size_t)ctx->ot_part_bytes))) return rc;

    // Pipeline: the host copies chunk k into pinned buffer k % 2 while the DMA
    // of chunk k - 1 runs; a buffer is refilled once its copy has completed.
    bool used[2] = {false, false};
    for (int64_t k = 0; k < nchunks; ++k) {
        const int64_t j0 = k * cols;
        const int64_t nc = (j0 + cols < m) ? cols : m - j0;
        const int b2 = (int)(k & 1);
        if (used[b2]) HIPCHK(ctx, hipEventSynchronize(ctx->pin_ev[b2]));
        parallel_copy(ctx->h_pin[b2], C + j0 * n, sizeof(double) * (size_t)(n * nc));
        double* part = ctx->ot_part + 2 * G * k;
        if (!f32) {
            double* dst = static_cast<double*>(ctx->C) + j0 * ld;
            HIPCHK(ctx, hipMemcpy2DAsync(dst, ld * 8, ctx->h_pin[b2], n * 8, n * 8, nc, hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(ctx, hipEventRecord(ctx->pin_ev[b2], ctx->stream));
            hipLaunchKernelGGL(ingest_kernel<false>, dim3((unsigned)G), dim3(256), 0, ctx->stream, dst, ld,
                               (float*)nullptr, ld, n, nc, part);
        } else {
            HIPCHK(ctx, hipMemcpyAsync(ctx->stage, ctx->h_pin[b2], sizeof(double) * (size_t)(n * nc), hipMemcpyHostToDevice,
                                       ctx->stream));
            HIPCHK(ctx, hipEventRecord(ctx->pin_ev[b2], ctx->stream));
            hipLaunchKernelGGL(ingest_kernel<true>, dim3((unsigned)G), dim3(256), 0, ctx->stream, ctx->stage, n,
                               static_cast<float*>(ctx->C) + j0 * ld, ld, n, nc, part);
        }
        HIPCHK(ctx, hipGetLastError());
        used[b2] = true;
    }
    HIPCHK(ctx, hipMemsetAsync(ctx->a, 0, sizeof(double) * ld, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->a, a, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->b, b, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
    std::vector<double> parts((size_t)(2 * G * nchunks));
    HIPCHK(ctx, hipMemcpyAsync(parts.data(), ctx->ot_part, sizeof(double) * parts.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));

    finish_ot_info(ctx, parts, a, b);
    ctx->uploaded = true;
    return PDOT_OK;
}

int pdot_upload_points(pdot_ctx* ctx, int32_t cost_kind, int32_t dim, const double* x, const double* y,
                       int32_t normalize, const double* a, const double* b) {
    if (!ctx) return PDOT_E_ARG;
    if (!x || !y || !a || !b) return fail(ctx, PDOT_E_ARG, "pdot_upload_points: null pointer");
    if (cost_kind != PDOT_COST_SQEUCLID && cost_kind != PDOT_COST_L1)
        return fail(ctx, PDOT_E_ARG, "pdot_upload_points: unknown cost kind");
    if (dim < 1 || dim > kMaxDim) return fail(ctx, PDOT_E_ARG, "pdot_upload_points: dim must be 1..3");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int64_t n = ctx->n, m = ctx->m, ld = ctx->ld;
    ctx->ot_ready = false;
    int rc;
    if (ctx->C) {  // no dense matrix in point mode
        hipFree(ctx->C);
        ctx->bytes -= (int64_t)(ctx->esz * (size_t)ld * (size_t)m);
        ctx->C = nullptr;
        free_graphs(ctx);
    }
    if (!ctx->px && (rc = dev_alloc(ctx, (void**)&ctx->px, sizeof(double) * (size_t)(kMaxDim * n)))) return rc;
    if (!ctx->py && (rc = dev_alloc(ctx, (void**)&ctx->py, sizeof(double) * (size_t)(kMaxDim * m)))) return rc;
    std::vector<double> xt((size_t)(dim * n));
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < dim; ++d) xt[(size_t)(d * n + i)] = x[i * dim + d];
    HIPCHK(ctx, hipMemcpyAsync(ctx->px, xt.data(), sizeof(double) * xt.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->py, y, sizeof(double) * (size_t)(m * dim), hipMemcpyHostToDevice, ctx->stream));
    const int cost = cost_kind == PDOT_COST_SQEUCLID ? kCostSqEuclid : kCostL1;
    // the norms of the matrix the points define, in the chunk order of a dense upload
    if ((rc = ensure_pinned(ctx))) return rc;
    const int64_t cols = ctx->pin_cols;
    const int64_t nchunks = (m + cols - 1) / cols;
    const int64_t G = cols < kIngestBlocks ? cols : kIngestBlocks;
    if (ctx->ot_part) {
        hipFree(ctx->ot_part);
        ctx->bytes -= ctx->ot_part_bytes;
        ctx->ot_part = nullptr;
    }
    ctx->ot_part_bytes = (int64_t)(sizeof(double) * 2 * G * nchunks);
    if ((rc = dev_alloc(ctx, (void**)&ctx->ot_part, (size_t)ctx->ot_part_bytes))) return rc;
    std::vector<double> parts((size_t)(2 * G * nchunks));
    auto stats_pass = [&](double cdiv) -> int {
        for (int64_t k = 0; k < nchunks; ++k) {
            const int64_t j0 = k * cols;
            const int64_t nc = (j0 + cols < m) ? cols : m - j0;
            if (cost == kCostSqEuclid)
                hipLaunchKernelGGL(point_stats_kernel<kCostSqEuclid>, dim3((unsigned)G), dim3(256), 0, ctx->stream, ctx->px,
                                   ctx->py, n, (int)dim, cdiv, j0, nc, ctx->ot_part + 2 * G * k);
            else
                hipLaunchKernelGGL(point_stats_kernel<kCostL1>, dim3((unsigned)G), dim3(256), 0, ctx->stream, ctx->px,
                                   ctx->py, n, (int)dim, cdiv, j0, nc, ctx->ot_part + 2 * G * k);
            HIPCHK(ctx, hipGetLastError());
        }
        HIPCHK(ctx, hipMemcpyAsync(parts.data(), ctx->ot_part, sizeof(double) * parts.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return PDOT_OK;
    };
    double cdiv = 1.0;
    if (normalize) {  // C = raw / max(raw), as a dense caller divides by C.max()
        if ((rc = stats_pass(1.0))) return rc;
        double mx = 0.0;
        for (size_t t = 0; t < parts.size(); t += 2) mx = host_nanmax(mx, parts[t]);
        if (!(mx > 0.0)) return fail(ctx, PDOT_E_ARG, "pdot_upload_points: cannot normalise a zero cost");
        // FP32 storage scales the cost as raw * (1 / max), which needs a normal reciprocal
        if (ctx->storage == PDOT_STORAGE_F32 && std::isfinite(mx) && (mx < 0x1p-1022 || mx > 0x1p+1022))
            return fail(ctx, PDOT_E_ARG,
                        "pdot_upload_points: FP32 storage needs the normalising maximum cost in [2^-1022, 2^1022]");
        cdiv = mx;
    }
    if ((rc = stats_pass(cdiv))) return rc;
    HIPCHK(ctx, hipMemsetAsync(ctx->a, 0, sizeof(double) * ld, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->a, a, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->b, b, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    finish_ot_info(ctx, parts, a, b);
    ctx->cost_kind = cost;
    ctx->dim = dim;
    ctx->cdiv = cdiv;
    if ((rc = configure_kernel(ctx))) return rc;  // the point kernel's shape and occupancy
    ctx->uploaded = true;
    return PDOT_OK;
}

int pdot_cached_ot_info(pdot_ctx* ctx, double* out4) {
    if (!ctx || !out4) return PDOT_E_ARG;
    if (!ctx->ot_ready) return fail(ctx, PDOT_E_STATE, "pdot_cached_ot_info before pdot_upload");
    for (int k = 0; k < 4; ++k) out4[k] = ctx->ot_info[k];
    return PDOT_OK;
}

int pdot_set_start(pdot_ctx* ctx, const double* u0, const double* v0, int32_t step_policy, double reduction_exponent,
                   double growth_exponent, double step_size, double primal_weight) {
    if (!ctx) return PDOT_E_ARG;
    if (!ctx->uploaded) return fail(ctx, PDOT_E_STATE, "pdot_set_start before pdot_upload");
    if ((u0 == nullptr) != (v0 == nullptr)) return fail(ctx, PDOT_E_ARG, "pdot_set_start: give both u0 and v0 or neither");
    if (step_policy != PDOT_POLICY_ADAPTIVE && step_policy != PDOT_POLICY_CONSTANT)
        return fail(ctx, PDOT_E_ARG, "pdot_set_start: unknown step policy");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t mat = ctx->esz * (size_t)ctx->ld * (size_t)ctx->m;
    int rc;
    for (int s = 0; s < 4; ++s) {
        HIPCHK(ctx, hipMemsetAsync(ctx->s.X[s], 0, mat, ctx->stream));
        if ((rc = fill_map(ctx, s, 0))) return rc;  // every row +0.0
        HIPCHK(ctx, hipMemsetAsync(ctx->s.u[s], 0, sizeof(double) * ctx->ld, ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(ctx->s.r[s], 0, sizeof(double) * ctx->ld, ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(ctx->s.v[s], 0, sizeof(double) * ctx->m, ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(ctx->s.c[s], 0, sizeof(double) * ctx->m, ctx->stream));
    }
    if (u0) {
        HIPCHK(ctx, hipMemcpyAsync(ctx->s.u[0], u0, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(ctx->s.v[0], v0, sizeof(double) * ctx->m, hipMemcpyHostToDevice, ctx->stream));
    }
    pdot_solver_state st;
    memset(&st, 0, sizeof(st));
    st.step_size = step_size;
    st.primal_weight = primal_weight;
    st.reduction_exponent = reduction_exponent;
    st.growth_exponent = growth_exponent;
    st.step_policy = step_policy;
    st.gzero = u0 ? 0 : 1;
    st.cur_slot = 0;
    st.avg_slot = 0;
    st.spare0 = 1;
    st.spare1 = 2;
    st.defer_ok = ctx->defer ? 1 : 0;
    if ((rc = write_state(ctx, &st))) return rc;
    ctx->started = true;
    return PDOT_OK;
}

int pdot_run_steps(pdot_ctx* ctx, int64_t count, pdot_solver_state* out) {
    if (!ctx) return PDOT_E_ARG;
    if (!ctx->started) return fail(ctx, PDOT_E_STATE, "pdot_run_steps before pdot_set_start");
    if (!ctx->evalvec) return fail(ctx, PDOT_E_STATE, "pdot_run_steps: launch buffers missing after a failed allocation");
    if (count < 0) return fail(ctx, PDOT_E_ARG, "pdot_run_steps: negative count");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    pdot_solver_state st0;
    int rc0 = read_state(ctx, &st0);
    if (rc0) return rc0;
    const bool first = first_pass_leaves(st0), dok = st0.defer_ok != 0;
    if (ctx->use_graph && count > 1) {
        // the kinds of the passes are part of the captured sequence
        const int64_t key = count * 4 + (first ? 2 : 0) + (dok ? 1 : 0);
        auto it = ctx->graphs.find(key);
        if (it == ctx->graphs.end()) {
            hipGraph_t g;
            HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_steps(ctx, count, first, dok);
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return fail(ctx, PDOT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            hipGraphExec_t ge;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return fail(ctx, PDOT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            it = ctx->graphs.emplace(key, ge).first;
        }
        HIPCHK(ctx, hipGraphLaunch(it->second, ctx->stream));
    } else {
        int rc = enqueue_steps(ctx, count, first, dok);
        if (rc) return rc;
    }
    int rc = check_pass_kinds(ctx);
    if (rc) return rc;
    return read_state(ctx, out);
}

int pdot_eval(pdot_ctx* ctx, pdot_point_stats* current, pdot_point_stats* average, pdot_solver_state* state) {
    if (!ctx) return PDOT_E_ARG;
    if (!ctx->started) return fail(ctx, PDOT_E_STATE, "pdot_eval before pdot_set_start");
    if (!ctx->evalvec) return fail(ctx, PDOT_E_STATE, "pdot_eval: launch buffers missing after a failed allocation");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_commit(ctx));
    HIPCHK(ctx, launch_step(ctx, true));
    HIPCHK(ctx, launch_reduce(ctx, true));
    hipLaunchKernelGGL(evalfin_kernel, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->evalpart, (int64_t)ctx->nstep,
                       ctx->evalvec, (int64_t)ctx->nred, ctx->evalout);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_eval, ctx->evalout, sizeof(double) * kEvalOut, hipMemcpyDeviceToHost, ctx->stream));
    pdot_solver_state st;
    int rc = read_state(ctx, &st);
    if (rc) return rc;
    const double* e = ctx->h_eval;
    const double* v = e + kMatStats;
    if (current) {
        current->primal_objective = e[0];
        current->x_inf = e[1];
        current->x_sq = e[2];
        current->rc_inf = e[3];
        current->rc_sq = e[4];
        current->p_inf = v[0];
        current->p_sq = v[1];
        current->dual_objective = v[2];
        current->u_inf = v[3];
        current->u_sq = v[4];
        current->y_sq = v[5];
    }
    if (average) {
        if (st.avg_count == 0) {
            memset(average, 0, sizeof(*average));
        } else {
            average->primal_objective = e[5];
            average->x_inf = e[6];
            average->x_sq = e[7];
            average->rc_inf = e[8];
            average->rc_sq = e[9];
            average->p_inf = v[6];
            average->p_sq = v[7];
            average->dual_objective = v[8];
            average->u_inf = v[9];
            average->u_sq = v[10];
            average->y_sq = v[11];
        }
    }
    if (state) *state = st;
    return PDOT_OK;
}

int pdot_discard(pdot_ctx* ctx) {
    if (!ctx) return PDOT_E_ARG;
    pdot_solver_state st;
    int rc = read_state(ctx, &st);
    if (rc) return rc;
    if (!st.pending) return PDOT_OK;
    st.pending = 0;
    return write_state(ctx, &st);
}

int pdot_restart(pdot_ctx* ctx, int32_t to_average, double new_primal_weight) {
    if (!ctx) return PDOT_E_ARG;
    if (!ctx->started) return fail(ctx, PDOT_E_STATE, "pdot_restart before pdot_set_start");
    pdot_solver_state st;
    int rc = read_state(ctx, &st);
    if (rc) return rc;
    st.pending = 0;
    if (to_average) {
        if (st.avg_count == 0) return fail(ctx, PDOT_E_STATE, "pdot_restart: restart to an empty average");
        if ((rc = materialize_average(ctx, st))) return rc;
        st.cur_slot = st.avg_slot;
        st.gzero = 0;
    }
    st.avg_count = 0;
    st.avg_pending = 0;  // the reset average has no lagging contribution
    st.sum_weights = 0.0;
    st.avg_slot = st.cur_slot;
    st.primal_weight = new_primal_weight;
    host_choose_spares(st);
    return write_state(ctx, &st);
}

int pdot_get_state(pdot_ctx* ctx, pdot_solver_state* out) {
    if (!ctx || !out) return PDOT_E_ARG;
    return read_state(ctx, out);
}

int pdot_set_state(pdot_ctx* ctx, const pdot_solver_state* in) {
    if (!ctx || !in) return PDOT_E_ARG;
    if (in->cur_slot < 0 || in->cur_slot > 3 || in->avg_slot < 0 || in->avg_slot > 3 || in->spare0 < 0 ||
        in->spare0 > 3 || in->spare1 < 0 || in->spare1 > 3)
        return fail(ctx, PDOT_E_ARG, "pdot_set_state: slot index out of range");
    if ((in->defer_ok != 0 && in->defer_ok != 1) || (in->avg_pending != 0 && in->avg_pending != 1))
        return fail(ctx, PDOT_E_ARG, "pdot_set_state: defer_ok and avg_pending must be 0 or 1");
    if (in->avg_pending && (in->avg_count == 0 || in->avg_slot == in->cur_slot))
        return fail(ctx, PDOT_E_ARG, "pdot_set_state: a lagging contribution needs a non-empty average in its own slot");
    return write_state(ctx, in);
}

int pdot_download(pdot_ctx* ctx, int32_t which, double* X, double* u, double* v) {
    if (!ctx) return PDOT_E_ARG;
    if (!ctx->started) return fail(ctx, PDOT_E_STATE, "pdot_download before pdot_set_start");
    pdot_solver_state st;
    int rc = read_state(ctx, &st);
    if (rc) return rc;
    int slot;
    if (which == PDOT_POINT_CURRENT) slot = st.cur_slot;
    else if (which == PDOT_POINT_AVERAGE) {
        if (st.avg_count == 0) return fail(ctx, PDOT_E_STATE, "pdot_download: the average is empty");
        if ((rc = materialize_average(ctx, st))) return rc;
        slot = st.avg_slot;
    } else return fail(ctx, PDOT_E_ARG, "pdot_download: unknown point");
    if (X && ctx->sparse && (rc = densify_slot(ctx, slot))) return rc;
    const int64_t n = ctx->n, m = ctx->m, ld = ctx->ld;
    if (X) {
        if (ctx->storage == PDOT_STORAGE_F64) {
            HIPCHK(ctx, hipMemcpy2DAsync(X, n * 8, ctx->s.X[slot], ld * 8, n * 8, m, hipMemcpyDeviceToHost, ctx->stream));
        } else {
            if ((rc = ensure_stage(ctx))) return rc;
            for (int64_t j0 = 0; j0 < m; j0 += ctx->stage_cols) {
                const int64_t nc = (j0 + ctx->stage_cols < m) ? ctx->stage_cols : m - j0;
                hipLaunchKernelGGL(widen_kernel<float>, dim3(2048), dim3(256), 0, ctx->stream,
                                   static_cast<const float*>(ctx->s.X[slot]) + j0 * ld, ctx->stage, n, ld, nc);
                HIPCHK(ctx, hipGetLastError());
                HIPCHK(ctx, hipMemcpyAsync(X + j0 * n, ctx->stage, sizeof(double) * n * nc, hipMemcpyDeviceToHost, ctx->stream));
            }
        }
    }
    if (u) HIPCHK(ctx, hipMemcpyAsync(u, ctx->s.u[slot], sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (v) HIPCHK(ctx, hipMemcpyAsync(v, ctx->s.v[slot], sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PDOT_OK;
}

int pdot_round(pdot_ctx* ctx, int32_t which, double* X_rounded) {
    if (!ctx || !X_rounded) return PDOT_E_ARG;
    if (!ctx->started) return fail(ctx, PDOT_E_STATE, "pdot_round before pdot_set_start");
    pdot_solver_state st;
    int rc = read_state(ctx, &st);
    if (rc) return rc;
    int slot;
    if (which == PDOT_POINT_CURRENT) slot = st.cur_slot;
    else if (which == PDOT_POINT_AVERAGE) {
        if (st.avg_count == 0) return fail(ctx, PDOT_E_STATE, "pdot_round: the average is empty");
        if ((rc = materialize_average(ctx, st))) return rc;
        slot = st.avg_slot;
    } else return fail(ctx, PDOT_E_ARG, "pdot_round: unknown point");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (ctx->sparse && (rc = densify_slot(ctx, slot))) return rc;
    double* Y = nullptr;
    if ((rc = dev_alloc(ctx, (void**)&Y, sizeof(double) * (size_t)(ctx->n * ctx->m)))) return rc;
    ctx->bytes -= (int64_t)(sizeof(double) * (size_t)(ctx->n * ctx->m));  // released below
    rc = pdot_round_plan(ctx->stream, ctx->s.X[slot], ctx->storage, ctx->ld, ctx->n, ctx->m, ctx->a, ctx->b, Y,
                         ctx->n, &ctx->err);
    if (rc == PDOT_OK) {
        hipError_t e = hipMemcpyAsync(X_rounded, Y, sizeof(double) * (size_t)(ctx->n * ctx->m), hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, PDOT_E_HIP, std::string("pdot_round: copy: ") + hipGetErrorString(e));
    }
    hipFree(Y);
    return rc;
}

int pdot_time_steps(pdot_ctx* ctx, int64_t count, double* step_kernel_ms, double* total_ms) {
    if (!ctx) return PDOT_E_ARG;
    if (!ctx->started) return fail(ctx, PDOT_E_STATE, "pdot_time_steps before pdot_set_start");
    if (!ctx->evalvec) return fail(ctx, PDOT_E_STATE, "pdot_time_steps: launch buffers missing after a failed allocation");
    if (count < 1) return fail(ctx, PDOT_E_ARG, "pdot_time_steps: count < 1");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t need = (size_t)(2 * count + 2);
    while (ctx->events.size() < need) {
        hipEvent_t e;
        HIPCHK(ctx, hipEventCreate(&e));
        ctx->events.push_back(e);
    }
    hipEvent_t* ev = ctx->events.data();
    pdot_solver_state st0;
    int rc0 = read_state(ctx, &st0);
    if (rc0) return rc0;
    const bool first = first_pass_leaves(st0), dok = st0.defer_ok != 0;
    HIPCHK(ctx, hipEventRecord(ev[0], ctx->stream));
    HIPCHK(ctx, launch_commit(ctx));
    for (int64_t k = 0; k < count; ++k) {
        HIPCHK(ctx, hipEventRecord(ev[2 + 2 * k], ctx->stream));
        const bool leave = pass_leaves(first, dok, k);
        HIPCHK(ctx, launch_step(ctx, false, leave));
        HIPCHK(ctx, hipEventRecord(ev[3 + 2 * k], ctx->stream));
        HIPCHK(ctx, launch_reduce(ctx, false));
    }
    HIPCHK(ctx, hipEventRecord(ev[1], ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ev[1]));
    if ((rc0 = check_pass_kinds(ctx))) return rc0;
    double ks = 0.0;
    for (int64_t k = 0; k < count; ++k) {
        float t = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&t, ev[2 + 2 * k], ev[3 + 2 * k]));
        ks += t;
    }
    float tt = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&tt, ev[0], ev[1]));
    if (step_kernel_ms) *step_kernel_ms = ks;
    if (total_ms) *total_ms = tt;
    return PDOT_OK;
}

int pdot_use_graphs(pdot_ctx* ctx, int32_t enable) {
    if (!ctx) return PDOT_E_ARG;
    ctx->use_graph = enable != 0;
    return PDOT_OK;
}

int pdot_set_defer(pdot_ctx* ctx, int32_t enable) {
    if (!ctx) return PDOT_E_ARG;
    if (enable != 0 && enable != 1) return fail(ctx, PDOT_E_ARG, "pdot_set_defer: enable must be 0 or 1");
    ctx->defer = enable != 0;
    if (!ctx->started) return PDOT_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    pdot_solver_state st;
    int rc = read_state(ctx, &st);
    if (rc) return rc;
    if (!enable && (rc = materialize_average(ctx, st))) return rc;
    st.defer_ok = enable;
    return write_state(ctx, &st);
}

int pdot_set_pass_kernels(pdot_ctx* ctx, int32_t enable) {
    if (!ctx) return PDOT_E_ARG;
    if (enable != 0 && enable != 1) return fail(ctx, PDOT_E_ARG, "pdot_set_pass_kernels: enable must be 0 or 1");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pass_kernels = enable != 0;
    free_graphs(ctx);  // captured sequences hold the kernels
    return PDOT_OK;
}

int pdot_set_stream_policy(pdot_ctx* ctx, int32_t mode) {
    if (!ctx) return PDOT_E_ARG;
    if (mode < -1 || mode > 1) return fail(ctx, PDOT_E_ARG, "pdot_set_stream_policy: mode must be -1, 0 or 1");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream_policy = mode;
    free_graphs(ctx);  // captured sequences hold the kernels
    return PDOT_OK;
}

int pdot_set_sparse(pdot_ctx* ctx, int32_t enable) {
    if (!ctx) return PDOT_E_ARG;
    if (enable != 0 && enable != 1) return fail(ctx, PDOT_E_ARG, "pdot_set_sparse: enable must be 0 or 1");
    if ((enable != 0) == ctx->sparse) return PDOT_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc;
    if (enable) {
        // dense passes kept no maps: every bit set, then kept from here on
        for (int s = 0; s < 4; ++s)
            if ((rc = fill_map(ctx, s, ~0ull))) return rc;
    } else {
        if ((rc = densify_all(ctx))) return rc;
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    free_graphs(ctx);  // the captured passes carry the mode
    ctx->sparse = enable != 0;
    return PDOT_OK;
}

int pdot_zero_share(pdot_ctx* ctx, int32_t which, double* share) {
    if (!ctx || !share) return PDOT_E_ARG;
    if (!ctx->started) return fail(ctx, PDOT_E_STATE, "pdot_zero_share before pdot_set_start");
    pdot_solver_state st;
    int rc = read_state(ctx, &st);
    if (rc) return rc;
    int slot;
    if (which == PDOT_POINT_CURRENT) slot = st.cur_slot;
    else if (which == PDOT_POINT_AVERAGE) {
        if (st.avg_count == 0) return fail(ctx, PDOT_E_STATE, "pdot_zero_share: the average is empty");
        slot = st.avg_slot;  // as the next averaging pass reads it (a lagging contribution stays lagging)
    } else return fail(ctx, PDOT_E_ARG, "pdot_zero_share: unknown point");
    *share = 0.0;
    if (!ctx->sparse || ctx->zg.words == 0) return PDOT_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    unsigned long long* d = nullptr;
    if ((rc = dev_alloc(ctx, (void**)&d, sizeof(unsigned long long)))) return rc;
    ctx->bytes -= (int64_t)sizeof(unsigned long long);  // released below
    unsigned long long live = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(live), ctx->stream);
    if (e == hipSuccess) {
        const int64_t blocks = std::min<int64_t>((ctx->zg.words + 255) / 256, 2048);
        hipLaunchKernelGGL(zcount_kernel, dim3(blocks), dim3(256), 0, ctx->stream, ctx->zmap[slot], ctx->zg, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&live, d, sizeof(live), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(ctx, PDOT_E_HIP, std::string("pdot_zero_share: ") + hipGetErrorString(e));
    const int64_t vec = 16 / ctx->esz;
    const double units = (double)((ctx->n + vec - 1) / vec) * (double)ctx->m;
    *share = 1.0 - (double)live / units;
    return PDOT_OK;
}

int pdot_set_controller_fold(pdot_ctx* ctx, int32_t mode) {
    if (!ctx) return PDOT_E_ARG;
    if (mode < -1 || mode > 1) return fail(ctx, PDOT_E_ARG, "pdot_set_controller_fold: mode must be -1, 0 or 1");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->fold_mode = mode;
    free_graphs(ctx);  // captured sequences hold the kernels
    return PDOT_OK;
}

int pdot_set_launch(pdot_ctx* ctx, int32_t col_blocks_hint, int32_t interleave) {
    if (!ctx) return PDOT_E_ARG;
    if (col_blocks_hint < 0) return fail(ctx, PDOT_E_ARG, "pdot_set_launch: negative hint");
    if (interleave != 0 && interleave != 1) return fail(ctx, PDOT_E_ARG, "pdot_set_launch: interleave must be 0 or 1");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    // the pending pass (if any) keeps its partials only while the geometry is unchanged
    pdot_solver_state st;
    if (ctx->started) {
        int rc = read_state(ctx, &st);
        if (rc) return rc;
        if (st.pending) return fail(ctx, PDOT_E_STATE, "pdot_set_launch: a pass is pending");
    }
    ctx->cb_hint = col_blocks_hint;
    ctx->interleave = interleave;
    return setup_geometry(ctx);
}

int pdot_get_launch(pdot_ctx* ctx, int32_t* rows_per_blk, int32_t* col_blocks, int32_t* cols_per_block, int32_t* blocks) {
    if (!ctx) return PDOT_E_ARG;
    if (rows_per_blk) *rows_per_blk = rows_per_block(ctx->storage, ctx->cost_kind);
    if (col_blocks) *col_blocks = ctx->Cb;
    if (cols_per_block) *cols_per_block = ctx->TC;
    if (blocks) *blocks = ctx->nstep;
    return PDOT_OK;
}

int64_t pdot_device_bytes(const pdot_ctx* ctx) { return ctx ? ctx->bytes : 0; }

int pdot_copy_probe(pdot_ctx* ctx, int64_t bytes, int32_t reps, double* ms_per_copy) {
    if (!ctx) return PDOT_E_ARG;
    if (bytes <= 0 || bytes % 16 != 0 || reps < 1) return fail(ctx, PDOT_E_ARG, "pdot_copy_probe: bytes must be a positive multiple of 16");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    void *src = nullptr, *dst = nullptr;
    int rc;
    if ((rc = dev_alloc(ctx, &src, (size_t)bytes))) return rc;
    if ((rc = dev_alloc(ctx, &dst, (size_t)bytes))) {
        hipFree(src);
        return rc;
    }
    HIPCHK(ctx, hipMemsetAsync(src, 0, (size_t)bytes, ctx->stream));
    const int64_t nv = bytes / 16;
    const unsigned grid = (unsigned)((nv + 255) / 256);
    hipEvent_t e0, e1;
    HIPCHK(ctx, hipEventCreate(&e0));
    HIPCHK(ctx, hipEventCreate(&e1));
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const dbl2*)src, (dbl2*)dst, nv);
    HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const dbl2*)src, (dbl2*)dst, nv);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(src);
    hipFree(dst);
    ctx->bytes -= 2 * bytes;
    if (ms_per_copy) *ms_per_copy = ms / reps;
    return PDOT_OK;
}

}  // extern "C"
