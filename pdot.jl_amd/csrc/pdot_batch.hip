// pdot_batch.hip -- many independent transport problems of uneven n x m
// advanced together on one MI355X (the batched form of the PDHG hot path).
//
// Layout: every instance has its own C, four point slots (X, u, v, r, c) and
// partial-sum buffers, carved out of four arenas (matrices in the storage
// type; vectors, per-pass partials and the per-instance pass reductions in
// FP64).  Per-instance descriptors (BInst) and solver states
// (pdot_solver_state) live in device arrays.
//
// Work lists: a pass over the batch is a list of step tiles (instance, row
// block, column block) and a list of reduce chunks (instance, 64 rows or
// columns).  The host rebuilds both lists only when the set of active
// instances changes; their lengths are read by the kernels from device
// memory, so a captured hipGraph of a whole evaluation interval stays valid.
// The step and reduce kernels are persistent: a fixed grid of resident
// blocks walks the lists (tiles sorted by size, largest first).
//
// Per pass and instance the arithmetic is exactly the single-instance path's
// (step_tile and reduce_chunk of pdot_common.h).  Each instance carries its
// own step size, primal weight, averaging weights, restart and termination
// flags, and a step budget: a pass is performed only while the budget is
// positive, so instances advance by different numbers of passes in one
// launch sequence.  After the reduce kernel, bcontroller_kernel runs each active
// instance's controller over its chunks' scalars (one wave per instance; the
// kernel boundary publishes them, so no chunk runs a device-scope fence, an
// L2 write-back and invalidate on this GPU).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pdot.h"
#include "pdot_common.h"

namespace {

// Batched step kernel: waves per block, each block one tile of 64 x VEC rows
// (256 in FP32).  Instances have n in the tens to thousands of rows; short
// tiles waste fewer lanes on the last row tile of an instance.  Measured at
// the 512-instance workload: one wave 2.35 ms per all-active pass, two
// 2.41, four 2.39; with about a dozen instances left 44, 45 and 53 us.
#ifndef PDOT_BATCH_WAVES
#define PDOT_BATCH_WAVES 1
#endif
constexpr int kBWaves = PDOT_BATCH_WAVES;
#ifndef PDOT_BATCH_PIPE
#define PDOT_BATCH_PIPE 0
#endif
constexpr int kBPipe = PDOT_BATCH_PIPE;  // column batches prefetched ahead in a tile (1 measured slower)
#ifndef PDOT_BATCH_UNROLL
#define PDOT_BATCH_UNROLL 2
#endif
constexpr int kBUnroll = PDOT_BATCH_UNROLL;  // columns per load batch (1 slower, 4 no faster)
constexpr int kBBlock = kWave * kBWaves;
// batched reduce kernel: waves per block (a chunk with many partials is
// split over them).  8 against 4: 16.3 against 18.1 us per pass with one
// instance left, all-active passes unchanged, the 512-instance batch
// 32.1 s against 33.2 s; 16 slower.
constexpr int kBRedWaves = 8;
constexpr int kBRedBlock = kWave * kBRedWaves;
constexpr int kMinColTiles = 32;           // column tiles per instance (at least)
constexpr int kHeavyParts = 32;            // chunks with more partials are reduced by a whole block

// columns per tile, at least.  8 makes tail passes about 0.5 us faster, but
// the tail instances' partial sums group differently and the slowest
// instance of the 512-instance batch then needs 824,448 iterations instead
// of 651,904 (its restart decisions are that sensitive to rounding).
constexpr int kMinTileCols = 4;

struct BInst {
    int64_t n, m, ld;
    int64_t C;                // element offsets into the matrix arena
    int64_t X[4];
    int64_t u[4], r[4];       // double offsets into the vector arena (length ld)
    int64_t v[4], c[4];       // (length m)
    int64_t a, b;
    int64_t rowpart, colpart, sqpart, evalpart;  // offsets into the partial arena
    int64_t Z[4];             // word offsets of each slot's zero map (R * waves * m words)
    int64_t redpart, evalvec;                    // offsets into the reduction arena
    int32_t R, Cb, TC, nstep, NBr, NBc, nred, sqchunk;
};

struct BTile {
    int32_t inst, rb, cb, pad;
};
struct BChunk {
    int32_t inst, rows, chunk, pad;
};

struct BParams {
    void* mat;
    double* vec;
    double* part;  // per-pass partial sums (layout follows the active set)
    double* red;   // per-instance pass reductions and eval vector partials (fixed layout)
    const BInst* inst;
    pdot_solver_state* st;
    const BTile* tiles;
    const BChunk* chunks;
    const int32_t* nwork;  // {ntiles, heavy chunks, light chunks, active instances}
    const int32_t* actlist;  // the active instances
    double* evalout;       // kEvalOut per instance
    uint64_t* zmap;        // the zero maps of every instance and slot
    int32_t sparse;        // passes read and write through them
};

template <typename T>
__device__ __forceinline__ void slot_ptrs(const BParams& p, const BInst& in, const pdot_solver_state& s,
                                          TileArgs<T>& a) {
    T* mat = static_cast<T*>(p.mat);
    const bool avg_empty = s.avg_count == 0;
    const int cs = s.cur_slot, as = avg_empty ? cs : s.avg_slot;
    a.C = mat + in.C;
    a.Xc = mat + in.X[cs];
    a.Xa = mat + in.X[as];
    a.Xn = mat + in.X[s.spare0];
    a.Xan = mat + in.X[s.spare1];
    a.uc = p.vec + in.u[cs];
    a.ua = p.vec + in.u[as];
    a.vc = p.vec + in.v[cs];
    a.va = p.vec + in.v[as];
    a.avg_empty = avg_empty;
}


// Persistent step kernel over the tile list.
template <typename T, bool EVAL>
__global__ __launch_bounds__(kBBlock) void bstep_kernel(BParams p) {
    constexpr int TR = kBWaves * kWave * VecT<T>::N;
    __shared__ pdot_solver_state sst;
    __shared__ BInst sinst;
    __shared__ BTile stl;
    __shared__ double red[kBBlock];
    __shared__ double colbuf[kBWaves * col_rows<T>()][kChunk];
    const int tid = threadIdx.x;
    const int ntiles = p.nwork[0];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (tid == 0) {
            stl = p.tiles[t];
            sinst = p.inst[stl.inst];
            sst = p.st[stl.inst];
        }
        __syncthreads();
        const bool do_step = !sst.numerical_error && (EVAL || sst.step_budget > 0);
        if (do_step || EVAL) {
            TileArgs<T> a;
            slot_ptrs<T>(p, sinst, sst, a);
            a.n = sinst.n;
            a.ld = sinst.ld;
            a.row0 = (int64_t)stl.rb * TR;
            a.row_end = (a.row0 + TR < sinst.n) ? a.row0 + TR : sinst.n;
            a.j0 = (int64_t)stl.cb * sinst.TC;
            a.jstep = 1;
            a.ncols = ((a.j0 + sinst.TC < sinst.m) ? a.j0 + sinst.TC : sinst.m) - a.j0;
            a.rowpart = p.part + sinst.rowpart + (int64_t)stl.cb * sinst.ld;
            a.colpart = p.part + sinst.colpart + (int64_t)stl.rb * sinst.m;
            a.eta = sst.step_size;
            a.omega = sst.primal_weight;
            avg_weights(sst, &a.nw, &a.aw);
            a.gzero = sst.gzero != 0;
            a.do_step = do_step;
            a.defer = pass_defers(sst, EVAL);
            a.pend = sst.avg_pending != 0;
            a.pnw = sst.pend_nw;
            a.paw = sst.pend_aw;
            double dd = 0.0, ev[kMatStats];
            // zero maps: word ((rb * waves + w) * m + j) of a slot's map covers
            // wave w's rows of column j of row tile rb (fixed tiles of TR rows,
            // whatever the column tiling)
            a.sparse = p.sparse != 0;
            const int cs = sst.cur_slot, as = sst.avg_count == 0 ? cs : sst.avg_slot;
            const int64_t zrow = (int64_t)stl.rb * kBWaves * sinst.m + a.j0;
            a.zc = p.zmap + sinst.Z[cs] + zrow;
            a.za = p.zmap + sinst.Z[as] + zrow;
            a.zn = p.zmap + sinst.Z[sst.spare0] + zrow;
            a.zan = p.zmap + sinst.Z[sst.spare1] + zrow;
            a.zstride = sinst.m;
            step_tile<T, EVAL, kBUnroll, kBWaves, kBPipe>(a, colbuf, red, &dd, ev);
            if (tid == 0) {
                const int64_t lb = stl.rb + (int64_t)stl.cb * sinst.R;
                if (do_step) p.part[sinst.sqpart + lb] = dd;
                if (EVAL) {
                    for (int k = 0; k < kMatStats; ++k) p.part[sinst.evalpart + lb * kMatStats + k] = ev[k];
                }
            }
        }
        __syncthreads();
    }
}

// The arguments of chunk `chunk` (64 rows, or 64 columns) of instance `in`
// in state `s`; *lc receives the chunk's index within the instance (rows
// first, then columns).
__device__ __forceinline__ ChunkArgs chunk_args(const BParams& p, const BInst& in, const pdot_solver_state& s,
                                                bool rows, int32_t chunk, bool do_step, bool eval, int64_t* lc) {
    const bool avg_empty = s.avg_count == 0;
    const int cs = s.cur_slot, as = avg_empty ? cs : s.avg_slot;
    const int64_t* yo = rows ? in.u : in.v;
    const int64_t* so = rows ? in.r : in.c;
    ChunkArgs c;
    c.rows = rows;
    c.k0 = (int64_t)chunk * kWave;
    c.len = rows ? in.n : in.m;
    c.part = p.part + (rows ? in.rowpart : in.colpart);
    c.pstride = rows ? in.ld : in.m;
    c.np = rows ? in.Cb : in.R;
    c.q = p.vec + (rows ? in.a : in.b);
    c.y_cur = p.vec + yo[cs];
    c.s_cur = p.vec + so[cs];
    c.y_avg = p.vec + yo[as];
    c.s_avg = p.vec + so[as];
    c.y_new = p.vec + yo[s.spare0];
    c.s_new = p.vec + so[s.spare0];
    c.y_anew = p.vec + yo[s.spare1];
    c.s_anew = p.vec + so[s.spare1];
    *lc = rows ? chunk : in.NBr + chunk;
    c.sqpart = p.part + in.sqpart;
    c.sq_lo = *lc * in.sqchunk;
    c.sq_hi = c.sq_lo + in.sqchunk < in.nstep ? c.sq_lo + in.sqchunk : in.nstep;
    c.eta = s.step_size;
    c.omega = s.primal_weight;
    avg_weights(s, &c.nw, &c.aw);
    c.avg_empty = avg_empty;
    c.do_step = do_step;
    c.defer = pass_defers(s, eval);
    c.pend = s.avg_pending != 0;
    c.pnw = s.pend_nw;
    c.paw = s.pend_aw;
    return c;
}

// Persistent reduce kernel over the chunk list: every chunk leaves its three
// scalars (and, with EVAL, its evaluation vector parts) for bcontroller_kernel.
template <bool EVAL>
__global__ __launch_bounds__(kBRedBlock) void breduce_kernel(BParams p) {
    __shared__ pdot_solver_state sst;
    __shared__ BInst sinst;
    __shared__ BChunk sch;
    __shared__ double psum[kBRedWaves][kWave];
    const int tid = threadIdx.x;
    const int nheavy = p.nwork[1], nlight = p.nwork[2];
    // chunks with many partials: the block's waves split the partials
    for (int t = blockIdx.x; t < nheavy; t += gridDim.x) {
        if (tid == 0) {
            sch = p.chunks[t];
            sinst = p.inst[sch.inst];
            sst = p.st[sch.inst];
        }
        __syncthreads();
        const bool do_step = !sst.numerical_error && (EVAL || sst.step_budget > 0);
        if (do_step || EVAL) {
            int64_t lc;
            const ChunkArgs c = chunk_args(p, sinst, sst, sch.rows != 0, sch.chunk, do_step, EVAL, &lc);
            double out3[3], ev[kVecStats];
            reduce_chunk<EVAL, kBRedWaves>(c, psum, out3, ev);
            if (tid == 0) {
                if (EVAL) {
                    for (int q = 0; q < kVecStats; ++q) p.red[sinst.evalvec + lc * kVecStats + q] = ev[q];
                }
                if (do_step) {
                    double* rp = p.red + sinst.redpart + 3 * lc;
                    rp[0] = out3[0];
                    rp[1] = out3[1];
                    rp[2] = out3[2];
                }
            }
        }
        __syncthreads();
    }
    // the other chunks: one wave each, no barriers.  The block order is
    // rotated so that the light chunks start at the first block without a
    // heavy chunk: when the grid covers all the work, no block does both.
    const int w = tid / kWave, lane = tid & (kWave - 1);
    const int vb = (int)((blockIdx.x + gridDim.x - (unsigned)nheavy % gridDim.x) % gridDim.x);
    for (int item = vb * kBRedWaves + w; item < nlight; item += gridDim.x * kBRedWaves) {
        const BChunk ch = p.chunks[nheavy + item];
        const BInst& in = p.inst[ch.inst];
        const pdot_solver_state& gs = p.st[ch.inst];
        const bool do_step = !gs.numerical_error && (EVAL || gs.step_budget > 0);
        if (!do_step && !EVAL) continue;
        int64_t lc;
        const ChunkArgs c = chunk_args(p, in, gs, ch.rows != 0, ch.chunk, do_step, EVAL, &lc);
        double out3[3], ev[kVecStats];
        reduce_chunk_wave<EVAL>(c, out3, ev);
        if (lane == 0) {
            if (EVAL) {
                for (int q = 0; q < kVecStats; ++q) p.red[in.evalvec + lc * kVecStats + q] = ev[q];
            }
            if (do_step) {
                double* rp = p.red + in.redpart + 3 * lc;
                rp[0] = out3[0];
                rp[1] = out3[1];
                rp[2] = out3[2];
            }
        }
    }
}

// Commits the pending pass of every instance (one wave per instance).
__global__ __launch_bounds__(kWave) void bcommit_kernel(BParams p, int32_t count) {
    __shared__ pdot_solver_state sst;
    const int k = blockIdx.x;
    if (k >= count) return;
    if (threadIdx.x == 0) sst = p.st[k];
    __syncthreads();
    if (!sst.pending) return;
    const BInst& in = p.inst[k];
    reduce_and_commit(&sst, p.red + in.redpart, in.nred);
    if (threadIdx.x == 0) p.st[k] = sst;
}

// Commits the pass of every active instance that stepped, from its chunks'
// scalars; an evaluation's pass is left pending for the host's decision
// instead (one wave per instance).
template <bool EVAL>
__global__ __launch_bounds__(kWave) void bcontroller_kernel(BParams p) {
    __shared__ pdot_solver_state sst;
    // the grid never exceeds the instance count, so the list entry is in
    // bounds and is loaded together with the list length
    const int nact = p.nwork[3];
    const int k = p.actlist[blockIdx.x];
    if ((int)blockIdx.x >= nact) return;
    if (threadIdx.x == 0) sst = p.st[k];
    __syncthreads();
    if (sst.numerical_error || (!EVAL && sst.step_budget <= 0)) return;
    if (EVAL) {
        if (threadIdx.x == 0) p.st[k].pending = 1;
        return;
    }
    const BInst& in = p.inst[k];
    reduce_and_commit(&sst, p.red + in.redpart, in.nred);
    if (threadIdx.x == 0) p.st[k] = sst;
}

// Final fixed-order reduction of each active instance's evaluation partials.
__global__ __launch_bounds__(kBBlock) void bevalfin_kernel(BParams p, const int32_t* active, int32_t count) {
    __shared__ double red[kBBlock];
    const int k = blockIdx.x;
    if (k >= count || !active[k]) return;
    const BInst in = p.inst[k];
    const int tid = threadIdx.x;
    double* out = p.evalout + (int64_t)k * kEvalOut;
    for (int q = 0; q < kMatStats; ++q) {
        const bool is_max = (q == 1 || q == 3 || q == 6 || q == 8);
        double acc = 0.0;
        for (int64_t b = tid; b < in.nstep; b += kBBlock) {
            const double v = p.part[in.evalpart + b * kMatStats + q];
            acc = is_max ? nanmax(acc, v) : acc + v;
        }
        acc = is_max ? block_max(acc, red) : block_sum(acc, red);
        if (tid == 0) out[q] = acc;
    }
    for (int q = 0; q < kVecStats; ++q) {
        const bool is_max = (q == 0 || q == 3 || q == 6 || q == 9);
        double acc = 0.0;
        for (int64_t b = tid; b < in.nred; b += kBBlock) {
            const double v = p.red[in.evalvec + b * kVecStats + q];
            acc = is_max ? nanmax(acc, v) : acc + v;
        }
        acc = is_max ? block_max(acc, red) : block_sum(acc, red);
        if (tid == 0) out[kMatStats + q] = acc;
    }
}

template <typename T>
__global__ void bnarrow_kernel(const double* __restrict__ src, T* __restrict__ dst, int64_t n, int64_t ld, int64_t ncols) {
    const int64_t total = n * ncols;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        dst[j * ld + i] = (T)src[t];
    }
}
template <typename T>
__global__ void bwiden_kernel(const T* __restrict__ src, double* __restrict__ dst, int64_t n, int64_t ld, int64_t ncols) {
    const int64_t total = n * ncols;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        dst[t] = (double)src[j * ld + i];
    }
}

}  // namespace

// ===================================================================== host

struct pdot_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t storage = PDOT_STORAGE_F32;
    size_t esz = 4;
    int32_t count = 0;
    std::vector<BInst> inst;
    void* mat = nullptr;
    double* vec = nullptr;
    double* part = nullptr;
    int64_t part_cap = 0;  // doubles
    double* red = nullptr;
    BInst* d_inst = nullptr;
    pdot_solver_state* d_st = nullptr;
    BTile* d_tiles = nullptr;
    BChunk* d_chunks = nullptr;
    int32_t* d_nwork = nullptr;
    int32_t* d_active = nullptr;
    int32_t* d_actlist = nullptr;
    double* d_evalout = nullptr;
    double* stage = nullptr;
    int64_t stage_elems = 0;
    pdot_solver_state* h_st = nullptr;  // pinned, count entries
    double* h_eval = nullptr;           // pinned, count * kEvalOut
    std::vector<int32_t> active;
    int64_t max_tiles = 0, max_chunks = 0, tile_cap = 0;
    int32_t ntiles = 0, nchunks = 0, nheavy = 0, nact = 0;
    int32_t cols_per_tile = 64;
    int32_t step_grid = 0, reduce_grid = 0, controller_grid = 0, grid_mult = 1;
    int32_t step_resident = 0, reduce_resident = 0;
    bool started = false;
    bool layout_ok = false;  // partial arenas and work lists match the active set
    bool use_graph = false;
    bool defer = true;  // ordinary passes alternate deferring / applying the averages' update
    bool sparse = true;  // passes skip the zero rows recorded in the zero maps
    uint64_t* zmap = nullptr;  // every instance's four zero maps (BInst::Z)
    int64_t zwords = 0;
    std::map<int64_t, hipGraphExec_t> graphs;
    std::vector<hipEvent_t> events;
    int64_t bytes = 0;
    std::string err;
};

namespace {

int bfail(pdot_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

#define BHIPCHK(b, call)                                                                              \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return bfail((b), PDOT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int balloc(pdot_batch* b, void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return bfail(b, PDOT_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    b->bytes += (int64_t)bytes;
    return PDOT_OK;
}

int rows_per_tile(int32_t storage) { return kBWaves * kWave * (storage == PDOT_STORAGE_F64 ? 2 : 4); }

BParams bparams(pdot_batch* b) {
    BParams p;
    p.mat = b->mat;
    p.vec = b->vec;
    p.part = b->part;
    p.red = b->red;
    p.inst = b->d_inst;
    p.st = b->d_st;
    p.tiles = b->d_tiles;
    p.chunks = b->d_chunks;
    p.nwork = b->d_nwork;
    p.actlist = b->d_actlist;
    p.evalout = b->d_evalout;
    p.zmap = b->zmap;
    p.sparse = (b->sparse && b->zmap) ? 1 : 0;
    return p;
}

void bfree_graphs(pdot_batch* b) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

// Reduction-arena layout (independent of the tile shape): the pass
// reductions a pending pass is committed from, and the evaluation vector
// partials.  Returns the arena size in doubles.
int64_t layout_reductions(pdot_batch* b) {
    int64_t off = 0;
    for (auto& in : b->inst) {
        in.NBr = (int32_t)((in.n + kWave - 1) / kWave);
        in.NBc = (int32_t)((in.m + kWave - 1) / kWave);
        in.nred = in.NBr + in.NBc;
        in.redpart = off; off += 3 * (int64_t)in.nred;
        in.evalvec = off; off += (int64_t)in.nred * kVecStats;
    }
    return off;
}

// Tile shape and partial-arena layout for the active set; returns the arena
// size in doubles.  Every active instance gets TC columns per tile with
//   TC = clamp(min(ceil(W / G), ceil(m / kMinColTiles)), kMinTileCols, cols_per_tile)
// where W = sum over active instances of (row tiles x columns) and G the
// resident step blocks: while many instances are active the tuned width
// applies; as they finish, the tiles narrow so that the remaining instances
// still spread over the whole grid instead of a few blocks walking their
// columns one load latency at a time.
int64_t layout_partials(pdot_batch* b) {
    const int TR = rows_per_tile(b->storage);
    const int64_t slots = std::max<int64_t>(1, (int64_t)b->step_resident * b->grid_mult);
    int64_t work = 0;
    for (int32_t k = 0; k < b->count; ++k) {
        BInst& in = b->inst[k];
        in.R = (int32_t)((in.n + TR - 1) / TR);
        if (b->active[k]) work += (int64_t)in.R * in.m;
    }
    const int64_t tct = std::max<int64_t>(1, (work + slots - 1) / slots);
    int64_t off = 0;
    b->max_tiles = 0;
    for (int32_t k = 0; k < b->count; ++k) {
        BInst& in = b->inst[k];
        const int64_t want = std::min<int64_t>(tct, (in.m + kMinColTiles - 1) / kMinColTiles);
        in.TC = (int32_t)std::min<int64_t>(b->cols_per_tile, std::max<int64_t>(kMinTileCols, want));
        in.Cb = (int32_t)((in.m + in.TC - 1) / in.TC);
        in.nstep = in.R * in.Cb;
        in.sqchunk = (in.nstep + in.nred - 1) / in.nred;
        if (!b->active[k]) continue;  // no kernel touches an inactive instance's partials
        in.rowpart = off; off += (int64_t)in.Cb * in.ld;
        in.colpart = off; off += (int64_t)in.R * in.m;
        in.sqpart = off; off += in.nstep;
        in.evalpart = off; off += (int64_t)in.nstep * kMatStats;
        b->max_tiles += in.nstep;
    }
    return off;
}

// Rebuild the tile and chunk lists from the active set (largest tiles first).
int rebuild_work(pdot_batch* b) {
    std::vector<BTile> tiles;
    std::vector<BChunk> chunks, heavy, light;
    const int TR = rows_per_tile(b->storage);
    std::vector<std::pair<int64_t, BTile>> weighted;
    std::vector<int32_t> actlist;
    for (int32_t k = 0; k < b->count; ++k) {
        if (!b->active[k]) continue;
        actlist.push_back(k);
        const BInst& in = b->inst[k];
        for (int32_t cb = 0; cb < in.Cb; ++cb) {
            const int64_t cols = std::min<int64_t>(in.TC, in.m - (int64_t)cb * in.TC);  // tile work for the sort
            for (int32_t rb = 0; rb < in.R; ++rb) {
                const int64_t rows = std::min<int64_t>(TR, in.n - (int64_t)rb * TR);
                weighted.push_back({rows * cols, BTile{k, rb, cb, 0}});
            }
        }
        // row chunks sum Cb partials, column chunks R
        for (int32_t c = 0; c < in.NBr; ++c) (in.Cb > kHeavyParts ? heavy : light).push_back(BChunk{k, 1, c, 0});
        for (int32_t c = 0; c < in.NBc; ++c) (in.R > kHeavyParts ? heavy : light).push_back(BChunk{k, 0, c, 0});
    }
    chunks.reserve(heavy.size() + light.size());
    chunks.insert(chunks.end(), heavy.begin(), heavy.end());
    chunks.insert(chunks.end(), light.begin(), light.end());
    std::stable_sort(weighted.begin(), weighted.end(),
                     [](const std::pair<int64_t, BTile>& x, const std::pair<int64_t, BTile>& y) { return x.first > y.first; });
    tiles.reserve(weighted.size());
    for (auto& w : weighted) tiles.push_back(w.second);
    b->ntiles = (int32_t)tiles.size();
    b->nchunks = (int32_t)chunks.size();
    b->nheavy = (int32_t)heavy.size();
    b->nact = (int32_t)actlist.size();
    int32_t nw[4] = {b->ntiles, b->nheavy, (int32_t)light.size(), b->nact};
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    if (!tiles.empty())
        BHIPCHK(b, hipMemcpy(b->d_tiles, tiles.data(), sizeof(BTile) * tiles.size(), hipMemcpyHostToDevice));
    if (!chunks.empty())
        BHIPCHK(b, hipMemcpy(b->d_chunks, chunks.data(), sizeof(BChunk) * chunks.size(), hipMemcpyHostToDevice));
    BHIPCHK(b, hipMemcpy(b->d_nwork, nw, sizeof(nw), hipMemcpyHostToDevice));
    if (!actlist.empty())
        BHIPCHK(b, hipMemcpy(b->d_actlist, actlist.data(), sizeof(int32_t) * actlist.size(), hipMemcpyHostToDevice));
    BHIPCHK(b, hipMemcpy(b->d_active, b->active.data(), sizeof(int32_t) * b->count, hipMemcpyHostToDevice));
    return PDOT_OK;
}

int setup_batch_partials(pdot_batch* b) {
    b->layout_ok = false;  // stays false if an allocation below fails
    const int64_t nd = layout_partials(b);
    int rc;
    if (nd > b->part_cap || !b->part) {
        bfree_graphs(b);
        if (b->part) {
            (void)hipFree(b->part);
            b->bytes -= (int64_t)sizeof(double) * b->part_cap;
        }
        b->part = nullptr;
        b->part_cap = 0;
        if ((rc = balloc(b, (void**)&b->part, sizeof(double) * (size_t)nd))) return rc;
        b->part_cap = nd;
    }
    if (b->max_tiles > b->tile_cap || !b->d_tiles) {
        bfree_graphs(b);
        if (b->d_tiles) {
            (void)hipFree(b->d_tiles);
            b->bytes -= (int64_t)sizeof(BTile) * b->tile_cap;
        }
        b->d_tiles = nullptr;
        b->tile_cap = 0;
        if ((rc = balloc(b, (void**)&b->d_tiles, sizeof(BTile) * (size_t)b->max_tiles))) return rc;
        b->tile_cap = b->max_tiles;
    }
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    BHIPCHK(b, hipMemcpy(b->d_inst, b->inst.data(), sizeof(BInst) * b->count, hipMemcpyHostToDevice));
    if ((rc = rebuild_work(b))) return rc;
    // persistent grids sized to the work: no idle blocks are launched once
    // few instances remain
    const int32_t sg = (int32_t)std::max<int64_t>(1, std::min<int64_t>((int64_t)b->step_resident * b->grid_mult, b->ntiles));
    const int32_t reduce_items = b->nheavy + (b->nchunks - b->nheavy + kBRedWaves - 1) / kBRedWaves;
    const int32_t rg = std::max(1, std::min(b->reduce_resident, reduce_items));
    const int32_t fg = std::max(1, b->nact);
    if (sg != b->step_grid || rg != b->reduce_grid || fg != b->controller_grid) bfree_graphs(b);  // captured launches hold the grid
    b->step_grid = sg;
    b->reduce_grid = rg;
    b->controller_grid = fg;
    b->layout_ok = true;
    return PDOT_OK;
}

hipError_t blaunch_step(pdot_batch* b, bool eval) {
    BParams p = bparams(b);
    if (b->storage == PDOT_STORAGE_F64) {
        if (eval) hipLaunchKernelGGL((bstep_kernel<double, true>), dim3(b->step_grid), dim3(kBBlock), 0, b->stream, p);
        else hipLaunchKernelGGL((bstep_kernel<double, false>), dim3(b->step_grid), dim3(kBBlock), 0, b->stream, p);
    } else {
        if (eval) hipLaunchKernelGGL((bstep_kernel<float, true>), dim3(b->step_grid), dim3(kBBlock), 0, b->stream, p);
        else hipLaunchKernelGGL((bstep_kernel<float, false>), dim3(b->step_grid), dim3(kBBlock), 0, b->stream, p);
    }
    return hipGetLastError();
}

hipError_t blaunch_reduce(pdot_batch* b, bool eval) {
    BParams p = bparams(b);
    if (eval) hipLaunchKernelGGL((breduce_kernel<true>), dim3(b->reduce_grid), dim3(kBRedBlock), 0, b->stream, p);
    else hipLaunchKernelGGL((breduce_kernel<false>), dim3(b->reduce_grid), dim3(kBRedBlock), 0, b->stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (eval) hipLaunchKernelGGL((bcontroller_kernel<true>), dim3(b->controller_grid), dim3(kWave), 0, b->stream, p);
    else hipLaunchKernelGGL((bcontroller_kernel<false>), dim3(b->controller_grid), dim3(kWave), 0, b->stream, p);
    return hipGetLastError();
}

hipError_t blaunch_commit(pdot_batch* b) {
    hipLaunchKernelGGL(bcommit_kernel, dim3(b->count), dim3(kWave), 0, b->stream, bparams(b), b->count);
    return hipGetLastError();
}

int benqueue(pdot_batch* b, int64_t passes) {
    BHIPCHK(b, blaunch_commit(b));
    for (int64_t k = 0; k < passes; ++k) {
        BHIPCHK(b, blaunch_step(b, false));
        BHIPCHK(b, blaunch_reduce(b, false));
    }
    return PDOT_OK;
}

int bread_states(pdot_batch* b) {
    BHIPCHK(b, hipMemcpyAsync(b->h_st, b->d_st, sizeof(pdot_solver_state) * b->count, hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return PDOT_OK;
}
int bwrite_states(pdot_batch* b) {
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    BHIPCHK(b, hipMemcpyAsync(b->d_st, b->h_st, sizeof(pdot_solver_state) * b->count, hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return PDOT_OK;
}

void bchoose_spares(pdot_solver_state& st) {
    int used0 = st.cur_slot, used1 = st.avg_count > 0 ? st.avg_slot : st.cur_slot;
    int k = 0;
    for (int s = 0; s < 4; ++s) {
        if (s == used0 || s == used1) continue;
        if (k == 0) st.spare0 = s;
        else if (k == 1) st.spare1 = s;
        ++k;
    }
}

// Writes instance k's lagging average contribution (if any) into its average
// slots, so that they hold the average itself; updates b->h_st[k] only (the
// caller writes the states back).
// Zero maps of instance k: the slot's segments expanded to the dense layout in
// place, and its map set to all ones (which describes that layout).
int bdensify(pdot_batch* b, int32_t k, int slot) {
    if (!b->sparse || !b->zmap) return PDOT_OK;
    const BInst& in = b->inst[k];
    ZGeom g;
    g.n = in.n;
    g.m = in.m;
    g.ld = in.ld;
    g.TRB = rows_per_tile(b->storage);
    g.R = (int32_t)((in.n + g.TRB - 1) / g.TRB);
    g.Cb = 1;  // one "column block" of all m columns: word (rb * waves + w) * m + j
    g.TC = (int32_t)in.m;
    g.interleave = 0;
    g.waves = kBWaves;
    g.vec = (int32_t)(16 / b->esz);
    g.tc = in.m;
    g.words = (int64_t)g.R * kBWaves * in.m;
    const int64_t blocks = std::min<int64_t>((g.words + 3) / 4, 4096);
    uint64_t* z = b->zmap + in.Z[slot];
    if (b->storage == PDOT_STORAGE_F64)
        hipLaunchKernelGGL(densify_kernel<double>, dim3(blocks), dim3(256), 0, b->stream,
                           static_cast<double*>(b->mat) + in.X[slot], z, g);
    else
        hipLaunchKernelGGL(densify_kernel<float>, dim3(blocks), dim3(256), 0, b->stream,
                           static_cast<float*>(b->mat) + in.X[slot], z, g);
    BHIPCHK(b, hipGetLastError());
    hipLaunchKernelGGL(fill_u64_kernel, dim3(std::min<int64_t>((g.words + 255) / 256, 1024)), dim3(256), 0, b->stream,
                       z, g.words, ~0ull);
    BHIPCHK(b, hipGetLastError());
    return PDOT_OK;
}

int bmaterialize(pdot_batch* b, int32_t k) {
    pdot_solver_state& st = b->h_st[k];
    if (!st.avg_pending) return PDOT_OK;
    const BInst& in = b->inst[k];
    const int64_t total = in.ld * in.m;
    int rc;
    if ((rc = bdensify(b, k, st.avg_slot)) || (rc = bdensify(b, k, st.cur_slot))) return rc;  // read and written whole
    if (b->storage == PDOT_STORAGE_F64)
        hipLaunchKernelGGL(materialize_kernel<double>, dim3(256), dim3(256), 0, b->stream,
                           static_cast<double*>(b->mat) + in.X[st.avg_slot],
                           static_cast<const double*>(b->mat) + in.X[st.cur_slot], total, st.pend_nw, st.pend_aw);
    else
        hipLaunchKernelGGL(materialize_kernel<float>, dim3(256), dim3(256), 0, b->stream,
                           static_cast<float*>(b->mat) + in.X[st.avg_slot],
                           static_cast<const float*>(b->mat) + in.X[st.cur_slot], total, st.pend_nw, st.pend_aw);
    BHIPCHK(b, hipGetLastError());
    VecAverages va;
    const int64_t* offs[4] = {in.u, in.r, in.v, in.c};
    for (int w = 0; w < 4; ++w) {
        va.avg[w] = b->vec + offs[w][st.avg_slot];
        va.cur[w] = b->vec + offs[w][st.cur_slot];
        va.len[w] = w < 2 ? in.n : in.m;
    }
    hipLaunchKernelGGL(materialize_vectors_kernel, dim3(8, 4), dim3(256), 0, b->stream, va, st.pend_nw, st.pend_aw);
    BHIPCHK(b, hipGetLastError());
    st.avg_pending = 0;
    return PDOT_OK;
}

int bensure_stage(pdot_batch* b, int64_t elems) {
    if (b->stage && b->stage_elems >= elems) return PDOT_OK;
    if (b->stage) {
        (void)hipFree(b->stage);
        b->bytes -= 8 * b->stage_elems;
        b->stage = nullptr;
    }
    b->stage_elems = elems;
    return balloc(b, (void**)&b->stage, sizeof(double) * (size_t)elems);
}

}  // namespace

extern "C" {

int pdot_batch_create(pdot_batch** out, int32_t count, const int64_t* n, const int64_t* m, int32_t storage,
                      int32_t device) {
    if (!out) return PDOT_E_ARG;
    *out = nullptr;
    if (count <= 0 || !n || !m || (storage != PDOT_STORAGE_F64 && storage != PDOT_STORAGE_F32)) return PDOT_E_ARG;
    for (int32_t k = 0; k < count; ++k)
        if (n[k] <= 0 || m[k] <= 0) return PDOT_E_ARG;
    pdot_batch* b = new pdot_batch();
    *out = b;
    b->count = count;
    b->storage = storage;
    b->esz = storage == PDOT_STORAGE_F64 ? 8 : 4;
    b->device = device;
    BHIPCHK(b, hipSetDevice(device));
    BHIPCHK(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->inst.resize(count);
    int64_t moff = 0, voff = 0;
    for (int32_t k = 0; k < count; ++k) {
        BInst& in = b->inst[k];
        memset(&in, 0, sizeof(in));
        in.n = n[k];
        in.m = m[k];
        in.ld = (n[k] + kLdAlign - 1) / kLdAlign * kLdAlign;
        const int64_t ms = in.ld * in.m;
        in.C = moff; moff += ms;
        for (int s = 0; s < 4; ++s) {
            in.X[s] = moff; moff += ms;
        }
        for (int s = 0; s < 4; ++s) {
            in.u[s] = voff; voff += in.ld;
            in.r[s] = voff; voff += in.ld;
            in.v[s] = voff; voff += in.m;
            in.c[s] = voff; voff += in.m;
        }
        in.a = voff; voff += in.ld;
        in.b = voff; voff += in.m;
        const int64_t rt = (in.n + rows_per_tile(storage) - 1) / rows_per_tile(storage);
        for (int s = 0; s < 4; ++s) {
            in.Z[s] = b->zwords;
            b->zwords += rt * kBWaves * in.m;
        }
    }
    int rc;
    if ((rc = balloc(b, &b->mat, b->esz * (size_t)moff))) return rc;
    if ((rc = balloc(b, (void**)&b->vec, sizeof(double) * (size_t)voff))) return rc;
    BHIPCHK(b, hipMemsetAsync(b->mat, 0, b->esz * (size_t)moff, b->stream));
    BHIPCHK(b, hipMemsetAsync(b->vec, 0, sizeof(double) * (size_t)voff, b->stream));
    // all bits set: the dense layout (the slots hold zeros, which every map describes)
    if ((rc = balloc(b, (void**)&b->zmap, sizeof(uint64_t) * (size_t)b->zwords))) return rc;
    BHIPCHK(b, hipMemsetAsync(b->zmap, 0xFF, sizeof(uint64_t) * (size_t)b->zwords, b->stream));
    if ((rc = balloc(b, (void**)&b->d_inst, sizeof(BInst) * count))) return rc;
    if ((rc = balloc(b, (void**)&b->d_st, sizeof(pdot_solver_state) * count))) return rc;
    if ((rc = balloc(b, (void**)&b->d_nwork, sizeof(int32_t) * 4))) return rc;
    if ((rc = balloc(b, (void**)&b->d_active, sizeof(int32_t) * count))) return rc;
    if ((rc = balloc(b, (void**)&b->d_actlist, sizeof(int32_t) * count))) return rc;
    if ((rc = balloc(b, (void**)&b->d_evalout, sizeof(double) * kEvalOut * count))) return rc;
    const int64_t nred_total = layout_reductions(b);
    if ((rc = balloc(b, (void**)&b->red, sizeof(double) * (size_t)nred_total))) return rc;
    b->max_chunks = 0;
    for (const auto& in : b->inst) b->max_chunks += in.nred;
    if ((rc = balloc(b, (void**)&b->d_chunks, sizeof(BChunk) * (size_t)b->max_chunks))) return rc;
    BHIPCHK(b, hipMemsetAsync(b->d_st, 0, sizeof(pdot_solver_state) * count, b->stream));
    BHIPCHK(b, hipHostMalloc((void**)&b->h_st, sizeof(pdot_solver_state) * count, hipHostMallocDefault));
    BHIPCHK(b, hipHostMalloc((void**)&b->h_eval, sizeof(double) * kEvalOut * count, hipHostMallocDefault));
    b->active.assign(count, 1);
    int ncu = 0, per_cu = 0, per_cu_r = 0;
    BHIPCHK(b, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
    if (storage == PDOT_STORAGE_F64)
        BHIPCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bstep_kernel<double, false>, kBBlock, 0));
    else
        BHIPCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bstep_kernel<float, false>, kBBlock, 0));
    BHIPCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_r, breduce_kernel<false>, kBRedBlock, 0));
    b->step_resident = std::max(1, per_cu) * std::max(1, ncu);
    b->reduce_resident = std::max(1, per_cu_r) * std::max(1, ncu);
    if ((rc = setup_batch_partials(b))) return rc;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return PDOT_OK;
}

void pdot_batch_destroy(pdot_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    bfree_graphs(b);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    void* ps[] = {b->mat, b->vec, b->part, b->red, b->d_inst, b->d_st, b->d_tiles, b->d_chunks,
                  b->d_nwork, b->d_active, b->d_actlist, b->d_evalout, b->stage, b->zmap};
    for (void* q : ps)
        if (q) (void)hipFree(q);
    if (b->h_st) (void)hipHostFree(b->h_st);
    if (b->h_eval) (void)hipHostFree(b->h_eval);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* pdot_batch_last_error(const pdot_batch* b) { return b ? b->err.c_str() : "null batch"; }

int pdot_batch_upload(pdot_batch* b, int32_t k, const double* C, const double* a, const double* bb) {
    if (!b) return PDOT_E_ARG;
    if (k < 0 || k >= b->count || !C || !a || !bb) return bfail(b, PDOT_E_ARG, "pdot_batch_upload: bad instance or pointer");
    BHIPCHK(b, hipSetDevice(b->device));
    const BInst& in = b->inst[k];
    if (b->storage == PDOT_STORAGE_F64) {
        BHIPCHK(b, hipMemcpy2DAsync(static_cast<double*>(b->mat) + in.C, in.ld * 8, C, in.n * 8, in.n * 8, in.m,
                                    hipMemcpyHostToDevice, b->stream));
    } else {
        int rc = bensure_stage(b, in.n * in.m);
        if (rc) return rc;
        BHIPCHK(b, hipMemcpyAsync(b->stage, C, sizeof(double) * in.n * in.m, hipMemcpyHostToDevice, b->stream));
        hipLaunchKernelGGL(bnarrow_kernel<float>, dim3(1024), dim3(256), 0, b->stream, b->stage,
                           static_cast<float*>(b->mat) + in.C, in.n, in.ld, in.m);
        BHIPCHK(b, hipGetLastError());
    }
    BHIPCHK(b, hipMemcpyAsync(b->vec + in.a, a, sizeof(double) * in.n, hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipMemcpyAsync(b->vec + in.b, bb, sizeof(double) * in.m, hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return PDOT_OK;
}

int pdot_batch_set_start(pdot_batch* b, int32_t step_policy, double reduction_exponent, double growth_exponent,
                         const double* step_size, const double* primal_weight) {
    if (!b) return PDOT_E_ARG;
    if (!step_size || !primal_weight) return bfail(b, PDOT_E_ARG, "pdot_batch_set_start: null pointer");
    if (step_policy != PDOT_POLICY_ADAPTIVE && step_policy != PDOT_POLICY_CONSTANT)
        return bfail(b, PDOT_E_ARG, "pdot_batch_set_start: unknown step policy");
    BHIPCHK(b, hipSetDevice(b->device));
    for (int32_t k = 0; k < b->count; ++k) {
        // slot 0 holds the start point; the other slots are written before they are read
        const BInst& in = b->inst[k];
        BHIPCHK(b, hipMemsetAsync(static_cast<char*>(b->mat) + b->esz * in.X[0], 0, b->esz * in.ld * in.m, b->stream));
        BHIPCHK(b, hipMemsetAsync(b->vec + in.u[0], 0, sizeof(double) * in.ld, b->stream));
        BHIPCHK(b, hipMemsetAsync(b->vec + in.r[0], 0, sizeof(double) * in.ld, b->stream));
        BHIPCHK(b, hipMemsetAsync(b->vec + in.v[0], 0, sizeof(double) * in.m, b->stream));
        BHIPCHK(b, hipMemsetAsync(b->vec + in.c[0], 0, sizeof(double) * in.m, b->stream));
        pdot_solver_state& st = b->h_st[k];
        memset(&st, 0, sizeof(st));
        st.step_size = step_size[k];
        st.primal_weight = primal_weight[k];
        st.reduction_exponent = reduction_exponent;
        st.growth_exponent = growth_exponent;
        st.step_policy = step_policy;
        st.gzero = 1;
        st.cur_slot = 0;
        st.avg_slot = 0;
        st.spare0 = 1;
        st.spare1 = 2;
        st.defer_ok = b->defer ? 1 : 0;
    }
    int rc = bwrite_states(b);
    if (rc) return rc;
    b->active.assign(b->count, 1);
    if ((rc = setup_batch_partials(b))) return rc;
    b->started = true;
    return PDOT_OK;
}

int pdot_batch_set_warm_start(pdot_batch* b, int32_t k, const double* u0, const double* v0) {
    if (!b) return PDOT_E_ARG;
    if (!b->started) return bfail(b, PDOT_E_STATE, "pdot_batch_set_warm_start before pdot_batch_set_start");
    if (k < 0 || k >= b->count || !u0 || !v0) return bfail(b, PDOT_E_ARG, "pdot_batch_set_warm_start: bad argument");
    int rc = bread_states(b);
    if (rc) return rc;
    pdot_solver_state& st = b->h_st[k];
    if (st.take_step_calls != 0 || st.pending) return bfail(b, PDOT_E_STATE, "pdot_batch_set_warm_start after passes ran");
    const BInst& in = b->inst[k];
    BHIPCHK(b, hipMemcpyAsync(b->vec + in.u[st.cur_slot], u0, sizeof(double) * in.n, hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipMemcpyAsync(b->vec + in.v[st.cur_slot], v0, sizeof(double) * in.m, hipMemcpyHostToDevice, b->stream));
    st.gzero = 0;
    return bwrite_states(b);
}

int pdot_batch_set_active(pdot_batch* b, const int32_t* active) {
    if (!b || !active) return PDOT_E_ARG;
    bool same = true;
    for (int32_t k = 0; k < b->count; ++k) {
        const int32_t a = active[k] ? 1 : 0;
        same = same && a == b->active[k];
        b->active[k] = a;
    }
    return same ? PDOT_OK : setup_batch_partials(b);
}

int pdot_batch_get_states(pdot_batch* b, pdot_solver_state* out) {
    if (!b || !out) return PDOT_E_ARG;
    int rc = bread_states(b);
    if (rc) return rc;
    memcpy(out, b->h_st, sizeof(pdot_solver_state) * b->count);
    return PDOT_OK;
}

int pdot_batch_set_states(pdot_batch* b, const pdot_solver_state* in) {
    if (!b || !in) return PDOT_E_ARG;
    for (int32_t k = 0; k < b->count; ++k) {
        const pdot_solver_state& s = in[k];
        const int32_t slots[4] = {s.cur_slot, s.avg_slot, s.spare0, s.spare1};
        for (int32_t q : slots)
            if (q < 0 || q > 3) return bfail(b, PDOT_E_ARG, "pdot_batch_set_states: slot index out of range");
        if ((s.defer_ok != 0 && s.defer_ok != 1) || (s.avg_pending != 0 && s.avg_pending != 1))
            return bfail(b, PDOT_E_ARG, "pdot_batch_set_states: defer_ok and avg_pending must be 0 or 1");
        if (s.avg_pending && (s.avg_count == 0 || s.avg_slot == s.cur_slot))
            return bfail(b, PDOT_E_ARG,
                         "pdot_batch_set_states: a lagging contribution needs a non-empty average in its own slot");
    }
    memcpy(b->h_st, in, sizeof(pdot_solver_state) * b->count);
    return bwrite_states(b);
}

int pdot_batch_run_steps(pdot_batch* b, const int64_t* counts, pdot_solver_state* out) {
    if (!b || !counts) return PDOT_E_ARG;
    if (!b->started) return bfail(b, PDOT_E_STATE, "pdot_batch_run_steps before pdot_batch_set_start");
    if (!b->layout_ok) return bfail(b, PDOT_E_STATE, "pdot_batch_run_steps: the last re-layout failed: " + b->err);
    BHIPCHK(b, hipSetDevice(b->device));
    int rc = bread_states(b);
    if (rc) return rc;
    int64_t passes = 0;
    for (int32_t k = 0; k < b->count; ++k) {
        if (counts[k] < 0) return bfail(b, PDOT_E_ARG, "pdot_batch_run_steps: negative count");
        pdot_solver_state& st = b->h_st[k];
        // the pending pass (if any) is committed first and consumes one unit
        st.step_budget = b->active[k] ? counts[k] + (st.pending ? 1 : 0) : 0;
        if (b->active[k]) passes = std::max<int64_t>(passes, counts[k]);
    }
    if ((rc = bwrite_states(b))) return rc;
    if (b->use_graph && passes > 1) {
        auto it = b->graphs.find(passes);
        if (it == b->graphs.end()) {
            hipGraph_t g;
            BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = benqueue(b, passes);
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return bfail(b, PDOT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            hipGraphExec_t ge;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return bfail(b, PDOT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            it = b->graphs.emplace(passes, ge).first;
        }
        BHIPCHK(b, hipGraphLaunch(it->second, b->stream));
    } else {
        if ((rc = benqueue(b, passes))) return rc;
    }
    if ((rc = bread_states(b))) return rc;
    if (out) memcpy(out, b->h_st, sizeof(pdot_solver_state) * b->count);
    return PDOT_OK;
}

int pdot_batch_eval(pdot_batch* b, pdot_point_stats* cur, pdot_point_stats* avg, pdot_solver_state* states) {
    if (!b) return PDOT_E_ARG;
    if (!b->started) return bfail(b, PDOT_E_STATE, "pdot_batch_eval before pdot_batch_set_start");
    if (!b->layout_ok) return bfail(b, PDOT_E_STATE, "pdot_batch_eval: the last re-layout failed: " + b->err);
    BHIPCHK(b, hipSetDevice(b->device));
    BHIPCHK(b, blaunch_commit(b));
    BHIPCHK(b, blaunch_step(b, true));
    BHIPCHK(b, blaunch_reduce(b, true));
    hipLaunchKernelGGL(bevalfin_kernel, dim3(b->count), dim3(kBBlock), 0, b->stream, bparams(b), b->d_active, b->count);
    BHIPCHK(b, hipGetLastError());
    BHIPCHK(b, hipMemcpyAsync(b->h_eval, b->d_evalout, sizeof(double) * kEvalOut * b->count, hipMemcpyDeviceToHost, b->stream));
    int rc = bread_states(b);
    if (rc) return rc;
    for (int32_t k = 0; k < b->count; ++k) {
        const double* e = b->h_eval + (int64_t)k * kEvalOut;
        const double* v = e + kMatStats;
        const bool on = b->active[k] != 0;
        if (cur) {
            pdot_point_stats& c = cur[k];
            if (!on) memset(&c, 0, sizeof(c));
            else {
                c.primal_objective = e[0]; c.x_inf = e[1]; c.x_sq = e[2]; c.rc_inf = e[3]; c.rc_sq = e[4];
                c.p_inf = v[0]; c.p_sq = v[1]; c.dual_objective = v[2]; c.u_inf = v[3]; c.u_sq = v[4]; c.y_sq = v[5];
            }
        }
        if (avg) {
            pdot_point_stats& a = avg[k];
            if (!on || b->h_st[k].avg_count == 0) memset(&a, 0, sizeof(a));
            else {
                a.primal_objective = e[5]; a.x_inf = e[6]; a.x_sq = e[7]; a.rc_inf = e[8]; a.rc_sq = e[9];
                a.p_inf = v[6]; a.p_sq = v[7]; a.dual_objective = v[8]; a.u_inf = v[9]; a.u_sq = v[10]; a.y_sq = v[11];
            }
        }
    }
    if (states) memcpy(states, b->h_st, sizeof(pdot_solver_state) * b->count);
    return PDOT_OK;
}

int pdot_batch_restart(pdot_batch* b, const int32_t* action, const double* new_primal_weight) {
    if (!b || !action || !new_primal_weight) return PDOT_E_ARG;
    int rc = bread_states(b);
    if (rc) return rc;
    for (int32_t k = 0; k < b->count; ++k) {
        if (!action[k]) continue;
        pdot_solver_state& st = b->h_st[k];
        if (action[k] != 1 && action[k] != 2) return bfail(b, PDOT_E_ARG, "pdot_batch_restart: action must be 0, 1 or 2");
        if (action[k] == 2 && st.avg_count == 0)
            return bfail(b, PDOT_E_STATE, "pdot_batch_restart: restart to an empty average");
    }
    for (int32_t k = 0; k < b->count; ++k) {
        if (!action[k]) continue;
        pdot_solver_state& st = b->h_st[k];
        st.pending = 0;
        if (action[k] == 2) {
            if ((rc = bmaterialize(b, k))) return rc;
            st.cur_slot = st.avg_slot;
            st.gzero = 0;
        }
        st.avg_count = 0;
        st.avg_pending = 0;  // the reset average has no lagging contribution
        st.sum_weights = 0.0;
        st.avg_slot = st.cur_slot;
        st.primal_weight = new_primal_weight[k];
        bchoose_spares(st);
    }
    return bwrite_states(b);
}

int pdot_batch_discard(pdot_batch* b, const int32_t* which) {
    if (!b || !which) return PDOT_E_ARG;
    int rc = bread_states(b);
    if (rc) return rc;
    for (int32_t k = 0; k < b->count; ++k)
        if (which[k]) b->h_st[k].pending = 0;
    return bwrite_states(b);
}

int pdot_batch_download(pdot_batch* b, int32_t k, int32_t which, double* X, double* u, double* v) {
    if (!b) return PDOT_E_ARG;
    if (k < 0 || k >= b->count) return bfail(b, PDOT_E_ARG, "pdot_batch_download: bad instance");
    int rc = bread_states(b);
    if (rc) return rc;
    const pdot_solver_state& st = b->h_st[k];
    int slot;
    if (which == PDOT_POINT_CURRENT) slot = st.cur_slot;
    else if (which == PDOT_POINT_AVERAGE) {
        if (st.avg_count == 0) return bfail(b, PDOT_E_STATE, "pdot_batch_download: the average is empty");
        if (st.avg_pending) {
            if ((rc = bmaterialize(b, k))) return rc;
            if ((rc = bwrite_states(b))) return rc;
        }
        slot = st.avg_slot;
    } else return bfail(b, PDOT_E_ARG, "pdot_batch_download: unknown point");
    const BInst& in = b->inst[k];
    if (X && (rc = bdensify(b, k, slot))) return rc;
    if (X) {
        if (b->storage == PDOT_STORAGE_F64) {
            BHIPCHK(b, hipMemcpy2DAsync(X, in.n * 8, static_cast<double*>(b->mat) + in.X[slot], in.ld * 8, in.n * 8, in.m,
                                        hipMemcpyDeviceToHost, b->stream));
        } else {
            if ((rc = bensure_stage(b, in.n * in.m))) return rc;
            hipLaunchKernelGGL(bwiden_kernel<float>, dim3(1024), dim3(256), 0, b->stream,
                               static_cast<const float*>(b->mat) + in.X[slot], b->stage, in.n, in.ld, in.m);
            BHIPCHK(b, hipGetLastError());
            BHIPCHK(b, hipMemcpyAsync(X, b->stage, sizeof(double) * in.n * in.m, hipMemcpyDeviceToHost, b->stream));
        }
    }
    if (u) BHIPCHK(b, hipMemcpyAsync(u, b->vec + in.u[slot], sizeof(double) * in.n, hipMemcpyDeviceToHost, b->stream));
    if (v) BHIPCHK(b, hipMemcpyAsync(v, b->vec + in.v[slot], sizeof(double) * in.m, hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return PDOT_OK;
}

int pdot_batch_time_steps(pdot_batch* b, int64_t count, double* step_kernel_ms, double* total_ms) {
    if (!b) return PDOT_E_ARG;
    if (!b->started) return bfail(b, PDOT_E_STATE, "pdot_batch_time_steps before pdot_batch_set_start");
    if (!b->layout_ok) return bfail(b, PDOT_E_STATE, "pdot_batch_time_steps: the last re-layout failed: " + b->err);
    if (count < 1) return bfail(b, PDOT_E_ARG, "pdot_batch_time_steps: count < 1");
    BHIPCHK(b, hipSetDevice(b->device));
    int rc = bread_states(b);
    if (rc) return rc;
    for (int32_t k = 0; k < b->count; ++k)
        b->h_st[k].step_budget = b->active[k] ? count + (b->h_st[k].pending ? 1 : 0) : 0;
    if ((rc = bwrite_states(b))) return rc;
    const size_t need = (size_t)(2 * count + 2);
    while (b->events.size() < need) {
        hipEvent_t e;
        BHIPCHK(b, hipEventCreate(&e));
        b->events.push_back(e);
    }
    hipEvent_t* ev = b->events.data();
    BHIPCHK(b, hipEventRecord(ev[0], b->stream));
    BHIPCHK(b, blaunch_commit(b));
    for (int64_t k = 0; k < count; ++k) {
        BHIPCHK(b, hipEventRecord(ev[2 + 2 * k], b->stream));
        BHIPCHK(b, blaunch_step(b, false));
        BHIPCHK(b, hipEventRecord(ev[3 + 2 * k], b->stream));
        BHIPCHK(b, blaunch_reduce(b, false));
    }
    BHIPCHK(b, hipEventRecord(ev[1], b->stream));
    BHIPCHK(b, hipEventSynchronize(ev[1]));
    double ks = 0.0;
    for (int64_t k = 0; k < count; ++k) {
        float t = 0.f;
        BHIPCHK(b, hipEventElapsedTime(&t, ev[2 + 2 * k], ev[3 + 2 * k]));
        ks += t;
    }
    float tt = 0.f;
    BHIPCHK(b, hipEventElapsedTime(&tt, ev[0], ev[1]));
    if (step_kernel_ms) *step_kernel_ms = ks;
    if (total_ms) *total_ms = tt;
    return PDOT_OK;
}

int pdot_batch_use_graphs(pdot_batch* b, int32_t enable) {
    if (!b) return PDOT_E_ARG;
    b->use_graph = enable != 0;
    return PDOT_OK;
}

int pdot_batch_set_defer(pdot_batch* b, int32_t enable) {
    if (!b) return PDOT_E_ARG;
    if (enable != 0 && enable != 1) return bfail(b, PDOT_E_ARG, "pdot_batch_set_defer: enable must be 0 or 1");
    b->defer = enable != 0;
    if (!b->started) return PDOT_OK;
    BHIPCHK(b, hipSetDevice(b->device));
    int rc = bread_states(b);
    if (rc) return rc;
    for (int32_t k = 0; k < b->count; ++k) {
        if (!enable && (rc = bmaterialize(b, k))) return rc;
        b->h_st[k].defer_ok = enable;
    }
    return bwrite_states(b);
}

int pdot_batch_set_sparse(pdot_batch* b, int32_t enable) {
    if (!b) return PDOT_E_ARG;
    if (enable != 0 && enable != 1) return bfail(b, PDOT_E_ARG, "pdot_batch_set_sparse: enable must be 0 or 1");
    if ((enable != 0) == b->sparse) return PDOT_OK;
    BHIPCHK(b, hipSetDevice(b->device));
    int rc;
    if (enable) {
        // dense passes kept no maps: every bit set, then kept from here on
        BHIPCHK(b, hipMemsetAsync(b->zmap, 0xFF, sizeof(uint64_t) * (size_t)b->zwords, b->stream));
    } else {
        for (int32_t k = 0; k < b->count; ++k)
            for (int s = 0; s < 4; ++s)
                if ((rc = bdensify(b, k, s))) return rc;
    }
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    bfree_graphs(b);  // the captured passes carry the mode
    b->sparse = enable != 0;
    return PDOT_OK;
}

int pdot_batch_set_launch(pdot_batch* b, int32_t cols_per_tile, int32_t grid_mult) {
    if (!b) return PDOT_E_ARG;
    if (cols_per_tile < 1 || grid_mult < 1 || grid_mult > 64)
        return bfail(b, PDOT_E_ARG, "pdot_batch_set_launch: cols_per_tile >= 1, 1 <= grid_mult <= 64");
    int rc = bread_states(b);
    if (rc) return rc;
    for (int32_t k = 0; k < b->count; ++k)
        if (b->h_st[k].pending) return bfail(b, PDOT_E_STATE, "pdot_batch_set_launch: a pass is pending");
    b->cols_per_tile = cols_per_tile;
    b->grid_mult = grid_mult;
    return setup_batch_partials(b);
}

int pdot_batch_get_launch(pdot_batch* b, int32_t* rows_per_tile_out, int32_t* cols_per_tile, int32_t* tiles,
                          int32_t* step_grid) {
    if (!b) return PDOT_E_ARG;
    if (rows_per_tile_out) *rows_per_tile_out = rows_per_tile(b->storage);
    if (cols_per_tile) *cols_per_tile = b->cols_per_tile;
    if (tiles) *tiles = b->ntiles;
    if (step_grid) *step_grid = b->step_grid;
    return PDOT_OK;
}

int64_t pdot_batch_device_bytes(const pdot_batch* b) { return b ? b->bytes : 0; }

}  // extern "C"
