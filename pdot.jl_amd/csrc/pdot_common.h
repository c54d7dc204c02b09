// pdot_common.h -- device code shared by the single-instance and the batched
// PDHG kernels of libpdot_hip.so (gfx950): Julia-semantics scalar helpers,
// fixed-order DPP/LDS reductions, the step-size controller, and the bodies of
// one step tile and one reduce chunk.  Included by pdot_hip.hip and
// pdot_batch.hip; everything lives in an anonymous namespace, one copy per
// translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pdot.h"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 512;  // small single-block finishing kernels
#ifndef PDOT_RED_WAVES
#define PDOT_RED_WAVES 8
#endif
constexpr int kRedWaves = PDOT_RED_WAVES;     // reduce kernel: waves per block
constexpr int kRedBlock = kWave * kRedWaves;  // each block owns 64 rows or columns
constexpr int kChunk = 64;        // columns per LDS flush of column sums
constexpr int kLdAlign = 64;      // leading-dimension alignment (elements)
constexpr int kMatStats = 10;     // per-block matrix eval reductions
constexpr int kVecStats = 12;     // per-block vector eval reductions
constexpr int kEvalOut = kMatStats + kVecStats;
// Shape of the single-problem step kernel per storage type (measured best on
// MI355X over whole 10k x 10k solves with zero maps): waves per block
// (stacked along rows), columns per load batch, and how many batches are
// prefetched ahead of the one computed.
#ifndef PDOT_F64_WAVES
#define PDOT_F64_WAVES 4
#endif
#ifndef PDOT_F64_UNROLL
#define PDOT_F64_UNROLL 1
#endif
#ifndef PDOT_F64_PIPE
#define PDOT_F64_PIPE 1
#endif
#ifndef PDOT_F32_WAVES
#define PDOT_F32_WAVES 8
#endif
#ifndef PDOT_F32_UNROLL
#define PDOT_F32_UNROLL 2
#endif
#ifndef PDOT_F32_PIPE
#define PDOT_F32_PIPE 0
#endif
#ifndef PDOT_F64P_WAVES
#define PDOT_F64P_WAVES PDOT_F64_WAVES  // the dense geometry: FP64 point solves equal dense ones bit for bit
#endif
#ifndef PDOT_F64P_UNROLL
#define PDOT_F64P_UNROLL 2  // two columns per batch, no prefetch: measured best with deferred averaging
#endif
#ifndef PDOT_F64P_PIPE
#define PDOT_F64P_PIPE 0
#endif
#ifndef PDOT_F32P_WAVES
#define PDOT_F32P_WAVES 4
#endif
#ifndef PDOT_F32P_UNROLL
#define PDOT_F32P_UNROLL 2
#endif
#ifndef PDOT_F32P_PIPE
#define PDOT_F32P_PIPE 0  // the cost's registers leave no room for a prefetched batch
#endif
// Columns per batch and batches prefetched in the kernel of the passes that
// leave the average alone (kKindLeave below), for the forms that have one;
// the waves per block and the launch geometry are the generic kernel's, so
// both kinds share one layout of the partial sums.
#ifndef PDOT_F32P_LEAVE_UNROLL
#define PDOT_F32P_LEAVE_UNROLL 4  // four columns per batch once the average's registers are free
#endif
#ifndef PDOT_F32P_LEAVE_PIPE
#define PDOT_F32P_LEAVE_PIPE 0
#endif

// Kinds of an ordinary pass: kKindAny decides at run time whether it reads and
// writes the running average; kKindLeave is a pass known to leave the average
// alone (its update deferred to the next pass), compiled without the
// average's loads, registers and stores.
enum { kKindAny = 0, kKindLeave = 1 };

// POINTS: the cost is formed from point clouds (more registers per lane, no
// C stream), which has its own best shape.
template <typename T, bool POINTS = false>
struct StepShape;
template <>
struct StepShape<double, false> {
    static constexpr int kWaves = PDOT_F64_WAVES, kUnroll = PDOT_F64_UNROLL, kPipe = PDOT_F64_PIPE;
};
template <>
struct StepShape<float, false> {
    static constexpr int kWaves = PDOT_F32_WAVES, kUnroll = PDOT_F32_UNROLL, kPipe = PDOT_F32_PIPE;
};
template <>
struct StepShape<double, true> {
    static constexpr int kWaves = PDOT_F64P_WAVES, kUnroll = PDOT_F64P_UNROLL, kPipe = PDOT_F64P_PIPE;
};
template <>
struct StepShape<float, true> {
    static constexpr int kWaves = PDOT_F32P_WAVES, kUnroll = PDOT_F32P_UNROLL, kPipe = PDOT_F32P_PIPE;
};
// The leave-alone kind's shape, and whether a form has a kernel of its own for
// it.  Measured at 10k x 10k per iteration (both kinds on the generic
// kernel's geometry): FP32 point clouds 4% faster with it (142 -> 126
// registers with four columns per batch); FP32 dense and FP64 point clouds
// within 0.5%, FP64 dense no faster (418 us against 420 us for that kind of
// pass), so those keep the one kernel.
template <typename T, bool POINTS = false>
struct LeaveShape {
    static constexpr bool kOwnKernel = false;
    static constexpr int kUnroll = StepShape<T, POINTS>::kUnroll, kPipe = StepShape<T, POINTS>::kPipe;
};
template <>
struct LeaveShape<float, true> {
    static constexpr bool kOwnKernel = true;
    static constexpr int kUnroll = PDOT_F32P_LEAVE_UNROLL, kPipe = PDOT_F32P_LEAVE_PIPE;
};
#ifndef PDOT_DENSE_LEAVE
#define PDOT_DENSE_LEAVE 1
#endif
template <>
struct LeaveShape<double, false> {
    static constexpr bool kOwnKernel = PDOT_DENSE_LEAVE;
    static constexpr int kUnroll = PDOT_F64_UNROLL, kPipe = PDOT_F64_PIPE;
};
template <>
struct LeaveShape<float, false> {
    static constexpr bool kOwnKernel = PDOT_DENSE_LEAVE;
    static constexpr int kUnroll = PDOT_F32_UNROLL, kPipe = PDOT_F32_PIPE;
};

// ------------------------------------------------------------------ helpers

// Julia's max(x, 0.0) / min(x, 0.0) / min(x, y) propagate NaN.
__device__ __forceinline__ double jl_max0(double x) { return x != x ? x : (x > 0.0 ? x : 0.0); }
__device__ __forceinline__ double jl_min0(double x) { return x != x ? x : (x < 0.0 ? x : 0.0); }
__device__ __forceinline__ double jl_min(double x, double y) {
    if (x != x || y != y) return __longlong_as_double(0x7ff8000000000000LL);
    return x < y ? x : y;
}
// max of non-negative values that keeps a NaN once seen
__device__ __forceinline__ double nanmax(double acc, double t) { return (acc != acc || t != t || t > acc) ? (acc != acc ? acc : t) : acc; }

// (k+1)^e as Julia's Int^Float64: integral exponents by repeated products.
__device__ inline double pow_int_base(double x, double e) {
    if (e == floor(e) && fabs(e) <= 64.0) {
        int k = (int)fabs(e);
        double r = 1.0;
        for (int t = 0; t < k; ++t) r *= x;
        return e < 0 ? 1.0 / r : r;
    }
    return pow(x, e);
}

// DPP lane movement of a double (two 32-bit halves); lanes outside row_mask
// or reading an invalid source receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double x) {
    int lo = __double2loint(x);
    int hi = __double2hiint(x);
    if constexpr (ROW_MASK == 0xF) {
        // every lane is written, from a lane of its own row: no old value to set
        lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
        hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    } else {
        // rows left out keep 0
        lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xF, false);
        hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xF, false);
    }
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_f64(double x, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave in a fixed DPP tree; every lane gets the
// total.  Must be called with all lanes active.
// Sum over each 16-lane row of a wave (the first four steps of wave_sum's
// tree); every lane of a row gets its row's total.  All lanes active.
constexpr int kWaveRows = 4;
__device__ __forceinline__ double row_sum(double x) {
    x += dpp_f64<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
    x += dpp_f64<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
    x += dpp_f64<0x141, 0xF>(x);  // row_half_mirror
    x += dpp_f64<0x140, 0xF>(x);  // row_mirror
    return x;
}
// The wave total from the four row totals, added in the order wave_sum's last
// two steps add them: (r3 + r2) + (r1 + r0).
__device__ __forceinline__ double rows_total(const double* r) {
    return (r[3] + r[2]) + (r[1] + r[0]);
}
// LDS rows of column sums per wave in step_tile: FP64 storage leaves the last
// two steps of the tree to the flush (fewer vector instructions per column);
// FP32 storage reduces each column to the wave total (measured faster there).
template <typename T>
constexpr int col_rows() { return sizeof(T) == 8 ? kWaveRows : 1; }
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_f64<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
    x += dpp_f64<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
    x += dpp_f64<0x141, 0xF>(x);  // row_half_mirror
    x += dpp_f64<0x140, 0xF>(x);  // row_mirror
    x += dpp_f64<0x142, 0xA>(x);  // row_bcast:15
    x += dpp_f64<0x143, 0xC>(x);  // row_bcast:31
    return readlane_f64(x, 63);
}
__device__ __forceinline__ double wave_max(double x) {
    x = nanmax(x, dpp_f64<0xB1, 0xF>(x));
    x = nanmax(x, dpp_f64<0x4E, 0xF>(x));
    x = nanmax(x, dpp_f64<0x141, 0xF>(x));
    x = nanmax(x, dpp_f64<0x140, 0xF>(x));
    x = nanmax(x, dpp_f64<0x142, 0xA>(x));
    x = nanmax(x, dpp_f64<0x143, 0xC>(x));
    return readlane_f64(x, 63);
}

// Deterministic block-wide sum/max (power-of-two block size); result valid in all threads.
__device__ inline double block_sum(double v, double* red) {
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = v;
    __syncthreads();
    for (int s = (int)blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = red[tid] + red[tid + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}
__device__ inline double block_max(double v, double* red) {
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = v;
    __syncthreads();
    for (int s = (int)blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = nanmax(red[tid], red[tid + s]);
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// --------------------------------------------------------------- controller

// The two slots a pass writes: the first slots holding neither the current
// point nor the (non-empty) average.
__device__ __forceinline__ void choose_spares(pdot_solver_state& st) {
    int used0 = st.cur_slot, used1 = st.avg_count > 0 ? st.avg_slot : st.cur_slot;
    int k = 0;
    for (int s = 0; s < 4; ++s) {
        if (s == used0 || s == used1) continue;
        if (k == 0) st.spare0 = s;
        else if (k == 1) st.spare1 = s;
        ++k;
    }
}

// Averaging weights of a pass from state st (saddle_point_gpu.jl:129-130,
// weight = the pass's step size): new = (avg + x*nw)*aw; nw = aw = 1 for the
// first point after a reset, so that the average equals it bit for bit.
__device__ __forceinline__ void avg_weights(const pdot_solver_state& st, double* nw, double* aw) {
    *nw = 1.0;
    *aw = 1.0;
    if (st.sum_weights > 0.0) {
        *aw = st.sum_weights / (st.sum_weights + st.step_size);
        *nw = st.step_size / st.sum_weights;
    }
}

// Whether an ordinary pass from state st defers the average's update: only
// where allowed (defer_ok), never for an evaluation's pass, and only
// when no contribution is already waiting and the average is not empty (the
// first pass after a reset writes no average anyway: it aliases the point).
// The step kernel and the commit decide alike from the same state.
__device__ __forceinline__ bool pass_defers(const pdot_solver_state& st, bool eval) {
    return st.defer_ok != 0 && !eval && st.avg_pending == 0 && st.avg_count > 0;
}

// Applies the average's lagging contribution in place: avg = (avg + cur*nw)*aw
// elementwise, stored as T, exactly as the next pass would have applied it
// in registers (saddle_point_gpu.jl:129-130).
template <typename T>
__global__ __launch_bounds__(256) void materialize_kernel(T* __restrict__ avg, const T* __restrict__ cur, int64_t total,
                                                          double nw, double aw) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
        const double t = (double)avg[k] + (double)cur[k] * nw;
        avg[k] = (T)(t * aw);
    }
}

// The same for the four vector averages (u, r of length n; v, c of length m).
struct VecAverages {
    double* avg[4];
    const double* cur[4];
    int64_t len[4];
};

__global__ __launch_bounds__(256) void materialize_vectors_kernel(VecAverages p, double nw, double aw) {
    const int w = blockIdx.y;
    double* __restrict__ avg = p.avg[w];
    const double* __restrict__ cur = p.cur[w];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < p.len[w]; k += (int64_t)gridDim.x * blockDim.x) {
        const double t = avg[k] + cur[k] * nw;
        avg[k] = t * aw;
    }
}

// Commit the pending pass: take_step!'s acceptance test and step-size
// recurrence (src/solver_gpu.jl:296-366), its replay of a rejected pass (the
// replay recomputes an identical pass, so only the scalars change), and the
// swap of the spare slots into the current point and the average
// (update_solution_in_solver_state!, :237-264, whose averaging weight is the
// pre-step eta).  A pass with interaction > 0 and movement == 0 sets
// numerical_error and is not applied (:338-344); a NaN step-size limit, on
// which the reference loops forever, is reported the same way.
__device__ inline void controller_commit(pdot_solver_state& st, double dot, double yy, double dx2) {
    const bool was_eval = st.pending != 0;
    st.pending = 0;
    if (st.step_budget > 0) st.step_budget -= 1;
    const double eta = st.step_size;
    const double omega = st.primal_weight;
    bool commit;
    st.take_step_calls += 1;
    st.last_dx2 = dx2;
    st.last_dot = dot;
    st.last_dy2 = yy;
    if (st.step_policy == PDOT_POLICY_CONSTANT) {
        st.cumulative_kkt_passes += 1;
        commit = true;
    } else {
        const double interaction = fabs(dot);
        const double nd = sqrt(dx2);
        const double ny = sqrt(yy);
        const double movement = 0.5 * omega * (nd * nd) + (0.5 / omega) * (ny * ny);
        st.last_interaction = interaction;
        st.last_movement = movement;
        double L = __longlong_as_double(0x7ff0000000000000LL);
        double eta_loc = eta;
        bool done = false, replay = false, ne = false;
        for (int pass = 0; pass < 4 && !done; ++pass) {
            st.total_number_iterations += 1;
            st.cumulative_kkt_passes += 1;
            if (!replay) {
                if (interaction > 0) {
                    L = movement / interaction;
                    if (movement == 0.0) {
                        ne = true;
                        break;
                    }
                } else {
                    L = __longlong_as_double(0x7ff0000000000000LL);
                }
            }
            if (eta_loc <= L) done = true;
            const double k1 = (double)(st.total_number_iterations + 1);
            const double first = (1 - 1 / pow_int_base(k1, st.reduction_exponent)) * L;
            const double second = (1 + 1 / pow_int_base(k1, st.growth_exponent)) * eta_loc;
            eta_loc = jl_min(first, second);
            replay = true;
            if (!done && eta_loc != eta_loc) {
                ne = true;
                break;
            }
        }
        if (!done) ne = true;
        st.last_step_limit = L;
        st.step_size = eta_loc;
        if (ne) {
            st.numerical_error = 1;
            st.error_call = st.take_step_calls;
        }
        commit = done;
    }
    if (commit) {
        if (pass_defers(st, was_eval)) {
            // the average stays in its slot; the new point's contribution waits
            // for the next pass (or for materialize_average) with this pass's
            // weights, formed from the pre-step eta as the step kernel formed them
            const double S = st.sum_weights;
            st.pend_nw = 1.0;
            st.pend_aw = 1.0;
            if (S > 0.0) {
                st.pend_aw = S / (S + eta);
                st.pend_nw = eta / S;
            }
            st.avg_pending = 1;
        } else {
            st.avg_slot = (st.avg_count == 0) ? st.spare0 : st.spare1;
            st.avg_pending = 0;
        }
        st.cur_slot = st.spare0;
        st.sum_weights += eta;
        st.avg_count += 1;
        st.gzero = 0;
        choose_spares(st);
    }
}

// Publishing a block's results to the last block of the same kernel without
// a device-scope fence (on gfx950 an L2 write-back, and an invalidate on the
// reading side): the results are written with agent-scope stores (written
// through to the agent's coherence point), the writer waits until its
// memory operations have completed, and only then counts itself done with an
// agent-scope atomic; the block that counts last reads the results with
// agent-scope loads.  This is the release sequence without the L2 write-back,
// which only plain stores need.
__device__ __forceinline__ void publish_f64(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double acquire_f64(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until this wave's earlier memory operations have completed; the
// signal fences keep the compiler from moving memory operations across it.
__device__ __forceinline__ void wait_memory_ops() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt(0) expcnt(0) lgkmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Counts one publisher done (after wait_memory_ops); true for the last of
// `total`, which also resets the counter for the next launch.
__device__ __forceinline__ bool count_done(uint32_t* counter, uint32_t total) {
    const uint32_t prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev != total - 1) return false;
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Reduce the per-block scalars of the reduce kernel (dot, ||dy||^2, ||dX||^2
// chunk) in a fixed order and commit.  Called by the 64 lanes of one wave;
// lane 0 updates *sst.  PUBLISHED: the scalars were written with
// publish_f64 in the same kernel and are read with acquire_f64, the loads
// of up to KU blocks per lane before the first sum (one memory round trip
// for up to 64 KU blocks, KU * 6 more registers).
template <bool PUBLISHED = false, int KU = 1>
__device__ inline void reduce_and_commit(pdot_solver_state* sst, const double* redpart, int nred) {
    const int lane = threadIdx.x & (kWave - 1);
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    if constexpr (PUBLISHED) {
        // the sums run in the same order as below
        constexpr int kU = KU;
        for (int k0 = lane; k0 < nred; k0 += kWave * kU) {
            double t0[kU], t1[kU], t2[kU];
#pragma unroll
            for (int q = 0; q < kU; ++q) {
                const int k = k0 + q * kWave;
                t0[q] = k < nred ? acquire_f64(&redpart[3 * k + 0]) : 0.0;
                t1[q] = k < nred ? acquire_f64(&redpart[3 * k + 1]) : 0.0;
                t2[q] = k < nred ? acquire_f64(&redpart[3 * k + 2]) : 0.0;
            }
#pragma unroll
            for (int q = 0; q < kU; ++q) {
                if (k0 + q * kWave < nred) {
                    d0 += t0[q];
                    d1 += t1[q];
                    d2 += t2[q];
                }
            }
        }
    } else {
        for (int k = lane; k < nred; k += kWave) {
            d0 += redpart[3 * k + 0];
            d1 += redpart[3 * k + 1];
            d2 += redpart[3 * k + 2];
        }
    }
    d0 = wave_sum(d0);
    d1 = wave_sum(d1);
    d2 = wave_sum(d2);
    if (lane == 0) controller_commit(*sst, d0, d1, d2);
}

// ---------------------------------------------------------------- vectors

typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T> struct VecT;
template <> struct VecT<double> {
    static constexpr int N = 2;
    using type = dbl2;
};
template <> struct VecT<float> {
    static constexpr int N = 4;
    using type = flt4;
};

// 16-byte loads/stores of the streamed matrices.  NT: with the nontemporal
// hint, for matrices touched once per pass and far larger than the caches;
// without it the lines may stay in the caches (the 256 MB Infinity Cache)
// for the next pass, which pays where a pass's matrices nearly fit there.
template <bool NT, typename V>
__device__ __forceinline__ V ld16(const V* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void st16(V v, V* p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <typename V, typename T>
__device__ __forceinline__ void unpack(const V& v, double* out);
template <>
__device__ __forceinline__ void unpack<dbl2, double>(const dbl2& v, double* o) {
    o[0] = v.x;
    o[1] = v.y;
}
template <>
__device__ __forceinline__ void unpack<flt4, float>(const flt4& v, double* o) {
    o[0] = (double)v.x;
    o[1] = (double)v.y;
    o[2] = (double)v.z;
    o[3] = (double)v.w;
}

template <bool NT>
__device__ __forceinline__ void store_vec(double* p, const double* o) {
    dbl2 v = {o[0], o[1]};
    st16<NT>(v, reinterpret_cast<dbl2*>(p));
}
template <bool NT>
__device__ __forceinline__ void store_vec(float* p, const double* o) {
    flt4 v = {(float)o[0], (float)o[1], (float)o[2], (float)o[3]};
    st16<NT>(v, reinterpret_cast<flt4*>(p));
}

// True unless every value is +0.0 (a -0.0 counts as a value).
template <int N>
__device__ __forceinline__ bool any_bits(const double* o) {
    uint64_t b = 0;
#pragma unroll
    for (int e = 0; e < N; ++e) b |= (uint64_t)__double_as_longlong(o[e]);
    return b != 0;
}

// Lane `lane`'s bit of a wave-uniform word.
__device__ __forceinline__ bool lane_bit(uint64_t word, int lane) {
    const uint32_t half = lane < 32 ? (uint32_t)word : (uint32_t)(word >> 32);
    return (half >> (lane & 31)) & 1;
}

// Set bits of a wave-uniform word below this lane: where this lane's 16
// bytes sit in a compacted segment (see TileArgs).
__device__ __forceinline__ int slot_in_segment(uint64_t word) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(word >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)word, 0u));
}

// Geometry of the zero maps: word ((rb * Cb + cb) * waves + w) * tc + t
// covers, in column j of tile column t (j = cb + t Cb interleaved, else
// cb TC + t), the vec rows row0 + (w * 64 + L) * vec of lane L, row0 =
// rb * TRB, for the lanes below the row block's end.
struct ZGeom {
    int64_t n = 0, m = 0;
    int64_t ld = 0;
    int32_t R = 0, Cb = 0, TRB = 0, TC = 0, interleave = 0, waves = 0, vec = 0;
    int64_t tc = 0;     // words per wave and tile
    int64_t words = 0;  // R * Cb * waves * tc
};

// Expands every wave segment of a slot to the dense layout in place: lane L
// takes its 16 bytes from position popcount(bits below L) of the segment, or
// +0.0 where its bit is clear, so that the slot's memory holds every value in
// column-major order (the map must then be set to all ones, which describes
// that layout).  One wave per word; every lane's load completes before any
// lane of the wave stores.
template <typename T>
__global__ __launch_bounds__(256) void densify_kernel(T* __restrict__ X, const uint64_t* __restrict__ z, ZGeom g) {
    using V = typename VecT<T>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t wi = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; wi < g.words; wi += nw) {
        const int64_t t = wi % g.tc;
        int64_t r = wi / g.tc;
        const int64_t w = r % g.waves;
        r /= g.waves;
        const int64_t cb = r % g.Cb, rb = r / g.Cb;
        int64_t ncols, j;
        if (g.interleave) {
            ncols = (g.m - cb + g.Cb - 1) / g.Cb;
            j = cb + t * g.Cb;
        } else {
            const int64_t j0 = cb * g.TC;
            ncols = ((j0 + g.TC < g.m) ? j0 + g.TC : g.m) - j0;
            j = j0 + t;
        }
        const int64_t row0 = rb * g.TRB;
        const int64_t row_end = (row0 + g.TRB < g.n) ? row0 + g.TRB : g.n;
        const int64_t seg = j * g.ld + row0 + w * kWave * g.vec;  // wave-uniform
        const uint64_t word = z[wi];
        if (t >= ncols || word == ~0ull) continue;  // wave-uniform: past the tile, or dense already
        const bool active = row0 + (w * kWave + lane) * g.vec < row_end;
        V v = {};
        if (active && lane_bit(word, lane)) v = *reinterpret_cast<const V*>(X + seg + (int64_t)slot_in_segment(word) * g.vec);
        wait_memory_ops();
        if (active) *reinterpret_cast<V*>(X + seg + (int64_t)lane * g.vec) = v;
    }
}

__global__ __launch_bounds__(256) void fill_u64_kernel(uint64_t* __restrict__ z, int64_t count, uint64_t value) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = value;
}


// ------------------------------------------------------------ step tile

// Where the cost C_ij comes from: the stored matrix, or formed in registers
// from point clouds (source x: dim x n, target y: m x dim) exactly as the
// dense generators form it, C_ij = (sum_d f(y_jd - x_id)) / cdiv, with
// f(t) = t*t (squared Euclidean) or |t| (L1), summed from 0 in d order.
enum { kCostDense = 0, kCostSqEuclid = 1, kCostL1 = 2 };
constexpr int kMaxDim = 3;

// a / s, correctly rounded, from y = RN(1/s): one multiply and two Markstein
// corrections (fused multiply-adds whose residuals a - q*s are exact).  With
// y within half an ulp of 1/s and the once-corrected quotient within one ulp
// of a/s, Markstein's theorem makes the second correction the correctly
// rounded quotient, the value of the IEEE division a / s.  Zero, tiny, huge
// and non-finite quotients, and scales whose reciprocal could lose precision,
// take the division itself, and so do numerators below 2^-960, whose
// residuals a - q*s could fall below the subnormal granularity 2^-1074 (the
// relative bound alone underflows when s < 2^-122).
__device__ __forceinline__ double exact_quotient(double a, double s, double y) {
    // the bounds depend on s only (loop-invariant in the step kernels):
    // a in [max(2^-960, 2^-900 s), 2^900 s], which a NaN fails
    const bool s_ok = s >= 0x1p-1000 && s <= 0x1p+1000;
    const double lo = fmax(0x1p-960, 0x1p-900 * s), hi = 0x1p+900 * s;
    if (!(s_ok && a >= lo && a <= hi)) return a / s;
    const double q0 = a * y;
    const double r0 = fma(-q0, s, a);
    const double q1 = fma(r0, y, q0);
    const double r1 = fma(-q1, s, a);
    return fma(r1, y, q1);
}

// crcp = 1.0 / cdiv (the IEEE division, once per thread).  EXACT: the
// quotient is the correctly rounded one, so the costs equal the dense
// generators' FP64 matrix bit for bit (FP64 storage).  Otherwise (FP32
// storage, whose dense form stores C rounded to FP32) the quotient is the
// product c * crcp, within about one FP64 ulp of c / cdiv, 2^29 times finer
// than the FP32 rounding, with no branch per element.  It equals c / cdiv
// for cdiv = 1, 0, +Inf and NaN; pdot_upload_points rejects the other scales
// whose reciprocal is not a normal number (finite cdiv outside
// [2^-1022, 2^1022]) for FP32 storage.
template <int COST, bool EXACT>
__device__ __forceinline__ double point_cost(const double* xi, const double* yj, int dim, double cdiv, double crcp) {
    // the sum starts from the first term: 0.0 + t*t and 0.0 + |t| equal the
    // term itself in every case (neither is ever -0.0), so the dense
    // generators' sum from 0.0 is reproduced without its first addition
    double c = (COST == kCostSqEuclid) ? (yj[0] - xi[0]) * (yj[0] - xi[0]) : fabs(yj[0] - xi[0]);
#pragma unroll
    for (int d = 1; d < kMaxDim; ++d) {
        if (d < dim) {
            const double t = yj[d] - xi[d];
            c += (COST == kCostSqEuclid) ? t * t : fabs(t);
        }
    }
    if constexpr (EXACT) {
        // x / 1 == x: no division for unscaled costs
        return cdiv == 1.0 ? c : exact_quotient(c, cdiv, crcp);
    } else {
        return c * crcp;
    }
}

// Everything one block needs to advance one tile of one problem by a PDHG
// pass: rows [row0, row_end) (at most WAVES*64*VEC rows; row_end - row0 a
// multiple of VEC unless row_end = n) of the columns j0 + t*jstep,
// t < ncols.
template <typename T>
struct TileArgs {
    const T* C;          // dense cost (kCostDense)
    const double* px;    // point costs: source coordinates, dim x n (leading dimension n)
    const double* py;    // target coordinates, m x dim
    double cdiv;
    int dim;
    const T* Xc;   // current point
    const T* Xa;   // running average (read when !avg_empty or EVAL)
    T* Xn;         // new point
    T* Xan;        // new average
    const double* uc;
    const double* ua;
    const double* vc;
    const double* va;
    double* rowpart;  // partial row sums of this tile, indexed by row
    double* colpart;  // partial column sums of this tile, indexed by column
    int64_t n, ld, row0, row_end, j0, jstep, ncols;
    double eta, omega, nw, aw;
    bool avg_empty, gzero, do_step;
    // The average's update may lag one pass behind: a
    // deferring pass neither reads nor writes the average; the next pass
    // first applies the lagging contribution of its input point, with the
    // weights pnw, paw, to the average it loads (in registers, rounded to
    // the storage type as a store would), then its own.
    bool defer, pend;
    double pnw, paw;
    // Zero maps (sparse passes): one 64-bit word per wave and tile column,
    // bit L for lane L's VEC rows (16 bytes).  A clear bit says those rows
    // of that column are +0.0 in that slot; the set bits' 16-byte units are
    // stored compacted at the front of the wave's segment of the column (its
    // 64 * VEC rows from the wave's first row), lane L's at position
    // popcount(bits below L), and nothing else of the segment is read or
    // written.  All bits set is the dense layout.  Wave w's word for tile
    // column t is at z* + w * zstride + t.
    bool sparse;
    const uint64_t* zc;  // current point's map
    const uint64_t* za;  // average's map
    uint64_t* zn;        // new point's map
    uint64_t* zan;       // new average's map
    int64_t zstride;
};

// One PDHG pass over a tile (src/solver_gpu.jl:125-174 and 237-264,
// src/saddle_point_gpu.jl:115-135): each lane owns VEC consecutive rows and
// walks the tile's columns.  Row sums accumulate in registers, column sums
// are reduced per column across the wave (DPP; FP64 storage: across each
// 16-lane row, the rows' totals added at the flush) and across the block's
// waves (LDS, every kChunk columns; colbuf holds WAVES * col_rows<T>() rows).  With EVAL the same loads also produce the
// evaluation reductions of the input point and of its average.  Called by
// all threads of the block; thread 0 receives ||Delta||^2 of the tile in
// *dd_out and the kMatStats evaluation reductions in ev_out.  NT: the
// matrices are streamed with the nontemporal hint (the arithmetic and its
// order do not depend on it).  ZMAP: the tile may use zero maps (a.sparse).
template <typename T, bool EVAL, int UNROLL, int WAVES, int PIPE, int COST = kCostDense, int KIND = kKindAny,
          bool NT = true, bool ZMAP = true>
__device__ __forceinline__ void step_tile(const TileArgs<T>& a, double (*colbuf)[kChunk], double* red, double* dd_out,
                                          double* ev_out) {
    static_assert(!(EVAL && KIND == kKindLeave), "an evaluation's pass always updates the average");
    constexpr int VEC = VecT<T>::N;
    using V = typename VecT<T>::type;
    constexpr int WROWS = kWave * VEC;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const int64_t n = a.n, ld = a.ld;
    // the wave's first row (wave-uniform) and this lane's offset from it
    const int64_t wrow0 = a.row0 + (int64_t)__builtin_amdgcn_readfirstlane(w) * WROWS;
    const uint32_t loff = (uint32_t)lane * VEC;
    const int64_t i0 = wrow0 + loff;
    const int64_t row_end = a.row_end;
    const bool lane_active = i0 < row_end;
    const bool avg_empty = a.avg_empty;
    const bool do_step = a.do_step;
    // a deferring pass leaves the average alone (never an evaluation); so does
    // every pass of the leave-alone kind, which has no lagging contribution
    const bool skip_avg = (KIND == kKindLeave) || avg_empty || (!EVAL && a.defer);
    const bool pend = (KIND != kKindLeave) && a.pend && !avg_empty;
    const double pnw = a.pnw, paw = a.paw;
    const T* __restrict__ C = a.C;
    const T* __restrict__ Xc = a.Xc;
    const T* __restrict__ Xa = a.Xa;
    T* __restrict__ Xn = a.Xn;
    T* __restrict__ Xan = a.Xan;
    const double* __restrict__ vc = a.vc;
    const double* __restrict__ va = a.va;
    const double eta = a.eta, omega = a.omega, nw = a.nw, aw = a.aw;
    const double tau = eta / omega;
    const bool gzero = a.gzero;
    const int64_t j0 = a.j0, jstep = a.jstep, ncols = a.ncols;
    const double crcp = COST != kCostDense ? 1.0 / a.cdiv : 1.0;  // point costs: the reciprocal of the scale
    // zero maps: the words of a batch's columns are read (scalar loads, into
    // zwc / zwa) while the batch before it is loaded, so they have arrived
    // by the time its loads need them.  All bits set: read every lane.
    const bool sparse = ZMAP && a.sparse;
    const int64_t zoff = ZMAP ? (int64_t)__builtin_amdgcn_readfirstlane(w) * a.zstride : 0;
    typedef const __attribute__((address_space(4))) uint64_t* ZPtr;
    const ZPtr zc = (ZPtr)(a.zc + zoff);
    const ZPtr za = (ZPtr)(a.za + zoff);
    uint64_t* __restrict__ zn = a.zn + zoff;
    uint64_t* __restrict__ zan = a.zan + zoff;
    uint64_t zwc[UNROLL], zwa[UNROLL];
    auto load_zwords = [&](int64_t tb) {  // block-uniform
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
            zwc[q] = ~0ull;
            zwa[q] = ~0ull;
            if (sparse && tb + q < ncols) {
                zwc[q] = zc[tb + q];
                if (!skip_avg) zwa[q] = za[tb + q];
            }
        }
    };
    double ui[VEC], uai[VEC];
    bool valid[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        valid[e] = (i0 + e) < row_end;
        ui[e] = (lane_active && valid[e]) ? a.uc[i0 + e] : 0.0;
        uai[e] = (EVAL && lane_active && valid[e]) ? a.ua[i0 + e] : 0.0;
        if (EVAL && pend) {  // the average's lagging contribution of u
            const double t = uai[e] + ui[e] * pnw;
            uai[e] = t * paw;
        }
    }
    // point costs: this lane's source coordinates stay in registers
    double xi[VEC][kMaxDim];
#pragma unroll
    for (int e = 0; e < VEC; ++e)
#pragma unroll
        for (int d = 0; d < kMaxDim; ++d)
            xi[e][d] = (COST != kCostDense && d < a.dim && lane_active && valid[e]) ? a.px[(int64_t)d * n + i0 + e] : 0.0;
    double rs[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) rs[e] = 0.0;
    double dd = 0.0;
    double ev[kMatStats];
#pragma unroll
    for (int k = 0; k < kMatStats; ++k) ev[k] = 0.0;

    // A batch is UNROLL consecutive tile columns whose loads are issued
    // together; kChunk is a multiple of UNROLL, so a batch never straddles a
    // column-sum flush.  With PIPE > 0 the loads of the next PIPE batches are
    // issued before a batch is computed (a ring of PIPE + 1 register sets), so
    // every wave keeps loads in flight while it computes and stores.
    struct Batch {
        V xv[UNROLL], cv[UNROLL], av[UNROLL];
        double vj[UNROLL], vaj[UNROLL];
        double yj[UNROLL][COST == kCostDense ? 1 : kMaxDim];
    };
    auto load_batch = [&](int64_t tb, Batch& B) {
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
            const bool colok = tb + q < ncols;
            const int64_t j = j0 + (tb + q) * jstep;
            B.vj[q] = colok ? vc[j] : 0.0;
            B.vaj[q] = (EVAL && colok) ? va[j] : 0.0;
            if constexpr (COST != kCostDense) {
#pragma unroll
                for (int d = 0; d < kMaxDim; ++d) B.yj[q][d] = (colok && d < a.dim) ? a.py[j * a.dim + d] : 0.0;
            }
            const uint64_t wc = zwc[q], wa = zwa[q];
            if (colok && lane_active) {
                const int64_t seg = j * ld + wrow0;  // the wave's segment of column j (wave-uniform)
                // a lane whose rows are recorded as +0.0 issues no load
                V xv = {};
                if (__builtin_amdgcn_inverse_ballot_w64(wc))
                    xv = ld16<NT>(reinterpret_cast<const V*>(Xc + seg) + (uint32_t)slot_in_segment(wc));
                B.xv[q] = xv;
                if constexpr (COST == kCostDense) B.cv[q] = ld16<NT>(reinterpret_cast<const V*>(C + seg + loff));
                if (!skip_avg) {
                    V av = {};
                    if (__builtin_amdgcn_inverse_ballot_w64(wa))
                        av = ld16<NT>(reinterpret_cast<const V*>(Xa + seg) + (uint32_t)slot_in_segment(wa));
                    B.av[q] = av;
                }
            }
        }
        load_zwords(tb + UNROLL);  // the next batch to be loaded
    };
    auto compute_batch = [&](int64_t tb, const Batch& B) {
        const int64_t tc = tb / kChunk * kChunk;
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
            const bool colok = tb + q < ncols;
            const int64_t j = j0 + (tb + q) * jstep;
            double csl = 0.0;
            double xn[VEC], xan[VEC];
            bool live = false, alive = false;
            if (colok && lane_active) {
                double x[VEC], c[VEC], xa[VEC];
                unpack<V, T>(B.xv[q], x);
                if constexpr (COST == kCostDense) {
                    unpack<V, T>(B.cv[q], c);
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) c[e] = point_cost<COST, sizeof(T) == 8>(xi[e], B.yj[q], a.dim, a.cdiv, crcp);
                }
                if (!skip_avg) {
                    unpack<V, T>(B.av[q], xa);
                    if (pend) {
                        // the lagging contribution of the input point (the previous
                        // pass's new point): avg = (avg + x*nw)*aw, stored as T
#pragma unroll
                        for (int e = 0; e < VEC; ++e) {
                            const double t = xa[e] + x[e] * pnw;
                            xa[e] = (double)(T)(t * paw);
                        }
                    }
                }
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    // G_ij = (C_ij - u_i) - v_j (solver_gpu.jl:249-250); 0 at start (:434)
                    double g = (gzero || !valid[e]) ? 0.0 : (c[e] - ui[e]) - B.vj[q];
                    if (EVAL) {
                        ev[0] += c[e] * x[e];
                        ev[1] = nanmax(ev[1], fabs(x[e]));
                        ev[2] += x[e] * x[e];
                        const double rc = jl_min0(g);
                        ev[3] = nanmax(ev[3], fabs(rc));
                        ev[4] += rc * rc;
                        if (!avg_empty) {
                            double vaj = B.vaj[q];
                            if (pend) {  // the average's lagging contribution of v
                                const double t = vaj + B.vj[q] * pnw;
                                vaj = t * paw;
                            }
                            const double ga = valid[e] ? (c[e] - uai[e]) - vaj : 0.0;
                            ev[5] += c[e] * xa[e];
                            ev[6] = nanmax(ev[6], fabs(xa[e]));
                            ev[7] += xa[e] * xa[e];
                            const double rca = jl_min0(ga);
                            ev[8] = nanmax(ev[8], fabs(rca));
                            ev[9] += rca * rca;
                        }
                    }
                    // Delta = max(X - (eta/omega) G, 0) - X  (solver_gpu.jl:139-141)
                    double d = x[e] - tau * g;
                    d = jl_max0(d);
                    double xnew;
                    if constexpr (sizeof(T) == 8) {
                        d -= x[e];
                        xnew = x[e] + d;  // X .+= Delta (solver_gpu.jl:242)
                    } else {
                        xnew = (double)(float)d;  // FP32 storage: Delta of the stored values
                        d = xnew - x[e];
                    }
                    xn[e] = xnew;
                    rs[e] += d;
                    csl += d;
                    dd += d * d;
                    // avg = (avg + x*nw)*aw (saddle_point_gpu.jl:129-130)
                    if (!skip_avg) {
                        double t = xa[e] + xnew * nw;
                        xan[e] = t * aw;
                    }
                }
                live = any_bits<VEC>(xn);
                alive = !skip_avg && any_bits<VEC>(xan);
            }
            // block-uniform: all lanes take part in the ballots
            const uint64_t bn = sparse ? __builtin_amdgcn_ballot_w64(live) : ~0ull;
            const uint64_t ban = sparse ? __builtin_amdgcn_ballot_w64(alive) : ~0ull;
            if (do_step && colok) {
                if (lane_active) {  // the new points compacted by their own maps
                    const int64_t seg = j * ld + wrow0;
                    if (__builtin_amdgcn_inverse_ballot_w64(bn))
                        store_vec<NT>(Xn + seg + (uint32_t)slot_in_segment(bn) * VEC, xn);
                    if (!skip_avg && __builtin_amdgcn_inverse_ballot_w64(ban))
                        store_vec<NT>(Xan + seg + (uint32_t)slot_in_segment(ban) * VEC, xan);
                }
                if (sparse && lane == 0) {
                    zn[tb + q] = bn;
                    if (!skip_avg) zan[tb + q] = ban;
                }
            }
            if constexpr (col_rows<T>() == kWaveRows) {
                // row totals to LDS; the flush finishes each wave's total with
                // the additions wave_sum would make
                const double s = row_sum(csl);
                if (colok && (lane & 15) == 0) colbuf[w * kWaveRows + (lane >> 4)][tb + q - tc] = s;
            } else {
                const double s = wave_sum(csl);
                if (colok && lane == 0) colbuf[w][tb + q - tc] = s;
            }
        }
        // flush the column sums after the chunk's last batch (block-uniform)
        const int64_t tend = (tc + kChunk < ncols) ? tc + kChunk : ncols;
        if (tb + UNROLL >= tend) {
            __syncthreads();
            if (do_step && tid < tend - tc) {
                double t;
                if constexpr (col_rows<T>() == kWaveRows) {
                    double r[kWaveRows];
#pragma unroll
                    for (int k = 0; k < kWaveRows; ++k) r[k] = colbuf[k][tid];
                    t = rows_total(r);
#pragma unroll
                    for (int ww = 1; ww < WAVES; ++ww) {
#pragma unroll
                        for (int k = 0; k < kWaveRows; ++k) r[k] = colbuf[ww * kWaveRows + k][tid];
                        t += rows_total(r);
                    }
                } else {
                    t = colbuf[0][tid];
#pragma unroll
                    for (int ww = 1; ww < WAVES; ++ww) t += colbuf[ww][tid];
                }
                a.colpart[j0 + (tc + tid) * jstep] = t;
            }
            __syncthreads();
        }
    };
    static_assert(kChunk % UNROLL == 0, "a batch must not straddle a column-sum flush");
    load_zwords(0);
    Batch ring[PIPE + 1];
#pragma unroll
    for (int k = 0; k < PIPE; ++k)
        if ((int64_t)k * UNROLL < ncols) load_batch((int64_t)k * UNROLL, ring[k]);
    for (int64_t tb = 0; tb < ncols; tb += (int64_t)(PIPE + 1) * UNROLL) {
#pragma unroll
        for (int s = 0; s <= PIPE; ++s) {
            const int64_t t = tb + (int64_t)s * UNROLL;
            if (t >= ncols) break;  // block-uniform
            const int64_t ahead = t + (int64_t)PIPE * UNROLL;
            if (ahead < ncols) load_batch(ahead, ring[(s + PIPE) % (PIPE + 1)]);
            compute_batch(t, ring[s]);
        }
    }

    if (do_step && lane_active) {
#pragma unroll
        for (int e = 0; e < VEC; ++e) a.rowpart[i0 + e] = rs[e];
    }
    const double ddb = do_step ? block_sum(dd, red) : 0.0;
    if (tid == 0) *dd_out = ddb;
    if (EVAL) {
#pragma unroll
        for (int k = 0; k < kMatStats; ++k) {
            const bool is_max = (k == 1 || k == 3 || k == 6 || k == 8);
            const double t = is_max ? block_max(ev[k], red) : block_sum(ev[k], red);
            if (tid == 0) ev_out[k] = t;
        }
    }
}

// --------------------------------------------------------- reduce chunk

// The rows (or columns) [k0, k0 + 64) of one problem: their partial sums,
// the slot vectors they update, and this chunk's share of the per-tile
// ||Delta||^2 values.
struct ChunkArgs {
    const double* part;  // partials: part[t * pstride + k], t < np
    int64_t pstride;
    int32_t np;
    int64_t k0, len;
    bool rows;           // rows (u, r, a) or columns (v, c, b)
    const double* q;     // a or b
    const double* y_cur;
    const double* s_cur;
    const double* y_avg;
    const double* s_avg;
    double* y_new;
    double* s_new;
    double* y_anew;
    double* s_anew;
    const double* sqpart;
    int64_t sq_lo, sq_hi;
    double eta, omega, nw, aw;
    bool avg_empty, do_step;
    // the averages lag as the matrix average does: a deferring pass writes
    // none; a pass with a lagging contribution applies it (weights pnw, paw)
    // to the average it loads before its own
    bool defer, pend;
    double pnw, paw;
};

// Finishes one pass for a chunk (src/solver_gpu.jl:172-173, 177-234,
// 267-288; src/saddle_point_gpu.jl:140-221): r+ = rowsum(Delta) + r,
// u+ = u + eta*omega*(a - 2 r+ + r), the averages of u, v, r, c, and the
// chunk's parts of dr'du + dc'dv, ||dy||^2 and ||Delta||^2.  The partials are
// split into RWAVES contiguous ranges (one per wave) summed with eight
// independent accumulators, then combined in a fixed order by wave 0.  With
// EVAL, also the vector parts of the evaluation of the input point and its
// average.  Called by all threads; thread 0 receives out3 = {dot, yy, dx}
// and ev_out[kVecStats].
// The chunk's work after the partial sums (one wave; `sum` is this lane's
// total of the partials): the dual update, the averages, the chunk's parts
// of dr'du + dc'dv, ||dy||^2 and ||Delta||^2, and with EVAL the vector parts
// of the evaluation.  Lane 0 receives out3 and ev_out.
template <bool EVAL>
__device__ __forceinline__ void finish_chunk(const ChunkArgs& c, int lane, double sum, double* out3, double* ev_out) {
    const int64_t k = c.k0 + lane;
    const bool do_step = c.do_step;
    const bool avg_empty = c.avg_empty;
    const double eta = c.eta, omega = c.omega, nw = c.nw, aw = c.aw;
    double dot = 0.0, yy = 0.0;
    double ev[kVecStats];
#pragma unroll
    for (int t = 0; t < kVecStats; ++t) ev[t] = 0.0;
    if (k < c.len) {
        const bool rows = c.rows;
        const double q = c.q[k];
        const double s_k = c.s_cur[k];
        const double y_k = c.y_cur[k];
        double sa = 0.0, ya = 0.0;
        if (!avg_empty) {
            sa = c.s_avg[k];
            ya = c.y_avg[k];
            if (c.pend) {  // the input point's lagging contribution
                const double t1 = sa + s_k * c.pnw;
                sa = t1 * c.paw;
                const double t2 = ya + y_k * c.pnw;
                ya = t2 * c.paw;
            }
        }
        if (do_step) {
            const double sn = sum + s_k;
            const double yn = y_k + eta * omega * (q - 2 * sn + s_k);
            const double ds = sn - s_k, dy = yn - y_k;
            dot = ds * dy;
            yy = dy * dy;
            c.s_new[k] = sn;
            c.y_new[k] = yn;
            if (!avg_empty && !(c.defer && !EVAL)) {
                const double t1 = sa + sn * nw;
                c.s_anew[k] = t1 * aw;
                const double t2 = ya + yn * nw;
                c.y_anew[k] = t2 * aw;
            }
        }
        if (EVAL) {
            const double t = q - s_k;
            ev[0] = fabs(t);
            ev[1] = t * t;
            ev[2] = q * y_k;
            ev[3] = rows ? fabs(y_k) : 0.0;
            ev[4] = rows ? y_k * y_k : 0.0;
            ev[5] = y_k * y_k;
            if (!avg_empty) {
                const double ta = q - sa;
                ev[6] = fabs(ta);
                ev[7] = ta * ta;
                ev[8] = q * ya;
                ev[9] = rows ? fabs(ya) : 0.0;
                ev[10] = rows ? ya * ya : 0.0;
                ev[11] = ya * ya;
            }
        }
    }
    if (EVAL) {
#pragma unroll
        for (int t = 0; t < kVecStats; ++t) {
            const bool is_max = (t == 0 || t == 3 || t == 6 || t == 9);
            const double v = is_max ? wave_max(ev[t]) : wave_sum(ev[t]);
            if (lane == 0) ev_out[t] = v;
        }
    }
    if (do_step) {
        double dx = 0.0;
        for (int64_t t = c.sq_lo + lane; t < c.sq_hi; t += kWave) dx += c.sqpart[t];
        dot = wave_sum(dot);
        yy = wave_sum(yy);
        dx = wave_sum(dx);
        if (lane == 0) {
            out3[0] = dot;
            out3[1] = yy;
            out3[2] = dx;
        }
    }
}

// Sum of partials t0 <= t < t1 at position k, with A independent
// accumulators (A loads in flight per lane) combined by a fixed pairwise
// tree; the last, partial group is loaded at once and added to the first
// accumulator in order.
template <int A = 8>
__device__ __forceinline__ double sum_partials(const ChunkArgs& c, int64_t k, int t0, int t1) {
    static_assert(A >= 2 && (A & (A - 1)) == 0, "accumulators: a power of two");
    double acc[A];
#pragma unroll
    for (int q = 0; q < A; ++q) acc[q] = 0.0;
    int t = t0;
    for (; t + A <= t1; t += A) {
#pragma unroll
        for (int q = 0; q < A; ++q) acc[q] += c.part[(int64_t)(t + q) * c.pstride + k];
    }
    if (t < t1) {
        double v[A];
#pragma unroll
        for (int q = 0; q < A; ++q) v[q] = (t + q < t1) ? c.part[(int64_t)(t + q) * c.pstride + k] : 0.0;
#pragma unroll
        for (int q = 0; q < A; ++q)
            if (t + q < t1) acc[0] += v[q];
    }
#pragma unroll
    for (int st = 1; st < A; st *= 2) {
#pragma unroll
        for (int q = 0; q + st < A; q += 2 * st) acc[q] += acc[q + st];
    }
    return acc[0];
}

// Finishes one pass for a chunk (src/solver_gpu.jl:172-173, 177-234,
// 267-288; src/saddle_point_gpu.jl:140-221): r+ = rowsum(Delta) + r,
// u+ = u + eta*omega*(a - 2 r+ + r), the averages of u, v, r, c, and the
// chunk's parts of dr'du + dc'dv, ||dy||^2 and ||Delta||^2.  The partials are
// split into RWAVES contiguous ranges (one per wave), then combined in a fixed
// order by wave 0, which finishes the chunk.  Called by all threads of the
// block; thread 0 receives out3 = {dot, yy, dx} and ev_out[kVecStats].
template <bool EVAL, int RWAVES, int A = 8>
__device__ __forceinline__ void reduce_chunk(const ChunkArgs& c, double (*psum)[kWave], double* out3, double* ev_out) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const int64_t k = c.k0 + lane;
    if (c.do_step) {
        double s = 0.0;
        if (k < c.len) {
            const int t0 = (int)(((int64_t)c.np * w) / RWAVES);
            const int t1 = (int)(((int64_t)c.np * (w + 1)) / RWAVES);
            s = sum_partials<A>(c, k, t0, t1);
        }
        psum[w][lane] = s;
    }
    __syncthreads();
    if (w == 0) {
        double sum = 0.0;
        if (c.do_step) {
            sum = psum[0][lane];
#pragma unroll
            for (int t = 1; t < RWAVES; ++t) sum += psum[t][lane];
        }
        finish_chunk<EVAL>(c, lane, sum, out3, ev_out);
    }
}

// The same for one wave alone (no barrier): the wave sums all partials of its
// chunk itself.  For chunks with few partials, so that the waves of a block
// finish independent chunks in parallel.
template <bool EVAL>
__device__ __forceinline__ void reduce_chunk_wave(const ChunkArgs& c, double* out3, double* ev_out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t k = c.k0 + lane;
    const double sum = (c.do_step && k < c.len) ? sum_partials(c, k, 0, c.np) : 0.0;
    finish_chunk<EVAL>(c, lane, sum, out3, ev_out);
}

}  // namespace
