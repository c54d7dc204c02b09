// pdot_mlp.hip -- the learned warm start's network on gfx950: a point-wise
// MLP with a mean-pooled set context that maps (point coordinates, scaled
// mass, side) to a dual potential, trained and evaluated with hand-written
// kernels.  Every matrix product runs on MFMA (v_mfma_f32_32x32x16_bf16:
// bf16 operands rounded to nearest even from FP32, FP32 accumulation); the
// small vector work (bias, ReLU, pooling, the context bias, loss, Adam) is
// plain FP32.
//
// Network (row-major weights, Y = X W), for N = ns + nt points (source rows
// first):
//   Z1 = X0 W1 + b1,          H1 = relu(Z1)                 X0: N x 4 features
//   cs = mean(H1[source]),    ct = mean(H1[target])          (1 x 128 each)
//   g  = cs W2b + ct W2c + b2                                 context bias
//   Z2 = H1 W2a + g,          H2 = relu(Z2)
//   Z3 = H2 W3 + b3,          H3 = relu(Z3)
//   out = H3 w4 + b4                                           N x 1
// Loss: mean((out - target)^2).
//
// Weight head (one scalar per instance, the log of the initial primal
// weight), from the pooled context and the size s = (log2 ns + log2 nt) / 24:
//   zh = cs Wh1a + ct Wh1b + s wh1s + bh1,  hh = relu(zh)       (1 x 32)
//   log_weight = hh . wh2 + bh2
// Loss: (log_weight - target)^2.  Its gradient reaches W1 and b1 through the
// pooling only when the trunk is trained with it.
//
// Parameters (flat FP32, this order):
//   W1 4x128, b1 128, W2a 128x128, W2b 128x128, W2c 128x128, b2 128,
//   W3 128x128, b3 128, w4 128, b4 1,
//   Wh1a 128x32, Wh1b 128x32, wh1s 32, bh1 32, wh2 32, bh2 1.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pdot.h"

namespace {

constexpr int kW = 128;   // hidden width
constexpr int kF = 4;     // input features
constexpr int kFP = 16;   // input features padded to the MFMA k-step
constexpr int kT = 32;    // MFMA output tile (32 x 32)
constexpr int kKs = 16;   // MFMA k-step
constexpr int kH = 32;    // weight-head width

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> bf16 bits, round to nearest even (NaN kept quiet).
__device__ __forceinline__ unsigned short bf16_bits(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ __bf16 to_bf16(float f) {
    unsigned short b = bf16_bits(f);
    return __builtin_bit_cast(__bf16, b);
}

// C[M x N] = op(A) op(B) with op(A)[r][k] = TA ? A[k*lda + r] : A[r*lda + k]
// and op(B)[k][c] = TB ? B[c*ldb + k] : B[k*ldb + c]; one wave per 32 x 32
// output tile, k in steps of 16.  MFMA 32x32x16 operand layout: lane l holds
// row (l % 32) of A and column (l % 32) of B, k = 8*(l / 32) + 0..7; the
// 16 accumulators of lane l hold column (l % 32), rows 8*(r/4) + 4*(l/32) + r%4.
// Split K: blockIdx.z takes k in [z*KS, min(K, (z+1)*KS)) and writes its
// partial product to C + z*slice (summed by splitk_sum_kernel); the products
// whose K is the point count (the weight gradients, M, N <= 128) would
// otherwise leave 16 waves each walking thousands of k-steps.
template <bool TA, bool TB>
__global__ __launch_bounds__(64) void gemm_bf16_kernel(const float* __restrict__ A, int64_t lda,
                                                       const float* __restrict__ B, int64_t ldb, float* __restrict__ C,
                                                       int64_t ldc, int64_t M, int64_t N, int64_t K, int64_t KS,
                                                       int64_t slice) {
    const int lane = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * kT, c0 = (int64_t)blockIdx.y * kT;
    const int li = lane & 31, kh = (lane >> 5) * 8;
    const int64_t kb = (int64_t)blockIdx.z * KS;
    const int64_t ke = (kb + KS < K) ? kb + KS : K;
    C += (int64_t)blockIdx.z * slice;
    f32x16 acc;
#pragma unroll
    for (int t = 0; t < 16; ++t) acc[t] = 0.f;
    for (int64_t k0 = kb; k0 < ke; k0 += kKs) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t r = r0 + li, c = c0 + li, k = k0 + kh + j;
            float av = 0.f, bv = 0.f;
            if (r < M && k < ke) av = TA ? A[k * lda + r] : A[r * lda + k];
            if (c < N && k < ke) bv = TB ? B[c * ldb + k] : B[k * ldb + c];
            a[j] = to_bf16(av);
            b[j] = to_bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int64_t r = r0 + 8 * (t / 4) + 4 * (lane >> 5) + (t % 4), c = c0 + li;
        if (r < M && c < N) C[r * ldc + c] = acc[t];
    }
}

// C[r*ldc + c] = sum over the S slices of part (row-major M x N each), in
// slice order.
__global__ void splitk_sum_kernel(const float* __restrict__ part, int S, int64_t M, int64_t N, float* __restrict__ C,
                                  int64_t ldc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * N) return;
    float s = 0.f;
    for (int z = 0; z < S; ++z) s += part[(int64_t)z * M * N + t];
    C[(t / N) * ldc + t % N] = s;
}

// Z = Y + bias (+ row-broadcast), H = relu(Z); Y may alias Z.
__global__ void bias_relu_kernel(const float* Y, const float* bias, float* Z, float* H, int64_t N, int w) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * w) return;
    const float z = Y[t] + bias[t % w];
    Z[t] = z;
    H[t] = z > 0.f ? z : 0.f;
}

// Column sums of rows [row0, row1) of D (row-major, width w) in two fixed-
// order stages: block b of colsum_part_kernel sums kSumRows rows per column
// into part[b*w + c]; colsum_finish_kernel sums the parts in block order and
// divides by div.
constexpr int kSumRows = 64;
__global__ void colsum_part_kernel(const float* __restrict__ D, int64_t row0, int64_t row1, int w,
                                   float* __restrict__ part) {
    const int c = threadIdx.x;
    if (c >= w) return;
    const int64_t r0 = row0 + (int64_t)blockIdx.x * kSumRows;
    const int64_t r1 = (r0 + kSumRows < row1) ? r0 + kSumRows : row1;
    float s = 0.f;
    for (int64_t i = r0; i < r1; ++i) s += D[i * w + c];
    part[(int64_t)blockIdx.x * w + c] = s;
}
__global__ void colsum_finish_kernel(const float* __restrict__ part, int nparts, int w, float div,
                                     float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w) return;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += part[(int64_t)b * w + c];
    out[c] = s / div;
}

// g = cs W2b + ct W2c + b2 (w threads).
__global__ void context_kernel(const float* cs, const float* ct, const float* W2b, const float* W2c, const float* b2,
                               float* g, int w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w) return;
    float s = 0.f;
    for (int k = 0; k < w; ++k) s += cs[k] * W2b[k * w + c];
    float t = 0.f;
    for (int k = 0; k < w; ++k) t += ct[k] * W2c[k * w + c];
    g[c] = (s + t) + b2[c];
}

// out[i] = Y[i] + b4 (Y is column 0 of a width-ld product); dout = 2(out - t)/N.
__global__ void head_loss_kernel(const float* Y, int64_t ldy, const float* b4, const float* target, float* out,
                                 float* dout, float* sq, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float o = Y[i * ldy] + b4[0];
    out[i] = o;
    if (target) {
        const float d = o - target[i];
        dout[i] = 2.f * d / (float)N;
        sq[i] = d * d;
    }
}

// dZ = dH * (Z > 0), in place on dH.
__global__ void relu_back_kernel(float* dH, const float* Z, int64_t total) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total && !(Z[t] > 0.f)) dH[t] = 0.f;
}

// Context backward: with dg = colsum(dZ2): dW2b = cs^T dg, dW2c = ct^T dg,
// dcs = W2b dg, dct = W2c dg; one thread per row k of W2b/W2c.
__global__ void context_back_kernel(const float* cs, const float* ct, const float* W2b, const float* W2c,
                                    const float* dg, float* dW2b, float* dW2c, float* dcs, float* dct, int w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= w) return;
    float s = 0.f, t = 0.f;
    for (int c = 0; c < w; ++c) {
        dW2b[k * w + c] = cs[k] * dg[c];
        dW2c[k * w + c] = ct[k] * dg[c];
        s += W2b[k * w + c] * dg[c];
        t += W2c[k * w + c] * dg[c];
    }
    dcs[k] = s;
    dct[k] = t;
}

// dH1 += dcs / ns on source rows and dct / nt on target rows.
__global__ void pool_back_kernel(float* dH1, const float* dcs, const float* dct, int64_t ns, int64_t nt, int w) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (ns + nt) * w) return;
    const int64_t i = t / w;
    const int c = (int)(t % w);
    dH1[t] += i < ns ? dcs[c] / (float)ns : dct[c] / (float)nt;
}

// Adam (bias-corrected), one thread per parameter.
__global__ void adam_kernel(float* p, const float* g, float* m, float* v, int64_t np, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// Pad features N x 4 -> N x 16 (zeros), and W1 4 x 128 -> 16 x 128.
__global__ void pad_kernel(const float* src, int64_t rows, int cols, float* dst, int dcols) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * dcols) return;
    const int64_t i = t / dcols;
    const int c = (int)(t % dcols);
    dst[t] = c < cols ? src[i * cols + c] : 0.f;
}

// The instance-size input of the weight head.
__device__ __forceinline__ float size_feature(int64_t ns, int64_t nt) {
    return (log2f((float)(ns > 0 ? ns : 1)) + log2f((float)(nt > 0 ? nt : 1))) * (1.f / 24.f);
}

// Weight head forward, one wave: lanes c < 32 form zh_c (saved) and
// hh_c = relu(zh_c); out[0] = sum_c hh_c wh2_c + bh2 (a shuffle tree over
// the 32 lanes, fixed order).  Plain FP32: the head is 8 K multiply-adds.
__global__ __launch_bounds__(64) void weight_head_kernel(const float* cs, const float* ct, const float* P,
                                                         int64_t oWa, int64_t oWb, int64_t ows, int64_t obh,
                                                         int64_t ow2, int64_t ob2, int64_t ns, int64_t nt, float* zh,
                                                         float* out) {
    const int c = threadIdx.x;
    float h = 0.f;
    if (c < kH) {
        float sa = 0.f, sb = 0.f;
        for (int k = 0; k < kW; ++k) sa += cs[k] * P[oWa + k * kH + c];
        for (int k = 0; k < kW; ++k) sb += ct[k] * P[oWb + k * kH + c];
        const float z = ((sa + sb) + size_feature(ns, nt) * P[ows + c]) + P[obh + c];
        zh[c] = z;
        h = (z > 0.f ? z : 0.f) * P[ow2 + c];
    }
    for (int off = 16; off >= 1; off >>= 1) h += __shfl_down(h, off, 64);
    if (c == 0) out[0] = h + P[ob2];
}

// Weight head backward for the loss (log_weight - target)^2, one block of
// kW threads: the head's gradients into G (their offsets), and, when dcs is
// not null, the gradient reaching the pooled context (dcs, dct).
__global__ __launch_bounds__(128) void weight_head_back_kernel(const float* cs, const float* ct, const float* P,
                                                               float* G, int64_t oWa, int64_t oWb, int64_t ows,
                                                               int64_t obh, int64_t ow2, int64_t ob2, int64_t ns,
                                                               int64_t nt, const float* zh, const float* out,
                                                               float target, float* dcs, float* dct) {
    __shared__ float dz[kH];
    const int t = threadIdx.x;
    const float dl = 2.f * (out[0] - target);
    if (t < kH) {
        const float z = zh[t];
        const float hh = z > 0.f ? z : 0.f;
        G[ow2 + t] = dl * hh;
        const float d = z > 0.f ? dl * P[ow2 + t] : 0.f;
        dz[t] = d;
        G[obh + t] = d;
        G[ows + t] = size_feature(ns, nt) * d;
    }
    if (t == 0) G[ob2] = dl;
    __syncthreads();
    // row k of Wh1a / Wh1b
    const int k = t;
    float sa = 0.f, sb = 0.f;
    for (int c = 0; c < kH; ++c) {
        G[oWa + k * kH + c] = cs[k] * dz[c];
        G[oWb + k * kH + c] = ct[k] * dz[c];
        sa += P[oWa + k * kH + c] * dz[c];
        sb += P[oWb + k * kH + c] * dz[c];
    }
    if (dcs) {
        dcs[k] = sa;
        dct[k] = sb;
    }
}

unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

struct pdot_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    // parameter offsets in the flat vector
    int64_t oW1, ob1, oW2a, oW2b, oW2c, ob2, oW3, ob3, ow4, ob4, np;
    int64_t oWha, oWhb, owhs, obh1, owh2, obh2;  // weight head
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // params, grads, Adam moments
    int64_t adam_t = 0;
    // activations for the current N
    int64_t cap = 0;
    float *X0 = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Z3 = nullptr, *H3 = nullptr;
    float *Y4 = nullptr, *out = nullptr, *tgt = nullptr, *dout = nullptr, *sq = nullptr, *D = nullptr, *E = nullptr;
    float *feat = nullptr;
    float *W1p = nullptr, *dW1p = nullptr, *w4p = nullptr;  // padded W1 (16x128), its gradient, w4 as 128 x 32
    float *cs = nullptr, *ct = nullptr, *gctx = nullptr, *dcs = nullptr, *dct = nullptr, *dg = nullptr;
    float *zh = nullptr, *logw = nullptr;  // weight head pre-activations and output
    float *sumpart = nullptr;              // column-sum parts (cap / kSumRows + 2 blocks of kW)
    float *kpart = nullptr;                // split-K partial products (kMaxSlices x kW x kW)
    std::vector<float> hbuf;
    std::string err;
};

namespace {

int mfail(pdot_mlp* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    return code;
}
#define MHIPCHK(mm, call)                                                                               \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return mfail((mm), PDOT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int malloc_f(pdot_mlp* m, float** p, int64_t n) {
    hipError_t e = hipMalloc((void**)p, sizeof(float) * (size_t)(n > 0 ? n : 1));
    if (e != hipSuccess) return mfail(m, PDOT_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return PDOT_OK;
}

constexpr int kMaxSlices = 64;      // split-K slices (partials of at most kW x kW)
constexpr int64_t kSliceK = 256;    // k per slice, at least

int ensure_cap(pdot_mlp* m, int64_t N) {
    if (N <= m->cap) return PDOT_OK;
    if (m->sumpart) (void)hipFree(m->sumpart);
    m->sumpart = nullptr;
    float** bufs[] = {&m->X0, &m->Z1, &m->H1, &m->Z2, &m->H2, &m->Z3, &m->H3, &m->D, &m->E};
    for (float** b : bufs)
        if (*b) (void)hipFree(*b);
    float** vecs[] = {&m->Y4, &m->out, &m->tgt, &m->dout, &m->sq};
    for (float** b : vecs)
        if (*b) (void)hipFree(*b);
    if (m->feat) (void)hipFree(m->feat);
    int rc;
    if ((rc = malloc_f(m, &m->X0, N * kFP))) return rc;
    float** wide[] = {&m->Z1, &m->H1, &m->Z2, &m->H2, &m->Z3, &m->H3, &m->D, &m->E};
    for (float** b : wide)
        if ((rc = malloc_f(m, b, N * kW))) return rc;
    if ((rc = malloc_f(m, &m->Y4, N * kT))) return rc;
    float** narrow[] = {&m->out, &m->tgt, &m->dout, &m->sq};
    for (float** b : narrow)
        if ((rc = malloc_f(m, b, N))) return rc;
    if ((rc = malloc_f(m, &m->feat, N * kF))) return rc;
    if ((rc = malloc_f(m, &m->sumpart, (N / kSumRows + 2) * kW))) return rc;
    m->cap = N;
    return PDOT_OK;
}

template <bool TA, bool TB>
hipError_t gemm(pdot_mlp* m, const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc,
                int64_t M, int64_t N, int64_t K) {
    const unsigned gx = (unsigned)((M + kT - 1) / kT), gy = (unsigned)((N + kT - 1) / kT);
    // split K when the output has few tiles and K is long
    int64_t S = 1;
    if ((int64_t)gx * gy < 64 && M <= kW && N <= kW && K >= 2 * kSliceK)
        S = std::min<int64_t>(kMaxSlices, (K + kSliceK - 1) / kSliceK);
    if (S == 1) {
        hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB>), dim3(gx, gy, 1), dim3(64), 0, m->stream, A, lda, B, ldb, C, ldc,
                           M, N, K, K, (int64_t)0);
        return hipGetLastError();
    }
    const int64_t KS = ((K + S - 1) / S + kKs - 1) / kKs * kKs;  // a multiple of the k-step
    S = (K + KS - 1) / KS;
    hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB>), dim3(gx, gy, (unsigned)S), dim3(64), 0, m->stream, A, lda, B, ldb,
                       m->kpart, N, M, N, K, KS, M * N);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(splitk_sum_kernel, dim3(blocks_for(M * N)), dim3(256), 0, m->stream, m->kpart, (int)S, M, N, C,
                       ldc);
    return hipGetLastError();
}

// Column sums (divided by div) of rows [row0, row1) of D (width w) into out.
hipError_t colsum_rows(pdot_mlp* m, const float* D, int64_t row0, int64_t row1, int w, float div, float* out) {
    const int64_t nb = std::max<int64_t>(1, (row1 - row0 + kSumRows - 1) / kSumRows);
    hipLaunchKernelGGL(colsum_part_kernel, dim3((unsigned)nb), dim3(kW), 0, m->stream, D, row0, row1, w, m->sumpart);
    hipLaunchKernelGGL(colsum_finish_kernel, dim3(1), dim3(kW), 0, m->stream, m->sumpart, (int)nb, w, div, out);
    return hipGetLastError();
}

// Layer 1 and the pooled context on the features already in m->feat.
int forward_trunk(pdot_mlp* m, int64_t ns, int64_t nt) {
    const int64_t N = ns + nt;
    float* P = m->p;
    hipLaunchKernelGGL(pad_kernel, dim3(blocks_for(N * kFP)), dim3(256), 0, m->stream, m->feat, N, kF, m->X0, kFP);
    // W1 (4 x 128) as a 16 x 128 matrix with zero rows 4..15
    MHIPCHK(m, hipMemsetAsync(m->W1p, 0, sizeof(float) * kFP * kW, m->stream));
    MHIPCHK(m, hipMemcpyAsync(m->W1p, P + m->oW1, sizeof(float) * kF * kW, hipMemcpyDeviceToDevice, m->stream));
    MHIPCHK(m, (gemm<false, false>(m, m->X0, kFP, m->W1p, kW, m->Z1, kW, N, kW, kFP)));
    hipLaunchKernelGGL(bias_relu_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->Z1, P + m->ob1, m->Z1,
                       m->H1, N, kW);
    MHIPCHK(m, colsum_rows(m, m->H1, 0, ns, kW, ns > 0 ? (float)ns : 1.f, m->cs));
    MHIPCHK(m, colsum_rows(m, m->H1, ns, ns + nt, kW, nt > 0 ? (float)nt : 1.f, m->ct));
    MHIPCHK(m, hipGetLastError());
    return PDOT_OK;
}

// The weight head after forward_trunk: m->zh, m->logw.
int forward_weight(pdot_mlp* m, int64_t ns, int64_t nt) {
    hipLaunchKernelGGL(weight_head_kernel, dim3(1), dim3(64), 0, m->stream, m->cs, m->ct, m->p, m->oWha, m->oWhb,
                       m->owhs, m->obh1, m->owh2, m->obh2, ns, nt, m->zh, m->logw);
    MHIPCHK(m, hipGetLastError());
    return PDOT_OK;
}

// Forward pass on the features already in m->feat; fills all activations.
int forward(pdot_mlp* m, int64_t ns, int64_t nt, const float* target_dev) {
    const int64_t N = ns + nt;
    float* P = m->p;
    int rc = forward_trunk(m, ns, nt);
    if (rc) return rc;
    hipLaunchKernelGGL(context_kernel, dim3(1), dim3(kW), 0, m->stream, m->cs, m->ct, P + m->oW2b, P + m->oW2c,
                       P + m->ob2, m->gctx, kW);
    MHIPCHK(m, (gemm<false, false>(m, m->H1, kW, P + m->oW2a, kW, m->Z2, kW, N, kW, kW)));
    hipLaunchKernelGGL(bias_relu_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->Z2, m->gctx, m->Z2,
                       m->H2, N, kW);
    MHIPCHK(m, (gemm<false, false>(m, m->H2, kW, P + m->oW3, kW, m->Z3, kW, N, kW, kW)));
    hipLaunchKernelGGL(bias_relu_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->Z3, P + m->ob3, m->Z3,
                       m->H3, N, kW);
    // w4 (128) as a 128 x 32 matrix with zero columns 1..31
    MHIPCHK(m, hipMemsetAsync(m->w4p, 0, sizeof(float) * kW * kT, m->stream));
    MHIPCHK(m, hipMemcpy2DAsync(m->w4p, sizeof(float) * kT, P + m->ow4, sizeof(float), sizeof(float), kW,
                                hipMemcpyDeviceToDevice, m->stream));
    MHIPCHK(m, (gemm<false, false>(m, m->H3, kW, m->w4p, kT, m->Y4, kT, N, kT, kW)));
    hipLaunchKernelGGL(head_loss_kernel, dim3(blocks_for(N)), dim3(256), 0, m->stream, m->Y4, (int64_t)kT,
                       P + m->ob4, target_dev, m->out, m->dout, m->sq, N);
    MHIPCHK(m, hipGetLastError());
    return PDOT_OK;
}

// Backward pass after forward(...) with a target: fills m->g.
int backward(pdot_mlp* m, int64_t ns, int64_t nt) {
    const int64_t N = ns + nt;
    float* P = m->p;
    float* G = m->g;
    // the dual loss does not reach the weight head
    MHIPCHK(m, hipMemsetAsync(G + m->oWha, 0, sizeof(float) * (m->np - m->oWha), m->stream));
    // head: dW4 = H3^T dout (128 x 1), db4 = sum dout, dH3 = dout w4^T
    MHIPCHK(m, (gemm<true, false>(m, m->H3, kW, m->dout, 1, G + m->ow4, 1, kW, 1, N)));
    MHIPCHK(m, colsum_rows(m, m->dout, 0, N, 1, 1.f, G + m->ob4));
    MHIPCHK(m, (gemm<false, true>(m, m->dout, 1, P + m->ow4, 1, m->D, kW, N, kW, 1)));
    hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->D, m->Z3, N * kW);
    // layer 3: dW3 = H2^T dZ3, db3 = colsum dZ3, dH2 = dZ3 W3^T
    MHIPCHK(m, (gemm<true, false>(m, m->H2, kW, m->D, kW, G + m->oW3, kW, kW, kW, N)));
    MHIPCHK(m, colsum_rows(m, m->D, 0, N, kW, 1.f, G + m->ob3));
    MHIPCHK(m, (gemm<false, true>(m, m->D, kW, P + m->oW3, kW, m->E, kW, N, kW, kW)));
    hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->E, m->Z2, N * kW);
    // layer 2: dW2a = H1^T dZ2, dg = colsum dZ2 (= db2), context weights, dH1 = dZ2 W2a^T + pooling
    MHIPCHK(m, (gemm<true, false>(m, m->H1, kW, m->E, kW, G + m->oW2a, kW, kW, kW, N)));
    MHIPCHK(m, colsum_rows(m, m->E, 0, N, kW, 1.f, G + m->ob2));
    hipLaunchKernelGGL(context_back_kernel, dim3(1), dim3(kW), 0, m->stream, m->cs, m->ct, P + m->oW2b, P + m->oW2c,
                       G + m->ob2, G + m->oW2b, G + m->oW2c, m->dcs, m->dct, kW);
    MHIPCHK(m, (gemm<false, true>(m, m->E, kW, P + m->oW2a, kW, m->D, kW, N, kW, kW)));
    hipLaunchKernelGGL(pool_back_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->D, m->dcs, m->dct, ns,
                       nt, kW);
    hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->D, m->Z1, N * kW);
    // layer 1: dW1 = X0^T dZ1 (16 x 128, rows 0..3 kept), db1 = colsum dZ1
    MHIPCHK(m, (gemm<true, false>(m, m->X0, kFP, m->D, kW, m->dW1p, kW, kFP, kW, N)));
    MHIPCHK(m, hipMemcpyAsync(G + m->oW1, m->dW1p, sizeof(float) * kF * kW, hipMemcpyDeviceToDevice, m->stream));
    MHIPCHK(m, colsum_rows(m, m->D, 0, N, kW, 1.f, G + m->ob1));
    MHIPCHK(m, hipGetLastError());
    return PDOT_OK;
}

// Gradient of (log_weight - target)^2 after forward_trunk + forward_weight:
// the head's entries of m->g and, if train_trunk, W1 and b1 through the
// pooling; every other entry is zero.
int backward_weight(pdot_mlp* m, int64_t ns, int64_t nt, float target, bool train_trunk) {
    const int64_t N = ns + nt;
    MHIPCHK(m, hipMemsetAsync(m->g, 0, sizeof(float) * m->np, m->stream));
    hipLaunchKernelGGL(weight_head_back_kernel, dim3(1), dim3(kW), 0, m->stream, m->cs, m->ct, m->p, m->g, m->oWha,
                       m->oWhb, m->owhs, m->obh1, m->owh2, m->obh2, ns, nt, m->zh, m->logw, target,
                       train_trunk ? m->dcs : nullptr, train_trunk ? m->dct : nullptr);
    if (train_trunk) {
        MHIPCHK(m, hipMemsetAsync(m->D, 0, sizeof(float) * N * kW, m->stream));
        hipLaunchKernelGGL(pool_back_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->D, m->dcs, m->dct,
                           ns, nt, kW);
        hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for(N * kW)), dim3(256), 0, m->stream, m->D, m->Z1, N * kW);
        MHIPCHK(m, (gemm<true, false>(m, m->X0, kFP, m->D, kW, m->dW1p, kW, kFP, kW, N)));
        MHIPCHK(m, hipMemcpyAsync(m->g + m->oW1, m->dW1p, sizeof(float) * kF * kW, hipMemcpyDeviceToDevice,
                                  m->stream));
        MHIPCHK(m, colsum_rows(m, m->D, 0, N, kW, 1.f, m->g + m->ob1));
    }
    MHIPCHK(m, hipGetLastError());
    return PDOT_OK;
}

int upload_inputs(pdot_mlp* m, const float* feats, int64_t N, const float* target) {
    int rc = ensure_cap(m, N);
    if (rc) return rc;
    MHIPCHK(m, hipMemcpyAsync(m->feat, feats, sizeof(float) * N * kF, hipMemcpyHostToDevice, m->stream));
    if (target) MHIPCHK(m, hipMemcpyAsync(m->tgt, target, sizeof(float) * N, hipMemcpyHostToDevice, m->stream));
    return PDOT_OK;
}

}  // namespace

extern "C" {

int pdot_mlp_create(pdot_mlp** out, int32_t device) {
    if (!out) return PDOT_E_ARG;
    pdot_mlp* m = new pdot_mlp();
    *out = m;
    m->device = device;
    MHIPCHK(m, hipSetDevice(device));
    MHIPCHK(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    int64_t o = 0;
    m->oW1 = o; o += kF * kW;
    m->ob1 = o; o += kW;
    m->oW2a = o; o += kW * kW;
    m->oW2b = o; o += kW * kW;
    m->oW2c = o; o += kW * kW;
    m->ob2 = o; o += kW;
    m->oW3 = o; o += kW * kW;
    m->ob3 = o; o += kW;
    m->ow4 = o; o += kW;
    m->ob4 = o; o += 1;
    m->oWha = o; o += kW * kH;
    m->oWhb = o; o += kW * kH;
    m->owhs = o; o += kH;
    m->obh1 = o; o += kH;
    m->owh2 = o; o += kH;
    m->obh2 = o; o += 1;
    m->np = o;
    int rc;
    float** pv[] = {&m->p, &m->g, &m->m, &m->v};
    for (float** q : pv) {
        if ((rc = malloc_f(m, q, m->np))) return rc;
        MHIPCHK(m, hipMemsetAsync(*q, 0, sizeof(float) * m->np, m->stream));
    }
    if ((rc = malloc_f(m, &m->W1p, kFP * kW))) return rc;
    if ((rc = malloc_f(m, &m->dW1p, kFP * kW))) return rc;
    if ((rc = malloc_f(m, &m->w4p, kW * kT))) return rc;
    if ((rc = malloc_f(m, &m->kpart, (int64_t)kMaxSlices * kW * kW))) return rc;
    float** sm[] = {&m->cs, &m->ct, &m->gctx, &m->dcs, &m->dct, &m->dg, &m->zh, &m->logw};
    for (float** q : sm)
        if ((rc = malloc_f(m, q, kW))) return rc;
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return PDOT_OK;
}

void pdot_mlp_destroy(pdot_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    float* all[] = {m->p, m->g, m->m, m->v, m->X0, m->Z1, m->H1, m->Z2, m->H2, m->Z3, m->H3, m->Y4, m->out, m->tgt,
                    m->dout, m->sq, m->D, m->E, m->feat, m->W1p, m->dW1p, m->w4p, m->cs, m->ct, m->gctx, m->dcs,
                    m->dct, m->dg, m->zh, m->logw, m->sumpart, m->kpart};
    for (float* q : all)
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* pdot_mlp_last_error(const pdot_mlp* m) { return m ? m->err.c_str() : "null network"; }

int64_t pdot_mlp_num_params(const pdot_mlp* m) { return m ? m->np : 0; }

int pdot_mlp_set_params(pdot_mlp* m, const float* params) {
    if (!m || !params) return PDOT_E_ARG;
    MHIPCHK(m, hipMemcpyAsync(m->p, params, sizeof(float) * m->np, hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipMemsetAsync(m->m, 0, sizeof(float) * m->np, m->stream));
    MHIPCHK(m, hipMemsetAsync(m->v, 0, sizeof(float) * m->np, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    return PDOT_OK;
}

int pdot_mlp_get_params(pdot_mlp* m, float* params) {
    if (!m || !params) return PDOT_E_ARG;
    MHIPCHK(m, hipMemcpyAsync(params, m->p, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return PDOT_OK;
}

int pdot_mlp_forward(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, float* out) {
    if (!m || !features || !out || nsrc < 0 || ntgt < 0 || nsrc + ntgt == 0) return PDOT_E_ARG;
    const int64_t N = nsrc + ntgt;
    int rc = upload_inputs(m, features, N, nullptr);
    if (rc) return rc;
    if ((rc = forward(m, nsrc, ntgt, nullptr))) return rc;
    MHIPCHK(m, hipMemcpyAsync(out, m->out, sizeof(float) * N, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return PDOT_OK;
}

int pdot_mlp_loss_grad(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, const float* target,
                       float* loss, float* grad) {
    if (!m || !features || !target || nsrc < 0 || ntgt < 0 || nsrc + ntgt == 0) return PDOT_E_ARG;
    const int64_t N = nsrc + ntgt;
    int rc = upload_inputs(m, features, N, target);
    if (rc) return rc;
    if ((rc = forward(m, nsrc, ntgt, m->tgt))) return rc;
    if ((rc = backward(m, nsrc, ntgt))) return rc;
    if (loss) {
        m->hbuf.resize(N);
        MHIPCHK(m, hipMemcpyAsync(m->hbuf.data(), m->sq, sizeof(float) * N, hipMemcpyDeviceToHost, m->stream));
    }
    if (grad) MHIPCHK(m, hipMemcpyAsync(grad, m->g, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    if (loss) {
        double s = 0.0;
        for (int64_t i = 0; i < N; ++i) s += m->hbuf[i];
        *loss = (float)(s / (double)N);
    }
    return PDOT_OK;
}

int pdot_mlp_predict_weight(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, float* log_weight) {
    if (!m || !features || !log_weight || nsrc < 1 || ntgt < 1) return PDOT_E_ARG;
    int rc = upload_inputs(m, features, nsrc + ntgt, nullptr);
    if (rc) return rc;
    if ((rc = forward_trunk(m, nsrc, ntgt))) return rc;
    if ((rc = forward_weight(m, nsrc, ntgt))) return rc;
    MHIPCHK(m, hipMemcpyAsync(log_weight, m->logw, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return PDOT_OK;
}

int pdot_mlp_weight_loss_grad(pdot_mlp* m, const float* features, int64_t nsrc, int64_t ntgt, float target,
                              int32_t train_trunk, float* log_weight, float* loss, float* grad) {
    if (!m || !features || nsrc < 1 || ntgt < 1 || !(target == target)) return PDOT_E_ARG;
    int rc = upload_inputs(m, features, nsrc + ntgt, nullptr);
    if (rc) return rc;
    if ((rc = forward_trunk(m, nsrc, ntgt))) return rc;
    if ((rc = forward_weight(m, nsrc, ntgt))) return rc;
    if ((rc = backward_weight(m, nsrc, ntgt, target, train_trunk != 0))) return rc;
    float lw = 0.f;
    MHIPCHK(m, hipMemcpyAsync(&lw, m->logw, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (grad) MHIPCHK(m, hipMemcpyAsync(grad, m->g, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    if (log_weight) *log_weight = lw;
    if (loss) *loss = (lw - target) * (lw - target);
    return PDOT_OK;
}

int pdot_mlp_adam_step(pdot_mlp* m, float lr, float beta1, float beta2, float eps) {
    if (!m) return PDOT_E_ARG;
    m->adam_t += 1;
    const float c1 = 1.f - powf(beta1, (float)m->adam_t), c2 = 1.f - powf(beta2, (float)m->adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(m->np)), dim3(256), 0, m->stream, m->p, m->g, m->m, m->v, m->np,
                       lr, beta1, beta2, eps, c1, c2);
    MHIPCHK(m, hipGetLastError());
    return PDOT_OK;
}

int pdot_mlp_gemm_probe(pdot_mlp* m, const float* A, const float* B, float* C, int64_t M, int64_t N, int64_t K,
                        int32_t trans_a, int32_t trans_b) {
    if (!m || !A || !B || !C || M <= 0 || N <= 0 || K <= 0) return PDOT_E_ARG;
    float *dA, *dB, *dC;
    int rc;
    if ((rc = malloc_f(m, &dA, M * K))) return rc;
    if ((rc = malloc_f(m, &dB, K * N))) return rc;
    if ((rc = malloc_f(m, &dC, M * N))) return rc;
    MHIPCHK(m, hipMemcpyAsync(dA, A, sizeof(float) * M * K, hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipMemcpyAsync(dB, B, sizeof(float) * K * N, hipMemcpyHostToDevice, m->stream));
    const int64_t lda = trans_a ? M : K, ldb = trans_b ? K : N;
    hipError_t e;
    if (!trans_a && !trans_b) e = gemm<false, false>(m, dA, lda, dB, ldb, dC, N, M, N, K);
    else if (trans_a && !trans_b) e = gemm<true, false>(m, dA, lda, dB, ldb, dC, N, M, N, K);
    else if (!trans_a && trans_b) e = gemm<false, true>(m, dA, lda, dB, ldb, dC, N, M, N, K);
    else e = gemm<true, true>(m, dA, lda, dB, ldb, dC, N, M, N, K);
    MHIPCHK(m, e);
    MHIPCHK(m, hipMemcpyAsync(C, dC, sizeof(float) * M * N, hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    return PDOT_OK;
}

}  // extern "C"
