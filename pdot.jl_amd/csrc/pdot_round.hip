// pdot_round.hip -- gfx950 kernels for the feasibility rounding of a
// transport plan (Altschuler-Weed-Rigollet), src/utils_gpu.jl:296-318:
//
//   s_i = min(a_i / rowsum_i(X), 1)          (0/0 -> 0)
//   t_j = min(b_j / colsum_j(diag(s) X), 1)  (0/0 -> 0)
//   X1  = (s_i X_ij) t_j
//   e_r = a - rowsum(X1), divided by ||e_r||_1 unless that is 0
//   e_c = b - colsum(X1)
//   Y   = X1 + e_r e_c^T
//
// Every element expression rounds as the reference's broadcasts do
// (-ffp-contract=off): Y_ij = fl(fl(fl(s_i X_ij) t_j) + fl(er_i ec_j)).  Only
// the summation order of the row and column sums differs from the
// reference's, so results agree to rounding and are bitwise reproducible.
//
// HBM traffic: four streaming reads of X (row sums; column sums of sX; both
// sums of X1; the final pass) and one write of Y.  X1 is recomputed from X in
// registers instead of being stored.  A pass is split into wave work units
// (64*kRpt rows x a column range): each lane keeps the running row sums of
// its kRpt rows in registers and a wave reduction forms each column's
// partial sum, so no LDS or barrier is needed; per-unit partials are finished
// in fixed order by small vector kernels.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pdot.h"
#include "pdot_common.h"
#include "pdot_round.h"

namespace {

constexpr int kRpt = 4;                // rows per lane
constexpr int kUnitRows = kWave * kRpt;
constexpr int kRoundBlock = 256;       // 4 independent waves per block
constexpr int kUnitsPerBlock = kRoundBlock / kWave;

struct RoundArgs {
    const void* X;
    const double* s;   // row scale or null
    const double* t;   // column scale or null
    const double* er;  // final pass only
    const double* ec;
    double* rowpart;   // [CB][n]
    double* colpart;   // [NG][m]
    double* Y;
    int64_t n, m, ld, ldy;
    int32_t NG, CB, TC;
};

__device__ __forceinline__ double load_as_double(const void* X, int64_t k, bool f32) {
    return f32 ? (double)static_cast<const float*>(X)[k] : static_cast<const double*>(X)[k];
}

// MODE 0: row sums of X; 1: column sums of sX; 2: row and column sums of
// (sX)t; 3: Y = (sX)t + er ec^T.
template <bool F32, int MODE>
__global__ __launch_bounds__(kRoundBlock) void round_pass_kernel(RoundArgs p) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t unit = (int64_t)blockIdx.x * kUnitsPerBlock + (threadIdx.x >> 6);
    const int64_t g = unit % p.NG;
    const int64_t cb = unit / p.NG;
    if (cb >= p.CB) return;  // whole wave leaves together
    const int64_t j0 = cb * p.TC;
    const int64_t j1 = (j0 + p.TC < p.m) ? j0 + p.TC : p.m;
    int64_t rows[kRpt];
    bool ok[kRpt];
    double s[kRpt], er[kRpt], racc[kRpt];
#pragma unroll
    for (int k = 0; k < kRpt; ++k) {
        rows[k] = g * kUnitRows + k * kWave + lane;
        ok[k] = rows[k] < p.n;
        s[k] = (MODE >= 1 && ok[k]) ? p.s[rows[k]] : 1.0;
        er[k] = (MODE == 3 && ok[k]) ? p.er[rows[k]] : 0.0;
        racc[k] = 0.0;
    }
    for (int64_t j = j0; j < j1; ++j) {
        const double tj = (MODE >= 2) ? p.t[j] : 1.0;
        const double ecj = (MODE == 3) ? p.ec[j] : 0.0;
        double cacc = 0.0;
#pragma unroll
        for (int k = 0; k < kRpt; ++k) {
            if (!ok[k]) continue;
            double y = load_as_double(p.X, rows[k] + j * p.ld, F32);
            if (MODE >= 1) y = s[k] * y;
            if (MODE >= 2) y = y * tj;
            if (MODE == 3) {
                const double e = er[k] * ecj;
                p.Y[rows[k] + j * p.ldy] = y + e;
            } else {
                racc[k] += y;
                cacc += y;
            }
        }
        if (MODE == 1 || MODE == 2) {
            cacc = wave_sum(cacc);
            if (lane == 0) p.colpart[g * p.m + j] = cacc;
        }
    }
    if (MODE == 0 || MODE == 2) {
#pragma unroll
        for (int k = 0; k < kRpt; ++k)
            if (ok[k]) p.rowpart[cb * p.n + rows[k]] = racc[k];
    }
}

// Fixed-order finish of per-unit partials: out_i = f(q_i, sum_k part[k][i]).
// RATIO: min(q / sum, 1) with 0/0 -> 0 (a/0 = Inf -> 1); else q - sum.
template <bool RATIO>
__global__ __launch_bounds__(256) void finish_kernel(const double* __restrict__ part, int32_t nparts, int64_t len,
                                                     const double* __restrict__ q, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double acc = 0.0;
    for (int k = 0; k < nparts; ++k) acc += part[(int64_t)k * len + i];
    if (RATIO) {
        double r = q[i] / acc;
        if (r != r) r = 0.0;
        out[i] = r < 1.0 ? r : 1.0;
    } else {
        out[i] = q[i] - acc;
    }
}

// er /= ||er||_1 unless the norm is 0 (one block, fixed order).
__global__ __launch_bounds__(1024) void l1_normalize_kernel(double* er, int64_t n) {
    __shared__ double red[1024];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += fabs(er[i]);
    const double l1 = block_sum(acc, red);
    if (l1 != 0.0)
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) er[i] = er[i] / l1;
}

template <bool F32>
hipError_t launch_pass(int mode, hipStream_t stream, const RoundArgs& p) {
    const int64_t units = (int64_t)p.NG * p.CB;
    const dim3 grid((unsigned)((units + kUnitsPerBlock - 1) / kUnitsPerBlock)), block(kRoundBlock);
    switch (mode) {
        case 0: hipLaunchKernelGGL((round_pass_kernel<F32, 0>), grid, block, 0, stream, p); break;
        case 1: hipLaunchKernelGGL((round_pass_kernel<F32, 1>), grid, block, 0, stream, p); break;
        case 2: hipLaunchKernelGGL((round_pass_kernel<F32, 2>), grid, block, 0, stream, p); break;
        default: hipLaunchKernelGGL((round_pass_kernel<F32, 3>), grid, block, 0, stream, p); break;
    }
    return hipGetLastError();
}

}  // namespace

namespace {

// Unit geometry of an n x m rounding on the current device: NG row groups of
// kUnitRows rows, CB column blocks of TC columns; about 16 waves per CU of
// work units, each streaming at least 16 columns.
struct RoundGeometry {
    int64_t NG, CB, TC, words;
    bool fits;
};

RoundGeometry round_geometry(int64_t n, int64_t m) {
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    RoundGeometry g;
    g.NG = (n + kUnitRows - 1) / kUnitRows;
    int64_t CB = ((int64_t)cus * 16 + g.NG - 1) / g.NG;
    const int64_t max_cb = (m + 15) / 16;
    if (CB > max_cb) CB = max_cb;
    if (CB < 1) CB = 1;
    g.TC = (m + CB - 1) / CB;
    g.CB = (m + g.TC - 1) / g.TC;
    g.fits = g.NG * g.CB / kUnitsPerBlock + 1 <= 0x7fffffffLL;
    g.words = g.CB * n + g.NG * m + 2 * n + 2 * m;
    return g;
}

}  // namespace

int64_t pdot_round_work_words(int64_t n, int64_t m) {
    if (n < 1 || m < 1) return -1;
    const RoundGeometry g = round_geometry(n, m);
    return g.fits ? g.words : -1;
}

int pdot_round_plan(hipStream_t stream, const void* X, int32_t storage, int64_t ld, int64_t n, int64_t m,
                    const double* a, const double* b, double* Y, int64_t ldy, std::string* err, double* work,
                    int64_t work_words) {
    auto failed = [&](int code, const std::string& msg) {
        if (err) *err = msg;
        return code;
    };
    if (n < 1 || m < 1 || ld < n || ldy < n || !X || !a || !b || !Y)
        return failed(PDOT_E_ARG, "pdot_round_plan: invalid arguments");
    if (storage != PDOT_STORAGE_F64 && storage != PDOT_STORAGE_F32)
        return failed(PDOT_E_ARG, "pdot_round_plan: unknown storage");
    const bool f32 = storage == PDOT_STORAGE_F32;
    if (f32 && static_cast<const void*>(Y) == X) return failed(PDOT_E_ARG, "pdot_round_plan: FP32 plan cannot be rounded in place");
    if (static_cast<const void*>(Y) == X && ldy != ld)
        return failed(PDOT_E_ARG, "pdot_round_plan: in-place rounding needs ldy == ld");
    if (static_cast<const void*>(Y) != X) {
        // a distinct output must not overlap the plan
        const char* x0 = static_cast<const char*>(X);
        const char* y0 = reinterpret_cast<const char*>(Y);
        const int64_t xb = (ld * (m - 1) + n) * (f32 ? 4 : 8), yb = (ldy * (m - 1) + n) * 8;
        if (y0 < x0 + xb && x0 < y0 + yb) return failed(PDOT_E_ARG, "pdot_round_plan: output overlaps the plan");
    }

    const RoundGeometry geo = round_geometry(n, m);
    if (!geo.fits) return failed(PDOT_E_ARG, "pdot_round_plan: grid too large");
    const int64_t NG = geo.NG, CB = geo.CB, TC = geo.TC;

    double* buf = work;
    const size_t words = (size_t)geo.words;
    hipError_t e = hipSuccess;
    if (work) {
        if (work_words < geo.words)
            return failed(PDOT_E_ARG, "pdot_round_plan: workspace of " + std::to_string(work_words) +
                                          " doubles is smaller than the " + std::to_string(geo.words) + " needed");
    } else {
        e = hipMalloc((void**)&buf, sizeof(double) * words);
        if (e != hipSuccess)
            return failed(PDOT_E_NOMEM, std::string("pdot_round_plan: hipMalloc: ") + hipGetErrorString(e));
    }
    RoundArgs p;
    p.X = X;
    p.rowpart = buf;
    p.colpart = p.rowpart + CB * n;
    double* s = p.colpart + NG * m;
    double* er = s + n;
    double* t = er + n;
    double* ec = t + m;
    p.s = s;
    p.t = t;
    p.er = er;
    p.ec = ec;
    p.Y = Y;
    p.n = n;
    p.m = m;
    p.ld = ld;
    p.ldy = ldy;
    p.NG = (int32_t)NG;
    p.CB = (int32_t)CB;
    p.TC = (int32_t)TC;
    auto pass = [&](int mode) { return f32 ? launch_pass<true>(mode, stream, p) : launch_pass<false>(mode, stream, p); };
    const dim3 blk(256), gn((unsigned)((n + 255) / 256)), gm((unsigned)((m + 255) / 256));
    int rc = PDOT_OK;
    std::string msg;
    do {
        if ((e = pass(0)) != hipSuccess) break;
        hipLaunchKernelGGL(finish_kernel<true>, gn, blk, 0, stream, p.rowpart, p.CB, n, a, s);
        if ((e = pass(1)) != hipSuccess) break;
        hipLaunchKernelGGL(finish_kernel<true>, gm, blk, 0, stream, p.colpart, p.NG, m, b, t);
        if ((e = pass(2)) != hipSuccess) break;
        hipLaunchKernelGGL(finish_kernel<false>, gn, blk, 0, stream, p.rowpart, p.CB, n, a, er);
        hipLaunchKernelGGL(finish_kernel<false>, gm, blk, 0, stream, p.colpart, p.NG, m, b, ec);
        hipLaunchKernelGGL(l1_normalize_kernel, dim3(1), dim3(1024), 0, stream, er, n);
        if ((e = pass(3)) != hipSuccess) break;
        e = hipStreamSynchronize(stream);
    } while (false);
    if (e != hipSuccess) {
        rc = PDOT_E_HIP;
        msg = std::string("pdot_round_plan: ") + hipGetErrorString(e);
        hipStreamSynchronize(stream);
    }
    if (!work) hipFree(buf);
    if (rc != PDOT_OK) return failed(rc, msg);
    return PDOT_OK;
}

// ------------------------------------------------------------------ C ABI

extern "C" int pdot_round_matrix(int32_t device, int64_t n, int64_t m, const double* a, const double* b, double* X,
                                 char* errbuf, int64_t errlen) {
    std::string err;
    auto done = [&](int code) {
        if (errbuf && errlen > 0) {
            const size_t k = err.size() < (size_t)(errlen - 1) ? err.size() : (size_t)(errlen - 1);
            err.copy(errbuf, k);
            errbuf[k] = '\0';
        }
        return code;
    };
    if (n < 1 || m < 1 || !a || !b || !X) {
        err = "pdot_round_matrix: invalid arguments";
        return done(PDOT_E_ARG);
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        err = std::string("pdot_round_matrix: hipSetDevice: ") + hipGetErrorString(e);
        return done(PDOT_E_HIP);
    }
    hipStream_t stream = nullptr;
    double *dX = nullptr, *da = nullptr, *db = nullptr;
    int rc = PDOT_OK;
    const size_t nm = (size_t)n * (size_t)m;
    do {
        if ((e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess) break;
        if ((e = hipMalloc((void**)&dX, sizeof(double) * nm)) != hipSuccess) {
            rc = PDOT_E_NOMEM;
            break;
        }
        if ((e = hipMalloc((void**)&da, sizeof(double) * (size_t)n)) != hipSuccess) break;
        if ((e = hipMalloc((void**)&db, sizeof(double) * (size_t)m)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(dX, X, sizeof(double) * nm, hipMemcpyHostToDevice, stream)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(da, a, sizeof(double) * n, hipMemcpyHostToDevice, stream)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(db, b, sizeof(double) * m, hipMemcpyHostToDevice, stream)) != hipSuccess) break;
        rc = pdot_round_plan(stream, dX, PDOT_STORAGE_F64, n, n, m, da, db, dX, n, &err);
        if (rc != PDOT_OK) break;
        if ((e = hipMemcpyAsync(X, dX, sizeof(double) * nm, hipMemcpyDeviceToHost, stream)) != hipSuccess) break;
        e = hipStreamSynchronize(stream);
    } while (false);
    if (e != hipSuccess && rc == PDOT_OK) rc = PDOT_E_HIP;
    if (e != hipSuccess && err.empty()) err = std::string("pdot_round_matrix: ") + hipGetErrorString(e);
    if (stream) hipStreamSynchronize(stream);
    if (dX) hipFree(dX);
    if (da) hipFree(da);
    if (db) hipFree(db);
    if (stream) hipStreamDestroy(stream);
    return done(rc);
}

// ------------------------------------------------ allocation-free variant

// Workspace of pdot_round_device_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  64 MiB covers the row and column
// partial sums of plans up to about 30000 x 30000 on a 256-CU device.
constexpr int64_t kNoallocScratchBytes = 64ll << 20;
constexpr int64_t kNoallocScratchWords = kNoallocScratchBytes / (int64_t)sizeof(double);
static __device__ __align__(256) double scratch[kNoallocScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of this device's copy of `scratch` (the current device
// must be `device`).  In host code the array's name is only a host-side
// stand-in; kernels get this address.
int scratch_address(int device, double** out, std::string* err) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            *err = std::string("pdot_round_device_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return PDOT_E_HIP;
        }
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return PDOT_OK;
}

// A device pointer of `device` whose allocation holds [p, p + bytes).
bool device_range_ok(const void* p, int64_t bytes, int device) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) return false;
    if (attr.type != hipMemoryTypeDevice || attr.device != device) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) return false;
    const char* b = static_cast<const char*>(base);
    const char* q = static_cast<const char*>(p);
    return q >= b && bytes >= 0 && (uint64_t)(q - b) + (uint64_t)bytes <= (uint64_t)size;
}

}  // namespace

extern "C" int pdot_round_device_noalloc(int32_t device, void* stream, int64_t n, int64_t m, const double* a,
                                         const double* b, const double* X, int64_t ldx, double* Y, int64_t ldy,
                                         char* errbuf, int64_t errlen) {
    std::string err;
    auto done = [&](int code) {
        if (errbuf && errlen > 0) {
            const size_t k = err.size() < (size_t)(errlen - 1) ? err.size() : (size_t)(errlen - 1);
            err.copy(errbuf, k);
            errbuf[k] = '\0';
        }
        return code;
    };
    if (n < 1 || m < 1 || ldx < n || ldy < n || !a || !b || !X || !Y) {
        err = "pdot_round_device_noalloc: invalid arguments";
        return done(PDOT_E_ARG);
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count || device >= kMaxDevices) {
        err = "pdot_round_device_noalloc: no such device";
        return done(PDOT_E_ARG);
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        err = std::string("pdot_round_device_noalloc: hipSetDevice: ") + hipGetErrorString(e);
        return done(PDOT_E_HIP);
    }
    double* work = nullptr;
    int rc = scratch_address(device, &work, &err);
    if (rc != PDOT_OK) return done(rc);
    const int64_t words = pdot_round_work_words(n, m);
    if (words < 0 || words > kNoallocScratchWords) {
        err = "pdot_round_device_noalloc: a " + std::to_string(n) + " x " + std::to_string(m) + " plan needs " +
              std::to_string(words < 0 ? -1 : words * 8) + " bytes of workspace, above the maximum of " +
              std::to_string(kNoallocScratchBytes);
        return done(PDOT_E_ARG);
    }
    const int64_t xbytes = (ldx * (m - 1) + n) * 8, ybytes = (ldy * (m - 1) + n) * 8;
    const bool ok = device_range_ok(X, xbytes, device) && device_range_ok(Y, ybytes, device) &&
                    device_range_ok(a, n * 8, device) && device_range_ok(b, m * 8, device);
    (void)hipGetLastError();  // a failed attribute query leaves no error for the launches below
    if (!ok) {
        err = "pdot_round_device_noalloc: X, Y, a and b must be device memory of this device holding the given sizes";
        return done(PDOT_E_ARG);
    }
    rc = pdot_round_plan(static_cast<hipStream_t>(stream), X, PDOT_STORAGE_F64, ldx, n, m, a, b, Y, ldy, &err, work,
                         kNoallocScratchWords);
    if (rc != PDOT_OK && err.rfind("pdot_round_plan", 0) == 0) err = "pdot_round_device_noalloc: " + err.substr(17);
    return done(rc);
}
