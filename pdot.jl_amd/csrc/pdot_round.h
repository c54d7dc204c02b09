// pdot_round.h -- internal interface of the rounding kernels (pdot_round.hip)
// shared with the solver context code (pdot_hip.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

// Feasibility rounding of a device-resident plan X (n x m, column-major,
// leading dimension ld, FP64 if storage == PDOT_STORAGE_F64 else FP32) onto
// the transport polytope of (a, b) (device FP64 vectors), written as FP64 to
// Y (leading dimension ldy; Y may alias X only for FP64 storage with
// ldy == ld).  Enqueued on `stream` and synchronised before returning.
// Returns PDOT_OK or a PDOT_E_* code with *err describing the failure.
// work/work_words: a caller-provided device workspace of doubles; with
// work == nullptr the workspace is allocated (and freed) here.  Nothing in it
// needs clearing beforehand.
int pdot_round_plan(hipStream_t stream, const void* X, int32_t storage, int64_t ld, int64_t n, int64_t m,
                    const double* a, const double* b, double* Y, int64_t ldy, std::string* err,
                    double* work = nullptr, int64_t work_words = 0);

// Doubles of workspace pdot_round_plan needs for an n x m plan on the current
// device (-1 if the launch would not fit).
int64_t pdot_round_work_words(int64_t n, int64_t m);
