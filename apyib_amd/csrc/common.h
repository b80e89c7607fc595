// Shared definitions for the apyib_amd HIP library (gfx950 / CDNA4 only).
//
// Complex numbers are complex128 stored interleaved (re, im) as double2, the
// memory layout of numpy.complex128 and torch.complex128.  Every entry point
// is stream-ordered: it only enqueues work on the stream it is given and
// returns a status code (APY_OK or a negative APY_E* value).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#include "../../include/apyib_amd.h"

typedef double v4d __attribute__((ext_vector_type(4)));

struct zc { double re, im; };

__host__ __device__ inline zc zmake(double r, double i) { zc z; z.re = r; z.im = i; return z; }
__host__ __device__ inline zc zadd(zc a, zc b) { return zmake(a.re + b.re, a.im + b.im); }
__host__ __device__ inline zc zsub(zc a, zc b) { return zmake(a.re - b.re, a.im - b.im); }
__host__ __device__ inline zc zmul(zc a, zc b) { return zmake(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__host__ __device__ inline zc zconj(zc a) { return zmake(a.re, -a.im); }
__host__ __device__ inline zc zscale(zc a, double s) { return zmake(a.re * s, a.im * s); }

// Complex division by a real denominator: exact componentwise scaling.
__host__ __device__ inline zc zdivr(zc a, double d) { return zmake(a.re / d, a.im / d); }

#define APY_HIP_TRY(expr)                                   \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return APY_E_HIP;                 \
  } while (0)

#define APY_LAUNCH_CHECK()                                  \
  do {                                                      \
    hipError_t _e = hipGetLastError();                      \
    if (_e != hipSuccess) return APY_E_HIP;                 \
  } while (0)

// Grid size for a 1-D elementwise launch over n items with 256-thread
// blocks, capped so that a grid-stride loop covers the rest.
inline unsigned apy_grid1d(int64_t n, int64_t cap = 65536) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Internal (C++) helpers shared between translation units.
// ev0 / ev1 (optional) are recorded just before and after the main GEMM
// kernel, excluding the split-K reduction.
int apy_zgemm_launch(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, hipStream_t s,
                     hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
size_t apy_zgemm_ws_bytes(const apy_gemm_desc* d);
// Workspace of the plan apy_zgemm_launch would run for d now (cached or model plan).
size_t apy_zgemm_plan_ws_bytes(const apy_gemm_desc* d);
// apy_cisd_so_step on a caller's workspace; exact = 1: the GEMM region sized
// for the plans that run now (apy_cisd_so_ws_exact bytes suffice).
int apy_cisd_so_step_ws(const apy_cisd_so_args* A, void* workspace, size_t ws_bytes, void* stream, int exact);
size_t apy_cisd_so_ws_exact(const apy_cisd_so_args* A);
int apy_zgemm_tune_impl(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, int reps, hipStream_t s,
                        int* best_cfg, int* best_split, double* best_ms);
