// Microbenchmarks of the two roofline denominators on the device the library
// runs on: the FP64 matrix-core rate (v_mfma_f64_16x16x4f64 chains from
// registers, no memory traffic) and the HBM copy bandwidth (one read and one
// write stream of 16-byte words).  The caller times a launch with events on
// its stream; flops and bytes per launch are stated with each entry point.
#include "common.h"

namespace {

constexpr int PK_BLOCK = 256;
constexpr int PK_CHAINS = 4;   // independent accumulators per wave

__global__ __launch_bounds__(PK_BLOCK) void mfma_f64_chains(int iters, double* __restrict__ out) {
  v4d acc[PK_CHAINS];
#pragma unroll
  for (int c = 0; c < PK_CHAINS; ++c) acc[c] = (v4d){0.0, 0.0, 0.0, 0.0};
  const double a = 1e-3 * (double)(threadIdx.x + 1), b = 1.0 + 1e-6 * (double)blockIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < PK_CHAINS; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < PK_CHAINS; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[(int64_t)blockIdx.x * PK_BLOCK + threadIdx.x] = s;     // keeps the chains live
}

constexpr int PK_UNROLL = 4;   // independent 16-byte loads in flight per thread before its stores

__global__ __launch_bounds__(PK_BLOCK) void copy_words(int64_t n, const double2* __restrict__ src,
                                                       double2* __restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * PK_BLOCK;
  int64_t e = blockIdx.x * (int64_t)PK_BLOCK + threadIdx.x;
  for (; e + (PK_UNROLL - 1) * stride < n; e += PK_UNROLL * stride) {
    double2 v[PK_UNROLL];
#pragma unroll
    for (int u = 0; u < PK_UNROLL; ++u) v[u] = src[e + u * stride];
#pragma unroll
    for (int u = 0; u < PK_UNROLL; ++u) dst[e + u * stride] = v[u];
  }
  for (; e < n; e += stride) dst[e] = src[e];
}

// Read stream: every thread sums its words (PK_UNROLL independent loads in flight) and writes one
// double, so the traffic is the n x 16 bytes read.
__global__ __launch_bounds__(PK_BLOCK) void read_words(int64_t n, const double2* __restrict__ src,
                                                       double* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * PK_BLOCK;
  int64_t e = blockIdx.x * (int64_t)PK_BLOCK + threadIdx.x;
  double acc[PK_UNROLL] = {};
  for (; e + (PK_UNROLL - 1) * stride < n; e += PK_UNROLL * stride) {
#pragma unroll
    for (int u = 0; u < PK_UNROLL; ++u) {
      const double2 v = src[e + u * stride];
      acc[u] += v.x + v.y;
    }
  }
  for (; e < n; e += stride) acc[0] += src[e].x + src[e].y;
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < PK_UNROLL; ++u) s += acc[u];
  out[blockIdx.x * (int64_t)PK_BLOCK + threadIdx.x] = s;
}

}  // namespace

extern "C" {

double apy_peak_mfma_f64_flops(int blocks, int iters) {
  if (blocks < 1 || iters < 1) return 0.0;
  // per wave and iteration PK_CHAINS MFMAs of 16 x 16 x 4 (2 flops per multiply-add)
  return (double)blocks * (PK_BLOCK / 64) * (double)iters * PK_CHAINS * 16.0 * 16.0 * 4.0 * 2.0;
}

int apy_peak_mfma_f64(int blocks, int iters, void* out, void* stream) {
  if (blocks < 1 || iters < 1 || !out) return APY_E_ARG;
  mfma_f64_chains<<<(unsigned)blocks, PK_BLOCK, 0, (hipStream_t)stream>>>(iters, (double*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_peak_copy(int64_t n, const void* src, void* dst, int blocks, void* stream) {
  if (n < 1 || !src || !dst || src == dst || blocks < 1) return APY_E_ARG;
  copy_words<<<(unsigned)blocks, PK_BLOCK, 0, (hipStream_t)stream>>>(n, (const double2*)src, (double2*)dst);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_peak_read(int64_t n, const void* src, void* out, int blocks, void* stream) {
  if (n < 1 || !src || !out || blocks < 1) return APY_E_ARG;
  read_words<<<(unsigned)blocks, PK_BLOCK, 0, (hipStream_t)stream>>>(n, (const double2*)src, (double*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
