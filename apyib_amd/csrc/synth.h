// Counter-based generator of the seeded synthetic integrals (device side).
// Same definition as apyib_amd/synthetic.py (splitmix64 + Box-Muller).
#pragma once
#include "common.h"

#define APY_GOLD 0x9E3779B97F4A7C15ULL
#define APY_C2 0xD1B54A32D192ED03ULL

__device__ inline uint64_t apy_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ inline double2 apy_cnormal(uint64_t seed, uint64_t tag, uint64_t counter, double sd) {
  uint64_t base = seed * APY_GOLD + tag * APY_C2;
  uint64_t h1 = apy_mix64(base + counter * 2ULL + 1ULL);
  uint64_t h2 = apy_mix64(h1 + APY_GOLD);
  double u1 = ((double)(h1 >> 11) + 0.5) * 0x1p-53;
  double u2 = (double)(h2 >> 11) * 0x1p-53;
  double r = sqrt(-2.0 * log(u1)) * (sd * 0.70710678118654752440);
  double s, c;
  sincos(2.0 * 3.14159265358979323846 * u2, &s, &c);
  return make_double2(r * c, r * s);
}

// <pq||rs> of the synthetic tensor over n spin orbitals.
struct SrcSynth {
  uint64_t seed, tag;
  int64_t n;
  double sd;
  int real_only;
  __device__ double2 g(int64_t p, int64_t q, int64_t r, int64_t s) const {
    uint64_t cnt = (((uint64_t)p * n + q) * n + r) * n + s;
    return apy_cnormal(seed, tag, cnt, sd);
  }
  __device__ double2 xsum(int64_t p, int64_t q, int64_t r, int64_t s) const {
    double2 a = g(p, q, r, s), b = g(q, p, s, r), c = g(q, p, r, s), d = g(p, q, s, r);
    return make_double2((a.x + b.x) - (c.x + d.x), (a.y + b.y) - (c.y + d.y));
  }
  __device__ double2 operator()(int64_t p, int64_t q, int64_t r, int64_t s) const {
    double2 x1 = xsum(p, q, r, s), x2 = xsum(r, s, p, q);
    double2 v = make_double2(0.125 * (x1.x + x2.x), 0.125 * (x1.y - x2.y));
    if (real_only) v.y = 0.0;
    return v;
  }
};
