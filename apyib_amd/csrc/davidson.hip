// Device-side control of the Davidson iteration for the lowest CISD root: the
// projected eigenproblem, the residual coefficients and the convergence test
// run on the GPU, so an iteration needs no host read and a sequence of them
// can be queued ahead of the GPU or replayed from a captured HIP graph.
//
// The projected matrix is kept column by column, g[j * ld + i] = <B_i|S_j>
// for i <= j (each expansion step writes one contiguous column with
// apy_zdot_many); the Hermitian matrix is its upper triangle mirrored.
#include "common.h"

#include <cmath>

namespace {

constexpr int JM = APY_MAXVEC;      // largest subspace
constexpr int EIG_NEG = 1 + JM;     // offset of -theta y in the eigen block

// Lowest eigenpair of the Hermitian k x k matrix by cyclic parallel Jacobi in
// one wavefront.  Each round of the circle-method tournament rotates n/2
// disjoint pairs (p, q); the rotation of a pair depends only on its own 2 x 2
// block, so all row rotations of a round and then all column rotations apply
// at once.  For h = A[p][q] = |h| e, the unitary Q = diag(1, conj e) R,
// R = [[c, s], [-s, c]] with t = s / c the root of t^2 + 2 tau t - 1 = 0,
// tau = (A[q][q] - A[p][p]) / (2 |h|), zeroes A[p][q] in Q^H A Q.  Sweeps stop
// when the off-diagonal norm is below 1e-14 of the Frobenius norm (5-6 sweeps
// at k = 16; the eigenvalue error is second order in it), or after 40; a
// tighter test would sit below the rounding floor of projected matrices whose
// eigenvalues span several orders and run every sweep.
__global__ __launch_bounds__(64) void heev_lowest_kernel(int k, const double2* __restrict__ g, int ld,
                                                         double2* __restrict__ out) {
  __shared__ double2 A[JM][JM + 1];
  __shared__ double2 Q[JM][JM + 1];
  __shared__ double rc[JM / 2], rs[JM / 2];
  __shared__ double2 re[JM / 2];
  __shared__ int rp[JM / 2], rq[JM / 2];
  __shared__ double red[64];
  const int t = threadIdx.x;
  const int n = k + (k & 1);
  for (int e = t; e < JM * JM; e += 64) {
    const int i = e / JM, j = e % JM;
    double2 v = make_double2(0.0, 0.0);
    if (i < k && j < k) {
      if (i <= j) {
        v = g[j * ld + i];
      } else {
        const double2 w = g[i * ld + j];
        v = make_double2(w.x, -w.y);
      }
      if (i == j) v.y = 0.0;
    }
    A[i][j] = v;
    Q[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int sweep = 0; sweep < 40 && k > 1; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int e = t; e < k * k; e += 64) {
      const int i = e / k, j = e % k;
      const double2 v = A[i][j];
      if (i == j) dia += v.x * v.x;
      else off += v.x * v.x + v.y * v.y;
    }
    red[t] = off;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    const double offs = red[0];
    __syncthreads();
    red[t] = dia;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    const double dias = red[0];
    __syncthreads();
    if (!(offs > 1e-28 * (dias + offs))) break;
    for (int round = 0; round < n - 1; ++round) {
      if (t < n / 2) {
        int p, q;
        if (t == 0) {
          p = round;
          q = n - 1;
        } else {
          p = (round + t) % (n - 1);
          q = (round - t + n - 1) % (n - 1);
        }
        if (p > q) { const int x = p; p = q; q = x; }
        double c = 1.0, s = 0.0;
        double2 ph = make_double2(1.0, 0.0);
        if (q < k) {
          const double2 h = A[p][q];
          const double ah = sqrt(h.x * h.x + h.y * h.y);
          if (ah > 0.0) {
            const double tau = (A[q][q].x - A[p][p].x) / (2.0 * ah);
            const double tt = fabs(tau) > 1e150 ? 0.5 / tau
                                                : (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tt * tt);
            s = tt * c;
            ph = make_double2(h.x / ah, h.y / ah);
          }
        }
        rp[t] = p;
        rq[t] = q < k ? q : p;                 // a pair with the padding index is the identity
        rc[t] = c;
        rs[t] = s;
        re[t] = ph;
      }
      __syncthreads();
      // rows: A[p][j] = c A[p][j] - s e A[q][j], A[q][j] = s A[p][j] + c e A[q][j]
      for (int e = t; e < (n / 2) * k; e += 64) {
        const int m = e / k, j = e % k;
        const int p = rp[m], q = rq[m];
        if (p == q) continue;
        const double c = rc[m], s = rs[m];
        const double2 ph = re[m];
        const double2 ap = A[p][j], aq = A[q][j];
        const double2 eaq = make_double2(ph.x * aq.x - ph.y * aq.y, ph.x * aq.y + ph.y * aq.x);
        A[p][j] = make_double2(c * ap.x - s * eaq.x, c * ap.y - s * eaq.y);
        A[q][j] = make_double2(s * ap.x + c * eaq.x, s * ap.y + c * eaq.y);
      }
      __syncthreads();
      // columns of A and Q: X[i][p] = c X[i][p] - s conj(e) X[i][q], X[i][q] = s X[i][p] + c conj(e) X[i][q]
      for (int e = t; e < 2 * (n / 2) * k; e += 64) {
        const int which = e / ((n / 2) * k), r = e % ((n / 2) * k);
        const int m = r / k, i = r % k;
        const int p = rp[m], q = rq[m];
        if (p == q) continue;
        const double c = rc[m], s = rs[m];
        const double2 ph = re[m];
        double2 (*X)[JM + 1] = which ? Q : A;
        const double2 xp = X[i][p], xq = X[i][q];
        const double2 cxq = make_double2(ph.x * xq.x + ph.y * xq.y, ph.x * xq.y - ph.y * xq.x);
        X[i][p] = make_double2(c * xp.x - s * cxq.x, c * xp.y - s * cxq.y);
        X[i][q] = make_double2(s * xp.x + c * cxq.x, s * xp.y + c * cxq.y);
      }
      __syncthreads();
    }
  }
  if (t == 0) {
    int m = 0;
    for (int i = 1; i < k; ++i)
      if (A[i][i].x < A[m][m].x) m = i;
    red[0] = A[m][m].x;
    red[1] = (double)m;
  }
  __syncthreads();
  const double theta = red[0];
  const int m = (int)red[1];
  if (t == 0) out[0] = make_double2(theta, 0.0);
  if (t < JM) {
    const double2 y = t < k ? Q[t][m] : make_double2(0.0, 0.0);
    out[1 + t] = y;
    out[EIG_NEG + t] = t < k ? make_double2(-theta * y.x, -theta * y.y) : make_double2(t == k ? 1.0 : 0.0, 0.0);
  }
  if (t == JM && k == JM) out[EIG_NEG + JM] = make_double2(1.0, 0.0);
}

// Convergence test of one iteration: ctl = [iterations, theta of the previous
// iteration, status (0 running, 1 converged), residual norm]; log[2 (it - 1)]
// = (theta, residual norm) while it <= cap.
__global__ void davidson_check_kernel(const double2* __restrict__ eig, const double2* __restrict__ rr, double e_conv,
                                      double r_conv, double* __restrict__ ctl, double* __restrict__ log, int cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int it = (int)ctl[0] + 1;
  const double theta = eig[0].x, rn = sqrt(fabs(rr[0].x));
  if (it <= cap) {
    log[2 * (it - 1)] = theta;
    log[2 * (it - 1) + 1] = rn;
  }
  if (ctl[2] == 0.0 && it > 1 && fabs(theta - ctl[1]) < e_conv && rn < r_conv) ctl[2] = 1.0;
  ctl[1] = theta;
  ctl[3] = rn;
  ctl[0] = (double)it;
}

constexpr int FNB = 512;   // partial-sum blocks of the fused passes

struct DvList {
  const double2* p[JM];
};

__device__ inline int64_t dv_prow(int64_t x, int64_t N) { return x * (2 * N - x - 1) / 2; }

// (x < y) of the packed pair index P among N
__device__ inline void dv_pdec(int64_t P, int64_t N, int64_t& x, int64_t& y) {
  const double b = 2.0 * (double)N - 1.0, disc = b * b - 8.0 * (double)P;
  int64_t r = (int64_t)((b - sqrt(disc > 0.0 ? disc : 0.0)) * 0.5);
  if (r < 0) r = 0;
  if (r > N - 2) r = N - 2;
  while (r > 0 && dv_prow(r, N) > P) --r;
  while (r + 1 <= N - 2 && dv_prow(r + 1, N) <= P) ++r;
  x = r;
  y = P - dv_prow(r, N) + r + 1;
}

// Fixed-order block sum of NV complex accumulators into part[blockIdx.x][NV].
template <int NV>
__device__ inline void block_sums(double (&ar)[NV], double (&ai)[NV], double2* __restrict__ part) {
  __shared__ double sr[256], si[256];
#pragma unroll 1
  for (int m = 0; m < NV; ++m) {
    sr[threadIdx.x] = ar[m];
    si[threadIdx.x] = ai[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * NV + m] = make_double2(sr[0], si[0]);
    __syncthreads();
  }
}

// One pass over packed vectors [c1 (O V) | c2 (i<j, a<b) | c0] of length len: the residual
// r = sum_j y_j S_j - theta sum_j y_j B_j of the Ritz pair in eig (apy_heev_lowest layout), the
// preconditioned correction d = r / (theta - D) written to d, and the block partials of
// [<B_j|d> (j < K) | <d|d> | <r|r>] (the first Gram-Schmidt dots and the residual norm).
template <int K>
__global__ __launch_bounds__(256) void resid_precond_partial(DvList S, DvList B, const double2* __restrict__ eig,
                                                             const double* __restrict__ eps, int64_t O, int64_t V,
                                                             int64_t len, double2* __restrict__ d,
                                                             double2* __restrict__ part) {
  const double theta = eig[0].x;
  double2 y[K];
#pragma unroll
  for (int j = 0; j < K; ++j) y[j] = eig[1 + j];
  double ar[K + 2], ai[K + 2];
#pragma unroll
  for (int m = 0; m < K + 2; ++m) { ar[m] = 0.0; ai[m] = 0.0; }
  const int64_t OV = O * V, Pv = V * (V - 1) / 2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x) {
    double2 bv[K];
    double rr = 0.0, ri = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const double2 sv = S.p[j][e];
      bv[j] = B.p[j][e];
      const double2 u = make_double2(sv.x - theta * bv[j].x, sv.y - theta * bv[j].y);
      rr += y[j].x * u.x - y[j].y * u.y;
      ri += y[j].x * u.y + y[j].y * u.x;
    }
    double D;
    if (e < OV) {
      D = eps[O + e % V] - eps[e / V];
    } else if (e < len - 1) {
      const int64_t q = e - OV;
      int64_t i, j, a, b;
      dv_pdec(q / Pv, O, i, j);
      dv_pdec(q % Pv, V, a, b);
      D = eps[O + a] + eps[O + b] - eps[i] - eps[j];
    } else {
      D = 0.0;
    }
    double den = theta - D;
    if (fabs(den) < 1e-8) den = den < 0.0 ? -1e-8 : 1e-8;
    const double2 dv = make_double2(rr / den, ri / den);
    d[e] = dv;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      ar[j] += bv[j].x * dv.x + bv[j].y * dv.y;
      ai[j] += bv[j].x * dv.y - bv[j].y * dv.x;
    }
    ar[K] += dv.x * dv.x + dv.y * dv.y;
    ar[K + 1] += rr * rr + ri * ri;
  }
  block_sums<K + 2>(ar, ai, part);
}

// d <- sum_j c_j B_j + c_K d (coefficients in device memory; d read before it is written at each
// index) and the block partials of [<B_j|d> (j < K) | <d|d>] of the result.
template <int K>
__global__ __launch_bounds__(256) void combine_dots_partial(DvList B, const double2* __restrict__ c, int64_t len,
                                                            double2* __restrict__ d, double2* __restrict__ part) {
  double2 cc[K + 1];
#pragma unroll
  for (int j = 0; j <= K; ++j) cc[j] = c[j];
  double ar[K + 1], ai[K + 1];
#pragma unroll
  for (int m = 0; m <= K; ++m) { ar[m] = 0.0; ai[m] = 0.0; }
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x) {
    double2 bv[K];
    const double2 x = d[e];
    double vr = cc[K].x * x.x - cc[K].y * x.y, vi = cc[K].x * x.y + cc[K].y * x.x;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      bv[j] = B.p[j][e];
      vr += cc[j].x * bv[j].x - cc[j].y * bv[j].y;
      vi += cc[j].x * bv[j].y + cc[j].y * bv[j].x;
    }
    d[e] = make_double2(vr, vi);
#pragma unroll
    for (int j = 0; j < K; ++j) {
      ar[j] += bv[j].x * vr + bv[j].y * vi;
      ai[j] += bv[j].x * vi - bv[j].y * vr;
    }
    ar[K] += vr * vr + vi * vi;
  }
  block_sums<K + 1>(ar, ai, part);
}

// value m = blockIdx.x: out[m] = sum of its nb partials, in a fixed order
__global__ void dv_finalize(const double2* __restrict__ part, int nb, int nv, double2* __restrict__ out) {
  __shared__ double sr[256], si[256];
  const int m = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) { a += part[(int64_t)q * nv + m].x; b += part[(int64_t)q * nv + m].y; }
  sr[threadIdx.x] = a;
  si[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = make_double2(sr[0], si[0]);
}

}  // namespace

extern "C" {

size_t apy_davidson_fused_workspace(void) { return (size_t)FNB * (JM + 2) * sizeof(double2); }

int apy_davidson_residual(int k, const void* const* S, const void* const* B, const void* eig, const void* eps,
                          int64_t O, int64_t V, void* d, void* dots, void* workspace, size_t ws_bytes, void* stream) {
  if (k < 1 || k > JM - 1 || !S || !B || !eig || !eps || !d || !dots || O < 2 || V < 2) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_davidson_fused_workspace()) return APY_E_WORKSPACE;
  DvList sl{}, bl{};
  for (int j = 0; j < k; ++j) {
    if (!S[j] || !B[j] || B[j] == d || S[j] == d) return APY_E_ARG;
    sl.p[j] = (const double2*)S[j];
    bl.p[j] = (const double2*)B[j];
  }
  const int64_t len = O * V + (O * (O - 1) / 2) * (V * (V - 1) / 2) + 1;
  hipStream_t st = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  switch (k) {
#define C(K)                                                                                                     \
  case K:                                                                                                        \
    resid_precond_partial<K><<<FNB, 256, 0, st>>>(sl, bl, (const double2*)eig, (const double*)eps, O, V, len,  \
                                                  (double2*)d, part);                                            \
    break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15)
#undef C
  }
  APY_LAUNCH_CHECK();
  dv_finalize<<<(unsigned)(k + 2), 256, 0, st>>>(part, FNB, k + 2, (double2*)dots);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_davidson_combine_dots(int k, const void* const* B, const void* coef, int64_t len, void* d, void* dots,
                              void* workspace, size_t ws_bytes, void* stream) {
  if (k < 1 || k > JM - 1 || !B || !coef || !d || !dots || len < 1) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_davidson_fused_workspace()) return APY_E_WORKSPACE;
  DvList bl{};
  for (int j = 0; j < k; ++j) {
    if (!B[j] || B[j] == d) return APY_E_ARG;
    bl.p[j] = (const double2*)B[j];
  }
  hipStream_t st = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  switch (k) {
#define C(K)                                                                                              \
  case K:                                                                                                 \
    combine_dots_partial<K><<<FNB, 256, 0, st>>>(bl, (const double2*)coef, len, (double2*)d, part);      \
    break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15)
#undef C
  }
  APY_LAUNCH_CHECK();
  dv_finalize<<<(unsigned)(k + 1), 256, 0, st>>>(part, FNB, k + 1, (double2*)dots);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_heev_lowest(int k, const void* g, int ld, void* eig, void* stream) {
  if (k < 1 || k > APY_MAXVEC || ld < k || !g || !eig) return APY_E_ARG;
  heev_lowest_kernel<<<1, 64, 0, (hipStream_t)stream>>>(k, (const double2*)g, ld, (double2*)eig);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_davidson_check(const void* eig, const void* rr, double e_conv, double r_conv, void* ctl, void* log, int cap,
                       void* stream) {
  if (!eig || !rr || !ctl || (cap > 0 && !log) || cap < 0) return APY_E_ARG;
  davidson_check_kernel<<<1, 64, 0, (hipStream_t)stream>>>((const double2*)eig, (const double2*)rr, e_conv, r_conv,
                                                           (double*)ctl, (double*)log, cap);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
