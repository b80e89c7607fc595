// Deterministic vector reductions for DIIS (utils.py:104-140) and the
// convergence test (ci_wfn.py:361-365): block partial sums in a fixed order,
// then one finishing block, so repeated calls give identical bits.
#include "common.h"

namespace {

constexpr int NB = 512;  // partial-sum blocks

struct VecList {
  const double2* p[APY_MAXVEC];
};
struct CoefList {
  double re[APY_MAXVEC], im[APY_MAXVEC];
};

// Segment blockIdx.y covers elements [seg * blockIdx.y, seg * blockIdx.y + len) of every vector; its
// partial sums go to partial[blockIdx.y][block][m] (one segment: the plain many-vector dot).
template <int K>
__global__ __launch_bounds__(256) void dot_many_partial(VecList xs, const double2* __restrict__ yv0, int64_t len,
                                                        int64_t seg, int conj_x, double2* __restrict__ partial_all) {
  const int64_t s0 = seg * blockIdx.y;
  const double2* y = yv0 + s0;
  double2* partial = partial_all + (int64_t)blockIdx.y * gridDim.x * K;
  double ar[K], ai[K];
#pragma unroll
  for (int m = 0; m < K; ++m) { ar[m] = 0.0; ai[m] = 0.0; }
  const double cs = conj_x ? -1.0 : 1.0;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < len; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 yv = y[l];
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const double2 x = xs.p[m][s0 + l];
      const double xi = cs * x.y;
      ar[m] += x.x * yv.x - xi * yv.y;
      ai[m] += x.x * yv.y + xi * yv.x;
    }
  }
  __shared__ double sr[256], si[256];
#pragma unroll 1
  for (int m = 0; m < K; ++m) {
    sr[threadIdx.x] = ar[m];
    si[threadIdx.x] = ai[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * K + m] = make_double2(sr[0], si[0]);
    __syncthreads();
  }
}

// segment blockIdx.x, value m = blockIdx.y: out[blockIdx.x][m] = sum of its nb partials (one block
// per value, each in the same fixed order)
__global__ void finalize_many(const double2* __restrict__ partial_all, int nb, int k, double2* out_all) {
  const double2* partial = partial_all + (int64_t)blockIdx.x * nb * k;
  double2* out = out_all + (int64_t)blockIdx.x * k;
  __shared__ double sr[256], si[256];
  const int m = blockIdx.y;
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) { a += partial[(int64_t)q * k + m].x; b += partial[(int64_t)q * k + m].y; }
  sr[threadIdx.x] = a;
  si[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = make_double2(sr[0], si[0]);
}

// sum (x - y)^2 (no conjugation) over [beg, end); with seglen > 0, blockIdx.y = 2 q + h selects
// [q seglen, q seglen + split) (h = 0) or [q seglen + split, (q + 1) seglen) (h = 1) instead
__global__ __launch_bounds__(256) void diffsq_partial(const double2* __restrict__ x, const double2* __restrict__ y,
                                                      int64_t beg, int64_t end, int64_t seglen, int64_t split,
                                                      double2* __restrict__ partial_all) {
  if (seglen > 0) {
    const int64_t q = blockIdx.y / 2, h = blockIdx.y % 2;
    beg = q * seglen + (h ? split : 0);
    end = q * seglen + (h ? seglen : split);
  }
  double2* partial = partial_all + (int64_t)blockIdx.y * gridDim.x;
  double a = 0.0, b = 0.0;
  for (int64_t l = beg + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < end; l += (int64_t)gridDim.x * blockDim.x) {
    const double dr = x[l].x - y[l].x, di = x[l].y - y[l].y;
    a += dr * dr - di * di;
    b += 2.0 * dr * di;
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = a;
  si[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(sr[0], si[0]);
}

template <int K>
__global__ __launch_bounds__(256) void combine_kernel(VecList xs, CoefList c, int64_t len, double2* __restrict__ y) {
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < len; l += (int64_t)gridDim.x * blockDim.x) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const double2 x = xs.p[m][l];
      a += c.re[m] * x.x - c.im[m] * x.y;
      b += c.re[m] * x.y + c.im[m] * x.x;
    }
    y[l] = make_double2(a, b);
  }
}

// y[q seglen + l] = sum_m c[q][m] x_m[q seglen + l], coefficients [nseg][K] in device memory
template <int K>
__global__ __launch_bounds__(256) void combine_seg_kernel(VecList xs, const double2* __restrict__ c, int64_t seglen,
                                                          int64_t total, double2* __restrict__ y,
                                                          const int* __restrict__ skip) {
  if (skip && skip[1] != 0) return;  // an iteration control block whose loop has stopped
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < total; l += (int64_t)gridDim.x * blockDim.x) {
    const double2* cq = c + (l / seglen) * K;
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const double2 x = xs.p[m][l];
      const double2 cm = cq[m];
      a += cm.x * x.x - cm.y * x.y;
      b += cm.x * x.y + cm.y * x.x;
    }
    y[l] = make_double2(a, b);
  }
}

// Gram-Schmidt coefficients from [<B_j|d> (j < k) | <d|d>] (one thread, fixed order)
__global__ void gs_coef_kernel(int k, const double2* __restrict__ dots, int normalize, double2* __restrict__ coef,
                               double2* __restrict__ status) {
  if (threadIdx.x != 0) return;
  double scale = 1.0;
  if (normalize) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += dots[j].x * dots[j].x + dots[j].y * dots[j].y;
    const double dd = dots[k].x, nn = dd - s;
    if (status) *status = make_double2(nn, dd);
    scale = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
  }
  for (int j = 0; j < k; ++j) coef[j] = make_double2(-dots[j].x * scale, -dots[j].y * scale);
  coef[k] = make_double2(scale, 0.0);
}

// CI-metric dots over the vector layout [c1 (OV) | c2 (O^2 V^2, weight 1/4) | c0]:
// sum_l w_l conj(x_m[l]) y[l] in one pass.
template <int K>
__global__ __launch_bounds__(256) void ci_dot_partial(VecList xs, const double2* __restrict__ y, int64_t ov,
                                                      int64_t len, double2* __restrict__ partial) {
  double ar[K], ai[K];
#pragma unroll
  for (int m = 0; m < K; ++m) { ar[m] = 0.0; ai[m] = 0.0; }
  const int64_t l2 = len - 1;  // c0 sits at the end
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < len; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 yv = y[l];
    const double w = (l >= ov && l < l2) ? 0.25 : 1.0;
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const double2 x = xs.p[m][l];
      ar[m] += w * (x.x * yv.x + x.y * yv.y);
      ai[m] += w * (x.x * yv.y - x.y * yv.x);
    }
  }
  __shared__ double sr[256], si[256];
#pragma unroll 1
  for (int m = 0; m < K; ++m) {
    sr[threadIdx.x] = ar[m];
    si[threadIdx.x] = ai[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * K + m] = make_double2(sr[0], si[0]);
    __syncthreads();
  }
}

template <int K>
void launch_ci_dot(VecList v, const double2* y, int64_t ov, int64_t len, double2* part, hipStream_t s) {
  ci_dot_partial<K><<<NB, 256, 0, s>>>(v, y, ov, len, part);
}

template <int K>
void launch_dot(VecList v, const double2* y, int64_t len, int64_t nseg, int conj, double2* part, hipStream_t s) {
  dot_many_partial<K><<<dim3(NB, (unsigned)nseg), 256, 0, s>>>(v, y, len, len, conj, part);
}

template <int K>
void launch_comb_seg(VecList v, const double2* c, int64_t seglen, int64_t nseg, double2* y, const int* skip,
                     hipStream_t s) {
  combine_seg_kernel<K><<<apy_grid1d(seglen * nseg, 8192), 256, 0, s>>>(v, c, seglen, seglen * nseg, y, skip);
}

// ------------------------------------------------------------------ Jacobi + DIIS control on the device
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ inline double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }

// DIIS of utils.py:104-140 over a ring of nslot stored (amplitude, residual) pairs.  row[q] = <r_q | r_slot>
// (conjugated dots of every slot's residual with the newest one, at slot `slot`); `valid` marks the slots of
// the window.  B [nslot][nslot] keeps the dots of earlier iterations (B[q][slot] = row[q], B[slot][q] =
// conj(row[q])); the bordered system [[B_w, -1], [-1, 0]] c = [0, .., 0, -1] over the window's slots (in
// slot order) is solved by Gaussian elimination with partial pivoting (largest |re| + |im|, as LAPACK's
// getrf), threads over rows; coef[q] = c of slot q, zero outside the window.  Nothing is done once the
// control block's status ctl[1] is set; a zero pivot sets it to 3 (singular DIIS system).
__global__ __launch_bounds__(256) void diis_slot_kernel(int nslot, int slot, unsigned valid,
                                                       const double2* __restrict__ row, double2* __restrict__ B,
                                                       double2* __restrict__ coef, int* __restrict__ ctl) {
  constexpr int NM = APY_MAXVEC + 1;
  __shared__ double2 M[NM][NM + 1];
  __shared__ int idx[APY_MAXVEC];
  __shared__ int nv_s, bad_s;
  if (ctl[1] != 0) return;
  const int tid = threadIdx.x;
  if (tid < nslot && ((valid >> tid) & 1u)) {
    B[tid * nslot + slot] = row[tid];
    B[slot * nslot + tid] = make_double2(row[tid].x, -row[tid].y);
  }
  __syncthreads();
  if (tid == 0) {
    int nv = 0;
    for (int q = 0; q < nslot; ++q)
      if ((valid >> q) & 1u) idx[nv++] = q;
    nv_s = nv;
    bad_s = 0;
  }
  __syncthreads();
  const int nv = nv_s, n = nv + 1;
  if (tid < n) {
    for (int b = 0; b < nv; ++b)
      M[tid][b] = tid < nv ? B[idx[tid] * nslot + idx[b]] : make_double2(-1.0, 0.0);
    M[tid][nv] = tid < nv ? make_double2(-1.0, 0.0) : make_double2(0.0, 0.0);
    M[tid][n] = tid < nv ? make_double2(0.0, 0.0) : make_double2(-1.0, 0.0);
  }
  __syncthreads();
  __shared__ double2 fac[NM], xs[NM];
  __shared__ int piv_s;
  for (int c = 0; c < n; ++c) {
    if (tid == 0) {
      int p = c;
      for (int r = c + 1; r < n; ++r)
        if (cabs1(M[r][c]) > cabs1(M[p][c])) p = r;
      if (cabs1(M[p][c]) == 0.0) bad_s = 1;
      piv_s = p;
    }
    __syncthreads();
    if (bad_s) break;
    const int p = piv_s;
    if (p != c && tid <= n - c) {                   // swap rows p and c, one column per thread
      const int j = c + tid;
      const double2 t = M[p][j];
      M[p][j] = M[c][j];
      M[c][j] = t;
    }
    __syncthreads();
    if (tid > c && tid < n) fac[tid] = cdiv(M[tid][c], M[c][c]);
    __syncthreads();
    const int w = n - c;                             // columns c+1..n of rows c+1..n-1, one element per thread
    for (int e = tid; e < (n - 1 - c) * w; e += blockDim.x) {
      const int r = c + 1 + e / w, j = c + 1 + e % w;
      const double2 m = cmul(fac[r], M[c][j]);
      M[r][j] = make_double2(M[r][j].x - m.x, M[r][j].y - m.y);
    }
    __syncthreads();
  }
  if (tid != 0) return;
  if (bad_s) {
    ctl[1] = 3;
    return;
  }
  for (int r = n - 1; r >= 0; --r) {
    double2 acc = M[r][n];
    for (int j = r + 1; j < n; ++j) {
      const double2 m = cmul(M[r][j], xs[j]);
      acc = make_double2(acc.x - m.x, acc.y - m.y);
    }
    xs[r] = cdiv(acc, M[r][r]);
  }
  for (int q = 0; q < nslot; ++q) coef[q] = make_double2(0.0, 0.0);
  for (int a = 0; a < nv; ++a) coef[idx[a]] = xs[a];
}

// Principal square root (the branch of numpy.sqrt on complex numbers), without cancellation.
__device__ inline double2 csqrt_principal(double2 z) {
  if (z.x == 0.0 && z.y == 0.0) return make_double2(0.0, z.y);
  const double r = hypot(z.x, z.y);
  if (z.x >= 0.0) {
    const double re = sqrt(0.5 * (r + z.x));
    return make_double2(re, z.y / (2.0 * re));
  }
  const double im = copysign(sqrt(0.5 * (r - z.x)), z.y);
  return make_double2(z.y / (2.0 * im), im);
}

// a < d for complex a and real d in NumPy's ordering of complex numbers (real part first).
__device__ inline bool np_less(double2 a, double d) { return a.x < d || (a.x == d && a.y < 0.0); }

// The convergence test of ci_wfn.py:361-378 after the next iteration has committed its amplitudes:
// it = ctl[0] + 1, E = energy[0], dE = E_prev - E, rms_k = sqrt(sums[k]) (the principal root of the
// un-conjugated sums of squares of the change); log[it] = (E, dE, rms1, rms2); status ctl[1] = 1 when
// it > 1, |dE| < e_conv and rms2 (and rms1 if check_rms1) < d_conv, else 2 when it reaches max_it.
__global__ void iter_check_kernel(const double2* __restrict__ energy, const double2* __restrict__ sums, int check_rms1,
                                  double e_conv, double d_conv, int max_it, double2* __restrict__ E_prev,
                                  double2* __restrict__ log, int cap, int* __restrict__ ctl) {
  if (threadIdx.x != 0 || ctl[1] != 0) return;
  const int it = ctl[0] + 1;
  ctl[0] = it;
  const double2 E = energy[0], Ep = *E_prev;
  const double2 dE = make_double2(Ep.x - E.x, Ep.y - E.y);
  *E_prev = E;
  const double2 r1 = csqrt_principal(sums[0]), r2 = csqrt_principal(sums[1]);
  if (it < cap) {
    log[4 * it] = E;
    log[4 * it + 1] = dE;
    log[4 * it + 2] = r1;
    log[4 * it + 3] = r2;
  }
  const bool conv = it > 1 && hypot(dE.x, dE.y) < e_conv && np_less(r2, d_conv) && (!check_rms1 || np_less(r1, d_conv));
  if (conv) ctl[1] = 1;
  else if (it >= max_it) ctl[1] = 2;
}

template <int K>
void launch_comb(VecList v, CoefList c, int64_t len, double2* y, hipStream_t s) {
  combine_kernel<K><<<apy_grid1d(len, 8192), 256, 0, s>>>(v, c, len, y);
}

}  // namespace

extern "C" {

size_t apy_reduce_workspace(int64_t len, int k) {
  (void)len;
  if (k < 1) k = 1;
  return (size_t)NB * (size_t)k * sizeof(double2);
}

int apy_zdot_many(int k, const void* const* xs, const void* y, int64_t len, int conj_x, void* out, void* workspace,
                  size_t ws_bytes, void* stream) {
  if (k < 1 || k > APY_MAXVEC || !xs || !y || !out || len < 0) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_reduce_workspace(len, k)) return APY_E_WORKSPACE;
  VecList v{};
  for (int m = 0; m < k; ++m) {
    if (!xs[m]) return APY_E_ARG;
    v.p[m] = (const double2*)xs[m];
  }
  hipStream_t s = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  switch (k) {
#define C(K) case K: launch_dot<K>(v, (const double2*)y, len, 1, conj_x, part, s); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
#undef C
  }
  APY_LAUNCH_CHECK();
  finalize_many<<<dim3(1, (unsigned)k), 256, 0, s>>>(part, NB, k, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_ci_dots(int64_t O, int64_t V, int k, const void* const* xs, const void* y, void* out, void* workspace,
                size_t ws_bytes, void* stream) {
  if (O < 1 || V < 1 || k < 1 || k > APY_MAXVEC || !xs || !y || !out) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_reduce_workspace(0, k)) return APY_E_WORKSPACE;
  const int64_t ov = O * V, len = ov + ov * ov + 1;
  VecList v{};
  for (int m = 0; m < k; ++m) {
    if (!xs[m]) return APY_E_ARG;
    v.p[m] = (const double2*)xs[m];
  }
  hipStream_t s = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  switch (k) {
#define C(K) case K: launch_ci_dot<K>(v, (const double2*)y, ov, len, part, s); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
#undef C
  }
  APY_LAUNCH_CHECK();
  finalize_many<<<dim3(1, (unsigned)k), 256, 0, s>>>(part, NB, k, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_zdiffsq_segments(const void* x, const void* y, int nseg, const int64_t* off, void* out, void* workspace,
                         size_t ws_bytes, void* stream) {
  if (!x || !y || !off || !out || nseg < 1 || nseg > APY_MAXVEC) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_reduce_workspace(0, 1)) return APY_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  for (int q = 0; q < nseg; ++q) {
    if (off[q + 1] < off[q]) return APY_E_ARG;
    diffsq_partial<<<NB, 256, 0, s>>>((const double2*)x, (const double2*)y, off[q], off[q + 1], 0, 0, part);
    APY_LAUNCH_CHECK();
    finalize_many<<<dim3(1, 1), 256, 0, s>>>(part, NB, 1, (double2*)out + q);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

int apy_zcombine(int k, const void* const* xs, const double* coef, int64_t len, void* y, void* stream) {
  if (k < 1 || k > APY_MAXVEC || !xs || !coef || !y || len < 0) return APY_E_ARG;
  VecList v{};
  CoefList c{};
  for (int m = 0; m < k; ++m) {
    if (!xs[m]) return APY_E_ARG;
    v.p[m] = (const double2*)xs[m];
    c.re[m] = coef[2 * m];
    c.im[m] = coef[2 * m + 1];
  }
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
#define C(K) case K: launch_comb<K>(v, c, len, (double2*)y, s); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
#undef C
  }
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_reduce_seg_workspace(int64_t nseg, int k) {
  if (k < 1) k = 1;
  if (nseg < 1) nseg = 1;
  return (size_t)nseg * NB * (size_t)k * sizeof(double2);
}

int apy_zdot_many_seg(int k, const void* const* xs, const void* y, int64_t nseg, int64_t seglen, int conj_x,
                      void* out, void* workspace, size_t ws_bytes, void* stream) {
  if (k < 1 || k > APY_MAXVEC || !xs || !y || !out || seglen < 0 || nseg < 1 || nseg > 65535) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_reduce_seg_workspace(nseg, k)) return APY_E_WORKSPACE;
  VecList v{};
  for (int m = 0; m < k; ++m) {
    if (!xs[m]) return APY_E_ARG;
    v.p[m] = (const double2*)xs[m];
  }
  hipStream_t s = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  switch (k) {
#define C(K) case K: launch_dot<K>(v, (const double2*)y, seglen, nseg, conj_x, part, s); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
#undef C
  }
  APY_LAUNCH_CHECK();
  finalize_many<<<dim3((unsigned)nseg, (unsigned)k), 256, 0, s>>>(part, NB, k, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_zdiffsq_batch(const void* x, const void* y, int64_t nseg, int64_t seglen, int64_t split, void* out,
                      void* workspace, size_t ws_bytes, void* stream) {
  if (!x || !y || !out || nseg < 1 || 2 * nseg > 65535 || seglen < 1 || split < 0 || split > seglen)
    return APY_E_ARG;
  if (!workspace || ws_bytes < apy_reduce_seg_workspace(2 * nseg, 1)) return APY_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  diffsq_partial<<<dim3(NB, (unsigned)(2 * nseg)), 256, 0, s>>>((const double2*)x, (const double2*)y, 0, 0, seglen,
                                                                split, part);
  APY_LAUNCH_CHECK();
  finalize_many<<<dim3((unsigned)(2 * nseg), 1), 256, 0, s>>>(part, NB, 1, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_zcombine_seg(int k, const void* const* xs, const void* coef, int64_t nseg, int64_t seglen, void* y,
                     void* stream) {
  if (k < 1 || k > APY_MAXVEC || !xs || !coef || !y || nseg < 1 || seglen < 1) return APY_E_ARG;
  VecList v{};
  for (int m = 0; m < k; ++m) {
    if (!xs[m]) return APY_E_ARG;
    v.p[m] = (const double2*)xs[m];
  }
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
#define C(K) case K: launch_comb_seg<K>(v, (const double2*)coef, seglen, nseg, (double2*)y, nullptr, s); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
#undef C
  }
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_zcombine_guarded(int k, const void* const* xs, const void* coef, int64_t len, void* y, const int* ctl,
                         void* stream) {
  if (k < 1 || k > APY_MAXVEC || !xs || !coef || !y || len < 1 || !ctl) return APY_E_ARG;
  VecList v{};
  for (int m = 0; m < k; ++m) {
    if (!xs[m]) return APY_E_ARG;
    v.p[m] = (const double2*)xs[m];
  }
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
#define C(K) case K: launch_comb_seg<K>(v, (const double2*)coef, len, 1, (double2*)y, ctl, s); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
#undef C
  }
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_diis_slot(int nslot, int slot, uint32_t valid, const void* row, void* B, void* coef, int* ctl, void* stream) {
  if (nslot < 1 || nslot > APY_MAXVEC || slot < 0 || slot >= nslot || !((valid >> slot) & 1u) ||
      (nslot < 32 && (valid >> nslot) != 0u) || !row || !B || !coef || !ctl)
    return APY_E_ARG;
  diis_slot_kernel<<<1, 256, 0, (hipStream_t)stream>>>(nslot, slot, valid, (const double2*)row, (double2*)B,
                                                      (double2*)coef, ctl);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_iter_check(const void* energy, const void* sums, int check_rms1, double e_conv, double d_conv, int max_it,
                   void* E_prev, void* log, int cap, int* ctl, void* stream) {
  if (!energy || !sums || !E_prev || !log || cap < 1 || !ctl) return APY_E_ARG;
  iter_check_kernel<<<1, 64, 0, (hipStream_t)stream>>>((const double2*)energy, (const double2*)sums, check_rms1,
                                                       e_conv, d_conv, max_it, (double2*)E_prev, (double2*)log, cap,
                                                       ctl);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_gs_coefficients(int k, const void* dots, int normalize, void* coef, void* status, void* stream) {
  if (k < 0 || k >= APY_MAXVEC || !dots || !coef) return APY_E_ARG;
  gs_coef_kernel<<<1, 64, 0, (hipStream_t)stream>>>(k, (const double2*)dots, normalize, (double2*)coef,
                                                    (double2*)status);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
