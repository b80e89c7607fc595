// Finite-difference atomic axial tensors: determinants of displaced-orbital
// overlaps contracted with CI/MP2 amplitudes (aats.py:558-1055 spatial,
// aats.py:120-554 spin-orbital), without forming determinant tensors.
//
// For an overlap S between bra orbitals (rows) and ket orbitals (columns)
// with occupied block P = S[o,o] (no x no, d = det P), every determinant the
// reference builds by replacing up to two occupied rows (i -> a) and up to two
// occupied columns (k -> c) is d times a minor of
//     Gamma = [[ -P^-1, Y ], [ X, W ]],   X = S[v,o] P^-1,  Y = P^-1 S[o,v],
//                                         W = S[v,v] - S[v,o] P^-1 S[o,v],
// (Jacobi's complementary-minor theorem): rows = replaced columns k then the
// new rows a, columns = replaced rows i then the new columns c, sign (-1)^(r c).
// Expanding those minors turns each amplitude-determinant sum of the
// reference into contractions of the amplitudes with x = X^T, y = Y, P^-1
// and W: O(o^2 v^2 (o+v)) per overlap class instead of (ov)^4 determinants.
//
// Kernels:
//   gamma_kernel   one workgroup per overlap: Gauss-Jordan inverse of P with
//                  partial pivoting (d = product of pivots), then x, y, P^-1
//                  (active block), W into a record.
//   axpby_kernel   amplitude combinations conj?(ca A + cb B) (normalised
//                  finite differences of the displaced amplitudes).
//   contract2      six o x v contractions of the doubles with x / y.
//   hmat_g/h       H(M) = P^T M W^T for four o x v matrices.
//   quarter        the four quarter transforms T = (P^T P^T W W) K2.
//   scalars        one workgroup per job: the 20 scalars of APY_AAT_NTERMS.
#include "common.h"

namespace {

constexpr int BLK = 256;

inline __host__ __device__ int64_t rnd16(int64_t n) { return (n + 15) / 16 * 16; }

struct GLayout {
  int64_t oa, nv, x, y, p, w, size;
};

inline __host__ __device__ GLayout glayout(int64_t no, int64_t nf, int64_t nv) {
  GLayout L;
  L.oa = no - nf;
  L.nv = nv;
  L.x = 16;                       // [0] = d, padded to 256 bytes
  L.y = L.x + L.oa * nv;
  L.p = L.y + L.oa * nv;
  L.w = L.p + L.oa * L.oa;
  L.size = rnd16(L.w + nv * nv);
  return L;
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline void cfma(double2& acc, double2 a, double2 b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
}
__device__ inline double2 czero() { return make_double2(0.0, 0.0); }

// ------------------------------------------------------------------ gamma
// ws: batch x no x 2no (Gauss-Jordan [P | I]); yw: batch x no x nv (full Y).
__global__ __launch_bounds__(BLK) void gamma_kernel(int64_t n, int64_t no, int64_t nf, const double2* __restrict__ S,
                                                    int64_t stride_s, double2* __restrict__ gam, double2* ws,
                                                    double2* yw) {
  const int64_t b = blockIdx.x;
  const int64_t nv = n - no, w2 = 2 * no;
  const GLayout L = glayout(no, nf, nv);
  const double2* Sb = S + b * stride_s;
  double2* M = ws + b * no * w2;
  double2* Y = yw + b * no * nv;
  double2* R = gam + b * L.size;
  __shared__ double smax[BLK];
  __shared__ int sidx[BLK];
  __shared__ double2 fac[BLK];
  __shared__ double2 det;
  __shared__ int singular;
  for (int64_t t = threadIdx.x; t < no * w2; t += BLK) {
    const int64_t r = t / w2, c = t % w2;
    M[t] = c < no ? Sb[r * n + c] : make_double2(r == c - no ? 1.0 : 0.0, 0.0);
  }
  if (threadIdx.x == 0) {
    det = make_double2(1.0, 0.0);
    singular = 0;
  }
  __syncthreads();
  for (int64_t col = 0; col < no; ++col) {
    double best = -1.0;
    int bi = -1;
    for (int64_t r = col + threadIdx.x; r < no; r += BLK) {
      const double2 v = M[r * w2 + col];
      const double m = v.x * v.x + v.y * v.y;
      if (m > best) { best = m; bi = (int)r; }
    }
    smax[threadIdx.x] = best;
    sidx[threadIdx.x] = bi;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
        const double o = smax[threadIdx.x + s];
        const int oi = sidx[threadIdx.x + s];
        if (o > smax[threadIdx.x] || (o == smax[threadIdx.x] && oi >= 0 && (sidx[threadIdx.x] < 0 || oi < sidx[threadIdx.x]))) {
          smax[threadIdx.x] = o;
          sidx[threadIdx.x] = oi;
        }
      }
      __syncthreads();
    }
    const int piv = sidx[0];
    const double pm = smax[0];
    __syncthreads();
    if (!(pm > 0.0)) {                 // exactly singular occupied block: d = 0
      if (threadIdx.x == 0) singular = 1;
      __syncthreads();
      break;
    }
    if (piv != col) {
      for (int64_t c = threadIdx.x; c < w2; c += BLK) {
        const double2 t0 = M[col * w2 + c];
        M[col * w2 + c] = M[(int64_t)piv * w2 + c];
        M[(int64_t)piv * w2 + c] = t0;
      }
    }
    __syncthreads();
    const double2 p = M[col * w2 + col];
    const double inv_den = 1.0 / (p.x * p.x + p.y * p.y);
    const double2 pinv = make_double2(p.x * inv_den, -p.y * inv_den);
    __syncthreads();
    if (threadIdx.x == 0) {
      double2 d = cmul(det, p);
      if (piv != col) d = make_double2(-d.x, -d.y);
      det = d;
    }
    for (int64_t c = threadIdx.x; c < w2; c += BLK) M[col * w2 + c] = cmul(M[col * w2 + c], pinv);
    for (int64_t r = threadIdx.x; r < no; r += BLK) fac[r] = (r == col) ? czero() : M[r * w2 + col];
    __syncthreads();
    for (int64_t t = threadIdx.x; t < no * w2; t += BLK) {
      const int64_t r = t / w2, c = t % w2;
      if (r != col) M[t] = csub(M[t], cmul(fac[r], M[col * w2 + c]));
    }
    __syncthreads();
  }
  if (singular) {
    for (int64_t t = threadIdx.x; t < L.size; t += BLK) R[t] = czero();
    return;
  }
  // full Y = P^-1 S[o, v]
  for (int64_t t = threadIdx.x; t < no * nv; t += BLK) {
    const int64_t m = t / nv, c = t % nv;
    double2 acc = czero();
    for (int64_t k = 0; k < no; ++k) cfma(acc, M[m * w2 + no + k], Sb[k * n + no + c]);
    Y[t] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) R[0] = det;
  const int64_t oa = L.oa;
  for (int64_t t = threadIdx.x; t < oa * nv; t += BLK) {
    const int64_t i = t / nv, a = t % nv;
    double2 acc = czero();                               // x[i][a] = (S[v,o] P^-1)[a][nf+i]
    for (int64_t m = 0; m < no; ++m) cfma(acc, Sb[(no + a) * n + m], M[m * w2 + no + nf + i]);
    R[L.x + t] = acc;
    R[L.y + t] = Y[(nf + i) * nv + a];
  }
  for (int64_t t = threadIdx.x; t < oa * oa; t += BLK) {
    const int64_t k = t / oa, i = t % oa;
    R[L.p + t] = M[(nf + k) * w2 + no + nf + i];
  }
  for (int64_t t = threadIdx.x; t < nv * nv; t += BLK) {
    const int64_t a = t / nv, c = t % nv;
    double2 acc = Sb[(no + a) * n + no + c];
    for (int64_t m = 0; m < no; ++m) acc = csub(acc, cmul(Sb[(no + a) * n + m], Y[m * nv + c]));
    R[L.w + t] = acc;
  }
}

// ------------------------------------------------------------------ axpby
__global__ __launch_bounds__(BLK) void axpby_kernel(const apy_aat_axpby_op* __restrict__ ops) {
  const apy_aat_axpby_op op = ops[blockIdx.y];
  const double2* A = (const double2*)op.a;
  const double2* B = (const double2*)op.b;
  double2* out = (double2*)op.out;
  for (int64_t e = blockIdx.x * (int64_t)BLK + threadIdx.x; e < op.n; e += (int64_t)gridDim.x * BLK) {
    double2 v = make_double2(op.ca * A[e].x, op.ca * A[e].y);
    if (B) { v.x += op.cb * B[e].x; v.y += op.cb * B[e].y; }
    if (op.conj) v.y = -v.y;
    out[e] = v;
  }
}

// ------------------------------------------------------------------ jobs
struct JLayout {
  int64_t ov, o2v2, small, per;
};
inline __host__ __device__ JLayout jlayout(int64_t oa, int64_t nv) {
  JLayout J;
  J.ov = oa * nv;
  J.o2v2 = J.ov * J.ov;
  J.small = rnd16(14 * J.ov);     // beta, beta', beta^, kA, kB, k^, G[4], H[4]
  J.per = J.small + 2 * rnd16(J.o2v2);
  return J;
}

__device__ inline int64_t i4(int64_t oa, int64_t nv, int64_t i, int64_t j, int64_t a, int64_t b) {
  return ((i * oa + j) * nv + a) * nv + b;
}

// which: 0 beta[i,a] = sum_jb B2[ijab] x[jb]      1 beta'[j,b] = sum_ia B2[ijab] x[ia]
//        2 beta^[i,a] = sum_jb A(B2)[ijab] x[jb]  3 kA[k,c] = sum_ld K2[klcd] y[ld]
//        4 kB[l,d] = sum_kc K2[klcd] y[kc]        5 k^[k,d] = sum_lc A(K2)[klcd] y[lc]
// with A(T) = (T - T^ab - T^ij + T^ijab) / 4.
__global__ __launch_bounds__(BLK) void contract2(int64_t oa, int64_t nv, const double2* __restrict__ gam, int64_t rec,
                                                 const apy_aat_job* __restrict__ jobs, double2* __restrict__ wsj) {
  const JLayout J = jlayout(oa, nv);
  const int64_t t = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (t >= 6 * J.ov) return;
  const apy_aat_job jb = jobs[blockIdx.y];
  const int which = (int)(t / J.ov);
  const int64_t e = t % J.ov, p = e / nv, q = e % nv;
  const GLayout L = glayout(oa, 0, nv);
  const double2* G = gam + jb.gamma * rec;
  const double2* x = G + L.x;
  const double2* y = G + L.y;
  const double2* T = (const double2*)(which < 3 ? jb.b2 : jb.k2);
  const double2* v = which < 3 ? x : y;
  double2 acc = czero();
  if (T) {
    for (int64_t r = 0; r < oa; ++r) {
      for (int64_t s = 0; s < nv; ++s) {
        double2 val;
        switch (which) {
          case 0: case 3: val = T[i4(oa, nv, p, r, q, s)]; break;
          case 1: case 4: val = T[i4(oa, nv, r, p, s, q)]; break;
          case 2:
            val = csub(csub(T[i4(oa, nv, p, r, q, s)], T[i4(oa, nv, p, r, s, q)]),
                       csub(T[i4(oa, nv, r, p, q, s)], T[i4(oa, nv, r, p, s, q)]));
            val = make_double2(0.25 * val.x, 0.25 * val.y);
            break;
          default:  // 5: k^[k=p, d=q] = sum_{l=r, c=s} A(K2)[k l c d] y[l c]
            val = csub(csub(T[i4(oa, nv, p, r, s, q)], T[i4(oa, nv, p, r, q, s)]),
                       csub(T[i4(oa, nv, r, p, s, q)], T[i4(oa, nv, r, p, q, s)]));
            val = make_double2(0.25 * val.x, 0.25 * val.y);
            break;
        }
        cfma(acc, val, v[r * nv + s]);
      }
    }
  }
  wsj[blockIdx.y * J.per + which * J.ov + e] = acc;
}

// G_w[k,b] = sum_d M_w[k,d] W[b,d]  for M = (K1, k^, kA, kB)
__global__ __launch_bounds__(BLK) void hmat_g(int64_t oa, int64_t nv, const double2* __restrict__ gam, int64_t rec,
                                              const apy_aat_job* __restrict__ jobs, double2* __restrict__ wsj) {
  const JLayout J = jlayout(oa, nv);
  const int64_t t = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (t >= 4 * J.ov) return;
  const apy_aat_job jb = jobs[blockIdx.y];
  const int w = (int)(t / J.ov);
  const int64_t e = t % J.ov, k = e / nv, bb = e % nv;
  const GLayout L = glayout(oa, 0, nv);
  const double2* W = gam + jb.gamma * rec + L.w;
  double2* base = wsj + blockIdx.y * J.per;
  const double2* M = w == 0 ? (const double2*)jb.k1 : base + (w == 1 ? 5 : (w == 2 ? 3 : 4)) * J.ov;
  double2 acc = czero();
  if (M)
    for (int64_t d = 0; d < nv; ++d) cfma(acc, M[k * nv + d], W[bb * nv + d]);
  base[(6 + w) * J.ov + e] = acc;
}

// H_w[i,b] = sum_k P[k,i] G_w[k,b]
__global__ __launch_bounds__(BLK) void hmat_h(int64_t oa, int64_t nv, const double2* __restrict__ gam, int64_t rec,
                                              const apy_aat_job* __restrict__ jobs, double2* __restrict__ wsj) {
  const JLayout J = jlayout(oa, nv);
  const int64_t t = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (t >= 4 * J.ov) return;
  const apy_aat_job jb = jobs[blockIdx.y];
  const int w = (int)(t / J.ov);
  const int64_t e = t % J.ov, i = e / nv, bb = e % nv;
  const GLayout L = glayout(oa, 0, nv);
  const double2* P = gam + jb.gamma * rec + L.p;
  double2* base = wsj + blockIdx.y * J.per;
  const double2* G = base + (6 + w) * J.ov;
  double2 acc = czero();
  for (int64_t k = 0; k < oa; ++k) cfma(acc, P[k * oa + i], G[k * nv + bb]);
  base[(10 + w) * J.ov + e] = acc;
}

// Quarter transforms of K2 into T = (P^T (x) P^T (x) W (x) W) K2, stage 0..3:
//   0: U1[k,l,c,b] = sum_d K2[k,l,c,d] W[b,d]     1: U2[k,l,a,b] = sum_c W[a,c] U1[k,l,c,b]
//   2: U1[k,j,a,b] = sum_l P[l,j] U2[k,l,a,b]     3: U2[i,j,a,b] = sum_k P[k,i] U1[k,j,a,b]
__global__ __launch_bounds__(BLK) void quarter(int stage, int64_t oa, int64_t nv, const double2* __restrict__ gam,
                                               int64_t rec, const apy_aat_job* __restrict__ jobs,
                                               double2* __restrict__ wsj) {
  const JLayout J = jlayout(oa, nv);
  const apy_aat_job jb = jobs[blockIdx.y];
  if (!jb.b2 || !jb.k2) return;
  const GLayout L = glayout(oa, 0, nv);
  const double2* W = gam + jb.gamma * rec + L.w;
  const double2* P = gam + jb.gamma * rec + L.p;
  double2* U1 = wsj + blockIdx.y * J.per + J.small;
  double2* U2 = U1 + rnd16(J.o2v2);
  const double2* K2 = (const double2*)jb.k2;
  for (int64_t e = blockIdx.x * (int64_t)BLK + threadIdx.x; e < J.o2v2; e += (int64_t)gridDim.x * BLK) {
    int64_t r0 = e;
    const int64_t b = r0 % nv; r0 /= nv;
    const int64_t a = r0 % nv; r0 /= nv;
    const int64_t j = r0 % oa;
    const int64_t i = r0 / oa;
    double2 acc = czero();
    switch (stage) {
      case 0:  // (i,j,a,b) = (k,l,c,b)
        for (int64_t d = 0; d < nv; ++d) cfma(acc, K2[i4(oa, nv, i, j, a, d)], W[b * nv + d]);
        U1[e] = acc;
        break;
      case 1:  // (k,l,a,b)
        for (int64_t c = 0; c < nv; ++c) cfma(acc, W[a * nv + c], U1[i4(oa, nv, i, j, c, b)]);
        U2[e] = acc;
        break;
      case 2:  // (k,j,a,b)
        for (int64_t l = 0; l < oa; ++l) cfma(acc, P[l * oa + j], U2[i4(oa, nv, i, l, a, b)]);
        U1[e] = acc;
        break;
      default:  // (i,j,a,b)
        for (int64_t k = 0; k < oa; ++k) cfma(acc, P[k * oa + i], U1[i4(oa, nv, k, j, a, b)]);
        U2[e] = acc;
        break;
    }
  }
}

__device__ double2 block_sum(double2 v, double2* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = cadd(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  const double2 r = sh[0];
  __syncthreads();
  return r;
}

__device__ double2 dot_ov(const double2* a, const double2* b, int64_t n) {
  double2 acc = czero();
  if (a && b)
    for (int64_t e = threadIdx.x; e < n; e += BLK) cfma(acc, a[e], b[e]);
  return acc;
}

__global__ __launch_bounds__(BLK) void scalars(int64_t oa, int64_t nv, const double2* __restrict__ gam, int64_t rec,
                                               const apy_aat_job* __restrict__ jobs, const double2* __restrict__ wsj,
                                               double2* __restrict__ out) {
  __shared__ double2 sh[BLK];
  const JLayout J = jlayout(oa, nv);
  const apy_aat_job jb = jobs[blockIdx.x];
  const GLayout L = glayout(oa, 0, nv);
  const double2* G = gam + jb.gamma * rec;
  const double2* x = G + L.x;
  const double2* y = G + L.y;
  const double2* base = wsj + blockIdx.x * J.per;
  const double2 *beta = base, *betap = base + J.ov, *betah = base + 2 * J.ov;
  const double2 *kA = base + 3 * J.ov, *kB = base + 4 * J.ov, *kh = base + 5 * J.ov;
  const double2 *H0 = base + 10 * J.ov, *H1 = base + 11 * J.ov, *H2 = base + 12 * J.ov, *H3 = base + 13 * J.ov;
  const double2* B1 = (const double2*)jb.b1;
  const double2* K1 = (const double2*)jb.k1;
  const double2* B2 = (const double2*)jb.b2;
  const double2* T = base + J.small + rnd16(J.o2v2);
  double2* o = out + blockIdx.x * (int64_t)APY_AAT_NTERMS;
  const int64_t n = J.ov;
  double2 r[APY_AAT_NTERMS - 1];
  r[0] = dot_ov(B1, x, n);        // sx
  r[1] = dot_ov(K1, y, n);        // sy
  r[2] = dot_ov(beta, x, n);      // bxx
  r[3] = dot_ov(betah, x, n);     // axx
  r[4] = dot_ov(kA, y, n);        // kyyp
  r[5] = dot_ov(kh, y, n);        // -kyy
  r[6] = dot_ov(B1, H0, n);       // s1
  r[7] = dot_ov(beta, H0, n);     // s2
  r[8] = dot_ov(betah, H0, n);    // s3
  r[9] = dot_ov(B1, H1, n);       // s4
  r[10] = dot_ov(B1, H2, n);      // s5
  r[11] = dot_ov(betah, H1, n);   // s6
  r[12] = dot_ov(betah, H2, n);   // s7
  r[13] = dot_ov(beta, H1, n);    // s8
  r[14] = dot_ov(betap, H1, n);   // s9
  r[15] = dot_ov(betap, H3, n);   // s10
  r[16] = dot_ov(beta, H2, n);    // s11
  double2 e1 = czero(), e2 = czero();
  if (B2 && jb.k2) {
    for (int64_t e = threadIdx.x; e < J.o2v2; e += BLK) {
      int64_t r0 = e;
      const int64_t b = r0 % nv; r0 /= nv;
      const int64_t a = r0 % nv; r0 /= nv;
      const int64_t j = r0 % oa;
      const int64_t i = r0 / oa;
      const double2 tv = T[e];
      double2 ta = csub(csub(tv, T[i4(oa, nv, i, j, b, a)]), csub(T[i4(oa, nv, j, i, a, b)], T[i4(oa, nv, j, i, b, a)]));
      ta = make_double2(0.25 * ta.x, 0.25 * ta.y);
      cfma(e1, B2[e], ta);
      cfma(e2, B2[e], tv);
    }
  }
  r[17] = e1;
  r[18] = e2;
  double2 red[APY_AAT_NTERMS - 1];
  for (int q = 0; q < APY_AAT_NTERMS - 1; ++q) red[q] = block_sum(r[q], sh);
  if (threadIdx.x == 0) {
    o[0] = G[0];
    for (int q = 0; q < APY_AAT_NTERMS - 1; ++q) o[1 + q] = red[q];
    o[6] = make_double2(-red[5].x, -red[5].y);   // kyy = -sum k^ y
  }
}

// Norm pieces of each amplitude set: |t1|^2, |t2|^2, Re sum conj(t2[ijab]) t2[ijba]
__global__ __launch_bounds__(BLK) void norms_kernel(int64_t oa, int64_t nv, const void* const* __restrict__ t1s,
                                                    const void* const* __restrict__ t2s, double* __restrict__ out) {
  __shared__ double2 sh[BLK];
  const double2* t1 = (const double2*)t1s[blockIdx.x];
  const double2* t2 = (const double2*)t2s[blockIdx.x];
  const int64_t ov = oa * nv;
  double n1 = 0.0, n2 = 0.0, nx = 0.0;
  if (t1)
    for (int64_t e = threadIdx.x; e < ov; e += BLK) n1 += t1[e].x * t1[e].x + t1[e].y * t1[e].y;
  if (t2)
    for (int64_t e = threadIdx.x; e < ov * ov; e += BLK) {
      const int64_t b = e % nv, a = (e / nv) % nv, ij = e / (nv * nv);
      const double2 u = t2[e], w = t2[(ij * nv + b) * nv + a];
      n2 += u.x * u.x + u.y * u.y;
      nx += u.x * w.x + u.y * w.y;
    }
  const double2 s12 = block_sum(make_double2(n1, n2), sh);
  const double2 s3 = block_sum(make_double2(nx, 0.0), sh);
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x + 0] = s12.x;
    out[3 * blockIdx.x + 1] = s12.y;
    out[3 * blockIdx.x + 2] = s3.x;
  }
}

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int64_t apy_aat_gamma_record(int64_t no, int64_t nf, int64_t nv) {
  if (no < 1 || nf < 0 || nf >= no || nv < 1) return 0;
  return glayout(no, nf, nv).size;
}

size_t apy_aat_gamma_workspace(int64_t n, int64_t no, int64_t batch) {
  if (n <= no || no < 1 || batch < 1) return 0;
  return al256((size_t)batch * no * 2 * no * sizeof(double2)) + al256((size_t)batch * no * (n - no) * sizeof(double2));
}

int apy_aat_gamma(int64_t n, int64_t no, int64_t nf, int64_t batch, const void* s, int64_t stride_s, void* gamma,
                  void* workspace, size_t ws_bytes, void* stream) {
  if (n <= no || no < 1 || no > BLK || nf < 0 || nf >= no || batch < 1 || batch > 2147483647 || !s || !gamma ||
      stride_s < n * n)
    return APY_E_ARG;
  const size_t need = apy_aat_gamma_workspace(n, no, batch);
  if (!workspace || ws_bytes < need) return APY_E_WORKSPACE;
  double2* ws = (double2*)workspace;
  double2* yw = (double2*)((char*)workspace + al256((size_t)batch * no * 2 * no * sizeof(double2)));
  gamma_kernel<<<(unsigned)batch, BLK, 0, (hipStream_t)stream>>>(n, no, nf, (const double2*)s, stride_s,
                                                                 (double2*)gamma, ws, yw);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_aat_axpby_workspace(int64_t count) {
  return count < 1 ? 0 : al256((size_t)count * sizeof(apy_aat_axpby_op));
}

int apy_aat_axpby(int64_t count, const apy_aat_axpby_op* ops, void* workspace, size_t ws_bytes, void* stream) {
  if (count < 1 || count > 65535 || !ops) return APY_E_ARG;
  int64_t nmax = 0;
  for (int64_t q = 0; q < count; ++q) {
    if (ops[q].n < 0 || (ops[q].n > 0 && (!ops[q].a || !ops[q].out))) return APY_E_ARG;
    if (ops[q].n > nmax) nmax = ops[q].n;
  }
  if (nmax == 0) return APY_OK;
  if (!workspace || ws_bytes < apy_aat_axpby_workspace(count)) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  APY_HIP_TRY(hipMemcpyAsync(workspace, ops, count * sizeof(apy_aat_axpby_op), hipMemcpyHostToDevice, st));
  dim3 grid(apy_grid1d(nmax, 4096), (unsigned)count);
  axpby_kernel<<<grid, BLK, 0, st>>>((const apy_aat_axpby_op*)workspace);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_aat_terms_workspace(int64_t oa, int64_t nv, int64_t njobs) {
  if (oa < 1 || nv < 1 || njobs < 1) return 0;
  const JLayout J = jlayout(oa, nv);
  return al256((size_t)njobs * sizeof(apy_aat_job)) + (size_t)njobs * J.per * sizeof(double2);
}

int apy_aat_terms(int64_t oa, int64_t nv, const void* gamma, int64_t rec, int64_t ngamma, int64_t njobs,
                  const apy_aat_job* jobs, void* out, void* workspace, size_t ws_bytes, void* stream) {
  if (oa < 1 || nv < 1 || njobs < 1 || njobs > 65535 || !gamma || !jobs || !out ||
      rec < glayout(oa, 0, nv).size)
    return APY_E_ARG;
  for (int64_t q = 0; q < njobs; ++q)
    if (jobs[q].gamma < 0 || jobs[q].gamma >= ngamma) return APY_E_ARG;
  const size_t need = apy_aat_terms_workspace(oa, nv, njobs);
  if (!workspace || ws_bytes < need) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  apy_aat_job* dj = (apy_aat_job*)workspace;
  double2* wsj = (double2*)((char*)workspace + al256((size_t)njobs * sizeof(apy_aat_job)));
  APY_HIP_TRY(hipMemcpyAsync(dj, jobs, njobs * sizeof(apy_aat_job), hipMemcpyHostToDevice, st));
  const JLayout J = jlayout(oa, nv);
  const double2* g = (const double2*)gamma;
  contract2<<<dim3((unsigned)((6 * J.ov + BLK - 1) / BLK), (unsigned)njobs), BLK, 0, st>>>(oa, nv, g, rec, dj, wsj);
  APY_LAUNCH_CHECK();
  hmat_g<<<dim3((unsigned)((4 * J.ov + BLK - 1) / BLK), (unsigned)njobs), BLK, 0, st>>>(oa, nv, g, rec, dj, wsj);
  APY_LAUNCH_CHECK();
  hmat_h<<<dim3((unsigned)((4 * J.ov + BLK - 1) / BLK), (unsigned)njobs), BLK, 0, st>>>(oa, nv, g, rec, dj, wsj);
  APY_LAUNCH_CHECK();
  for (int stage = 0; stage < 4; ++stage) {
    quarter<<<dim3(apy_grid1d(J.o2v2, 2048), (unsigned)njobs), BLK, 0, st>>>(stage, oa, nv, g, rec, dj, wsj);
    APY_LAUNCH_CHECK();
  }
  scalars<<<(unsigned)njobs, BLK, 0, st>>>(oa, nv, g, rec, dj, wsj, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_aat_norms_workspace(int64_t nset) {
  return nset < 1 ? 0 : 2 * al256((size_t)nset * sizeof(void*));
}

int apy_aat_norms(int64_t oa, int64_t nv, int64_t nset, const void* const* t1, const void* const* t2, void* out,
                  void* workspace, size_t ws_bytes, void* stream) {
  if (oa < 1 || nv < 1 || nset < 1 || nset > 2147483647 || !t1 || !t2 || !out) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_aat_norms_workspace(nset)) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const void** d1 = (const void**)workspace;
  const void** d2 = (const void**)((char*)workspace + al256((size_t)nset * sizeof(void*)));
  APY_HIP_TRY(hipMemcpyAsync(d1, t1, nset * sizeof(void*), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(d2, t2, nset * sizeof(void*), hipMemcpyHostToDevice, st));
  norms_kernel<<<(unsigned)nset, BLK, 0, st>>>(oa, nv, d1, d2, (double*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
