// Spin-adapted closed-shell CISD / CID (the reference's solve_CISD and
// solve_CID, ci_wfn.py:420-574 and :51-167) on spatial orbitals: o occupied
// and v virtual spatial orbitals, Dirac integrals <pq|rs> = (pr|qs) of the
// complex MO tensor, amplitudes t1[i][a] and t2[i][j][a][b] with the pair
// symmetry t2[i][j][a][b] = t2[j][i][b][a].
//
// The doubles residual of ci_wfn.py:466-482 is split into one half R and
// its pair image, R_ijab + R_jiba, wherever the reference writes both:
//   <ab|cd> t_ijcd (PPL) is split into pair-symmetric and antisymmetric
//     parts: with W+[(a<=b),(c<=d)] = <ab|cd> + <ab|dc> (c<d; <ab|cc> on the
//     diagonal), W-[(a<b),(c<d)] = <ab|cd> - <ab|dc>, T+[(i<=j),(c<=d)] =
//     (t_ijcd + t_ijdc)/2, T-[(i<j),(c<d)] = (t_ijcd - t_ijdc)/2,
//     P+ = T+ W+^T and P- = T- W-^T give sum_cd <ab|cd> t_ijcd =
//     P+[(ij)][(ab)] + s_ij s_ab P-[(i<j)][(a<b)] (s = sign of the order),
//     two GEMMs of a quarter of the plain o^2 v^4 flops between them;
//   <kl|ij> t_klab - F_ki t_kjab - F_kj t_ikab (HHL, one pair operator);
//   F_bc t_ijac (Y) and its image F_ac t_ijcb;
//   the six ring einsums: with T[(i,a),(k,c)] = t_ikac, U[(i,a),(k,c)] = t_ikca,
//     X1 = T W2 - U W1,  W2[(k,c),(j,b)] = 2<kb|cj> - <kb|jc>, W1 = <kb|cj>,
//     X2 = U W3,         W3[(k,c),(j,a)] = <ka|jc>,
//     r += X1[(ia),(jb)] + X1[(jb),(ia)] - X2[(ib),(ja)] - X2[(ja),(ib)];
//   singles into doubles: Q1[i][(j,a,b)] = sum_c <ab|cj> t_ic,
//     Q2[(i,j,a)][b] = sum_k <ak|ij> t_kb,  r += Q1 - Q2 + images.
// The epilogue (sp_epi_doubles) reads every operand along b, the contiguous
// index of the output, except the pair images, which it sums along a into an
// LDS tile and reads back transposed.
// Each product is one complex GEMM on the FP64 matrix cores (zgemm.hip);
// the two ladder GEMMs dominate.  The singles residual (ci_wfn.py:
// 457-464) is six GEMMs into one accumulator; the energy (:495) one
// deterministic reduction.  The CID residual of ci_wfn.py:83-93 (halved
// terms, then r + r^T over the pair) is the doubles part of the same
// expression with t1 = 0.
#include "common.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace {

enum SpKind { S_VVVS = 0, S_VVVA, S_OOOO, S_W21, S_W3, S_VVOO, S_L, S_VVVO, S_VOOO, S_S1, S_W4, S_W5, S_W4T, S_COUNT };

// pairs x <= y (symmetric) and x < y (antisymmetric) among N, row-major
__host__ __device__ inline int64_t nsym(int64_t N) { return N * (N + 1) / 2; }
__host__ __device__ inline int64_t nanti(int64_t N) { return N * (N - 1) / 2; }
__host__ __device__ inline int64_t isym(int64_t x, int64_t y, int64_t N) { return x * (2 * N - x + 1) / 2 + y - x; }
__host__ __device__ inline int64_t ianti(int64_t x, int64_t y, int64_t N) { return x * (2 * N - x - 1) / 2 + y - x - 1; }

__host__ __device__ inline int64_t sp_size(int kind, int64_t o, int64_t v) {
  const int64_t ov = o * v;
  switch (kind) {
    case S_VVVS: return nsym(v) * nsym(v);
    case S_VVVA: return nanti(v) * nanti(v);
    case S_OOOO: return o * o * o * o;
    case S_W21: return 2 * ov * ov;
    case S_W3: return ov * ov;
    case S_VVOO: return ov * ov;
    case S_L: return ov * ov;
    case S_VVVO: return v * v * o * v;
    case S_VOOO: return v * o * o * o;
    case S_S1: return ov * ov;
    case S_W4: return v * o * v * v;
    case S_W5: return o * o * o * v;
    case S_W4T: return v * v * o * v;
  }
  return 0;
}

void* sp_ptr(const apy_sp_ints* I, int kind) {
  switch (kind) {
    case S_VVVS: return I->vvvs;
    case S_VVVA: return I->vvva;
    case S_OOOO: return I->oooo;
    case S_W21: return I->w21;
    case S_W3: return I->w3;
    case S_VVOO: return I->vvoo;
    case S_L: return I->l;
    case S_VVVO: return I->vvvo;
    case S_VOOO: return I->vooo;
    case S_S1: return I->s1;
    case S_W4: return I->w4;
    case S_W5: return I->w5;
    case S_W4T: return I->w4t;
  }
  return nullptr;
}

// <pq|rs> = (pr|qs) of the chemists' tensor E[m][m][m][m]
struct Dirac {
  const double2* E;
  int64_t m;
  __device__ double2 operator()(int64_t p, int64_t q, int64_t r, int64_t s) const {
    return E[((p * m + r) * m + q) * m + s];
  }
};

__device__ inline double2 d2lin(double a, double2 x, double b, double2 y) {
  return make_double2(a * x.x + b * y.x, a * x.y + b * y.y);
}

__global__ __launch_bounds__(256) void sp_build_kernel(Dirac g, int kind, int64_t o, int64_t v, double2* out,
                                                       int64_t total) {
  const int64_t ov = o * v;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    double2 val;
    switch (kind) {
      case S_VVVS:
      case S_VVVA: {  // handled by sp_ladder_kernel
        continue;
      }
      case S_OOOO: {  // [(i,j)][(k,l)] = <kl|ij>
        const int64_t l = t % o, k = (t / o) % o, j = (t / (o * o)) % o, i = t / (o * o * o);
        val = g(k, l, i, j);
        break;
      }
      case S_W21: {  // rows (k,c): 2<kb|cj> - <kb|jc>; rows ov + (k,c): -<kb|cj>; columns (j,b)
        const int64_t row = t / ov, col = t % ov;
        const int64_t j = col / v, b = col % v;
        const int64_t r = row % ov, k = r / v, c = r % v;
        const double2 x = g(k, o + b, o + c, j);
        val = row < ov ? d2lin(2.0, x, -1.0, g(k, o + b, j, o + c)) : make_double2(-x.x, -x.y);
        break;
      }
      case S_W3: {  // [(k,c)][(j,a)] = <ka|jc>
        const int64_t k = t / (v * ov), c = (t / ov) % v, j = (t / v) % o, a = t % v;
        val = g(k, o + a, j, o + c);
        break;
      }
      case S_VVOO: {  // [(i,j)][(a,b)] = <ab|ij>
        const int64_t b = t % v, a = (t / v) % v, j = (t / (v * v)) % o, i = t / (v * v * o);
        val = g(o + a, o + b, i, j);
        break;
      }
      case S_L: {  // [(i,j)][(a,b)] = 2<ij|ab> - <ij|ba>
        const int64_t b = t % v, a = (t / v) % v, j = (t / (v * v)) % o, i = t / (v * v * o);
        val = d2lin(2.0, g(i, j, o + a, o + b), -1.0, g(i, j, o + b, o + a));
        break;
      }
      case S_VVVO: {  // [(j,a,b)][c] = <ab|cj>
        const int64_t c = t % v, b = (t / v) % v, a = (t / (v * v)) % v, j = t / (v * v * v);
        val = g(o + a, o + b, o + c, j);
        break;
      }
      case S_VOOO: {  // [(i,j,a)][k] = <ak|ij>
        const int64_t k = t % o, a = (t / o) % v, j = (t / ov) % o, i = t / (ov * o);
        val = g(o + a, k, i, j);
        break;
      }
      case S_S1: {  // [(i,a)][(j,b)] = 2<ja|bi> - <ja|ib>
        const int64_t row = t / ov, col = t % ov;
        const int64_t i = row / v, a = row % v, j = col / v, b = col % v;
        val = d2lin(2.0, g(j, o + a, o + b, i), -1.0, g(j, o + a, i, o + b));
        break;
      }
      case S_W4: {  // [a][(j,b,c)] = 2<aj|bc> - <aj|cb>
        const int64_t c = t % v, b = (t / v) % v, j = (t / (v * v)) % o, a = t / (v * v * o);
        val = d2lin(2.0, g(o + a, j, o + b, o + c), -1.0, g(o + a, j, o + c, o + b));
        break;
      }
      case S_W4T: {  // [b][a][(j,c)] = 2<aj|bc> - <aj|cb>
        const int64_t c = t % v, j = (t / v) % o, a = (t / ov) % v, b = t / (ov * v);
        val = d2lin(2.0, g(o + a, j, o + b, o + c), -1.0, g(o + a, j, o + c, o + b));
        break;
      }
      default: {  // S_W5 [i][(j,k,b)] = 2<kj|ib> - <kj|bi>
        const int64_t b = t % v, k = (t / v) % o, j = (t / ov) % o, i = t / (ov * o);
        val = d2lin(2.0, g(k, j, i, o + b), -1.0, g(k, j, o + b, i));
        break;
      }
    }
    out[t] = val;
  }
}

// W+ and W- from one pass over (a,b,c,d) in v^4 (threads outside a<=b, c<=d exit)
__global__ __launch_bounds__(256) void sp_ladder_kernel(Dirac g, int64_t o, int64_t v, double2* __restrict__ ws,
                                                        double2* __restrict__ wa) {
  const int64_t total = v * v * v * v, sv = nsym(v), av = nanti(v);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = t % v, c = (t / v) % v, b = (t / (v * v)) % v, a = t / (v * v * v);
    if (a > b || c > d) continue;
    const double2 x = g(o + a, o + b, o + c, o + d);
    if (c == d) {
      if (ws) ws[isym(a, b, v) * sv + isym(c, d, v)] = x;
      continue;
    }
    const double2 y = g(o + a, o + b, o + d, o + c);
    if (ws) ws[isym(a, b, v) * sv + isym(c, d, v)] = make_double2(x.x + y.x, x.y + y.y);
    if (wa && a < b) wa[ianti(a, b, v) * av + ianti(c, d, v)] = make_double2(x.x - y.x, x.y - y.y);
  }
}

// T+[(i<=j)][(c<=d)] = (t_ijcd + t_ijdc)/2, T-[(i<j)][(c<d)] = (t_ijcd - t_ijdc)/2
// T+ / T- of nb amplitude vectors (stride L, doubles at offset o*v) into [nb][(i<=j)][(c<=d)] and
// [nb][(i<j)][(c<d)]
__global__ __launch_bounds__(256) void sp_pack_kernel(int64_t o, int64_t v, int64_t nb, int64_t L,
                                                      const double2* __restrict__ tv, double2* __restrict__ ts,
                                                      double2* __restrict__ ta) {
  const int64_t per = o * o * v * v, total = nb * per, sv = nsym(v), av = nanti(v), so = nsym(o), ao = nanti(o);
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = u / per, t = u % per;
    const int64_t d = t % v, c = (t / v) % v, j = (t / (v * v)) % o, i = t / (v * v * o);
    if (i > j || c > d) continue;
    const double2* t2 = tv + q * L + o * v;
    const double2 x = t2[t], y = t2[((i * o + j) * v + d) * v + c];
    ts[q * so * sv + isym(i, j, o) * sv + isym(c, d, v)] =
        c == d ? x : make_double2(0.5 * (x.x + y.x), 0.5 * (x.y + y.y));
    if (i < j && c < d)
      ta[q * ao * av + ianti(i, j, o) * av + ianti(c, d, v)] = make_double2(0.5 * (x.x - y.x), 0.5 * (x.y - y.y));
  }
}

// Hoo[(i,j)][(k,l)] = <kl|ij> - F[k][i] d(l,j) - F[l][j] d(k,i);  g = [2 F_ov | -F_ov]
__global__ __launch_bounds__(256) void sp_prepare_kernel(int64_t o, int64_t v, const double2* __restrict__ F,
                                                         const double2* __restrict__ oooo, double2* __restrict__ hoo,
                                                         double2* __restrict__ gvec) {
  const int64_t n = o + v, o4 = o * o * o * o, ov = o * v;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < o4 + 2 * ov;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < o4) {
      const int64_t l = t % o, k = (t / o) % o, j = (t / (o * o)) % o, i = t / (o * o * o);
      double2 h = oooo[t];
      if (l == j) { const double2 f = F[k * n + i]; h.x -= f.x; h.y -= f.y; }
      if (k == i) { const double2 f = F[l * n + j]; h.x -= f.x; h.y -= f.y; }
      hoo[t] = h;
    } else {
      const int64_t q = t - o4, e = q % ov, i = e / v, a = e % v;
      const double2 f = F[i * n + o + a];
      gvec[q] = q < ov ? make_double2(2.0 * f.x, 2.0 * f.y) : make_double2(-f.x, -f.y);
    }
  }
}

// TU[q][(i,a)][(k,c)] = t_ikac, TU[q][(i,a)][ov + (k,c)] = t_ikca = t_kiac for the nb vectors (stride L)
__global__ __launch_bounds__(256) void sp_tu_kernel(int64_t o, int64_t v, int64_t nb, int64_t L,
                                                    const double2* __restrict__ tv, double2* __restrict__ TU) {
  const int64_t ov = o * v, per = 2 * ov * ov, total = nb * per;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = u / per, t = u % per;
    const int64_t row = t / (2 * ov), col = t % (2 * ov);
    const int64_t i = row / v, a = row % v;
    const int64_t cc = col % ov, k = cc / v, c = cc % v;
    const double2* t2 = tv + q * L + ov;
    // t_ikca = t_kiac by the pair symmetry: both halves read along c
    TU[u] = col < ov ? t2[((i * o + k) * v + a) * v + c] : t2[((k * o + i) * v + a) * v + c];
  }
}

// w21_f[q] = w21[q] + delta_kj F_q[o+b][o+c] on the rows (k,c) < ov (the W2 half) of nb ring operators:
// X1 = T w21_f then carries the F_vv term Y[(i,a),(j,b)] = sum_c t_ijac F_bc, and its pair image
__global__ __launch_bounds__(256) void sp_fold_kernel(int64_t o, int64_t v, int64_t nb, const double2* __restrict__ F,
                                                      int64_t sF, const double2* __restrict__ w21,
                                                      double2* __restrict__ w21_f) {
  const int64_t n = o + v, ov = o * v, per = 2 * ov * ov;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nb * per; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = u / per, t = u % per, row = t / ov, col = t % ov;
    double2 x = w21[u];
    if (row < ov && row / v == col / v) {
      const double2 f = F[q * sF + (o + col % v) * n + o + row % v];
      x.x += f.x;
      x.y += f.y;
    }
    w21_f[u] = x;
  }
}

// Spin-orbital amplitudes of nb closed-shell solutions (spin orbital P = 2p + spin, spin 0 = alpha, as
// utils.py:283-365 orders them), with t2 taken pair-symmetrised (t_ijab + t_jiba)/2:
// T1[I][A] = t1[i][a] for equal spins; T2[I][J][A][B] = t_ijab where
// (I, J) and (A, B) carry the spins (s, s') with s != s', -t_ijba for (s, s') against (s', s), and
// t_ijab - t_ijba when all four spins agree; zero where spin is not conserved.
__global__ __launch_bounds__(256) void sp_to_so_kernel(int64_t o, int64_t v, int64_t nb, int singles,
                                                       const double2* __restrict__ t, double2* __restrict__ T1,
                                                       double2* __restrict__ T2) {
  const int64_t O = 2 * o, V = 2 * v, ov = o * v, L = ov + ov * ov, n1 = O * V, n2 = n1 * n1, per = n1 + n2;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nb * per; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = u / per, w = u % per;
    const double2* tq = t + q * L;
    if (w < n1) {
      const int64_t I = w / V, A = w % V;
      T1[q * n1 + w] = (singles && (I & 1) == (A & 1)) ? tq[(I >> 1) * v + (A >> 1)] : make_double2(0.0, 0.0);
      continue;
    }
    const int64_t e = w - n1;
    const int64_t B = e % V, A = (e / V) % V, J = (e / (V * V)) % O, I = e / (V * V * O);
    const int64_t i = I >> 1, j = J >> 1, a = A >> 1, b = B >> 1;
    const int si = I & 1, sj = J & 1, sa = A & 1, sb = B & 1;
    const double2* t2 = tq + ov;
    // the pair-symmetrised t_ijab and t_ijba: T comes out exactly antisymmetric in (I,J) and in (A,B)
    const double2 p0 = t2[((i * o + j) * v + a) * v + b], p1 = t2[((j * o + i) * v + b) * v + a];
    const double2 q0 = t2[((i * o + j) * v + b) * v + a], q1 = t2[((j * o + i) * v + a) * v + b];
    const double2 x = make_double2(0.5 * (p0.x + p1.x), 0.5 * (p0.y + p1.y));
    const double2 y = make_double2(0.5 * (q0.x + q1.x), 0.5 * (q0.y + q1.y));
    double2 r = make_double2(0.0, 0.0);
    if (si == sa && sj == sb) { r.x += x.x; r.y += x.y; }
    if (si == sb && sj == sa) { r.x -= y.x; r.y -= y.y; }
    T2[q * n2 + e] = r;
  }
}

// T1[q][(i,a)] = t1 of vector q (stride L): the singles of nb vectors stacked for one product
__global__ __launch_bounds__(256) void sp_t1_stack(int64_t ov, int64_t nb, int64_t L, const double2* __restrict__ tv,
                                                   double2* __restrict__ T1) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nb * ov; u += (int64_t)gridDim.x * blockDim.x)
    T1[u] = tv[(u / ov) * L + u % ov];
}

// Epilogue operands; every per-instance array is instance 0's, the others follow at the region
// strides (so*sv for Ps, ao*av for Pa, (ov)^2 for R, Y, X1, X2, Q1, Q2; L for the vectors).
struct SpEpi {
  int64_t o, v, nb, L;
  const double* eps;
  const double2* vvoo;
  const double2* Ps;  // P+ [(i<=j)][(a<=b)]
  const double2* Pa;  // P- [(i<j)][(a<b)]
  const double2* R;
  const double2* Y;
  const double2* X1;
  const double2* X2;
  const double2* Q1;
  const double2* Q2;
  const double2* t2;
  const double2* energy;
  double2* t2_new;
  double2* r2;
  const double2* rb;  // optional response term (doubles part): r += rb - (re[q] + rs + es[q]) rt
  const double2* rt;  // shared by the instances
  const double2* re;
  double2 rs;
  const double2* es;
  int64_t s_eps, s_vvoo;  // per-instance strides of eps and vvoo (0: shared by the instances)
};

// scalar of the response term of instance q: re[q] + rs + es[q] (null arrays read as zero)
__device__ __forceinline__ double2 resp_scalar(const double2* re, double2 rs, const double2* es, int64_t q) {
  if (re) { const double2 x = re[q]; rs.x += x.x; rs.y += x.y; }
  if (es) { const double2 x = es[q]; rs.x += x.x; rs.y += x.y; }
  return rs;
}

constexpr int SET = 32;  // edge of the (a, b) tile of the doubles epilogue

// The doubles epilogue over 32 x 32 tiles of (a, b) for one (instance, i, j): the terms whose
// operands are stored (b, a)-transposed (the pair images of Y, X1, Q1, Q2, the X2 term and the
// a > b half of the ladder) are summed while reading them along a, into one LDS tile, and read
// back along b; every other operand is read along b directly.  All global reads and writes are
// along the contiguous index.
__global__ __launch_bounds__(256) void sp_epi_doubles(SpEpi e) {
  __shared__ double2 T[SET][SET + 1];  // T[b - b0][a - a0]: the transposed terms of output (a, b)
  const int64_t o = e.o, v = e.v, ov = o * v, vv = v * v, per = ov * ov;
  const int64_t nt = (v + SET - 1) / SET, tiles = e.nb * o * o * nt * nt;
  const int64_t so = nsym(o), sv = nsym(v), ao = nanti(o), av = nanti(v);
  const int tx = threadIdx.x % SET, ty = threadIdx.x / SET;
  for (int64_t w = blockIdx.x; w < tiles; w += gridDim.x) {
    int64_t r = w;
    const int64_t tb = r % nt; r /= nt;
    const int64_t ta = r % nt; r /= nt;
    const int64_t j = r % o; r /= o;
    const int64_t i = r % o;
    const int64_t q = r / o, qd = q * per, qv = q * e.L;
    const int64_t a0 = ta * SET, b0 = tb * SET;
    const int64_t i0 = i < j ? i : j, j0 = i < j ? j : i;
    const double2* Ps = e.Ps + q * so * sv + isym(i0, j0, o) * sv;
    const double2* Pa = (i != j) ? e.Pa + q * ao * av + ianti(i0, j0, o) * av : nullptr;
    const double sgn_lower = i < j ? -1.0 : 1.0;  // sign of P- at a > b
    const double2* Y = e.Y ? e.Y + qd : nullptr;
    const double2* X1 = e.X1 + qd;
    const double2* X2 = e.X2 + qd;
    const double2* Q1 = e.Q1 ? e.Q1 + qd : nullptr;
    const double2* Q2 = e.Q2 ? e.Q2 + qd : nullptr;
    __syncthreads();  // T of the previous tile is no longer read
#pragma unroll
    for (int k = 0; k < SET / 8; ++k) {  // element (b, a) = (b0 + ty + 8k, a0 + tx), read along a
      const int64_t b = b0 + ty + 8 * k, a = a0 + tx;
      double2 s = make_double2(0.0, 0.0);
      if (a < v && b < v) {
        const double2 x1t = X1[(j * v + b) * ov + i * v + a];
        const double2 x2 = X2[(i * v + b) * ov + j * v + a];
        s.x = x1t.x - x2.x;
        s.y = x1t.y - x2.y;
        if (Y) {
          const double2 y2 = Y[((j * o + i) * v + b) * v + a];
          s.x += y2.x;
          s.y += y2.y;
        }
        if (Q1) {
          const double2 q1t = Q1[j * (o * vv) + (i * v + b) * v + a];
          const double2 q2t = Q2[((j * o + i) * v + b) * v + a];
          s.x += q1t.x - q2t.x;
          s.y += q1t.y - q2t.y;
        }
        if (b < a) {
          const double2 ps = Ps[isym(b, a, v)];
          s.x += ps.x; s.y += ps.y;
          if (Pa) {
            const double2 pa = Pa[ianti(b, a, v)];
            s.x += sgn_lower * pa.x; s.y += sgn_lower * pa.y;
          }
        }
      }
      T[ty + 8 * k][tx] = s;
    }
    __syncthreads();
    const double2 E = e.energy[q];
    const double* ep = e.eps + q * e.s_eps;
    const double eij = ep[i] + ep[j];
    const double2 rsc = e.rb ? resp_scalar(e.re, e.rs, e.es, q) : make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < SET / 8; ++k) {  // output (a, b) = (a0 + ty + 8k, b0 + tx), read along b
      const int64_t a = a0 + ty + 8 * k, b = b0 + tx;
      if (a >= v || b >= v) continue;
      const int64_t t = ((i * o + j) * v + a) * v + b;
      double2 rr = e.vvoo[q * e.s_vvoo + t];
      const double2 tt = T[tx][ty + 8 * k];
      const double2 p = e.R[qd + t];
      const double2 x1 = X1[(i * v + a) * ov + j * v + b];
      const double2 x2t = X2[(j * v + a) * ov + i * v + b];
      rr.x += tt.x + p.x + x1.x - x2t.x;
      rr.y += tt.y + p.y + x1.y - x2t.y;
      if (Y) {
        const double2 y1 = Y[t];
        rr.x += y1.x;
        rr.y += y1.y;
      }
      if (Q1) {
        const double2 q1 = Q1[i * (o * vv) + (j * v + a) * v + b];
        const double2 q2 = Q2[t];
        rr.x += q1.x - q2.x;
        rr.y += q1.y - q2.y;
      }
      if (a <= b) {
        const double2 ps = Ps[isym(a, b, v)];
        rr.x += ps.x; rr.y += ps.y;
        if (Pa && a < b) {
          const double2 pa = Pa[ianti(a, b, v)];
          rr.x -= sgn_lower * pa.x; rr.y -= sgn_lower * pa.y;
        }
      }
      const double2 x = e.t2[qv + t];
      rr.x -= E.x * x.x - E.y * x.y;
      rr.y -= E.x * x.y + E.y * x.x;
      if (e.rb) {
        const double2 bb = e.rb[qv + t], y = e.rt[t];
        rr.x += bb.x - (rsc.x * y.x - rsc.y * y.y);
        rr.y += bb.y - (rsc.x * y.y + rsc.y * y.x);
      }
      const double D = eij - ep[o + a] - ep[o + b];
      e.r2[qv + t] = rr;
      e.t2_new[qv + t] = make_double2(x.x + rr.x / D, x.y + rr.y / D);
    }
  }
}

// r1 = F[a][i] + acc - E t1 (ci_wfn.py:457-464); t1_new = t1 + r1 / D_ia, per instance (acc [nb][ov],
// vectors L apart)
__global__ __launch_bounds__(256) void sp_epi_singles(int64_t o, int64_t v, int64_t nb, int64_t L,
                                                      const double* __restrict__ eps, const double2* __restrict__ F,
                                                      const double2* __restrict__ acc, const double2* __restrict__ t1,
                                                      const double2* __restrict__ energy, double2* __restrict__ t1_new,
                                                      double2* __restrict__ r1, const double2* __restrict__ rb,
                                                      const double2* __restrict__ rt, const double2* __restrict__ re,
                                                      double2 rs, const double2* __restrict__ es, int64_t s_eps,
                                                      int64_t s_fock) {
  const int64_t n = o + v, ov = o * v, total = nb * ov;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = u / ov, t = u % ov, qv = q * L;
    const int64_t i = t / v, a = t % v;
    const double2 E = energy[q];
    const double2 f = F[q * s_fock + (o + a) * n + i], s = acc[u], x = t1[qv + t];
    double2 r = make_double2(f.x + s.x - (E.x * x.x - E.y * x.y), f.y + s.y - (E.x * x.y + E.y * x.x));
    if (rb) {
      const double2 c = resp_scalar(re, rs, es, q), bb = rb[qv + t], y = rt[t];
      r.x += bb.x - (c.x * y.x - c.y * y.y);
      r.y += bb.y - (c.x * y.y + c.y * y.x);
    }
    const double D = eps[q * s_eps + i] - eps[q * s_eps + o + a];
    r1[qv + t] = r;
    t1_new[qv + t] = make_double2(x.x + r.x / D, x.y + r.y / D);
  }
}

// zero the first n entries of nb vectors L apart in a and b (the CID singles block)
__global__ __launch_bounds__(256) void sp_zero_kernel(int64_t n, int64_t nb, int64_t L, double2* __restrict__ a,
                                                      double2* __restrict__ b) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nb * n; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t at = (u / n) * L + u % n;
    a[at] = make_double2(0.0, 0.0);
    b[at] = make_double2(0.0, 0.0);
  }
}

constexpr int EN_BLOCKS = 256;

// E = sum g[0:ov] t1 + sum L t2 (no conjugation; g[0:ov] = 2 F_ov); instance blockIdx.y of vectors
// `stride` apart, partial sums [instance][EN_BLOCKS]
__global__ __launch_bounds__(256) void sp_energy_partial(int64_t ov, int singles, const double2* __restrict__ gv,
                                                         const double2* __restrict__ Lv, const double2* __restrict__ tv,
                                                         int64_t stride, int64_t s_g, int64_t s_l,
                                                         double2* __restrict__ partial) {
  const int64_t n2 = ov * ov;
  const double2* t = tv + blockIdx.y * stride;
  const double2* gvec = gv ? gv + blockIdx.y * s_g : nullptr;
  const double2* L = Lv + blockIdx.y * s_l;
  double sr = 0.0, si = 0.0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ov + n2; q += (int64_t)gridDim.x * blockDim.x) {
    double2 w, x = t[q];
    if (q < ov) {
      if (!singles) continue;
      w = gvec[q];
    } else {
      w = L[q - ov];
    }
    sr += w.x * x.x - w.y * x.y;
    si += w.x * x.y + w.y * x.x;
  }
  __shared__ double hr[256], hi[256];
  hr[threadIdx.x] = sr;
  hi[threadIdx.x] = si;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { hr[threadIdx.x] += hr[threadIdx.x + w]; hi[threadIdx.x] += hi[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(hr[0], hi[0]);
}

// out[blockIdx.x] = sum of that instance's np partials
__global__ __launch_bounds__(256) void sp_energy_final(const double2* __restrict__ partial_all, int np, double2* out_all) {
  const double2* partial = partial_all + (int64_t)blockIdx.x * np;
  double2* out = out_all + blockIdx.x;
  __shared__ double hr[256], hi[256];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < np; q += blockDim.x) { a += partial[q].x; b += partial[q].y; }
  hr[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { hr[threadIdx.x] += hr[threadIdx.x + w]; hi[threadIdx.x] += hi[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(hr[0], hi[0]);
}

apy_gemm_desc gd(int64_t m, int64_t nn, int64_t k, int al, const void* a, int64_t lda, int bl, const void* b,
                 int64_t ldb, void* c, int64_t ldc, double alpha, double beta, int64_t batch = 1, int64_t sa = 0,
                 int64_t sb = 0, int64_t sc = 0, int reduce = 0) {
  apy_gemm_desc d{};
  d.m = m; d.n = nn; d.k = k;
  d.a_layout = al; d.b_layout = bl;
  d.alpha_re = alpha; d.beta_re = beta;
  d.a = a; d.lda = lda; d.b = b; d.ldb = ldb; d.c = c; d.ldc = ldc;
  d.batch = batch;
  d.reduce_batch = reduce;
  if (batch > 1) { d.stride_a = sa; d.stride_b = sb; d.stride_c = sc; }
  return d;
}

// Workspace regions: each holds nb instances one after another (R, Y, X1, X2, Q1, Q2: (ov)^2 each;
// TU: 2 (ov)^2; Ts/Ps: so*sv; Ta/Pa: ao*av; acc, T1: ov), then the GEMM scratch.
struct SpWS {
  size_t R, Y, X1, X2, TU, Ts, Ta, Ps, Pa, Q1, Q2, acc, t1, gemm, gemm2, total;
};

// One side stream per device (default priority).  A step's products fall into two chains with no
// data between them once TU and T+/- exist (below): the second runs on this stream, so the small
// products of a single vector (ladders with M = o(o+1)/2 rows, rings with M = ov) fill the chip
// together instead of one after another.
hipStream_t sp_side_stream() {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = streams.find(dev);
  if (it != streams.end()) return it->second;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  streams[dev] = s;
  return s;
}

constexpr int NSPG = 14;

// Element strides between the instances of a step over independent systems (0 everywhere: the instances
// share fock, eps, hoo, g and the integral blocks).
struct SpInst {
  int64_t fock, eps, hoo, g, blk[S_COUNT];
};

SpInst sp_inst_strides(int64_t o, int64_t v) {
  SpInst S{};
  const int64_t n = o + v;
  S.fock = n * n;
  S.eps = n;
  S.hoo = o * o * o * o;
  S.g = 2 * o * v;
  for (int k = 0; k < S_COUNT; ++k) S.blk[k] = sp_size(k, o, v);
  return S;
}

// The GEMMs of one step over nb vectors L apart, in order (pointers null when only sizing).  The
// ladder and ring GEMMs stack the instances along M; the ones reading the vectors directly run as
// batched GEMMs with the vector stride.
int sp_gemms(int64_t o, int64_t v, int singles, int64_t nb, int64_t L, const apy_cisd_sp_args* A, char* base,
             const SpWS* w, apy_gemm_desc g[NSPG], const SpInst* S, bool use_w4t = false) {
  // S null: shared integrals, the instances stacked along M of the ladder and ring products; S set: every
  // product batched over the instances with the operands' own strides.
  const SpInst Z{};
  const SpInst& st = S ? *S : Z;
  const int64_t n = o + v, ov = o * v, oo = o * o, vv = v * v, d2 = ov * ov;
  auto P = [&](size_t off) -> double2* { return base ? (double2*)(base + off) : nullptr; };
  const double2* t1 = A ? (const double2*)A->t : nullptr;
  const double2* t2 = t1 ? t1 + ov : nullptr;
  const double2* F = A ? (const double2*)A->fock : nullptr;
  const double2* Fvv = F ? F + o * n + o : nullptr;
  const apy_sp_ints* I = A ? A->ints : nullptr;
  double2* R = P(w->R);
  double2* TU = P(w->TU);
  double2* acc = P(w->acc);
  int q = 0;
  const int64_t so = nsym(o), ao = nanti(o), sv = nsym(v), av = nanti(v);
  // the three-product complex form (apy_cisd_sp_args.gauss3m) for the ladder and ring products
  const int g3 = A && A->gauss3m ? 1 : 0;
  if (S) {  // ladder (P+, P-) and ring (X1, X2) per instance
    g[q++] = gd(so, sv, sv, APY_MK, P(w->Ts), sv, APY_NK, I ? I->vvvs : nullptr, sv, P(w->Ps), sv, 1.0, 0.0, nb,
                so * sv, st.blk[S_VVVS], so * sv);
    if (ao > 0 && av > 0)
      g[q++] = gd(ao, av, av, APY_MK, P(w->Ta), av, APY_NK, I ? I->vvva : nullptr, av, P(w->Pa), av, 1.0, 0.0, nb,
                  ao * av, st.blk[S_VVVA], ao * av);
  } else {
    g[q++] = gd(nb * so, sv, sv, APY_MK, P(w->Ts), sv, APY_NK, I ? I->vvvs : nullptr, sv, P(w->Ps), sv, 1.0, 0.0);  // P+
    if (ao > 0 && av > 0)
      g[q++] = gd(nb * ao, av, av, APY_MK, P(w->Ta), av, APY_NK, I ? I->vvva : nullptr, av, P(w->Pa), av, 1.0, 0.0);  // P-
  }
  for (int e = 0; e < q; ++e) g[e].gauss3m = g3;
  g[q++] = gd(oo, vv, oo, APY_MK, A ? A->hoo : nullptr, oo, APY_KN, t2, vv, R, vv, 1.0, 0.0, nb, st.hoo, L, d2);  // HHL
  // F_vv t2 (Y), unless folded into the ring operator (w21_f = w21 + delta_kj F_bc, apy_cisd_sp_fold_fock)
  const bool fold = A && A->w21_f;
  const void* w21 = fold ? A->w21_f : (I ? I->w21 : nullptr);
  if (!fold)
    g[q++] = gd(oo * v, v, v, APY_MK, t2, v, APY_NK, Fvv, n, P(w->Y), v, 1.0, 0.0, nb, L, st.fock, d2);
  if (S) {
    g[q++] = gd(ov, ov, 2 * ov, APY_MK, TU, 2 * ov, APY_KN, w21, ov, P(w->X1), ov, 1.0, 0.0, nb, 2 * d2,
                st.blk[S_W21], d2);
    g[q++] = gd(ov, ov, ov, APY_MK, TU ? TU + ov : nullptr, 2 * ov, APY_KN, I ? I->w3 : nullptr, ov, P(w->X2), ov,
                1.0, 0.0, nb, 2 * d2, st.blk[S_W3], d2);
  } else {
    g[q++] = gd(nb * ov, ov, 2 * ov, APY_MK, TU, 2 * ov, APY_KN, w21, ov, P(w->X1), ov, 1.0, 0.0);
    g[q++] = gd(nb * ov, ov, ov, APY_MK, TU ? TU + ov : nullptr, 2 * ov, APY_KN, I ? I->w3 : nullptr, ov, P(w->X2),
                ov, 1.0, 0.0);
  }
  g[q - 1].gauss3m = g[q - 2].gauss3m = g3;
  // shared integrals: the t1 of the nb vectors stacked (sp_t1_stack) into one Q1 product of nb*o rows, and
  // the w4 product taken from the T half of TU, rows (q,i), with w4t [b][a][(j,c)] as a sum over b
  const bool stack = !S && nb > 1;
  const bool w4t = !S && use_w4t;
  if (singles) {
    if (stack)
      g[q++] = gd(nb * o, vv * o, v, APY_MK, P(w->t1), v, APY_NK, I ? I->vvvo : nullptr, v, P(w->Q1), vv * o, 1.0,
                  0.0);
    else
      g[q++] = gd(o, vv * o, v, APY_MK, t1, v, APY_NK, I ? I->vvvo : nullptr, v, P(w->Q1), vv * o, 1.0, 0.0, nb, L,
                  st.blk[S_VVVO], d2);
    g[q++] = gd(v * oo, v, o, APY_MK, I ? I->vooo : nullptr, o, APY_KN, t1, v, P(w->Q2), v, 1.0, 0.0, nb,
                st.blk[S_VOOO], L, d2);
    g[q++] = gd(ov, 1, ov, APY_MK, I ? I->s1 : nullptr, ov, APY_KN, t1, 1, acc, 1, 1.0, 0.0, nb, st.blk[S_S1], L, ov);
    g[q++] = gd(ov, 1, 2 * ov, APY_MK, TU, 2 * ov, APY_KN, A ? A->g : nullptr, 1, acc, 1, 1.0, 1.0, nb, 2 * d2, st.g,
                ov);
    if (w4t)
      g[q++] = gd(nb * o, v, ov, APY_MK, TU, v * 2 * ov, APY_NK, I ? I->w4t : nullptr, ov, acc, v, 1.0, 1.0, v, 2 * ov,
                  v * ov, 0, 1);
    else
      g[q++] = gd(o, v, ov * v, APY_MK, t2, ov * v, APY_NK, I ? I->w4 : nullptr, ov * v, acc, v, 1.0, 1.0, nb, L,
                  st.blk[S_W4], ov);
    g[q++] = gd(o, v, oo * v, APY_MK, I ? I->w5 : nullptr, oo * v, APY_KN, t2, v, acc, v, -1.0, 1.0, nb, st.blk[S_W5],
                L, ov);
    g[q++] = gd(o, v, o, APY_KM, F, n, APY_KN, t1, v, acc, v, -1.0, 1.0, nb, st.fock, L, ov);   // -F_oo^T t1
    g[q++] = gd(o, v, v, APY_MK, t1, v, APY_NK, Fvv, n, acc, v, 1.0, 1.0, nb, L, st.fock, ov);  // t1 F_vv^T
  }
  return q;
}

SpWS sp_carve(int64_t o, int64_t v, int singles, int64_t nb, bool multi = false) {
  const size_t z = 16, ov = (size_t)(o * v), d2 = ov * ov * z, B = (size_t)nb;
  auto rnd = [](size_t b) { return (b + 255) / 256 * 256; };
  SpWS w{};
  size_t off = 0;
  w.R = off; off += rnd(B * d2);
  w.Y = off; off += rnd(B * d2);
  w.X1 = off; off += rnd(B * d2);
  w.X2 = off; off += rnd(B * d2);
  w.TU = off; off += rnd(B * 2 * d2);
  const size_t so = (size_t)nsym(o), ao = (size_t)nanti(o), sv = (size_t)nsym(v), av = (size_t)nanti(v);
  w.Ts = off; off += rnd(B * so * sv * z);
  w.Ta = off; off += rnd(B * ao * av * z + z);
  w.Ps = off; off += rnd(B * so * sv * z);
  w.Pa = off; off += rnd(B * ao * av * z + z);
  if (singles) {
    w.Q1 = off; off += rnd(B * d2);
    w.Q2 = off; off += rnd(B * d2);
    w.acc = off; off += rnd(B * ov * z);
    w.t1 = off; off += rnd(B * ov * z);
  }
  w.gemm = off;
  apy_gemm_desc g[NSPG];
  const SpInst S = sp_inst_strides(o, v);
  size_t m = 0;
  for (int form = 0; form < (multi ? 1 : 2); ++form) {   // with and without the w4t form of the w4 product
    const int q = sp_gemms(o, v, singles, nb, (int64_t)(ov + ov * ov), nullptr, nullptr, &w, g, multi ? &S : nullptr,
                           form == 1);
    for (int k = 0; k < q; ++k) m = std::max(m, apy_zgemm_ws_bytes(&g[k]));
  }
  w.gemm2 = off + rnd(std::max(m, (size_t)EN_BLOCKS * z));   // split-K region of the side chain
  w.total = w.gemm2 + std::max(m, (size_t)256);
  return w;
}

// [p, p + n) and [q, q + m) overlap (byte ranges)
bool overlap(const void* p, size_t n, const void* q, size_t m) {
  const char *a = (const char*)p, *b = (const char*)q;
  return a && b && a < b + m && b < a + n;
}

int sp_step_impl(const apy_cisd_sp_args* A, int64_t nb, const void* resp_shift, void* workspace, size_t ws_bytes,
                 void* stream, bool multi = false) {
  if (!A || !A->ints || !A->fock || !A->eps || !A->hoo || !A->g || !A->t || !A->t_new || !A->out || !A->energy ||
      nb < 1)
    return APY_E_ARG;
  const int64_t o = A->o, v = A->v;
  const apy_sp_ints* I = A->ints;
  if (o < 1 || v < 1 || I->o != o || I->v != v) return APY_E_ARG;
  const int64_t ov = o * v, L = ov + ov * ov;
  const size_t vb = (size_t)(nb * L) * 16, one = (size_t)L * 16;
  // t_new may be t itself (in-place update); no other overlap among the vectors written and read
  if (overlap(A->out, vb, A->t, vb) || overlap(A->t_new, vb, A->out, vb)) return APY_E_ARG;
  if (A->t_new != A->t && overlap(A->t_new, vb, A->t, vb)) return APY_E_ARG;
  if (A->resp_b && (!A->resp_t || overlap(A->resp_b, vb, A->out, vb) || overlap(A->resp_b, vb, A->t_new, vb) ||
                    overlap(A->resp_t, one, A->out, vb) || overlap(A->resp_t, one, A->t_new, vb)))
    return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  if (multi && (A->resp_b || resp_shift)) return APY_E_ARG;  // the response term needs shared integrals
  if (!I->vvvs || (v > 1 && !I->vvva) || !I->w21 || !I->w3 || !I->vvoo) return APY_E_ARG;
  if (singles && (!I->vvvo || !I->vooo || !I->s1 || !I->w4 || !I->w5)) return APY_E_ARG;
  const SpWS w = sp_carve(o, v, singles, nb, multi);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  const SpInst S = sp_inst_strides(o, v);
  hipStream_t st = (hipStream_t)stream;
  char* base = (char*)workspace;
  const double2* t = (const double2*)A->t;
  sp_tu_kernel<<<apy_grid1d(nb * 2 * ov * ov), 256, 0, st>>>(o, v, nb, L, t, (double2*)(base + w.TU));
  APY_LAUNCH_CHECK();
  sp_pack_kernel<<<apy_grid1d(nb * ov * ov), 256, 0, st>>>(o, v, nb, L, t, (double2*)(base + w.Ts),
                                                            (double2*)(base + w.Ta));
  APY_LAUNCH_CHECK();
  if (singles && !multi && nb > 1) {
    sp_t1_stack<<<apy_grid1d(nb * ov), 256, 0, st>>>(ov, nb, L, t, (double2*)(base + w.t1));
    APY_LAUNCH_CHECK();
  }
  apy_gemm_desc g[NSPG];
  const int q = sp_gemms(o, v, singles, nb, L, A, base, &w, g, multi ? &S : nullptr, !multi && I->w4t);
  // Two chains with no data between them: P+, HHL (+ Y), X1, Q1 on the caller's stream; P-, X2, Q2 and
  // the singles accumulator products (one after another into acc) forked onto the side stream and
  // joined before the epilogues.  Under stream capture everything stays on the caller's stream.
  auto on_side = [&](const apy_gemm_desc& d) {
    const char* c = (const char*)d.c;
    return c == base + w.Pa || c == base + w.X2 || (singles && (c == base + w.Q2 || c == base + w.acc));
  };
  hipStream_t ss = nullptr;
  {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) ss = sp_side_stream();
  }
  hipEvent_t evf = nullptr, evj = nullptr;
  if (ss) {
    if (hipEventCreateWithFlags(&evf, hipEventDisableTiming) != hipSuccess) return APY_E_HIP;
    if (hipEventCreateWithFlags(&evj, hipEventDisableTiming) != hipSuccess) {
      (void)hipEventDestroy(evf);
      return APY_E_HIP;
    }
    APY_HIP_TRY(hipEventRecord(evf, st));
    APY_HIP_TRY(hipStreamWaitEvent(ss, evf, 0));
  }
  int status = APY_OK;
  const size_t gb = w.gemm2 - w.gemm;
  for (int k = 0; k < q && status == APY_OK; ++k)
    if (ss && on_side(g[k])) status = apy_zgemm_launch(&g[k], base + w.gemm2, ws_bytes - w.gemm2, ss);
  if (ss && hipEventRecord(evj, ss) != hipSuccess && status == APY_OK) status = APY_E_HIP;
  for (int k = 0; k < q && status == APY_OK; ++k)
    if (!ss || !on_side(g[k])) status = apy_zgemm_launch(&g[k], base + w.gemm, gb, st);
  if (ss) {
    // join even after a failed launch, so no later work on the caller's stream races the side stream
    if (hipStreamWaitEvent(st, evj, 0) != hipSuccess && status == APY_OK) status = APY_E_HIP;
    (void)hipEventDestroy(evf);
    (void)hipEventDestroy(evj);
  }
  if (status != APY_OK) return status;
  const double2* rb = (const double2*)A->resp_b;
  const double2* rt = (const double2*)A->resp_t;
  const double2* re = (const double2*)A->resp_energy;
  const double2* es = (const double2*)resp_shift;
  const double2 rs = make_double2(A->resp_shift_re, A->resp_shift_im);
  SpEpi e{o, v, nb, L, A->eps, (const double2*)I->vvoo, (const double2*)(base + w.Ps), (const double2*)(base + w.Pa),
          (const double2*)(base + w.R), A->w21_f ? nullptr : (const double2*)(base + w.Y),
          (const double2*)(base + w.X1), (const double2*)(base + w.X2),
          singles ? (const double2*)(base + w.Q1) : nullptr, singles ? (const double2*)(base + w.Q2) : nullptr,
          t + ov, (const double2*)A->energy, (double2*)A->t_new + ov, (double2*)A->out + ov,
          rb ? rb + ov : nullptr, rb ? rt + ov : nullptr, re, rs, es, multi ? S.eps : 0,
          multi ? S.blk[S_VVOO] : 0};
  const int64_t nt = (v + SET - 1) / SET;
  sp_epi_doubles<<<(unsigned)std::min<int64_t>(nb * o * o * nt * nt, 1 << 20), 256, 0, st>>>(e);
  APY_LAUNCH_CHECK();
  if (singles) {
    sp_epi_singles<<<apy_grid1d(nb * ov), 256, 0, st>>>(o, v, nb, L, A->eps, (const double2*)A->fock,
                                                        (const double2*)(base + w.acc), t, (const double2*)A->energy,
                                                        (double2*)A->t_new, (double2*)A->out, rb, rt, re, rs, es,
                                                        multi ? S.eps : 0, multi ? S.fock : 0);
  } else {  // CID keeps the singles block of the vectors at zero
    sp_zero_kernel<<<apy_grid1d(nb * ov), 256, 0, st>>>(ov, nb, L, (double2*)A->t_new, (double2*)A->out);
  }
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int sp_energy_impl(int64_t o, int64_t v, int singles, int64_t nb, const void* g, const void* L, const void* t,
                   void* energy, void* workspace, size_t ws_bytes, void* stream, int64_t s_g, int64_t s_l) {
  if (o < 1 || v < 1 || nb < 1 || nb > 65535 || !L || !t || !energy || (singles && !g)) return APY_E_ARG;
  if (!workspace || ws_bytes < (size_t)nb * apy_cisd_sp_energy_workspace()) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t ov = o * v;
  sp_energy_partial<<<dim3(EN_BLOCKS, (unsigned)nb), 256, 0, st>>>(ov, singles ? 1 : 0, (const double2*)g,
                                                                    (const double2*)L, (const double2*)t,
                                                                    ov + ov * ov, s_g, s_l, (double2*)workspace);
  APY_LAUNCH_CHECK();
  sp_energy_final<<<(unsigned)nb, 256, 0, st>>>((const double2*)workspace, EN_BLOCKS, (double2*)energy);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // namespace

extern "C" {

int apy_sp_ints_sizes(int64_t o, int64_t v, int64_t sizes_out[13]) {
  if (o < 1 || v < 1 || !sizes_out) return APY_E_ARG;
  for (int k = 0; k < S_COUNT; ++k) sizes_out[k] = sp_size(k, o, v);
  return APY_OK;
}

int apy_sp_ints_from_mo(int64_t m, const void* eri_mo, const apy_sp_ints* ints, void* stream) {
  if (!eri_mo || !ints || ints->o < 1 || ints->v < 1 || ints->o + ints->v != m) return APY_E_ARG;
  Dirac g{(const double2*)eri_mo, m};
  if (ints->vvvs || ints->vvva) {
    const int64_t v = ints->v;
    sp_ladder_kernel<<<apy_grid1d(v * v * v * v, 1 << 20), 256, 0, (hipStream_t)stream>>>(
        g, ints->o, v, (double2*)ints->vvvs, (double2*)ints->vvva);
    APY_LAUNCH_CHECK();
  }
  for (int kind = S_OOOO; kind < S_COUNT; ++kind) {
    double2* out = (double2*)sp_ptr(ints, kind);
    if (!out) continue;
    const int64_t total = sp_size(kind, ints->o, ints->v);
    sp_build_kernel<<<apy_grid1d(total, 1 << 20), 256, 0, (hipStream_t)stream>>>(g, kind, ints->o, ints->v, out,
                                                                                 total);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

int apy_cisd_sp_fold_fock(int64_t o, int64_t v, int64_t nb, const void* fock, const void* w21, void* w21_f,
                          void* stream) {
  if (o < 1 || v < 1 || nb < 1 || !fock || !w21 || !w21_f || w21 == w21_f) return APY_E_ARG;
  const int64_t n = o + v, total = nb * 2 * (o * v) * (o * v);
  sp_fold_kernel<<<apy_grid1d(total, 1 << 20), 256, 0, (hipStream_t)stream>>>(
      o, v, nb, (const double2*)fock, n * n, (const double2*)w21, (double2*)w21_f);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_sp_amplitudes_to_so(int64_t o, int64_t v, int64_t nb, int singles, const void* t, void* t1_so, void* t2_so,
                            void* stream) {
  if (o < 1 || v < 1 || nb < 1 || !t || !t1_so || !t2_so) return APY_E_ARG;
  const int64_t n1 = 4 * o * v, total = nb * (n1 + n1 * n1);
  sp_to_so_kernel<<<apy_grid1d(total, 1 << 20), 256, 0, (hipStream_t)stream>>>(
      o, v, nb, singles ? 1 : 0, (const double2*)t, (double2*)t1_so, (double2*)t2_so);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_cisd_sp_prepare(int64_t o, int64_t v, const void* fock, const apy_sp_ints* ints, void* hoo, void* g,
                        void* stream) {
  if (o < 1 || v < 1 || !fock || !ints || !ints->oooo || !hoo || !g || ints->o != o || ints->v != v)
    return APY_E_ARG;
  const int64_t total = o * o * o * o + 2 * o * v;
  sp_prepare_kernel<<<apy_grid1d(total), 256, 0, (hipStream_t)stream>>>(
      o, v, (const double2*)fock, (const double2*)ints->oooo, (double2*)hoo, (double2*)g);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_cisd_sp_workspace(int64_t o, int64_t v, int singles) {
  if (o < 1 || v < 1) return 0;
  return sp_carve(o, v, singles ? 1 : 0, 1).total;
}

int apy_cisd_sp_step(const apy_cisd_sp_args* A, void* workspace, size_t ws_bytes, void* stream) {
  return sp_step_impl(A, 1, nullptr, workspace, ws_bytes, stream);
}

size_t apy_cisd_sp_batch_workspace(int64_t o, int64_t v, int singles, int64_t nb) {
  if (o < 1 || v < 1 || nb < 1) return 0;
  return sp_carve(o, v, singles ? 1 : 0, nb).total;
}

int apy_cisd_sp_step_batch(const apy_cisd_sp_args* A, int64_t nb, const void* resp_shift, void* workspace,
                           size_t ws_bytes, void* stream) {
  return sp_step_impl(A, nb, resp_shift, workspace, ws_bytes, stream);
}

size_t apy_cisd_sp_energy_workspace(void) { return (size_t)EN_BLOCKS * 16; }

int apy_cisd_sp_energy(int64_t o, int64_t v, int singles, const void* g, const void* L, const void* t, void* energy,
                       void* workspace, size_t ws_bytes, void* stream) {
  return apy_cisd_sp_energy_batch(o, v, singles, 1, g, L, t, energy, workspace, ws_bytes, stream);
}

int apy_cisd_sp_energy_batch(int64_t o, int64_t v, int singles, int64_t nb, const void* g, const void* L,
                             const void* t, void* energy, void* workspace, size_t ws_bytes, void* stream) {
  return sp_energy_impl(o, v, singles, nb, g, L, t, energy, workspace, ws_bytes, stream, 0, 0);
}

int apy_cisd_sp_tune(const apy_cisd_sp_args* A, int64_t nb, int multi, void* workspace, size_t ws_bytes, int reps,
                     void* stream) {
  if (!A || !A->ints || !A->fock || !A->hoo || !A->g || !A->t || nb < 1) return APY_E_ARG;
  const int64_t o = A->o, v = A->v;
  const apy_sp_ints* I = A->ints;
  if (o < 1 || v < 1 || I->o != o || I->v != v) return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  if (!I->vvvs || (v > 1 && !I->vvva) || !I->w21 || !I->w3) return APY_E_ARG;
  if (singles && (!I->vvvo || !I->vooo || !I->s1 || !I->w4 || !I->w5)) return APY_E_ARG;
  const SpWS w = sp_carve(o, v, singles, nb, multi != 0);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  const SpInst S = sp_inst_strides(o, v);
  const int64_t ov = o * v, L = ov + ov * ov;
  char* base = (char*)workspace;
  apy_gemm_desc g[NSPG];
  const int q = sp_gemms(o, v, singles, nb, L, A, base, &w, g, multi ? &S : nullptr, !multi && I->w4t);
  for (int k = 0; k < q; ++k) {
    const int s = apy_zgemm_tune_impl(&g[k], base + w.gemm, ws_bytes - w.gemm, reps, (hipStream_t)stream, nullptr,
                                      nullptr, nullptr);
    if (s != APY_OK) return s;
  }
  return APY_OK;
}

int apy_cisd_sp_energy_multi(int64_t o, int64_t v, int singles, int64_t nb, const void* g, const void* L,
                             const void* t, void* energy, void* workspace, size_t ws_bytes, void* stream) {
  return sp_energy_impl(o, v, singles, nb, g, L, t, energy, workspace, ws_bytes, stream, 2 * o * v,
                        sp_size(S_L, o, v));
}

size_t apy_cisd_sp_multi_workspace(int64_t o, int64_t v, int singles, int64_t nb) {
  if (o < 1 || v < 1 || nb < 1) return 0;
  return sp_carve(o, v, singles ? 1 : 0, nb, true).total;
}

int apy_cisd_sp_step_multi(const apy_cisd_sp_args* A, int64_t nb, void* workspace, size_t ws_bytes, void* stream) {
  return sp_step_impl(A, nb, nullptr, workspace, ws_bytes, stream, true);
}

}  // extern "C"
