// Electron-repulsion integrals over contracted Cartesian Gaussians on the GPU (McMurchie-Davidson),
// the AO integrals psi4's MintsHelper gives the reference (hamiltonian.py:33 ao_eri, analytic_aats.py:94
// mo_tei_deriv1 before its MO transform).  A function is one or two terms sharing a centre and the
// exponents, sum_term sum_p c_term,p x^l y^m z^n exp(-a_p r^2): an s/p basis function is one term, its
// derivative with respect to its centre two (l+1 and l-1).  One thread per output element; every
// primitive quartet is expanded in Hermite Gaussians (E coefficients), the Coulomb integrals R_tuv come
// from the Boys function by the Hermite recursion, and the quartet sum is
//   (ab|cd) = 2 pi^(5/2) / (p q sqrt(p+q)) sum_tuv E^ab_tuv sum_t'u'v' (-1)^(t'+u'+v') E^cd_t'u'v'
//             R_(t+t')(u+u')(v+v')(alpha, P - Q).
#include "common.h"

#include <mutex>

namespace {

constexpr int LF = APY_GTO_MAXL;          // largest angular momentum of one term along all axes together
constexpr int LQ = 3 * 2 + LF;            // largest total of a quartet: three d functions and one f term
constexpr int NE = 2 * LF + 1;            // E_t^{ij}, t <= i + j
constexpr int NRB = 64;                   // (Lx+1)(Ly+1)(Lz+1) <= 64 for Lx+Ly+Lz <= 9

// F_n(T), n = 0..nmax, by the series e^-T sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)) for the top order and
// downward recursion F_n = (2T F_{n+1} + e^-T) / (2n+1) (all terms positive: no cancellation).
__device__ void boys_series(int nmax, double T, double* F) {
  double term = 1.0 / (2 * nmax + 1), sum = term;
  for (int k = 1; k < 400; ++k) {
    term *= 2.0 * T / (2 * nmax + 2 * k + 1);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  const double e = exp(-T);
  F[nmax] = e * sum;
  for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + e) / (2 * n + 1);
}

// F_n(T_k), T_k = k / BOYS_DENS below BOYS_TMAX, orders 0..BOYS_NORD-1: filled once per device by
// boys_table_kernel (boys_series at every grid point), read by boys_fn.
constexpr int BOYS_NORD = LQ + 7;
constexpr int BOYS_DENS = 20;
constexpr double BOYS_TMAX = 36.0;
constexpr int BOYS_NGRID = 36 * BOYS_DENS + 2;
__device__ double g_boys[BOYS_NGRID][BOYS_NORD];

__global__ __launch_bounds__(256) void boys_table_kernel() {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= BOYS_NGRID) return;
  double F[BOYS_NORD];
  boys_series(BOYS_NORD - 1, (double)k / BOYS_DENS, F);
  for (int n = 0; n < BOYS_NORD; ++n) g_boys[k][n] = F[n];
}

// F_n(T), n = 0..nmax (nmax <= LQ).  Below BOYS_TMAX the seven-term Taylor expansion about the nearest grid
// point (dF_n/dT = -F_{n+1}):  F_n(T) = sum_j F_{n+j}(T_k) (T_k - T)^j / j!,  |T_k - T| <= 0.025, remainder
// below 1e-15 F_n.  Above, F_0 from erf and upward recursion F_{n+1} = ((2n+1) F_n - e^-T) / (2T) (e^-T is
// then far below every (2n+1) F_n: no cancellation).
__device__ void boys_fn(int nmax, double T, double* F) {
  if (T < BOYS_TMAX) {
    const double x = T * BOYS_DENS;
    const int k = (int)(x + 0.5);
    const double d = ((double)k - x) / BOYS_DENS;
    const double* g = g_boys[k];
    for (int n = 0; n <= nmax; ++n) {
      double f = g[n + 6];
      f = g[n + 5] + d * f * (1.0 / 6.0);
      f = g[n + 4] + d * f * (1.0 / 5.0);
      f = g[n + 3] + d * f * (1.0 / 4.0);
      f = g[n + 2] + d * f * (1.0 / 3.0);
      f = g[n + 1] + d * f * (1.0 / 2.0);
      F[n] = g[n] + d * f;
    }
  } else {
    const double e = exp(-T);
    F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - e) / (2.0 * T);
  }
}

// E_t^{ij} (t = 0..i+j) along one axis for exponents a, b and Q = A - B (i, j <= LF):
//   E^{00}_0 = exp(-ab/(a+b) Q^2),
//   E^{i+1,j}_t = E^{ij}_{t-1} / 2p + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1},   X_PA = -b Q / p,
//   E^{i,j+1}_t = E^{ij}_{t-1} / 2p + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1},   X_PB =  a Q / p.
__device__ void herm_E(int i, int j, double Q, double a, double b, double* out) {
  const double p = a + b, h = 0.5 / p, xpa = -b * Q / p, xpb = a * Q / p;
  double e[LF + 1][LF + 1][NE];
  for (int ii = 0; ii <= i; ++ii)
    for (int jj = 0; jj <= j; ++jj)
      for (int t = 0; t < NE; ++t) e[ii][jj][t] = 0.0;
  e[0][0][0] = exp(-a * b / p * Q * Q);
  for (int ii = 0; ii <= i; ++ii) {
    for (int jj = 0; jj <= j; ++jj) {
      if (ii == 0 && jj == 0) continue;
      const int pi = ii > 0 ? ii - 1 : ii, pj = ii > 0 ? jj : jj - 1;
      const double x = ii > 0 ? xpa : xpb;
      for (int t = 0; t <= ii + jj; ++t) {
        double v = x * e[pi][pj][t];
        if (t > 0) v += h * e[pi][pj][t - 1];
        if (t + 1 <= pi + pj) v += (t + 1) * e[pi][pj][t + 1];
        e[ii][jj][t] = v;
      }
    }
  }
  for (int t = 0; t <= i + j; ++t) out[t] = e[i][j][t];
}

__device__ inline int lsum(const int* l) { return l[0] + l[1] + l[2]; }

#define kNaN __longlong_as_double(0x7ff8000000000000LL)

// A table entry inside the kernels' limits (centre index below ncen, 1..APY_GTO_MAXPRIM primitives, one or
// two terms of angular momentum <= LF); entries outside them give NaN, never an out-of-bounds access.
__device__ bool gto_ok(const apy_gto& g, int64_t ncen) {
  if (g.atom < 0 || g.atom >= ncen || g.nprim < 1 || g.nprim > APY_GTO_MAXPRIM || g.nterm < 1 || g.nterm > 2)
    return false;
  for (int t = 0; t < g.nterm; ++t)
    if (g.lmn[t][0] < 0 || g.lmn[t][1] < 0 || g.lmn[t][2] < 0 || lsum(g.lmn[t]) > LF) return false;
  return true;
}

__host__ __device__ inline int apy_oei_ncomp(int which) {
  if (which <= 0 || which > 31) return 0;
  return (which & 1) + ((which >> 1) & 1) + ((which >> 2) & 1) + 3 * ((which >> 3) & 1) + 3 * ((which >> 4) & 1);
}

// Hermite Coulomb integrals R^0_tuv(alpha, PQ) for t <= Lt[0], u <= Lt[1], v <= Lt[2] (L = the sum, <= LQ;
// (Lt[0]+1)(Lt[1]+1)(Lt[2]+1) <= NRB), from F = F_n(alpha |PQ|^2), n <= L:
//   R^n_000 = (-2 alpha)^n F_n,  R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + PQ_x R^{n+1}_{tuv}  (and alike for u, v).
// Level n is built from level n + 1 in the two alternating buffers Rb; the result is Rb[returned index]
// at [t sy + u sz + v], sz = Lt[2]+1, sy = (Lt[1]+1) sz.
__device__ int herm_R(int L, const int* Lt, double alpha, const double* PQ, const double* F, double (*Rb)[NRB]) {
  const int sz = Lt[2] + 1, sy = (Lt[1] + 1) * sz;
  int cur = 0;
  for (int n = L; n >= 0; --n) {
    const int nxt = cur ^ 1;  // level n + 1 lives in Rb[cur]
    double* out = Rb[nxt];
    const double* prev = Rb[cur];
    const int lim = L - n;
    for (int t = 0; t <= Lt[0]; ++t)
      for (int u = 0; u <= Lt[1]; ++u)
        for (int v = 0; v <= Lt[2]; ++v) {
          if (t + u + v > lim) continue;
          double r;
          if (t > 0) {
            r = PQ[0] * prev[(t - 1) * sy + u * sz + v];
            if (t > 1) r += (t - 1) * prev[(t - 2) * sy + u * sz + v];
          } else if (u > 0) {
            r = PQ[1] * prev[u * sz - sz + v];
            if (u > 1) r += (u - 1) * prev[(u - 2) * sz + v];
          } else if (v > 0) {
            r = PQ[2] * prev[v - 1];
            if (v > 1) r += (v - 1) * prev[v - 2];
          } else {
            double s = 1.0;
            for (int k = 0; k < n; ++k) s *= -2.0 * alpha;
            r = s * F[n];
          }
          out[t * sy + u * sz + v] = r;
        }
    cur = nxt;
  }
  return cur;
}

// One primitive quartet of (a b | c d): terms (ta, tb, tc, td), primitives (pa, pb, pc, pd), centres ca..cd,
// qab = A - B, qcd = C - D.  The caller has checked the table entries and that the quartet's total angular
// momentum is at most LQ.
__device__ double eri_prim(const apy_gto& A, const apy_gto& B, const apy_gto& C, const apy_gto& D, int ta, int tb,
                           int tc, int td, int pa, int pb, int pc, int pd, const double* ca, const double* cb,
                           const double* cc, const double* cd, const double* qab, const double* qcd) {
  const int* la = A.lmn[ta];
  const int* lb = B.lmn[tb];
  const int* lc = C.lmn[tc];
  const int* ld = D.lmn[td];
  int L1[3], L2[3], Lt[3];
  for (int x = 0; x < 3; ++x) { L1[x] = la[x] + lb[x]; L2[x] = lc[x] + ld[x]; Lt[x] = L1[x] + L2[x]; }
  const int L = Lt[0] + Lt[1] + Lt[2];
  const int sz = Lt[2] + 1, sy = (Lt[1] + 1) * sz;  // R index (t, u, v) -> t sy + u sz + v
  const double a = A.exps[pa], b = B.exps[pb], p = a + b;
  const double c = C.exps[pc], d = D.exps[pd], q = c + d;
  double Eab[3][NE], Ecd[3][NE], PQ[3];
  for (int x = 0; x < 3; ++x) {
    herm_E(la[x], lb[x], qab[x], a, b, Eab[x]);
    herm_E(lc[x], ld[x], qcd[x], c, d, Ecd[x]);
    PQ[x] = (a * ca[x] + b * cb[x]) / p - (c * cc[x] + d * cd[x]) / q;
  }
  const double alpha = p * q / (p + q);
  double F[LQ + 1];
  boys_fn(L, alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]), F);
  double Rb[2][NRB];
  const double* R = Rb[herm_R(L, Lt, alpha, PQ, F, Rb)];
  double val = 0.0;
  for (int t = 0; t <= L1[0]; ++t)
    for (int u = 0; u <= L1[1]; ++u)
      for (int v = 0; v <= L1[2]; ++v) {
        const double e1 = Eab[0][t] * Eab[1][u] * Eab[2][v];
        double s2 = 0.0;
        for (int tt = 0; tt <= L2[0]; ++tt)
          for (int uu = 0; uu <= L2[1]; ++uu)
            for (int vv = 0; vv <= L2[2]; ++vv) {
              const double e2 = Ecd[0][tt] * Ecd[1][uu] * Ecd[2][vv];
              const double sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
              s2 += sg * e2 * R[(t + tt) * sy + (u + uu) * sz + (v + vv)];
            }
        val += e1 * s2;
      }
  const double pref0 = 2.0 * M_PI * M_PI * sqrt(M_PI);
  return A.coefs[ta][pa] * B.coefs[tb][pb] * C.coefs[tc][pc] * D.coefs[td][pd] * pref0 / (p * q * sqrt(p + q)) * val;
}

// Whether every term combination of the quartet stays within the total angular momentum LQ.
__device__ bool quartet_ok(const apy_gto& A, const apy_gto& B, const apy_gto& C, const apy_gto& D) {
  for (int ta = 0; ta < A.nterm; ++ta)
    for (int tb = 0; tb < B.nterm; ++tb)
      for (int tc = 0; tc < C.nterm; ++tc)
        for (int td = 0; td < D.nterm; ++td)
          if (lsum(A.lmn[ta]) + lsum(B.lmn[tb]) + lsum(C.lmn[tc]) + lsum(D.lmn[td]) > LQ) return false;
  return true;
}

// x = hi (hi + 1) / 2 + lo with 0 <= lo <= hi: the row-packed lower triangle.
__device__ inline void tri_split(int64_t x, int64_t& hi, int64_t& lo) {
  int64_t h = (int64_t)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
  while (h * (h + 1) / 2 > x) --h;
  while ((h + 1) * (h + 2) / 2 <= x) ++h;
  hi = h;
  lo = x - h * (h + 1) / 2;
}

// Which permutational symmetry of the real integrals (ab|cd) the launch uses: ERI_ALL when the four
// tables are one table (8 images per computed quartet: i >= j, k >= l, (ij) >= (kl)), ERI_KET when the
// ket tables are one table ((ab|cd) = (ab|dc): 2 images, k >= l), ERI_NONE otherwise.
enum EriSym { ERI_NONE = 0, ERI_KET = 1, ERI_ALL = 2 };

// One wavefront per computed element (wave-uniform control flow): the lanes split the element's
// (term, primitive) quartets and a fixed-order butterfly sums them; lane 0 writes the value to every
// image of the element under the launch's symmetry.  A table entry outside the limits (angular
// momentum, primitive count, term count, atom index) gives NaN before any indexing.
__global__ __launch_bounds__(256) void eri_kernel(int64_t ngeom, int64_t natom, const double* __restrict__ centers,
                                                  const apy_gto* __restrict__ fa, int64_t na,
                                                  const apy_gto* __restrict__ fb, int64_t nb,
                                                  const apy_gto* __restrict__ fc, int64_t nc,
                                                  const apy_gto* __restrict__ fd, int64_t nd, int sym,
                                                  double2* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t full = na * nb * nc * nd;
  const int64_t pc = nc * (nc + 1) / 2;
  const int64_t per = sym == ERI_ALL ? pc * (pc + 1) / 2 : sym == ERI_KET ? na * nb * pc : full;
  const int64_t total = ngeom * per;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t e = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; e < total; e += nwave) {
    const int64_t g = e / per, r = e % per;
    int64_t i, j, k, l;
    if (sym == ERI_ALL) {
      int64_t bra, ket;
      tri_split(r, bra, ket);
      tri_split(bra, i, j);
      tri_split(ket, k, l);
    } else if (sym == ERI_KET) {
      tri_split(r % pc, k, l);
      j = (r / pc) % nb;
      i = r / (pc * nb);
    } else {
      l = r % nd; k = (r / nd) % nc; j = (r / (nd * nc)) % nb; i = r / (nd * nc * nb);
    }
    double2* o = out + g * full;
    // element (p, q, s, t) of the [na][nb][nc][nd] output
    auto at = [&](int64_t p, int64_t q, int64_t s2, int64_t t) { return ((p * nb + q) * nc + s2) * nd + t; };
    auto put = [&](double v) {
      const double2 w = make_double2(v, 0.0);
      o[at(i, j, k, l)] = w;
      if (sym == ERI_NONE) return;
      o[at(i, j, l, k)] = w;
      if (sym == ERI_KET) return;
      o[at(j, i, k, l)] = w;
      o[at(j, i, l, k)] = w;
      o[at(k, l, i, j)] = w;
      o[at(l, k, i, j)] = w;
      o[at(k, l, j, i)] = w;
      o[at(l, k, j, i)] = w;
    };
    const apy_gto& A = fa[i];
    const apy_gto& B = fb[j];
    const apy_gto& C = fc[k];
    const apy_gto& D = fd[l];
    if (!gto_ok(A, natom) || !gto_ok(B, natom) || !gto_ok(C, natom) || !gto_ok(D, natom) ||
        !quartet_ok(A, B, C, D)) {
      if (lane == 0) put(kNaN);
      continue;
    }
    const double* cen = centers + 3 * natom * g;
    const double* ca = cen + 3 * A.atom;
    const double* cb = cen + 3 * B.atom;
    const double* cc = cen + 3 * C.atom;
    const double* cd = cen + 3 * D.atom;
    const double qab[3] = {ca[0] - cb[0], ca[1] - cb[1], ca[2] - cb[2]};
    const double qcd[3] = {cc[0] - cd[0], cc[1] - cd[1], cc[2] - cd[2]};
    // quartet w = ((((((ta nbt + tb) nct + tc) ndt + td) npa + pa) npb + pb) npc + pc) npd + pd
    const int nt = A.nterm * B.nterm * C.nterm * D.nterm;
    const int np = A.nprim * B.nprim * C.nprim * D.nprim;
    double acc = 0.0;
    for (int w = lane; w < nt * np; w += 64) {
      int x = w;
      const int pd = x % D.nprim; x /= D.nprim;
      const int pc = x % C.nprim; x /= C.nprim;
      const int pb = x % B.nprim; x /= B.nprim;
      const int pa = x % A.nprim; x /= A.nprim;
      const int td = x % D.nterm; x /= D.nterm;
      const int tc = x % C.nterm; x /= C.nterm;
      const int tb = x % B.nterm; x /= B.nterm;
      const int ta = x;
      acc += eri_prim(A, B, C, D, ta, tb, tc, td, pa, pb, pc, pd, ca, cb, cc, cd, qab, qcd);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) put(acc);
  }
}

// out[m,n,l,s] = on(m) D[m,n,l,s] + on(n) D[n,m,l,s] + on(l) D[l,s,m,n] + on(s) D[s,l,m,n] with
// D = (d m/dx n | l s) and on(p) = 1 if function p sits on the displaced atom (the derivative of (mn|ls)
// when that atom moves, each function following its atom).
__global__ __launch_bounds__(256) void eri_deriv_assemble_kernel(int64_t n, const int* __restrict__ atom_of,
                                                                 int atom, const double2* __restrict__ D,
                                                                 double2* __restrict__ out) {
  const int64_t n2 = n * n, n3 = n2 * n, total = n3 * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e % n, l = (e / n) % n, q = (e / n2) % n, m = e / n3;
    double re = 0.0;
    if (atom_of[m] == atom) re += D[((m * n + q) * n + l) * n + s].x;
    if (atom_of[q] == atom) re += D[((q * n + m) * n + l) * n + s].x;
    if (atom_of[l] == atom) re += D[((l * n + s) * n + m) * n + q].x;
    if (atom_of[s] == atom) re += D[((s * n + l) * n + m) * n + q].x;
    out[e] = make_double2(re, 0.0);
  }
}


// ------------------------------------------------------------------------------ one-electron integrals
constexpr int OJ = LF + 3;  // ket powers j <= LF + 2 (the kinetic operator raises j by two)
constexpr int NOEI = 9;     // S, T, V, Dx, Dy, Dz, Lx, Ly, Lz

// 1-D overlaps S_ij = int x_A^i x_B^j exp(-a x_A^2 - b x_B^2) dx for i <= imax (<= LF), j <= jmax (< OJ),
// Obara-Saika: S_00 = sqrt(pi/p) exp(-ab/p Q^2) with Q = A - B, p = a + b, and
//   S_{i+1,j} = X_PA S_ij + (i S_{i-1,j} + j S_{i,j-1}) / 2p,   X_PA = -b Q / p,
//   S_{i,j+1} = X_PB S_ij + (i S_{i-1,j} + j S_{i,j-1}) / 2p,   X_PB =  a Q / p.
__device__ void ovl_1d(int imax, int jmax, double Q, double a, double b, double (*S)[OJ]) {
  const double p = a + b, h = 0.5 / p, xpa = -b * Q / p, xpb = a * Q / p;
  S[0][0] = sqrt(M_PI / p) * exp(-a * b / p * Q * Q);
  for (int j = 0; j < jmax; ++j) S[0][j + 1] = xpb * S[0][j] + (j > 0 ? j * h * S[0][j - 1] : 0.0);
  for (int i = 0; i < imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      double v = xpa * S[i][j];
      if (i > 0) v += i * h * S[i - 1][j];
      if (j > 0) v += j * h * S[i][j - 1];
      S[i + 1][j] = v;
    }
}

// The integrals `which` selects between bra function A at ra and ket function B at rb, accumulated into
// v[NOEI].  Per axis, from the 1-D overlaps S_ij (i = bra power, j = ket power):
//   kinetic    t = -1/2 (j(j-1) S_{i,j-2} - 2b(2j+1) S_ij + 4b^2 S_{i,j+2}),
//   moment     m = S_{i,j+1} + (B - O) S_ij             (x - O = x_B + B - O),
//   derivative d = j S_{i,j-1} - 2b S_{i,j+1}           (d/dx acting on the ket);
//   T = t_x s_y s_z + s_x t_y s_z + s_x s_y t_z,  D_x = m_x s_y s_z,  L_x = (m_y d_z - m_z d_y) s_x (cyclic).
// The nuclear attraction uses the Hermite expansion of the product:
//   V = -sum_C Z_C 2 pi / p sum_tuv E^x_t E^y_u E^z_v R_tuv(p, P - C).
// The (term, primitive) pairs w = ((ta nbt + tb) npa + pa) npb + pb from `first` in steps of `step` (the
// lanes of a wavefront split them).
__device__ void oei_pair(const apy_gto& A, const double* ra, const apy_gto& B, const double* rb, int which,
                         const double* origin, const double* nuc, const double* Z, int64_t nnuc, int first, int step,
                         double* v) {
  const double Q[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
  const int nw = A.nterm * B.nterm * A.nprim * B.nprim;
  for (int w = first; w < nw; w += step) {
    const int pb = w % B.nprim, pa = (w / B.nprim) % A.nprim, tb = (w / (B.nprim * A.nprim)) % B.nterm,
              ta = w / (B.nprim * A.nprim * B.nterm);
    const int* la = A.lmn[ta];
    const int* lb = B.lmn[tb];
    const double a = A.exps[pa], b = B.exps[pb], p = a + b;
    const double c = A.coefs[ta][pa] * B.coefs[tb][pb];
    double S[3][LF + 1][OJ], s[3], t[3], m[3], d[3];
    for (int x = 0; x < 3; ++x) {
      const int i = la[x], j = lb[x];
      ovl_1d(i, j + 2, Q[x], a, b, S[x]);
      s[x] = S[x][i][j];
      t[x] = -0.5 * ((j > 1 ? j * (j - 1) * S[x][i][j - 2] : 0.0) - 2.0 * b * (2 * j + 1) * S[x][i][j] +
                     4.0 * b * b * S[x][i][j + 2]);
      m[x] = S[x][i][j + 1] + (rb[x] - origin[x]) * S[x][i][j];
      d[x] = (j > 0 ? j * S[x][i][j - 1] : 0.0) - 2.0 * b * S[x][i][j + 1];
    }
    if (which & APY_OEI_OVERLAP) v[0] += c * s[0] * s[1] * s[2];
    if (which & APY_OEI_KINETIC) v[1] += c * (t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2]);
    if (which & APY_OEI_DIPOLE) {
      v[3] += c * m[0] * s[1] * s[2];
      v[4] += c * s[0] * m[1] * s[2];
      v[5] += c * s[0] * s[1] * m[2];
    }
    if (which & APY_OEI_ANGMOM) {
      v[6] += c * (m[1] * d[2] - m[2] * d[1]) * s[0];
      v[7] += c * (m[2] * d[0] - m[0] * d[2]) * s[1];
      v[8] += c * (m[0] * d[1] - m[1] * d[0]) * s[2];
    }
    if (which & APY_OEI_POTENTIAL) {
      double E[3][NE], P[3];
      int Lt[3];
      for (int x = 0; x < 3; ++x) {
        herm_E(la[x], lb[x], Q[x], a, b, E[x]);
        P[x] = (a * ra[x] + b * rb[x]) / p;
        Lt[x] = la[x] + lb[x];
      }
      const int L = Lt[0] + Lt[1] + Lt[2];
      const int sz = Lt[2] + 1, sy = (Lt[1] + 1) * sz;
      double acc = 0.0;
      for (int64_t C = 0; C < nnuc; ++C) {
        const double PC[3] = {P[0] - nuc[3 * C], P[1] - nuc[3 * C + 1], P[2] - nuc[3 * C + 2]};
        double F[LQ + 1], Rb[2][NRB];
        boys_fn(L, p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), F);
        const double* R = Rb[herm_R(L, Lt, p, PC, F, Rb)];
        double sum = 0.0;
        for (int tt = 0; tt <= Lt[0]; ++tt)
          for (int u = 0; u <= Lt[1]; ++u)
            for (int vv = 0; vv <= Lt[2]; ++vv) sum += E[0][tt] * E[1][u] * E[2][vv] * R[tt * sy + u * sz + vv];
        acc -= Z[C] * sum;
      }
      v[2] += c * 2.0 * M_PI / p * acc;
    }
  }
}

// One wavefront per output element (wave-uniform control flow): the lanes split the (term, primitive)
// pairs and a fixed-order butterfly sums each component.
__global__ __launch_bounds__(256) void oei_kernel(apy_oei_args p) {
  const int lane = threadIdx.x & 63;
  const int64_t per = p.na * p.nb, total = p.nset * per;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t e = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; e < total; e += nwave) {
    const int64_t s = e / per, r = e % per, i = r / p.nb, j = r % p.nb;
    const apy_gto& A = p.fa[i];
    const apy_gto& B = p.fb[j];
    double v[NOEI];
    for (int k = 0; k < NOEI; ++k) v[k] = 0.0;
    if (gto_ok(A, p.nca) && gto_ok(B, p.ncb)) {
      oei_pair(A, p.ca + 3 * (s * p.nca + A.atom), B, p.cb + 3 * (s * p.ncb + B.atom), p.which, p.origin,
               p.nuclei ? p.nuclei + 3 * s * p.nnuc : nullptr, p.charges, p.nnuc, lane, 64, v);
#pragma unroll
      for (int k = 0; k < NOEI; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    } else {
      for (int k = 0; k < NOEI; ++k) v[k] = kNaN;
    }
    if (lane != 0) continue;
    // selected components in the order S, T, V, Dx, Dy, Dz, Lx, Ly, Lz
    constexpr int bit_of[NOEI] = {APY_OEI_OVERLAP, APY_OEI_KINETIC, APY_OEI_POTENTIAL, APY_OEI_DIPOLE, APY_OEI_DIPOLE,
                                  APY_OEI_DIPOLE,  APY_OEI_ANGMOM,  APY_OEI_ANGMOM,    APY_OEI_ANGMOM};
    int64_t c = 0;
    const int64_t ncomp = apy_oei_ncomp(p.which);
#pragma unroll
    for (int k = 0; k < NOEI; ++k)
      if (p.which & bit_of[k]) p.out[((s * ncomp + c++) * p.na + i) * p.nb + j] = v[k];
  }
}
// Fill g_boys on the current device before its first integral launch; the stream is synchronised once so
// that launches on other streams see the table.
int ensure_boys_table(hipStream_t st) {
  static std::mutex mu;
  static bool ready[256] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 256) return APY_E_HIP;
  std::lock_guard<std::mutex> lk(mu);
  if (ready[dev]) return APY_OK;
  boys_table_kernel<<<(BOYS_NGRID + 255) / 256, 256, 0, st>>>();
  APY_LAUNCH_CHECK();
  if (hipStreamSynchronize(st) != hipSuccess) return APY_E_HIP;
  ready[dev] = true;
  return APY_OK;
}

}  // namespace

extern "C" {

int apy_eri_gto(int64_t ngeom, int64_t natom, const double* centers, const apy_gto* fa, int64_t na,
                const apy_gto* fb, int64_t nb, const apy_gto* fc, int64_t nc, const apy_gto* fd, int64_t nd,
                void* out, void* stream) {
  if (ngeom < 1 || natom < 1 || !centers || !fa || !fb || !fc || !fd || !out || na < 1 || nb < 1 || nc < 1 ||
      nd < 1)
    return APY_E_ARG;
  // the same table in every slot: 8-fold permutational symmetry; the same ket tables: (ab|cd) = (ab|dc)
  const int sym = (fa == fb && fb == fc && fc == fd && na == nb && nb == nc && nc == nd) ? ERI_ALL
                  : (fc == fd && nc == nd)                                             ? ERI_KET
                                                                                       : ERI_NONE;
  const int64_t pc = nc * (nc + 1) / 2;
  const int64_t per = sym == ERI_ALL ? pc * (pc + 1) / 2 : sym == ERI_KET ? na * nb * pc : na * nb * nc * nd;
  const int64_t total = ngeom * per;
  if (const int rc = ensure_boys_table((hipStream_t)stream); rc != APY_OK) return rc;
  eri_kernel<<<apy_grid1d(64 * total), 256, 0, (hipStream_t)stream>>>(ngeom, natom, centers, fa, na, fb, nb, fc, nc, fd,
                                                                  nd, sym, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_eri_deriv_assemble(int64_t n, const int* atom_of, int atom, const void* D, void* out, void* stream) {
  if (n < 1 || !atom_of || !D || !out || D == out) return APY_E_ARG;
  const int64_t total = n * n * n * n;
  eri_deriv_assemble_kernel<<<apy_grid1d(total), 256, 0, (hipStream_t)stream>>>(n, atom_of, atom, (const double2*)D,
                                                                                 (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_oei_components(int which) { return apy_oei_ncomp(which); }

int apy_oei_gto(const apy_oei_args* a, void* stream) {
  if (!a || apy_oei_ncomp(a->which) == 0 || a->nset < 1 || !a->fa || !a->fb || a->na < 1 || a->nb < 1 || !a->ca ||
      !a->cb || a->nca < 1 || a->ncb < 1 || !a->out)
    return APY_E_ARG;
  if ((a->which & APY_OEI_POTENTIAL) && (!a->nuclei || !a->charges || a->nnuc < 1)) return APY_E_ARG;
  const int64_t total = a->nset * a->na * a->nb;
  if ((a->which & APY_OEI_POTENTIAL) != 0)
    if (const int rc = ensure_boys_table((hipStream_t)stream); rc != APY_OK) return rc;
  oei_kernel<<<apy_grid1d(64 * total), 256, 0, (hipStream_t)stream>>>(*a);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
