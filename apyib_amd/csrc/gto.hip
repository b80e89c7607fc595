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

namespace {

constexpr int LF = 2;                     // largest angular momentum of one term along all axes together
constexpr int LQ = 5;                     // largest total angular momentum of a quartet
constexpr int NE = 2 * LF + 1;            // E_t^{ij}, t <= i + j
constexpr int NRB = 18;                   // (Lx+1)(Ly+1)(Lz+1) <= 18 for Lx+Ly+Lz <= 5

// F_n(T), n = 0..nmax (nmax <= LQ): the series e^-T sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)) for the top
// order and downward recursion F_n = (2T F_{n+1} + e^-T) / (2n+1) below T = 30; above, F_0 from erf and
// upward recursion F_{n+1} = ((2n+1) F_n - e^-T) / (2T) (no cancellation there for n <= 5).
__device__ void boys_fn(int nmax, double T, double* F) {
  if (T < 30.0) {
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    const double e = exp(-T);
    F[nmax] = e * sum;
    for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + e) / (2 * n + 1);
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

// Contracted integral (a b | c d) at one geometry.  Returns NaN for a table entry outside the limits
// (angular momentum, primitive count, term count, atom index), so a bad table cannot index out of bounds.
__device__ double eri_quartet(const apy_gto& A, const apy_gto& B, const apy_gto& C, const apy_gto& D,
                              const double* cen, int64_t natom) {
  const apy_gto* f[4] = {&A, &B, &C, &D};
  for (int k = 0; k < 4; ++k) {
    const apy_gto& g = *f[k];
    if (g.atom < 0 || g.atom >= natom || g.nprim < 1 || g.nprim > APY_GTO_MAXPRIM || g.nterm < 1 || g.nterm > 2)
      return __longlong_as_double(0x7ff8000000000000LL);
    for (int t = 0; t < g.nterm; ++t)
      if (g.lmn[t][0] < 0 || g.lmn[t][1] < 0 || g.lmn[t][2] < 0 || lsum(g.lmn[t]) > LF)
        return __longlong_as_double(0x7ff8000000000000LL);
  }
  const double* ca = cen + 3 * A.atom;
  const double* cb = cen + 3 * B.atom;
  const double* cc = cen + 3 * C.atom;
  const double* cd = cen + 3 * D.atom;
  const double qab[3] = {ca[0] - cb[0], ca[1] - cb[1], ca[2] - cb[2]};
  const double qcd[3] = {cc[0] - cd[0], cc[1] - cd[1], cc[2] - cd[2]};
  const double pref0 = 2.0 * M_PI * M_PI * sqrt(M_PI);
  double total = 0.0;
  for (int ta = 0; ta < A.nterm; ++ta)
    for (int tb = 0; tb < B.nterm; ++tb)
      for (int tc = 0; tc < C.nterm; ++tc)
        for (int td = 0; td < D.nterm; ++td) {
          const int* la = A.lmn[ta];
          const int* lb = B.lmn[tb];
          const int* lc = C.lmn[tc];
          const int* ld = D.lmn[td];
          int L1[3], L2[3], Lt[3];
          for (int x = 0; x < 3; ++x) { L1[x] = la[x] + lb[x]; L2[x] = lc[x] + ld[x]; Lt[x] = L1[x] + L2[x]; }
          const int L = Lt[0] + Lt[1] + Lt[2];
          if (L > LQ) return __longlong_as_double(0x7ff8000000000000LL);
          const int sz = Lt[2] + 1, sy = (Lt[1] + 1) * sz;       // R index (t, u, v) -> t sy + u sz + v
          for (int pa = 0; pa < A.nprim; ++pa)
            for (int pb = 0; pb < B.nprim; ++pb) {
              const double a = A.exps[pa], b = B.exps[pb], p = a + b;
              const double cab = A.coefs[ta][pa] * B.coefs[tb][pb];
              double Eab[3][NE];
              for (int x = 0; x < 3; ++x) herm_E(la[x], lb[x], qab[x], a, b, Eab[x]);
              double P[3];
              for (int x = 0; x < 3; ++x) P[x] = (a * ca[x] + b * cb[x]) / p;
              for (int pc = 0; pc < C.nprim; ++pc)
                for (int pd = 0; pd < D.nprim; ++pd) {
                  const double c = C.exps[pc], d = D.exps[pd], q = c + d;
                  const double ccd = C.coefs[tc][pc] * D.coefs[td][pd];
                  double Ecd[3][NE];
                  for (int x = 0; x < 3; ++x) herm_E(lc[x], ld[x], qcd[x], c, d, Ecd[x]);
                  double PQ[3];
                  for (int x = 0; x < 3; ++x) PQ[x] = P[x] - (c * cc[x] + d * cd[x]) / q;
                  const double alpha = p * q / (p + q);
                  double F[LQ + 1];
                  boys_fn(L, alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]), F);
                  // Hermite Coulomb integrals: level n holds R^(n)_tuv for t + u + v <= L - n
                  double Rb[2][NRB];
                  int cur = 0;
                  for (int n = L; n >= 0; --n) {
                    const int nxt = cur ^ 1;                          // level n + 1 lives in Rb[cur]
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
                  const double* R = Rb[cur];
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
                  total += cab * ccd * pref0 / (p * q * sqrt(p + q)) * val;
                }
            }
        }
  return total;
}

__global__ __launch_bounds__(256) void eri_kernel(int64_t ngeom, int64_t natom, const double* __restrict__ centers,
                                                  const apy_gto* __restrict__ fa, int64_t na,
                                                  const apy_gto* __restrict__ fb, int64_t nb,
                                                  const apy_gto* __restrict__ fc, int64_t nc,
                                                  const apy_gto* __restrict__ fd, int64_t nd,
                                                  double2* __restrict__ out) {
  const int64_t per = na * nb * nc * nd, total = ngeom * per;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e / per, r = e % per;
    const int64_t l = r % nd, k = (r / nd) % nc, j = (r / (nd * nc)) % nb, i = r / (nd * nc * nb);
    out[e] = make_double2(eri_quartet(fa[i], fb[j], fc[k], fd[l], centers + 3 * natom * g, natom), 0.0);
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

}  // namespace

extern "C" {

int apy_eri_gto(int64_t ngeom, int64_t natom, const double* centers, const apy_gto* fa, int64_t na,
                const apy_gto* fb, int64_t nb, const apy_gto* fc, int64_t nc, const apy_gto* fd, int64_t nd,
                void* out, void* stream) {
  if (ngeom < 1 || natom < 1 || !centers || !fa || !fb || !fc || !fd || !out || na < 1 || nb < 1 || nc < 1 ||
      nd < 1)
    return APY_E_ARG;
  const int64_t total = ngeom * na * nb * nc * nd;
  eri_kernel<<<apy_grid1d(total), 256, 0, (hipStream_t)stream>>>(ngeom, natom, centers, fa, na, fb, nb, fc, nc, fd,
                                                                  nd, (double2*)out);
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

}  // extern "C"
