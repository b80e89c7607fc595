// Davidson preconditioner and the one-particle density (the energy
// gradient with respect to F) of a spin-orbital CISD vector.
//
// Vector layout (length L + 1, L = O V + O^2 V^2): [c1 (O x V) | c2 dense
// antisymmetric (O x O x V x V) | c0].  The first L entries are the layout
// apy_cisd_so_step reads and writes.
//
// Density of a normalised vector (ci_wfn.py:380-394, the reference's
// adjoint sketch; d E = sum_pq D[p,q] d F[p,q] for the variational CISD):
//   D_oo[i,j] = -sum_a conj(c1[j,a]) c1[i,a] - 1/2 sum_kab conj(c2[j,k,a,b]) c2[i,k,a,b]
//   D_vv[a,b] =  sum_i conj(c1[i,a]) c1[i,b] + 1/2 sum_ijc conj(c2[i,j,a,c]) c2[i,j,b,c]
//   D_ov[i,a] =  conj(c0) c1[i,a] + sum_jb conj(c1[j,b]) c2[i,j,a,b]
//   D_vo[a,i] =  conj(c1[i,a]) c0 + sum_jb conj(c2[i,j,a,b]) c1[j,b]
// D_oo and D_vv are GEMMs on the matrix cores; the ov/vo blocks (O^2 V^2
// work each) are one elementwise-reduction kernel.
//
// Two-particle density (the gradient with respect to the antisymmetrised
// integrals, the reference's TPD sketch at ci_wfn.py:396-409 restated for
// <pq||rs> as the sigma equations read it), in the packed block layouts of
// apy_so_ints; each stored entry carries the sum over the full-tensor
// positions it stands for:
//   vvvv_p[(a<b),(c<d)] = sum_{i<j} conj(c2[i,j,a,b]) c2[i,j,c,d]            GEMM Pv x Pv x Po
//   oooo_p[(i<j),(k<l)] = sum_{a<b} conj(c2[i,j,a,b]) c2[k,l,a,b]            GEMM Po x Po x Pv
//   ring[(j,b),(k,c)]   = conj(c1[j,b]) c1[k,c]
//                         + sum_ia conj(c2[i,j,a,b]) c2[i,k,a,c]               GEMM OV x OV x (OV+1)
//   vovv_p[a,j,(c<b)]   = sum_i conj(c1[i,a]) c2[i,j,c,b]                     GEMM V x O Pv x O
//   ooov[k,j,i,b]       = -1/2 sum_a conj(c1[i,a]) c2[k,j,a,b]
//   ovoo_p[k,b,(i<j)]   = -sum_a conj(c2[i,j,a,b]) c1[k,a]
//   vvvo_p[(a<b),j,c]   = sum_i conj(c2[i,j,a,b]) c1[i,c]
//   vvoo_t[i,j,a,b]     = 1/4 c0 conj(c2[i,j,a,b]),   oovv[i,j,a,b] = 1/4 conj(c0) c2[i,j,a,b]
// so that dE = sum over blocks of G_blk . d(block) for every perturbation of
// <pq||rs> antisymmetric in (p,q) and (r,s).
#include "common.h"

#include <algorithm>

namespace {

__global__ __launch_bounds__(256) void precond_kernel(int64_t O, int64_t V, const double* __restrict__ eps,
                                                      double theta, const double2* __restrict__ r,
                                                      double2* __restrict__ out) {
  const int64_t OV = O * V, L = OV + OV * OV;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= L; e += (int64_t)gridDim.x * blockDim.x) {
    double D;
    if (e < OV) {
      D = eps[O + e % V] - eps[e / V];
    } else if (e < L) {
      int64_t q = e - OV;
      const int64_t b = q % V; q /= V;
      const int64_t a = q % V; q /= V;
      const int64_t j = q % O;
      const int64_t i = q / O;
      if (i == j || a == b) { out[e] = make_double2(0.0, 0.0); continue; }
      D = eps[O + a] + eps[O + b] - eps[i] - eps[j];
    } else {
      D = 0.0;
    }
    double den = theta - D;
    if (fabs(den) < 1e-8) den = den < 0.0 ? -1e-8 : 1e-8;
    out[e] = make_double2(r[e].x / den, r[e].y / den);
  }
}

// D_ov and D_vo: one thread per (i, a), sums over (j, b).
__global__ __launch_bounds__(256) void opd_ov_kernel(int64_t O, int64_t V, const double2* __restrict__ c,
                                                     double2* __restrict__ D) {
  const int64_t OV = O * V, L = OV + OV * OV, n = O + V;
  const double2* c1 = c;
  const double2* c2 = c + OV;
  const double2 c0 = c[L];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < OV; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / V, a = e % V;
    const double2 x = c1[e];
    // conj(c0) c1[i,a] and conj(c1[i,a]) c0
    double ovr = c0.x * x.x + c0.y * x.y, ovi = c0.x * x.y - c0.y * x.x;
    double vor = x.x * c0.x + x.y * c0.y, voi = x.x * c0.y - x.y * c0.x;
    for (int64_t j = 0; j < O; ++j) {
      const double2* row = c2 + ((i * O + j) * V + a) * V;
      for (int64_t b = 0; b < V; ++b) {
        const double2 y = c1[j * V + b], t = row[b];
        ovr += y.x * t.x + y.y * t.y;  // conj(c1[j,b]) c2[i,j,a,b]
        ovi += y.x * t.y - y.y * t.x;
        vor += t.x * y.x + t.y * y.y;  // conj(c2[i,j,a,b]) c1[j,b]
        voi += t.x * y.y - t.y * y.x;
      }
    }
    D[i * n + O + a] = make_double2(ovr, ovi);
    D[(O + a) * n + i] = make_double2(vor, voi);
  }
}

apy_gemm_desc desc(int64_t m, int64_t nn, int64_t k, int al, const void* a, int64_t lda, int bl, const void* b,
                   int64_t ldb, void* c, int64_t ldc, double alpha, double beta, int conj_a, int conj_b) {
  apy_gemm_desc d{};
  d.m = m; d.n = nn; d.k = k;
  d.a_layout = al; d.b_layout = bl;
  d.conj_a = conj_a; d.conj_b = conj_b;
  d.alpha_re = alpha; d.beta_re = beta;
  d.a = a; d.lda = lda; d.b = b; d.ldb = ldb; d.c = c; d.ldc = ldc;
  d.batch = 1;
  return d;
}

// The four GEMMs of D_oo and D_vv (pointers null when only sizing).
void opd_gemms(int64_t O, int64_t V, const double2* c, double2* D, apy_gemm_desc g[4]) {
  const int64_t OV = O * V, n = O + V;
  const double2* c1 = c;
  const double2* c2 = c ? c + OV : nullptr;
  double2* Doo = D;
  double2* Dvv = D ? D + O * n + O : nullptr;
  // D_oo[i,j] = -c1[i,:] . conj(c1[j,:])
  g[0] = desc(O, O, V, APY_MK, c1, V, APY_NK, c1, V, Doo, n, -1.0, 0.0, 0, 1);
  // D_oo[i,j] -= 1/2 c2[i,(kab)] . conj(c2[j,(kab)])
  g[1] = desc(O, O, O * V * V, APY_MK, c2, O * V * V, APY_NK, c2, O * V * V, Doo, n, -0.5, 1.0, 0, 1);
  // D_vv[a,b] = conj(c1[:,a]) . c1[:,b]
  g[2] = desc(V, V, O, APY_KM, c1, V, APY_KN, c1, V, Dvv, n, 1.0, 0.0, 1, 0);
  // D_vv[a,b] += 1/2 sum_(ij) conj(c2[(ij),a,:]) . c2[(ij),b,:]
  g[3] = desc(V, V, V, APY_MK, c2, V, APY_NK, c2, V, Dvv, n, 0.5, 1.0, 1, 0);
  g[3].batch = O * O; g[3].stride_a = V * V; g[3].stride_b = V * V; g[3].stride_c = 0; g[3].reduce_batch = 1;
}

__host__ __device__ inline int64_t tp_npair(int64_t N) { return N * (N - 1) / 2; }
// packed index of x < y among N
__host__ __device__ inline int64_t tp_pidx(int64_t x, int64_t y, int64_t N) { return x * (2 * N - x - 1) / 2 + y - x - 1; }

__device__ inline void tp_pdec(int64_t P, int64_t N, int64_t& x, int64_t& y) {
  // row x holds pairs [x*(2N-x-1)/2, (x+1)*(2N-x-2)/2)
  double b = 2.0 * (double)N - 1.0;
  double disc = b * b - 8.0 * (double)P;
  int64_t r = (int64_t)((b - sqrt(disc > 0.0 ? disc : 0.0)) * 0.5);
  if (r < 0) r = 0;
  if (r > N - 2) r = N - 2;
  while (r > 0 && tp_pidx(r, r + 1, N) > P) --r;
  while (r + 1 <= N - 2 && tp_pidx(r + 1, r + 2, N) <= P) ++r;
  x = r;
  y = P - tp_pidx(r, r + 1, N) + r + 1;
}

__device__ inline double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Cp[(i<j)][(a<b)] = c2[i,j,a,b]; Xr[(i,a)][(j,b)] = c2[i,j,a,b], Xr[OV][(j,b)] = c1[j,b]
__global__ __launch_bounds__(256) void tpd_pack_kernel(int64_t O, int64_t V, const double2* __restrict__ c,
                                                       double2* __restrict__ Cp, double2* __restrict__ Xr) {
  const int64_t OV = O * V, Po = tp_npair(O), Pv = tp_npair(V);
  const double2* c2 = c + OV;
  const int64_t nx = (OV + 1) * OV, total = nx + Po * Pv;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    if (t < nx) {
      const int64_t row = t / OV, col = t % OV;
      if (row == OV) {
        Xr[t] = c[col];
      } else {
        const int64_t i = row / V, a = row % V, j = col / V, b = col % V;
        Xr[t] = c2[((i * O + j) * V + a) * V + b];
      }
    } else {
      const int64_t u = t - nx;
      int64_t i, j, a, b;
      tp_pdec(u / Pv, O, i, j);
      tp_pdec(u % Pv, V, a, b);
      Cp[u] = c2[((i * O + j) * V + a) * V + b];
    }
  }
}

// The O^3 V^2 / O^2 V^3 blocks with K = O or V: ooov, ovoo_p, vvvo_p, and
// the elementwise vvoo_t / oovv.  One thread per output entry.
__global__ __launch_bounds__(256) void tpd_small_kernel(int64_t O, int64_t V, const double2* __restrict__ c,
                                                        double2* __restrict__ ooov, double2* __restrict__ ovoo_p,
                                                        double2* __restrict__ vvvo_p, double2* __restrict__ vvoo_t,
                                                        double2* __restrict__ oovv) {
  const int64_t OV = O * V, L = OV + OV * OV, Po = tp_npair(O), Pv = tp_npair(V);
  const double2* c1 = c;
  const double2* c2 = c + OV;
  const double2 c0 = c[L];
  const int64_t n1 = ooov ? O * O * O * V : 0;
  const int64_t n2 = ovoo_p ? O * V * Po : 0;
  const int64_t n3 = vvvo_p ? Pv * O * V : 0;
  const int64_t n4 = (vvoo_t || oovv) ? OV * OV : 0;
  const int64_t total = n1 + n2 + n3 + n4;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t u = t;
    if (u < n1) {  // ooov[k][j][i][b], b fastest
      const int64_t b = u % V, i = (u / V) % O, j = (u / OV) % O, k = u / (OV * O);
      double sr = 0.0, si = 0.0;
      const double2* row = c2 + (k * O + j) * V * V + b;
      for (int64_t a = 0; a < V; ++a) {
        const double2 z = cmulc(c1[i * V + a], row[a * V]);
        sr += z.x; si += z.y;
      }
      ooov[u] = make_double2(-0.5 * sr, -0.5 * si);
      continue;
    }
    u -= n1;
    if (u < n2) {  // thread order (k, P, b) with b fastest; stored at [k][b][P]
      const int64_t b = u % V, P = (u / V) % Po, k = u / (V * Po);
      int64_t i, j;
      tp_pdec(P, O, i, j);
      const double2* row = c2 + (i * O + j) * V * V + b;
      double sr = 0.0, si = 0.0;
      for (int64_t a = 0; a < V; ++a) {
        const double2 z = cmulc(row[a * V], c1[k * V + a]);
        sr += z.x; si += z.y;
      }
      ovoo_p[(k * V + b) * Po + P] = make_double2(-sr, -si);
      continue;
    }
    u -= n2;
    if (u < n3) {  // vvvo_p[P][j][c], c fastest
      const int64_t cc = u % V, j = (u / V) % O, P = u / OV;
      int64_t a, b;
      tp_pdec(P, V, a, b);
      double sr = 0.0, si = 0.0;
      for (int64_t i = 0; i < O; ++i) {
        const double2 z = cmulc(c2[((i * O + j) * V + a) * V + b], c1[i * V + cc]);
        sr += z.x; si += z.y;
      }
      vvvo_p[u] = make_double2(sr, si);
      continue;
    }
    u -= n3;
    {
      const double2 x = c2[u];
      if (vvoo_t) vvoo_t[u] = make_double2(0.25 * (c0.x * x.x + c0.y * x.y), 0.25 * (c0.y * x.x - c0.x * x.y));
      if (oovv) oovv[u] = make_double2(0.25 * (c0.x * x.x + c0.y * x.y), 0.25 * (c0.x * x.y - c0.y * x.x));
    }
  }
}

// Th[i][j][(c<b)] = c2[i][j][c][b] (the dense c2 packed over its virtual pair).
__global__ __launch_bounds__(256) void tpd_half_pack_kernel(int64_t O, int64_t V, const double2* __restrict__ c2,
                                                            double2* __restrict__ th) {
  const int64_t Pv = tp_npair(V), total = O * O * Pv;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t P = t % Pv, ij = t / Pv;
    int64_t c, b;
    tp_pdec(P, V, c, b);
    th[t] = c2[(ij * V + c) * V + b];
  }
}

// The four GEMM blocks (pointers null when only sizing).
void tpd_gemms(int64_t O, int64_t V, const double2* c, const double2* Cp, const double2* Xr, const double2* Th,
               const apy_so_ints* g, apy_gemm_desc d[4], int* used) {
  const int64_t OV = O * V, Po = tp_npair(O), Pv = tp_npair(V);
  const double2* c1 = c;
  const double2* c2 = c ? c + OV : nullptr;
  int q = 0;
  if (!g || g->vvvv_p)
    d[q++] = desc(Pv, Pv, Po, APY_KM, Cp, Pv, APY_KN, Cp, Pv, g ? g->vvvv_p : nullptr, Pv, 1.0, 0.0, 1, 0);
  if (!g || g->oooo_p)
    d[q++] = desc(Po, Po, Pv, APY_MK, Cp, Pv, APY_NK, Cp, Pv, g ? g->oooo_p : nullptr, Po, 1.0, 0.0, 1, 0);
  if (!g || g->ring)
    d[q++] = desc(OV, OV, OV + 1, APY_KM, Xr, OV, APY_KN, Xr, OV, g ? g->ring : nullptr, OV, 1.0, 0.0, 1, 0);
  if (!g || g->vovv_p)
    d[q++] = desc(V, O * Pv, O, APY_KM, c1, V, APY_KN, Th, O * Pv, g ? g->vovv_p : nullptr, O * Pv, 1.0, 0.0, 1, 0);
  *used = q;
}

struct TpdWS {
  size_t cp, xr, th, gemm, total;
};

TpdWS tpd_carve(int64_t O, int64_t V) {
  const int64_t OV = O * V, Po = tp_npair(O), Pv = tp_npair(V);
  TpdWS w{};
  w.cp = 0;
  w.xr = ((size_t)(Po * Pv) * 16 + 255) / 256 * 256;
  w.th = w.xr + ((size_t)((OV + 1) * OV) * 16 + 255) / 256 * 256;
  w.gemm = w.th + ((size_t)(O * O * Pv) * 16 + 255) / 256 * 256;
  apy_gemm_desc d[4];
  int used = 0;
  tpd_gemms(O, V, nullptr, nullptr, nullptr, nullptr, nullptr, d, &used);
  size_t m = 0;
  for (int q = 0; q < used; ++q) m = std::max(m, apy_zgemm_ws_bytes(&d[q]));
  w.total = w.gemm + m;
  return w;
}

// Whole-tensor gradient G[p][q][r][s] from the block gradients: each
// full-tensor position takes its share (1/4, 1/2 or 1) of the stored entry
// it maps to, with the antisymmetry sign; p == q or r == s gives 0.
__global__ __launch_bounds__(256) void tpd_full_kernel(int64_t O, int64_t V, apy_so_ints g, double2* __restrict__ G) {
  const int64_t n = O + V, Po = tp_npair(O), Pv = tp_npair(V);
  const int64_t total = n * n * n * n;
  const double2* vvvv_p = (const double2*)g.vvvv_p;
  const double2* ring = (const double2*)g.ring;
  const double2* oooo_p = (const double2*)g.oooo_p;
  const double2* ovoo_p = (const double2*)g.ovoo_p;
  const double2* vvvo_p = (const double2*)g.vvvo_p;
  const double2* vovv_p = (const double2*)g.vovv_p;
  const double2* ooov = (const double2*)g.ooov;
  const double2* vvoo_t = (const double2*)g.vvoo_t;
  const double2* oovv = (const double2*)g.oovv;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t % n, r = (t / n) % n, q = (t / (n * n)) % n, p = t / (n * n * n);
    if (p == q || r == s) { G[t] = make_double2(0.0, 0.0); continue; }
    const int code = ((p >= O) << 3) | ((q >= O) << 2) | ((r >= O) << 1) | (int)(s >= O);
    double f = 1.0;
    double2 v = make_double2(0.0, 0.0);
    auto srt = [&](int64_t& x, int64_t& y) { if (x > y) { int64_t z = x; x = y; y = z; f = -f; } };
    switch (code) {
      case 0x0: {  // <kl||ij> -> oooo_p[(ij)][(kl)]
        int64_t k = p, l = q, i = r, j = s;
        srt(k, l); srt(i, j);
        v = oooo_p[tp_pidx(i, j, O) * Po + tp_pidx(k, l, O)]; f *= 0.25; break;
      }
      case 0xF: {
        int64_t a = p - O, b = q - O, c = r - O, d = s - O;
        srt(a, b); srt(c, d);
        v = vvvv_p[tp_pidx(a, b, V) * Pv + tp_pidx(c, d, V)]; f *= 0.25; break;
      }
      case 0x6: v = ring[((s * V + (q - O)) * O + p) * V + (r - O)]; f = 0.25; break;    // <kb||cj>
      case 0xA: v = ring[((s * V + (p - O)) * O + q) * V + (r - O)]; f = -0.25; break;   // <bk||cj>
      case 0x5: v = ring[((r * V + (q - O)) * O + p) * V + (s - O)]; f = -0.25; break;   // <kb||jc>
      case 0x9: v = ring[((r * V + (p - O)) * O + q) * V + (s - O)]; f = 0.25; break;    // <bk||jc>
      case 0x4: case 0x8: {  // <kb||ij>, <bk||ij>
        const int64_t k = code == 0x4 ? p : q, b = (code == 0x4 ? q : p) - O;
        int64_t i = r, j = s;
        f = code == 0x4 ? 0.25 : -0.25;
        srt(i, j);
        v = ovoo_p[(k * V + b) * Po + tp_pidx(i, j, O)]; break;
      }
      case 0xE: case 0xD: {  // <ab||cj>, <ab||jc>
        int64_t a = p - O, b = q - O;
        const int64_t c = (code == 0xE ? r : s) - O, j = code == 0xE ? s : r;
        f = code == 0xE ? 0.25 : -0.25;
        srt(a, b);
        v = vvvo_p[(tp_pidx(a, b, V) * O + j) * V + c]; break;
      }
      case 0xB: case 0x7: {  // <aj||cb>, <ja||cb> -> vovv_p[a][j][(c<b)]
        const int64_t a = (code == 0xB ? p : q) - O, j = code == 0xB ? q : p;
        int64_t c = r - O, b = s - O;
        f = code == 0xB ? 0.25 : -0.25;
        srt(c, b);
        v = vovv_p[(a * O + j) * Pv + tp_pidx(c, b, V)]; break;
      }
      case 0x1: v = ooov[((p * O + q) * O + r) * V + (s - O)]; f = 0.5; break;                  // <kj||ib>
      case 0x2: v = ooov[((p * O + q) * O + s) * V + (r - O)]; f = -0.5; break;                 // <kj||bi>
      case 0xC: v = vvoo_t[((r * O + s) * V + (p - O)) * V + (q - O)]; break;                     // <ab||ij>
      default: v = oovv[((p * O + q) * V + (r - O)) * V + (s - O)]; break;                        // 0x3 <ij||ab>
    }
    G[t] = make_double2(f * v.x, f * v.y);
  }
}

// Packed CI vectors [c1 (OV) | c2 over (i<j, a<b) (Po Pv) | c0]: the antisymmetric doubles stored once,
// so the CI metric (dense c2 weighted 1/4) is the Euclidean inner product.
__global__ __launch_bounds__(256) void ci_pack_kernel(int64_t O, int64_t V, const double2* __restrict__ dense,
                                                      double2* __restrict__ packed) {
  const int64_t OV = O * V, Po = tp_npair(O), Pv = tp_npair(V), L = OV + OV * OV, Lp = OV + Po * Pv + 1;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < Lp; e += (int64_t)gridDim.x * blockDim.x) {
    double2 v;
    if (e < OV) {
      v = dense[e];
    } else if (e < Lp - 1) {
      const int64_t q = e - OV;
      int64_t i, j, a, b;
      tp_pdec(q / Pv, O, i, j);
      tp_pdec(q % Pv, V, a, b);
      v = dense[OV + ((i * O + j) * V + a) * V + b];
    } else {
      v = dense[L];
    }
    packed[e] = v;
  }
}

__global__ __launch_bounds__(256) void ci_unpack_kernel(int64_t O, int64_t V, const double2* __restrict__ packed,
                                                        double2* __restrict__ dense) {
  const int64_t OV = O * V, Pv = tp_npair(V), L = OV + OV * OV, Lp = OV + tp_npair(O) * Pv + 1;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= L; e += (int64_t)gridDim.x * blockDim.x) {
    double2 v;
    if (e < OV) {
      v = packed[e];
    } else if (e < L) {
      int64_t q = e - OV;
      const int64_t b = q % V; q /= V;
      const int64_t a = q % V; q /= V;
      const int64_t j = q % O, i = q / O;
      if (i == j || a == b) {
        v = make_double2(0.0, 0.0);
      } else {
        const double sg = ((i < j) == (a < b)) ? 1.0 : -1.0;
        const double2 p = packed[OV + tp_pidx(i < j ? i : j, i < j ? j : i, O) * Pv + tp_pidx(a < b ? a : b, a < b ? b : a, V)];
        v = make_double2(sg * p.x, sg * p.y);
      }
    } else {
      v = packed[Lp - 1];
    }
    dense[e] = v;
  }
}

// r / (theta - D) on the packed layout (D_ia, D_ijab; 0 for c0), denominators kept away from zero as in
// precond_kernel
__global__ __launch_bounds__(256) void precond_packed_kernel(int64_t O, int64_t V, const double* __restrict__ eps,
                                                             double theta, const double2* __restrict__ r,
                                                             double2* __restrict__ out,
                                                             const double* __restrict__ theta_dev) {
  if (theta_dev) theta = *theta_dev;
  const int64_t OV = O * V, Pv = tp_npair(V), Lp = OV + tp_npair(O) * Pv + 1;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < Lp; e += (int64_t)gridDim.x * blockDim.x) {
    double D;
    if (e < OV) {
      D = eps[O + e % V] - eps[e / V];
    } else if (e < Lp - 1) {
      const int64_t q = e - OV;
      int64_t i, j, a, b;
      tp_pdec(q / Pv, O, i, j);
      tp_pdec(q % Pv, V, a, b);
      D = eps[O + a] + eps[O + b] - eps[i] - eps[j];
    } else {
      D = 0.0;
    }
    double den = theta - D;
    if (fabs(den) < 1e-8) den = den < 0.0 ? -1e-8 : 1e-8;
    out[e] = make_double2(r[e].x / den, r[e].y / den);
  }
}

}  // namespace

extern "C" {

int64_t apy_ci_packed_length(int64_t O, int64_t V) {
  if (O < 2 || V < 2) return 0;
  return O * V + tp_npair(O) * tp_npair(V) + 1;
}

int apy_ci_pack(int64_t O, int64_t V, const void* dense, void* packed, void* stream) {
  if (O < 2 || V < 2 || !dense || !packed || dense == packed) return APY_E_ARG;
  ci_pack_kernel<<<apy_grid1d(apy_ci_packed_length(O, V)), 256, 0, (hipStream_t)stream>>>(O, V, (const double2*)dense,
                                                                                          (double2*)packed);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_ci_unpack(int64_t O, int64_t V, const void* packed, void* dense, void* stream) {
  if (O < 2 || V < 2 || !dense || !packed || dense == packed) return APY_E_ARG;
  const int64_t L = O * V + O * O * V * V;
  ci_unpack_kernel<<<apy_grid1d(L + 1), 256, 0, (hipStream_t)stream>>>(O, V, (const double2*)packed,
                                                                       (double2*)dense);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_ci_precondition_packed(int64_t O, int64_t V, const void* eps, double theta, const void* r, void* out,
                               void* stream) {
  if (O < 2 || V < 2 || !eps || !r || !out) return APY_E_ARG;
  precond_packed_kernel<<<apy_grid1d(apy_ci_packed_length(O, V)), 256, 0, (hipStream_t)stream>>>(
      O, V, (const double*)eps, theta, (const double2*)r, (double2*)out, nullptr);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_ci_precondition_packed_dev(int64_t O, int64_t V, const void* eps, const void* theta, const void* r, void* out,
                                   void* stream) {
  if (O < 2 || V < 2 || !eps || !theta || !r || !out) return APY_E_ARG;
  precond_packed_kernel<<<apy_grid1d(apy_ci_packed_length(O, V)), 256, 0, (hipStream_t)stream>>>(
      O, V, (const double*)eps, 0.0, (const double2*)r, (double2*)out, (const double*)theta);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_ci_precondition(int64_t O, int64_t V, const void* eps, double theta, const void* r, void* out,
                        void* stream) {
  if (O < 1 || V < 1 || !eps || !r || !out) return APY_E_ARG;
  const int64_t L = O * V + O * O * V * V;
  precond_kernel<<<apy_grid1d(L + 1), 256, 0, (hipStream_t)stream>>>(O, V, (const double*)eps, theta,
                                                                      (const double2*)r, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_cisd_so_opd_workspace(int64_t O, int64_t V) {
  if (O < 1 || V < 1) return 0;
  apy_gemm_desc g[4];
  opd_gemms(O, V, nullptr, nullptr, g);
  size_t m = 0;
  for (int q = 0; q < 4; ++q) {
    const size_t b = apy_zgemm_ws_bytes(&g[q]);
    if (b > m) m = b;
  }
  return m;
}

int apy_cisd_so_opd(int64_t O, int64_t V, const void* c, void* D, void* workspace, size_t ws_bytes, void* stream) {
  if (O < 1 || V < 1 || !c || !D) return APY_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  apy_gemm_desc g[4];
  opd_gemms(O, V, (const double2*)c, (double2*)D, g);
  for (int q = 0; q < 4; ++q) {
    const int s = apy_zgemm_launch(&g[q], workspace, ws_bytes, st);
    if (s != APY_OK) return s;
  }
  opd_ov_kernel<<<apy_grid1d(O * V), 256, 0, st>>>(O, V, (const double2*)c, (double2*)D);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_cisd_so_tpd_workspace(int64_t O, int64_t V) {
  if (O < 2 || V < 2) return 0;
  return tpd_carve(O, V).total;
}

int apy_cisd_so_tpd(int64_t O, int64_t V, const void* c, const apy_so_ints* grad, void* workspace, size_t ws_bytes,
                    void* stream) {
  if (O < 2 || V < 2 || !c || !grad || grad->O != O || grad->V != V) return APY_E_ARG;
  const TpdWS w = tpd_carve(O, V);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* base = (char*)workspace;
  double2* Cp = (double2*)(base + w.cp);
  double2* Xr = (double2*)(base + w.xr);
  const int64_t OV = O * V;
  const double2* cv = (const double2*)c;
  if (grad->vvvv_p || grad->oooo_p || grad->ring) {
    const int64_t total = (OV + 1) * OV + tp_npair(O) * tp_npair(V);
    tpd_pack_kernel<<<apy_grid1d(total), 256, 0, st>>>(O, V, cv, Cp, Xr);
    APY_LAUNCH_CHECK();
  }
  apy_gemm_desc d[4];
  int used = 0;
  double2* Th = (double2*)(base + w.th);
  if (grad->vovv_p) {
    tpd_half_pack_kernel<<<apy_grid1d(O * O * tp_npair(V)), 256, 0, st>>>(O, V, cv + OV, Th);
    APY_LAUNCH_CHECK();
  }
  tpd_gemms(O, V, cv, Cp, Xr, Th, grad, d, &used);
  for (int q = 0; q < used; ++q) {
    const int s = apy_zgemm_launch(&d[q], base + w.gemm, ws_bytes - w.gemm, st);
    if (s != APY_OK) return s;
  }
  const int64_t Po = tp_npair(O), Pv = tp_npair(V);
  const int64_t total = (grad->ooov ? O * O * OV : 0) + (grad->ovoo_p ? OV * Po : 0) + (grad->vvvo_p ? Pv * OV : 0) +
                        ((grad->vvoo_t || grad->oovv) ? OV * OV : 0);
  if (total > 0) {
    tpd_small_kernel<<<apy_grid1d(total), 256, 0, st>>>(O, V, cv, (double2*)grad->ooov, (double2*)grad->ovoo_p,
                                                        (double2*)grad->vvvo_p, (double2*)grad->vvoo_t,
                                                        (double2*)grad->oovv);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

int apy_so_ints_grad_to_full(int64_t n, const apy_so_ints* grad, void* full, void* stream) {
  if (!grad || !full || grad->O < 1 || grad->V < 1 || grad->O + grad->V != n) return APY_E_ARG;
  if (!grad->vvvv_p || !grad->ring || !grad->oooo_p || !grad->ovoo_p || !grad->vvvo_p || !grad->vovv_p ||
      !grad->ooov || !grad->vvoo_t || !grad->oovv)
    return APY_E_ARG;
  const int64_t total = n * n * n * n;
  tpd_full_kernel<<<apy_grid1d(total), 256, 0, (hipStream_t)stream>>>(grad->O, grad->V, *grad, (double2*)full);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
