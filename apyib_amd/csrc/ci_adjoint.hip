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
#include "common.h"

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

}  // namespace

extern "C" {

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

}  // extern "C"
