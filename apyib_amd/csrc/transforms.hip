// AO -> MO transforms, Fock build, MO overlaps and phase alignment
// (utils.py:217-279, 370-446).
//
// compute_ERI_MO (utils.py:274-277) is four quarter transforms, each one
// (batched) complex GEMM on the FP64 matrix cores:
//   X1[(m n l), s] = sum_g (mn|lg) C[g,s]                M=nbf^3 N=nact K=nbf
//   X2[m n][r, s]  = sum_l conj(C[l,r]) X1[m n][l, s]    batch nbf^2 of nact x nact x nbf
//   X3[m][q, (r s)] = sum_n C[n,q] X2[m][n, (r s)]        batch nbf of nact x nact^2 x nbf
//   X4[p, (q r s)] = sum_m conj(C[m,p]) X3[m, (q r s)]    M=nact N=nact^3 K=nbf
// The Fock build (utils.py:249, hf_wfn.py:81) is a two-index contraction of
// the AO integrals with the density, one block per output element pair.
#include "common.h"

namespace {

// F[m,n] = h[m,n] + sum_{l,s} D[l,s] (2 (mn|ls) - (ml|ns))  (times `scale`
// on the two-electron part).  One block per (m, n).
// Instance blockIdx.y of a batch: h, eri, D and F advance by their strides (eri's may be 0: shared).
__global__ __launch_bounds__(256) void fock_kernel(int64_t nb, const double2* __restrict__ h0,
                                                   const double2* __restrict__ eri0, int64_t stride_eri,
                                                   const double2* __restrict__ D0, double scale,
                                                   double2* __restrict__ F0) {
  const int64_t g = blockIdx.y, n2 = nb * nb;
  const double2* h = h0 ? h0 + g * n2 : nullptr;
  const double2* eri = eri0 + g * stride_eri;
  const double2* D = D0 + g * n2;
  double2* F = F0 + g * n2;
  const int64_t mn = blockIdx.x;
  const int64_t m = mn / nb, n = mn % nb;
  double ar = 0.0, ai = 0.0;
  for (int64_t ls = threadIdx.x; ls < nb * nb; ls += blockDim.x) {
    const int64_t l = ls / nb, s = ls % nb;
    const double2 j = eri[(m * nb + n) * nb * nb + ls];
    const double2 k = eri[((m * nb + l) * nb + n) * nb + s];
    const double wr = 2.0 * j.x - k.x, wi = 2.0 * j.y - k.y;
    const double2 d = D[ls];
    ar += d.x * wr - d.y * wi;
    ai += d.x * wi + d.y * wr;
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double2 hv = h ? h[mn] : make_double2(0.0, 0.0);
    F[mn] = make_double2(hv.x + scale * sr[0], hv.y + scale * si[0]);
  }
}

// C'[:, m] = C[:, m] / (S_mm / |S_mm|), S = C0^H S_AO C (utils.py:435-444).
__global__ void phase_kernel(int64_t nb, int64_t ncol, const double2* __restrict__ S, int64_t lds,
                             const double2* __restrict__ C, double2* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nb * ncol; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = t % ncol;
    const double2 s = S[col * lds + col];
    const double a = sqrt(s.x * s.x + s.y * s.y);
    // divide by p = s / |s|:  x / p = x * conj(p) / |p|^2 = x * conj(s) / |s|   (|p| = 1)
    const double pr = s.x / a, pi = s.y / a;
    const double2 x = C[t];
    out[t] = make_double2(x.x * pr + x.y * pi, x.y * pr - x.x * pi);
  }
}

// out[p,q,r,s] = eri[p/2,q/2,r/2,s/2] delta(sp,sq) delta(sr,ss)  (utils.py:332-363)
__global__ __launch_bounds__(256) void eri_so_kernel(int64_t m, const double2* __restrict__ eri,
                                                     double2* __restrict__ out) {
  const int64_t n = 2 * m, total = n * n * n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t r0 = t;
    const int64_t s = r0 % n; r0 /= n;
    const int64_t r = r0 % n; r0 /= n;
    const int64_t q = r0 % n;
    const int64_t p = r0 / n;
    double2 v = make_double2(0.0, 0.0);
    if (((p ^ q) & 1) == 0 && ((r ^ s) & 1) == 0) v = eri[(((p >> 1) * m + (q >> 1)) * m + (r >> 1)) * m + (s >> 1)];
    out[t] = v;
  }
}

apy_gemm_desc gd(int64_t m, int64_t n, int64_t k, int al, const void* a, int64_t lda, int bl, const void* b,
                 int64_t ldb, void* c, int64_t ldc) {
  apy_gemm_desc d{};
  d.m = m; d.n = n; d.k = k;
  d.a_layout = al; d.b_layout = bl;
  d.alpha_re = 1.0;
  d.a = a; d.lda = lda; d.b = b; d.ldb = ldb; d.c = c; d.ldc = ldc;
  d.batch = 1;
  return d;
}

struct EriPlan {
  apy_gemm_desc g[4];
  size_t x_bytes;     // each intermediate buffer
  size_t gemm_bytes;  // split-K workspace
};

EriPlan eri_plan(int64_t nb, int64_t na, const void* eri, const void* C, void* x1, void* x2, void* out) {
  EriPlan P;
  // X1[(mnl), s] = sum_g ERI[(mnl), g] C[g, s]
  P.g[0] = gd(nb * nb * nb, na, nb, APY_MK, eri, nb, APY_KN, C, na, x1, na);
  // X2[mn][r, s] = sum_l conj(C[l, r]) X1[mn][l, s]
  P.g[1] = gd(na, na, nb, APY_KM, C, na, APY_KN, x1, na, x2, na);
  P.g[1].conj_a = 1;
  P.g[1].batch = nb * nb; P.g[1].stride_a = 0; P.g[1].stride_b = nb * na; P.g[1].stride_c = na * na;
  // X3[m][q, (rs)] = sum_n C[n, q] X2[m][n, (rs)]
  P.g[2] = gd(na, na * na, nb, APY_KM, C, na, APY_KN, x2, na * na, x1, na * na);
  P.g[2].batch = nb; P.g[2].stride_a = 0; P.g[2].stride_b = nb * na * na; P.g[2].stride_c = na * na * na;
  // out[p, (qrs)] = sum_m conj(C[m, p]) X3[m, (qrs)]
  P.g[3] = gd(na, na * na * na, nb, APY_KM, C, na, APY_KN, x1, na * na * na, out, na * na * na);
  P.g[3].conj_a = 1;
  const int64_t big = nb * nb * nb * na > nb * na * na * na ? nb * nb * nb * na : nb * na * na * na;
  P.x_bytes = (size_t)big * sizeof(double2);
  P.gemm_bytes = 0;
  for (int q = 0; q < 4; ++q) {
    size_t w = apy_zgemm_ws_bytes(&P.g[q]);
    if (w > P.gemm_bytes) P.gemm_bytes = w;
  }
  return P;
}

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

size_t apy_eri_mo_workspace(int64_t nbf, int64_t nact) {
  if (nbf < 1 || nact < 1) return 0;
  EriPlan P = eri_plan(nbf, nact, nullptr, nullptr, nullptr, nullptr, nullptr);
  return 2 * al256(P.x_bytes) + al256(P.gemm_bytes);
}

int apy_eri_mo(int64_t nbf, int64_t nact, const void* eri_ao, const void* C, void* eri_mo, void* workspace,
               size_t ws_bytes, void* stream) {
  if (nbf < 1 || nact < 1 || nact > nbf || !eri_ao || !C || !eri_mo) return APY_E_ARG;
  EriPlan P0 = eri_plan(nbf, nact, nullptr, nullptr, nullptr, nullptr, nullptr);
  const size_t need = 2 * al256(P0.x_bytes) + al256(P0.gemm_bytes);
  if (!workspace || ws_bytes < need) return APY_E_WORKSPACE;
  char* w = (char*)workspace;
  void* x1 = w;
  void* x2 = w + al256(P0.x_bytes);
  void* gw = w + 2 * al256(P0.x_bytes);
  EriPlan P = eri_plan(nbf, nact, eri_ao, C, x1, x2, eri_mo);
  for (int q = 0; q < 4; ++q) {
    int s = apy_zgemm_launch(&P.g[q], gw, P0.gemm_bytes, (hipStream_t)stream);
    if (s != APY_OK) return s;
  }
  return APY_OK;
}

int apy_eri_so_from_mo(int64_t m, const void* eri_mo, void* eri_so, void* stream) {
  if (m < 1 || !eri_mo || !eri_so) return APY_E_ARG;
  const int64_t n = 2 * m;
  eri_so_kernel<<<apy_grid1d(n * n * n * n, 1 << 20), 256, 0, (hipStream_t)stream>>>(m, (const double2*)eri_mo,
                                                                                    (double2*)eri_so);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_fock_build(int64_t nbf, const void* h, const void* eri_ao, const void* D, double scale, void* F,
                   void* stream) {
  return apy_fock_build_batch(nbf, 1, h, eri_ao, 0, D, scale, F, stream);
}

int apy_fock_build_batch(int64_t nbf, int64_t batch, const void* h, const void* eri_ao, int64_t stride_eri,
                         const void* D, double scale, void* F, void* stream) {
  if (nbf < 1 || batch < 1 || batch > 65535 || !eri_ao || !D || !F || stride_eri < 0) return APY_E_ARG;
  fock_kernel<<<dim3((unsigned)(nbf * nbf), (unsigned)batch), 256, 0, (hipStream_t)stream>>>(
      nbf, (const double2*)h, (const double2*)eri_ao, stride_eri, (const double2*)D, scale, (double2*)F);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_mo_overlap_workspace(int64_t nbf, int64_t nmo, int64_t batch) {
  if (nbf < 1 || nmo < 1 || batch < 1) return 0;
  return al256((size_t)batch * nbf * nmo * sizeof(double2));
}

int apy_mo_overlap(int64_t nbf, int64_t nmo, int64_t batch, const void* c_bra, int64_t stride_bra, const void* s_ao,
                   int64_t stride_s, const void* c_ket, int64_t stride_ket, void* out, void* workspace,
                   size_t ws_bytes, void* stream) {
  if (nbf < 1 || nmo < 1 || batch < 1 || !c_bra || !s_ao || !c_ket || !out) return APY_E_ARG;
  const size_t need = apy_mo_overlap_workspace(nbf, nmo, batch);
  if (!workspace || ws_bytes < need) return APY_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  // Y_b = S_b C_ket,b   (nbf x nmo)
  apy_gemm_desc a = gd(nbf, nmo, nbf, APY_MK, s_ao, nbf, APY_KN, c_ket, nmo, workspace, nmo);
  a.batch = batch; a.stride_a = stride_s; a.stride_b = stride_ket; a.stride_c = nbf * nmo;
  a.split_k = 1;  // no partial-sum workspace
  int st = apy_zgemm_launch(&a, nullptr, 0, s);
  if (st == APY_E_WORKSPACE) return APY_E_ARG;
  if (st != APY_OK) return st;
  // out_b = C_bra,b^H Y_b
  apy_gemm_desc b = gd(nmo, nmo, nbf, APY_KM, c_bra, nmo, APY_KN, workspace, nmo, out, nmo);
  b.conj_a = 1;
  b.batch = batch; b.stride_a = stride_bra; b.stride_b = nbf * nmo; b.stride_c = nmo * nmo;
  b.split_k = 1;
  st = apy_zgemm_launch(&b, nullptr, 0, s);
  if (st == APY_E_WORKSPACE) return APY_E_ARG;
  return st;
}

int apy_phase_align(int64_t nbf, int64_t ncol, const void* s_mo, int64_t lds, const void* C, void* out,
                    void* stream) {
  if (nbf < 1 || ncol < 1 || !s_mo || !C || !out) return APY_E_ARG;
  phase_kernel<<<apy_grid1d(nbf * ncol), 256, 0, (hipStream_t)stream>>>(nbf, ncol, (const double2*)s_mo, lds,
                                                                        (const double2*)C, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
