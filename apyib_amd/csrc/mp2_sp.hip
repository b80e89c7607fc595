// Spin-adapted (closed-shell) MP2 energy derivatives, the reference's forms
// of compute_MP2_Hessian (analytic.py:307-335) on the device.
//
// t2[i][j][a][b] are closed-shell MP2 amplitudes (mp2_wfn.py:54), pair
// symmetric (t_ijab = t_jiba, as <ab|ij> / D is); n = o + v; a Fock-like
// matrix f is [n][n] row-major.  The four Fock terms of a t2 response come in
// the reference's two index orders:
//   form 0 (analytic.py:327-330, analytic_aats.py:447-452, nuclear):
//     - f_ik t_kjab - t_ikab f_kj + t_ijcb f_ac + t_ijac f_cb
//   form 1 (analytic_aats.py:346-350, magnetic field):
//     - f_ki t_kjab - f_kj t_ikab + f_ac t_ijcb + f_bc t_ijac
// With the pair symmetry every term is a row product of T or the pair image
// (X^P_ijab = X_jiba) of one: with Bt = T f_vv^T and Bf = T f_vv over rows
// (i,j,a), Ct = f_oo^T T and Cf = f_oo T over rows i,
//   form 0 = Bt^P + Bf - Cf - Ct^P,   form 1 = Bt^P + Bt - Ct - Ct^P,
// four GEMMs (two for form 1) per direction and one elementwise pass.
//   apy_mp2_sp_response:    t2_grad = (eri_grad + Fock terms) / D        (analytic.py:326-331)
//   apy_mp2_sp_densities:   dE/dERI_grad = 2u, dE/df_oo = -2 T U^T over (j,a,b),
//                           dE/df_vv = 2 sum_ijb t_ijab u_ijcb, u = 2t - t^T (:320-324)
//   apy_mp2_sp_directional: E_grad of many directions from those densities, by
//                           deterministic two-pass reductions.
#include "common.h"

#include <algorithm>

namespace {

constexpr int DIR_BLOCKS = 64;   // partial sums per direction

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

// The response GEMMs of nb directions into W[0..3] ([nb][o^2 v^2] each): Bt = T f_vv^T, Ct = f_oo^T T and,
// for form 0, Bf = T f_vv, Cf = f_oo T.  Returns the number of GEMMs.
int response_gemms(int64_t o, int64_t v, int64_t nb, int form, const double2* t2, const double2* f, double2* const W[4],
                   apy_gemm_desc g[4]) {
  const int64_t n = o + v, d2 = o * o * v * v, ov2 = o * v * v;
  const double2* fvv = f ? f + o * n + o : nullptr;
  g[0] = gd(o * o * v, v, v, APY_MK, t2, v, APY_NK, fvv, n, W[0], v, 1.0, 0.0, nb, 0, n * n, d2);
  g[1] = gd(o, ov2, o, APY_KM, f, n, APY_KN, t2, ov2, W[1], ov2, 1.0, 0.0, nb, n * n, 0, d2);
  if (form == 1) return 2;
  g[2] = gd(o * o * v, v, v, APY_MK, t2, v, APY_KN, fvv, n, W[2], v, 1.0, 0.0, nb, 0, n * n, d2);
  g[3] = gd(o, ov2, o, APY_MK, f, n, APY_KN, t2, ov2, W[3], ov2, 1.0, 0.0, nb, n * n, 0, d2);
  return 4;
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// t2_grad[q] = (eri_grad[q] + Bt^P + B - C - Ct^P) / D over every (i,j,a,b), with B, C = Bf, Cf (form 0)
// or Bt, Ct (form 1)
__global__ __launch_bounds__(256) void sp_response_epi(int64_t o, int64_t v, int64_t nb,
                                                       const double2* __restrict__ eri_grad,
                                                       const double2* __restrict__ Bt, const double2* __restrict__ Ct,
                                                       const double2* __restrict__ B, const double2* __restrict__ C,
                                                       const double* __restrict__ eps, double2* __restrict__ out) {
  const int64_t d2 = o * o * v * v, total = nb * d2;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = u / d2, t = u % d2;
    const int64_t b = t % v, a = (t / v) % v, j = (t / (v * v)) % o, i = t / (v * v * o);
    const int64_t tp = ((j * o + i) * v + b) * v + a;   // the pair image (j,i,b,a)
    double2 r = eri_grad ? eri_grad[u] : make_double2(0.0, 0.0);
    const double2 b1 = B[q * d2 + t], b2 = Bt[q * d2 + tp], c1 = C[q * d2 + t], c2 = Ct[q * d2 + tp];
    r.x += b1.x + b2.x - c1.x - c2.x;
    r.y += b1.y + b2.y - c1.y - c2.y;
    const double D = eps[i] + eps[j] - eps[o + a] - eps[o + b];
    out[u] = make_double2(r.x / D, r.y / D);
  }
}

// u = 2 t - t^T (a <-> b) into U, and d_eri = 2 u
__global__ __launch_bounds__(256) void sp_u_kernel(int64_t o, int64_t v, const double2* __restrict__ t2,
                                                   double2* __restrict__ U, double2* __restrict__ d_eri) {
  const int64_t total = o * o * v * v;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t % v, a = (t / v) % v, ij = t / (v * v);
    const double2 x = t2[t], y = t2[(ij * v + b) * v + a];
    const double2 uu = make_double2(2.0 * x.x - y.x, 2.0 * x.y - y.y);
    U[t] = uu;
    if (d_eri) d_eri[t] = make_double2(2.0 * uu.x, 2.0 * uu.y);
  }
}

// partial[q][blk] of sum d_eri.eri_grad[q] + d_oo.f[q]_oo + d_vv.f[q]_vv (plain products)
__global__ __launch_bounds__(256) void sp_dir_partial(int64_t o, int64_t v, const double2* __restrict__ d_eri,
                                                      const double2* __restrict__ d_oo,
                                                      const double2* __restrict__ d_vv,
                                                      const double2* __restrict__ f,
                                                      const double2* __restrict__ eri_grad,
                                                      double2* __restrict__ partial) {
  __shared__ double2 red[256];
  const int64_t q = blockIdx.y, n = o + v, d2 = o * o * v * v, nf = o * o + v * v;
  const int64_t total = (eri_grad ? d2 : 0) + (f ? nf : 0);
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)DIR_BLOCKS * blockDim.x) {
    double2 x, y;
    const int64_t ef = eri_grad ? e - d2 : e;
    if (eri_grad && e < d2) {
      x = d_eri[e];
      y = eri_grad[q * d2 + e];
    } else if (ef < o * o) {
      const int64_t i = ef / o, k = ef % o;
      x = d_oo[ef];
      y = f[q * n * n + i * n + k];
    } else {
      const int64_t w = ef - o * o, a = w / v, c = w % v;
      x = d_vv[w];
      y = f[q * n * n + (o + a) * n + o + c];
    }
    acc.x += x.x * y.x - x.y * y.y;
    acc.y += x.x * y.y + x.y * y.x;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x].x += red[threadIdx.x + w].x;
      red[threadIdx.x].y += red[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[q * DIR_BLOCKS + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(DIR_BLOCKS) void sp_dir_final(const double2* __restrict__ partial,
                                                          double2* __restrict__ out) {
  __shared__ double2 red[DIR_BLOCKS];
  red[threadIdx.x] = partial[(int64_t)blockIdx.x * DIR_BLOCKS + threadIdx.x];
  __syncthreads();
  for (int w = DIR_BLOCKS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x].x += red[threadIdx.x + w].x;
      red[threadIdx.x].y += red[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace

extern "C" {

size_t apy_mp2_sp_response_workspace(int64_t o, int64_t v, int64_t nb) {
  if (o < 1 || v < 1 || nb < 1) return 0;
  const size_t reg = round256((size_t)(nb * o * o * v * v) * 16);
  apy_gemm_desc g[4];
  double2* const W[4] = {nullptr, nullptr, nullptr, nullptr};
  const int ng = response_gemms(o, v, nb, 0, nullptr, nullptr, W, g);
  size_t m = 0;
  for (int k = 0; k < ng; ++k) m = std::max(m, apy_zgemm_ws_bytes(&g[k]));
  return 4 * reg + m;
}

int apy_mp2_sp_response(int64_t o, int64_t v, int64_t nb, int form, const void* t2, const void* f_grad,
                        const void* eri_grad, const double* eps, void* t2_grad, void* workspace, size_t ws_bytes,
                        void* stream) {
  if (o < 1 || v < 1 || nb < 1 || nb > 65535 || (form != 0 && form != 1) || !t2 || !f_grad || !eps || !t2_grad)
    return APY_E_ARG;
  const size_t d2b = (size_t)(o * o * v * v) * 16;
  if (t2_grad == t2 || (eri_grad && eri_grad == t2)) return APY_E_ARG;
  const size_t reg = round256((size_t)nb * d2b);
  if (!workspace || ws_bytes < apy_mp2_sp_response_workspace(o, v, nb)) return APY_E_WORKSPACE;
  char* w = (char*)workspace;
  double2* const W[4] = {(double2*)w, (double2*)(w + reg), (double2*)(w + 2 * reg), (double2*)(w + 3 * reg)};
  hipStream_t st = (hipStream_t)stream;
  apy_gemm_desc g[4];
  const int ng = response_gemms(o, v, nb, form, (const double2*)t2, (const double2*)f_grad, W, g);
  for (int k = 0; k < ng; ++k) {
    const int s = apy_zgemm_launch(&g[k], w + 4 * reg, ws_bytes - 4 * reg, st);
    if (s != APY_OK) return s;
  }
  const double2* B = form == 0 ? W[2] : W[0];
  const double2* C = form == 0 ? W[3] : W[1];
  sp_response_epi<<<apy_grid1d(nb * o * o * v * v, 1 << 20), 256, 0, st>>>(
      o, v, nb, (const double2*)eri_grad, W[0], W[1], B, C, eps, (double2*)t2_grad);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_mp2_sp_densities_workspace(int64_t o, int64_t v) {
  if (o < 1 || v < 1) return 0;
  const size_t reg = round256((size_t)(o * o * v * v) * 16);
  const int64_t ov2 = o * v * v;
  apy_gemm_desc g0 = gd(o, o, ov2, APY_MK, nullptr, ov2, APY_NK, nullptr, ov2, nullptr, o, -2.0, 0.0);
  apy_gemm_desc g1 = gd(v, v, v, APY_MK, nullptr, v, APY_NK, nullptr, v, nullptr, v, 2.0, 0.0, o * o, v * v, v * v,
                        0, 1);
  return reg + std::max(apy_zgemm_ws_bytes(&g0), apy_zgemm_ws_bytes(&g1));
}

int apy_mp2_sp_densities(int64_t o, int64_t v, const void* t2, void* d_eri, void* d_oo, void* d_vv, void* workspace,
                         size_t ws_bytes, void* stream) {
  if (o < 1 || v < 1 || !t2 || !d_oo || !d_vv || d_eri == t2) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_mp2_sp_densities_workspace(o, v)) return APY_E_WORKSPACE;
  const size_t reg = round256((size_t)(o * o * v * v) * 16);
  char* w = (char*)workspace;
  double2* U = (double2*)w;
  hipStream_t st = (hipStream_t)stream;
  sp_u_kernel<<<apy_grid1d(o * o * v * v, 1 << 20), 256, 0, st>>>(o, v, (const double2*)t2, U, (double2*)d_eri);
  APY_LAUNCH_CHECK();
  const int64_t ov2 = o * v * v;
  // d_oo[i][k] = -2 sum_(j,a,b) t[i][(jab)] u[k][(jab)]
  apy_gemm_desc g0 = gd(o, o, ov2, APY_MK, t2, ov2, APY_NK, U, ov2, d_oo, o, -2.0, 0.0);
  // d_vv[a][c] = 2 sum_(ij) sum_b t_ij[a][b] u_ij[c][b]
  apy_gemm_desc g1 = gd(v, v, v, APY_MK, t2, v, APY_NK, U, v, d_vv, v, 2.0, 0.0, o * o, v * v, v * v, 0, 1);
  int s = apy_zgemm_launch(&g0, w + reg, ws_bytes - reg, st);
  if (s != APY_OK) return s;
  s = apy_zgemm_launch(&g1, w + reg, ws_bytes - reg, st);
  return s;
}

size_t apy_mp2_sp_directional_workspace(int64_t nd) {
  return nd < 1 ? 0 : (size_t)nd * DIR_BLOCKS * 16;
}

int apy_mp2_sp_directional(int64_t o, int64_t v, int64_t nd, const void* d_eri, const void* d_oo, const void* d_vv,
                           const void* f_grad, const void* eri_grad, void* out, void* workspace, size_t ws_bytes,
                           void* stream) {
  if (o < 1 || v < 1 || nd < 1 || nd > 65535 || !out || (!f_grad && !eri_grad)) return APY_E_ARG;
  if ((eri_grad && !d_eri) || (f_grad && (!d_oo || !d_vv))) return APY_E_ARG;
  if (!workspace || ws_bytes < apy_mp2_sp_directional_workspace(nd)) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  double2* part = (double2*)workspace;
  sp_dir_partial<<<dim3(DIR_BLOCKS, (unsigned)nd), 256, 0, st>>>(
      o, v, (const double2*)d_eri, (const double2*)d_oo, (const double2*)d_vv, (const double2*)f_grad,
      (const double2*)eri_grad, part);
  APY_LAUNCH_CHECK();
  sp_dir_final<<<(unsigned)nd, DIR_BLOCKS, 0, st>>>(part, (double2*)out);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
