// Orbital-invariant spin-orbital MP2 and its adjoint (lambda) on the GPU.
//
//   R(x) = s * drv + P(ab) x F_vv^T - P(ij) F_oo^T x        (ci_wfn.py:211-212)
// with (drv, s, F) = (<ab||ij>, 1, F) for the amplitudes t and
// (<ij||ab>, 1/4, F^T) for the adjoint lambda.  One step is two GEMMs on
// the matrix cores,
//   Y[(ija), b] = sum_e x[(ija), e] F_vv[b, e]      (F^T: F_vv[e, b])
//   Z[i][j, (ab)] = sum_m F_oo[m, j] x[i][m, (ab)]   (F^T: F_oo[j, m])
// and one elementwise pass  r = s drv + Y_ijab - Y_ijba - Z_ijab + Z_jiab,
// x <- x + r/D, accumulating sum |r|^2.  Gradient densities:
//   dE/dF_vv[b,c] =  2 sum_(ija) lambda[(ija),b] t[(ija),c]
//   dE/dF_oo[k,j] = -2 sum_i sum_(ab) t[i][k,(ab)] lambda[i][j,(ab)]
#include "common.h"

#include <algorithm>

namespace {

constexpr int NB = 512;

__global__ __launch_bounds__(256) void nc_update(int64_t O, int64_t V, double s, const double2* __restrict__ drv,
                                                 const double2* __restrict__ Y, const double2* __restrict__ Z,
                                                 const double* __restrict__ eps, const double2* x, double2* xn,
                                                 double* __restrict__ partial) {
  const int64_t n = O * O * V * V;
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r0 = e;
    const int64_t b = r0 % V; r0 /= V;
    const int64_t a = r0 % V; r0 /= V;
    const int64_t j = r0 % O;
    const int64_t i = r0 / O;
    const int64_t ijba = ((i * O + j) * V + b) * V + a;
    const int64_t jiab = ((j * O + i) * V + a) * V + b;
    const double2 d = drv[e];
    double rr = s * d.x + Y[e].x - Y[ijba].x - Z[e].x + Z[jiab].x;
    double ri = s * d.y + Y[e].y - Y[ijba].y - Z[e].y + Z[jiab].y;
    if (i == j || a == b) { rr = 0.0; ri = 0.0; }
    const double D = eps[i] + eps[j] - eps[O + a] - eps[O + b];
    const double2 xv = x[e];
    xn[e] = make_double2(xv.x + rr / D, xv.y + ri / D);
    acc += rr * rr + ri * ri;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void sum_real(const double* __restrict__ partial, int np, double2* out) {
  __shared__ double sh[256];
  double a = 0.0;
  for (int q = threadIdx.x; q < np; q += blockDim.x) a += partial[q];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(sh[0], 0.0);
}

__global__ void zero_kernel(double2* p, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = make_double2(0.0, 0.0);
}

apy_gemm_desc mk(int64_t m, int64_t n, int64_t k, int al, const void* a, int64_t lda, int bl, const void* b,
                 int64_t ldb, void* c, int64_t ldc, double alpha, double beta) {
  apy_gemm_desc d{};
  d.m = m; d.n = n; d.k = k;
  d.a_layout = al; d.b_layout = bl;
  d.alpha_re = alpha; d.beta_re = beta;
  d.a = a; d.lda = lda; d.b = b; d.ldb = ldb; d.c = c; d.ldc = ldc;
  d.batch = 1;
  return d;
}

void step_gemms(int64_t O, int64_t V, const double2* F, int tf, const double2* x, double2* Y, double2* Z,
                apy_gemm_desc g[2]) {
  const int64_t n = O + V;
  // Y[(ija), b] = sum_e x[(ija), e] Fvv(b, e)
  g[0] = mk(O * O * V, V, V, APY_MK, x, V, tf ? APY_KN : APY_NK, F ? F + O * n + O : nullptr, n, Y, V, 1.0, 0.0);
  // Z[i][j, ab] = sum_m Foo(m, j) x[i][m, ab]: A(j, m) = F[m][j] (KM) or F[j][m] (MK, transposed F)
  g[1] = mk(O, V * V, O, tf ? APY_MK : APY_KM, F, n, APY_KN, x, V * V, Z, V * V, 1.0, 0.0);
  g[1].batch = O; g[1].stride_a = 0; g[1].stride_b = O * V * V; g[1].stride_c = O * V * V;
}

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t ws_layout(int64_t O, int64_t V, size_t* off_z, size_t* off_p, size_t* off_g) {
  const size_t t = (size_t)O * O * V * V * sizeof(double2);
  apy_gemm_desc g[2];
  step_gemms(O, V, nullptr, 0, nullptr, nullptr, nullptr, g);
  size_t gb = std::max(apy_zgemm_ws_bytes(&g[0]), apy_zgemm_ws_bytes(&g[1]));
  apy_gemm_desc g2[2];
  step_gemms(O, V, nullptr, 1, nullptr, nullptr, nullptr, g2);
  gb = std::max(gb, std::max(apy_zgemm_ws_bytes(&g2[0]), apy_zgemm_ws_bytes(&g2[1])));
  *off_z = al256(t);
  *off_p = *off_z + al256(t);
  *off_g = *off_p + al256(NB * sizeof(double));
  return *off_g + al256(gb);
}

}  // namespace

extern "C" {

size_t apy_mp2_nc_workspace(int64_t O, int64_t V) {
  if (O < 1 || V < 1) return 0;
  size_t a, b, c;
  return ws_layout(O, V, &a, &b, &c);
}

int apy_mp2_nc_step(const apy_mp2_nc_args* A, void* workspace, size_t ws_bytes, void* stream) {
  if (!A || A->O < 1 || A->V < 1 || !A->fock || !A->eps || !A->drv || !A->x || !A->x_new || !A->rnorm)
    return APY_E_ARG;
  size_t oz, op, og;
  const size_t need = ws_layout(A->O, A->V, &oz, &op, &og);
  if (!workspace || ws_bytes < need) return APY_E_WORKSPACE;
  char* w = (char*)workspace;
  double2* Y = (double2*)w;
  double2* Z = (double2*)(w + oz);
  double* part = (double*)(w + op);
  void* gw = w + og;
  hipStream_t st = (hipStream_t)stream;
  apy_gemm_desc g[2];
  step_gemms(A->O, A->V, (const double2*)A->fock, A->transpose_f, (const double2*)A->x, Y, Z, g);
  for (int q = 0; q < 2; ++q) {
    int s = apy_zgemm_launch(&g[q], gw, need - og, st);
    if (s != APY_OK) return s;
  }
  nc_update<<<NB, 256, 0, st>>>(A->O, A->V, A->drv_scale, (const double2*)A->drv, Y, Z, A->eps,
                                (const double2*)A->x, (double2*)A->x_new, part);
  APY_LAUNCH_CHECK();
  sum_real<<<1, 256, 0, st>>>(part, NB, (double2*)A->rnorm);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_mp2_nc_density_workspace(int64_t O, int64_t V) {
  if (O < 1 || V < 1) return 0;
  const int64_t n = O + V;
  apy_gemm_desc a = mk(V, V, O * O * V, APY_KM, nullptr, V, APY_KN, nullptr, V, nullptr, n, 2.0, 0.0);
  apy_gemm_desc b = mk(O, O, V * V, APY_MK, nullptr, V * V, APY_NK, nullptr, V * V, nullptr, n, -2.0, 0.0);
  b.batch = O; b.reduce_batch = 1;
  return std::max(apy_zgemm_ws_bytes(&a), apy_zgemm_ws_bytes(&b));
}

int apy_mp2_nc_densities(int64_t O, int64_t V, const void* t, const void* lam, void* grad_f, void* workspace,
                         size_t ws_bytes, void* stream) {
  if (O < 1 || V < 1 || !t || !lam || !grad_f) return APY_E_ARG;
  const int64_t n = O + V;
  hipStream_t st = (hipStream_t)stream;
  zero_kernel<<<apy_grid1d(n * n), 256, 0, st>>>((double2*)grad_f, n * n);
  APY_LAUNCH_CHECK();
  double2* G = (double2*)grad_f;
  // dE/dF_vv[b,c] = 2 sum_(ija) lam[(ija), b] t[(ija), c]
  apy_gemm_desc a = mk(V, V, O * O * V, APY_KM, lam, V, APY_KN, t, V, G + O * n + O, n, 2.0, 0.0);
  int s = apy_zgemm_launch(&a, workspace, ws_bytes, st);
  if (s != APY_OK) return s;
  // dE/dF_oo[k,j] = -2 sum_i sum_(ab) t[i][k,(ab)] lam[i][j,(ab)]
  apy_gemm_desc b = mk(O, O, V * V, APY_MK, t, V * V, APY_NK, lam, V * V, G, n, -2.0, 0.0);
  b.batch = O; b.stride_a = O * V * V; b.stride_b = O * V * V; b.stride_c = 0; b.reduce_batch = 1;
  return apy_zgemm_launch(&b, workspace, ws_bytes, st);
}

}  // extern "C"

namespace {

// lam = conj(t) / 4: the lambda amplitudes of a Hermitian system (F = F^H, <ab||ij> = conj(<ij||ab>))
__global__ __launch_bounds__(256) void adj_lambda_hermitian(int64_t n, const double2* __restrict__ t,
                                                            double2* __restrict__ lam) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = t[e];
    lam[e] = make_double2(0.25 * v.x, -0.25 * v.y);
  }
}

__global__ __launch_bounds__(256) void adj_combine_kernel(int64_t total, const double2* __restrict__ d,
                                                          double2* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const double2 f = d[e], w = d[total + e], v = d[2 * total + e];
    out[e] = make_double2(f.x + 0.25 * w.x + v.x, f.y + 0.25 * w.y + v.y);
  }
}

// Workspace: dense t and lambda of the largest item, the [3][total] directional dots, then the scratch
// of the density GEMMs / dot reductions.
struct AdjWS {
  size_t t, lam, dots, scratch, scratch_bytes, total;
};

bool adj_carve(int64_t nitem, const apy_mp2_adjoint_item* it, AdjWS& w, int64_t& ndir_total) {
  auto rnd = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t dense = 0, scratch = 0;
  ndir_total = 0;
  for (int64_t b = 0; b < nitem; ++b) {
    const int64_t O = it[b].O, V = it[b].V, n = O + V;
    if (O < 1 || V < 1 || it[b].ndir < 0 || !it[b].t || !it[b].grad_f) return false;
    if (it[b].ndir > 0 && !it[b].dF) return false;
    dense = std::max(dense, (size_t)(O * O * V * V) * sizeof(double2));
    scratch = std::max(scratch, apy_mp2_nc_density_workspace(O, V));
    scratch = std::max(scratch, apy_reduce_workspace(std::max(n * n, O * O * V * V), APY_MAXVEC));
    ndir_total += it[b].ndir;
  }
  size_t off = 0;
  w.t = off; off += rnd(dense);
  w.lam = off; off += rnd(dense);
  w.dots = off; off += rnd((size_t)std::max<int64_t>(ndir_total, 1) * 3 * sizeof(double2));
  w.scratch = off; w.scratch_bytes = rnd(scratch); off += w.scratch_bytes;
  w.total = off;
  return true;
}

}  // namespace

extern "C" {

size_t apy_mp2_adjoint_sums_workspace(int64_t nitem, const apy_mp2_adjoint_item* items) {
  AdjWS w{};
  int64_t nd = 0;
  if (nitem < 1 || !items || !adj_carve(nitem, items, w, nd)) return 0;
  return w.total;
}

int apy_mp2_adjoint_sums(int64_t nitem, const apy_mp2_adjoint_item* items, void* out, void* workspace,
                         size_t ws_bytes, void* stream) {
  AdjWS w{};
  int64_t total = 0;
  if (nitem < 1 || !items || !out || !adj_carve(nitem, items, w, total)) return APY_E_ARG;
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* base = (char*)workspace;
  double2* T = (double2*)(base + w.t);
  double2* L = (double2*)(base + w.lam);
  double2* D = (double2*)(base + w.dots);
  void* scr = base + w.scratch;
  if (total > 0) {
    zero_kernel<<<apy_grid1d(3 * total), 256, 0, st>>>(D, 3 * total);
    APY_LAUNCH_CHECK();
  }
  int64_t off = 0;
  for (int64_t b = 0; b < nitem; ++b) {
    const apy_mp2_adjoint_item& I = items[b];
    const int64_t O = I.O, V = I.V, n = O + V, d2 = O * O * V * V;
    int s = apy_t2_unpack(O, V, I.t, T, stream);
    if (s == APY_OK && I.lam) s = apy_t2_unpack(O, V, I.lam, L, stream);
    if (s == APY_OK && !I.lam) {
      adj_lambda_hermitian<<<apy_grid1d(d2), 256, 0, st>>>(d2, T, L);
      APY_LAUNCH_CHECK();
    }
    if (s == APY_OK) s = apy_mp2_nc_densities(O, V, T, L, I.grad_f, scr, w.scratch_bytes, stream);
    if (s != APY_OK) return s;
    // sum G.dF_q, sum t.dW_q, sum lambda.dV_q into the three rows of the dot block
    const void* const* xs[3] = {I.dF, I.dW, I.dV};
    const void* ys[3] = {I.grad_f, T, L};
    const int64_t lens[3] = {n * n, d2, d2};
    for (int term = 0; term < 3; ++term) {
      if (!xs[term]) continue;
      for (int64_t s0 = 0; s0 < I.ndir; s0 += APY_MAXVEC) {
        const int k = (int)std::min<int64_t>(APY_MAXVEC, I.ndir - s0);
        s = apy_zdot_many(k, xs[term] + s0, ys[term], lens[term], 0, D + term * total + off + s0, scr,
                          w.scratch_bytes, stream);
        if (s != APY_OK) return s;
      }
    }
    off += I.ndir;
  }
  if (total > 0) {
    adj_combine_kernel<<<apy_grid1d(total), 256, 0, st>>>(total, D, (double2*)out);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

}  // extern "C"
