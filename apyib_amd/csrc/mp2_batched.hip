// Batched orbital-invariant spin-orbital MP2 over independent instances of
// possibly different sizes (the +/- displaced systems of a finite-difference
// set), one launch per iteration for the whole set.
//
// Per instance k (O_k, V_k, F_k, eps_k, drv_k) the iteration is the Jacobi
// step of the non-canonical MP2 equations (ci_wfn.py:210-212 with the
// driver <ab||ij>, solve_MP2_SO at diagonal F, mp2_wfn.py:64-87):
//   r_ij[a,b] = s drv_ij[a,b] + Y_ij[a,b] - Y_ij[b,a] - Z_ij[a,b] + Z_ji[a,b]
//   Y_ij[a,b] = sum_e x_ij[a,e] Fvv(b,e),   Z_ij[a,b] = sum_m Foo(m,j) x_im[a,b]
//   x_new = x + r / (e_i + e_j - e_a - e_b)
// with Fvv(b,e) = F[O+b][O+e], Foo(m,j) = F[m][j]; transpose_f uses F^T
// (the adjoint lambda equations).  Per step, over all instances at once:
// nc_z_kernel forms Z (one workgroup per (instance, i, slice of the V^2
// elements): the O amplitudes x_im[e] of an element are read once,
// coalesced, for all j); nc_w_kernel forms W_ij = Y_ij - Y_ij^T (one
// workgroup per pair, x_ij and Fvv staged in LDS, 2 x 2 register blocks);
// nc_update_kernel applies the update element by element (coalesced) and
// writes each pair's squared residual norm; per_instance_sum reduces them
// per instance (deterministic order).  Z and W live at the end of the
// workspace.  All tables live in the
// workspace, written once by apy_mp2_nc_batched_setup, so a sequence of steps
// can be captured in a HIP graph.
#include "common.h"

#include <map>
#include <mutex>
#include <vector>

namespace {

constexpr int BLK = 256;
constexpr int LDS_V = 66;          // x_ij and Fvv staged in LDS when V <= LDS_V
constexpr int ZMAX_O = 32;         // Z kernel keeps an element's O amplitudes in registers up to this O
constexpr int GEMM_MIN_V = 48;     // from this V on, Y = x Fvv^T of an instance is one MFMA GEMM (M = O^2 V)

struct Item {
  int32_t inst, i, j, pad;
};

struct Header {
  int64_t n, items;
};

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace: header, instance table, pair table, per-instance pair offsets, per-instance offsets of the Z
// blocks (in double2), pair norms, then the Z blocks and the W blocks (each the sum of O^2 V^2 double2; W
// starts zoff[n] elements after Z).  Every offset but the sizes of the last two regions follows from
// (n, items), so the step needs no host table.
constexpr int EN_B = 64;           // energy partial sums per instance

struct WsLayout {
  size_t inst, item, off, zoff, part, epart, z;
};

inline WsLayout layout(int n, int64_t items) {
  WsLayout L;
  L.inst = al256(sizeof(Header));
  L.item = L.inst + al256((size_t)n * sizeof(apy_mp2_nc_instance));
  L.off = L.item + al256((size_t)items * sizeof(Item));
  L.zoff = L.off + al256((size_t)(n + 1) * sizeof(int64_t));
  L.part = L.zoff + al256((size_t)(n + 1) * sizeof(int64_t));
  L.epart = L.part + al256((size_t)items * sizeof(double));
  L.z = L.epart + al256((size_t)n * EN_B * sizeof(double2));
  return L;
}

inline int64_t z_elements(int n, const apy_mp2_nc_instance* inst) {
  int64_t t = 0;
  for (int k = 0; k < n; ++k) t += inst[k].O * inst[k].O * inst[k].V * inst[k].V;
  return t;
}

__device__ inline void cfma(double2& acc, double2 a, double2 b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
}

// Z_k[i][j][e] = sum_m Foo(m, j) x_k[i][m][e] for every instance k (Foo(m, j) = F[m][j], or F[j][m] for F^T).
// Workgroup (k, i, s) of the pair table covers slice s of the V^2 elements (O slices): a thread loads the O
// amplitudes x_im[e] of its element once (coalesced over e) and writes Z_ij[e] for every j.
__global__ __launch_bounds__(BLK) void nc_z_kernel(const apy_mp2_nc_instance* __restrict__ insts,
                                                   const Item* __restrict__ items, const int64_t* __restrict__ zoff,
                                                   int transpose_f, int reverse, double2* __restrict__ zbase) {
  __shared__ double2 fo[ZMAX_O * ZMAX_O];
  const Item it = items[blockIdx.x];
  const apy_mp2_nc_instance I = insts[it.inst];
  const int64_t O = I.O, V = I.V, n = O + V, VV = V * V;
  const int64_t i = it.i;
  const double2* F = (const double2*)I.fock;
  const double2* x = (const double2*)(reverse ? I.x_new : I.x);
  double2* Z = zbase + zoff[it.inst] + i * O * VV;
  const int64_t chunk = (VV + O - 1) / O, e0 = it.j * chunk, e1 = e0 + chunk < VV ? e0 + chunk : VV;
  auto foo = [&](int64_t m, int64_t j) { return transpose_f ? F[j * n + m] : F[m * n + j]; };
  if (O <= ZMAX_O) {
    for (int64_t q = threadIdx.x; q < O * O; q += BLK) fo[q] = foo(q / O, q % O);  // fo[m O + j]
    __syncthreads();
    for (int64_t e = e0 + threadIdx.x; e < e1; e += BLK) {
      double2 xv[ZMAX_O];
#pragma unroll
      for (int m = 0; m < ZMAX_O; ++m) xv[m] = m < O ? x[(i * O + m) * VV + e] : make_double2(0.0, 0.0);
      for (int64_t j = 0; j < O; ++j) {
        double2 z = make_double2(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < ZMAX_O; ++m)
          if (m < O) cfma(z, fo[m * O + j], xv[m]);
        Z[j * VV + e] = z;
      }
    }
  } else {
    for (int64_t e = e0 + threadIdx.x; e < e1; e += BLK)
      for (int64_t j = 0; j < O; ++j) {
        double2 z = make_double2(0.0, 0.0);
        for (int64_t m = 0; m < O; ++m) cfma(z, foo(m, j), x[(i * O + m) * VV + e]);
        Z[j * VV + e] = z;
      }
  }
}

// Y_ij = x_ij Fvv^T for every pair i != j of the instances whose F_vv term is not a GEMM (V > GEMM_MIN_V does
// not hold): one (instance, i, j) pair per workgroup, Y formed in 2 x 2 blocks (a thread owns the block pair
// {(A,B), (B,A)}, A <= B); Fvv is staged in LDS when V <= LDS_V, x_ij is read through the caches.
__global__ __launch_bounds__(BLK) void nc_w_kernel(const apy_mp2_nc_instance* __restrict__ insts,
                                                   const Item* __restrict__ items, int transpose_f, int reverse,
                                                   const int64_t* __restrict__ zoff, int ninst,
                                                   double2* __restrict__ zbase) {
  __shared__ double2 fs[LDS_V * LDS_V];
  const Item it = items[blockIdx.x];
  if (it.i == it.j) return;  // block-uniform: the diagonal pair keeps its amplitudes
  const apy_mp2_nc_instance I = insts[it.inst];
  const int64_t O = I.O, V = I.V, n = O + V, VV = V * V;
  if (V >= GEMM_MIN_V) return;  // formed by the instance's GEMM
  const int64_t i = it.i, j = it.j;
  const double2* F = (const double2*)I.fock;
  const double2* xij = (const double2*)(reverse ? I.x_new : I.x) + (i * O + j) * VV;
  double2* W = zbase + zoff[ninst] + zoff[it.inst] + (i * O + j) * VV;
  const bool lds = V <= LDS_V;
  // Fvv(b, q): F[O+b][O+q] (or F[O+q][O+b] for F^T), staged as fs[b V + q]
  const int64_t fr0 = transpose_f ? 1 : n, fc0 = transpose_f ? n : 1;
  if (lds)
    for (int64_t e = threadIdx.x; e < VV; e += BLK) fs[e] = F[(O + e / V) * fr0 + (O + e % V) * fc0];
  __syncthreads();
  const double2* X = xij;   // rows of x_ij stay in L1/L2: with Fvv alone in LDS two workgroups fit a CU
  const int64_t fr = lds ? V : fr0, fc = lds ? 1 : fc0;
  const double2* Fv = lds ? (const double2*)fs : F + O * n + O;
  const int64_t nb = (V + 1) / 2;
  const int64_t npair = nb * (nb + 1) / 2;
  for (int64_t p = threadIdx.x; p < npair; p += BLK) {
    // unordered block pair (A, B), A <= B, from the triangular index p
    int64_t A = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (A * (A + 1) / 2 > p) --A;
    while ((A + 1) * (A + 2) / 2 <= p) ++A;
    const int64_t Bq = p - A * (A + 1) / 2;          // Bq <= A
    const int64_t a0 = 2 * Bq, b0 = 2 * A;            // rows of block 1 (a0..a0+1), block 2 (b0..b0+1)
    const bool a1ok = a0 + 1 < V, b1ok = b0 + 1 < V;
    double2 yab[2][2], yba[2][2];
    for (int u = 0; u < 2; ++u)
      for (int w = 0; w < 2; ++w) { yab[u][w] = make_double2(0.0, 0.0); yba[u][w] = make_double2(0.0, 0.0); }
    for (int64_t q = 0; q < V; ++q) {
      const double2 xa0 = X[a0 * V + q], xa1 = a1ok ? X[(a0 + 1) * V + q] : make_double2(0.0, 0.0);
      const double2 xb0 = X[b0 * V + q], xb1 = b1ok ? X[(b0 + 1) * V + q] : make_double2(0.0, 0.0);
      const double2 fa0 = Fv[a0 * fr + q * fc], fa1 = a1ok ? Fv[(a0 + 1) * fr + q * fc] : make_double2(0.0, 0.0);
      const double2 fb0 = Fv[b0 * fr + q * fc], fb1 = b1ok ? Fv[(b0 + 1) * fr + q * fc] : make_double2(0.0, 0.0);
      cfma(yab[0][0], xa0, fb0); cfma(yab[0][1], xa0, fb1);
      cfma(yab[1][0], xa1, fb0); cfma(yab[1][1], xa1, fb1);
      cfma(yba[0][0], xb0, fa0); cfma(yba[0][1], xb0, fa1);
      cfma(yba[1][0], xb1, fa0); cfma(yba[1][1], xb1, fa1);
    }
    // yab[u][w] = Y[a0+u][b0+w], yba[u][w] = Y[b0+u][a0+w]
    for (int u = 0; u < 2; ++u)
      for (int w = 0; w < 2; ++w) {
        if (a0 + u < V && b0 + w < V) W[(a0 + u) * V + b0 + w] = yab[u][w];
        if (b0 + u < V && a0 + w < V) W[(b0 + u) * V + a0 + w] = yba[u][w];
      }
  }
}

// The Jacobi update of one (instance, i, j) pair per workgroup, every element coalesced:
//   r = s drv + W_ij - Z_ij + Z_ji  (a != b; the diagonal elements and the i == j pairs keep x),
//   x_new = x + r / (e_i + e_j - e_a - e_b),  part[pair] = sum |r|^2.
__global__ __launch_bounds__(BLK) void nc_update_kernel(const apy_mp2_nc_instance* __restrict__ insts,
                                                        const Item* __restrict__ items, double s, int reverse,
                                                        const int64_t* __restrict__ zoff, int ninst,
                                                        const double2* __restrict__ zbase,
                                                        double* __restrict__ part) {
  __shared__ double red[BLK];
  const Item it = items[blockIdx.x];
  const apy_mp2_nc_instance I = insts[it.inst];
  const int64_t O = I.O, V = I.V, VV = V * V;
  const int64_t i = it.i, j = it.j, ij = (i * O + j) * VV;
  const double* eps = (const double*)I.eps;
  const double2* x = (const double2*)(reverse ? I.x_new : I.x) + ij;
  double2* xn = (double2*)(reverse ? I.x : I.x_new) + ij;
  double acc = 0.0;
  if (i == j) {
    for (int64_t e = threadIdx.x; e < VV; e += BLK) xn[e] = x[e];
  } else {
    const double2* drv = (const double2*)I.drv + ij;
    const double2* Yij = zbase + zoff[ninst] + zoff[it.inst] + ij;   // Y_ij = x_ij Fvv^T
    const double2* Zij = zbase + zoff[it.inst] + ij;
    const double2* Zji = zbase + zoff[it.inst] + (j * O + i) * VV;
    const double eij = eps[i] + eps[j];
    for (int64_t e = threadIdx.x; e < VV; e += BLK) {
      const int64_t a = e / V, b = e % V;
      const double2 xv = x[e];
      if (a == b) {
        xn[e] = xv;
        continue;
      }
      const double2 d = drv[e], y1 = Yij[e], y2 = Yij[b * V + a], zij = Zij[e], zji = Zji[e];
      const double2 w = make_double2(y1.x - y2.x, y1.y - y2.y);
      const double rr = s * d.x + w.x - zij.x + zji.x;
      const double ri = s * d.y + w.y - zij.y + zji.y;
      const double D = eij - eps[O + a] - eps[O + b];
      xn[e] = make_double2(xv.x + rr / D, xv.y + ri / D);
      acc += rr * rr + ri * ri;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(BLK) void per_instance_sum(const int64_t* __restrict__ off,
                                                        const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[BLK];
  const int64_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
  double acc = 0.0;
  for (int64_t q = b + threadIdx.x; q < e; q += BLK) acc += part[q];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// energy[k] = scale * sum w_k x_k (bilinear), in two fixed-order passes: EN_B partial sums per instance
// (blockIdx.y = instance), then one block per instance adds its partials.
__global__ __launch_bounds__(BLK) void nc_energy_partial(const apy_mp2_nc_instance* __restrict__ insts, int use_x_new,
                                                         double2* __restrict__ epart) {
  __shared__ double2 red[BLK];
  const apy_mp2_nc_instance I = insts[blockIdx.y];
  const double2* w = (const double2*)I.w;
  const double2* x = (const double2*)(use_x_new ? I.x_new : I.x);
  const int64_t N = I.O * I.O * I.V * I.V;
  double2 acc = make_double2(0.0, 0.0);
  if (w)
    for (int64_t e = blockIdx.x * (int64_t)BLK + threadIdx.x; e < N; e += (int64_t)EN_B * BLK) cfma(acc, w[e], x[e]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int q = BLK / 2; q > 0; q >>= 1) {
    if ((int)threadIdx.x < q) {
      red[threadIdx.x].x += red[threadIdx.x + q].x;
      red[threadIdx.x].y += red[threadIdx.x + q].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) epart[(int64_t)blockIdx.y * EN_B + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(EN_B) void nc_energy_final(const double2* __restrict__ epart, double scale,
                                                       double2* __restrict__ out) {
  __shared__ double2 red[EN_B];
  red[threadIdx.x] = epart[(int64_t)blockIdx.x * EN_B + threadIdx.x];
  __syncthreads();
  for (int q = EN_B / 2; q > 0; q >>= 1) {
    if ((int)threadIdx.x < q) {
      red[threadIdx.x].x += red[threadIdx.x + q].x;
      red[threadIdx.x].y += red[threadIdx.x + q].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = make_double2(scale * red[0].x, scale * red[0].y);
}

// Host copy of each workspace's instance table (set up by apy_mp2_nc_batched_setup), from which the step
// launches the per-instance F_vv GEMMs.
std::mutex g_tab_mu;
std::map<const void*, std::vector<apy_mp2_nc_instance>> g_tabs;

int64_t count_items(int n, const apy_mp2_nc_instance* inst) {
  int64_t items = 0;
  for (int k = 0; k < n; ++k) items += inst[k].O * inst[k].O;
  return items;
}

}  // namespace

extern "C" {

size_t apy_mp2_nc_batched_workspace(int n, const apy_mp2_nc_instance* inst) {
  if (n < 1 || !inst) return 0;
  return layout(n, count_items(n, inst)).z + 2 * (size_t)z_elements(n, inst) * sizeof(double2);
}

int apy_mp2_nc_batched_setup(int n, const apy_mp2_nc_instance* inst, void* workspace, size_t ws_bytes,
                             int64_t* n_items, void* stream) {
  if (n < 1 || n > 65535 || !inst || !n_items) return APY_E_ARG;
  for (int k = 0; k < n; ++k) {
    const apy_mp2_nc_instance& I = inst[k];
    if (I.O < 1 || I.V < 1 || I.O > 32767 || !I.fock || !I.eps || !I.drv || !I.x || !I.x_new || I.x == I.x_new)
      return APY_E_ARG;
  }
  const int64_t items = count_items(n, inst);
  if (items > 2147483647) return APY_E_ARG;
  const WsLayout L = layout(n, items);
  if (!workspace || ws_bytes < L.z + 2 * (size_t)z_elements(n, inst) * sizeof(double2)) return APY_E_WORKSPACE;
  std::vector<Item> it;
  it.reserve(items);
  std::vector<int64_t> off(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    off[k] = (int64_t)it.size();
    for (int32_t i = 0; i < (int32_t)inst[k].O; ++i)
      for (int32_t j = 0; j < (int32_t)inst[k].O; ++j) it.push_back(Item{k, i, j, 0});
  }
  off[n] = (int64_t)it.size();
  std::vector<int64_t> zoff(n + 1, 0);
  for (int k = 0; k < n; ++k) zoff[k + 1] = zoff[k] + inst[k].O * inst[k].O * inst[k].V * inst[k].V;
  Header h{n, items};
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  // pageable sources: each copy returns once the host buffer has been read
  APY_HIP_TRY(hipMemcpyAsync(w, &h, sizeof(h), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.inst, inst, n * sizeof(apy_mp2_nc_instance), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.item, it.data(), items * sizeof(Item), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.zoff, zoff.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipStreamSynchronize(st));
  {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    g_tabs[workspace] = std::vector<apy_mp2_nc_instance>(inst, inst + n);
  }
  *n_items = items;
  return APY_OK;
}

int apy_mp2_nc_batched_step(int n, int64_t n_items, int transpose_f, double drv_scale, int reverse, void* rnorm,
                            void* workspace, void* stream) {
  if (n < 1 || n > 65535 || n_items < 1 || n_items > 2147483647 || !rnorm || !workspace) return APY_E_ARG;
  const WsLayout L = layout(n, n_items);
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)(w + L.part);
  const int64_t* zoff = (const int64_t*)(w + L.zoff);
  double2* zbase = (double2*)(w + L.z);
  const auto* insts = (const apy_mp2_nc_instance*)(w + L.inst);
  const auto* items = (const Item*)(w + L.item);
  std::vector<apy_mp2_nc_instance> tab;
  {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto f = g_tabs.find(workspace);
    if (f == g_tabs.end() || (int)f->second.size() != n) return APY_E_ARG;  // not set up on this workspace
    tab = f->second;
  }
  // the Y blocks start zoff[n] elements after the Z blocks (the kernels read that entry themselves)
  nc_z_kernel<<<(unsigned)n_items, BLK, 0, st>>>(insts, items, zoff, transpose_f, reverse, zbase);
  APY_LAUNCH_CHECK();
  nc_w_kernel<<<(unsigned)n_items, BLK, 0, st>>>(insts, items, transpose_f, reverse, zoff, n, zbase);
  APY_LAUNCH_CHECK();
  // Y_k[(i,j,a)][b] = sum_q x_k[(i,j,a)][q] Fvv(b, q) on the matrix cores, one GEMM per instance with V >=
  // GEMM_MIN_V (M = O^2 V, N = K = V; Fvv(b, q) = F[O+b][O+q], or F[O+q][O+b] for F^T)
  double2* ybase = zbase;
  int64_t yoff = 0;
  for (int k = 0; k < n; ++k) yoff += tab[k].O * tab[k].O * tab[k].V * tab[k].V;
  ybase += yoff;
  int64_t off = 0;
  for (int k = 0; k < n; ++k) {
    const apy_mp2_nc_instance& I = tab[k];
    const int64_t O = I.O, V = I.V, nn = O + V;
    if (V >= GEMM_MIN_V) {
      apy_gemm_desc d{};
      d.m = O * O * V; d.n = V; d.k = V;
      d.a_layout = APY_MK; d.a = reverse ? I.x_new : I.x; d.lda = V;
      d.b_layout = transpose_f ? APY_KN : APY_NK;
      d.b = (const double2*)I.fock + O * nn + O; d.ldb = nn;
      d.c = ybase + off; d.ldc = V;
      d.alpha_re = 1.0; d.batch = 1; d.split_k = 1;
      const int rc = apy_zgemm_launch(&d, nullptr, 0, st);
      if (rc != APY_OK) return rc;
    }
    off += O * O * V * V;
  }
  nc_update_kernel<<<(unsigned)n_items, BLK, 0, st>>>(insts, items, drv_scale, reverse, zoff, n, zbase, part);
  APY_LAUNCH_CHECK();
  per_instance_sum<<<(unsigned)n, BLK, 0, st>>>((const int64_t*)(w + L.off), part, (double*)rnorm);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_mp2_nc_batched_energy(int n, int64_t n_items, int use_x_new, double scale, void* energy,
                              void* workspace, void* stream) {
  if (n < 1 || n > 65535 || n_items < 1 || !energy || !workspace) return APY_E_ARG;
  const WsLayout L = layout(n, n_items);
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  double2* epart = (double2*)(w + L.epart);
  nc_energy_partial<<<dim3(EN_B, (unsigned)n), BLK, 0, st>>>((const apy_mp2_nc_instance*)(w + L.inst), use_x_new,
                                                              epart);
  APY_LAUNCH_CHECK();
  nc_energy_final<<<(unsigned)n, EN_B, 0, st>>>(epart, scale, (double2*)energy);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
