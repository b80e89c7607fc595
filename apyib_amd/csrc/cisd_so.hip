// Spin-orbital CISD / CID operator on MI355X.
//
// The doubles residual of ci_wfn.py:319-334 is regrouped into GEMMs on the
// FP64 matrix cores plus one fused elementwise epilogue:
//   PPL + HHL + F_oo (ci_wfn.py:326-329)  in (i<j)x(a<b) pair space:
//        R_p  = T_p . vvvv_p^T            [Po x Pv x Pv]   (t2 and <ab||cd>
//        R_p += Hoo . T_p                 [Po x Pv x Po]    antisymmetric, so
//                                                            1/2 sum_cd = sum_c<d)
//   ring (ci_wfn.py:330-333) as one product and its P(ij)P(ab) images:
//        X[(ia),(jb)] = sum_kc t2[i,k,a,c] <kb||cj>    [OV x OV x OV]
//   F_vv (ci_wfn.py:324-325):   Y[ija,b] = sum_c t2[ija,c] F_bc
//   T1 -> T2 (ci_wfn.py:320-323): U[a,b,(ij)] = sum_k t1[k,a] <kb||ij>,
//                                 Q[(ab),j,i] = sum_c <ab||cj> t1[i,c]
// and the epilogue, per packed (i<j, a<b), sums
//   <ab||ij> + R_p + P(ij)P(ab) X + P(ab) Y - P(ab) U + P(ij) Q
// then applies r = s - E t, t += r/D, writing the four antisymmetric images.
// The singles residual (ci_wfn.py:309-316) is six small GEMMs into one
// accumulator and an elementwise update.
#include "common.h"
#include "synth.h"

#include <algorithm>
#include <vector>

namespace {

__host__ __device__ inline int64_t npair(int64_t N) { return N * (N - 1) / 2; }
__host__ __device__ inline int64_t prow(int64_t p, int64_t N) { return p * (2 * N - p - 1) / 2; }

__device__ inline void pdec(int64_t P, int64_t N, int64_t& p, int64_t& q) {
  double b = 2.0 * (double)N - 1.0;
  double disc = b * b - 8.0 * (double)P;
  int64_t pp = (int64_t)((b - sqrt(disc > 0.0 ? disc : 0.0)) * 0.5);
  if (pp < 0) pp = 0;
  if (pp > N - 2) pp = N - 2;
  while (pp > 0 && prow(pp, N) > P) --pp;
  while (pp + 1 <= N - 2 && prow(pp + 1, N) <= P) ++pp;
  p = pp;
  q = P - prow(pp, N) + pp + 1;
}

// ---------------------------------------------------------------- sources
struct SrcFull {
  const double2* A;
  int64_t n;
  __device__ double2 operator()(int64_t p, int64_t q, int64_t r, int64_t s) const {
    return A[((p * n + q) * n + r) * n + s];
  }
};

// <pq||rs> = <pq|rs> - <pq|sr>, <pq|rs> = (p/2 r/2 | q/2 s/2) delta(sp,sr) delta(sq,ss)
struct SrcMO {
  const double2* E;
  int64_t m;
  __device__ double2 chem(int64_t P, int64_t Q, int64_t R, int64_t S) const {
    return E[((P * m + Q) * m + R) * m + S];
  }
  __device__ double2 operator()(int64_t p, int64_t q, int64_t r, int64_t s) const {
    double2 v = make_double2(0.0, 0.0);
    if (((p ^ r) & 1) == 0 && ((q ^ s) & 1) == 0) v = chem(p >> 1, r >> 1, q >> 1, s >> 1);
    if (((p ^ s) & 1) == 0 && ((q ^ r) & 1) == 0) {
      double2 w = chem(p >> 1, s >> 1, q >> 1, r >> 1);
      v.x -= w.x;
      v.y -= w.y;
    }
    return v;
  }
};

enum Kind { K_VVVV_P = 0, K_RING, K_OOOO_P, K_OVOO_P, K_VVVO_P, K_VOVV, K_OOOV, K_VVOO_T, K_OOVV, K_COUNT };

__host__ __device__ inline int64_t kind_size(int kind, int64_t O, int64_t V) {
  const int64_t Po = npair(O), Pv = npair(V);
  switch (kind) {
    case K_VVVV_P: return Pv * Pv;
    case K_RING: return O * V * O * V;
    case K_OOOO_P: return Po * Po;
    case K_OVOO_P: return O * V * Po;
    case K_VVVO_P: return Pv * O * V;
    case K_VOVV: return V * O * V * V;
    case K_OOOV: return O * O * O * V;
    case K_VVOO_T: return O * O * V * V;
    case K_OOVV: return O * O * V * V;
  }
  return 0;
}

// Map flat index t of block `kind` to spin-orbital indices (p,q,r,s) of <pq||rs>.
__device__ inline void kind_index(int kind, int64_t t, int64_t O, int64_t V, int64_t& p, int64_t& q, int64_t& r,
                                  int64_t& s) {
  const int64_t Po = npair(O), Pv = npair(V);
  int64_t a, b, c, d, i, j, k;
  switch (kind) {
    case K_VVVV_P: {
      pdec(t / Pv, V, a, b);
      pdec(t % Pv, V, c, d);
      p = O + a; q = O + b; r = O + c; s = O + d;
      break;
    }
    case K_RING: {  // [(j,b),(k,c)] = <kb||cj>
      c = t % V; t /= V; k = t % O; t /= O; b = t % V; j = t / V;
      p = k; q = O + b; r = O + c; s = j;
      break;
    }
    case K_OOOO_P: {  // [(i<j)][(k<l)] = <kl||ij>
      int64_t l;
      pdec(t / Po, O, i, j);
      pdec(t % Po, O, k, l);
      p = k; q = l; r = i; s = j;
      break;
    }
    case K_OVOO_P: {  // [k][b][(i<j)] = <kb||ij>
      int64_t P = t % Po; t /= Po;
      b = t % V; k = t / V;
      pdec(P, O, i, j);
      p = k; q = O + b; r = i; s = j;
      break;
    }
    case K_VVVO_P: {  // [(a<b)][j][c] = <ab||cj>
      c = t % V; t /= V; j = t % O; int64_t P = t / O;
      pdec(P, V, a, b);
      p = O + a; q = O + b; r = O + c; s = j;
      break;
    }
    case K_VOVV: {  // [a][j][c][b] = <aj||cb>
      b = t % V; t /= V; c = t % V; t /= V; j = t % O; a = t / O;
      p = O + a; q = j; r = O + c; s = O + b;
      break;
    }
    case K_OOOV: {  // [k][j][i][b] = <kj||ib>
      b = t % V; t /= V; i = t % O; t /= O; j = t % O; k = t / O;
      p = k; q = j; r = i; s = O + b;
      break;
    }
    case K_VVOO_T: {  // [i][j][a][b] = <ab||ij>
      b = t % V; t /= V; a = t % V; t /= V; j = t % O; i = t / O;
      p = O + a; q = O + b; r = i; s = j;
      break;
    }
    default: {  // K_OOVV [i][j][a][b] = <ij||ab>
      b = t % V; t /= V; a = t % V; t /= V; j = t % O; i = t / O;
      p = i; q = j; r = O + a; s = O + b;
      break;
    }
  }
}

template <class Src>
__global__ __launch_bounds__(256) void build_block(Src src, int kind, int64_t O, int64_t V, double2* out,
                                                   int64_t total, int accumulate, double scale) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t p, q, r, s;
    kind_index(kind, t, O, V, p, q, r, s);
    double2 v = src(p, q, r, s);
    if (accumulate) {
      double2 o = out[t];
      out[t] = make_double2(o.x + scale * v.x, o.y + scale * v.y);
    } else {
      out[t] = v;
    }
  }
}

void* block_ptr(const apy_so_ints* I, int kind) {
  switch (kind) {
    case K_VVVV_P: return I->vvvv_p;
    case K_RING: return I->ring;
    case K_OOOO_P: return I->oooo_p;
    case K_OVOO_P: return I->ovoo_p;
    case K_VVVO_P: return I->vvvo_p;
    case K_VOVV: return I->vovv;
    case K_OOOV: return I->ooov;
    case K_VVOO_T: return I->vvoo_t;
    case K_OOVV: return I->oovv;
  }
  return nullptr;
}

template <class Src>
int build_all(const Src& src, const apy_so_ints* I, int accumulate, double scale, hipStream_t st) {
  if (!I || I->O < 0 || I->V < 0) return APY_E_ARG;
  for (int kind = 0; kind < K_COUNT; ++kind) {
    double2* out = (double2*)block_ptr(I, kind);
    if (!out) continue;
    int64_t total = kind_size(kind, I->O, I->V);
    if (total <= 0) continue;
    build_block<Src><<<apy_grid1d(total, 1 << 20), 256, 0, st>>>(src, kind, I->O, I->V, out, total, accumulate,
                                                                 scale);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

__global__ void so_fock_kernel(int64_t m, const double2* fmo, double2* fso) {
  const int64_t n = 2 * m;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * n; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t p = t / n, q = t % n;
    fso[t] = ((p ^ q) & 1) ? make_double2(0.0, 0.0) : fmo[(p >> 1) * m + (q >> 1)];
  }
}

// ------------------------------------------------------------ sigma pieces
__global__ __launch_bounds__(256) void pack_t2(const double2* __restrict__ t2, double2* __restrict__ tp, int64_t O,
                                               int64_t V) {
  const int64_t Po = npair(O), Pv = npair(V), total = Po * Pv;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t i, j, a, b;
    pdec(t / Pv, O, i, j);
    pdec(t % Pv, V, a, b);
    tp[t] = t2[((i * O + j) * V + a) * V + b];
  }
}

// T'[(i,a),(k,c)] = t2[i,k,a,c]
__global__ __launch_bounds__(256) void ring_perm(const double2* __restrict__ t2, double2* __restrict__ tr, int64_t O,
                                                 int64_t V) {
  const int64_t total = O * V * O * V;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = t;
    int64_t c = r % V; r /= V;
    int64_t k = r % O; r /= O;
    int64_t a = r % V;
    int64_t i = r / V;
    tr[t] = t2[((i * O + k) * V + a) * V + c];
  }
}

__global__ void pair_oo_kernel(const double2* __restrict__ oooo_p, const double2* __restrict__ F, int64_t n,
                               int64_t O, double2* __restrict__ hoo) {
  const int64_t Po = npair(O), total = Po * Po;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t i, j, k, l;
    pdec(t / Po, O, i, j);
    pdec(t % Po, O, k, l);
    double2 v = oooo_p ? oooo_p[t] : make_double2(0.0, 0.0);
    // -sum_m F_mj t2[i,m] - sum_m F_mi t2[m,j] expressed on t2[(k<l)]
    if (k == i) { double2 f = F[l * n + j]; v.x -= f.x; v.y -= f.y; }
    if (l == i) { double2 f = F[k * n + j]; v.x += f.x; v.y += f.y; }
    if (l == j) { double2 f = F[k * n + i]; v.x -= f.x; v.y -= f.y; }
    if (k == j) { double2 f = F[l * n + i]; v.x += f.x; v.y += f.y; }
    hoo[t] = v;
  }
}

struct EpiArgs {
  int64_t O, V, Po, Pv, n;
  const double2* drv;
  const double2* Rp;
  const double2* X;
  const double2* Y;
  const double2* U;
  const double2* Q;
  int64_t ldq;             // row stride of Q[((ab),j)][i] (O for one instance)
  const double2* F;
  const double2* t1;
  int fov;
  const double* eps;
  const double2* E;
  double c0r, c0i;
  const double2* c0p;      // when set: the driver weight from device memory
  const double2* t2;
  double2* t2n;
  double2* out2;
  int mode;
};

__device__ inline void acc(double2& s, double2 v) { s.x += v.x; s.y += v.y; }
__device__ inline void dec(double2& s, double2 v) { s.x -= v.x; s.y -= v.y; }
__device__ inline double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

__global__ __launch_bounds__(256) void epi_doubles(EpiArgs e) {
  const int64_t total = e.Po * e.Pv;
  const int64_t O = e.O, V = e.V, OV = O * V;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t Pij = t / e.Pv, Pab = t % e.Pv;
    int64_t i, j, a, b;
    pdec(Pij, O, i, j);
    pdec(Pab, V, a, b);
    const int64_t ijab = ((i * O + j) * V + a) * V + b;
    const int64_t jiab = ((j * O + i) * V + a) * V + b;
    const int64_t ijba = ((i * O + j) * V + b) * V + a;
    const int64_t jiba = ((j * O + i) * V + b) * V + a;
    double2 d = e.drv[ijab];
    const double2 c0 = e.c0p ? *e.c0p : make_double2(e.c0r, e.c0i);
    double2 s = make_double2(c0.x * d.x - c0.y * d.y, c0.x * d.y + c0.y * d.x);
    acc(s, e.Rp[t]);
    acc(s, e.X[(i * V + a) * OV + j * V + b]);
    dec(s, e.X[(j * V + a) * OV + i * V + b]);
    dec(s, e.X[(i * V + b) * OV + j * V + a]);
    acc(s, e.X[(j * V + b) * OV + i * V + a]);
    acc(s, e.Y[ijab]);
    dec(s, e.Y[ijba]);
    if (e.U) {
      dec(s, e.U[(a * V + b) * e.Po + Pij]);
      acc(s, e.U[(b * V + a) * e.Po + Pij]);
    }
    if (e.Q) {
      acc(s, e.Q[(Pab * O + j) * e.ldq + i]);
      dec(s, e.Q[(Pab * O + i) * e.ldq + j]);
    }
    if (e.fov) {
      const int64_t n = e.n;
      acc(s, cm(e.F[(O + b) * n + j], e.t1[i * V + a]));
      dec(s, cm(e.F[(O + b) * n + i], e.t1[j * V + a]));
      dec(s, cm(e.F[(O + a) * n + j], e.t1[i * V + b]));
      acc(s, cm(e.F[(O + a) * n + i], e.t1[j * V + b]));
    }
    if (e.mode == APY_MODE_RESIDUAL) {
      const double2 tv = e.t2[ijab];
      const double2 E = *e.E;
      double2 r = s;
      dec(r, cm(E, tv));
      const double D = e.eps[i] + e.eps[j] - e.eps[O + a] - e.eps[O + b];
      const double2 tn = make_double2(tv.x + r.x / D, tv.y + r.y / D);
      const double2 nr = make_double2(-r.x, -r.y), ntn = make_double2(-tn.x, -tn.y);
      e.out2[ijab] = r; e.out2[jiba] = r; e.out2[jiab] = nr; e.out2[ijba] = nr;
      e.t2n[ijab] = tn; e.t2n[jiba] = tn; e.t2n[jiab] = ntn; e.t2n[ijba] = ntn;
    } else {
      const double2 ns = make_double2(-s.x, -s.y);
      e.out2[ijab] = s; e.out2[jiba] = s; e.out2[jiab] = ns; e.out2[ijba] = ns;
    }
  }
}

__global__ void epi_singles(int64_t O, int64_t V, int64_t n, const double2* __restrict__ F,
                            const double* __restrict__ eps, const double2* E, double c0r, double c0i,
                            const double2* __restrict__ s1, const double2* t1, double2* t1n, double2* out1,
                            int mode, const double2* c0p) {
  const int64_t total = O * V;
  if (c0p) { c0r = c0p->x; c0i = c0p->y; }
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / V, a = t % V;
    double2 f = F[(O + a) * n + i];
    double2 s = s1[t];
    s.x += c0r * f.x - c0i * f.y;
    s.y += c0r * f.y + c0i * f.x;
    if (mode == APY_MODE_RESIDUAL) {
      const double2 tv = t1[t];
      const double2 Ev = *E;
      double2 r = s;
      r.x -= Ev.x * tv.x - Ev.y * tv.y;
      r.y -= Ev.x * tv.y + Ev.y * tv.x;
      const double D = eps[i] - eps[O + a];
      t1n[t] = make_double2(tv.x + r.x / D, tv.y + r.y / D);
      out1[t] = r;
    } else {
      out1[t] = s;
    }
  }
}

// Energy partial sums: terms l < OV are t1[i,a] f[i,O+a]; the rest 1/4 t2*oovv.
__global__ __launch_bounds__(256) void energy_partial(int64_t O, int64_t V, int64_t n, int singles,
                                                      const double2* __restrict__ F, const double2* __restrict__ oovv,
                                                      const double2* __restrict__ t, double2* partial) {
  const int64_t OV = O * V, L2 = O * O * V * V;
  double sr = 0.0, si = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (singles) {
    for (int64_t q = l; q < OV; q += stride) {
      const int64_t i = q / V, a = q % V;
      double2 x = t[q], f = F[i * n + O + a];
      sr += x.x * f.x - x.y * f.y;
      si += x.x * f.y + x.y * f.x;
    }
  }
  double qr = 0.0, qi = 0.0;
  for (int64_t q = l; q < L2; q += stride) {
    double2 x = t[OV + q], w = oovv[q];
    qr += x.x * w.x - x.y * w.y;
    qi += x.x * w.y + x.y * w.x;
  }
  sr += 0.25 * qr;
  si += 0.25 * qi;
  __shared__ double rs[256], is[256];
  rs[threadIdx.x] = sr;
  is[threadIdx.x] = si;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { rs[threadIdx.x] += rs[threadIdx.x + w]; is[threadIdx.x] += is[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(rs[0], is[0]);
}

__global__ void finalize_sum(const double2* __restrict__ partial, int np, double2* out) {
  __shared__ double rs[256], is[256];
  double sr = 0.0, si = 0.0;
  for (int q = threadIdx.x; q < np; q += blockDim.x) { sr += partial[q].x; si += partial[q].y; }
  rs[threadIdx.x] = sr;
  is[threadIdx.x] = si;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { rs[threadIdx.x] += rs[threadIdx.x + w]; is[threadIdx.x] += is[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(rs[0], is[0]);
}

constexpr int ENERGY_BLOCKS = 1024;

// Workspace carve-up of one sigma step.
struct SigmaWS {
  double2 *tp, *rp, *tr, *X, *Y, *U, *Q, *s1;
  void* gemm;
  size_t gemm_bytes;
  size_t total;
};

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// GEMM descriptors of one step (pointers filled where known).
struct StepGemms {
  apy_gemm_desc g[12];
  int count;
};

void fill(apy_gemm_desc& d, int64_t m, int64_t n, int64_t k, int al, const void* a, int64_t lda, int bl,
          const void* b, int64_t ldb, void* c, int64_t ldc, double alr, double ber) {
  d = apy_gemm_desc{};
  d.m = m; d.n = n; d.k = k;
  d.a_layout = al; d.b_layout = bl;
  d.alpha_re = alr; d.beta_re = ber;
  d.a = a; d.lda = lda; d.b = b; d.ldb = ldb; d.c = c; d.ldc = ldc;
  d.batch = 1;
}

// Build the GEMM list; with null data pointers it is used only to size the workspace.
void step_gemms(int64_t O, int64_t V, int singles, const apy_so_ints* I, const double2* F, const void* hoo,
                const double2* t, const SigmaWS& w, StepGemms& G) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, n = O + V;
  const double2* t1 = t;
  const double2* t2 = t ? t + OV : nullptr;
  G.count = 0;
  // PPL: R_p = T_p . vvvv_p^T
  fill(G.g[G.count++], Po, Pv, Pv, APY_MK, w.tp, Pv, APY_NK, I ? I->vvvv_p : nullptr, Pv, w.rp, Pv, 1.0, 0.0);
  // HHL + F_oo: R_p += Hoo . T_p
  fill(G.g[G.count++], Po, Pv, Po, APY_MK, hoo, Po, APY_KN, w.tp, Pv, w.rp, Pv, 1.0, 1.0);
  // ring: X = T' . ring^T
  fill(G.g[G.count++], OV, OV, OV, APY_MK, w.tr, OV, APY_NK, I ? I->ring : nullptr, OV, w.X, OV, 1.0, 0.0);
  // F_vv: Y[(ija),b] = sum_c t2[(ija),c] F[O+b,O+c]
  fill(G.g[G.count++], O * O * V, V, V, APY_MK, t2, V, APY_NK, F ? F + O * n + O : nullptr, n, w.Y, V, 1.0, 0.0);
  if (!singles) return;
  // U[a,(b,ij)] = sum_k t1[k,a] ovoo_p[k,(b,ij)]
  fill(G.g[G.count++], V, V * Po, O, APY_KM, t1, V, APY_KN, I ? I->ovoo_p : nullptr, V * Po, w.U, V * Po, 1.0, 0.0);
  // Q[((ab),j),i] = sum_c vvvo_p[((ab),j),c] t1[i,c]
  fill(G.g[G.count++], Pv * O, O, V, APY_MK, I ? I->vvvo_p : nullptr, V, APY_NK, t1, V, w.Q, O, 1.0, 0.0);
  // singles: s1 = t1 F_vv^T
  fill(G.g[G.count++], O, V, V, APY_MK, t1, V, APY_NK, F ? F + O * n + O : nullptr, n, w.s1, V, 1.0, 0.0);
  // s1 -= F_oo^T t1
  fill(G.g[G.count++], O, V, O, APY_KM, F, n, APY_KN, t1, V, w.s1, V, -1.0, 1.0);
  // s1 += ring . t1  (as a matrix-vector product)
  fill(G.g[G.count++], OV, 1, OV, APY_MK, I ? I->ring : nullptr, OV, APY_KN, t1, 1, w.s1, 1, 1.0, 1.0);
  // s1 += sum_jb F[j,b] t2[i,j,a,b] = -sum_j sum_b t2[j][(ia)][b] F[j][O+b]
  {
    apy_gemm_desc& d = G.g[G.count++];
    fill(d, OV, 1, V, APY_MK, t2, V, APY_KN, F ? F + O : nullptr, 1, w.s1, 1, -1.0, 1.0);
    d.batch = O; d.stride_a = OV * V; d.stride_b = n; d.stride_c = 0; d.reduce_batch = 1;
  }
  // s1 += 1/2 sum_(jcb) t2[i,(jcb)] vovv[a,(jcb)]
  fill(G.g[G.count++], O, V, OV * V, APY_MK, t2, OV * V, APY_NK, I ? I->vovv : nullptr, OV * V, w.s1, V, 0.5, 1.0);
  // s1 -= 1/2 sum_(kj) sum_b ooov[kj][i][b] t2[kj][a][b]
  {
    apy_gemm_desc& d = G.g[G.count++];
    fill(d, O, V, V, APY_MK, I ? I->ooov : nullptr, V, APY_NK, t2, V, w.s1, V, -0.5, 1.0);
    d.batch = O * O; d.stride_a = OV; d.stride_b = V * V; d.stride_c = 0; d.reduce_batch = 1;
  }
}

SigmaWS carve(int64_t O, int64_t V, int singles, void* base) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V;
  SigmaWS w{};
  size_t off = 0;
  char* b = (char*)base;
  auto take = [&](size_t elems) -> double2* {
    double2* p = b ? (double2*)(b + off) : nullptr;
    off += al256(elems * sizeof(double2));
    return p;
  };
  w.tp = take(Po * Pv);
  w.rp = take(Po * Pv);
  w.tr = take(OV * OV);
  w.X = take(OV * OV);
  w.Y = take(O * O * V * V);
  if (singles) {
    w.U = take(V * V * Po);
    w.Q = take(Pv * O * O);
    w.s1 = take(OV);
  }
  // GEMM split-K workspace: the largest any of the step's GEMMs asks for.
  StepGemms G;
  step_gemms(O, V, singles, nullptr, nullptr, nullptr, nullptr, w, G);
  size_t g = 0;
  for (int q = 0; q < G.count; ++q) {
    size_t need = apy_zgemm_ws_bytes(&G.g[q]);
    if (need > g) g = need;
  }
  w.gemm = b ? (void*)(b + off) : nullptr;
  w.gemm_bytes = g;
  off += al256(g);
  w.total = off;
  return w;
}

}  // namespace

extern "C" {

int apy_so_ints_sizes(int64_t O, int64_t V, int64_t sizes_out[9]) {
  if (O < 0 || V < 0 || !sizes_out) return APY_E_ARG;
  for (int k = 0; k < K_COUNT; ++k) sizes_out[k] = kind_size(k, O, V);
  return APY_OK;
}

int apy_so_ints_from_full(int64_t n, const void* A, const apy_so_ints* ints, void* stream) {
  if (!A || !ints || ints->O + ints->V != n) return APY_E_ARG;
  SrcFull src{(const double2*)A, n};
  return build_all(src, ints, 0, 0.0, (hipStream_t)stream);
}

int apy_so_ints_from_mo(int64_t m, const void* eri_mo, const apy_so_ints* ints, void* stream) {
  if (!eri_mo || !ints || ints->O + ints->V != 2 * m) return APY_E_ARG;
  SrcMO src{(const double2*)eri_mo, m};
  return build_all(src, ints, 0, 0.0, (hipStream_t)stream);
}

int apy_synth_so_ints(uint64_t seed, uint64_t tag, double sd, int accumulate, double scale, int real_only,
                      const apy_so_ints* ints, void* stream) {
  if (!ints) return APY_E_ARG;
  SrcSynth src{seed, tag, ints->O + ints->V, sd, real_only};
  return build_all(src, ints, accumulate, scale, (hipStream_t)stream);
}

int apy_so_fock_from_mo(int64_t m, const void* f_mo, void* f_so, void* stream) {
  if (m <= 0 || !f_mo || !f_so) return APY_E_ARG;
  so_fock_kernel<<<apy_grid1d(4 * m * m), 256, 0, (hipStream_t)stream>>>(m, (const double2*)f_mo, (double2*)f_so);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_cisd_so_pair_operator(int64_t O, int64_t V, const void* fock, const void* oooo_p, void* hoo, void* stream) {
  if (O < 0 || V < 0 || !fock || !hoo) return APY_E_ARG;
  const int64_t Po = npair(O);
  if (Po == 0) return APY_OK;
  pair_oo_kernel<<<apy_grid1d(Po * Po), 256, 0, (hipStream_t)stream>>>(
      (const double2*)oooo_p, (const double2*)fock, O + V, O, (double2*)hoo);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_cisd_so_workspace(int64_t O, int64_t V, int singles) {
  if (O < 0 || V < 0) return 0;
  return carve(O, V, singles ? 1 : 0, nullptr).total;
}

int apy_cisd_so_step(const apy_cisd_so_args* A, void* workspace, size_t ws_bytes, void* stream) {
  if (!A || !A->ints || !A->fock || !A->eps || !A->t || !A->out || !A->hoo) return APY_E_ARG;
  if (A->out == A->t || (A->mode == APY_MODE_RESIDUAL && (!A->t_new || A->t_new == A->out))) return APY_E_ARG;
  const int64_t O = A->O, V = A->V;
  if (O < 2 || V < 2 || A->ints->O != O || A->ints->V != V) return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  if (A->mode != APY_MODE_RESIDUAL && A->mode != APY_MODE_SIGMA) return APY_E_ARG;
  if (A->mode == APY_MODE_RESIDUAL && !A->energy) return APY_E_ARG;
  const apy_so_ints* I = A->ints;
  if (!I->vvvv_p || !I->ring || !I->vvoo_t) return APY_E_ARG;
  if (singles && (!I->ovoo_p || !I->vvvo_p || !I->vovv || !I->ooov)) return APY_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  SigmaWS w = carve(O, V, singles, workspace);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;

  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, n = O + V;
  const double2* t = (const double2*)A->t;
  double2* tn = (double2*)A->t_new;
  double2* out = (double2*)A->out;
  const double2* F = (const double2*)A->fock;

  pack_t2<<<apy_grid1d(Po * Pv, 1 << 20), 256, 0, st>>>(t + OV, w.tp, O, V);
  APY_LAUNCH_CHECK();
  ring_perm<<<apy_grid1d(OV * OV, 1 << 20), 256, 0, st>>>(t + OV, w.tr, O, V);
  APY_LAUNCH_CHECK();

  StepGemms G;
  step_gemms(O, V, singles, I, F, A->hoo, t, w, G);
  for (int q = 0; q < G.count; ++q) {
    // G.g[0] is the particle-particle ladder, the dominant kernel.
    int s = q == 0 ? apy_zgemm_launch(&G.g[q], w.gemm, w.gemm_bytes, st, (hipEvent_t)A->ev_begin,
                                      (hipEvent_t)A->ev_end)
                   : apy_zgemm_launch(&G.g[q], w.gemm, w.gemm_bytes, st);
    if (s != APY_OK) return s;
  }

  const double c0r = A->mode == APY_MODE_RESIDUAL ? 1.0 : A->c0_re;
  const double c0i = A->mode == APY_MODE_RESIDUAL ? 0.0 : A->c0_im;
  EpiArgs e;
  e.O = O; e.V = V; e.Po = Po; e.Pv = Pv; e.n = n;
  e.drv = (const double2*)I->vvoo_t;
  e.Rp = w.rp; e.X = w.X; e.Y = w.Y;
  e.U = singles ? w.U : nullptr;
  e.Q = singles ? w.Q : nullptr;
  e.ldq = O;
  e.F = F;
  e.t1 = t;
  e.fov = (singles && A->fov_coupling) ? 1 : 0;
  e.eps = A->eps;
  e.E = (const double2*)A->energy;
  e.c0r = c0r; e.c0i = c0i;
  const double2* c0p = A->mode == APY_MODE_SIGMA ? (const double2*)A->c0_dev : nullptr;
  e.c0p = c0p;
  e.t2 = t + OV;
  e.t2n = tn ? tn + OV : nullptr;
  e.out2 = out + OV;
  e.mode = A->mode;
  epi_doubles<<<apy_grid1d(Po * Pv, 1 << 20), 256, 0, st>>>(e);
  APY_LAUNCH_CHECK();
  if (singles) {
    epi_singles<<<apy_grid1d(OV), 256, 0, st>>>(O, V, n, F, A->eps, (const double2*)A->energy, c0r, c0i, w.s1, t, tn,
                                                out, A->mode, c0p);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

int apy_cisd_so_tune(const apy_cisd_so_args* A, void* workspace, size_t ws_bytes, int reps, void* stream) {
  if (!A || !A->ints || !A->fock || !A->t || !A->hoo) return APY_E_ARG;
  const int64_t O = A->O, V = A->V;
  if (O < 2 || V < 2 || A->ints->O != O || A->ints->V != V) return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  SigmaWS w = carve(O, V, singles, workspace);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  StepGemms G;
  step_gemms(O, V, singles, A->ints, (const double2*)A->fock, A->hoo, (const double2*)A->t, w, G);
  for (int q = 0; q < G.count; ++q) {
    if (!G.g[q].a || !G.g[q].b) continue;
    int s = apy_zgemm_tune_impl(&G.g[q], w.gemm, w.gemm_bytes, reps, (hipStream_t)stream, nullptr, nullptr, nullptr);
    if (s != APY_OK) return s;
  }
  return APY_OK;
}

int apy_cisd_so_energy(int64_t O, int64_t V, int singles, const void* fock, const void* oovv, const void* t,
                       void* energy, void* workspace, size_t ws_bytes, void* stream) {
  if (O < 0 || V < 0 || !fock || !oovv || !t || !energy) return APY_E_ARG;
  if (!workspace || ws_bytes < ENERGY_BLOCKS * sizeof(double2)) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  energy_partial<<<ENERGY_BLOCKS, 256, 0, st>>>(O, V, O + V, singles, (const double2*)fock, (const double2*)oovv,
                                                (const double2*)t, (double2*)workspace);
  APY_LAUNCH_CHECK();
  finalize_sum<<<1, 256, 0, st>>>((const double2*)workspace, ENERGY_BLOCKS, (double2*)energy);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batched sigma
// Sigma products of nb instances that share the integrals (the field-
// displaced copies of one system): the particle-particle ladder, the ring
// term, the vvvo.t1 term and the ring.t1 singles term each run as ONE GEMM
// with the instances stacked along M (or N), so the integral blocks are read
// once per batch; the Fock-dependent products and the epilogues run per
// instance.
namespace {

struct BatchWS {
  double2 *tp, *rp, *tr, *X, *Y, *U, *Q, *s1, *t1;
  void* gemm;
  size_t gemm_bytes, total;
};

void batch_gemms(int64_t O, int64_t V, int nb, int singles, const apy_so_ints* I, const double2* const* F,
                 const double2* const* hoo, const double2* const* t, const BatchWS& w, std::vector<apy_gemm_desc>& G) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, n = O + V, L2 = O * O * V * V;
  G.clear();
  apy_gemm_desc d;
  // folded over instances
  fill(d, nb * Po, Pv, Pv, APY_MK, w.tp, Pv, APY_NK, I ? I->vvvv_p : nullptr, Pv, w.rp, Pv, 1.0, 0.0);
  G.push_back(d);                                                              // [0] PPL
  fill(d, nb * OV, OV, OV, APY_MK, w.tr, OV, APY_NK, I ? I->ring : nullptr, OV, w.X, OV, 1.0, 0.0);
  G.push_back(d);                                                              // [1] ring
  if (singles) {
    fill(d, Pv * O, nb * O, V, APY_MK, I ? I->vvvo_p : nullptr, V, APY_NK, w.t1, V, w.Q, nb * O, 1.0, 0.0);
    G.push_back(d);                                                            // [2] Q
  }
  for (int k = 0; k < nb; ++k) {
    const double2* tk = t ? t[k] : nullptr;
    const double2* t1 = tk;
    const double2* t2 = tk ? tk + OV : nullptr;
    const double2* Fk = F ? F[k] : nullptr;
    fill(d, Po, Pv, Po, APY_MK, hoo ? hoo[k] : nullptr, Po, APY_KN, w.tp ? w.tp + k * Po * Pv : nullptr, Pv,
         w.rp ? w.rp + k * Po * Pv : nullptr, Pv, 1.0, 1.0);
    G.push_back(d);                                                            // HHL + F_oo
    fill(d, O * O * V, V, V, APY_MK, t2, V, APY_NK, Fk ? Fk + O * n + O : nullptr, n, w.Y ? w.Y + k * L2 : nullptr, V,
         1.0, 0.0);
    G.push_back(d);                                                            // F_vv
    if (!singles) continue;
    double2* s1 = w.s1 ? w.s1 + k * OV : nullptr;
    fill(d, V, V * Po, O, APY_KM, t1, V, APY_KN, I ? I->ovoo_p : nullptr, V * Po, w.U ? w.U + k * V * V * Po : nullptr,
         V * Po, 1.0, 0.0);
    G.push_back(d);                                                            // U
    fill(d, O, V, V, APY_MK, t1, V, APY_NK, Fk ? Fk + O * n + O : nullptr, n, s1, V, 1.0, 0.0);
    G.push_back(d);                                                            // s1 = t1 F_vv^T
    fill(d, O, V, O, APY_KM, Fk, n, APY_KN, t1, V, s1, V, -1.0, 1.0);
    G.push_back(d);                                                            // s1 -= F_oo^T t1
    fill(d, OV, 1, V, APY_MK, t2, V, APY_KN, Fk ? Fk + O : nullptr, 1, s1, 1, -1.0, 1.0);
    d.batch = O; d.stride_a = OV * V; d.stride_b = n; d.stride_c = 0; d.reduce_batch = 1;
    G.push_back(d);                                                            // s1 += F_jb t2
    fill(d, O, V, OV * V, APY_MK, t2, OV * V, APY_NK, I ? I->vovv : nullptr, OV * V, s1, V, 0.5, 1.0);
    G.push_back(d);                                                            // s1 += 1/2 t2 vovv
    fill(d, O, V, V, APY_MK, I ? I->ooov : nullptr, V, APY_NK, t2, V, s1, V, -0.5, 1.0);
    d.batch = O * O; d.stride_a = OV; d.stride_b = V * V; d.stride_c = 0; d.reduce_batch = 1;
    G.push_back(d);                                                            // s1 -= 1/2 ooov t2
  }
  if (singles) {
    fill(d, nb, OV, OV, APY_MK, w.t1, OV, APY_NK, I ? I->ring : nullptr, OV, w.s1, OV, 1.0, 1.0);
    G.push_back(d);                                                            // s1 += ring . t1 (all instances)
  }
}

BatchWS carve_batch(int64_t O, int64_t V, int nb, int singles, void* base) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V;
  BatchWS w{};
  size_t off = 0;
  char* b = (char*)base;
  auto take = [&](size_t elems) -> double2* {
    double2* p = b ? (double2*)(b + off) : nullptr;
    off += al256(elems * sizeof(double2));
    return p;
  };
  w.tp = take((size_t)nb * Po * Pv);
  w.rp = take((size_t)nb * Po * Pv);
  w.tr = take((size_t)nb * OV * OV);
  w.X = take((size_t)nb * OV * OV);
  w.Y = take((size_t)nb * O * O * V * V);
  if (singles) {
    w.U = take((size_t)nb * V * V * Po);
    w.Q = take((size_t)Pv * O * nb * O);
    w.s1 = take((size_t)nb * OV);
    w.t1 = take((size_t)nb * OV);
  }
  std::vector<apy_gemm_desc> G;
  BatchWS shape = w;
  if (!b) {  // sizing pass: non-null dummies so every descriptor is generated
    shape.tp = shape.rp = shape.tr = shape.X = shape.Y = shape.U = shape.Q = shape.s1 = shape.t1 = nullptr;
  }
  batch_gemms(O, V, nb, singles, nullptr, nullptr, nullptr, nullptr, shape, G);
  size_t g = 0;
  for (const apy_gemm_desc& d : G) g = std::max(g, apy_zgemm_ws_bytes(&d));
  w.gemm = b ? (void*)(b + off) : nullptr;
  w.gemm_bytes = g;
  off += al256(g);
  w.total = off;
  return w;
}

}  // namespace

extern "C" {

size_t apy_cisd_so_sigma_batch_workspace(int64_t O, int64_t V, int nb, int singles) {
  if (O < 2 || V < 2 || nb < 1) return 0;
  return carve_batch(O, V, nb, singles ? 1 : 0, nullptr).total;
}

int apy_cisd_so_sigma_batch(const apy_cisd_so_sigma_batch_args* A, void* workspace, size_t ws_bytes, void* stream) {
  if (!A || !A->ints || !A->fock || !A->hoo || !A->c0 || !A->t || !A->out) return APY_E_ARG;
  const int64_t O = A->O, V = A->V;
  const int nb = A->nb;
  if (O < 2 || V < 2 || nb < 1 || nb > 4096 || A->ints->O != O || A->ints->V != V) return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  const apy_so_ints* I = A->ints;
  if (!I->vvvv_p || !I->ring || !I->vvoo_t) return APY_E_ARG;
  if (singles && (!I->ovoo_p || !I->vvvo_p || !I->vovv || !I->ooov)) return APY_E_ARG;
  for (int k = 0; k < nb; ++k)
    if (!A->fock[k] || !A->hoo[k] || !A->t[k] || !A->out[k] || A->out[k] == A->t[k]) return APY_E_ARG;
  BatchWS w = carve_batch(O, V, nb, singles, workspace);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, n = O + V;
  std::vector<const double2*> F(nb), H(nb), T(nb);
  for (int k = 0; k < nb; ++k) {
    F[k] = (const double2*)A->fock[k];
    H[k] = (const double2*)A->hoo[k];
    T[k] = (const double2*)A->t[k];
    pack_t2<<<apy_grid1d(Po * Pv, 1 << 20), 256, 0, st>>>(T[k] + OV, w.tp + k * Po * Pv, O, V);
    APY_LAUNCH_CHECK();
    ring_perm<<<apy_grid1d(OV * OV, 1 << 20), 256, 0, st>>>(T[k] + OV, w.tr + k * OV * OV, O, V);
    APY_LAUNCH_CHECK();
    if (singles) APY_HIP_TRY(hipMemcpyAsync(w.t1 + k * OV, T[k], OV * sizeof(double2), hipMemcpyDeviceToDevice, st));
  }
  std::vector<apy_gemm_desc> G;
  batch_gemms(O, V, nb, singles, I, F.data(), H.data(), T.data(), w, G);
  for (size_t q = 0; q < G.size(); ++q) {
    const int s = q == 0 ? apy_zgemm_launch(&G[q], w.gemm, w.gemm_bytes, st, (hipEvent_t)A->ev_begin,
                                            (hipEvent_t)A->ev_end)
                         : apy_zgemm_launch(&G[q], w.gemm, w.gemm_bytes, st);
    if (s != APY_OK) return s;
  }
  for (int k = 0; k < nb; ++k) {
    double2* out = (double2*)A->out[k];
    EpiArgs e;
    e.O = O; e.V = V; e.Po = Po; e.Pv = Pv; e.n = n;
    e.drv = (const double2*)I->vvoo_t;
    e.Rp = w.rp + k * Po * Pv; e.X = w.X + k * OV * OV; e.Y = w.Y + k * O * O * V * V;
    e.U = singles ? w.U + k * V * V * Po : nullptr;
    e.Q = singles ? w.Q + k * O : nullptr;
    e.ldq = (int64_t)nb * O;
    e.F = F[k];
    e.t1 = T[k];
    e.fov = (singles && A->fov_coupling) ? 1 : 0;
    e.eps = nullptr;
    e.E = nullptr;
    e.c0r = A->c0[2 * k]; e.c0i = A->c0[2 * k + 1];
    e.c0p = nullptr;
    e.t2 = T[k] + OV;
    e.t2n = nullptr;
    e.out2 = out + OV;
    e.mode = APY_MODE_SIGMA;
    epi_doubles<<<apy_grid1d(Po * Pv, 1 << 20), 256, 0, st>>>(e);
    APY_LAUNCH_CHECK();
    if (singles) {
      epi_singles<<<apy_grid1d(OV), 256, 0, st>>>(O, V, n, F[k], nullptr, nullptr, e.c0r, e.c0i, w.s1 + k * OV,
                                                  T[k], nullptr, out, APY_MODE_SIGMA, nullptr);
      APY_LAUNCH_CHECK();
    }
  }
  return APY_OK;
}

}  // extern "C"
