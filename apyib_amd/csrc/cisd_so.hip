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
    case K_VOVV: return V * O * Pv;
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
    case K_VOVV: {  // [a][j][(c<b)] = <aj||cb>
      int64_t P = t % Pv; t /= Pv;
      j = t % O; a = t / O;
      pdec(P, V, c, b);
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
    case K_VOVV: return I->vovv_p;
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
// Every sigma works from the packed doubles T_p[(i<j)][(a<b)] (the packed
// input itself, or pack_t2 of a dense input); the dense images the GEMMs
// need are expanded from it with the antisymmetry signs.
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

__device__ inline int64_t pidx(int64_t p, int64_t q, int64_t N) { return prow(p, N) + q - p - 1; }

// T'[(i,a),(k,c)] = t2[i,k,a,c] from T_p: +T_p[(ik),(ac)] when i<k and a<c
// (or both reversed), minus when one pair is reversed, zero on i == k or a == c.
// One workgroup per (i, k, 32 x 32 tile of (a, c)): the packed row (ik) is read in two passes, the
// a < c half with c fastest and the a > c half with a fastest, so that both read along the packed
// row; the tile goes through LDS and out along c.
constexpr int RT = 32;
__global__ __launch_bounds__(256) void ring_perm_packed(const double2* __restrict__ tp, double2* __restrict__ tr,
                                                        int64_t O, int64_t V) {
  __shared__ double2 T[RT][RT + 1];
  const int64_t Pv = npair(V), OV = O * V, nt = (V + RT - 1) / RT, tiles = O * O * nt * nt;
  const int tx = threadIdx.x % RT, ty = threadIdx.x / RT;
  for (int64_t w = blockIdx.x; w < tiles; w += gridDim.x) {
    const int64_t tc = w % nt, ta = (w / nt) % nt, k = (w / (nt * nt)) % O, i = w / (nt * nt * O);
    const int64_t a0 = ta * RT, c0 = tc * RT;
    const double2* row = i != k ? tp + pidx(i < k ? i : k, i < k ? k : i, O) * Pv : nullptr;
    const double su = i < k ? 1.0 : -1.0;     // sign of the a < c half; the a > c half has -su
    __syncthreads();                            // the previous tile has been written out
#pragma unroll
    for (int r = 0; r < RT / 8; ++r) {
      const int ra = ty + 8 * r;
      // upper pass (c fastest): a < c, and every slot the lower pass does not cover (zeros)
      const int64_t a = a0 + ra, c = c0 + tx;
      if (!(row && a < V && c < V && a > c)) {
        double2 x = make_double2(0.0, 0.0);
        if (row && a < V && c < V && a < c) {
          const double2 y = row[pidx(a, c, V)];
          x = make_double2(su * y.x, su * y.y);
        }
        T[ra][tx] = x;
      }
      // lower pass (a fastest): a > c
      const int64_t al = a0 + tx, cl = c0 + ra;
      if (row && al < V && cl < V && al > cl) {
        const double2 y = row[pidx(cl, al, V)];
        T[tx][ra] = make_double2(-su * y.x, -su * y.y);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RT / 8; ++r) {
      const int64_t a = a0 + ty + 8 * r, c = c0 + tx;
      if (a < V && c < V) tr[(i * V + a) * OV + k * V + c] = T[ty + 8 * r][tx];
    }
  }
}

// Th[i][j][(c<b)] = t2[i,j,c,b] from T_p (zero rows for i == j).
__global__ __launch_bounds__(256) void half_pack_packed(const double2* __restrict__ tp, double2* __restrict__ th,
                                                        int64_t O, int64_t Pv) {
  const int64_t total = O * O * Pv;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t P = t % Pv, r = t / Pv;
    const int64_t j = r % O, i = r / O;
    double2 v = make_double2(0.0, 0.0);
    if (i != j) {
      const double2 w = tp[pidx(i < j ? i : j, i < j ? j : i, O) * Pv + P];
      v = i < j ? w : make_double2(-w.x, -w.y);
    }
    th[t] = v;
  }
}

// ovT[(i<j)][k][b] = ovoo_p[k][b][(i<j)] = <kb||ij>
__global__ __launch_bounds__(256) void ovoo_transpose(const double2* __restrict__ ov, double2* __restrict__ ovT,
                                                      int64_t O, int64_t V) {
  const int64_t Po = npair(O), total = Po * O * V;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t % V, r = t / V;
    const int64_t k = r % O, P = r / O;
    ovT[t] = ov[(k * V + b) * Po + P];
  }
}

// ooovT[k][i][(j,b)] = ooov[k][j][i][b] = <kj||ib>
__global__ __launch_bounds__(256) void ooov_transpose(const double2* __restrict__ x, double2* __restrict__ y,
                                                      int64_t O, int64_t V) {
  const int64_t total = O * O * O * V;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = t;
    const int64_t b = r % V; r /= V;
    const int64_t j = r % O; r /= O;
    const int64_t i = r % O;
    const int64_t k = r / O;
    y[t] = x[((k * O + j) * O + i) * V + b];
  }
}

__global__ __launch_bounds__(256) void ring_fock_kernel(int64_t O, int64_t V, const double2* __restrict__ F,
                                                        const double2* __restrict__ ring, double2* __restrict__ rf) {
  const int64_t n = O + V, OV = O * V, total = OV * OV;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = t / OV, col = t % OV;
    double2 v = ring[t];
    if (row / V == col / V) {
      const double2 f = F[(O + row % V) * n + O + col % V];
      v.x += 0.5 * f.x;
      v.y += 0.5 * f.y;
    }
    rf[t] = v;
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

__device__ inline void acc(double2& s, double2 v) { s.x += v.x; s.y += v.y; }
__device__ inline void dec(double2& s, double2 v) { s.x -= v.x; s.y -= v.y; }
__device__ inline double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline void cfma(double2& s, double2 a, double2 b) {
  s.x += a.x * b.x - a.y * b.y;
  s.y += a.x * b.y + a.y * b.x;
}

// ------------------------------------------------------------ doubles epilogue
// One workgroup per (i<j) pair and pair of 32-wide virtual tiles (A <= B);
// thread (ty, tx) owns a = a0 + ty + 8q (q = 0..3), b = b0 + tx, a < b.  It sums
//   c0 <ab||ij> + R_p + X[(ia),(jb)] - X[(ja),(ib)] - X[(ib),(ja)] + X[(jb),(ia)]
//   + Q_T[j][i][(ab)] - Q_T[i][j][(ab)] - U_ab + U_ba   (U_ab = sum_k t1[k,a] <kb||ij>)
//   + the f_bj t_ia coupling (Hermitian operator only)
// -- the reference's doubles terms of ci_wfn.py:319-334 regrouped (F_vv is in X).
// The two X blocks read transposed go through LDS so every global access is a
// row segment.  Output: packed s (sigma) or, in residual mode, r = s - E t and
// t + r/D; dense outputs write the four antisymmetric images, the two
// transposed ones through LDS again.
constexpr int ET = 32;   // virtual tile edge
constexpr int EKC = 8;   // occupied chunk of the in-kernel U contraction

struct Epi2Args {
  int64_t O, V, Po, Pv, n;
  int ntile;
  const double2* drv;      // vvoo_t [i][j][a][b] = <ab||ij> (NULL: no driver term)
  const double2* Rp;       // [(ij)][(ab)]
  const double2* X;        // [(i,a)][(j,b)]
  const double2* QT;       // Q_T[j][i][(ab)] at QT + j*qj + i*Pv (NULL: no T1 -> T2 terms)
  int64_t qj;
  const double2* ovT;      // [(ij)][k][b] = <kb||ij>
  const double2* t1;       // [i][a]
  const double2* F;        // [n][n]
  int fov;
  const double* eps;
  const double2* E;
  double c0r, c0i;
  const double2* c0p;
  const double2* Tp;       // packed t2 (residual mode)
  double2* out;            // packed [(ij)][(ab)] or dense [i][j][a][b]
  double2* tn;             // residual mode: t + r/D, same layout as out
  int mode;
  int packed;
};

__device__ inline void epi_dense_transposed(const Epi2Args& e, double2 (*L)[ET + 1], double2* base, int64_t i,
                                            int64_t j, int64_t a0, int64_t b0, int ty, int tx) {
  // base[i][j][b][a] = -v(a,b), base[j][i][b][a] = +v(a,b) for the tile's
  // values v(a,b) held in L[a - a0][b - b0]; rows b, lanes over a.
  const int64_t O = e.O, V = e.V;
  for (int q = 0; q < ET / 8; ++q) {
    const int lb = ty + 8 * q;
    const int64_t b = b0 + lb, a = a0 + tx;
    if (b < V && a < b) {
      const double2 v = L[tx][lb];
      base[((i * O + j) * V + b) * V + a] = make_double2(-v.x, -v.y);
      base[((j * O + i) * V + b) * V + a] = v;
    }
  }
}

__global__ __launch_bounds__(256) void epi2_doubles(Epi2Args e) {
  __shared__ double2 L[ET][ET + 1];
  __shared__ double2 sA[EKC][ET], sB[EKC][ET], oA[EKC][ET], oB[EKC][ET];
  const int tid = threadIdx.x, tx = tid & (ET - 1), ty = tid >> 5;
  const int64_t O = e.O, V = e.V, OV = O * V, Pv = e.Pv;
  const int ntp = e.ntile * (e.ntile + 1) / 2;
  const int64_t Pij = blockIdx.x / ntp;
  int tp = (int)(blockIdx.x % ntp), ta = 0;
  while (tp >= e.ntile - ta) { tp -= e.ntile - ta; ++ta; }
  const int tb = ta + tp;
  const int64_t a0 = (int64_t)ta * ET, b0 = (int64_t)tb * ET;
  int64_t i, j;
  pdec(Pij, O, i, j);

  // transposed X blocks: L[lb][la] = -X[(i,b),(j,a)] + X[(j,b),(i,a)]
  for (int q = 0; q < ET / 8; ++q) {
    const int lb = ty + 8 * q;
    const int64_t b = b0 + lb, a = a0 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (b < V && a < b) {
      const double2 x3 = e.X[(i * V + b) * OV + j * V + a], x4 = e.X[(j * V + b) * OV + i * V + a];
      v = make_double2(x4.x - x3.x, x4.y - x3.y);
    }
    L[lb][tx] = v;
  }

  double2 s[ET / 8];
  const double2 c0 = e.c0p ? *e.c0p : make_double2(e.c0r, e.c0i);
  for (int q = 0; q < ET / 8; ++q) {
    const int la = ty + 8 * q;
    const int64_t a = a0 + la, b = b0 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (b < V && a < b) {
      if (e.drv) v = cm(c0, e.drv[((i * O + j) * V + a) * V + b]);
      const int64_t Pab = pidx(a, b, V);
      acc(v, e.Rp[Pij * Pv + Pab]);
      acc(v, e.X[(i * V + a) * OV + j * V + b]);
      dec(v, e.X[(j * V + a) * OV + i * V + b]);
      if (e.QT) {
        acc(v, e.QT[j * e.qj + i * Pv + Pab]);
        dec(v, e.QT[i * e.qj + j * Pv + Pab]);
      }
      if (e.fov) {
        const int64_t n = e.n;
        cfma(v, e.F[(O + b) * n + j], e.t1[i * V + a]);
        cfma(v, e.F[(O + a) * n + i], e.t1[j * V + b]);
        const double2 m1 = cm(e.F[(O + b) * n + i], e.t1[j * V + a]);
        const double2 m2 = cm(e.F[(O + a) * n + j], e.t1[i * V + b]);
        v.x -= m1.x + m2.x;
        v.y -= m1.y + m2.y;
      }
    }
    s[q] = v;
  }

  // T1 -> T2 through <kb||ij>: -U_ab + U_ba, U_ab = sum_k t1[k,a] <kb||ij>
  if (e.QT) {
    double2 u[ET / 8];
    for (int q = 0; q < ET / 8; ++q) u[q] = make_double2(0.0, 0.0);
    const double2* ov = e.ovT + Pij * O * V;
    for (int64_t k0 = 0; k0 < O; k0 += EKC) {
      __syncthreads();
      {
        const int kk = ty;                     // 8 rows x 32 lanes: one element of each array per thread
        const int64_t k = k0 + kk, a = a0 + tx, b = b0 + tx;
        const bool kv = k < O;
        sA[kk][tx] = (kv && a < V) ? e.t1[k * V + a] : make_double2(0.0, 0.0);
        sB[kk][tx] = (kv && b < V) ? e.t1[k * V + b] : make_double2(0.0, 0.0);
        oA[kk][tx] = (kv && a < V) ? ov[k * V + a] : make_double2(0.0, 0.0);
        oB[kk][tx] = (kv && b < V) ? ov[k * V + b] : make_double2(0.0, 0.0);
      }
      __syncthreads();
      for (int q = 0; q < ET / 8; ++q) {
        const int la = ty + 8 * q;
        for (int kk = 0; kk < EKC; ++kk) {
          const double2 x = cm(sA[kk][la], oB[kk][tx]);   // t1[k,a] <kb||ij>
          const double2 y = cm(sB[kk][tx], oA[kk][la]);   // t1[k,b] <ka||ij>
          u[q].x += y.x - x.x;
          u[q].y += y.y - x.y;
        }
      }
    }
    for (int q = 0; q < ET / 8; ++q) acc(s[q], u[q]);
  }
  __syncthreads();                             // L complete
  for (int q = 0; q < ET / 8; ++q) acc(s[q], L[tx][ty + 8 * q]);

  double2 tnv[ET / 8];
  if (e.mode == APY_MODE_RESIDUAL) {
    const double2 E = *e.E;
    for (int q = 0; q < ET / 8; ++q) {
      const int64_t a = a0 + ty + 8 * q, b = b0 + tx;
      if (b < V && a < b) {
        const double2 tv = e.Tp[Pij * Pv + pidx(a, b, V)];
        dec(s[q], cm(E, tv));
        const double D = e.eps[i] + e.eps[j] - e.eps[O + a] - e.eps[O + b];
        tnv[q] = make_double2(tv.x + s[q].x / D, tv.y + s[q].y / D);
      } else {
        tnv[q] = make_double2(0.0, 0.0);
      }
    }
  }

  if (e.packed) {
    for (int q = 0; q < ET / 8; ++q) {
      const int64_t a = a0 + ty + 8 * q, b = b0 + tx;
      if (b < V && a < b) {
        const int64_t o = Pij * Pv + pidx(a, b, V);
        e.out[o] = s[q];
        if (e.mode == APY_MODE_RESIDUAL) e.tn[o] = tnv[q];
      }
    }
    return;
  }
  // dense: direct images [i][j][a][b] = v, [j][i][a][b] = -v
  for (int q = 0; q < ET / 8; ++q) {
    const int64_t a = a0 + ty + 8 * q, b = b0 + tx;
    if (b < V && a < b) {
      e.out[((i * O + j) * V + a) * V + b] = s[q];
      e.out[((j * O + i) * V + a) * V + b] = make_double2(-s[q].x, -s[q].y);
      if (e.mode == APY_MODE_RESIDUAL) {
        e.tn[((i * O + j) * V + a) * V + b] = tnv[q];
        e.tn[((j * O + i) * V + a) * V + b] = make_double2(-tnv[q].x, -tnv[q].y);
      }
    }
  }
  // transposed images through LDS
  __syncthreads();
  for (int q = 0; q < ET / 8; ++q) L[ty + 8 * q][tx] = s[q];
  __syncthreads();
  epi_dense_transposed(e, L, e.out, i, j, a0, b0, ty, tx);
  if (e.mode == APY_MODE_RESIDUAL) {
    __syncthreads();
    for (int q = 0; q < ET / 8; ++q) L[ty + 8 * q][tx] = tnv[q];
    __syncthreads();
    epi_dense_transposed(e, L, e.tn, i, j, a0, b0, ty, tx);
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

// Energy partial sums: terms l < OV are t1[i,a] f[i,O+a]; the rest 1/4 t2*oovv
// (dense) or t2p*oovv over (i<j, a<b) (packed).
__global__ __launch_bounds__(256) void energy_partial(int64_t O, int64_t V, int64_t n, int singles, int packed,
                                                      const double2* __restrict__ F, const double2* __restrict__ oovv,
                                                      const double2* __restrict__ t, double2* partial) {
  const int64_t OV = O * V;
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
  if (!oovv) {
    // one-electron part only: no doubles term
  } else if (packed) {
    const int64_t Pv = npair(V), L2 = npair(O) * Pv;
    for (int64_t q = l; q < L2; q += stride) {
      int64_t i, j, a, b;
      pdec(q / Pv, O, i, j);
      pdec(q % Pv, V, a, b);
      const double2 x = t[OV + q], w = oovv[((i * O + j) * V + a) * V + b];
      qr += x.x * w.x - x.y * w.y;
      qi += x.x * w.y + x.y * w.x;
    }
  } else {
    const int64_t L2 = O * O * V * V;
    for (int64_t q = l; q < L2; q += stride) {
      double2 x = t[OV + q], w = oovv[q];
      qr += x.x * w.x - x.y * w.y;
      qi += x.x * w.y + x.y * w.x;
    }
    qr *= 0.25;
    qi *= 0.25;
  }
  sr += qr;
  si += qi;
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

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

void fill(apy_gemm_desc& d, int64_t m, int64_t n, int64_t k, int al, const void* a, int64_t lda, int bl,
          const void* b, int64_t ldb, void* c, int64_t ldc, double alr, double ber) {
  d = apy_gemm_desc{};
  d.m = m; d.n = n; d.k = k;
  d.a_layout = al; d.b_layout = bl;
  d.alpha_re = alr; d.beta_re = ber;
  d.a = a; d.lda = lda; d.b = b; d.ldb = ldb; d.c = c; d.ldc = ldc;
  d.batch = 1;
}

// Workspace of nb instances (nb = 1: one sigma step).  Regions hold
// [nb][...] per buffer; tp only for dense inputs.
struct SigmaWS {
  double2 *tp, *rp, *tr, *X, *QT, *th, *ovT, *ooT, *s1, *t1;
  void* gemm;
  size_t gemm_bytes;
  size_t total;
};

// The GEMMs of a step over nb instances (nb = 1 for apy_cisd_so_step).
//   t[k]      the instance's vector ([t1 | t2...]; only t1 is read here)
//   F[k]      its Fock matrix, hoo[k] its pair operator
//   ring_f    (nb == 1 only, optional) the ring block with F_vv folded in
// With null pointers the list is built only to size the workspace.
//   g3        the PPL and ring products in the three-product complex form
void sigma_gemms(int64_t O, int64_t V, int nb, int singles, const apy_so_ints* I, const double2* const* F,
                 const void* const* hoo, const double2* const* t, const void* ring_f, const SigmaWS& w,
                 std::vector<apy_gemm_desc>& G, int g3 = 0) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, n = O + V;
  const int64_t PP = Po * Pv, R2 = OV * OV;
  G.clear();
  apy_gemm_desc d;
  auto off = [](const double2* p, int64_t o) -> const double2* { return p ? p + o : nullptr; };
  auto offw = [](double2* p, int64_t o) -> double2* { return p ? p + o : nullptr; };
  // [0] PPL over all instances: R_p = T_p . vvvv_p^T
  fill(d, nb * Po, Pv, Pv, APY_MK, w.tp, Pv, APY_NK, I ? I->vvvv_p : nullptr, Pv, w.rp, Pv, 1.0, 0.0);
  d.gauss3m = g3;
  G.push_back(d);
  // [1] ring over all instances: X = T' . ring^T (ring_f: F_vv folded in)
  fill(d, nb * OV, OV, OV, APY_MK, w.tr, OV, APY_NK, ring_f ? ring_f : (I ? I->ring : nullptr), OV, w.X, OV, 1.0, 0.0);
  d.gauss3m = g3;
  G.push_back(d);
  if (singles) {
    // Q_T[j][(k,i)][(ab)] = sum_c t1_k[i,c] <ab||cj>: batched over j, all instances' t1 stacked along M
    fill(d, nb * O, Pv, V, APY_MK, w.t1, V, APY_NK, I ? I->vvvo_p : nullptr, O * V, w.QT, Pv, 1.0, 0.0);
    d.batch = O; d.stride_a = 0; d.stride_b = V; d.stride_c = (int64_t)nb * O * Pv;
    G.push_back(d);
    // s1 += 1/2 sum_(jcb) t2 <aj||cb> = sum_j sum_(c<b) Th[(k,i)][(j,cb)] vovv_p[a][(j,cb)], all instances
    fill(d, nb * O, V, O * Pv, APY_MK, w.th, O * Pv, APY_NK, I ? I->vovv_p : nullptr, O * Pv, w.s1, V, 1.0, 0.0);
    G.push_back(d);
    // s1 += ring . t1 (all instances: M = nb)
    fill(d, nb, OV, OV, APY_MK, w.t1, OV, APY_NK, I ? I->ring : nullptr, OV, w.s1, OV, 1.0, 1.0);
    G.push_back(d);
  }
  for (int k = 0; k < nb; ++k) {
    const double2* Fk = F ? F[k] : nullptr;
    const double2* trk = off(w.tr, k * R2);
    double2* s1 = offw(w.s1, k * OV);
    // HHL + F_oo: R_p += Hoo . T_p
    fill(d, Po, Pv, Po, APY_MK, hoo ? hoo[k] : nullptr, Po, APY_KN, off(w.tp, k * PP), Pv, offw(w.rp, k * PP), Pv,
         1.0, 1.0);
    G.push_back(d);
    if (!ring_f) {
      // F_vv into X: X[(i,a,j)][b] += 1/2 sum_c T'[(i,a,j)][c] F[O+b][O+c]
      fill(d, O * V * O, V, V, APY_MK, trk, V, APY_NK, Fk ? Fk + O * n + O : nullptr, n, offw(w.X, k * R2), V, 0.5,
           1.0);
      G.push_back(d);
    }
    if (!singles) continue;
    const double2* t1 = t ? t[k] : nullptr;
    // s1 += t1 F_vv^T
    fill(d, O, V, V, APY_MK, t1, V, APY_NK, Fk ? Fk + O * n + O : nullptr, n, s1, V, 1.0, 1.0);
    G.push_back(d);
    // s1 -= F_oo^T t1
    fill(d, O, V, O, APY_KM, Fk, n, APY_KN, t1, V, s1, V, -1.0, 1.0);
    G.push_back(d);
    // s1[(ia)] += sum_j sum_b T'[(ia),(jb)] F[j][O+b]
    fill(d, OV, 1, V, APY_MK, trk, OV, APY_KN, Fk ? Fk + O : nullptr, 1, s1, 1, 1.0, 1.0);
    d.batch = O; d.stride_a = V; d.stride_b = n; d.stride_c = 0; d.reduce_batch = 1;
    G.push_back(d);
    // s1[i][a] -= 1/2 sum_k sum_(jb) <kj||ib> T'[(k,a),(j,b)]
    fill(d, O, V, OV, APY_MK, w.ooT, OV, APY_NK, trk, OV, s1, V, -0.5, 1.0);
    d.batch = O; d.stride_a = O * OV; d.stride_b = V * OV; d.stride_c = 0; d.reduce_batch = 1;
    G.push_back(d);
  }
}

// exact = 1 sizes the GEMM region for the plans that would run now instead of
// every plan the tuner may pick (the no-allocation entry point).
SigmaWS carve(int64_t O, int64_t V, int nb, int singles, int dense_in, const void* ring_f, void* base,
              int exact = 0) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V;
  SigmaWS w{};
  size_t off = 0;
  char* b = (char*)base;
  auto take = [&](size_t elems) -> double2* {
    double2* p = b ? (double2*)(b + off) : nullptr;
    off += al256(elems * sizeof(double2));
    return p;
  };
  if (dense_in) w.tp = take((size_t)nb * Po * Pv);
  w.rp = take((size_t)nb * Po * Pv);
  w.tr = take((size_t)nb * OV * OV);
  w.X = take((size_t)nb * OV * OV);
  if (singles) {
    w.QT = take((size_t)O * nb * O * Pv);
    w.th = take((size_t)nb * O * O * Pv);
    w.ovT = take((size_t)Po * O * V);
    w.ooT = take((size_t)O * O * OV);
    w.s1 = take((size_t)nb * OV);
    w.t1 = take((size_t)nb * OV);
  }
  // GEMM split-K workspace: the largest any of the step's GEMMs asks for.
  std::vector<apy_gemm_desc> G;
  sigma_gemms(O, V, nb, singles, nullptr, nullptr, nullptr, nullptr, ring_f, w, G);
  size_t g = 0;
  for (const apy_gemm_desc& d : G) g = std::max(g, exact ? apy_zgemm_plan_ws_bytes(&d) : apy_zgemm_ws_bytes(&d));
  w.gemm = b ? (void*)(b + off) : nullptr;
  w.gemm_bytes = g;
  off += al256(g);
  w.total = off;
  return w;
}

// Everything of a sigma over nb instances up to the epilogues: the packed
// doubles (tp: the caller's packed inputs when tp_in is given), the dense
// images, the transposed integral pieces and all GEMMs.
int sigma_body(int64_t O, int64_t V, int nb, int singles, const apy_so_ints* I, const double2* const* F,
               const void* const* hoo, const double2* const* t, int packed_in, const void* ring_f, SigmaWS& w,
               hipEvent_t ev0, hipEvent_t ev1, hipStream_t st, int g3) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, PP = Po * Pv;
  if (packed_in) {
    if (nb != 1) return APY_E_ARG;
    w.tp = (double2*)(t[0] + OV);      // T_p is the packed input itself
  }
  if (singles && nb == 1) w.t1 = (double2*)t[0];   // one instance: t1 read in place
  for (int k = 0; k < nb; ++k) {
    double2* tpk = w.tp + k * PP;
    if (!packed_in) {
      pack_t2<<<apy_grid1d(PP, 1 << 20), 256, 0, st>>>(t[k] + OV, tpk, O, V);
      APY_LAUNCH_CHECK();
    }
    ring_perm_packed<<<(unsigned)std::min<int64_t>(O * O * ((V + RT - 1) / RT) * ((V + RT - 1) / RT), 1 << 20), 256, 0, st>>>(
        tpk, w.tr + k * OV * OV, O, V);
    APY_LAUNCH_CHECK();
    if (singles) {
      half_pack_packed<<<apy_grid1d(O * O * Pv, 1 << 20), 256, 0, st>>>(tpk, w.th + k * O * O * Pv, O, Pv);
      APY_LAUNCH_CHECK();
      if (nb > 1)
        APY_HIP_TRY(hipMemcpyAsync(w.t1 + k * OV, t[k], OV * sizeof(double2), hipMemcpyDeviceToDevice, st));
    }
  }
  if (singles) {
    ovoo_transpose<<<apy_grid1d(Po * OV), 256, 0, st>>>((const double2*)I->ovoo_p, w.ovT, O, V);
    APY_LAUNCH_CHECK();
    ooov_transpose<<<apy_grid1d(O * O * OV), 256, 0, st>>>((const double2*)I->ooov, w.ooT, O, V);
    APY_LAUNCH_CHECK();
  }
  std::vector<apy_gemm_desc> G;
  sigma_gemms(O, V, nb, singles, I, F, hoo, t, ring_f, w, G, g3);
  for (size_t q = 0; q < G.size(); ++q) {
    const int s = q == 0 ? apy_zgemm_launch(&G[q], w.gemm, w.gemm_bytes, st, ev0, ev1)
                         : apy_zgemm_launch(&G[q], w.gemm, w.gemm_bytes, st);
    if (s != APY_OK) return s;
  }
  return APY_OK;
}

Epi2Args epi_args(int64_t O, int64_t V, int nb, int k, int singles, const apy_so_ints* I, const SigmaWS& w,
                  const double2* F, const double2* t, int fov) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V;
  Epi2Args e{};
  e.O = O; e.V = V; e.Po = Po; e.Pv = Pv; e.n = O + V;
  e.ntile = (int)((V + ET - 1) / ET);
  e.drv = (const double2*)I->vvoo_t;
  e.Rp = w.rp + k * Po * Pv;
  e.X = w.X + k * OV * OV;
  e.QT = singles ? w.QT + k * O * Pv : nullptr;
  e.qj = (int64_t)nb * O * Pv;
  e.ovT = singles ? w.ovT : nullptr;
  e.t1 = t;
  e.F = F;
  e.fov = (singles && fov) ? 1 : 0;
  e.Tp = w.tp + k * Po * Pv;
  return e;
}

inline unsigned epi_grid(int64_t O, int64_t V) {
  const int64_t nt = (V + ET - 1) / ET;
  return (unsigned)(npair(O) * (nt * (nt + 1) / 2));
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

int apy_cisd_so_ring_fock(int64_t O, int64_t V, const void* fock, const void* ring, void* ring_f, void* stream) {
  if (O < 1 || V < 1 || !fock || !ring || !ring_f || ring == ring_f) return APY_E_ARG;
  const int64_t OV = O * V;
  ring_fock_kernel<<<apy_grid1d(OV * OV, 1 << 20), 256, 0, (hipStream_t)stream>>>(
      O, V, (const double2*)fock, (const double2*)ring, (double2*)ring_f);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_cisd_so_workspace(int64_t O, int64_t V, int singles) {
  if (O < 0 || V < 0) return 0;
  // sized for a dense input and no ring_f: covers every call of this (O, V)
  return carve(O, V, 1, singles ? 1 : 0, 1, nullptr, nullptr).total;
}

}  // extern "C"

// The step on a caller's workspace; exact = 1 sizes the GEMM region for the
// plans that run now (see carve).  Shared by the entry points below and by
// other translation units (common.h).
int apy_cisd_so_step_ws(const apy_cisd_so_args* A, void* workspace, size_t ws_bytes, void* stream, int exact) {
  if (!A || !A->ints || !A->fock || !A->eps || !A->t || !A->out || !A->hoo) return APY_E_ARG;
  if (A->out == A->t || (A->mode == APY_MODE_RESIDUAL && (!A->t_new || A->t_new == A->out))) return APY_E_ARG;
  const int64_t O = A->O, V = A->V;
  if (O < 2 || V < 2 || A->ints->O != O || A->ints->V != V) return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  if (A->mode != APY_MODE_RESIDUAL && A->mode != APY_MODE_SIGMA) return APY_E_ARG;
  if (A->mode == APY_MODE_RESIDUAL && !A->energy) return APY_E_ARG;
  const int packed = A->packed ? 1 : 0;
  if (packed && A->mode != APY_MODE_SIGMA) return APY_E_ARG;
  const apy_so_ints* I = A->ints;
  if (!I->vvvv_p || !I->ring || !I->vvoo_t) return APY_E_ARG;
  if (singles && (!I->ovoo_p || !I->vvvo_p || !I->vovv_p || !I->ooov)) return APY_E_ARG;
  if (A->ring_f && (A->ring_f == I->ring || A->ring_f == A->out)) return APY_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  SigmaWS w = carve(O, V, 1, singles, !packed, A->ring_f, workspace, exact);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;

  const int64_t OV = O * V, n = O + V;
  const double2* t = (const double2*)A->t;
  double2* tn = (double2*)A->t_new;
  double2* out = (double2*)A->out;
  const double2* F = (const double2*)A->fock;
  const double2* Fk[1] = {F};
  const void* Hk[1] = {A->hoo};
  const double2* Tk[1] = {t};
  int s = sigma_body(O, V, 1, singles, I, Fk, Hk, Tk, packed, A->ring_f, w, (hipEvent_t)A->ev_begin,
                     (hipEvent_t)A->ev_end, st, A->gauss3m ? 1 : 0);
  if (s != APY_OK) return s;

  const double c0r = A->mode == APY_MODE_RESIDUAL ? 1.0 : A->c0_re;
  const double c0i = A->mode == APY_MODE_RESIDUAL ? 0.0 : A->c0_im;
  const double2* c0p = A->mode == APY_MODE_SIGMA ? (const double2*)A->c0_dev : nullptr;
  Epi2Args e = epi_args(O, V, 1, 0, singles, I, w, F, t, A->fov_coupling);
  e.eps = A->eps;
  e.E = (const double2*)A->energy;
  e.c0r = c0r; e.c0i = c0i; e.c0p = c0p;
  e.out = out + OV;
  e.tn = tn ? tn + OV : nullptr;
  e.mode = A->mode;
  e.packed = packed;
  epi2_doubles<<<epi_grid(O, V), 256, 0, st>>>(e);
  APY_LAUNCH_CHECK();
  if (singles) {
    epi_singles<<<apy_grid1d(OV), 256, 0, st>>>(O, V, n, F, A->eps, (const double2*)A->energy, c0r, c0i, w.s1, t, tn,
                                                out, A->mode, c0p);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

size_t apy_cisd_so_ws_exact(const apy_cisd_so_args* A) {
  if (!A || A->O < 2 || A->V < 2) return 0;
  return carve(A->O, A->V, 1, A->singles ? 1 : 0, A->packed ? 0 : 1, A->ring_f, nullptr, 1).total;
}

extern "C" {

int apy_cisd_so_step(const apy_cisd_so_args* A, void* workspace, size_t ws_bytes, void* stream) {
  return apy_cisd_so_step_ws(A, workspace, ws_bytes, stream, 0);
}

int apy_cisd_so_tune(const apy_cisd_so_args* A, void* workspace, size_t ws_bytes, int reps, void* stream) {
  if (!A || !A->ints || !A->fock || !A->t || !A->hoo) return APY_E_ARG;
  const int64_t O = A->O, V = A->V;
  if (O < 2 || V < 2 || A->ints->O != O || A->ints->V != V) return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  const int packed = A->packed ? 1 : 0;
  SigmaWS w = carve(O, V, 1, singles, !packed, A->ring_f, workspace);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  if (packed) w.tp = (double2*)((const double2*)A->t + O * V);
  const double2* Fk[1] = {(const double2*)A->fock};
  const void* Hk[1] = {A->hoo};
  const double2* Tk[1] = {(const double2*)A->t};
  std::vector<apy_gemm_desc> G;
  sigma_gemms(O, V, 1, singles, A->ints, Fk, Hk, Tk, A->ring_f, w, G, A->gauss3m ? 1 : 0);
  for (const apy_gemm_desc& d : G) {
    if (!d.a || !d.b) continue;
    int s = apy_zgemm_tune_impl(&d, w.gemm, w.gemm_bytes, reps, (hipStream_t)stream, nullptr, nullptr, nullptr);
    if (s != APY_OK) return s;
  }
  return APY_OK;
}

static int energy_impl(int64_t O, int64_t V, int singles, int packed, const void* fock, const void* oovv,
                       const void* t, void* energy, void* workspace, size_t ws_bytes, void* stream) {
  if (O < 0 || V < 0 || !fock || !oovv || !t || !energy) return APY_E_ARG;
  if (!workspace || ws_bytes < ENERGY_BLOCKS * sizeof(double2)) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  energy_partial<<<ENERGY_BLOCKS, 256, 0, st>>>(O, V, O + V, singles, packed, (const double2*)fock,
                                                (const double2*)oovv, (const double2*)t, (double2*)workspace);
  APY_LAUNCH_CHECK();
  finalize_sum<<<1, 256, 0, st>>>((const double2*)workspace, ENERGY_BLOCKS, (double2*)energy);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_cisd_so_energy(int64_t O, int64_t V, int singles, const void* fock, const void* oovv, const void* t,
                       void* energy, void* workspace, size_t ws_bytes, void* stream) {
  return energy_impl(O, V, singles, 0, fock, oovv, t, energy, workspace, ws_bytes, stream);
}

int apy_cisd_so_energy_packed(int64_t O, int64_t V, int singles, const void* fock, const void* oovv,
                              const void* t, void* energy, void* workspace, size_t ws_bytes, void* stream) {
  if (O < 2 || V < 2) return APY_E_ARG;
  return energy_impl(O, V, singles, 1, fock, oovv, t, energy, workspace, ws_bytes, stream);
}

// ------------------------------------------------------------ batched sigma
// Sigma products of nb instances that share the integrals (the field-
// displaced copies of one system): the particle-particle ladder, the ring
// term, the vvvo.t1, t2.vovv and ring.t1 terms each run as ONE GEMM with the
// instances stacked along M, so the integral blocks are read once per batch;
// the Fock-dependent products and the epilogues run per instance.
size_t apy_cisd_so_sigma_batch_workspace(int64_t O, int64_t V, int nb, int singles) {
  if (O < 2 || V < 2 || nb < 1) return 0;
  return carve(O, V, nb, singles ? 1 : 0, 1, nullptr, nullptr).total;
}

int apy_cisd_so_sigma_batch(const apy_cisd_so_sigma_batch_args* A, void* workspace, size_t ws_bytes, void* stream) {
  if (!A || !A->ints || !A->fock || !A->hoo || !A->c0 || !A->t || !A->out) return APY_E_ARG;
  const int64_t O = A->O, V = A->V;
  const int nb = A->nb;
  if (O < 2 || V < 2 || nb < 1 || nb > 4096 || A->ints->O != O || A->ints->V != V) return APY_E_ARG;
  const int singles = A->singles ? 1 : 0;
  const apy_so_ints* I = A->ints;
  if (!I->vvvv_p || !I->ring || !I->vvoo_t) return APY_E_ARG;
  if (singles && (!I->ovoo_p || !I->vvvo_p || !I->vovv_p || !I->ooov)) return APY_E_ARG;
  for (int k = 0; k < nb; ++k)
    if (!A->fock[k] || !A->hoo[k] || !A->t[k] || !A->out[k] || A->out[k] == A->t[k]) return APY_E_ARG;
  SigmaWS w = carve(O, V, nb, singles, 1, nullptr, workspace);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t OV = O * V, n = O + V;
  std::vector<const double2*> F(nb), T(nb);
  std::vector<const void*> H(nb);
  for (int k = 0; k < nb; ++k) {
    F[k] = (const double2*)A->fock[k];
    H[k] = A->hoo[k];
    T[k] = (const double2*)A->t[k];
  }
  int s = sigma_body(O, V, nb, singles, I, F.data(), H.data(), T.data(), 0, nullptr, w, (hipEvent_t)A->ev_begin,
                     (hipEvent_t)A->ev_end, st, A->gauss3m ? 1 : 0);
  if (s != APY_OK) return s;
  for (int k = 0; k < nb; ++k) {
    double2* out = (double2*)A->out[k];
    Epi2Args e = epi_args(O, V, nb, k, singles, I, w, F[k], T[k], A->fov_coupling);
    e.c0r = A->c0[2 * k]; e.c0i = A->c0[2 * k + 1];
    e.out = out + OV;
    e.mode = APY_MODE_SIGMA;
    e.packed = 0;
    epi2_doubles<<<epi_grid(O, V), 256, 0, st>>>(e);
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

// ------------------------------------------------------------ one-electron sigma
// The Fock-linear part L(F) of the Hermitian CISD sigma on a packed vector
// [c1 | c2 over (i<j, a<b) | c0]: the terms of ci_wfn.py:309-334 that carry
// F_SO (r_T1: F_vo c0, -F_oo^T c1, c1 F_vv^T, F_ov.c2; r_T2: the two F_vv and
// the two F_oo contractions) plus the Hermitian f_ov couplings (c0 row
// F_ov.c1, doubles P(ij)P(ab) f_bj c_ia).  For a one-electron perturbation
// dF of a Hamiltonian whose two-electron integrals stay fixed (the field-
// displaced copies of fin_diff.py:335-370), H(F0 + dF) = H(F0) + L(dF)
// exactly, so the sigma of every perturbed instance follows from one full
// sigma and this O^2 V^3 product.  The products are the step's own GEMMs
// (ring permutation, 1/2 T' F_vv^T into X, pair operator of F alone) and its
// epilogues with the two-electron pieces left out.
namespace {

struct FockWS {
  double2 *tr, *X, *rp, *s1, *part;
  void* gemm;
  size_t gemm_bytes;
  size_t total;
};

void fock_gemms(int64_t O, int64_t V, const double2* F, const void* hoo, const double2* t, const FockWS& w,
                std::vector<apy_gemm_desc>& G) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, n = O + V;
  G.clear();
  apy_gemm_desc d;
  const double2* tp = t ? t + OV : nullptr;
  // X[(i,a,j)][b] = 1/2 sum_c T'[(i,a,j)][c] F[O+b][O+c]
  fill(d, O * V * O, V, V, APY_MK, w.tr, V, APY_NK, F ? F + O * n + O : nullptr, n, w.X, V, 0.5, 0.0);
  G.push_back(d);
  // R_p = Hoo(F) . T_p  (the F_oo terms over occupied pairs)
  fill(d, Po, Pv, Po, APY_MK, hoo, Po, APY_KN, tp, Pv, w.rp, Pv, 1.0, 0.0);
  G.push_back(d);
  // s1 = c1 F_vv^T - F_oo^T c1 + sum_jb T'[(ia),(jb)] F[j][O+b]
  fill(d, O, V, V, APY_MK, t, V, APY_NK, F ? F + O * n + O : nullptr, n, w.s1, V, 1.0, 0.0);
  G.push_back(d);
  fill(d, O, V, O, APY_KM, F, n, APY_KN, t, V, w.s1, V, -1.0, 1.0);
  G.push_back(d);
  fill(d, OV, 1, V, APY_MK, w.tr, OV, APY_KN, F ? F + O : nullptr, 1, w.s1, 1, 1.0, 1.0);
  d.batch = O; d.stride_a = V; d.stride_b = n; d.stride_c = 0; d.reduce_batch = 1;
  G.push_back(d);
}

FockWS carve_fock(int64_t O, int64_t V, void* base) {
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V;
  FockWS w{};
  size_t off = 0;
  char* b = (char*)base;
  auto take = [&](size_t elems) -> double2* {
    double2* p = b ? (double2*)(b + off) : nullptr;
    off += al256(elems * sizeof(double2));
    return p;
  };
  w.tr = take((size_t)OV * OV);
  w.X = take((size_t)OV * OV);
  w.rp = take((size_t)Po * Pv);
  w.s1 = take((size_t)OV);
  w.part = take((size_t)ENERGY_BLOCKS);
  std::vector<apy_gemm_desc> G;
  fock_gemms(O, V, nullptr, nullptr, nullptr, w, G);
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

size_t apy_cisd_so_fock_sigma_workspace(int64_t O, int64_t V) {
  if (O < 2 || V < 2) return 0;
  return carve_fock(O, V, nullptr).total;
}

int apy_cisd_so_fock_sigma(const apy_cisd_so_fock_sigma_args* A, void* workspace, size_t ws_bytes, void* stream) {
  if (!A || !A->fock || !A->hoo || !A->t || !A->out) return APY_E_ARG;
  const int64_t O = A->O, V = A->V;
  const int nf = A->nf;
  if (O < 2 || V < 2 || nf < 1 || nf > 4096) return APY_E_ARG;
  for (int f = 0; f < nf; ++f)
    if (!A->fock[f] || !A->hoo[f] || !A->out[f] || A->out[f] == A->t) return APY_E_ARG;
  FockWS w = carve_fock(O, V, workspace);
  if (!workspace || ws_bytes < w.total) return APY_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t Po = npair(O), Pv = npair(V), OV = O * V, n = O + V, PP = Po * Pv;
  const double2* t = (const double2*)A->t;
  const double2* c0p = t + OV + PP;
  ring_perm_packed<<<(unsigned)std::min<int64_t>(O * O * ((V + RT - 1) / RT) * ((V + RT - 1) / RT), 1 << 20), 256, 0, st>>>(
      t + OV, w.tr, O, V);
  APY_LAUNCH_CHECK();
  std::vector<apy_gemm_desc> G;
  for (int f = 0; f < nf; ++f) {
    const double2* F = (const double2*)A->fock[f];
    double2* out = (double2*)A->out[f];
    fock_gemms(O, V, F, A->hoo[f], t, w, G);
    for (const apy_gemm_desc& d : G) {
      const int s = apy_zgemm_launch(&d, w.gemm, w.gemm_bytes, st);
      if (s != APY_OK) return s;
    }
    Epi2Args e{};
    e.O = O; e.V = V; e.Po = Po; e.Pv = Pv; e.n = n;
    e.ntile = (int)((V + ET - 1) / ET);
    e.drv = nullptr;
    e.Rp = w.rp;
    e.X = w.X;
    e.QT = nullptr;
    e.qj = 0;
    e.ovT = nullptr;
    e.t1 = t;
    e.F = F;
    e.fov = 1;
    e.c0p = c0p;
    e.Tp = t + OV;
    e.out = out + OV;
    e.mode = APY_MODE_SIGMA;
    e.packed = 1;
    epi2_doubles<<<epi_grid(O, V), 256, 0, st>>>(e);
    APY_LAUNCH_CHECK();
    epi_singles<<<apy_grid1d(OV), 256, 0, st>>>(O, V, n, F, nullptr, nullptr, 0.0, 0.0, w.s1, t, nullptr, out,
                                                APY_MODE_SIGMA, c0p);
    APY_LAUNCH_CHECK();
    energy_partial<<<ENERGY_BLOCKS, 256, 0, st>>>(O, V, n, 1, 1, F, nullptr, t, w.part);
    APY_LAUNCH_CHECK();
    finalize_sum<<<1, 256, 0, st>>>(w.part, ENERGY_BLOCKS, out + OV + PP);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

}  // extern "C"
