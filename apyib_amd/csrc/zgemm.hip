// Complex128 GEMM on CDNA4 FP64 matrix cores.
//
// One complex multiply-accumulate is four real MFMA products on separate
// real/imaginary planes (re += ar*br - ai*bi, im += ar*bi + ai*br), i.e.
// 8 real flops per complex MAC, executed with v_mfma_f64_16x16x4f64.
//
// Tiling: 256 threads = 4 waves arranged 2x2; each wave owns a WM x WN
// complex output tile built from 16x16 MFMA blocks.  K is staged through
// LDS BK complex elements at a time, split into separate real and imaginary
// planes.  Global loads for stage s+1 are issued into registers before the
// MFMAs of stage s run (software pipelining with a single LDS buffer).
//
// MFMA f64 16x16x4 lane layout (checked on gfx950):
//   A operand: lane l holds A[l%16][l/16]       (16 x 4)
//   B operand: lane l holds B[l/16][l%16]       (4 x 16)
//   D result : lane l, element r holds D[l/16 + 4r][l%16]
//
// Work decomposition: the flat block id is remapped so that consecutive
// logical tiles land on the same XCD (workgroups are dealt round-robin
// over the 8 XCDs), with the M tile fastest: the blocks that share a B
// tile share that XCD's L2.
#include "common.h"
#include <algorithm>

namespace {

constexpr int BK = 16;        // complex K elements per LDS stage
constexpr int BKP = BK + 1;   // padded LDS row length (doubles) for k-contiguous tiles
constexpr int NTHR = 256;

struct GemmArgs {
  int64_t m, n, k;
  const double2* a; int64_t lda, sa;
  const double2* b; int64_t ldb, sb;
  double2* c; int64_t ldc, sc;
  double alr, ali, ber, bei;
  int conj_a, conj_b, use_beta;
  int64_t batch;
  int split;
  int64_t kchunk;
  int64_t mtiles, ntiles;
  int64_t total;
  double2* ws;      // partial sums [npart][m][n] or nullptr
};

__device__ inline int64_t xcd_remap(int64_t bid, int64_t total) {
  int64_t xcd = bid & 7, slot = bid >> 3;
  int64_t per = total >> 3, rem = total & 7;
  return (xcd < rem) ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

template <int BM, int BN, int WM, int WN, bool A_MK, bool B_NK>
__global__ __launch_bounds__(NTHR) void zgemm_mfma(GemmArgs p) {
  constexpr int RB = WM / 16, CB = WN / 16;
  constexpr int A_PER = BM * BK / NTHR;
  constexpr int B_PER = BN * BK / NTHR;
  constexpr int B_LDS = B_NK ? BN * BKP : BK * BN;
  static_assert(BM == 2 * WM && BN == 2 * WN, "2x2 wave grid");
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small");

  __shared__ double As_re[BM * BKP], As_im[BM * BKP];
  __shared__ double Bs_re[B_LDS], Bs_im[B_LDS];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  int64_t logical = xcd_remap(blockIdx.x, p.total);
  const int64_t mt = logical % p.mtiles; logical /= p.mtiles;
  const int64_t nt = logical % p.ntiles; logical /= p.ntiles;
  const int s = (int)(logical % p.split);
  const int64_t bidx = logical / p.split;

  const int64_t m0 = mt * BM, n0 = nt * BN;
  const int64_t kbeg = (int64_t)s * p.kchunk;
  const int64_t kend = min(p.k, kbeg + p.kchunk);

  const double2* A = p.a + bidx * p.sa;
  const double2* B = p.b + bidx * p.sb;

  v4d acc_re[RB][CB], acc_im[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) { acc_re[i][j] = (v4d){0, 0, 0, 0}; acc_im[i][j] = (v4d){0, 0, 0, 0}; }

  double2 ra[A_PER], rb[B_PER];
  const double csa = p.conj_a ? -1.0 : 1.0, csb = p.conj_b ? -1.0 : 1.0;

  auto load = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      int idx = tid + NTHR * q;
      int r, kk;
      if (A_MK) { r = idx / BK; kk = idx % BK; } else { kk = idx / BM; r = idx % BM; }
      int64_t gm = m0 + r, gk = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (gm < p.m && gk < kend) v = A_MK ? A[gm * p.lda + gk] : A[gk * p.lda + gm];
      v.y *= csa;
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      int idx = tid + NTHR * q;
      int c, kk;
      if (B_NK) { c = idx / BK; kk = idx % BK; } else { kk = idx / BN; c = idx % BN; }
      int64_t gn = n0 + c, gk = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (gn < p.n && gk < kend) v = B_NK ? B[gn * p.ldb + gk] : B[gk * p.ldb + gn];
      v.y *= csb;
      rb[q] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      int idx = tid + NTHR * q;
      int r, kk;
      if (A_MK) { r = idx / BK; kk = idx % BK; } else { kk = idx / BM; r = idx % BM; }
      As_re[r * BKP + kk] = ra[q].x;
      As_im[r * BKP + kk] = ra[q].y;
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      int idx = tid + NTHR * q;
      int c, kk;
      if (B_NK) { c = idx / BK; kk = idx % BK; } else { kk = idx / BN; c = idx % BN; }
      int o = B_NK ? c * BKP + kk : kk * BN + c;
      Bs_re[o] = rb[q].x;
      Bs_im[o] = rb[q].y;
    }
  };

  if (kbeg < kend) load(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();  // previous stage's MFMA reads are done
    store();
    __syncthreads();
    if (k0 + BK < kend) load(k0 + BK);  // in flight while the MFMAs below run
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      double are[RB], aim[RB], naim[RB], bre[CB], bim[CB];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        int row = wm * WM + i * 16 + (lane & 15);
        are[i] = As_re[row * BKP + kk];
        aim[i] = As_im[row * BKP + kk];
        naim[i] = -aim[i];
      }
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        int col = wn * WN + j * 16 + (lane & 15);
        int o = B_NK ? col * BKP + kk : kk * BN + col;
        bre[j] = Bs_re[o];
        bim[j] = Bs_im[o];
      }
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[i], bre[j], acc_re[i][j], 0, 0, 0);
          acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[i], bim[j], acc_im[i][j], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(naim[i], bim[j], acc_re[i][j], 0, 0, 0);
          acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[i], bre[j], acc_im[i][j], 0, 0, 0);
        }
    }
  }

  // Epilogue.
  double2* Cb = p.c + bidx * p.sc;
  double2* W = p.ws ? p.ws + (bidx * p.split + s) * (p.m * p.n) : nullptr;
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t row = m0 + wm * WM + i * 16 + (lane >> 4) + 4 * r;
        int64_t col = n0 + wn * WN + j * 16 + (lane & 15);
        if (row < p.m && col < p.n) {
          double vr = acc_re[i][j][r], vi = acc_im[i][j][r];
          double o_r = p.alr * vr - p.ali * vi, o_i = p.alr * vi + p.ali * vr;
          if (W) {
            W[row * p.n + col] = make_double2(o_r, o_i);
          } else {
            double2* cp = Cb + row * p.ldc + col;
            if (p.use_beta) {
              double2 old = *cp;
              o_r += p.ber * old.x - p.bei * old.y;
              o_i += p.ber * old.y + p.bei * old.x;
            }
            *cp = make_double2(o_r, o_i);
          }
        }
      }
}

// C_g = beta*C_g + sum over `group` consecutive partials of ws.
__global__ __launch_bounds__(256) void zgemm_reduce(const double2* __restrict__ ws, int64_t group, int64_t ngroups,
                                                     int64_t m, int64_t n, double2* c, int64_t ldc, int64_t sc,
                                                     double ber, double bei, int use_beta) {
  const int64_t mn = m * n;
  const int64_t total = ngroups * mn;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t g = t / mn, e = t % mn;
    int64_t row = e / n, col = e % n;
    const double2* w = ws + g * group * mn + e;
    double sr = 0.0, si = 0.0;
    for (int64_t q = 0; q < group; ++q) { double2 v = w[q * mn]; sr += v.x; si += v.y; }
    double2* cp = c + g * sc + row * ldc + col;
    if (use_beta) {
      double2 old = *cp;
      sr += ber * old.x - bei * old.y;
      si += ber * old.y + bei * old.x;
    }
    *cp = make_double2(sr, si);
  }
}

struct Plan {
  int cfg;        // 0: 64x64 tiles, 1: 32x64 tiles
  int bm, bn;
  int split;
  int64_t kchunk;
  int64_t mtiles, ntiles, total;
  bool use_ws;
  int64_t npart;
};

Plan make_plan(const apy_gemm_desc* d) {
  Plan pl;
  pl.cfg = (d->m <= 32) ? 1 : 0;
  pl.bm = pl.cfg == 1 ? 32 : 64;
  pl.bn = 64;
  pl.mtiles = (d->m + pl.bm - 1) / pl.bm;
  pl.ntiles = (d->n + pl.bn - 1) / pl.bn;
  int64_t tiles = pl.mtiles * pl.ntiles * (d->reduce_batch ? d->batch : d->batch);
  int split = d->split_k;
  if (split <= 0) {
    // Aim for >= ~2048 resident-work units on 256 CUs while keeping each
    // K slice at least 8 stages long.
    split = 1;
    int64_t max_split = std::max<int64_t>(1, d->k / (8 * BK));
    while (tiles * split < 2048 && split < max_split && split < 64) ++split;
  }
  int64_t kchunk = (d->k + split - 1) / split;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  if (kchunk < BK) kchunk = BK;
  split = (int)((d->k + kchunk - 1) / kchunk);
  if (split < 1) split = 1;
  pl.split = split;
  pl.kchunk = kchunk;
  pl.total = pl.mtiles * pl.ntiles * split * d->batch;
  pl.use_ws = (split > 1) || d->reduce_batch;
  pl.npart = split * d->batch;
  return pl;
}

template <int BM, int BN, int WM, int WN>
void launch_cfg(const GemmArgs& a, int a_layout, int b_layout, hipStream_t s) {
  dim3 grid((unsigned)a.total), block(NTHR);
  if (a_layout == APY_MK && b_layout == APY_NK) zgemm_mfma<BM, BN, WM, WN, true, true><<<grid, block, 0, s>>>(a);
  else if (a_layout == APY_MK) zgemm_mfma<BM, BN, WM, WN, true, false><<<grid, block, 0, s>>>(a);
  else if (b_layout == APY_NK) zgemm_mfma<BM, BN, WM, WN, false, true><<<grid, block, 0, s>>>(a);
  else zgemm_mfma<BM, BN, WM, WN, false, false><<<grid, block, 0, s>>>(a);
}

}  // namespace

size_t apy_zgemm_ws_bytes(const apy_gemm_desc* d) {
  if (!d || d->m <= 0 || d->n <= 0 || d->k <= 0 || d->batch <= 0) return 0;
  Plan pl = make_plan(d);
  if (!pl.use_ws) return 0;
  return (size_t)pl.npart * (size_t)d->m * (size_t)d->n * sizeof(double2);
}

int apy_zgemm_launch(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, hipStream_t s) {
  if (!d) return APY_E_ARG;
  if (d->m < 0 || d->n < 0 || d->k < 0 || d->batch < 1) return APY_E_ARG;
  if (d->a_layout != APY_MK && d->a_layout != APY_KM) return APY_E_ARG;
  if (d->b_layout != APY_KN && d->b_layout != APY_NK) return APY_E_ARG;
  if (d->m == 0 || d->n == 0) return APY_OK;
  if (!d->c || ((uintptr_t)d->c & 15)) return APY_E_ARG;
  if (d->k > 0 && (!d->a || !d->b || ((uintptr_t)d->a & 15) || ((uintptr_t)d->b & 15))) return APY_E_ARG;
  const int use_beta = (d->beta_re != 0.0 || d->beta_im != 0.0) ? 1 : 0;
  if (d->k == 0) {
    // An empty sum contributes zero: C = beta*C.
    int64_t nb = d->reduce_batch ? 1 : d->batch;
    int64_t tot = nb * d->m * d->n;
    // Reuse the reduce kernel with an empty group.
    zgemm_reduce<<<apy_grid1d(tot), 256, 0, s>>>(nullptr, 0, nb, d->m, d->n, (double2*)d->c, d->ldc, d->stride_c,
                                                 d->beta_re, d->beta_im, use_beta);
    APY_LAUNCH_CHECK();
    return APY_OK;
  }
  Plan pl = make_plan(d);
  size_t need = pl.use_ws ? (size_t)pl.npart * (size_t)d->m * (size_t)d->n * sizeof(double2) : 0;
  if (need > ws_bytes || (need && !workspace)) return APY_E_WORKSPACE;
  if (pl.total > 0x7fffffffLL) return APY_E_ARG;

  GemmArgs a;
  a.m = d->m; a.n = d->n; a.k = d->k;
  a.a = (const double2*)d->a; a.lda = d->lda; a.sa = d->stride_a;
  a.b = (const double2*)d->b; a.ldb = d->ldb; a.sb = d->stride_b;
  a.c = (double2*)d->c; a.ldc = d->ldc; a.sc = d->stride_c;
  a.alr = d->alpha_re; a.ali = d->alpha_im; a.ber = d->beta_re; a.bei = d->beta_im;
  a.conj_a = d->conj_a; a.conj_b = d->conj_b; a.use_beta = use_beta;
  a.batch = d->batch; a.split = pl.split; a.kchunk = pl.kchunk;
  a.mtiles = pl.mtiles; a.ntiles = pl.ntiles; a.total = pl.total;
  a.ws = pl.use_ws ? (double2*)workspace : nullptr;

  if (pl.cfg == 1) launch_cfg<32, 64, 16, 32>(a, d->a_layout, d->b_layout, s);
  else launch_cfg<64, 64, 32, 32>(a, d->a_layout, d->b_layout, s);
  APY_LAUNCH_CHECK();

  if (pl.use_ws) {
    int64_t group = d->reduce_batch ? pl.npart : pl.split;
    int64_t ngroups = d->reduce_batch ? 1 : d->batch;
    int64_t tot = ngroups * d->m * d->n;
    zgemm_reduce<<<apy_grid1d(tot), 256, 0, s>>>((const double2*)workspace, group, ngroups, d->m, d->n,
                                                 (double2*)d->c, d->ldc, d->stride_c, d->beta_re, d->beta_im,
                                                 use_beta);
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

extern "C" {

int apy_version(void) { return 100; }

size_t apy_zgemm_workspace(const apy_gemm_desc* d) { return apy_zgemm_ws_bytes(d); }

int apy_zgemm(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, void* stream) {
  return apy_zgemm_launch(d, workspace, ws_bytes, (hipStream_t)stream);
}

}  // extern "C"
