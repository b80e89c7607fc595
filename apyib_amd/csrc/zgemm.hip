// Complex128 GEMM on CDNA4 FP64 matrix cores.
//
// One complex multiply-accumulate is four real MFMA products on separate
// real/imaginary planes (re += ar*br - ai*bi, im += ar*bi + ai*br), i.e.
// 8 real flops per complex MAC, executed with v_mfma_f64_16x16x4f64.
//
// Tiling: 256 threads = 4 waves arranged 2x2; each wave owns a WM x WN
// complex output tile built from 16x16 MFMA blocks.  K is staged through a
// double-buffered LDS tile, BK complex elements at a time, with two register
// stages in front of it: while the MFMAs consume LDS buffer s, stage s+1
// waits in registers for the other buffer and stage s+2 is in flight from
// memory; each stage needs a single barrier.  The loads of a stage carry
// no data dependence (clamped addresses; bounds mask and conjugation are
// applied when the registers go to LDS) and the loop body has no
// conditional compute, so the wait counts before each LDS store leave the
// newest stage's loads in flight.  The BK = 8 tiles (configs 8, 9) halve
// the LDS per block, which raises the resident blocks per CU; on the
// 190 x 12720 x 12720 ladder product they run at 0.85 of the FP64 MFMA
// peak.
//
// MFMA f64 16x16x4 lane layout (checked on gfx950):
//   A operand: lane l holds A[l%16][l/16]       (16 x 4)
//   B operand: lane l holds B[l/16][l%16]       (4 x 16)
//   D result : lane l, element r holds D[l/16 + 4r][l%16]
//
// Work decomposition: (M tile, N tile, K slice, batch) units on a flat grid
// whose block id is remapped so consecutive logical units land on the same
// XCD (workgroups are dealt round-robin over the 8 XCDs), M tile fastest:
// blocks sharing a B tile share that XCD's L2.  K slices (split-K) write
// partial products to a workspace that one reduction kernel sums in a fixed
// order, so results are deterministic for a given plan.
//
// Plans (tile config, split) come from a cost model over the device's
// resident-block count, or from the tune cache filled by apy_zgemm_tune,
// which times the candidate plans on the caller's operands.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

namespace {

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
  int g3;           // Gauss three-product form (apy_gemm_desc.gauss3m)
};

__device__ inline int64_t xcd_remap(int64_t bid, int64_t total) {
  int64_t xcd = bid & 7, slot = bid >> 3;
  int64_t per = total >> 3, rem = total & 7;
  return (xcd < rem) ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// v (conjugated when sg holds the sign bit) if in, else 0; branch-free.
__device__ inline double2 masked(double2 v, bool in, unsigned long long sg) {
  const unsigned long long m = in ? ~0ull : 0ull;
  return make_double2(__longlong_as_double(__double_as_longlong(v.x) & m),
                      __longlong_as_double((__double_as_longlong(v.y) ^ sg) & m));
}

// G3: the Gauss form of the complex product, three real products per complex multiply-add
// (T1 = ar br, T2 = ai bi, T3 = (ar + ai)(br + bi); re = T1 - T2, im = T3 - T1 - T2): three MFMAs
// instead of four, one more accumulator plane, the operand sums formed in registers.
template <int BM, int BN, int WM, int WN, int BK, bool A_MK, bool B_NK, bool G3>
__global__ __launch_bounds__(NTHR) void zgemm_mfma(GemmArgs p) {
  constexpr int RB = WM / 16, CB = WN / 16;
  constexpr int BKP = BK + 1;  // padded row (complex elements) of k-contiguous LDS tiles
  constexpr int A_PER = BM * BK / NTHR;
  constexpr int B_PER = BN * BK / NTHR;
  constexpr int A_LDS = BM * BKP;
  constexpr int B_LDS = B_NK ? BN * BKP : BK * BN;
  static_assert(BM == 2 * WM && BN == 2 * WN, "2x2 wave grid");
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small");

  // Interleaved (re, im) tiles: one 16-byte LDS read yields both planes of
  // an MFMA operand.  Row strides of BKP complex (272 B for BK=16) keep the
  // 8-lane groups of a 128-bit LDS access on distinct banks.
  __shared__ double2 As[2][A_LDS];
  __shared__ double2 Bs[2][B_LDS];

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

  v4d acc_re[RB][CB], acc_im[RB][CB], acc_3[G3 ? RB : 1][G3 ? CB : 1];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) { acc_re[i][j] = (v4d){0, 0, 0, 0}; acc_im[i][j] = (v4d){0, 0, 0, 0}; }
  if (G3) {
#pragma unroll
    for (int i = 0; i < (G3 ? RB : 1); ++i)
#pragma unroll
      for (int j = 0; j < (G3 ? CB : 1); ++j) acc_3[i][j] = (v4d){0, 0, 0, 0};
  }

  // Two register stages: stage s+2 is in flight while stage s+1 waits in
  // registers and stage s is consumed from LDS.
  double2 ra[2][A_PER], rb[2][B_PER];
  // conjugation as a sign-bit flip of the imaginary part
  const unsigned long long sga = p.conj_a ? 0x8000000000000000ull : 0ull;
  const unsigned long long sgb = p.conj_b ? 0x8000000000000000ull : 0ull;

  // `set` / `buf` arguments are std::integral_constant so that every
  // register-array index is a compile-time constant (no scratch memory).
  // Loads read clamped (always in-bounds) addresses and carry no data
  // dependence, so a stage's loads stay in flight while the previous stage
  // computes; the bounds mask and the conjugation sign are applied when the
  // registers are written to LDS one stage later.
  auto load = [&](int64_t k0, auto setc) {
    constexpr int set = decltype(setc)::value;
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int idx = tid + NTHR * q;
      int r, kk;
      if (A_MK) { r = idx / BK; kk = idx % BK; } else { kk = idx / BM; r = idx % BM; }
      const int64_t gm = min(m0 + r, p.m - 1), gk = min(k0 + kk, kend - 1);
      ra[set][q] = A_MK ? A[gm * p.lda + gk] : A[gk * p.lda + gm];
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int idx = tid + NTHR * q;
      int c, kk;
      if (B_NK) { c = idx / BK; kk = idx % BK; } else { kk = idx / BN; c = idx % BN; }
      const int64_t gn = min(n0 + c, p.n - 1), gk = min(k0 + kk, kend - 1);
      rb[set][q] = B_NK ? B[gn * p.ldb + gk] : B[gk * p.ldb + gn];
    }
  };
  auto store = [&](auto bufc, int64_t k0) {
    constexpr int buf = decltype(bufc)::value, set = buf;
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int idx = tid + NTHR * q;
      int r, kk;
      if (A_MK) { r = idx / BK; kk = idx % BK; } else { kk = idx / BM; r = idx % BM; }
      const bool in = (m0 + r < p.m) && (k0 + kk < kend);
      As[buf][r * BKP + kk] = masked(ra[set][q], in, sga);
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int idx = tid + NTHR * q;
      int c, kk;
      if (B_NK) { c = idx / BK; kk = idx % BK; } else { kk = idx / BN; c = idx % BN; }
      const bool in = (n0 + c < p.n) && (k0 + kk < kend);
      Bs[buf][B_NK ? c * BKP + kk : kk * BN + c] = masked(rb[set][q], in, sgb);
    }
  };

  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  // An even number of stages (a padded last stage reads zeros) keeps the
  // two-stage loop body free of conditional compute, so the accumulators
  // stay in the matrix-core registers across iterations.
  const int64_t nreal = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  const int64_t nstages = (nreal + 1) & ~(int64_t)1;
  if (nstages > 0) {
    load(kbeg, I0{});
    load(kbeg + BK, I1{});
    store(I0{}, kbeg);
  }
  __syncthreads();
  auto stage = [&](int64_t st, auto curc) {
    constexpr int cur = decltype(curc)::value;
    using Other = std::integral_constant<int, cur ^ 1>;
    // set `cur` was stored last stage; past the end the clamped loads
    // re-read the last stage's lines and are masked at the store.  No branch
    // here: the wait counts at the stores then cover only the older loads.
    load(kbeg + (st + 2) * BK, curc);
    const double2* Ab = As[cur];
    const double2* Bb = Bs[cur];
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      double2 av[RB], bv[CB];
#pragma unroll
      for (int i = 0; i < RB; ++i) av[i] = Ab[(wm * WM + i * 16 + (lane & 15)) * BKP + kk];
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        const int col = wn * WN + j * 16 + (lane & 15);
        bv[j] = Bb[B_NK ? col * BKP + kk : kk * BN + col];
      }
      if (G3) {
        // acc_re = T1, acc_im = T2, acc_3 = T3
        double sa[RB], sb[CB];
#pragma unroll
        for (int i = 0; i < RB; ++i) sa[i] = av[i].x + av[i].y;
#pragma unroll
        for (int j = 0; j < CB; ++j) sb[j] = bv[j].x + bv[j].y;
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < CB; ++j) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].x, bv[j].x, acc_re[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].y, bv[j].y, acc_im[i][j], 0, 0, 0);
            acc_3[G3 ? i : 0][G3 ? j : 0] =
                __builtin_amdgcn_mfma_f64_16x16x4f64(sa[i], sb[j], acc_3[G3 ? i : 0][G3 ? j : 0], 0, 0, 0);
          }
      } else {
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < CB; ++j) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].x, bv[j].x, acc_re[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].x, bv[j].y, acc_im[i][j], 0, 0, 0);
          }
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < CB; ++j) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[i].y, bv[j].y, acc_re[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].y, bv[j].x, acc_im[i][j], 0, 0, 0);
          }
      }
    }
    store(Other{}, kbeg + (st + 1) * BK);  // past the end: writes zeros
    __syncthreads();
  };
  for (int64_t st = 0; st < nstages; st += 2) {
    stage(st, I0{});
    stage(st + 1, I1{});
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
        const int64_t row = m0 + wm * WM + i * 16 + (lane >> 4) + 4 * r;
        const int64_t col = n0 + wn * WN + j * 16 + (lane & 15);
        if (row < p.m && col < p.n) {
          double vr = acc_re[i][j][r], vi = acc_im[i][j][r];
          if (G3) {
            const double t1 = vr, t2 = vi, t3 = acc_3[G3 ? i : 0][G3 ? j : 0][r];
            vr = t1 - t2;
            vi = t3 - t1 - t2;
          }
          double o_r = p.alr * vr - p.ali * vi, o_i = p.alr * vi + p.ali * vr;
          if (W) {
            W[row * p.n + col] = make_double2(o_r, o_i);
          } else {
            double2* cp = Cb + row * p.ldc + col;
            if (p.use_beta) {
              const double2 old = *cp;
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
    const int64_t g = t / mn, e = t % mn;
    const int64_t row = e / n, col = e % n;
    double sr = 0.0, si = 0.0;
    if (group > 0) {
      const double2* w = ws + g * group * mn + e;
      for (int64_t q = 0; q < group; ++q) { const double2 v = w[q * mn]; sr += v.x; si += v.y; }
    }
    double2* cp = c + g * sc + row * ldc + col;
    if (use_beta) {
      const double2 old = *cp;
      sr += ber * old.x - bei * old.y;
      si += ber * old.y + bei * old.x;
    }
    *cp = make_double2(sr, si);
  }
}

// The same sum for small outputs with many K slices (skinny products split thousands of ways): a
// block takes 32 consecutive outputs and 8 slice lanes; lane r sums slices r, r+8, ... in order (four
// interleaved accumulators), then the 8 lane sums are added in lane order.  Deterministic for a plan.
constexpr int RW_E = 32, RW_R = 8;
__global__ __launch_bounds__(RW_E * RW_R) void zgemm_reduce_wide(const double2* __restrict__ ws, int64_t group,
                                                                 int64_t ngroups, int64_t m, int64_t n, double2* c,
                                                                 int64_t ldc, int64_t sc, double ber, double bei,
                                                                 int use_beta) {
  const int64_t mn = m * n, total = ngroups * mn;
  const int j = threadIdx.x % RW_E, r = threadIdx.x / RW_E;
  const int64_t t = blockIdx.x * (int64_t)RW_E + j;
  __shared__ double sr[RW_R][RW_E], si[RW_R][RW_E];
  double ar[4] = {0.0, 0.0, 0.0, 0.0}, ai[4] = {0.0, 0.0, 0.0, 0.0};
  if (t < total) {
    const int64_t g = t / mn, e = t % mn;
    const double2* w = ws + g * group * mn + e;
    int64_t q = r;
    for (; q + 3 * RW_R < group; q += 4 * RW_R) {
      const double2 v0 = w[q * mn], v1 = w[(q + RW_R) * mn], v2 = w[(q + 2 * RW_R) * mn], v3 = w[(q + 3 * RW_R) * mn];
      ar[0] += v0.x; ai[0] += v0.y; ar[1] += v1.x; ai[1] += v1.y;
      ar[2] += v2.x; ai[2] += v2.y; ar[3] += v3.x; ai[3] += v3.y;
    }
    for (int u = 0; q < group; q += RW_R, ++u) { const double2 v = w[q * mn]; ar[u] += v.x; ai[u] += v.y; }
  }
  sr[r][j] = (ar[0] + ar[1]) + (ar[2] + ar[3]);
  si[r][j] = (ai[0] + ai[1]) + (ai[2] + ai[3]);
  __syncthreads();
  if (r == 0 && t < total) {
    double xr = 0.0, xi = 0.0;
    for (int k = 0; k < RW_R; ++k) { xr += sr[k][j]; xi += si[k][j]; }
    const int64_t g = t / mn, e = t % mn, row = e / n, col = e % n;
    double2* cp = c + g * sc + row * ldc + col;
    if (use_beta) {
      const double2 old = *cp;
      xr += ber * old.x - bei * old.y;
      xi += ber * old.y + bei * old.x;
    }
    *cp = make_double2(xr, xi);
  }
}

// ------------------------------------------------------------ tile configs
struct Cfg {
  int bm, bn, wm, wn, bk;
  double eff;  // relative MFMA efficiency used by the cost model
};
constexpr int NCFG = 9;
constexpr Cfg CFGS[NCFG + 1] = {
    // eff: measured fraction of the FP64 MFMA peak on large products
    {0, 0, 0, 0, 0, 0.0},          // 0: automatic
    {64, 64, 32, 32, 16, 0.83},    // 1
    {32, 64, 16, 32, 16, 0.82},    // 2
    {64, 128, 32, 64, 16, 0.72},   // 3
    {128, 64, 64, 32, 16, 0.60},   // 4
    {96, 64, 48, 32, 16, 0.71},    // 5
    {64, 64, 32, 32, 32, 0.69},    // 6
    {32, 32, 16, 16, 16, 0.76},    // 7
    {64, 64, 32, 32, 8, 0.855},    // 8: half the LDS of 1, more blocks per CU
    {32, 64, 16, 32, 8, 0.81},     // 9
};

template <int CI, bool AM, bool BN_>
struct KernelOf {
  static constexpr Cfg c = CFGS[CI];
  static void* ptr() { return (void*)zgemm_mfma<c.bm, c.bn, c.wm, c.wn, c.bk, AM, BN_, false>; }
  static void launch(const GemmArgs& a, hipStream_t s) {
    if (a.g3)
      zgemm_mfma<c.bm, c.bn, c.wm, c.wn, c.bk, AM, BN_, true><<<dim3((unsigned)a.total), dim3(NTHR), 0, s>>>(a);
    else
      zgemm_mfma<c.bm, c.bn, c.wm, c.wn, c.bk, AM, BN_, false><<<dim3((unsigned)a.total), dim3(NTHR), 0, s>>>(a);
  }
};

template <int CI>
void launch_layout(const GemmArgs& a, bool am, bool bnk, hipStream_t s) {
  if (am && bnk) KernelOf<CI, true, true>::launch(a, s);
  else if (am) KernelOf<CI, true, false>::launch(a, s);
  else if (bnk) KernelOf<CI, false, true>::launch(a, s);
  else KernelOf<CI, false, false>::launch(a, s);
}

template <int CI>
void* kernel_ptr(bool am, bool bnk) {
  if (am && bnk) return KernelOf<CI, true, true>::ptr();
  if (am) return KernelOf<CI, true, false>::ptr();
  if (bnk) return KernelOf<CI, false, true>::ptr();
  return KernelOf<CI, false, false>::ptr();
}

void launch_cfg(int cfg, const GemmArgs& a, bool am, bool bnk, hipStream_t s) {
  switch (cfg) {
    case 1: launch_layout<1>(a, am, bnk, s); break;
    case 2: launch_layout<2>(a, am, bnk, s); break;
    case 3: launch_layout<3>(a, am, bnk, s); break;
    case 4: launch_layout<4>(a, am, bnk, s); break;
    case 5: launch_layout<5>(a, am, bnk, s); break;
    case 6: launch_layout<6>(a, am, bnk, s); break;
    case 8: launch_layout<8>(a, am, bnk, s); break;
    case 9: launch_layout<9>(a, am, bnk, s); break;
    default: launch_layout<7>(a, am, bnk, s); break;
  }
}

void* cfg_kernel(int cfg, bool am, bool bnk) {
  switch (cfg) {
    case 1: return kernel_ptr<1>(am, bnk);
    case 2: return kernel_ptr<2>(am, bnk);
    case 3: return kernel_ptr<3>(am, bnk);
    case 4: return kernel_ptr<4>(am, bnk);
    case 5: return kernel_ptr<5>(am, bnk);
    case 6: return kernel_ptr<6>(am, bnk);
    case 8: return kernel_ptr<8>(am, bnk);
    case 9: return kernel_ptr<9>(am, bnk);
    default: return kernel_ptr<7>(am, bnk);
  }
}

// Device facts, queried once per device.
struct DevInfo {
  int cus = 256;
  int occ[NCFG + 1][4] = {};
  bool ready = false;
};
std::mutex g_mu;
std::map<int, DevInfo> g_dev;

using Key = std::tuple<int64_t, int64_t, int64_t, int, int, int64_t, int, int>;  // m,n,k,al,bl,batch,reduce,device
std::map<Key, std::pair<int, int>> g_tuned;
// cost-model choices already made (the model's search over tiles and splits is deterministic per shape)
std::map<Key, std::pair<int, int>> g_model;

int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) d = 0;
  return d;
}

const DevInfo& dev_info() {
  std::lock_guard<std::mutex> lk(g_mu);
  int d = current_device();
  DevInfo& I = g_dev[d];
  if (!I.ready) {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, d) == hipSuccess) I.cus = pr.multiProcessorCount;
    for (int c = 1; c <= NCFG; ++c)
      for (int l = 0; l < 4; ++l) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)cfg_kernel(c, l & 1, l >> 1), NTHR, 0) !=
                hipSuccess ||
            nb < 1)
          nb = 1;
        I.occ[c][l] = nb;
      }
    I.ready = true;
  }
  return I;
}

Key key_of(const apy_gemm_desc* d) {
  return Key(d->m, d->n, d->k, d->a_layout, d->b_layout, d->batch, d->reduce_batch, current_device());
}

struct Plan {
  int cfg;
  int split;
  int64_t kchunk;
  int64_t mtiles, ntiles, total;
  bool use_ws;
  int64_t npart;
  double cost;
};

// Largest split the workspace query reserves for: partial sums are capped
// at ~768 MB; small outputs with long K (a few tiles) may use up to 256
// slices so that the grid still fills the chip.
int max_split(const apy_gemm_desc* d) {
  const double mn_bytes = (double)d->m * (double)d->n * 16.0 * (double)d->batch;
  const double cap = mn_bytes < 4.0e6 ? 256.0 : 16.0;
  int s = (int)std::min(cap, std::max(1.0, std::floor(768.0e6 / std::max(mn_bytes, 1.0))));
  return std::max(1, s);
}

Plan plan_for(const apy_gemm_desc* d, int cfg, int split, const DevInfo& I) {
  const Cfg& c = CFGS[cfg];
  Plan pl;
  pl.cfg = cfg;
  pl.mtiles = (d->m + c.bm - 1) / c.bm;
  pl.ntiles = (d->n + c.bn - 1) / c.bn;
  int64_t kchunk = (d->k + split - 1) / split;
  kchunk = ((kchunk + c.bk - 1) / c.bk) * c.bk;
  if (kchunk < c.bk) kchunk = c.bk;
  split = (int)((d->k + kchunk - 1) / kchunk);
  if (split < 1) split = 1;
  pl.split = split;
  pl.kchunk = kchunk;
  pl.total = pl.mtiles * pl.ntiles * split * d->batch;
  pl.use_ws = split > 1 || d->reduce_batch;
  pl.npart = (int64_t)split * d->batch;
  // Cost model: rounds of resident blocks x stages per block x stage time,
  // plus the partial-sum traffic of the reduction.
  const int lay = (d->a_layout == APY_MK ? 1 : 0) | (d->b_layout == APY_NK ? 2 : 0);
  const double slots = (double)I.cus * (double)I.occ[cfg][lay];
  const double rounds = std::ceil((double)pl.total / slots);
  const double stages = std::ceil((double)kchunk / c.bk);
  const double peak_cu = 77.0e12 / 256.0;
  const double t_stage = (double)I.occ[cfg][lay] * c.bm * c.bn * c.bk * 8.0 / (peak_cu * c.eff);
  double t = rounds * stages * t_stage + 3.0e-6;
  if (pl.use_ws) t += (double)(pl.npart + 2) * d->m * d->n * 16.0 / 4.0e12 + 4.0e-6;
  pl.cost = t;
  return pl;
}

Plan plan_for(const apy_gemm_desc* d, int cfg, int split) { return plan_for(d, cfg, split, dev_info()); }

bool cfg_fits(const apy_gemm_desc* d, int cfg) {
  const Cfg& c = CFGS[cfg];
  // Tiles much larger than the matrix waste most of the MFMA work.
  if (c.bm > 32 && d->m <= c.bm / 2) return false;
  if (c.bn > 32 && d->n <= c.bn / 2) return false;
  return true;
}

Plan model_plan(const apy_gemm_desc* d) {
  const Key key = key_of(d);
  int mcfg = 0, msplit = 0;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_model.find(key);
    if (it != g_model.end()) { mcfg = it->second.first; msplit = it->second.second; }
  }
  // plan_for takes g_mu itself (dev_info), so it runs after the lock above is released
  if (mcfg > 0) return plan_for(d, mcfg, msplit);
  const DevInfo& I = dev_info();
  Plan best;
  best.cost = 1e300;
  best.cfg = 0;
  const int smax = max_split(d);
  for (int cfg = 1; cfg <= NCFG; ++cfg) {
    if (!cfg_fits(d, cfg) && cfg != 7) continue;
    for (int s = 1; s <= smax; ++s) {
      if (s > 1 && (d->k + s - 1) / s < 4 * CFGS[cfg].bk) break;
      Plan p = plan_for(d, cfg, s, I);
      if (p.cost < best.cost) best = p;
    }
  }
  if (best.cfg == 0) best = plan_for(d, 7, 1, I);
  std::lock_guard<std::mutex> lk(g_mu);
  g_model[key] = {best.cfg, best.split};
  return best;
}

Plan make_plan(const apy_gemm_desc* d) {
  if (d->tile >= 1 && d->tile <= NCFG) {
    int s = d->split_k > 0 ? std::min(d->split_k, max_split(d)) : 0;
    if (s > 0) return plan_for(d, d->tile, s);
    Plan best;
    best.cost = 1e300;
    const DevInfo& I = dev_info();
    for (int q = 1; q <= max_split(d); ++q) {
      Plan p = plan_for(d, d->tile, q, I);
      if (p.cost < best.cost) best = p;
    }
    return best;
  }
  int tcfg = 0, tsplit = 0;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tuned.find(key_of(d));
    if (it != g_tuned.end()) { tcfg = it->second.first; tsplit = it->second.second; }
  }
  if (tcfg > 0) return plan_for(d, tcfg, d->split_k > 0 ? std::min(d->split_k, max_split(d)) : tsplit);
  if (d->split_k > 0) {
    Plan m = model_plan(d);
    return plan_for(d, m.cfg, std::min(d->split_k, max_split(d)));
  }
  return model_plan(d);
}

size_t ws_need(const apy_gemm_desc* d, const Plan& pl) {
  return pl.use_ws ? (size_t)pl.npart * (size_t)d->m * (size_t)d->n * sizeof(double2) : 0;
}

int run_plan(const apy_gemm_desc* d, const Plan& pl, void* workspace, size_t ws_bytes, hipStream_t s,
             hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  const int use_beta = (d->beta_re != 0.0 || d->beta_im != 0.0) ? 1 : 0;
  if (ws_need(d, pl) > ws_bytes || (ws_need(d, pl) && !workspace)) return APY_E_WORKSPACE;
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
  a.g3 = d->gauss3m ? 1 : 0;
  if (ev0) APY_HIP_TRY(hipEventRecord(ev0, s));
  launch_cfg(pl.cfg, a, d->a_layout == APY_MK, d->b_layout == APY_NK, s);
  APY_LAUNCH_CHECK();
  if (ev1) APY_HIP_TRY(hipEventRecord(ev1, s));
  if (pl.use_ws) {
    const int64_t group = d->reduce_batch ? pl.npart : pl.split;
    const int64_t ngroups = d->reduce_batch ? 1 : d->batch;
    const int64_t tot = ngroups * d->m * d->n;
    if (tot <= (1 << 18) && group >= 4 * RW_R) {
      zgemm_reduce_wide<<<(unsigned)((tot + RW_E - 1) / RW_E), RW_E * RW_R, 0, s>>>(
          (const double2*)workspace, group, ngroups, d->m, d->n, (double2*)d->c, d->ldc, d->stride_c, d->beta_re,
          d->beta_im, use_beta);
    } else {
      zgemm_reduce<<<apy_grid1d(tot), 256, 0, s>>>((const double2*)workspace, group, ngroups, d->m, d->n,
                                                   (double2*)d->c, d->ldc, d->stride_c, d->beta_re, d->beta_im,
                                                   use_beta);
    }
    APY_LAUNCH_CHECK();
  }
  return APY_OK;
}

int check_desc(const apy_gemm_desc* d) {
  if (!d) return APY_E_ARG;
  if (d->m < 0 || d->n < 0 || d->k < 0 || d->batch < 1) return APY_E_ARG;
  if (d->a_layout != APY_MK && d->a_layout != APY_KM) return APY_E_ARG;
  if (d->b_layout != APY_KN && d->b_layout != APY_NK) return APY_E_ARG;
  if (d->tile < 0 || d->tile > NCFG || d->split_k < 0) return APY_E_ARG;
  if (d->m > 0 && d->n > 0 && (!d->c || ((uintptr_t)d->c & 15))) return APY_E_ARG;
  if (d->m > 0 && d->n > 0 && d->k > 0 && (!d->a || !d->b || ((uintptr_t)d->a & 15) || ((uintptr_t)d->b & 15)))
    return APY_E_ARG;
  return APY_OK;
}

}  // namespace

size_t apy_zgemm_ws_bytes(const apy_gemm_desc* d) {
  if (!d || d->m <= 0 || d->n <= 0 || d->k <= 0 || d->batch <= 0) return 0;
  const int smax = max_split(d);
  if (smax <= 1 && !d->reduce_batch) return 0;
  return (size_t)smax * (size_t)d->batch * (size_t)d->m * (size_t)d->n * sizeof(double2);
}

size_t apy_zgemm_plan_ws_bytes(const apy_gemm_desc* d) {
  if (!d || d->m <= 0 || d->n <= 0 || d->k <= 0 || d->batch <= 0) return 0;
  return ws_need(d, make_plan(d));
}

int apy_zgemm_launch(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, hipStream_t s, hipEvent_t ev0,
                     hipEvent_t ev1) {
  int st = check_desc(d);
  if (st != APY_OK) return st;
  if (d->m == 0 || d->n == 0) return APY_OK;
  if (d->k == 0) {
    // An empty sum contributes zero: C = beta*C.
    const int use_beta = (d->beta_re != 0.0 || d->beta_im != 0.0) ? 1 : 0;
    const int64_t nb = d->reduce_batch ? 1 : d->batch;
    zgemm_reduce<<<apy_grid1d(nb * d->m * d->n), 256, 0, s>>>(nullptr, 0, nb, d->m, d->n, (double2*)d->c, d->ldc,
                                                               d->stride_c, d->beta_re, d->beta_im, use_beta);
    APY_LAUNCH_CHECK();
    return APY_OK;
  }
  return run_plan(d, make_plan(d), workspace, ws_bytes, s, ev0, ev1);
}

int apy_zgemm_tune_impl(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, int reps, hipStream_t s,
                        int* best_cfg, int* best_split, double* best_ms) {
  int st = check_desc(d);
  if (st != APY_OK) return st;
  if (d->m == 0 || d->n == 0 || d->k == 0) return APY_OK;
  if (reps < 1) reps = 1;
  hipEvent_t e0, e1;
  APY_HIP_TRY(hipEventCreate(&e0));
  APY_HIP_TRY(hipEventCreate(&e1));
  double bms = 1e300;
  int bc = 0, bs = 0;
  const int smax = max_split(d);
  for (int cfg = 1; cfg <= NCFG; ++cfg) {
    if (!cfg_fits(d, cfg)) continue;
    // the model's best split for this tile, and its neighbours
    Plan m;
    m.cost = 1e300;
    for (int q = 1; q <= smax; ++q) {
      if (q > 1 && (d->k + q - 1) / q < 4 * CFGS[cfg].bk) break;
      Plan p = plan_for(d, cfg, q);
      if (p.cost < m.cost) m = p;
    }
    // the model's split, its neighbours, half and double, and the small splits that fill a skinny grid
    const int raw[8] = {m.split, m.split - 1, m.split + 1, m.split / 2, m.split * 2, 1, 2, 4};
    int cands[8], nc = 0;
    for (int q : raw) {
      if (q < 1 || q > smax) continue;
      if (q > 1 && (d->k + q - 1) / q < 4 * CFGS[cfg].bk) continue;
      const int eff = plan_for(d, cfg, q).split;   // splits that round to the same K slices time once
      bool dup = false;
      for (int c = 0; c < nc; ++c) dup = dup || cands[c] == eff;
      if (!dup) cands[nc++] = eff;
    }
    for (int ci = 0; ci < nc; ++ci) {
      Plan p = plan_for(d, cfg, cands[ci]);
      if (ws_need(d, p) > ws_bytes) continue;
      st = run_plan(d, p, workspace, ws_bytes, s);  // warm-up
      if (st != APY_OK) break;
      APY_HIP_TRY(hipEventRecord(e0, s));
      for (int r = 0; r < reps && st == APY_OK; ++r) st = run_plan(d, p, workspace, ws_bytes, s);
      APY_HIP_TRY(hipEventRecord(e1, s));
      APY_HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.f;
      APY_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      if (st != APY_OK) break;
      if (ms / reps < bms) { bms = ms / reps; bc = cfg; bs = p.split; }
    }
    if (st != APY_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != APY_OK) return st;
  if (bc > 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuned[key_of(d)] = std::make_pair(bc, bs);
  }
  if (best_cfg) *best_cfg = bc;
  if (best_split) *best_split = bs;
  if (best_ms) *best_ms = bms;
  return APY_OK;
}

extern "C" {

int apy_version(void) { return 100; }

size_t apy_zgemm_workspace(const apy_gemm_desc* d) { return apy_zgemm_ws_bytes(d); }

int apy_zgemm(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, void* stream) {
  return apy_zgemm_launch(d, workspace, ws_bytes, (hipStream_t)stream);
}

int apy_zgemm_plan(const apy_gemm_desc* d, int* tile, int* split) {
  int st = check_desc(d);
  if (st != APY_OK) return st;
  if (d->m == 0 || d->n == 0 || d->k == 0) {
    if (tile) *tile = 0;
    if (split) *split = 0;
    return APY_OK;
  }
  Plan p = make_plan(d);
  if (tile) *tile = p.cfg;
  if (split) *split = p.split;
  return APY_OK;
}

int apy_zgemm_tune(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, int reps, void* stream, int* tile,
                   int* split, double* ms) {
  return apy_zgemm_tune_impl(d, workspace, ws_bytes, reps, (hipStream_t)stream, tile, split, ms);
}

int apy_zgemm_tune_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuned.clear();
  return APY_OK;
}

int apy_zgemm_tune_set(const apy_gemm_desc* d, int tile, int split) {
  // shape-only: the descriptor's pointers are not used
  if (!d || d->m < 1 || d->n < 1 || d->k < 1 || d->batch < 1) return APY_E_ARG;
  if ((d->a_layout != APY_MK && d->a_layout != APY_KM) || (d->b_layout != APY_KN && d->b_layout != APY_NK))
    return APY_E_ARG;
  if (tile < 1 || tile > NCFG || split < 1) return APY_E_ARG;
  const int s = std::min(split, max_split(d));
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuned[key_of(d)] = std::make_pair(tile, s);
  return APY_OK;
}

int apy_zgemm_tune_export(int64_t* rows, int64_t capacity, int64_t* count) {
  if (!count || capacity < 0 || (capacity > 0 && !rows)) return APY_E_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  const int dev = current_device();
  int64_t q = 0;
  for (const auto& kv : g_tuned) {
    const Key& k = kv.first;
    if (std::get<7>(k) != dev) continue;
    if (q < capacity) {
      int64_t* r = rows + 9 * q;
      r[0] = std::get<0>(k); r[1] = std::get<1>(k); r[2] = std::get<2>(k);
      r[3] = std::get<3>(k); r[4] = std::get<4>(k); r[5] = std::get<5>(k); r[6] = std::get<6>(k);
      r[7] = kv.second.first; r[8] = kv.second.second;
    }
    ++q;
  }
  *count = q;
  return APY_OK;
}

}  // extern "C"
