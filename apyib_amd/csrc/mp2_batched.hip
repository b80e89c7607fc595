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
// with Fvv(b,e) = F[O+b][O+e], Foo(m,j) = F[m][j]; an instance with
// transpose_f set uses F^T (the adjoint lambda equations), and s is the
// instance's driver scale, so amplitude and lambda solves share one batch.
// The amplitudes are antisymmetric in (i,j) and in (a,b) and are stored
// packed, x[(i<j)][(a<b)], a quarter of the dense tensor; the residual is
// formed only for i<j, a<b.
//
// Two kernels per step cover every instance.  nc_z_kernel forms
// Z_ij[(ab)] = sum_m Foo(m,j) x_im[(ab)] for all ordered pairs (i,j): a
// workgroup takes one (instance, i, slice of the (a<b) pairs), holds the O
// amplitudes x_im of each element in registers (read once, coalesced) and
// writes Z_ij for every j.  ncp_step_kernel then gives a workgroup one
// (instance, pair i<j, 32 x 32 tile pair A <= B of the virtual plane): it
// forms Y_ij[a,b] - Y_ij[b,a] over the tile on the FP64 matrix cores with
// K = e staged in 8-wide chunks through LDS (the tile's rows of x_ij expanded
// from the packed row with the antisymmetry signs, and the matching columns
// of Fvv), adds the driver and -Z_ij + Z_ji (two coalesced reads), applies the
// update and writes its partial residual norm; per_instance_sum reduces
// those per instance in a fixed order.  All tables live in the workspace, written once
// by apy_mp2_nc_batched_setup, so a sequence of steps can be captured in a
// HIP graph.
#include "common.h"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int BLK = 256;
constexpr int MT = 32;       // virtual tile edge
#ifndef APY_MP2_MKC
#define APY_MP2_MKC 8
#endif
constexpr int MKC = APY_MP2_MKC;   // e chunk of the Y product (a multiple of 8)
static_assert(MKC % 8 == 0 && MKC <= 32, "e chunk");
constexpr int SBLK = 256;                          // step kernel: 4 waves, 16 x 16 of the 32 x 32 tile each
constexpr int EN_B = 64;     // energy partial sums per instance

struct Item {
  int32_t inst, pij, ta, tb;
};

constexpr int ZS = 2 * BLK;  // (a<b) elements per Z work item
constexpr int ZMAX_O = 32;   // largest O of the matrix-core Z path (two 16-row j blocks, Foo in LDS)

struct ZItem {
  int32_t inst, i, p0, pad;
};

struct Header {
  int64_t n, items, zitems;
  int64_t zitem_off, z_off;   // byte offsets of the Z work items and of the Z blocks
};

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

__host__ __device__ inline int64_t npair(int64_t N) { return N * (N - 1) / 2; }
__host__ __device__ inline int64_t prow(int64_t p, int64_t N) { return p * (2 * N - p - 1) / 2; }
__host__ __device__ inline int64_t pidx(int64_t p, int64_t q, int64_t N) { return prow(p, N) + q - p - 1; }
inline int64_t ntiles(int64_t V) { return (V + MT - 1) / MT; }

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

// Workspace: header, instance table, item table, per-instance item offsets,
// item partial norms, energy partial sums, per-instance offsets of the Z
// blocks, then (at byte offsets the header holds, so the step needs only
// (n, items) on the host) the Z work items and the Z blocks themselves
// (sum_k O_k^2 Pv_k complex).
struct WsLayout {
  size_t inst, item, off, part, epart, zoff, zitem, z, total;
};

inline WsLayout layout(int n, int64_t items, int64_t zitems = 0, int64_t zelems = 0) {
  WsLayout L;
  L.inst = al256(sizeof(Header));
  L.item = L.inst + al256((size_t)n * sizeof(apy_mp2_nc_instance));
  L.off = L.item + al256((size_t)items * sizeof(Item));
  L.part = L.off + al256((size_t)(n + 1) * sizeof(int64_t));
  L.epart = L.part + al256((size_t)items * sizeof(double));
  L.zoff = L.epart + al256((size_t)n * EN_B * sizeof(double2));
  L.zitem = L.zoff + al256((size_t)(n + 1) * sizeof(int64_t));
  L.z = L.zitem + al256((size_t)zitems * sizeof(ZItem));
  L.total = L.z + al256((size_t)zelems * sizeof(double2));
  return L;
}

int64_t count_zitems(int n, const apy_mp2_nc_instance* inst, int64_t* zelems) {
  int64_t z = 0, e = 0;
  for (int k = 0; k < n; ++k) {
    if (inst[k].O < 2 || inst[k].V < 2) continue;
    const int64_t Pv = npair(inst[k].V);
    z += inst[k].O * ((Pv + ZS - 1) / ZS);
    e += inst[k].O * inst[k].O * Pv;
  }
  if (zelems) *zelems = e;
  return z;
}

int64_t count_items(int n, const apy_mp2_nc_instance* inst) {
  int64_t items = 0;
  for (int k = 0; k < n; ++k) {
    const int64_t nt = ntiles(inst[k].V);
    if (inst[k].O >= 2 && inst[k].V >= 2) items += npair(inst[k].O) * (nt * (nt + 1) / 2);
  }
  return items;
}

// acc += a b as four fused multiply-adds (written out: the compiler may not reassociate acc + (ab - cd))
__device__ inline void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// x_(pq)[c, d] of the packed amplitudes for any p != q, c != d (zero otherwise)
__device__ inline double2 xget(const double2* x, int64_t O, int64_t Pv, int64_t V, int64_t p, int64_t q, int64_t c,
                               int64_t d) {
  if (p == q || c == d) return make_double2(0.0, 0.0);
  const double2 v = x[pidx(p < q ? p : q, p < q ? q : p, O) * Pv + pidx(c < d ? c : d, c < d ? d : c, V)];
  return ((p < q) == (c < d)) ? v : make_double2(-v.x, -v.y);
}

// Z of one work item on the FP64 matrix cores: Z_ij[P] = sum_m Foo(m,j) x_im[P] as 16 x 16 blocks
// D[j][P] = sum_m A[j][m] B[m][P] with A = Foo^T (from LDS) and B the rows x_im (read along P, 16
// consecutive P per 16 lanes); each wave takes ZS / 64 blocks of 16 P, every j block (O <= 32) and
// K = m in steps of 4.  MFMA f64 16x16x4 lanes: A[l%16][l/16], B[l/16][l%16], D[l/16 + 4r][l%16].
__device__ __forceinline__ void z_item_mfma(const double2* __restrict__ fo, const double2* __restrict__ x,
                                            double2* Z, int O, int64_t Pv, int i, int64_t p0) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nj = (O + 15) / 16, nk = (O + 3) / 4;
  for (int pb = wave; pb < ZS / 16; pb += BLK / 64) {
    const int64_t P = p0 + 16 * pb + (lane & 15);
    v4d re[2], im[2];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) { re[jb] = (v4d){0, 0, 0, 0}; im[jb] = (v4d){0, 0, 0, 0}; }
    double2 bvs[ZMAX_O / 4];                            // B[m][P] = x_im[P] for every k step, loaded up front
#pragma unroll
    for (int kb = 0; kb < ZMAX_O / 4; ++kb) {
      const int m = 4 * kb + (lane >> 4);
      double2 bv = make_double2(0.0, 0.0);
      if (kb < nk && m < O && m != i && P < Pv) {
        bv = x[pidx(i < m ? i : m, i < m ? m : i, O) * Pv + P];
        if (m < i) bv = make_double2(-bv.x, -bv.y);
      }
      bvs[kb] = bv;
    }
#pragma unroll
    for (int kb = 0; kb < ZMAX_O / 4; ++kb) {
      if (kb >= nk) break;
      const int m = 4 * kb + (lane >> 4);
      const double2 bv = bvs[kb];
#pragma unroll
      for (int jb = 0; jb < 2; ++jb) {
        if (jb >= nj) break;
        const int j = 16 * jb + (lane & 15);
        const double2 av = (m < O && j < O) ? fo[m * O + j] : make_double2(0.0, 0.0);   // A[j][m] = Foo(m, j)
        re[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, re[jb], 0, 0, 0);
        im[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, im[jb], 0, 0, 0);
        re[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.y, bv.y, re[jb], 0, 0, 0);
        im[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.x, im[jb], 0, 0, 0);
      }
    }
    if (P < Pv) {
#pragma unroll
      for (int jb = 0; jb < 2; ++jb) {
        if (jb >= nj) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = 16 * jb + (lane >> 4) + 4 * r;
          if (j < O) Z[j * Pv + P] = make_double2(re[jb][r], im[jb][r]);
        }
      }
    }
  }
}

// Z_k[i][j][(ab)] = sum_m Foo(m,j) x_im[(ab)] for every instance, row i and slice of
// ZS pairs (ab) (grid-stride over the Z work list, whose length is in the header).
__device__ __forceinline__ void nc_z_phase(char* __restrict__ ws, const apy_mp2_nc_instance* __restrict__ insts,
                                           const int64_t* __restrict__ zoff, int reverse) {
  __shared__ double2 fo[ZMAX_O * ZMAX_O];
  const Header* hdr = (const Header*)ws;
  const int64_t nz = hdr->zitems;
  const ZItem* zitems = (const ZItem*)(ws + hdr->zitem_off);
  double2* zbase = (double2*)(ws + hdr->z_off);
  for (int64_t w = blockIdx.x; w < nz; w += gridDim.x) {
    const ZItem it = zitems[w];
    const apy_mp2_nc_instance I = insts[it.inst];
    const int64_t O = I.O, V = I.V, n = O + V, Pv = npair(V), i = it.i;
    const int transpose_f = I.transpose_f;
    const double2* F = (const double2*)I.fock;
    const double2* x = (const double2*)(reverse ? I.x_new : I.x);
    double2* Z = zbase + zoff[it.inst] + i * O * Pv;
    auto foo = [&](int64_t m, int64_t j) { return transpose_f ? F[j * n + m] : F[m * n + j]; };
    if (O <= ZMAX_O) {
      __syncthreads();                                   // fo of the previous item is no longer read
      for (int64_t q = threadIdx.x; q < O * O; q += BLK) fo[q] = foo(q / O, q % O);   // fo[m O + j]
      __syncthreads();
      z_item_mfma(fo, x, Z, (int)O, Pv, (int)i, it.p0);
    } else {
      for (int q = 0; q < ZS / BLK; ++q) {
        const int64_t P = it.p0 + threadIdx.x + (int64_t)BLK * q;
        if (P >= Pv) continue;
        for (int64_t j = 0; j < O; ++j) {
          double2 z = make_double2(0.0, 0.0);
          for (int64_t m = 0; m < O; ++m) {
            if (m == i) continue;
            double2 v = x[pidx(i < m ? i : m, i < m ? m : i, O) * Pv + P];
            if (m < i) v = make_double2(-v.x, -v.y);
            cfma(z, foo(m, j), v);
          }
          Z[j * Pv + P] = z;
        }
      }
    }
  }
}

__global__ __launch_bounds__(BLK) void nc_z_kernel(char* __restrict__ ws, const apy_mp2_nc_instance* __restrict__ insts,
                                                   const int64_t* __restrict__ zoff, int reverse) {
  nc_z_phase(ws, insts, zoff, reverse);
}

// One work item per workgroup of four waves: a 32 x 32 tile of the residual, each wave a 16 x 16
// quarter of it on the FP64 matrix cores.  Off the diagonal (A < B) both products go into one
// complex accumulator,
//   r[a][b] += sum_e XA[a][e] GB[e][b] - sum_e GA[e][a] XB[b][e],
// as eight v_mfma_f64_16x16x4f64 per four values of e (real and imaginary planes; the second
// product with its A operand negated).  On a diagonal tile pair (A == B) the two products are one
// matrix and its transpose, so each wave forms only Y = XA GA over its quarter (four MFMAs per
// four e) and r = Y - Y^T is assembled through LDS.  MFMA f64 16x16x4 lanes: A[l%16][l/16],
// B[l/16][l%16], D[l/16 + 4r][l%16].
__device__ __forceinline__ void ncp_step_item(int64_t b, const apy_mp2_nc_instance* __restrict__ insts,
                                              const Item* __restrict__ items, int reverse,
                                              const char* __restrict__ ws, const int64_t* __restrict__ zoff,
                                              double* __restrict__ part) {
  constexpr int TILE_ELEMS = 2 * MT * (MKC + 1) + 2 * MKC * (MT + 1);
  constexpr int SH_ELEMS = TILE_ELEMS > MT * (MT + 1) ? TILE_ELEMS : MT * (MT + 1);
  __shared__ double2 sh[SH_ELEMS];
  __shared__ double red[SBLK];
  auto XA = reinterpret_cast<double2 (*)[MKC + 1]>(sh);
  auto XB = reinterpret_cast<double2 (*)[MKC + 1]>(sh + MT * (MKC + 1));
  auto GA = reinterpret_cast<double2 (*)[MT + 1]>(sh + 2 * MT * (MKC + 1));
  auto GB = reinterpret_cast<double2 (*)[MT + 1]>(sh + 2 * MT * (MKC + 1) + MKC * (MT + 1));
  auto Ysh = reinterpret_cast<double2 (*)[MT + 1]>(sh);    // diagonal tiles, after the products
  __syncthreads();                                   // LDS of the previous item is no longer read
  const Item it = items[b];
  const apy_mp2_nc_instance I = insts[it.inst];
  const int64_t O = I.O, V = I.V, n = O + V, Pv = npair(V);
  const int transpose_f = I.transpose_f;
  const double s = I.drv_scale;
  const int tid = threadIdx.x, lane = tid & 63, wm = (tid >> 6) & 1, wn = tid >> 7;
  int64_t i, j;
  pdec(it.pij, O, i, j);
  const double2* F = (const double2*)I.fock;
  const double2* x = (const double2*)(reverse ? I.x_new : I.x);
  double2* xn = (double2*)(reverse ? I.x : I.x_new);
  const double2* row = x + (int64_t)it.pij * Pv;
  const int64_t a0 = (int64_t)it.ta * MT, b0 = (int64_t)it.tb * MT;
  // Fvv(b, e) (F^T: F[O+e][O+b])
  auto fvv = [&](int64_t b, int64_t e) { return transpose_f ? F[(O + e) * n + O + b] : F[(O + b) * n + O + e]; };

  v4d acc_re = (v4d){0, 0, 0, 0}, acc_im = (v4d){0, 0, 0, 0};
  // diagonal tile pair (A == B): XB = XA and GB = GA, so every wave forms only Y = x Fvv^T over its
  // quarter and the residual Y[a][b] - Y[b][a] is assembled through LDS after the products
  const bool diag = it.ta == it.tb;                          // uniform across the workgroup
  // one e chunk of x_ij rows a0.., b0.. and of Fvv columns, fetched into registers while the previous
  // chunk is multiplied (NQ elements of each of the four tiles per thread), then stored to LDS
  constexpr int NQ = MT * MKC / SBLK;
  const int Vi = (int)V;
  const uint32_t V2 = 2u * (uint32_t)Vi;
  // x_ij[p][e] is row[pidx(p, e)] above the diagonal (contiguous along e) and -row[pidx(e, p)] below it
  // (contiguous along p): each tile element is fetched by one of two passes, the upper one with e fastest
  // across the threads and the lower one with p fastest, so that both read along the packed row.
  double2 sxa_u[NQ], sxa_l[NQ], sxb_u[NQ], sxb_l[NQ], sga[NQ], sgb[NQ];
  auto upper = [&](int p, int e) -> double2 {        // p <= e or out of range (zero there)
    if (p >= Vi || e >= Vi || p >= e) return make_double2(0.0, 0.0);
    return row[(uint32_t)p * (V2 - (uint32_t)p - 1u) / 2u + (uint32_t)(e - p - 1)];
  };
  auto lower = [&](int p, int e) -> double2 {        // p > e in range
    if (p >= Vi || e >= Vi || p <= e) return make_double2(0.0, 0.0);
    const double2 v = row[(uint32_t)e * (V2 - (uint32_t)e - 1u) / 2u + (uint32_t)(p - e - 1)];
    return make_double2(-v.x, -v.y);
  };
  auto fetch = [&](int e0) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int q = tid + SBLK * k;
      const int r = q / MKC, c = q % MKC;            // upper pass: XA[r][c] = x_ij[a0+r, e0+c], c fastest
      sxa_u[k] = upper((int)a0 + r, e0 + c);
      const int rl = q % MT, cl = q / MT;            // lower pass: r fastest
      sxa_l[k] = lower((int)a0 + rl, e0 + cl);
      // GA[rr][cc] = Fvv(a0+cc, e0+rr), read along the contiguous index of F (e, or b for F^T)
      const int rr = transpose_f ? q / MT : q % MKC, cc = transpose_f ? q % MT : q / MKC;
      const int e2 = e0 + rr, ca = (int)a0 + cc, cb = (int)b0 + cc;
      sga[k] = (e2 < Vi && ca < Vi) ? fvv(ca, e2) : make_double2(0.0, 0.0);
      if (!diag) {
        sxb_u[k] = upper((int)b0 + r, e0 + c);
        sxb_l[k] = lower((int)b0 + rl, e0 + cl);
        sgb[k] = (e2 < Vi && cb < Vi) ? fvv(cb, e2) : make_double2(0.0, 0.0);
      }
    }
  };
  fetch(0);
  for (int e0 = 0; e0 < Vi; e0 += MKC) {
    __syncthreads();                                 // the previous chunk is no longer read
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int q = tid + SBLK * k;
      const int r = q / MKC, c = q % MKC, rl = q % MT, cl = q / MT;
      const int pa = (int)a0 + r, pb = (int)b0 + r, e = e0 + c, el = e0 + cl;
      // each element is stored once: by the lower pass where p > e is in range, else by the upper pass
      if (!(pa > e && pa < Vi && e < Vi)) XA[r][c] = sxa_u[k];
      if ((int)a0 + rl > el && (int)a0 + rl < Vi && el < Vi) XA[rl][cl] = sxa_l[k];
      const int rr = transpose_f ? q / MT : q % MKC, cc = transpose_f ? q % MT : q / MKC;
      GA[rr][cc] = sga[k];
      if (!diag) {
        if (!(pb > e && pb < Vi && e < Vi)) XB[r][c] = sxb_u[k];
        if ((int)b0 + rl > el && (int)b0 + rl < Vi && el < Vi) XB[rl][cl] = sxb_l[k];
        GB[rr][cc] = sgb[k];
      }
    }
    __syncthreads();
    if (e0 + MKC < Vi) fetch(e0 + MKC);              // in flight during the products below
#pragma unroll
    for (int ks = 0; ks < MKC / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4), ra = wm * 16 + (lane & 15), cb = wn * 16 + (lane & 15);
      const double2 xa = XA[ra][kk], gb = diag ? GA[kk][cb] : GB[kk][cb];   // sum_e x[a,e] Fvv(b,e)
      acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(xa.x, gb.x, acc_re, 0, 0, 0);
      acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(xa.x, gb.y, acc_im, 0, 0, 0);
      acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(-xa.y, gb.y, acc_re, 0, 0, 0);
      acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(xa.y, gb.x, acc_im, 0, 0, 0);
      if (!diag) {
        const double2 ga = GA[kk][ra], xb = XB[cb][kk];   // - sum_e Fvv(a,e) x[b,e]
        acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(-ga.x, xb.x, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(-ga.x, xb.y, acc_im, 0, 0, 0);
        acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(ga.y, xb.y, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(-ga.y, xb.x, acc_im, 0, 0, 0);
      }
    }
  }
  if (diag) {
    // every wave holds Y over its quarter of the tile (the lower-left one too): r = Y[a][b] - Y[b][a]
    __syncthreads();                                 // the last chunk's tiles (aliased by Ysh) are read
#pragma unroll
    for (int r = 0; r < 4; ++r)
      Ysh[wm * 16 + (lane >> 4) + 4 * r][wn * 16 + (lane & 15)] = make_double2(acc_re[r], acc_im[r]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double2 yt = Ysh[wn * 16 + (lane & 15)][wm * 16 + (lane >> 4) + 4 * r];
      acc_re[r] -= yt.x;
      acc_im[r] -= yt.y;
    }
  }

  const double* eps = (const double*)I.eps;
  const double2* drv = (const double2*)I.drv;
  const double2* Z = (const double2*)(ws + ((const Header*)ws)->z_off) + zoff[it.inst];
  const double eij = eps[i] + eps[j];
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
      const int64_t a = a0 + wm * 16 + (lane >> 4) + 4 * r, b = b0 + wn * 16 + (lane & 15);
      if (b >= V || a >= b) continue;
      const int64_t Pab = pidx(a, b, V);
      const double2 d = drv[((i * O + j) * V + a) * V + b];
      // -Z_ij + Z_ji: -sum_m Foo(m,j) x_im[a,b] + sum_m Foo(m,i) x_jm[a,b]
      const double2 zij = Z[(i * O + j) * Pv + Pab], zji = Z[(j * O + i) * Pv + Pab];
      const double2 rr = make_double2(s * d.x + acc_re[r] - zij.x + zji.x, s * d.y + acc_im[r] - zij.y + zji.y);
      const double D = eij - eps[O + a] - eps[O + b];
      const double2 xv = row[Pab];
      xn[(int64_t)it.pij * Pv + Pab] = make_double2(xv.x + rr.x / D, xv.y + rr.y / D);
      acc += rr.x * rr.x + rr.y * rr.y;
    }
  red[tid] = acc;
  __syncthreads();
  for (int w = SBLK / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) part[b] = red[0];
}

__global__ __launch_bounds__(SBLK) void ncp_step_kernel(const apy_mp2_nc_instance* __restrict__ insts,
                                                        const Item* __restrict__ items, int reverse,
                                                        const char* __restrict__ ws, const int64_t* __restrict__ zoff,
                                                        double* __restrict__ part) {
  ncp_step_item(blockIdx.x, insts, items, reverse, ws, zoff, part);
}

__device__ __forceinline__ void instance_sum(int64_t k, const int64_t* __restrict__ off,
                                             const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[BLK];
  __syncthreads();
  const int64_t b = off[k], e = off[k + 1];
  double acc = 0.0;
  for (int64_t q = b + threadIdx.x; q < e; q += BLK) acc += part[q];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

__global__ __launch_bounds__(BLK) void per_instance_sum(const int64_t* __restrict__ off,
                                                        const double* __restrict__ part, double* __restrict__ out) {
  instance_sum(blockIdx.x, off, part, out);
}

// The whole solve in one cooperative launch: every resident workgroup loops over the Z work items,
// the step items and the per-instance sums of an iteration with a grid barrier after each phase, then
// reads the n residual norms and stops when sqrt(max) < tol (the test of MP2NCBatch.solve after every
// step) or after maxit steps.  Each phase is the same per-item code as the three kernels of
// apy_mp2_nc_batched_step, so the iterates are bit-identical to the step-by-step solve.  An odd step
// count leaves the amplitudes in x_new; they are copied back to x before the kernel ends.
__global__ __launch_bounds__(SBLK) void nc_persistent_kernel(char* __restrict__ ws,
                                                             const apy_mp2_nc_instance* __restrict__ insts,
                                                             const Item* __restrict__ items,
                                                             const int64_t* __restrict__ off,
                                                             const int64_t* __restrict__ zoff,
                                                             double* __restrict__ part, double* __restrict__ rnorm,
                                                             int n, int64_t n_items, double tol, int maxit,
                                                             int64_t* __restrict__ iters) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  int rev = 0, it = 0;
  while (it < maxit) {
    nc_z_phase(ws, insts, zoff, rev);
    grid.sync();
    for (int64_t b = blockIdx.x; b < n_items; b += gridDim.x) ncp_step_item(b, insts, items, rev, ws, zoff, part);
    grid.sync();
    for (int64_t k = blockIdx.x; k < n; k += gridDim.x) instance_sum(k, off, part, rnorm);
    grid.sync();
    ++it;
    rev ^= 1;
    double mx = 0.0;
    for (int k = 0; k < n; ++k) mx = fmax(mx, rnorm[k]);
    if (sqrt(mx) < tol) break;
  }
  if (rev) {
    for (int k = 0; k < n; ++k) {
      const apy_mp2_nc_instance I = insts[k];
      const int64_t len = npair(I.O) * npair(I.V);
      const double2* src = (const double2*)I.x_new;
      double2* dst = (double2*)I.x;
      for (int64_t e = blockIdx.x * (int64_t)SBLK + threadIdx.x; e < len; e += (int64_t)gridDim.x * SBLK)
        dst[e] = src[e];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *iters = it;
}

// energy[k] = scale * sum_dense w_k x_k = 4 scale * sum_(i<j, a<b) w_k x_k (w and x
// antisymmetric), in two fixed-order passes: EN_B partial sums per instance
// (blockIdx.y = instance), then one block per instance adds its partials.
__global__ __launch_bounds__(BLK) void nc_energy_partial(const apy_mp2_nc_instance* __restrict__ insts, int use_x_new,
                                                         double2* __restrict__ epart) {
  __shared__ double2 red[BLK];
  const apy_mp2_nc_instance I = insts[blockIdx.y];
  const double2* w = (const double2*)I.w;
  const double2* x = (const double2*)(use_x_new ? I.x_new : I.x);
  const int64_t O = I.O, V = I.V, Pv = npair(V), N = npair(O) * Pv;
  double2 acc = make_double2(0.0, 0.0);
  if (w)
    for (int64_t e = blockIdx.x * (int64_t)BLK + threadIdx.x; e < N; e += (int64_t)EN_B * BLK) {
      int64_t i, j, a, b;
      pdec(e / Pv, O, i, j);
      pdec(e % Pv, V, a, b);
      cfma(acc, w[((i * O + j) * V + a) * V + b], x[e]);
    }
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
  if (threadIdx.x == 0) out[blockIdx.x] = make_double2(4.0 * scale * red[0].x, 4.0 * scale * red[0].y);
}

__global__ __launch_bounds__(BLK) void t2_pack_kernel(int64_t O, int64_t V, const double2* __restrict__ dense,
                                                      double2* __restrict__ packed) {
  const int64_t Pv = npair(V), total = npair(O) * Pv;
  for (int64_t t = blockIdx.x * (int64_t)BLK + threadIdx.x; t < total; t += (int64_t)gridDim.x * BLK) {
    int64_t i, j, a, b;
    pdec(t / Pv, O, i, j);
    pdec(t % Pv, V, a, b);
    packed[t] = dense[((i * O + j) * V + a) * V + b];
  }
}

__global__ __launch_bounds__(BLK) void t2_unpack_kernel(int64_t O, int64_t V, const double2* __restrict__ packed,
                                                        double2* __restrict__ dense) {
  const int64_t total = O * O * V * V, Pv = npair(V);
  for (int64_t t = blockIdx.x * (int64_t)BLK + threadIdx.x; t < total; t += (int64_t)gridDim.x * BLK) {
    int64_t r = t;
    const int64_t b = r % V; r /= V;
    const int64_t a = r % V; r /= V;
    const int64_t j = r % O;
    const int64_t i = r / O;
    dense[t] = xget(packed, O, Pv, V, i, j, a, b);
  }
}

}  // namespace

extern "C" {

int64_t apy_t2_packed_length(int64_t O, int64_t V) {
  if (O < 0 || V < 0) return 0;
  return npair(O) * npair(V);
}

int apy_t2_pack(int64_t O, int64_t V, const void* dense, void* packed, void* stream) {
  if (O < 0 || V < 0 || !dense || !packed || dense == packed) return APY_E_ARG;
  const int64_t total = npair(O) * npair(V);
  if (total == 0) return APY_OK;
  t2_pack_kernel<<<apy_grid1d(total, 1 << 20), BLK, 0, (hipStream_t)stream>>>(O, V, (const double2*)dense,
                                                                              (double2*)packed);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_t2_unpack(int64_t O, int64_t V, const void* packed, void* dense, void* stream) {
  if (O < 0 || V < 0 || !dense || !packed || dense == packed) return APY_E_ARG;
  const int64_t total = O * O * V * V;
  if (total == 0) return APY_OK;
  t2_unpack_kernel<<<apy_grid1d(total, 1 << 20), BLK, 0, (hipStream_t)stream>>>(O, V, (const double2*)packed,
                                                                                (double2*)dense);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

size_t apy_mp2_nc_batched_workspace(int n, const apy_mp2_nc_instance* inst) {
  if (n < 1 || !inst) return 0;
  int64_t ze = 0;
  const int64_t zi = count_zitems(n, inst, &ze);
  return layout(n, count_items(n, inst), zi, ze).total;
}

int apy_mp2_nc_batched_setup(int n, const apy_mp2_nc_instance* inst, void* workspace, size_t ws_bytes,
                             int64_t* n_items, void* stream) {
  if (n < 1 || n > 65535 || !inst || !n_items) return APY_E_ARG;
  for (int k = 0; k < n; ++k) {
    const apy_mp2_nc_instance& I = inst[k];
    if (I.O < 1 || I.V < 1 || I.O > 32767 || I.V > 32767 || !I.fock || !I.eps || !I.drv || !I.x || !I.x_new ||
        I.x == I.x_new)
      return APY_E_ARG;
  }
  const int64_t items = count_items(n, inst);
  int64_t zelems = 0;
  const int64_t zitems = count_zitems(n, inst, &zelems);
  if (items < 1 || items > 2147483647 || zitems > 2147483647) return APY_E_ARG;
  const WsLayout L = layout(n, items, zitems, zelems);
  if (!workspace || ws_bytes < L.total) return APY_E_WORKSPACE;
  std::vector<Item> it;
  it.reserve(items);
  std::vector<ZItem> zit;
  zit.reserve(zitems);
  std::vector<int64_t> off(n + 1, 0), zoff(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    off[k] = (int64_t)it.size();
    zoff[k + 1] = zoff[k];
    if (inst[k].O < 2 || inst[k].V < 2) continue;
    const int32_t nt = (int32_t)ntiles(inst[k].V);
    const int64_t Pv = npair(inst[k].V);
    for (int32_t p = 0; p < (int32_t)npair(inst[k].O); ++p)
      for (int32_t ta = 0; ta < nt; ++ta)
        for (int32_t tb = ta; tb < nt; ++tb) it.push_back(Item{k, p, ta, tb});
    for (int32_t i = 0; i < (int32_t)inst[k].O; ++i)
      for (int64_t p0 = 0; p0 < Pv; p0 += ZS) zit.push_back(ZItem{k, i, (int32_t)p0, 0});
    zoff[k + 1] = zoff[k] + inst[k].O * inst[k].O * Pv;
  }
  off[n] = (int64_t)it.size();
  Header h{n, items, zitems, (int64_t)L.zitem, (int64_t)L.z};
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  // pageable sources: each copy returns once the host buffer has been read
  APY_HIP_TRY(hipMemcpyAsync(w, &h, sizeof(h), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.inst, inst, n * sizeof(apy_mp2_nc_instance), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.item, it.data(), items * sizeof(Item), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipMemcpyAsync(w + L.zoff, zoff.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (zitems > 0)
    APY_HIP_TRY(hipMemcpyAsync(w + L.zitem, zit.data(), zitems * sizeof(ZItem), hipMemcpyHostToDevice, st));
  APY_HIP_TRY(hipStreamSynchronize(st));
  *n_items = items;
  return APY_OK;
}

int apy_mp2_nc_batched_step(int n, int64_t n_items, int reverse, void* rnorm, void* workspace, void* stream) {
  if (n < 1 || n > 65535 || n_items < 1 || n_items > 2147483647 || !rnorm || !workspace) return APY_E_ARG;
  const WsLayout L = layout(n, n_items);
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)(w + L.part);
  const auto* insts = (const apy_mp2_nc_instance*)(w + L.inst);
  const int64_t* zoff = (const int64_t*)(w + L.zoff);
  // the Z work list is grid-strided over at most 4096 workgroups
  nc_z_kernel<<<(unsigned)std::min<int64_t>(n_items, 4096), BLK, 0, st>>>(w, insts, zoff, reverse);
  APY_LAUNCH_CHECK();
  ncp_step_kernel<<<(unsigned)n_items, SBLK, 0, st>>>(insts, (const Item*)(w + L.item), reverse, w, zoff, part);
  APY_LAUNCH_CHECK();
  per_instance_sum<<<(unsigned)n, BLK, 0, st>>>((const int64_t*)(w + L.off), part, (double*)rnorm);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

int apy_mp2_nc_batched_solve(int n, int64_t n_items, double tol, int maxit, void* rnorm, void* iterations,
                             void* workspace, void* stream) {
  if (n < 1 || n > 65535 || n_items < 1 || n_items > 2147483647 || maxit < 1 || !(tol >= 0.0) || !rnorm ||
      !iterations || !workspace)
    return APY_E_ARG;
  const WsLayout L = layout(n, n_items);
  char* w = (char*)workspace;
  int dev = 0, ncu = 0, per_cu = 0;
  APY_HIP_TRY(hipGetDevice(&dev));
  APY_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  APY_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, nc_persistent_kernel, SBLK, 0));
  if (per_cu < 1 || ncu < 1) return APY_E_HIP;
  // every workgroup resident at once (a cooperative launch refuses a larger grid)
  const int64_t grid = std::min<int64_t>((int64_t)per_cu * ncu, n_items);
  const apy_mp2_nc_instance* insts = (const apy_mp2_nc_instance*)(w + L.inst);
  const Item* items = (const Item*)(w + L.item);
  const int64_t* off = (const int64_t*)(w + L.off);
  const int64_t* zoff = (const int64_t*)(w + L.zoff);
  double* part = (double*)(w + L.part);
  double* rn = (double*)rnorm;
  int64_t* it = (int64_t*)iterations;
  void* args[] = {&w, &insts, &items, &off, &zoff, &part, &rn, &n, &n_items, &tol, &maxit, &it};
  APY_HIP_TRY(hipLaunchCooperativeKernel((const void*)nc_persistent_kernel, dim3((unsigned)grid), dim3(SBLK), args, 0,
                                         (hipStream_t)stream));
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
