// MP2 amplitudes and energies (mp2_wfn.py:42-87).
//
// Spin-orbital:  t[i,j,a,b] = <ab||ij> / D_ijab,  E = 1/4 sum <ij||ab> t
// Spatial:       t[i,j,a,b] = (ai|bj) / D_ijab,
//                E = sum (2 (ia|jb) - (ib|ja)) t        (chemists' MO integrals)
// D_ijab = e_i + e_j - e_a - e_b.  Both are one pass over O^2V^2 elements:
// read the integrals, write t, reduce E -- HBM-bound.
//
// Batched form: instances of different (O, V) in one launch (up to 40 per
// launch).
// Each instance's element range is cut into chunks of CHUNK elements; the
// grid walks a flat chunk list (chunk -> instance by binary search over the
// chunk offsets), so a block never straddles two instances and its energy
// partial sum belongs to one instance.  A second kernel adds each
// instance's partials in a fixed order (deterministic).
#include "common.h"

namespace {

constexpr int CHUNK = 4096;  // elements per block-chunk (16 per thread)

struct Mp2Inst {
  int64_t O, V;
  const double* eps;
  const double2* vvoo_t;
  const double2* oovv;
  double2* t2;
  int64_t chunk0, nchunk;  // first chunk index and chunk count
};

__device__ inline double2 block_sum(double2 v) {
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = v.x;
  si[threadIdx.x] = v.y;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sr[threadIdx.x] += sr[threadIdx.x + w]; si[threadIdx.x] += si[threadIdx.x + w]; }
    __syncthreads();
  }
  double2 r = make_double2(sr[0], si[0]);
  __syncthreads();
  return r;
}

// Up to MAXI instances travel by value in the kernel arguments (no
// device-side table, no host synchronisation); larger batches are launched
// in groups of MAXI.
constexpr int MAXI = 40;
struct Mp2Batch {
  int n;
  int64_t nchunks;
  Mp2Inst inst[MAXI];
};

__global__ __launch_bounds__(256) void mp2_so_chunks(Mp2Batch B, double2* __restrict__ partial) {
  for (int64_t c = blockIdx.x; c < B.nchunks; c += gridDim.x) {
    // instance owning chunk c
    int lo = 0, hi = B.n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (B.inst[mid].chunk0 <= c) lo = mid; else hi = mid - 1;
    }
    const Mp2Inst& I = B.inst[lo];
    const int64_t O = I.O, V = I.V, n = O * O * V * V;
    const int64_t beg = (c - I.chunk0) * CHUNK;
    const int64_t end = min(n, beg + CHUNK);
    double er = 0.0, ei = 0.0;
    for (int64_t e = beg + threadIdx.x; e < end; e += blockDim.x) {
      int64_t r = e;
      const int64_t b = r % V; r /= V;
      const int64_t a = r % V; r /= V;
      const int64_t j = r % O;
      const int64_t i = r / O;
      const double D = I.eps[i] + I.eps[j] - I.eps[O + a] - I.eps[O + b];
      const double2 d = I.vvoo_t[e];
      const double2 t = make_double2(d.x / D, d.y / D);
      I.t2[e] = t;
      const double2 w = I.oovv[e];
      er += w.x * t.x - w.y * t.y;
      ei += w.x * t.y + w.y * t.x;
    }
    double2 s = block_sum(make_double2(0.25 * er, 0.25 * ei));
    if (threadIdx.x == 0) partial[c] = s;
  }
}

__global__ void mp2_finalize(Mp2Batch B, const double2* __restrict__ partial, double2* energies) {
  for (int k = blockIdx.x; k < B.n; k += gridDim.x) {
    const Mp2Inst& I = B.inst[k];
    double sr = 0.0, si = 0.0;
    for (int64_t q = threadIdx.x; q < I.nchunk; q += blockDim.x) { sr += partial[I.chunk0 + q].x; si += partial[I.chunk0 + q].y; }
    double2 s = block_sum(make_double2(sr, si));
    if (threadIdx.x == 0) energies[k] = s;
  }
}

// Spatial closed-shell MP2 from chemists' MO integrals eri[m][m][m][m]
// (active orbitals; occupied first): t = (ai|bj)/D, E = sum (2(ia|jb)-(ib|ja)) t.
__global__ __launch_bounds__(256) void mp2_spatial_kernel(int64_t O, int64_t V, const double* __restrict__ eps,
                                                          const double2* __restrict__ eri, double2* __restrict__ t2,
                                                          double2* __restrict__ partial) {
  const int64_t m = O + V, n = O * O * V * V;
  double er = 0.0, ei = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = e;
    const int64_t b = r % V; r /= V;
    const int64_t a = r % V; r /= V;
    const int64_t j = r % O;
    const int64_t i = r / O;
    const int64_t A = O + a, B = O + b;
    const double D = eps[i] + eps[j] - eps[A] - eps[B];
    const double2 d = eri[((A * m + i) * m + B) * m + j];  // (ai|bj)
    const double2 t = make_double2(d.x / D, d.y / D);
    t2[e] = t;
    const double2 x = eri[((i * m + A) * m + j) * m + B];  // (ia|jb)
    const double2 y = eri[((i * m + B) * m + j) * m + A];  // (ib|ja)
    const double wr = 2.0 * x.x - y.x, wi = 2.0 * x.y - y.y;
    er += wr * t.x - wi * t.y;
    ei += wr * t.y + wi * t.x;
  }
  double2 s = block_sum(make_double2(er, ei));
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void sum_partials(const double2* __restrict__ partial, int np, double2* out) {
  double sr = 0.0, si = 0.0;
  for (int q = threadIdx.x; q < np; q += blockDim.x) { sr += partial[q].x; si += partial[q].y; }
  double2 s = block_sum(make_double2(sr, si));
  if (threadIdx.x == 0) *out = s;
}

constexpr int SPATIAL_BLOCKS = 1024;

int64_t total_chunks(const apy_mp2_instance* inst, int n) {
  int64_t nch = 0;
  for (int k = 0; k < n; ++k) {
    const int64_t e = inst[k].O * inst[k].O * inst[k].V * inst[k].V;
    nch += (e + CHUNK - 1) / CHUNK;
  }
  return nch;
}

}  // namespace

extern "C" {

size_t apy_mp2_so_batched_workspace(int n, const apy_mp2_instance* inst) {
  if (n < 1 || !inst) return 0;
  const int64_t nch = total_chunks(inst, n);
  return (size_t)(nch > 0 ? nch : 1) * sizeof(double2);
}

int apy_mp2_so_batched(int n, const apy_mp2_instance* inst, void* energies, void* workspace, size_t ws_bytes,
                       void* stream) {
  if (n < 1 || !inst || !energies) return APY_E_ARG;
  for (int k = 0; k < n; ++k) {
    const apy_mp2_instance& I = inst[k];
    if (I.O < 1 || I.V < 1 || !I.eps || !I.vvoo_t || !I.oovv || !I.t2) return APY_E_ARG;
  }
  if (!workspace || ws_bytes < apy_mp2_so_batched_workspace(n, inst)) return APY_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double2* partial = (double2*)workspace;
  for (int g0 = 0; g0 < n; g0 += MAXI) {
    const int gn = n - g0 < MAXI ? n - g0 : MAXI;
    Mp2Batch B;
    B.n = gn;
    int64_t c0 = 0;
    for (int k = 0; k < gn; ++k) {
      const apy_mp2_instance& I = inst[g0 + k];
      const int64_t e = I.O * I.O * I.V * I.V;
      B.inst[k] = Mp2Inst{I.O, I.V, I.eps, (const double2*)I.vvoo_t, (const double2*)I.oovv, (double2*)I.t2, c0,
                          (e + CHUNK - 1) / CHUNK};
      c0 += B.inst[k].nchunk;
    }
    B.nchunks = c0;
    const unsigned grid = (unsigned)(c0 < 8192 ? (c0 > 0 ? c0 : 1) : 8192);
    mp2_so_chunks<<<grid, 256, 0, s>>>(B, partial);
    APY_LAUNCH_CHECK();
    mp2_finalize<<<(unsigned)gn, 256, 0, s>>>(B, partial, (double2*)energies + g0);
    APY_LAUNCH_CHECK();
    // the next group may reuse `partial` only after this group's finalize: stream order
  }
  return APY_OK;
}

int apy_mp2_so(int64_t O, int64_t V, const double* eps, const void* vvoo_t, const void* oovv, void* t2,
               void* energy, void* workspace, size_t ws_bytes, void* stream) {
  apy_mp2_instance I{O, V, eps, vvoo_t, oovv, t2};
  return apy_mp2_so_batched(1, &I, energy, workspace, ws_bytes, stream);
}

size_t apy_mp2_spatial_workspace(void) { return SPATIAL_BLOCKS * sizeof(double2); }

int apy_mp2_spatial(int64_t O, int64_t V, const double* eps, const void* eri_mo, void* t2, void* energy,
                    void* workspace, size_t ws_bytes, void* stream) {
  if (O < 1 || V < 1 || !eps || !eri_mo || !t2 || !energy) return APY_E_ARG;
  if (!workspace || ws_bytes < SPATIAL_BLOCKS * sizeof(double2)) return APY_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  mp2_spatial_kernel<<<SPATIAL_BLOCKS, 256, 0, s>>>(O, V, eps, (const double2*)eri_mo, (double2*)t2,
                                                    (double2*)workspace);
  APY_LAUNCH_CHECK();
  sum_partials<<<1, 256, 0, s>>>((const double2*)workspace, SPATIAL_BLOCKS, (double2*)energy);
  APY_LAUNCH_CHECK();
  return APY_OK;
}

}  // extern "C"
