/*
 * apyib_amd — C ABI of the MI355X (gfx950) correlated-wavefunction hot path.
 *
 * The reference (apyib) is pure Python/NumPy; it has no FFI.  Each entry
 * point below replaces one NumPy/opt_einsum operation or loop of the
 * reference and cites it (file:line in the apyib source tree).  The Python
 * classes in apyib_amd/ (ci_wfn, mp2_wfn, utils, energy) keep the
 * reference's call signatures and reach the GPU only through these symbols.
 *
 * Conventions
 *  - complex128 is interleaved (re, im) doubles, row-major, indices in the
 *    reference's order ([i,j,a,b], [a,b,c,d], ...).
 *  - All pointers are device pointers owned by the caller, except where a
 *    parameter says "host".
 *  - `stream` is a hipStream_t (NULL = legacy default stream).  Calls only
 *    enqueue work; they never synchronise unless documented.
 *  - Return value: APY_OK (0) or a negative APY_E_* code.  No C++ exception
 *    crosses this boundary.
 *  - Spin-orbital (SO) ordering is the reference's: occupied first,
 *    alpha/beta interleaved (utils.py:294-311, 332-363).
 */
#ifndef APYIB_AMD_H
#define APYIB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APY_OK 0
#define APY_E_ARG (-1)       /* invalid argument (sizes, layout, null)   */
#define APY_E_HIP (-2)       /* a HIP runtime call or launch failed       */
#define APY_E_WORKSPACE (-3) /* caller workspace smaller than required    */
#define APY_E_SINGULAR (-4)  /* singular linear system (DIIS)             */

/* Library version (major*10000 + minor*100 + patch). */
int apy_version(void);

/* ------------------------------------------------------------------------
 * Complex GEMM on FP64 MFMA (v_mfma_f64_16x16x4f64) as real-block products.
 * Replaces every opt_einsum.contract that is a plain matrix product once
 * indices are grouped: ci_wfn.py:310-333, 466-482; utils.py:252, 274-277,
 * 381 (mo overlap), 121 and 138 (DIIS).
 *
 *   C_b = alpha * op(A_b) op(B_b) + beta * C_b     for b in [0, batch)
 *   reduce_batch = 1:  C = alpha * sum_b op(A_b) op(B_b) + beta * C
 *
 * A layout: APY_MK  A(m,k) = A[m*lda + k];  APY_KM  A(m,k) = A[k*lda + m]
 * B layout: APY_KN  B(k,n) = B[k*ldb + n];  APY_NK  B(k,n) = B[n*ldb + k]
 * conj_a / conj_b conjugate the operand as it is read.
 * ---------------------------------------------------------------------- */
#define APY_MK 0
#define APY_KM 1
#define APY_KN 0
#define APY_NK 1

typedef struct apy_gemm_desc {
  int64_t m, n, k;
  int a_layout, b_layout;
  int conj_a, conj_b;
  double alpha_re, alpha_im;
  double beta_re, beta_im;
  const void* a; int64_t lda; int64_t stride_a;
  const void* b; int64_t ldb; int64_t stride_b;
  void* c; int64_t ldc; int64_t stride_c;
  int64_t batch;      /* >= 1 */
  int reduce_batch;   /* 0 or 1 */
  int split_k;        /* 0 = choose automatically, else number of K slices */
} apy_gemm_desc;

/* Workspace (bytes) apy_zgemm needs for this descriptor (0 if none). */
size_t apy_zgemm_workspace(const apy_gemm_desc* d);
int apy_zgemm(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* APYIB_AMD_H */
