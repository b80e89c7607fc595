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
#define APY_E_TOO_LARGE (-5) /* the problem needs more than the library's
                                static scratch (no-allocation entry point) */

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
  int tile;           /* 0 = choose automatically, else tile config 1..9 */
  int gauss3m;        /* 1: three real products per complex multiply-add
                         (re = ar br - ai bi, im = (ar + ai)(br + bi) - ar br
                         - ai bi): 3/4 of the matrix-core work, with the
                         imaginary part's rounding relative to |a||b| rather
                         than to its own size (for operands whose imaginary
                         parts are not small next to their real parts) */
} apy_gemm_desc;

/* Workspace (bytes) apy_zgemm needs for this descriptor (0 if none); it
 * covers every plan the tuner may pick for the shape. */
size_t apy_zgemm_workspace(const apy_gemm_desc* d);
int apy_zgemm(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, void* stream);
/* The (tile config, split) apy_zgemm would use for d. Host only. */
int apy_zgemm_plan(const apy_gemm_desc* d, int* tile, int* split);
/* Autotuner: time the candidate plans for d's shape on d's operands (C is
 * overwritten), keep the fastest in a process-wide cache (mutex guarded)
 * that later calls with the same shape use.  Synchronises the stream.
 * Outputs (may be NULL): chosen tile, split, and its time in ms. */
int apy_zgemm_tune(const apy_gemm_desc* d, void* workspace, size_t ws_bytes, int reps, void* stream, int* tile,
                   int* split, double* ms);
int apy_zgemm_tune_clear(void);
/* Set the cached plan for d's shape (m, n, k, layouts, batch, reduce) on
 * the current device, e.g. from a saved table.  Host only. */
int apy_zgemm_tune_set(const apy_gemm_desc* d, int tile, int split);
/* Export the current device's cached plans, 9 int64 per entry:
 * m, n, k, a_layout, b_layout, batch, reduce_batch, tile, split.  *count is
 * the number of entries (rows are written up to capacity).  Host only. */
int apy_zgemm_tune_export(int64_t* rows, int64_t capacity, int64_t* count);

/* ------------------------------------------------------------------------
 * Antisymmetrized spin-orbital integral blocks <pq||rs> in the layouts the
 * sigma kernels read.  O = occupied, V = virtual spin orbitals (n = O+V,
 * occupied first), Po = O(O-1)/2, Pv = V(V-1)/2.  A pair index (p<q) of N
 * orbitals is  p*(2N-p-1)/2 + (q-p-1)  (row-major upper triangle).
 * The reference forms each block per term as ERI_SO[..] - ERI_SO.swapaxes
 * (2,3)[..] (ci_wfn.py:312-333, 355); here each is built once.
 * Any pointer may be NULL when the caller does not need that block.
 * ---------------------------------------------------------------------- */
typedef struct apy_so_ints {
  int64_t O, V;
  void* vvvv_p;  /* [Pv][Pv]      <ab||cd>, a<b rows, c<d columns  (ci_wfn.py:329)      */
  void* ring;    /* [O][V][O][V]  ring[(j,b),(k,c)] = <kb||cj>      (ci_wfn.py:312,330-333) */
  void* oooo_p;  /* [Po][Po]      <kl||ij>, (i<j) rows, (k<l) cols  (ci_wfn.py:328)      */
  void* ovoo_p;  /* [O][V][Po]    <kb||ij>, i<j                     (ci_wfn.py:320-321)  */
  void* vvvo_p;  /* [Pv][O][V]    <ab||cj> stored at [(a<b)][j][c]  (ci_wfn.py:322-323)  */
  void* vovv_p;  /* [V][O][Pv]    <aj||cb> stored at [a][j][(c<b)]  (ci_wfn.py:314)      */
  void* ooov;    /* [O][O][O][V]  <kj||ib>                          (ci_wfn.py:315)      */
  void* vvoo_t;  /* [O][O][V][V]  <ab||ij> stored at [i][j][a][b]   (ci_wfn.py:319)      */
  void* oovv;    /* [O][O][V][V]  <ij||ab>                          (ci_wfn.py:355)      */
} apy_so_ints;

/* Number of complex elements of each block, in the order of the struct
 * (vvvv_p, ring, oooo_p, ovoo_p, vvvo_p, vovv_p, ooov, vvoo_t, oovv). Host. */
int apy_so_ints_sizes(int64_t O, int64_t V, int64_t sizes_out[9]);

/* Blocks from a full antisymmetrized tensor A[n][n][n][n] = <pq||rs>. */
int apy_so_ints_from_full(int64_t n, const void* A, const apy_so_ints* ints, void* stream);

/* Blocks from spatial MO integrals in chemists' order, eri_mo[m][m][m][m]
 * = (pq|rs) over the active orbitals (utils.py:258-279), fused with the
 * spin expansion of compute_ERI_SO (utils.py:317-365), the Dirac swap
 * (ci_wfn.py:284) and the antisymmetrization of each term.  SO p is
 * spatial p/2 with spin p%2; O = occupied SOs.  Requires ints->O+ints->V
 * == 2m. */
int apy_so_ints_from_mo(int64_t m, const void* eri_mo, const apy_so_ints* ints, void* stream);

/* Blocks of the seeded synthetic <pq||rs> (definition in
 * apyib_amd/synthetic.py): std = complex std of the draw.  With
 * accumulate = 1 the blocks are incremented by scale * (draw) instead of
 * overwritten; real_only keeps the real part of the draw (a real,
 * symmetric-orbital displacement).  tag selects the stream (1 = base). */
int apy_synth_so_ints(uint64_t seed, uint64_t tag, double std, int accumulate, double scale,
                      int real_only, const apy_so_ints* ints, void* stream);

/* Spin-orbital Fock matrix F_SO[p,q] = F_MO[p/2,q/2] delta(spin p, spin q)
 * (compute_F_SO, utils.py:283-313): f_mo is [m][m], f_so [2m][2m]. */
int apy_so_fock_from_mo(int64_t m, const void* f_mo, void* f_so, void* stream);

/* ------------------------------------------------------------------------
 * Spin-adapted closed-shell CISD / CID (ci_wfn.py:420-574 and :51-167) on
 * spatial orbitals: o occupied, v virtual; Dirac integrals <pq|rs> =
 * (pr|qs) of the chemists' MO tensor (utils.py:258-279, ci_wfn.py:433).
 * Amplitude vector [t1 (o*v) | t2 (o*o*v*v)], t2[i][j][a][b] with
 * t2[i][j][a][b] = t2[j][i][b][a] (the solver keeps it so; the doubles
 * residual is assembled as a half and its pair image).
 * ---------------------------------------------------------------------- */
typedef struct apy_sp_ints {
  int64_t o, v;
  void* vvvs;  /* [(a<=b)][(c<=d)]    <ab|cd> + <ab|dc> (c<d), <ab|cc>     (ci_wfn.py:475) */
  void* vvva;  /* [(a<b)][(c<d)]      <ab|cd> - <ab|dc>                   (ci_wfn.py:475) */
  void* oooo;  /* [(i,j)][(k,l)]      <kl|ij>                             (ci_wfn.py:474) */
  void* w21;   /* [2ov][(j,b)]  rows (k,c): 2<kb|cj> - <kb|jc>; rows ov+(k,c): -<kb|cj>  (:476-481) */
  void* w3;    /* [(k,c)][(j,a)]      <ka|jc>                             (ci_wfn.py:478, 482) */
  void* vvoo;  /* [(i,j)][(a,b)]      <ab|ij>  (doubles driver)            (ci_wfn.py:466) */
  void* l;     /* [(i,j)][(a,b)]      2<ij|ab> - <ij|ba>  (energy)          (ci_wfn.py:495) */
  void* vvvo;  /* [(j,a,b)][c]        <ab|cj>                             (ci_wfn.py:467-468) */
  void* vooo;  /* [(i,j,a)][k]        <ak|ij>                             (ci_wfn.py:469-470) */
  void* s1;    /* [(i,a)][(j,b)]      2<ja|bi> - <ja|ib>                  (ci_wfn.py:460) */
  void* w4;    /* [a][(j,b,c)]        2<aj|bc> - <aj|cb>                  (ci_wfn.py:462) */
  void* w5;    /* [i][(j,k,b)]        2<kj|ib> - <kj|bi>                  (ci_wfn.py:463) */
  void* w4t;   /* [b][a][(j,c)]       2<aj|bc> - <aj|cb>, optional: w4 reordered so that a step over
                  several vectors sharing these integrals forms the w4 term of all of them in one
                  product (rows (vector, i), summed over b); NULL: the per-vector w4 product */
} apy_sp_ints;

/* Complex elements of each block, in struct order.  Pairs are packed row
 * by row: (x<=y) at x(2N-x+1)/2 + y - x, (x<y) at x(2N-x-1)/2 + y - x - 1.
 * Host. */
int apy_sp_ints_sizes(int64_t o, int64_t v, int64_t sizes_out[13]);
/* Blocks from eri_mo[m][m][m][m] = (pq|rs), m = o + v (NULL blocks skipped). */
int apy_sp_ints_from_mo(int64_t m, const void* eri_mo, const apy_sp_ints* ints, void* stream);
/* Per-solve operands that depend on the Fock matrix fock[(o+v)][(o+v)]:
 * hoo[(i,j)][(k,l)] = <kl|ij> - F[k][i] d(l,j) - F[l][j] d(k,i) (o^4) and
 * g = [2 F_ov | -F_ov] (2 o v). */
int apy_cisd_sp_prepare(int64_t o, int64_t v, const void* fock, const apy_sp_ints* ints, void* hoo, void* g,
                        void* stream);

typedef struct apy_cisd_sp_args {
  int64_t o, v;
  const void* fock;         /* [n][n] MO Fock matrix, n = o + v           */
  const double* eps;        /* [n] orbital energies for D_ia, D_ijab       */
  const apy_sp_ints* ints;
  const void* hoo;          /* from apy_cisd_sp_prepare                    */
  const void* g;            /* from apy_cisd_sp_prepare                    */
  int singles;              /* 1 CISD, 0 CID (t1 block kept at zero)       */
  const void* energy;       /* device complex scalar E of the iteration    */
  const void* t;            /* [o*v + o*o*v*v] amplitudes                  */
  void* t_new;              /* t + r / D (must differ from out)            */
  void* out;                /* residual r (must differ from t)             */
  /* Optional linear-response term (resp_b null: off), added to r before
   * the update:  r += resp_b - (*resp_energy + resp_shift) * resp_t.
   * resp_b, resp_t: vectors in the amplitude layout; resp_energy: device
   * complex scalar (null reads as zero).  With resp_b = the perturbed
   * Hamiltonian's residual of (1, t) minus this Hamiltonian's driver and
   * resp_t = t, a step on dt is one iteration of the first-order amplitude
   * equations of analytic_aats.py:785-826 (CID :1582-1610). */
  const void* resp_b;
  const void* resp_t;
  const void* resp_energy;
  double resp_shift_re, resp_shift_im;
  /* Optional ring operator with the F_vv term folded in (apy_cisd_sp_fold_fock
   * of this fock and ints->w21; for apy_cisd_sp_step_multi one per instance,
   * 2 (ov)^2 apart): the step then runs one GEMM fewer.  NULL: off. */
  const void* w21_f;
  int gauss3m;              /* 1: the ladder and ring products in the three-
                               product complex form (apy_gemm_desc.gauss3m) */
} apy_cisd_sp_args;

/* Spin-orbital amplitudes of nb closed-shell (spin-adapted) solutions t [nb][o v + o^2 v^2] in the
 * spin-orbital layout of ci_wfn.py:263-416 (spin orbital 2p + spin, alpha first, utils.py:283-365):
 * t1_so [nb][2o][2v] (zero when singles = 0), t2_so [nb][2o][2o][2v][2v] with
 * T_{IJAB} = t_ijab for opposite-spin pairs matched (I,A),(J,B), -t_ijba matched crosswise, and
 * t_ijab - t_ijba for four equal spins.  For an RHF reference these are the solutions of the
 * spin-orbital CISD/CID equations (the spin-orbital Hamiltonian is spin-free). */
int apy_sp_amplitudes_to_so(int64_t o, int64_t v, int64_t nb, int singles, const void* t, void* t1_so, void* t2_so,
                            void* stream);

/* w21_f[q] = w21[q] + delta_kj F_q[o+b][o+c] on the rows (k,c) < ov, for nb
 * instances (fock + q n^2, w21 and w21_f + q 2 (ov)^2): the ring operator with
 * the F_vv term of the doubles residual (ci_wfn.py:471-472) folded in. */
int apy_cisd_sp_fold_fock(int64_t o, int64_t v, int64_t nb, const void* fock, const void* w21, void* w21_f,
                          void* stream);

/* One Jacobi step of ci_wfn.py:457-486 (CID :83-96):  r = sigma(t) - E t
 * (+ the response term above), t_new = t + r / D. */
size_t apy_cisd_sp_workspace(int64_t o, int64_t v, int singles);
int apy_cisd_sp_step(const apy_cisd_sp_args* a, void* workspace, size_t ws_bytes, void* stream);
/* E = sum 2 F_ov t1 + sum (2<ij|ab> - <ij|ba>) t2 (ci_wfn.py:495, CID :112);
 * deterministic two-pass sum into the device scalar `energy`. */
size_t apy_cisd_sp_energy_workspace(void);
int apy_cisd_sp_energy(int64_t o, int64_t v, int singles, const void* g, const void* L, const void* t, void* energy,
                       void* workspace, size_t ws_bytes, void* stream);

/* The same step for nb instances sharing fock, eps, ints, hoo and g (the
 * reference solves one perturbation at a time, analytic_aats.py:780-890,
 * :1027-1143; here the first-order amplitudes of all perturbations of a
 * molecule iterate together):  a->t, a->t_new, a->out and a->resp_b point
 * to nb vectors of the amplitude layout stored one after another;
 * a->energy and a->resp_energy to nb device complex scalars; resp_shift to
 * nb device complex scalars added to the response scalar (may be null);
 * a->resp_t is one vector shared by all instances.  The ladder and ring
 * GEMMs stack the instances along M.  nb = 1 is apy_cisd_sp_step. */
size_t apy_cisd_sp_batch_workspace(int64_t o, int64_t v, int singles, int64_t nb);
int apy_cisd_sp_step_batch(const apy_cisd_sp_args* a, int64_t nb, const void* resp_shift, void* workspace,
                           size_t ws_bytes, void* stream);
/* energy[q] of nb vectors t stored one after another (1 <= nb <= 65535);
 * workspace nb * apy_cisd_sp_energy_workspace() bytes. */
int apy_cisd_sp_energy_batch(int64_t o, int64_t v, int singles, int64_t nb, const void* g, const void* L,
                             const void* t, void* energy, void* workspace, size_t ws_bytes, void* stream);

/* The same step for nb independent systems of one (o, v) — the displaced
 * geometries and fields of a finite-difference set, which fin_diff.py:267-372
 * solves one after another (ci_wfn.py:420-574 each) — every instance with
 * its own operands, stored one instance after another: fock + q n^2,
 * eps + q n, hoo + q o^4, g + q 2ov, each block of ints + q times its
 * apy_sp_ints_sizes entry (a->ints points at instance 0's blocks), t, t_new,
 * out + q (ov + o^2 v^2), energy + q.  Every product is batched over the
 * instances.  The response term is not available (resp_b must be null).
 * apy_cisd_sp_energy_multi: energy[q] with g and L per instance (strides 2ov
 * and o^2 v^2). */
size_t apy_cisd_sp_multi_workspace(int64_t o, int64_t v, int singles, int64_t nb);
int apy_cisd_sp_step_multi(const apy_cisd_sp_args* a, int64_t nb, void* workspace, size_t ws_bytes, void* stream);
int apy_cisd_sp_energy_multi(int64_t o, int64_t v, int singles, int64_t nb, const void* g, const void* L,
                             const void* t, void* energy, void* workspace, size_t ws_bytes, void* stream);

/* apy_cisd_sp_tune: time the candidate tile/split plans of every GEMM of one
 * spin-adapted step (nb instances; multi = 0 for apy_cisd_sp_step_batch's
 * shared-integral form, 1 for apy_cisd_sp_step_multi) and keep the fastest
 * in the process-wide plan cache that later steps of the same shapes use.
 * The arguments are those of the step; workspace as for that step.  Writes
 * only the workspace.  No counterpart in the reference (its einsums have no
 * launch plans). */
int apy_cisd_sp_tune(const apy_cisd_sp_args* a, int64_t nb, int multi, void* workspace, size_t ws_bytes, int reps,
                     void* stream);

/* ------------------------------------------------------------------------
 * CISD / CID in spin orbitals: one iteration of the reference's Jacobi
 * step, or the plain sigma product (for Davidson).
 *   residual mode (ci_wfn.py:309-337, CID :210-221):
 *       r = sigma(1, t) - E t ;  t_new = t + r / D
 *   sigma mode:  out = sigma(c0, t)   (out must not alias t)
 * t and r are one vector [t1 (O*V) | t2 (O*O*V*V)], the reference's DIIS
 * vector layout (ci_wfn.py:345-346); CID uses only the t2 part but keeps
 * the same layout with a zero t1.  t2 must be antisymmetric in (i,j) and
 * in (a,b); the diagonal entries (i==j or a==b) are never written.
 * fov_coupling = 0 reproduces the reference operator, which omits the
 * P(ij)P(ab) f_bj t_ia doubles-singles coupling (exact for canonical
 * Hartree-Fock orbitals); 1 adds it, making the operator Hermitian.
 * ---------------------------------------------------------------------- */
#define APY_MODE_RESIDUAL 0
#define APY_MODE_SIGMA 1

typedef struct apy_cisd_so_args {
  int64_t O, V;
  const void* fock;        /* [n][n] spin-orbital Fock matrix           */
  const double* eps;       /* [n] orbital energies for D (ci_wfn.py:268-271) */
  const apy_so_ints* ints;
  const void* hoo;         /* [Po][Po] from apy_cisd_so_pair_operator    */
  int singles;             /* 1 CISD, 0 CID                             */
  int mode;                /* APY_MODE_RESIDUAL or APY_MODE_SIGMA       */
  int fov_coupling;
  const void* energy;      /* device complex scalar E (residual mode)   */
  double c0_re, c0_im;     /* driver weight (sigma mode; 1 in residual) */
  const void* t;           /* [O*V + O^2*V^2] input amplitudes           */
  void* t_new;             /* residual mode: t + r/D (may equal t)       */
  void* out;               /* residual (residual mode) or sigma; != t    */
  void* ev_begin;          /* optional hipEvent_t recorded just before and */
  void* ev_end;            /* just after the PPL GEMM kernel (profiling)   */
  const void* c0_dev;      /* sigma mode, optional: the driver weight read
                              from this device complex scalar instead of
                              c0_re/c0_im (no host read before the call)  */
  const void* ring_f;      /* optional [OV][OV] from apy_cisd_so_ring_fock:
                              the ring block with the F_vv term folded in,
                              which removes one GEMM from the step         */
  int packed;              /* 0: t, t_new, out are [t1 | t2 dense] (length
                              O*V + O^2*V^2); 1: [t1 | t2 over (i<j, a<b)]
                              (length O*V + Po*Pv), sigma mode only        */
  int gauss3m;             /* 1: the PPL and ring products in the three-
                              product complex form (apy_gemm_desc.gauss3m) */
} apy_cisd_so_args;

/* apy_cisd_so_step without a workspace argument and without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): no counterpart
 * in the reference.  The workspace is the library's static per-device array
 * of APY_NOALLOC_SCRATCH_BYTES; problems that need more (see
 * apy_cisd_so_noalloc_bytes, the bytes this call would use with the GEMM
 * plans in effect) return APY_E_TOO_LARGE.  All calls on one device share
 * the array: calls that overlap on one GPU race.  Same arguments and
 * results as apy_cisd_so_step. */
#define APY_NOALLOC_SCRATCH_BYTES (256u << 20)
size_t apy_cisd_so_noalloc_bytes(const apy_cisd_so_args* a);
int apy_cisd_so_step_noalloc(const apy_cisd_so_args* a, void* stream);

/* Pair-space occupied operator: <kl||ij> plus the F_oo terms of the doubles
 * residual (ci_wfn.py:326-328), [(i<j)][(k<l)]; built once per solve. */
int apy_cisd_so_pair_operator(int64_t O, int64_t V, const void* fock, const void* oooo_p, void* hoo,
                              void* stream);
/* ring_f[(j,b)][(k,c)] = <kb||cj> + 1/2 delta_jk F[O+b][O+c]: the ring block
 * with the F_vv terms of the doubles residual (ci_wfn.py:324-325) folded in.
 * With it the ring product X = T' ring_f^T carries 1/2 sum_c t_ijac F_bc in
 * X[(i,a),(j,b)], whose four P(ij)P(ab) images sum to P(ab) sum_c t_ijac F_bc.
 * Built once per Fock matrix; ring_f must not alias ring. */
int apy_cisd_so_ring_fock(int64_t O, int64_t V, const void* fock, const void* ring, void* ring_f, void* stream);
size_t apy_cisd_so_workspace(int64_t O, int64_t V, int singles);
int apy_cisd_so_step(const apy_cisd_so_args* a, void* workspace, size_t ws_bytes, void* stream);
/* Autotune every GEMM of a step for this (O, V) on the operands in `a`
 * (apy_zgemm_tune; only workspace is written).  Synchronises the stream. */
int apy_cisd_so_tune(const apy_cisd_so_args* a, void* workspace, size_t ws_bytes, int reps, void* stream);

/* Sigma products (SIGMA mode) of nb instances sharing `ints` and (O, V),
 * each with its own Fock matrix (the +/- field-displaced copies of one
 * system, fin_diff.py:335-370).  The particle-particle ladder, ring, vvvo.t1
 * and ring.t1 products run as single GEMMs with the instances stacked, so
 * each integral block is read once per batch.  Host arrays of nb device
 * pointers; c0 is a host array of nb (re, im) pairs.  The optional events
 * bracket the stacked PPL kernel. */
typedef struct apy_cisd_so_sigma_batch_args {
  int64_t O, V;
  int nb;
  int singles, fov_coupling;
  const apy_so_ints* ints;
  const void* const* fock;
  const void* const* hoo;
  const double* c0;
  const void* const* t;
  void* const* out;
  void* ev_begin;
  void* ev_end;
  int gauss3m;             /* as in apy_cisd_so_args */
} apy_cisd_so_sigma_batch_args;
size_t apy_cisd_so_sigma_batch_workspace(int64_t O, int64_t V, int nb, int singles);
int apy_cisd_so_sigma_batch(const apy_cisd_so_sigma_batch_args* a, void* workspace, size_t ws_bytes, void* stream);

/* The Fock-linear part L(F) of the Hermitian CISD sigma (SIGMA mode with
 * fov_coupling) on one packed vector [c1 | c2 over (i<j, a<b) | c0], for nf
 * one-electron matrices at once: out_f = L(fock[f]) t.  L collects the F_SO
 * terms of ci_wfn.py:309-334 (r_T1: F_vo c0, F_oo, F_vv, F_ov.t2; r_T2: the
 * two F_vv and the two F_oo contractions) and the Hermitian f_ov couplings,
 * so for a perturbation dF that leaves <pq||rs> unchanged (the field-
 * displaced instances of fin_diff.py:335-370) sigma(F0 + dF) = sigma(F0) +
 * L(dF) exactly.  hoo[f] is apy_cisd_so_pair_operator(O, V, fock[f], NULL)
 * (the F_oo part alone).  No counterpart in the reference: it lets the
 * instances of a field set share one subspace and one full sigma per
 * direction. */
typedef struct apy_cisd_so_fock_sigma_args {
  int64_t O, V;
  int nf;
  const void* const* fock;  /* nf device pointers, [n][n] each            */
  const void* const* hoo;   /* nf device pointers, [Po][Po] each          */
  const void* t;            /* packed input, apy_ci_packed_length(O, V)   */
  void* const* out;         /* nf packed outputs, none equal to t         */
} apy_cisd_so_fock_sigma_args;
size_t apy_cisd_so_fock_sigma_workspace(int64_t O, int64_t V);
int apy_cisd_so_fock_sigma(const apy_cisd_so_fock_sigma_args* a, void* workspace, size_t ws_bytes, void* stream);

/* E = sum_ia t1 f_ia + 1/4 sum t2 <ij||ab> (ci_wfn.py:355; CID :237 with
 * singles = 0), written to the device complex scalar `energy`. */
int apy_cisd_so_energy(int64_t O, int64_t V, int singles, const void* fock, const void* oovv, const void* t,
                       void* energy, void* workspace, size_t ws_bytes, void* stream);
/* The same energy of a packed vector [t1 | t2 over (i<j, a<b)]: the four
 * antisymmetric images of each packed pair give 1/4 sum = sum over i<j, a<b. */
int apy_cisd_so_energy_packed(int64_t O, int64_t V, int singles, const void* fock, const void* oovv,
                              const void* t, void* energy, void* workspace, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * AO -> MO transforms, Fock build, overlaps, phase (utils.py:217-279, 370-446)
 * AO arrays are complex128 (a real AO tensor is passed with zero imaginary
 * parts).  C is [nbf][ncol] row-major.
 * ---------------------------------------------------------------------- */
/* (pq|rs) = sum conj(C_mp) C_nq conj(C_lr) C_gs (mn|lg)  (compute_ERI_MO,
 * utils.py:258-279): four quarter transforms on the matrix cores.
 * eri_ao [nbf]^4, C [nbf][nact] (active columns), eri_mo [nact]^4. */
size_t apy_eri_mo_workspace(int64_t nbf, int64_t nact);
int apy_eri_mo(int64_t nbf, int64_t nact, const void* eri_ao, const void* C, void* eri_mo, void* workspace,
               size_t ws_bytes, void* stream);
/* Spin-orbital chemists' integrals (compute_ERI_SO, utils.py:317-365):
 * eri_so[p,q,r,s] = eri_mo[p/2,q/2,r/2,s/2] delta(sp,sq) delta(sr,ss),
 * eri_mo [m]^4, eri_so [2m]^4 (not antisymmetrized). */
int apy_eri_so_from_mo(int64_t m, const void* eri_mo, void* eri_so, void* stream);
/* F = h + scale * sum_ls D_ls (2 (mn|ls) - (ml|ns))  (utils.py:241,249 with
 * scale 1; hf_wfn.py:81 with scale 1/2).  h may be NULL (zero). */
int apy_fock_build(int64_t nbf, const void* h, const void* eri_ao, const void* D, double scale, void* F,
                   void* stream);
/* The same for `batch` SCF instances iterated together (the displaced
 * geometries and fields of fin_diff.py:267-372, each running hf_wfn.py:81):
 * h, D, F are [batch][nbf][nbf]; eri_ao advances by stride_eri elements per
 * instance (0: one tensor shared).  Each instance's F is bit-identical to
 * apy_fock_build on it. */
int apy_fock_build_batch(int64_t nbf, int64_t batch, const void* h, const void* eri_ao, int64_t stride_eri,
                         const void* D, double scale, void* F, void* stream);
/* out_b = C_bra,b^H S_b C_ket,b for b < batch (compute_mo_overlap,
 * utils.py:381); C_* are [nbf][nmo], S [nbf][nbf], out [nmo][nmo]; strides
 * are element counts between batch entries (0 = shared). */
size_t apy_mo_overlap_workspace(int64_t nbf, int64_t nmo, int64_t batch);
int apy_mo_overlap(int64_t nbf, int64_t nmo, int64_t batch, const void* c_bra, int64_t stride_bra, const void* s_ao,
                   int64_t stride_s, const void* c_ket, int64_t stride_ket, void* out, void* workspace,
                   size_t ws_bytes, void* stream);
/* out[:, m] = C[:, m] / (S_mm / |S_mm|)  (compute_phase, utils.py:427-446);
 * S is the MO overlap C0^H S_AO C with leading dimension lds. */
int apy_phase_align(int64_t nbf, int64_t ncol, const void* s_mo, int64_t lds, const void* C, void* out,
                    void* stream);

/* ------------------------------------------------------------------------
 * MP2 (mp2_wfn.py:42-87).
 * Spin orbitals (solve_MP2_SO, :64-87):  t = <ab||ij>/D,  E = 1/4 sum <ij||ab> t
 * Spatial (solve_MP2, :42-59), from chemists' MO integrals (pq|rs) over the
 *   active orbitals [m][m][m][m], occupied first:
 *   t[i,j,a,b] = (ai|bj)/D,  E = sum (2 (ia|jb) - (ib|ja)) t
 * D_ijab = eps_i + eps_j - eps_a - eps_b (mp2_wfn.py:39, 66-68).
 * `energy` / `energies` are device complex scalars.
 * ---------------------------------------------------------------------- */
typedef struct apy_mp2_instance {
  int64_t O, V;
  const double* eps;    /* [O+V] */
  const void* vvoo_t;   /* [O][O][V][V] <ab||ij> at [i][j][a][b] */
  const void* oovv;     /* [O][O][V][V] <ij||ab> */
  void* t2;             /* [O][O][V][V] output */
} apy_mp2_instance;

/* n instances of any (O, V) in one launch per 40 instances: the batched
 * finite-difference form (each instance is one displaced wavefunction).
 * `inst` is a host array; energies[k] receives instance k's energy. */
size_t apy_mp2_so_batched_workspace(int n, const apy_mp2_instance* inst);
int apy_mp2_so_batched(int n, const apy_mp2_instance* inst, void* energies, void* workspace, size_t ws_bytes,
                       void* stream);
int apy_mp2_so(int64_t O, int64_t V, const double* eps, const void* vvoo_t, const void* oovv, void* t2,
               void* energy, void* workspace, size_t ws_bytes, void* stream);
size_t apy_mp2_spatial_workspace(void);
int apy_mp2_spatial(int64_t O, int64_t V, const double* eps, const void* eri_mo, void* t2, void* energy,
                    void* workspace, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Orbital-invariant MP2 and its adjoint (lambda) — the energy derivative of
 * analytic.py:307-338 without unrolling iterations.
 *   R(x) = s*drv + P(ab) x F_vv^T - P(ij) F_oo^T x   (ci_wfn.py:211-212)
 *   amplitudes: drv = <ab||ij> at [i][j][a][b], s = 1,   F
 *   adjoint:    drv = <ij||ab>,                s = 1/4, F^T (transpose_f)
 * One Jacobi step x_new = x + R(x)/D; rnorm (device complex scalar) gets
 * (sum |R|^2, 0).  x_new may equal x.  E = 1/4 sum <ij||ab> t; at diagonal
 * F this is solve_MP2_SO (mp2_wfn.py:64-87).
 * ---------------------------------------------------------------------- */
typedef struct apy_mp2_nc_args {
  int64_t O, V;
  const void* fock;     /* [n][n] */
  const double* eps;    /* [n] preconditioner D_ijab */
  int transpose_f;
  double drv_scale;
  const void* drv;      /* [O][O][V][V] */
  const void* x;
  void* x_new;
  void* rnorm;
} apy_mp2_nc_args;
size_t apy_mp2_nc_workspace(int64_t O, int64_t V);
int apy_mp2_nc_step(const apy_mp2_nc_args* a, void* workspace, size_t ws_bytes, void* stream);
/* grad_f [n][n] = dE/dF: vv block 2 sum lam t, oo block -2 sum t lam, rest 0. */
size_t apy_mp2_nc_density_workspace(int64_t O, int64_t V);
int apy_mp2_nc_densities(int64_t O, int64_t V, const void* t, const void* lam, void* grad_f, void* workspace,
                         size_t ws_bytes, void* stream);
/* Directional MP2 energy derivatives of several base systems in one call
 * (the gradient step the reference's finite-difference loop replaces,
 * fin_diff.py:285-370, in the Lagrangian form of ci_wfn.py:210-212 /
 * analytic.py:307-338): for item b with converged packed amplitudes t and
 * lambda (x[(i<j)][(a<b)], apy_t2_packed_length elements each), grad_f
 * receives dE/dF as apy_mp2_nc_densities gives it, and for each of its ndir
 * directions q
 *   out[off_b + q] = sum G dF_q + 1/4 sum t dW_q + sum lambda dV_q
 * (plain products; dF_q [n][n], dW_q and dV_q [O][O][V][V] dense, t and
 * lambda taken dense; off_b = ndir of the items before b).  dF, dW, dV are
 * host arrays of ndir device pointers (dW or dV null: that term is zero).
 * lam null: the system is Hermitian (F = F^H, <ab||ij> = conj(<ij||ab>)),
 * whose lambda equations are the conjugated amplitude equations scaled by
 * 1/4, so lambda = conj(t) / 4 is used without a separate solve (the
 * reference's MP2 gradient, analytic.py:320-324, likewise uses t alone).
 * Workspace apy_mp2_adjoint_sums_workspace(nitem, items) bytes. */
typedef struct apy_mp2_adjoint_item {
  int64_t O, V;
  const void* t;
  const void* lam;
  void* grad_f;
  int64_t ndir;
  const void* const* dF;
  const void* const* dW;
  const void* const* dV;
} apy_mp2_adjoint_item;
size_t apy_mp2_adjoint_sums_workspace(int64_t nitem, const apy_mp2_adjoint_item* items);
int apy_mp2_adjoint_sums(int64_t nitem, const apy_mp2_adjoint_item* items, void* out, void* workspace,
                         size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Spin-adapted (closed-shell) MP2 energy derivatives in the reference's own
 * forms (compute_MP2_Hessian, analytic.py:307-335).  t2 [o][o][v][v] are the
 * closed-shell amplitudes of apy_mp2_spatial (mp2_wfn.py:54), pair symmetric
 * (t_ijab = t_jiba); f-like matrices [n][n] row-major, n = o + v; a
 * direction q holds f_grad + q n^2 and eri_grad + q o^2 v^2 (<ij|ab> at
 * [i][j][a][b]).  Plain products throughout, no conjugation.
 *   apy_mp2_sp_response: the t2 response of nb directions, t2_grad =
 *     (eri_grad + Fock terms) / D with the Fock terms in the reference's two
 *     index orders: form 0 (analytic.py:326-331, analytic_aats.py:447-452)
 *     - f_ik t_kjab - t_ikab f_kj + t_ijcb f_ac + t_ijac f_cb; form 1
 *     (analytic_aats.py:346-350) - f_ki t_kjab - f_kj t_ikab + f_ac t_ijcb
 *     + f_bc t_ijac.  eri_grad may be null (zero).
 *   apy_mp2_sp_densities: the coefficients of E_grad (analytic.py:320-324):
 *     d_eri = 2 (2 t - t^T) [o][o][v][v] (may be null), d_oo [o][o] =
 *     -2 sum_jab t_ijab u_kjab, d_vv [v][v] = 2 sum_ijb t_ijab u_ijcb with
 *     u = 2 t - t^T (each f term of the reference counted twice by the pair
 *     symmetry).
 *   apy_mp2_sp_directional: out[q] = sum d_eri eri_grad[q] + sum d_oo
 *     f_grad[q]_oo + sum d_vv f_grad[q]_vv, the E_grad of nd directions
 *     (either of f_grad, eri_grad may be null; out: nd complex).
 * ---------------------------------------------------------------------- */
size_t apy_mp2_sp_response_workspace(int64_t o, int64_t v, int64_t nb);
int apy_mp2_sp_response(int64_t o, int64_t v, int64_t nb, int form, const void* t2, const void* f_grad,
                        const void* eri_grad, const double* eps, void* t2_grad, void* workspace, size_t ws_bytes,
                        void* stream);
size_t apy_mp2_sp_densities_workspace(int64_t o, int64_t v);
int apy_mp2_sp_densities(int64_t o, int64_t v, const void* t2, void* d_eri, void* d_oo, void* d_vv, void* workspace,
                         size_t ws_bytes, void* stream);
size_t apy_mp2_sp_directional_workspace(int64_t nd);
int apy_mp2_sp_directional(int64_t o, int64_t v, int64_t nd, const void* d_eri, const void* d_oo, const void* d_vv,
                           const void* f_grad, const void* eri_grad, void* out, void* workspace, size_t ws_bytes,
                           void* stream);

/* Packed CI vectors [c1 (O*V) | c2 over (i<j, a<b) (Po*Pv) | c0], length
 * apy_ci_packed_length(O, V): the antisymmetric doubles stored once, so the
 * CI metric of the dense form (c2 weighted 1/4) is the Euclidean product and
 * a Davidson subspace takes a quarter of the memory traffic.  apy_ci_pack
 * reads the (i<j, a<b) entries of a dense [c1 | c2 | c0] vector;
 * apy_ci_unpack writes all four antisymmetric images (zero diagonals);
 * apy_ci_precondition_packed is apy_ci_precondition on the packed layout.
 * O, V >= 2; dense and packed must differ. */
int64_t apy_ci_packed_length(int64_t O, int64_t V);
int apy_ci_pack(int64_t O, int64_t V, const void* dense, void* packed, void* stream);
int apy_ci_unpack(int64_t O, int64_t V, const void* packed, void* dense, void* stream);
int apy_ci_precondition_packed(int64_t O, int64_t V, const void* eps, double theta, const void* r, void* out,
                               void* stream);
/* The same with theta read from device memory (a double, e.g. the first
 * entry of an apy_heev_lowest block), so a device-controlled Davidson step
 * needs no host value. */
int apy_ci_precondition_packed_dev(int64_t O, int64_t V, const void* eps, const void* theta, const void* r, void* out,
                                   void* stream);
/* ------------------------------------------------------------------------
 * Davidson support and the CISD energy gradient with respect to F.
 * Vectors have length L + 1, L = O V + O^2 V^2: [c1 | c2 (dense,
 * antisymmetric) | c0] -- the apy_cisd_so_step layout plus the reference
 * weight.
 *   apy_ci_precondition: out = r / (theta - D) with D = e_a - e_i (singles),
 *     e_a + e_b - e_i - e_j (doubles, 0 where i == j or a == b), 0 (c0).
 *   apy_cisd_so_opd: D[p][q] = dE/dF[p][q] of the normalised vector c (the
 *     one-particle density of the adjoint sketch at ci_wfn.py:380-394;
 *     the lowest root of the Hermitian CISD matrix is variational, so this
 *     is the exact gradient): D is [(O+V)][(O+V)], fully written.
 * ---------------------------------------------------------------------- */
int apy_ci_precondition(int64_t O, int64_t V, const void* eps, double theta, const void* r, void* out,
                        void* stream);
/* CI-metric inner products of Davidson vectors (layout [c1 | c2 | c0],
 * length O V + O^2 V^2 + 1): out[m] = sum_l w_l conj(x_m[l]) y[l] with
 * w = 1/4 on the dense c2 block (each determinant counted once), 1 else;
 * k <= APY_MAXVEC, workspace apy_reduce_workspace(0, k); deterministic. */
int apy_ci_dots(int64_t O, int64_t V, int k, const void* const* xs, const void* y, void* out, void* workspace,
                size_t ws_bytes, void* stream);
size_t apy_cisd_so_opd_workspace(int64_t O, int64_t V);
int apy_cisd_so_opd(int64_t O, int64_t V, const void* c, void* D, void* workspace, size_t ws_bytes, void* stream);

/* Two-particle density of the normalised vector c: the gradient of the
 * lowest-root energy with respect to the antisymmetrised integrals
 * <pq||rs> (the TPD of the adjoint sketch at ci_wfn.py:396-409, there for
 * the Dirac integrals), written into `grad` in the block layouts of
 * apy_so_ints (any block pointer may be NULL to skip it).  Each stored
 * entry is the derivative with respect to that stored entry, i.e. the sum
 * over the full-tensor positions it stands for, so that
 *   dE = sum_blocks sum_entries grad_blk * d(ints_blk)
 * for every perturbation antisymmetric in (p,q) and (r,s); with the OPD,
 * E = sum F D + sum_blocks ints_blk . grad_blk.  vvvv_p, oooo_p, ring and
 * vovv_p are GEMMs on the matrix cores.  O, V >= 2. */
size_t apy_cisd_so_tpd_workspace(int64_t O, int64_t V);
int apy_cisd_so_tpd(int64_t O, int64_t V, const void* c, const apy_so_ints* grad, void* workspace, size_t ws_bytes,
                    void* stream);
/* Whole-tensor gradient G[n][n][n][n] (n = O + V) from block gradients (all
 * nine blocks required): the adjoint of apy_so_ints_from_full for
 * antisymmetric perturbations, G antisymmetric in (p,q) and (r,s), zero on
 * p == q or r == s, so that dE = sum G dA. */
int apy_so_ints_grad_to_full(int64_t n, const apy_so_ints* grad, void* full, void* stream);

/* ------------------------------------------------------------------------
 * The same non-canonical MP2 (or lambda) iteration batched over independent
 * instances of possibly different (O, V): the +/- displaced systems of a
 * finite-difference set (fin_diff.py:267-372 solves them one at a time).
 * The amplitudes are antisymmetric in (i,j) and (a,b) and stored packed,
 * x[(i<j)][(a<b)] (apy_t2_packed_length elements; apy_t2_pack / apy_t2_unpack
 * convert from / to the dense [O][O][V][V] tensor).  One fused kernel per
 * step covers every instance (a workgroup per instance, pair i<j and 32 x 32
 * tile pair of the virtual plane).  setup uploads the instance and work
 * tables into the workspace (host arrays; synchronises the stream) and
 * returns the item count; step and energy touch only device memory (they
 * write scratch regions of the workspace), so a sequence of them can be
 * captured in a HIP graph.  reverse = 1 iterates x_new -> x (ping-pong
 * without re-setup).  rnorm[k] = sum over (i<j, a<b) |r_k|^2 (double);
 * energy[k] = scale * sum_dense w_k x_k (complex, w antisymmetric; w may be
 * NULL).  Instances with O < 2 or V < 2 have no pairs and keep x.
 * ---------------------------------------------------------------------- */
int64_t apy_t2_packed_length(int64_t O, int64_t V);
int apy_t2_pack(int64_t O, int64_t V, const void* dense, void* packed, void* stream);
int apy_t2_unpack(int64_t O, int64_t V, const void* packed, void* dense, void* stream);
typedef struct {
  int64_t O, V;
  const void* fock;  /* [(O+V)][(O+V)] complex */
  const void* eps;   /* [O+V] double: Jacobi denominators */
  const void* drv;   /* [O][O][V][V] driver (<ab||ij> at [i,j,a,b], or <ij||ab>), antisymmetric */
  void* x;           /* [Po][Pv] packed amplitudes */
  void* x_new;       /* [Po][Pv] (distinct from x) */
  const void* w;     /* [O][O][V][V] <ij||ab> for the energy, or NULL */
  int transpose_f;   /* 1: the lambda equations (F^T) */
  double drv_scale;  /* s: the driver enters as s * drv */
} apy_mp2_nc_instance;
size_t apy_mp2_nc_batched_workspace(int n, const apy_mp2_nc_instance* inst);
int apy_mp2_nc_batched_setup(int n, const apy_mp2_nc_instance* inst, void* workspace, size_t ws_bytes,
                             int64_t* n_items, void* stream);
int apy_mp2_nc_batched_step(int n, int64_t n_items, int reverse, void* rnorm, void* workspace, void* stream);
/* The whole batched solve in ONE persistent cooperative launch (the
 * per-displacement loop of fin_diff.py:285-370 as one kernel): the resident
 * workgroups iterate apy_mp2_nc_batched_step's three phases with a grid
 * barrier after each, test sqrt(max_k rnorm[k]) < tol on the device after
 * every step, and stop there or after maxit steps, leaving the amplitudes in
 * x (copied back from x_new after an odd step count).  iterations (device
 * int64) receives the step count; the iterates are bit-identical to stepping
 * from x with apy_mp2_nc_batched_step and testing after each step.  The grid
 * is the number of workgroups that fit on the device at once. */
int apy_mp2_nc_batched_solve(int n, int64_t n_items, double tol, int maxit, void* rnorm, void* iterations,
                             void* workspace, void* stream);
int apy_mp2_nc_batched_energy(int n, int64_t n_items, int use_x_new, double scale, void* energy,
                              void* workspace, void* stream);

/* ------------------------------------------------------------------------
 * AO electron-repulsion integrals (replace psi4's MintsHelper.ao_eri of
 * hamiltonian.py:33 and the two-electron part of mo_tei_deriv1,
 * analytic_aats.py:94, before its MO transform).
 * A function is one or two terms sharing a centre (atom) and exponents:
 *   sum_t sum_p coefs[t][p] x^lmn[t][0] y^lmn[t][1] z^lmn[t][2] e^{-exps[p] r^2}
 * (coefficients include the primitive normalisation).  Limits: nprim <=
 * APY_GTO_MAXPRIM, lmn[t] summing to <= APY_GTO_MAXL, quartet total <= 9 (s/p/d basis
 * functions and their first derivatives); an entry outside them yields NaN.
 * ---------------------------------------------------------------------- */
#define APY_GTO_MAXPRIM 6
#define APY_GTO_MAXL 3
typedef struct apy_gto {
  int32_t atom;                               /* index into the centres       */
  int32_t nprim;
  int32_t nterm;                              /* 1 or 2                       */
  int32_t lmn[2][3];
  int32_t pad;                                /* exps at byte offset 40       */
  double exps[APY_GTO_MAXPRIM];
  double coefs[2][APY_GTO_MAXPRIM];
} apy_gto;                                    /* 184 bytes                    */

/* out[g][i][j][k][l] = (fa_i fb_j | fc_k fd_l) at geometry g (chemists'
 * order, complex128 with zero imaginary part), centres [ngeom][natom][3]
 * bohr; tables and centres in device memory. */
int apy_eri_gto(int64_t ngeom, int64_t natom, const double* centers, const apy_gto* fa, int64_t na,
                const apy_gto* fb, int64_t nb, const apy_gto* fc, int64_t nc, const apy_gto* fd, int64_t nd,
                void* out, void* stream);
/* d (mn|ls) / dR_{atom,x} from D[m,n,l,s] = (dm/dx n|l s) (apy_eri_gto with
 * the x-derivative functions in fa): out = on(m) D[m,n,l,s] + on(n) D[n,m,l,s]
 * + on(l) D[l,s,m,n] + on(s) D[s,l,m,n], on(p) = (atom_of[p] == atom);
 * complex128 [n]^4, atom_of int32 [n] in device memory. */
int apy_eri_deriv_assemble(int64_t n, const int* atom_of, int atom, const void* D, void* out, void* stream);

/* One-electron integrals over two apy_gto tables (replace psi4's
 * MintsHelper ao_overlap, ao_kinetic, ao_potential, ao_dipole and
 * ao_angular_momentum of hamiltonian.py:29-35, 49, 55 and the mixed-geometry
 * ao_overlap(bra_basis, ket_basis) of utils.py:371-376), for nset sets at
 * once; every array in device memory:
 *   out[s][c][i][j], c over the selected components in the order
 *   S, T, V, Dx, Dy, Dz, Lx, Ly, Lz:
 *   S  = <a_i|b_j>,  T = <a_i| -1/2 nabla^2 |b_j>,
 *   V  = <a_i| -sum_C Z_C / |r - C| |b_j>   (nuclei[s][C], charges[C]),
 *   D  = <a_i| r - origin |b_j>,
 *   L  = <a_i| (r - origin) x nabla |b_j>   (real, antisymmetric for a = b).
 * Bra function i sits at ca[s][fa[i].atom], ket function j at
 * cb[s][fb[j].atom] (bohr), so bra and ket may belong to different
 * geometries.  Limits as for apy_eri_gto (lmn summing to <= 3); an entry
 * outside them yields NaN. */
#define APY_OEI_OVERLAP 1
#define APY_OEI_KINETIC 2
#define APY_OEI_POTENTIAL 4
#define APY_OEI_DIPOLE 8
#define APY_OEI_ANGMOM 16
typedef struct apy_oei_args {
  int64_t nset;
  const apy_gto* fa;                          /* bra table [na]               */
  int64_t na;
  const double* ca;                           /* bra centres [nset][nca][3]   */
  int64_t nca;
  const apy_gto* fb;                          /* ket table [nb]               */
  int64_t nb;
  const double* cb;                           /* ket centres [nset][ncb][3]   */
  int64_t ncb;
  const double* nuclei;                       /* [nset][nnuc][3] (V only)     */
  const double* charges;                      /* [nnuc]          (V only)     */
  int64_t nnuc;
  double origin[3];                           /* for D and L                  */
  int32_t which;                              /* OR of APY_OEI_* bits         */
  int32_t pad;
  double* out;                                /* [nset][ncomp][na][nb]        */
} apy_oei_args;                               /* 136 bytes                    */
/* Number of output components that `which` selects (0 if invalid). */
int apy_oei_components(int which);
int apy_oei_gto(const apy_oei_args* args, void* stream);

/* ------------------------------------------------------------------------
 * Vector reductions used by DIIS and the convergence test.
 * ---------------------------------------------------------------------- */
#define APY_MAXVEC 16
/* out[m] = sum_l conj?(x_m[l]) * y[l] for m < k   (utils.py:121) */
int apy_zdot_many(int k, const void* const* xs, const void* y, int64_t len, int conj_x, void* out,
                  void* workspace, size_t ws_bytes, void* stream);
/* out[seg] = sum_l (x[l]-y[l])^2 over each of nseg segments [off[s], off[s+1])
 * (un-conjugated, ci_wfn.py:361-365).  off is a host array of nseg+1. */
int apy_zdiffsq_segments(const void* x, const void* y, int nseg, const int64_t* off, void* out,
                         void* workspace, size_t ws_bytes, void* stream);
/* y = sum_m c_m x_m for m < k, c host complex (re,im pairs) (utils.py:138) */
int apy_zcombine(int k, const void* const* xs, const double* coef, int64_t len, void* y, void* stream);
size_t apy_reduce_workspace(int64_t len, int k);

/* The same for nseg segments of seglen elements stored one after another
 * (one DIIS per segment, utils.py:104-140, for sets of problems iterated
 * together); each segment's sums are bit-identical to the single-vector
 * calls on it.
 *   apy_zdot_many_seg:  out[s][m] = sum_l conj?(x_m[s][l]) y[s][l]
 *   apy_zdiffsq_batch:  out[s][0], out[s][1] = sum (x-y)^2 over [0, split)
 *                       and [split, seglen) of segment s
 *   apy_zcombine_seg:   y[s] = sum_m c[s][m] x_m[s], c [nseg][k] complex in
 *                       device memory
 * Workspace apy_reduce_seg_workspace(nseg, k) (nseg -> 2 nseg for diffsq). */
size_t apy_reduce_seg_workspace(int64_t nseg, int k);
/* Gram-Schmidt coefficients on the device from dots = [<B_j|d> (j < k) |
 * <d|d>]: normalize = 0 gives coef = [-<B_j|d> | 1]; normalize = 1 gives
 * coef = [-<B_j|d> | 1] / sqrt(nn), nn = <d|d> - sum_j |<B_j|d>|^2 (the norm
 * of the projected d), and status = (nn, <d|d>) (coef all 0 if nn <= 0).
 * One thread; k < APY_MAXVEC.  Lets a Davidson expansion run without host
 * reads (apply with apy_zcombine_seg, nseg = 1). */
int apy_gs_coefficients(int k, const void* dots, int normalize, void* coef, void* status, void* stream);
/* The projected eigenproblem of the Davidson iteration on the device (the
 * host numpy.linalg.eigh of the lowest root in davidson.py; the reference
 * has no Davidson solver, its CISD is the Jacobi iteration of
 * ci_wfn.py:302-378).  g holds the k x k Hermitian projected matrix by its
 * upper triangle, column-major: g[j * ld + i] = <B_i|S_j> for i <= j
 * (k <= APY_MAXVEC).  eig (device, 2 APY_MAXVEC + 2 complex) receives
 * [theta | y (APY_MAXVEC, zero past k) | -theta y (k) | 1 | 0 ...]: the
 * lowest eigenvalue (real part of entry 0), its unit eigenvector, and the
 * coefficients that form the residual sum_j y_j S_j - theta sum_j y_j B_j
 * with two apy_zcombine_seg calls (the second appending the partial sum
 * with weight 1).  Cyclic parallel Jacobi rotations in one wavefront. */
int apy_heev_lowest(int k, const void* g, int ld, void* eig, void* stream);
/* Convergence test of a device-controlled Davidson iteration: ctl (device,
 * double[4]) = [iterations, previous theta, status (0 running, 1
 * converged), residual norm]; with theta = eig[0] and the residual norm
 * sqrt(rr[0]) the status becomes 1 when |theta - previous| < e_conv and the
 * norm < r_conv (not on the first iteration); log[2 (it - 1) + {0, 1}] =
 * (theta, norm) for it <= cap. */
int apy_davidson_check(const void* eig, const void* rr, double e_conv, double r_conv, void* ctl, void* log, int cap,
                       void* stream);
/* Fused passes of the device-controlled Davidson step over packed vectors
 * (apy_ci_packed_length(O, V)), k <= APY_MAXVEC - 1 subspace vectors, each
 * one read of the operands instead of three or four calls:
 *   apy_davidson_residual: d = r / (theta - D) with r = sum_j y_j S_j - theta
 *     sum_j y_j B_j for the Ritz pair in eig (apy_heev_lowest) and D the
 *     packed diagonal of apy_ci_precondition_packed; dots (device, k + 2
 *     complex) = [<B_j|d> (j < k) | <d|d> | <r|r>] (the input of
 *     apy_gs_coefficients, then the residual norm for apy_davidson_check).
 *   apy_davidson_combine_dots: d <- sum_j coef_j B_j + coef_k d (coef in
 *     device memory, e.g. apy_gs_coefficients) and dots = [<B_j|d> (j < k) |
 *     <d|d>] of the result (the second Gram-Schmidt pass's input).
 * d must differ from every S_j and B_j.  Workspace
 * apy_davidson_fused_workspace() bytes; deterministic fixed-order sums. */
size_t apy_davidson_fused_workspace(void);
int apy_davidson_residual(int k, const void* const* S, const void* const* B, const void* eig, const void* eps,
                          int64_t O, int64_t V, void* d, void* dots, void* workspace, size_t ws_bytes, void* stream);
int apy_davidson_combine_dots(int k, const void* const* B, const void* coef, int64_t len, void* d, void* dots,
                              void* workspace, size_t ws_bytes, void* stream);
int apy_zdot_many_seg(int k, const void* const* xs, const void* y, int64_t nseg, int64_t seglen, int conj_x,
                      void* out, void* workspace, size_t ws_bytes, void* stream);
int apy_zdiffsq_batch(const void* x, const void* y, int64_t nseg, int64_t seglen, int64_t split, void* out,
                      void* workspace, size_t ws_bytes, void* stream);
int apy_zcombine_seg(int k, const void* const* xs, const void* coef, int64_t nseg, int64_t seglen, void* y,
                     void* stream);

/* The Jacobi + DIIS iteration of ci_wfn.py:302-378 controlled from the
 * device, so that several iterations can be queued (or replayed from a
 * captured HIP graph) between host reads.  The control block ctl (int32[2],
 * device) holds the iterations performed and the status: 0 running,
 * 1 converged, 2 max_iterations reached, 3 singular DIIS system; once the
 * status is set, apy_diis_slot, apy_zcombine_guarded and apy_iter_check do
 * nothing, so the amplitudes of the last iteration stay where it put them.
 *
 * apy_diis_slot: DIIS of utils.py:104-140 over a ring of nslot stored
 * (amplitude, residual) pairs.  row[q] = <r_q | r_slot> (conjugated dots with
 * the newest residual, apy_zdot_many over every slot); valid = bit mask of
 * the window's slots (slot included).  B [nslot][nslot] (device) keeps the
 * dots of earlier iterations; coef[q] (device, zero outside the window)
 * solves the bordered system [[B_w, -1], [-1, 0]] c = [0, .., 0, -1].
 * apy_zcombine_guarded: y = sum_q coef[q] x_q (coef in device memory).
 * apy_iter_check: iteration it = ctl[0] + 1 done: E = energy[0],
 * dE = E_prev - E, rms_k = sqrt(sums[k]) (principal root; sums are the
 * un-conjugated sums of squares of the change of t1 and t2); log[4 it + j]
 * = (E, dE, rms1, rms2) while it < cap; converged (ci_wfn.py:371-374) when
 * it > 1, |dE| < e_conv and rms2 (and rms1 if check_rms1) < d_conv in
 * NumPy's complex order; status 2 when it reaches max_it.  E_prev <- E. */
int apy_diis_slot(int nslot, int slot, uint32_t valid, const void* row, void* B, void* coef, int* ctl, void* stream);
int apy_zcombine_guarded(int k, const void* const* xs, const void* coef, int64_t len, void* y, const int* ctl,
                         void* stream);
int apy_iter_check(const void* energy, const void* sums, int check_rms1, double e_conv, double d_conv, int max_it,
                   void* E_prev, void* log, int cap, int* ctl, void* stream);

/* ------------------------------------------------------------------------
 * Finite-difference atomic axial tensors (replaces AAT.compute_all_dets,
 * aats.py:558-642, and the determinant-amplitude sums of
 * compute_spatial_aats, aats.py:646-1055, and compute_SO_I_*,
 * aats.py:120-516).
 *
 * Overlaps S (n x n, rows = bra orbitals, columns = ket orbitals, the first
 * no of each occupied, the first nf of those frozen) are reduced to a record
 *   [d = det S[o,o]] [x: oa x nv] [y: oa x nv] [P: oa x oa] [W: nv x nv]
 * (oa = no - nf) with P = S[o,o]^-1 (active block), x[i][a] = (S[v,o] P)[a][i],
 * y = (P S[o,v])[active rows], W = S[v,v] - S[v,o] P S[o,v]: every
 * determinant with up to two occupied rows and two occupied columns
 * replaced is d times a minor built from these.  d = 0 (exactly singular
 * occupied block) zeroes the whole record.
 * ---------------------------------------------------------------------- */
#define APY_AAT_NTERMS 20
/* double2 elements per record (a multiple of 16) */
int64_t apy_aat_gamma_record(int64_t no, int64_t nf, int64_t nv);
size_t apy_aat_gamma_workspace(int64_t n, int64_t no, int64_t batch);
/* gamma[b * record] for S[b * stride_s], b < batch; no <= 256 */
int apy_aat_gamma(int64_t n, int64_t no, int64_t nf, int64_t batch, const void* s, int64_t stride_s, void* gamma,
                  void* workspace, size_t ws_bytes, void* stream);

/* out = conj?(ca * a + cb * b) over n complex elements; b may be NULL.
 * Forms the normalised finite-difference amplitudes of aats.py:680-704. */
typedef struct {
  int64_t n;
  const void* a;
  const void* b;
  void* out;
  double ca, cb;
  int conj;
} apy_aat_axpby_op;
size_t apy_aat_axpby_workspace(int64_t count);
int apy_aat_axpby(int64_t count, const apy_aat_axpby_op* ops, void* workspace, size_t ws_bytes, void* stream);

/* One overlap class: bra amplitudes (b1 [oa,nv], b2 [oa,oa,nv,nv], already
 * conjugated) and ket amplitudes (k1, k2) against record `gamma`; NULL
 * amplitudes are zero.  out[job * APY_AAT_NTERMS + q], with
 * A(t) = (t - t^ab - t^ij + t^ijab)/4, beta = sum_jb b2[ijab] x[jb],
 * beta' = sum_ia b2[ijab] x[ia], beta^ = sum_jb A(b2)[ijab] x[jb],
 * kA = sum_ld k2[klcd] y[ld], kB = sum_kc k2[klcd] y[kc],
 * k^ = sum_lc A(k2)[klcd] y[lc], H(M) = P^T M W^T, T = (P^T P^T W W) k2:
 *   q: 0 d  1 b1.x  2 k1.y  3 beta.x  4 beta^.x  5 kA.y  6 -k^.y
 *      7 b1.H(k1)  8 beta.H(k1)  9 beta^.H(k1)  10 b1.H(k^)  11 b1.H(kA)
 *      12 beta^.H(k^)  13 beta^.H(kA)  14 beta.H(k^)  15 beta'.H(k^)
 *      16 beta'.H(kB)  17 beta.H(kA)  18 b2.A(T)  19 b2.T
 * (bilinear sums, no conjugation).  The job table is a host array; each
 * job's gamma index must be below ngamma, the number of records. */
typedef struct {
  int64_t gamma;
  const void* b1;
  const void* b2;
  const void* k1;
  const void* k2;
} apy_aat_job;
size_t apy_aat_terms_workspace(int64_t oa, int64_t nv, int64_t njobs);
int apy_aat_terms(int64_t oa, int64_t nv, const void* gamma, int64_t record, int64_t ngamma, int64_t njobs,
                  const apy_aat_job* jobs, void* out, void* workspace, size_t ws_bytes, void* stream);

/* out[3 s + (0,1,2)] = sum |t1|^2, sum |t2|^2, Re sum conj(t2[ijab]) t2[ijba]
 * for each amplitude set s (t1[s] may be NULL), normalisations of
 * aats.py:134-157, 652-669.  t1, t2 are host arrays of device pointers. */
size_t apy_aat_norms_workspace(int64_t nset);
int apy_aat_norms(int64_t oa, int64_t nv, int64_t nset, const void* const* t1, const void* const* t2, void* out,
                  void* workspace, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Roofline denominators measured on the device (no counterpart in the
 * reference; they give the measured side of the FP64 matrix-core and HBM
 * peaks the throughput figures are quoted against).
 *   apy_peak_mfma_f64: blocks x 256 threads, each wave iterating 4
 *     independent v_mfma_f64_16x16x4f64 accumulator chains `iters` times from
 *     registers; apy_peak_mfma_f64_flops gives its flop count (host); out
 *     receives blocks x 256 doubles (keeps the chains live).
 *   apy_peak_copy: dst[e] = src[e] for n complex128 words (16 n bytes read
 *     and 16 n written), grid-stride over `blocks` x 256 threads.
 *   apy_peak_read: the n words of src read (16 n bytes), each thread's sum
 *     written to out (blocks x 256 doubles).
 * The caller times a launch with events on the stream. */
double apy_peak_mfma_f64_flops(int blocks, int iters);
int apy_peak_mfma_f64(int blocks, int iters, void* out, void* stream);
int apy_peak_copy(int64_t n, const void* src, void* dst, int blocks, void* stream);
int apy_peak_read(int64_t n, const void* src, void* out, int blocks, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* APYIB_AMD_H */
