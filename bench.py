#!/usr/bin/env python
"""Benchmark of the spin-orbital CISD sigma on one MI355X.

Workload (one GPU): synthetic complex128 CISD at O=20 occupied and V=160
virtual spin orbitals (seeded <pq||rs>, complex std 0.05 Eh; Fock matrix
with U[-1.5,-0.3] / U[0.2,3.0] Eh diagonal and 1e-3 Hermitian noise), the
lowest root by Davidson iteration on the Hermitian CISD matrix.  A step is
one Davidson iteration: the sigma of the new vector (and the restart sigma
every max_space iterations), the CI-metric projections, Ritz pair,
residual, preconditioner and Gram-Schmidt.  Inputs are generated on the
device before timing; nothing crosses PCIe in the timed region except the
small projected matrices and scalars the iteration reads back.

Prints one JSON line.  `value` is the executed FP64 work of the step's
sigma contractions (8 flops per complex multiply-add, antisymmetric pairs
packed) divided by the measured time per step.  `aat_set` holds the second
metric: finite-difference instance sets solved batched and by the adjoint
against the plain per-instance loop on the same kernels.
"""
import argparse
import ctypes
import json
import os
import sys
import time

METRIC = "CISD sigma TFLOP/s (cplx FP64); AAT set wall-time batched+adjoint vs plain loop"
PEAK_FP64 = 78.6  # MI355X dense FP64 matrix, TFLOP/s (published)
ROOT = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """Raw hipEvent_t handles (timing on the stream the kernels use)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        if self.hip.hipEventCreate(ctypes.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e.value

    def elapsed_ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        ms = ctypes.c_float()
        if self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value

    def destroy(self, e):
        self.hip.hipEventDestroy(e)

    def record(self, e, stream):
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        if self.hip.hipEventRecord(e, stream) != 0:
            raise RuntimeError("hipEventRecord failed")


def measured_peaks(dev, hip, reps=4):
    """The two roofline denominators measured on this device with the
    library's microbenchmarks: FP64 matrix-core rate (apy_peak_mfma_f64,
    MFMA chains from registers) and HBM copy bandwidth (apy_peak_copy, 1 GiB
    read + 1 GiB written).  Best of reps launches each, timed with HIP events
    on the launch stream."""
    import torch
    from apyib_amd import _lib
    lib = _lib.load()
    stream = torch.cuda.current_stream(dev).cuda_stream
    ncu = torch.cuda.get_device_properties(dev).multi_processor_count
    blocks, iters = 8 * ncu, 20000
    out = torch.empty(blocks * 256, dtype=torch.float64, device=dev)
    flops = float(lib.apy_peak_mfma_f64_flops(blocks, iters))
    n = 1 << 26                                      # complex128 words: 1 GiB per buffer
    src = torch.ones(n, dtype=torch.complex128, device=dev)
    dst = torch.empty_like(src)

    def best(launch, amount):
        rate = 0.0
        for _ in range(reps + 1):                    # the first launch warms up
            e0, e1 = hip.create(), hip.create()
            hip.record(e0, stream)
            _lib.check(launch(), "peak microbenchmark")
            hip.record(e1, stream)
            torch.cuda.synchronize(dev)
            rate = max(rate, amount / (hip.elapsed_ms(e0, e1) * 1e-3))
            hip.destroy(e0)
            hip.destroy(e1)
        return rate
    mf = best(lambda: lib.apy_peak_mfma_f64(blocks, iters, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)),
              flops) / 1e12
    bw = best(lambda: lib.apy_peak_copy(n, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                        8 * ncu, ctypes.c_void_p(stream)), 32.0 * n) / 1e12
    rd_out = torch.empty(8 * ncu * 256, dtype=torch.float64, device=dev)
    rd = best(lambda: lib.apy_peak_read(n, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(rd_out.data_ptr()),
                                        8 * ncu, ctypes.c_void_p(stream)), 16.0 * n) / 1e12
    del out, src, dst, rd_out
    res = {"what": "build-owned microbenchmarks on this device: v_mfma_f64_16x16x4f64 chains from registers "
                   f"({blocks} workgroups x {iters} iterations), a 1 GiB complex128 copy (bytes read + written) "
                   "and a 1 GiB read stream",
           "fp64_mfma_tflops": round(mf, 2), "hbm_copy_tbps": round(bw, 3), "hbm_read_tbps": round(rd, 3),
           "published": {"fp64_mfma_tflops": PEAK_FP64, "hbm_tbps": 8.0}}
    log(f"measured peaks: {res}")
    return res


def read_traffic():
    """HBM bytes per PPL launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "ppl_o20_v160_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(solver, O, V, flops_step, seed, reps=3):
    """The reference algorithm's sigma (oracle NumPy restatement of
    ci_wfn.py:309-334, dense einsums as written, BLAS through
    einsum(optimize=True)) on the host cores, on this workload's integrals
    copied back from the device: one warm-up call, then `reps` timed calls,
    median reported.  The CPU sigma is compared with the GPU's on the same
    amplitudes (the reference's form, no f_vo coupling)."""
    import numpy as np
    import torch
    from apyib_amd.so_ci import SOCI
    from oracle import so
    from threadpoolctl import threadpool_info

    t0 = time.perf_counter()
    host = {k: solver.ints[k].cpu().numpy() for k in solver.ints.blocks}
    W = so.blocks_from_packed(host, O, V)
    del host
    F = solver.fock.cpu().numpy()
    rng = np.random.default_rng(seed)
    t1 = 0.01 * (rng.standard_normal((O, V)) + 1j * rng.standard_normal((O, V)))
    t2 = 0.01 * (rng.standard_normal((O, O, V, V)) + 1j * rng.standard_normal((O, O, V, V)))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    log(f"cpu baseline: host blocks ready in {time.perf_counter() - t0:.1f} s")
    so.cisd_so_sigma_blocks(F, W, t1, t2, O)                       # warm-up (BLAS threads, page faults)
    secs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        s1, s2 = so.cisd_so_sigma_blocks(F, W, t1, t2, O)
        secs.append(time.perf_counter() - t0)
    sec = float(np.median(secs))
    blas = [i for i in threadpool_info() if i.get("user_api") == "blas"]
    threads = max([i.get("num_threads", 1) for i in blas] or [1])
    vendor = ", ".join(f"{i.get('internal_api')} {i.get('version')} ({i.get('architecture')})" for i in blas) or "?"
    # the GPU sigma of the same amplitudes, reference form
    ref = SOCI(solver.ints, solver.fock, solver.eps.cpu().numpy(), singles=True, fov_coupling=False)
    x = torch.from_numpy(np.concatenate([t1.ravel(), t2.ravel()])).to(solver.device)
    out = ref.new_vector()
    ref.sigma(x, out)
    g = out.cpu().numpy()
    scale = max(np.abs(s1).max(), np.abs(s2).max())
    diff = max(np.abs(g[:O * V] - s1.ravel()).max(), np.abs(g[O * V:] - s2.ravel()).max()) / scale
    log(f"cpu baseline: sigma {[round(x, 2) for x in secs]} s on {threads} BLAS threads ({vendor}); "
        f"GPU vs CPU sigma max rel diff {diff:.2e}")
    return {
        "value": flops_step / sec / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"CISD sigma (NumPy restatement of ci_wfn.py:309-334, dense einsums as written) at O={O},V={V} "
                   f"on this workload's integrals: median of {reps} after one warm-up, "
                   f"{sec:.2f} s per sigma ({', '.join(f'{x:.2f}' for x in secs)}); rate = the GPU step's executed "
                   f"flops / CPU time"),
        "cpu_model": _cpu_model(),
        "blas": vendor,
        "seconds_per_sigma": secs,
        "max_rel_diff_gpu_vs_cpu_sigma": float(diff),
    }


def aat_set_measurement(dev, seed, tol=1e-12):
    """Second metric: a finite-difference tensor set solved batched (one
    fused launch per iteration for all instances) and by the adjoint,
    against the plain per-displacement loop (the same kernels, one instance
    at a time from a cold start, host convergence check per iteration: the
    shape of fin_diff.py:285-370).

    Set A: 24 instances at O=10, V=38 (3 atoms: 9 nuclear + 3 field
    directions, +/- 1e-4).  Every leg is the median of 3 runs after a
    warm-up of each path.  Set B: 36 instances at O~16, V~60 with seeded
    uneven sizes (18 directions, 14 distinct systems).  MP2 (orbital-
    invariant spin-orbital).  Batched legs: cold starts with direct
    launches, the whole solve as one persistent cooperative launch, the
    same replayed from a captured graph (capture included),
    and warm starts (the undisplaced systems solved first, one batch, then
    every displaced instance from its system's amplitudes).  Adjoint: the
    amplitude and lambda equations of every undisplaced system in one batch,
    then the densities and all directional sums (apy_mp2_adjoint_sums);
    the Hermitian leg solves the amplitudes only (lambda = conj(t)/4 for
    these Hermitian systems)."""
    import numpy as np
    import torch
    from apyib_amd.batched import MP2NCBatch, fd_mp2_instances, mp2_adjoint_set, plain_solve, warm_batch_solve

    def timed(fn, reps=3):
        """(result of the last run, median wall time of reps runs)"""
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return out, float(np.median(ts))

    res = {}
    rng = np.random.default_rng(seed + 17)
    sets = {
        "A": dict(O=10, V=38, natom=3, sizes=None),
        "B": dict(O=16, V=60, natom=5, sizes=[(int(rng.integers(14, 19)), int(rng.integers(54, 67)))
                                              for _ in range(18)]),
    }
    h = 1e-4
    for name, cfg in sets.items():
        insts, dirs, bases, base_of = fd_mp2_instances(cfg["O"], cfg["V"], seed, cfg["natom"], h=h, device=dev,
                                                       sizes=cfg["sizes"], with_bases=True)
        base_of_inst = [base_of[k // 2] for k in range(len(insts))]
        plain_solve(insts[:2], tol)                                   # warm every path once
        MP2NCBatch(insts[:2]).solve(tol)
        (Ep, its), t_plain = timed(lambda: plain_solve(insts, tol))

        def batched(graph):
            b = MP2NCBatch(insts)
            b.solve(tol, maxit=400, graph=graph)
            return b.energies().clone(), b.iterations
        (Eb, itb), t_batch = timed(lambda: batched(False))
        (Eg, itg), t_graph = timed(lambda: batched(True))

        def persistent():
            b = MP2NCBatch(insts)
            b.solve(tol, maxit=400, persistent=True)
            return b.energies().clone(), b.iterations
        (Epst, itpst), t_pers = timed(persistent)
        # the same batch solved again from zero amplitudes, replaying the graph captured above
        rb = MP2NCBatch(insts)
        rb.solve(tol, maxit=400, graph=True)

        def replay():
            rb.reset()
            rb.solve(tol, maxit=400, graph=True)
            return rb.energies().clone(), rb.iterations
        (Er, itr), t_replay = timed(replay)
        del rb
        (Ew, itwb, itw), t_warm = timed(lambda: warm_batch_solve(insts, bases, base_of_inst, tol))
        nd = len(dirs)
        dF = [((insts[2 * k].F - insts[2 * k + 1].F) / (2 * h)).reshape(-1) for k in range(nd)]
        dW = [(insts[2 * k].w - insts[2 * k + 1].w) / (2 * h) for k in range(nd)]
        dV = [(insts[2 * k].drv - insts[2 * k + 1].drv) / (2 * h) for k in range(nd)]
        groups = [[d for d in range(nd) if base_of[d] == b] for b in range(len(bases))]
        directions = [([dF[d] for d in g], [dW[d] for d in g], [dV[d] for d in g]) for g in groups]
        mp2_adjoint_set(bases[:1], directions[:1], tol)               # warm
        (an, dens), t_adj = timed(lambda: mp2_adjoint_set(bases, directions, tol))
        # the base systems are Hermitian (F = F^H, <ab||ij> = conj(<ij||ab>)): lambda = conj(t)/4 without its solve
        mp2_adjoint_set(bases[:1], directions[:1], tol, hermitian=True)
        (anh, densh), t_adj_h = timed(lambda: mp2_adjoint_set(bases, directions, tol, hermitian=True))
        fd = ((Ep[0::2] - Ep[1::2]) / (2 * h)).cpu().numpy()
        errs, errs_h = [], []
        for b, g in enumerate(groups):
            G, t, lam = dens[b]
            for q, d in enumerate(g):
                # error relative to the natural scale of each directional sum, sum |dE/dx| |dx|
                scale = float((torch.sum((G * dF[d]).abs()) + 0.25 * torch.sum((t * dW[d]).abs())
                               + torch.sum((lam * dV[d]).abs())).cpu())
                errs.append(abs(an[b][q] - fd[d]) / scale)
                errs_h.append(abs(anh[b][q] - fd[d]) / scale)
        diff_h = max(float(np.abs(anh[b] - an[b]).max()) for b in range(len(groups)))
        dE = lambda E: float(torch.abs(E - Ep).max().cpu())
        entry = {"instances": len(insts), "directions": nd, "systems": len(bases),
                 "sizes": sorted({(I.O, I.V) for I in insts}),
                 "plain_loop_s": round(t_plain, 5), "batched_s": round(t_batch, 5),
                 "batched_graph_s": round(t_graph, 5), "batched_warm_s": round(t_warm, 5),
                 "batched_persistent_s": round(t_pers, 5), "iterations_persistent": int(itpst),
                 "speedup_batched_persistent": round(t_plain / t_pers, 2),
                 "max_abs_dE_persistent_vs_plain": dE(Epst),
                 "batched_graph_replay_s": round(t_replay, 5), "iterations_graph_replay": int(itr),
                 "speedup_batched_graph_replay": round(t_plain / t_replay, 2),
                 "max_abs_dE_graph_replay_vs_plain": dE(Er),
                 "speedup_batched": round(t_plain / t_batch, 2), "speedup_batched_graph": round(t_plain / t_graph, 2),
                 "speedup_batched_warm": round(t_plain / t_warm, 2),
                 "iterations_plain_max": int(max(its)), "iterations_batched": int(itb),
                 "iterations_warm": [int(itwb), int(itw)],
                 "max_abs_dE_batched_vs_plain": dE(Eb), "max_abs_dE_graph_vs_plain": dE(Eg),
                 "max_abs_dE_warm_vs_plain": dE(Ew),
                 "adjoint_s": round(t_adj, 5), "speedup_adjoint_vs_plain_fd": round(t_plain / t_adj, 2),
                 "max_err_adjoint_vs_fd_rel_to_scale": float(max(errs)), "fd_step": h,
                 "adjoint_hermitian_s": round(t_adj_h, 5),
                 "speedup_adjoint_hermitian_vs_plain_fd": round(t_plain / t_adj_h, 2),
                 "max_err_adjoint_hermitian_vs_fd_rel_to_scale": float(max(errs_h)),
                 "max_abs_diff_adjoint_hermitian_vs_solved_lambda": diff_h}
        res[name] = entry
        log(f"aat set {name}: {entry}")
    return res


def cisd_field_set_measurement(dev, seed, O=24, V=200, h=1e-4, max_space=8):
    """CISD field-derivative set sharing one integral set: 6 instances (3
    field components x +/- h, i x antisymmetric one-electron
    perturbations) of the Hermitian CISD lowest root.  Plain loop: one
    Davidson solve per instance on the same kernels (batch of one);
    batched: all 6 in one multi-instance Davidson (stacked PPL/ring GEMMs);
    field set: the root of the undisplaced F first, then all 6 in one
    shared subspace (one full sigma per new direction, the +/- pairs
    sharing theirs; L(dF) products for the rest), against the same loop on
    the packed-vector Davidson that it runs on; adjoint: one solve at the
    base point plus the one-particle density (dE/dF), giving all three
    derivatives."""
    import numpy as np
    import torch
    from apyib_amd import synthetic
    from apyib_amd.davidson import CISDDavidson, FieldSetDavidson, MultiDavidson
    from apyib_amd.so_ci import SOCI, SOIntegrals

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    t0 = time.perf_counter()
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    F, eps = synthetic.fock(O, V, seed)
    n = O + V
    F0 = torch.from_numpy(F).to(dev)
    dirs = [torch.from_numpy(synthetic.fock_perturbation(n, seed * 1000 + 500 + k, 'field')).to(dev) for k in range(3)]
    focks = [F0 + sgn * h * dF for dF in dirs for sgn in (+1, -1)]
    torch.cuda.synchronize()
    log(f"cisd field set: integrals {ints.nbytes() / 1e9:.2f} GB in {time.perf_counter() - t0:.1f} s")
    conv = dict(e_conv=1e-12, r_conv=1e-6)

    def plain():
        E, its = [], []
        for Fk in focks:
            md = MultiDavidson(ints, [Fk], eps, max_space)
            E.append(md.solve(**conv)[0])
            its.append(md.dav[0].iterations)
            del md
        return np.array(E), its
    (Ep, its_p), t_plain = timed(plain)
    log(f"cisd field set: plain loop {t_plain:.2f} s, iterations {its_p}")

    def batched():
        md = MultiDavidson(ints, focks, eps, max_space)
        E = md.solve(**conv)
        its = [d.iterations for d in md.dav]
        del md
        return E, its
    (Eb, its_b), t_batch = timed(batched)
    log(f"cisd field set: batched {t_batch:.2f} s, iterations {its_b}")

    def warm(batch):
        """The undisplaced root first, then the displaced instances from it
        (batched, or one at a time)."""
        md = MultiDavidson(ints, [F0], eps, max_space)
        md.solve(**conv)
        c0 = md.dav[0].c.clone()
        base_its = md.dav[0].iterations
        del md
        if batch:
            md = MultiDavidson(ints, focks, eps, max_space)
            E = md.solve(**conv, guess=[c0] * len(focks))
            its = [d.iterations for d in md.dav]
            del md
            return np.asarray(E), [base_its] + its
        E, its = [], [base_its]
        for Fk in focks:
            md = MultiDavidson(ints, [Fk], eps, max_space)
            E.append(md.solve(**conv, guess=[c0])[0])
            its.append(md.dav[0].iterations)
            del md
        return np.array(E), its
    (Ew, its_w), t_warm = timed(lambda: warm(True))
    log(f"cisd field set: batched warm {t_warm:.2f} s, iterations {its_w}")
    (Epw, its_pw), t_plain_warm = timed(lambda: warm(False))
    log(f"cisd field set: one at a time warm {t_plain_warm:.2f} s, iterations {its_pw}")

    def plain_packed():
        E, its, ns = [], [], 0
        for Fk in focks:
            dav = CISDDavidson(SOCI(ints, Fk, eps, singles=True, fov_coupling=True), max_space)
            E.append(dav.solve(**conv))
            its.append(dav.iterations)
            ns += dav.nsigma
            del dav
        return np.array(E), its, ns
    (Epp, its_pp, ns_pp), t_plain_packed = timed(plain_packed)
    log(f"cisd field set: plain loop (packed vectors) {t_plain_packed:.2f} s, iterations {its_pp}, sigmas {ns_pp}")

    def field_set():
        inst = [(d, sgn * h) for d in range(len(dirs)) for sgn in (+1, -1)]
        fs = FieldSetDavidson(ints, F0, eps, dirs, inst, max_space=16, base_space=max_space)
        E = fs.solve(**conv)
        out = (E, fs.base_iterations, fs.iterations, fs.nsigma, fs.nfock, fs.converged)
        del fs
        return out
    (Efs, its_fb, its_fs, ns_fs, nf_fs, conv_fs), t_field = timed(field_set)
    log(f"cisd field set: shared subspace {t_field:.2f} s, base iterations {its_fb}, shared {its_fs}, "
        f"sigmas {ns_fs}, L(dF) products {nf_fs}")

    def adjoint():
        md = MultiDavidson(ints, [F0], eps, max_space)
        md.solve(**conv)
        D = md.dav[0].gradient_fock()
        g = torch.stack([torch.sum(D * dF) for dF in dirs]).cpu().numpy()
        scale = np.array([float(torch.sum((D * dF).abs()).cpu()) for dF in dirs])
        del md
        return g, scale
    (g, scale), t_adj = timed(adjoint)
    fd_b = (Eb[0::2] - Eb[1::2]) / (2 * h)
    fd_p = (Ep[0::2] - Ep[1::2]) / (2 * h)
    res = {"workload": f"Hermitian CISD lowest root, O={O} V={V}, 6 instances (3 field components x +/-{h}) "
                       f"sharing one seeded <pq||rs>", "instances": 6,
           "plain_loop_s": round(t_plain, 3), "batched_s": round(t_batch, 3), "adjoint_s": round(t_adj, 3),
           "batched_warm_s": round(t_warm, 3), "plain_warm_s": round(t_plain_warm, 3),
           "speedup_batched": round(t_plain / t_batch, 2), "speedup_batched_warm": round(t_plain / t_warm, 2),
           "speedup_plain_warm": round(t_plain / t_plain_warm, 2),
           "speedup_adjoint_vs_plain_fd": round(t_plain / t_adj, 2),
           "plain_packed_s": round(t_plain_packed, 3), "sigmas_plain_packed": int(ns_pp),
           "field_set_s": round(t_field, 3), "sigmas_field_set": int(ns_fs), "fock_products_field_set": int(nf_fs),
           "iterations_field_set_base_then_shared": [int(its_fb), int(its_fs)], "field_set_converged": bool(conv_fs),
           "speedup_field_set_vs_plain_packed": round(t_plain_packed / t_field, 2),
           "speedup_field_set_vs_plain": round(t_plain / t_field, 2),
           "max_abs_dE_field_set_vs_plain": float(np.abs(Efs - Ep).max()),
           "max_abs_dE_plain_packed_vs_plain": float(np.abs(Epp - Ep).max()),
           "iterations_plain": its_p, "iterations_batched": its_b,
           "iterations_warm_base_then_instances": its_w,
           "max_abs_dE_warm_vs_plain": float(np.abs(Ew - Ep).max()),
           "max_abs_dE_batched_vs_plain": float(np.abs(Eb - Ep).max()),
           "max_err_adjoint_vs_fd_rel_to_scale": float((np.abs(g.real - fd_b) / scale).max()),
           "max_err_fd_batched_vs_plain_rel_to_scale": float((np.abs(fd_b - fd_p) / scale).max())}
    log(f"cisd field set: {res}")
    return res


def eri_so_measurement(dev, seed, m=90):
    """compute_ERI_SO (utils.py:317-365) at a C1-sized molecule, m = 90
    spatial orbitals -> 180 spin orbitals (1.05 GB in, 16.8 GB out): the GPU
    expansion (apy_eri_so_from_mo) at full size, the reference's four-index
    Python loop on a subsample of one outer index (extrapolated by 2m), and
    the vectorized NumPy form at m/2 (scaled by the element count); the GPU
    and loop results agree bit for bit on the subsample."""
    import numpy as np
    import torch
    from apyib_amd import transforms as tr
    from oracle import transforms as otr
    g = torch.Generator(device=dev).manual_seed(seed + 90)
    x = torch.complex(torch.randn((m,) * 4, generator=g, dtype=torch.float64, device=dev),
                      torch.randn((m,) * 4, generator=g, dtype=torch.float64, device=dev))
    out = tr.eri_so(x)                                    # warm
    del out
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(3):
        e0.record()
        out = tr.eri_so(x)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
        if len(ts) < 3:
            del out
    gpu_ms = min(ts)
    n = 2 * m
    host = x.cpu().numpy()
    t0 = time.perf_counter()
    loop = otr.compute_ERI_SO_loop(host, p_range=[0])     # p = 0: n^3 iterations of the loop body
    t_loop = time.perf_counter() - t0
    exact = bool(np.array_equal(loop[0], out[0].cpu().numpy()))
    del out, loop
    mh = m // 2
    t0 = time.perf_counter()
    vec = otr.compute_ERI_SO(host[:mh, :mh, :mh, :mh].copy())
    t_vec = time.perf_counter() - t0
    del vec, host, x
    torch.cuda.empty_cache()
    bytes_moved = 16.0 * (m ** 4 + n ** 4)
    res = {"workload": f"compute_ERI_SO at m={m} spatial -> {n} spin orbitals (16.8 GB complex128 out)",
           "gpu_ms": round(gpu_ms, 3), "gpu_tbps": round(bytes_moved / (gpu_ms * 1e-3) / 1e12, 3),
           "cpu_loop_s_sample": round(t_loop, 2), "cpu_loop_sample": f"outer index p = 0 of {n} ({n ** 3} iterations)",
           "cpu_loop_s_extrapolated": round(t_loop * n, 1),
           "cpu_vectorized_s_sample": round(t_vec, 2), "cpu_vectorized_sample": f"m = {mh} (1/16 of the elements)",
           "cpu_vectorized_s_extrapolated": round(t_vec * (m / mh) ** 4, 1),
           "gpu_equals_loop_on_sample": exact, "cores": 1}
    log(f"eri_so: {res}")
    return res


def gauss3m_measurement(dev, ints, F, eps, hip, fl, max_space, K, W, theta_ref):
    """The headline Davidson iterations with the PPL and ring products in the
    three-product complex form (SOCI gauss3m: 3 real products per complex
    multiply-add instead of 4).  Reported as a time-to-solution option beside
    the four-product headline: the effective rate counts the same 8 flops per
    complex multiply-add as value, the executed rate the matrix-core flops
    actually issued (6 per complex multiply-add in those two products)."""
    import torch
    from apyib_amd.davidson import CISDDavidson
    from apyib_amd.so_ci import SOCI, executed_flops
    solver = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True, gauss3m=True)
    dav = CISDDavidson(solver, max_space=max_space)
    gen = dav.iterate_device()
    next(gen)
    for _ in range(W):
        next(gen)
    torch.cuda.synchronize()
    evs = [(hip.create(), hip.create()) for _ in range(K)]
    n0 = dav.nsigma
    t0 = time.perf_counter()
    for k in range(K):
        solver.ev_dominant = evs[k]
        next(gen)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    solver.ev_dominant = None
    _, theta, rn, _ = dav.read_state()
    gen.close()
    nsig = dav.nsigma - n0
    ppl_ms = sum(hip.elapsed_ms(a, b) for a, b in evs) / K
    for a, b in evs:
        hip.destroy(a)
        hip.destroy(b)
    ms = elapsed * 1e3 / K
    eff = sum(fl.values()) * nsig / K / (ms * 1e-3) / 1e12
    exe = executed_flops(fl, gauss3m=True) * nsig / K / (ms * 1e-3) / 1e12
    log(f"three-product step {ms:.3f} ms, PPL {ppl_ms:.3f} ms, effective {eff:.1f} / executed {exe:.1f} TFLOP/s, "
        f"E_corr = {theta:.12f} (four-product {theta_ref:.12f})")
    del dav, solver
    return {"what": "the headline Davidson iterations with the PPL and ring GEMMs in the three-product complex form "
                    "(T1 = ar br, T2 = ai bi, T3 = (ar+ai)(br+bi); 3 FP64 MFMA products per complex multiply-add "
                    "instead of 4); an opt-in for operands whose imaginary parts are not small",
            "ms_per_step": round(ms, 4), "sigma_per_step": round(nsig / K, 3),
            "effective_tflops": round(eff, 3), "executed_tflops": round(exe, 3),
            "ppl_kernel_ms": round(ppl_ms, 4),
            "ppl_executed_tflops": round(0.75 * fl['ppl'] / (ppl_ms * 1e-3) / 1e12, 3),
            "ppl_effective_tflops": round(fl['ppl'] / (ppl_ms * 1e-3) / 1e12, 3),
            "E_corr_after_steps": theta, "residual_after_steps": rn,
            "E_corr_minus_four_product": theta - theta_ref}


def spatial_step_flops(o, v):
    """Executed flops of one spin-adapted CISD step (apy_cisd_sp_step, F_vv folded into the ring
    operator): the two ladder halves, HHL, the ring X1 (K = 2ov) and X2, singles->doubles and the six
    singles products."""
    ov, oo, vv = o * v, o * o, v * v
    so, ao, sv, av = o * (o + 1) // 2, o * (o - 1) // 2, v * (v + 1) // 2, v * (v - 1) // 2
    ladder = 8.0 * (so * sv * sv + ao * av * av)
    return ladder, ladder + 8.0 * (oo * vv * oo + 3 * ov ** 3 + o * vv * o * v + v * oo * v * o + ov * ov
                                   + 2 * ov * ov + o * v * ov * v + o * v * oo * v + o * v * o + o * v * v)


def spatial_step_flops_gauss3m(o, v):
    """Matrix-core flops one spin-adapted step issues with the ladder and ring
    products in the three-product form (6 instead of 8 per complex
    multiply-add in those products)."""
    ladder, total = spatial_step_flops(o, v)
    return total - 0.25 * (ladder + 8.0 * 3 * (o * v) ** 3)


def spatial_cisd_measurement(dev, seed, o=10, v=80, steps=10, tune=False, gauss3m=True):
    """Spin-adapted closed-shell CISD residual steps (apy_cisd_sp_step) at
    o=10, v=80 spatial orbitals (the closed-shell counterpart of the
    O=20, V=160 spin-orbital workload), seeded chemists' integrals with the
    particle-exchange symmetry, generated on the host and copied once."""
    import numpy as np
    import torch
    from apyib_amd.spatial_ci import SPIntegrals, SpatialCI

    rng = np.random.default_rng(seed + 31)
    n = o + v
    E4 = 0.05 * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    E4 = torch.from_numpy(0.5 * (E4 + E4.transpose(2, 3, 0, 1))).to(dev)
    eps = np.concatenate([np.sort(rng.uniform(-1.5, -0.3, o)), np.sort(rng.uniform(0.2, 3.0, v))])
    F = torch.from_numpy(np.diag(eps).astype(complex)).to(dev)
    ci = SpatialCI(SPIntegrals.from_mo(E4, o), F, eps)
    del E4
    if tune:
        ci.tune()
    t = ci.initial_guess(ci.new_vector())
    tn, r = ci.new_vector(), ci.new_vector()
    ci.energy(t)
    for _ in range(2):
        ci.step(t, tn, r)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ci.step(t, tn, r)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    ov, oo, vv = o * v, o * o, v * v
    ladder, flops = spatial_step_flops(o, v)
    res = {"workload": f"spin-adapted CISD residual step (ci_wfn.py:457-486), o={o} v={v} spatial orbitals",
           "ms_per_step": round(ms, 3), "executed_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
           "ladder_flops": ladder, "step_flops": flops,
           "plain_ladder_flops": 8.0 * oo * vv * vv}
    if not gauss3m:
        log(f"spatial cisd: {res}")
        return res
    # the same steps with the ladder and ring products in the three-product form
    ci3 = SpatialCI(ci.ints, F, eps, gauss3m=True)
    tn3, r3 = ci3.new_vector(), ci3.new_vector()
    ci.step(t, tn, r)
    ci3.step(t, tn3, r3, energy=ci.energy_dev)     # the same E in both residuals
    diff3 = float((r3 - r).abs().max().cpu()) / max(1e-300, float(r.abs().max().cpu()))
    for _ in range(2):
        ci3.step(t, tn3, r3, energy=ci.energy_dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ci3.step(t, tn3, r3, energy=ci.energy_dev)
    torch.cuda.synchronize()
    ms3 = (time.perf_counter() - t0) / steps * 1e3
    del ci3
    res.update({"gauss3m": {"what": "the same step with the ladder and ring products in the three-product complex form",
                       "ms_per_step": round(ms3, 3),
                       "effective_tflops": round(flops / (ms3 * 1e-3) / 1e12, 2),
                       "executed_tflops": round(spatial_step_flops_gauss3m(o, v) / (ms3 * 1e-3) / 1e12, 2),
                       "max_rel_diff_residual_vs_four_product": diff3}})
    log(f"spatial cisd: {res}")
    return res


def _response_step_gauss3m(rs, X, Xn, R, sv, steps, sync_time):
    """The batched response step with the ladder and ring products in the
    three-product form: (time of `steps` steps, largest residual difference
    from the four-product step relative to its largest element)."""
    rs.gauss3m = False
    rs._step(X, Xn, R, sv)
    R4 = R.clone()
    rs.gauss3m = True
    rs._step(X, Xn, R, sv)
    diff3 = float((R - R4).abs().max().cpu()) / max(1e-300, float(R4.abs().max().cpu()))
    for _ in range(2):
        rs._step(X, Xn, R, sv)
    _, t = sync_time(lambda: [rs._step(X, Xn, R, sv) for _ in range(steps)])
    rs.gauss3m = False
    return t, diff3


def analytic_response_measurement(dev, seed, o=10, v=80, nb=15, scale=0.01, nfd=2, h=1e-5, tune=False, steps=10,
                                  gauss3m=True):
    """First-order CISD amplitudes of `nb` perturbations (the 3N+3 field
    and nuclear perturbations of a 4-atom molecule in the analytic AAT) of
    one spin-adapted CISD solution at o=10, v=80: seeded chemists' integrals
    at coupling scale `scale` (small enough for the reference's Jacobi
    iteration to converge), seeded pair-symmetric derivative integrals and
    antisymmetric derivative Fock matrices, all generated on the device.
    Legs: the batched set (ResponseSet), the same solves one at a time
    (ResponseCI), and central differences (two CISD solves with H +- h dH,
    integral blocks included) timed on `nfd` perturbations and scaled to
    `nb`; the differences also check the projected derivatives."""
    import numpy as np
    import torch
    from apyib_amd.spatial_ci import ResponseCI, ResponseSet, SPIntegrals, SpatialCI

    def sync_time(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    n = o + v
    g = torch.Generator(device=dev)

    def pair_sym(k, amp):
        g.manual_seed(seed * 1000 + k)
        x = torch.randn((n,) * 4, generator=g, device=dev, dtype=torch.float64)
        y = torch.randn((n,) * 4, generator=g, device=dev, dtype=torch.float64)
        e = amp * torch.complex(x, y)
        return 0.5 * (e + e.permute(2, 3, 0, 1))

    params = dict(e_convergence=1e-10, d_convergence=1e-10, DIIS=True, max_iterations=100)
    rng = np.random.default_rng(seed + 77)
    eps = np.concatenate([np.sort(rng.uniform(-1.5, -0.3, o)), np.sort(rng.uniform(0.2, 3.0, v))])
    F = torch.from_numpy(np.diag(eps).astype(complex)).to(dev)
    E4 = pair_sym(0, scale)
    ints = SPIntegrals.from_mo(E4, o)
    ci = SpatialCI(ints, F, eps)
    if tune:
        ci.tune()
        ci.tune(nb)
    (E, _, _, it_cisd), t_cisd = sync_time(lambda: ci.solve(params))
    t_ref = ci.t.clone()

    def pert(k):
        dE4 = pair_sym(k + 1, scale)
        g.manual_seed(seed * 1000 + 500 + k)
        a = 0.02 * torch.randn((n, n), generator=g, device=dev, dtype=torch.float64)
        return dE4, torch.complex(a - a.T, torch.zeros_like(a))

    def build_set():
        rs = ResponseSet(ints, F, eps, True, t_ref, E)
        for k in range(nb):
            dE4, dF = pert(k)
            rs.add(SPIntegrals.from_mo(dE4, o), dF)
        return rs
    rs, t_setup = sync_time(build_set)
    (dE_set, _, _, its), t_set = sync_time(lambda: rs.solve(params))
    # one batched response step (apy_cisd_sp_step_batch over the nb vectors), timed alone
    X = torch.zeros((nb, rs.L), dtype=torch.complex128, device=dev)
    Xn, R = torch.zeros_like(X), torch.zeros_like(X)
    sv = torch.zeros(nb, dtype=torch.complex128, device=dev)
    rs._step(X, Xn, R, sv)
    X.copy_(Xn)
    for _ in range(2):
        rs._step(X, Xn, R, sv)
    _, t_steps = sync_time(lambda: [rs._step(X, Xn, R, sv) for _ in range(steps)])
    t_steps3 = diff3 = None
    if gauss3m:
        t_steps3, diff3 = _response_step_gauss3m(rs, X, Xn, R, sv, steps, sync_time)
    del rs, X, Xn, R

    def one_by_one():
        out = []
        for k in range(nb):
            dE4, dF = pert(k)
            r = ResponseCI(ints, F, eps, True, t_ref, E, SPIntegrals.from_mo(dE4, o), dF)
            out.append(r.solve(params)[0])
        return out
    dE_seq, t_seq = sync_time(one_by_one)

    def fd():
        out = []
        for k in range(nfd):
            dE4, dF = pert(k)
            Es = []
            for sg in (1.0, -1.0):
                c = SpatialCI(SPIntegrals.from_mo(E4 + sg * h * dE4, o), F + sg * h * dF, eps)
                Es.append(c.solve(params)[0])
            out.append((Es[0] - Es[1]) / (2 * h))
        return out
    dE_fd, t_fd = sync_time(fd)
    err_fd = max(abs(a - b) / max(1.0, abs(b)) for a, b in zip(dE_set[:nfd], dE_fd))
    err_seq = max(abs(a - b) / max(1.0, abs(b)) for a, b in zip(dE_set, dE_seq))
    fd_scaled = t_fd / nfd * nb
    res = {"workload": f"first-order CISD amplitudes of {nb} perturbations of one spin-adapted CISD solution "
                       f"(analytic AAT response, analytic_aats.py:733-890), o={o} v={v}, coupling scale {scale}",
           "cisd_solve_ms": round(t_cisd * 1e3, 1), "cisd_iterations": it_cisd,
           "set_setup_ms": round(t_setup * 1e3, 1), "set_solve_ms": round(t_set * 1e3, 1),
           "set_iterations_max": int(max(its)),
           "response_step_ms": round(t_steps / steps * 1e3, 3),
           "response_step_executed_tflops": round(nb * spatial_step_flops(o, v)[1] / (t_steps / steps) / 1e12, 2),
           "response_step_gauss3m": None if t_steps3 is None else {
               "what": "the same batched step with the ladder and ring products in the three-product complex form",
               "ms": round(t_steps3 / steps * 1e3, 3),
               "effective_tflops": round(nb * spatial_step_flops(o, v)[1] / (t_steps3 / steps) / 1e12, 2),
               "executed_tflops": round(nb * spatial_step_flops_gauss3m(o, v) / (t_steps3 / steps) / 1e12, 2),
               "max_rel_diff_residual_vs_four_product": diff3},
           "one_by_one_ms": round(t_seq * 1e3, 1),
           "central_differences_ms": round(fd_scaled * 1e3, 1),
           "central_differences_measured": f"{nfd} perturbations, scaled to {nb}",
           "speedup_set_vs_one_by_one": round(t_seq / (t_setup + t_set), 2),
           "speedup_set_vs_differences": round(fd_scaled / (t_setup + t_set), 2),
           "max_rel_diff_set_vs_one_by_one": float(err_seq),
           "max_rel_diff_vs_differences": float(err_fd)}
    log(f"analytic response: {res}")
    return res


H2O2 = """
H -1.780954530308296   1.411647335546379   0.872055376436941
H  1.780954530308296  -1.411647335546379   0.872055376436941
O -1.371214332646589  -0.115525249760340  -0.054947416764017
O  1.371214332646589   0.115525249760340  -0.054947416764017
no_com
no_reorient
symmetry c1
units bohr
"""


def h2o2_aat_measurement(dev, h=1e-4):
    """End-to-end CISD atomic axial tensor of H2O2/STO-3G (12 basis
    functions, the molecule of the reference's analytic AAT tests) by the
    two routes the reference offers: analytic (analytic_aats.py:553-1391:
    SCF, CISD, CPHF, 15 first-order CI solves as one set) and finite
    differences (parallel.py:16-48: 30 displaced SCF + CISD solves with the
    Hartree-Fock parts batched, then the determinant-based AAT).  The 30
    displaced solves alone are timed batched and one by one.  Both AATs are
    compared."""
    import contextlib
    import copy
    import numpy as np
    import torch
    from apyib_amd.analytic_aats import analytic_derivative
    from apyib_amd.energy import energy, phase_corrected_energies, phase_corrected_energy
    from apyib_amd.molecule import Molecule
    from apyib_amd.parallel import compute_parallel_aats

    params = {'geom': H2O2, 'basis': 'STO-3G', 'method': 'CISD', 'e_convergence': 1e-13, 'd_convergence': 1e-13,
              'DIIS': True, 'freeze_core': False, 'F_el': [0.0, 0.0, 0.0], 'F_mag': [0.0, 0.0, 0.0],
              'max_iterations': 120}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with contextlib.redirect_stdout(sys.stderr):
            r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    timed(lambda: energy(copy.deepcopy(params)))                      # warm the kernels and plans
    I_an, t_an = timed(lambda: analytic_derivative(copy.deepcopy(params)).compute_CISD_AATs('full', 'non-canonical'))
    I_fd, t_fd = timed(lambda: compute_parallel_aats(copy.deepcopy(params), h, h, 'full'))

    (_, _, C0, b0), _ = timed(lambda: energy(copy.deepcopy(params)))
    mol = Molecule.from_string(H2O2)
    plist = []
    for sign in (1, -1):
        for alpha in range(3 * mol.natom()):
            g = mol.coords.copy()
            g[alpha // 3][alpha % 3] += sign * h
            q = copy.deepcopy(params)
            q['geom'] = mol.with_coords(g).geometry_string()
            plist.append(q)
    for sign in (1, -1):
        for beta in range(3):
            q = copy.deepcopy(params)
            q['F_mag'][beta] += sign * h
            plist.append(q)
    _, t_seq = timed(lambda: [phase_corrected_energy(q, b0, C0) for q in plist])
    _, t_bat = timed(lambda: phase_corrected_energies(plist, b0, C0))
    diff = float(np.abs(np.asarray(I_an) - np.asarray(I_fd)).max())
    res = {"workload": "CISD atomic axial tensor of H2O2/STO-3G (12 basis functions), analytic vs finite "
                       f"differences (h = {h})",
           "analytic_s": round(t_an, 3), "finite_differences_s": round(t_fd, 3),
           "displaced_solves_one_by_one_s": round(t_seq, 3), "displaced_solves_batched_s": round(t_bat, 3),
           "max_abs_diff_analytic_vs_differences": diff, "max_abs_aat": float(np.abs(I_an).max())}
    log(f"h2o2 aat: {res}")
    return res


def mp2_adjoint_set_measurement(dev, seed, O=30, V=270, ninst=8, ndir=200, nfd=2, h=1e-5, tol=1e-10):
    """MP2 adjoint set: `ninst` orbital-invariant spin-orbital MP2 instances
    at O=30, V=270 (only <ij||ab>, <ab||ij> and F), each with `ndir` seeded
    Hermitian directions (the last 4 on <pq||rs>: real draws with its pair
    antisymmetry, the rest on F).  Adjoint: one amplitude solve, one lambda
    solve and the gradient densities per instance give all directional
    derivatives (the Hermitian leg takes lambda = conj(t)/4 instead of its
    solve).  Plain loop: central differences, two solves per
    direction, timed on `nfd` directions per instance (one on F, one on
    <pq||rs>) and scaled to `ndir`; those directions also check the adjoint
    against the differences."""
    import numpy as np
    import torch
    from apyib_amd import synthetic
    from apyib_amd.adjoint import MP2Adjoint
    from apyib_amd.kernels import zdot_many
    from apyib_amd.so_ci import MP2_BLOCKS, SOIntegrals

    def sync_time(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    n, N2 = O + V, O * O * V * V
    t_adj = t_fd = t_adj_h = diff_h = 0.0
    errs, gens = [], 0.0
    for k in range(ninst):
        sk = seed * 100 + k
        F, eps = synthetic.fock(O, V, sk)
        F = torch.from_numpy(F).to(dev)
        ints = SOIntegrals.synthetic(O, V, sk, device=dev, blocks=MP2_BLOCKS)
        W, Vd = ints["oovv"].view(N2), ints["vvoo_t"].view(N2)
        dFs = [torch.from_numpy(synthetic.fock_perturbation(n, sk * 1000 + d, "nuclear" if d % 2 else "field"))
               .to(dev) for d in range(ndir - 4)]
        if k == 0:                                                   # warm the kernels and plans once
            MP2Adjoint(F, W, Vd, eps, O, tol=tol).gradients()

        def adjoint(hermitian=False):
            s = MP2Adjoint(F, W, Vd, eps, O, tol=tol, hermitian=hermitian)
            s.energy()
            G, gW, gV = s.gradients()
            out = list(zdot_many(dFs, G))
            for d in range(4):                                       # <pq||rs> directions, generated on the device
                dI = SOIntegrals(O, V, dev, MP2_BLOCKS)
                dI.add_synthetic(sk * 1000 + 900 + d, tag=300 + d, std=1.0, accumulate=False, real_only=True)
                out.append(zdot_many([dI["oovv"]], gW)[0] + zdot_many([dI["vvoo_t"]], gV)[0])
                del dI
            return np.array(out), G, gW, gV
        (g, G, gW, gV), ta = sync_time(adjoint)
        t_adj += ta
        # the same with lambda = conj(t)/4 (these seeded instances are Hermitian)
        (gh, _, _, _), tah = sync_time(lambda: adjoint(hermitian=True))
        t_adj_h += tah
        diff_h = max(diff_h, float(np.abs(gh - g).max() / max(1e-300, np.abs(g).max())))

        def fd_pair(which):
            Ev = []
            for sgn in (+1, -1):
                if which == "F":
                    Ev.append(complex(MP2Adjoint(F + sgn * h * dFs[0], W, Vd, eps, O, tol=tol).energy()))
                else:
                    P = SOIntegrals.synthetic(O, V, sk, device=dev, blocks=MP2_BLOCKS)
                    P.add_synthetic(sk * 1000 + 900, tag=300, std=1.0, scale=sgn * h, accumulate=True,
                                    real_only=True)
                    Ev.append(complex(MP2Adjoint(F, P["oovv"].view(N2), P["vvoo_t"].view(N2), eps, O,
                                                 tol=tol).energy()))
                    del P
            return (Ev[0] - Ev[1]) / (2 * h)
        (fF, fW), tf = sync_time(lambda: (fd_pair("F"), fd_pair("W")))
        t_fd += tf
        scF = float(torch.sum((G.view(-1) * dFs[0].view(-1)).abs()))
        dI = SOIntegrals(O, V, dev, MP2_BLOCKS)
        dI.add_synthetic(sk * 1000 + 900, tag=300, std=1.0, accumulate=False, real_only=True)
        scW = float(torch.sum((gW * dI["oovv"].view(N2)).abs()) + torch.sum((gV * dI["vvoo_t"].view(N2)).abs()))
        del dI
        errs += [abs(g[0].real - fF.real) / scF, abs(g[ndir - 4].real - fW.real) / scW]
        log(f"mp2 adjoint set: instance {k}: adjoint {ta:.3f} s, 2 FD pairs {tf:.3f} s, "
            f"err {errs[-2]:.2e} {errs[-1]:.2e}")
        del ints, W, Vd, G, gW, gV
        torch.cuda.empty_cache()
    plain = t_fd * ndir / nfd
    res = {"workload": f"orbital-invariant spin-orbital MP2 energy, {ninst} instances O={O} V={V}, {ndir} seeded "
                       f"Hermitian directions each ({ndir - 4} on F, 4 on <pq||rs>)",
           "instances": ninst, "directions_per_instance": ndir,
           "adjoint_s": round(t_adj, 3),
           "plain_fd_s_measured": round(t_fd, 3), "plain_fd_directions_measured": nfd * ninst,
           "plain_fd_s_scaled": round(plain, 2),
           "speedup_adjoint_vs_plain_fd": round(plain / t_adj, 1),
           "fd_step": h, "max_err_adjoint_vs_fd_rel_to_scale": float(max(errs)),
           "adjoint_hermitian_s": round(t_adj_h, 3),
           "speedup_adjoint_hermitian_vs_plain_fd": round(plain / t_adj_h, 1),
           "max_rel_diff_adjoint_hermitian_vs_solved_lambda": diff_h}
    log(f"mp2 adjoint set: {res}")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--O", type=int, default=20)
    ap.add_argument("--V", type=int, default=160)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-tune", action="store_true", help="use neither saved plans nor the autotuner (cost model)")
    ap.add_argument("--tune", action="store_true", help="re-run the GEMM autotuner instead of loading saved plans")
    ap.add_argument("--plans", default=os.path.join(ROOT, "apyib_amd", "plans_gfx950.json"),
                    help="saved GEMM plans (loaded unless --tune or --no-tune)")
    ap.add_argument("--save-plans", default=None, help="write the tuned plans here (with --tune)")
    ap.add_argument("--max-space", type=int, default=12, help="Davidson subspace size before restart")
    ap.add_argument("--host-control", action="store_true",
                    help="Davidson with the projected eigenproblem on the host (numpy eigh, one host read per "
                         "iteration) instead of on the device")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph-replayed Davidson timings (fresh-sigma and combined restart)")
    ap.add_argument("--no-gauss3m", action="store_true",
                    help="skip the three-product (Gauss) complex GEMM legs (headline iterations, spin-adapted "
                         "and response steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-aat-set", action="store_true", help="skip the batched finite-difference set timing")
    ap.add_argument("--no-cisd-set", action="store_true", help="skip the CISD field-derivative set timing")
    ap.add_argument("--no-mp2-adjoint-set", action="store_true",
                    help="skip the MP2 adjoint set (8 instances O=30 V=270)")
    ap.add_argument("--no-spatial", action="store_true", help="skip the spin-adapted CISD step timing")
    ap.add_argument("--no-eri-so", action="store_true", help="skip the compute_ERI_SO timing (GPU and CPU sample)")
    ap.add_argument("--no-analytic-response", action="store_true",
                    help="skip the first-order CISD amplitude set timing")
    ap.add_argument("--no-h2o2-aat", action="store_true",
                    help="skip the end-to-end H2O2 AAT timing (analytic and finite differences)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} (WORLD_SIZE={world}) is not supported: "
              "sharding the instance sets and the ladder product over GPUs is designed but not built",
              file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, ROOT)
    import numpy as np
    import torch
    from apyib_amd import synthetic
    from apyib_amd.davidson import CISDDavidson
    from apyib_amd.so_ci import SOCI, SOIntegrals, dense_flops, step_flops

    dev = torch.device("cuda:0")
    O, V = args.O, args.V
    t0 = time.perf_counter()
    ints = SOIntegrals.synthetic(O, V, args.seed, device=dev)
    F, eps = synthetic.fock(O, V, args.seed)
    solver = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True)
    torch.cuda.synchronize()
    log(f"setup: integrals {ints.nbytes() / 1e9:.2f} GB generated on device in {time.perf_counter() - t0:.2f} s")
    from apyib_amd.kernels import plans_clear, plans_export, plans_load
    plans = "model"
    plans_clear()   # the solver installed the saved plans; --no-tune and --tune start from none
    if args.no_tune or args.tune:
        os.environ["APYIB_AMD_PLANS"] = "none"   # and later solvers install none either
    if not args.no_tune:
        if not args.tune and os.path.exists(args.plans):
            plans_load(json.load(open(args.plans)))
            plans = os.path.relpath(args.plans, ROOT)
            log(f"GEMM plans loaded from {plans}")
        else:
            t0 = time.perf_counter()
            solver.tune(reps=5)
            torch.cuda.synchronize()
            plans = "autotuned in this run"
            log(f"autotune: {time.perf_counter() - t0:.1f} s")
            if args.save_plans:
                json.dump(plans_export(), open(args.save_plans, "w"), indent=1)

    fl = step_flops(O, V, True)
    flops_sigma = sum(fl.values())
    hip = HipEvents()
    peaks = measured_peaks(dev, hip)
    K, W = args.steps, args.warmup
    evs = [(hip.create(), hip.create()) for _ in range(K)]

    # a step = one Davidson iteration of the lowest root: sigma of the new
    # vector (plus the restart sigma every max_space iterations), CI-metric
    # projections, Ritz pair, residual, preconditioner, Gram-Schmidt
    dav = CISDDavidson(solver, max_space=args.max_space)
    # the sigma calls of the timed iterations, bracketed by events on their stream
    sigma_events = []
    sigma_plain = dav._sigma
    stream = torch.cuda.current_stream(dev).cuda_stream

    def sigma_timed(q):
        if not timing[0]:
            return sigma_plain(q)
        e0, e1 = hip.create(), hip.create()
        hip.record(e0, stream)
        sigma_plain(q)
        hip.record(e1, stream)
        sigma_events.append((e0, e1))
    timing = [False]
    dav._sigma = sigma_timed
    device_control = not args.host_control
    # device control: Ritz pair (apy_heev_lowest), residual and convergence test on the GPU, no host
    # read inside an iteration; host control: numpy eigh and one host read per iteration
    gen = dav.iterate_device() if device_control else dav.iterate()
    next(gen)                       # reference vector, its sigma and the first Ritz pair
    for _ in range(W):
        next(gen)
    torch.cuda.synchronize()
    n0 = dav.nsigma
    t_start = time.perf_counter()
    theta = rn = None
    timing[0] = True
    for k in range(K):
        solver.ev_dominant = evs[k]
        out = next(gen)
        if not device_control:
            theta, rn = out
    torch.cuda.synchronize()
    timing[0] = False
    solver.ev_dominant = None
    elapsed = time.perf_counter() - t_start
    if device_control:
        _, theta, rn, _ = dav.read_state()
    sigma_ms = [hip.elapsed_ms(a, b) for a, b in sigma_events]
    for a, b in sigma_events:
        hip.destroy(a)
        hip.destroy(b)
    nsig = dav.nsigma - n0
    gen.close()
    ms_step = elapsed * 1e3 / K
    ppl_ms = [hip.elapsed_ms(a, b) for a, b in evs]
    ppl_avg = sum(ppl_ms) / len(ppl_ms)
    for a, b in evs:
        hip.destroy(a)
        hip.destroy(b)
    achieved = fl['ppl'] / (ppl_avg * 1e-3) / 1e12
    flops_step = flops_sigma * nsig / K
    value = flops_step / (ms_step * 1e-3) / 1e12
    log(f"step {ms_step:.3f} ms ({nsig} sigma in {K} iterations), PPL kernel {ppl_avg:.3f} ms "
        f"({achieved:.1f} TFLOP/s), E_corr = {theta:.12f}, residual {rn:.2e}")

    graph_leg = None
    if device_control and not args.no_graph:
        # the same iterations replayed from HIP graphs (one per subspace size, captured before timing)
        dav._sigma = sigma_plain
        dg = CISDDavidson(solver, max_space=args.max_space)
        gg = dg.iterate_device(graph=True)
        next(gg)
        for _ in range(W):
            next(gg)
        torch.cuda.synchronize()
        g0 = dg.nsigma
        tg = time.perf_counter()
        for _ in range(K):
            next(gg)
        torch.cuda.synchronize()
        eg = time.perf_counter() - tg
        gsig = dg.nsigma - g0
        _, theta_g, rn_g, _ = dg.read_state()
        gg.close()
        ms_g = eg * 1e3 / K
        graph_leg = {"what": "the same Davidson iterations replayed from captured HIP graphs (one per subspace "
                             "size, captured before the timed region)",
                     "ms_per_step": round(ms_g, 4), "sigma_per_step": round(gsig / K, 3),
                     "tflops": round(flops_sigma * gsig / K / (ms_g * 1e-3) / 1e12, 3),
                     "E_corr_after_steps": theta_g, "residual_after_steps": rn_g}
        log(f"graph-replayed step {ms_g:.3f} ms ({gsig} sigma in {K} iterations), E_corr = {theta_g:.12f}")
        del dg

    restart_leg = None
    if device_control and not args.no_graph:
        # the same iterations with the restart sigma formed from the stored products (restart="combine")
        dav._sigma = sigma_plain
        dc = CISDDavidson(solver, max_space=args.max_space, restart="combine")
        gc = dc.iterate_device(graph=True)
        next(gc)
        for _ in range(W):
            next(gc)
        torch.cuda.synchronize()
        c0 = dc.nsigma
        tc = time.perf_counter()
        for _ in range(K):
            next(gc)
        torch.cuda.synchronize()
        ec = time.perf_counter() - tc
        csig = dc.nsigma - c0
        _, theta_c, rn_c, _ = dc.read_state()
        gc.close()
        ms_c = ec * 1e3 / K
        restart_leg = {"what": "the graph-replayed iterations with restart=\"combine\": at a restart the Ritz vector's "
                               "sigma is sum_j y_j S_j of the stored products instead of a fresh sigma",
                       "ms_per_step": round(ms_c, 4), "sigma_per_step": round(csig / K, 3),
                       "tflops": round(flops_sigma * csig / K / (ms_c * 1e-3) / 1e12, 3),
                       "E_corr_after_steps": theta_c, "residual_after_steps": rn_c,
                       "E_corr_minus_fresh_sigma_restart": theta_c - theta}
        log(f"combined-restart step {ms_c:.3f} ms ({csig} sigma in {K} iterations), E_corr = {theta_c:.12f}")
        del dc

    gauss3m = None
    if not args.no_gauss3m:
        gauss3m = gauss3m_measurement(dev, ints, F, eps, hip, fl, args.max_space, K, W, theta)

    aat_set = None
    if not args.no_aat_set:
        aat_set = aat_set_measurement(dev, args.seed)
        if not args.no_cisd_set:
            aat_set["C"] = cisd_field_set_measurement(dev, args.seed)
        if not args.no_mp2_adjoint_set:
            aat_set["D"] = mp2_adjoint_set_measurement(dev, args.seed)
    spatial = None if args.no_spatial else spatial_cisd_measurement(dev, args.seed, tune=args.tune,
                                                                    gauss3m=not args.no_gauss3m)
    eri_so = None if args.no_eri_so else eri_so_measurement(dev, args.seed)
    response = None if args.no_analytic_response else analytic_response_measurement(dev, args.seed,
                                                                                     tune=args.tune,
                                                                                     gauss3m=not args.no_gauss3m)
    h2o2 = None if args.no_h2o2_aat else h2o2_aat_measurement(dev)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(solver, O, V, flops_sigma, args.seed)
        except MemoryError as ex:  # host without ~20 GB free
            log(f"cpu baseline skipped: {ex!r}")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic",
        "config": {
            "workload": (f"CISD lowest root (spin-orbital, Hermitian) by Davidson, seeded complex128 Hamiltonian, "
                         f"O={O} V={V}; one step = one Davidson iteration"),
            "O": O, "V": V, "seed": args.seed, "solver": f"davidson (max_space {args.max_space})",
            "control": ("device: Ritz pair (apy_heev_lowest), residual and convergence test on the GPU, no host read "
                        "inside an iteration" if device_control else "host: numpy eigh, one host read per iteration"),
            "gemm_plans": plans,
            "sigma_per_step": nsig / K,
            "step_flops_executed": flops_step,
            "sigma_flops_dense_equivalent": dense_flops(O, V),
            "E_corr_after_steps": theta, "residual_after_steps": rn,
        },
        "dense_equivalent_tflops": round(dense_flops(O, V) * nsig / K / (ms_step * 1e-3) / 1e12, 3),
        "graph": graph_leg,
        "restart_combined": restart_leg,
        "gauss3m": gauss3m,
        "sigma": {"what": "the sigma calls of the timed iterations alone (HIP events on their stream around each "
                          "apy_cisd_so_step + reference-row energy); executed flops as value",
                  "calls": len(sigma_ms), "ms": round(sum(sigma_ms) / max(1, len(sigma_ms)), 4),
                  "tflops": round(flops_sigma / (sum(sigma_ms) / max(1, len(sigma_ms)) * 1e-3) / 1e12, 3),
                  "frac_of_fp64_mfma_peak": round(flops_sigma / (sum(sigma_ms) / max(1, len(sigma_ms)) * 1e-3)
                                                  / 1e12 / PEAK_FP64, 4)},
        "roofline": {
            "kernel": "zgemm_mfma (particle-particle ladder, packed pairs)",
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP64,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64, 4),
            "traffic": read_traffic(),
            "frac_of_measured_peak": round(achieved / peaks["fp64_mfma_tflops"], 4),
        },
        "measured_peaks": peaks,
        "cpu_baseline": cpu,
        "aat_set": aat_set,
        "spatial_cisd": spatial,
        "eri_so": eri_so,
        "analytic_response": response,
        "h2o2_aat": h2o2,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
