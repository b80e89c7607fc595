"""Spin-orbital CISD / CID on the GPU: device integrals and the reference's
Jacobi + DIIS iteration (ci_wfn.py:171-416, utils.py:104-140).

All tensor work runs in libapyib_amd.so.  With DIIS the bordered system and
the convergence test are solved on the device as well (solve_device), so
Python only queues iterations and reads a control block; solve_host keeps
them on the host with a handful of scalars copied back per iteration.
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .kernels import ensure_default_plans, stream_of, workspace

CDT = torch.complex128


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def npair(n):
    return n * (n - 1) // 2


ALL_BLOCKS = _lib.SO_BLOCKS
CID_BLOCKS = ("vvvv_p", "ring", "oooo_p", "vvoo_t", "oovv")
MP2_BLOCKS = ("vvoo_t", "oovv")


class SOIntegrals:
    """Antisymmetrized spin-orbital blocks <pq||rs> resident on one device,
    in the layouts documented for apy_so_ints in include/apyib_amd.h."""

    def __init__(self, O, V, device="cuda", blocks=ALL_BLOCKS):
        lib = _lib.load()
        self.O, self.V = int(O), int(V)
        self.device = torch.device(device)
        sizes = (ctypes.c_int64 * 9)()
        _lib.check(lib.apy_so_ints_sizes(self.O, self.V, sizes), "apy_so_ints_sizes")
        self.blocks = {}
        for k, name in enumerate(_lib.SO_BLOCKS):
            if name in blocks:
                self.blocks[name] = torch.empty(int(sizes[k]), dtype=CDT, device=self.device)
        self.struct = _lib.SoInts(self.O, self.V, *[
            self.blocks[nm].data_ptr() if nm in self.blocks else None for nm in _lib.SO_BLOCKS])

    def __getitem__(self, name):
        return self.blocks[name]

    def nbytes(self):
        return sum(t.numel() * 16 for t in self.blocks.values())

    @classmethod
    def from_full(cls, A, O, blocks=ALL_BLOCKS):
        """From a full antisymmetrized tensor A[n,n,n,n] = <pq||rs> (device)."""
        n = A.shape[0]
        self = cls(O, n - O, A.device, blocks)
        A = A.contiguous()
        _lib.check(_lib.load().apy_so_ints_from_full(n, _vp(A), ctypes.byref(self.struct), stream_of(A.device)),
                   "apy_so_ints_from_full")
        return self

    @classmethod
    def from_mo(cls, eri_mo, O, blocks=ALL_BLOCKS):
        """From spatial chemists' MO integrals (pq|rs) over the active space
        (utils.py:258-279), spin-expanded and antisymmetrized on the fly
        (utils.py:317-365, ci_wfn.py:284, 312-333)."""
        m = eri_mo.shape[0]
        self = cls(O, 2 * m - O, eri_mo.device, blocks)
        eri_mo = eri_mo.contiguous()
        _lib.check(_lib.load().apy_so_ints_from_mo(m, _vp(eri_mo), ctypes.byref(self.struct),
                                                   stream_of(eri_mo.device)), "apy_so_ints_from_mo")
        return self

    @classmethod
    def synthetic(cls, O, V, seed, std=0.05, device="cuda", blocks=ALL_BLOCKS):
        """Seeded synthetic blocks (generator of apyib_amd.synthetic)."""
        self = cls(O, V, device, blocks)
        self.add_synthetic(seed, tag=1, std=std, scale=1.0, accumulate=False)
        return self

    def add_synthetic(self, seed, tag=1, std=0.05, scale=1.0, accumulate=True, real_only=False):
        _lib.check(_lib.load().apy_synth_so_ints(int(seed), int(tag), float(std), int(accumulate), float(scale),
                                                 int(real_only), ctypes.byref(self.struct), stream_of(self.device)),
                   "apy_synth_so_ints")
        return self


def so_fock_from_mo(f_mo):
    """F_SO[p,q] = F_MO[p/2,q/2] delta(spin) (utils.py:283-313) on the device."""
    m = f_mo.shape[0]
    f_mo = f_mo.contiguous()
    out = torch.empty((2 * m, 2 * m), dtype=CDT, device=f_mo.device)
    _lib.check(_lib.load().apy_so_fock_from_mo(m, _vp(f_mo), _vp(out), stream_of(f_mo.device)),
               "apy_so_fock_from_mo")
    return out


def _converged(iteration, dE, rms_list, params):
    """ci_wfn.py:371-374: from iteration 2 on, |dE| < e_convergence and every
    rms (a complex number: the square root of an un-conjugated sum of
    squares) below d_convergence in NumPy's complex ordering."""
    if iteration <= 1 or not abs(dE) < params['e_convergence']:
        return False
    return all(np.complex128(r) < params['d_convergence'] for r in rms_list)


class SOCI:
    """One spin-orbital CISD (singles=True) or CID problem on one device.

    fock: [n,n] complex128 device tensor (spin-orbital Fock matrix)
    eps:  [n] float64 orbital energies used in the denominators
    ints: SOIntegrals with at least the blocks the method reads
    fov_coupling: add the P(ij)P(ab) f_bj t_ia term the reference omits
        (Hermitian operator; identical results for canonical HF orbitals).
    fold_fvv: keep a copy of the ring block with the F_vv terms folded in
        (apy_cisd_so_ring_fock, O^2 V^2 elements) so each step runs one GEMM
        fewer; off for instances whose sigma is only formed in batches.
    gauss3m: run the PPL and ring products in the three-product complex
        form (3 real products instead of 4, 3/4 of their matrix-core work).
        Its rounding error in the imaginary part scales with |re| + |im|
        rather than |im|, so it is for operands whose imaginary parts are
        not small against their real parts; real-orbital molecules in a
        weak field keep the four-product form (the default).
    """

    def __init__(self, ints, fock, eps, singles=True, fov_coupling=False, fold_fvv=True, gauss3m=False):
        self.lib = _lib.load()
        self.ints = ints
        self.O, self.V = ints.O, ints.V
        self.n = self.O + self.V
        self.device = ints.device
        self.singles = bool(singles)
        self.fov = bool(fov_coupling)
        self.gauss3m = bool(gauss3m)
        ensure_default_plans(self.device)
        self.fock = fock.to(self.device, CDT).contiguous()
        self.eps = torch.as_tensor(eps, dtype=torch.float64).to(self.device).contiguous()
        O, V = self.O, self.V
        self.OV = O * V
        self.L = O * V + O * O * V * V
        self.hoo = torch.zeros(npair(O) * npair(O), dtype=CDT, device=self.device)
        oooo = ints.blocks.get("oooo_p")
        _lib.check(self.lib.apy_cisd_so_pair_operator(O, V, _vp(self.fock), _vp(oooo), _vp(self.hoo),
                                                      stream_of(self.device)), "apy_cisd_so_pair_operator")
        self.ring_f = None
        if fold_fvv and "ring" in ints.blocks:
            self.ring_f = torch.empty(O * V * O * V, dtype=CDT, device=self.device)
            _lib.check(self.lib.apy_cisd_so_ring_fock(O, V, _vp(self.fock), _vp(ints["ring"]), _vp(self.ring_f),
                                                      stream_of(self.device)), "apy_cisd_so_ring_fock")
        self.Lp = O * V + npair(O) * npair(V)
        self.ws_bytes = int(self.lib.apy_cisd_so_workspace(O, V, int(self.singles)))
        self.red_bytes = int(self.lib.apy_reduce_workspace(self.L, _lib.APY_MAXVEC))
        self.energy_dev = torch.zeros(1, dtype=CDT, device=self.device)
        self.zero_energy = torch.zeros(1, dtype=CDT, device=self.device)
        self.scal = torch.zeros(_lib.APY_MAXVEC + 4, dtype=CDT, device=self.device)
        # optional (begin, end) hipEvent_t pair recorded around the PPL kernel
        self.ev_dominant = None

    # ------------------------------------------------------------ primitives
    def new_vector(self):
        return torch.zeros(self.L, dtype=CDT, device=self.device)

    def split(self, t):
        O, V = self.O, self.V
        return t[:self.OV].view(O, V), t[self.OV:].view(O, O, V, V)

    def step(self, t, t_new, out, energy=None, mode=_lib.APY_MODE_RESIDUAL, c0=1.0, ev_dominant=None, c0_dev=None,
             packed=False, noalloc=False):
        """One call of apy_cisd_so_step (residual update or sigma product).
        c0_dev: optional device address of the sigma-mode driver weight;
        packed: t and out are packed vectors [t1 | t2 over (i<j, a<b)];
        noalloc: apy_cisd_so_step_noalloc (the library's static scratch, no
        workspace allocated here)."""
        c0 = complex(c0)
        ev = ev_dominant or self.ev_dominant or (None, None)
        args = _lib.CisdSoArgs(
            self.O, self.V, self.fock.data_ptr(), self.eps.data_ptr(), ctypes.pointer(self.ints.struct),
            self.hoo.data_ptr(), int(self.singles), int(mode), int(self.fov),
            (energy if energy is not None else self.energy_dev).data_ptr(), c0.real, c0.imag,
            t.data_ptr(), t_new.data_ptr() if t_new is not None else None, out.data_ptr(), ev[0], ev[1], c0_dev,
            self.ring_f.data_ptr() if self.ring_f is not None else None, int(bool(packed)), int(self.gauss3m))
        if noalloc:
            _lib.check(self.lib.apy_cisd_so_step_noalloc(ctypes.byref(args), stream_of(self.device)),
                       "apy_cisd_so_step_noalloc")
            return
        ws = workspace(self.ws_bytes, self.device)
        _lib.check(self.lib.apy_cisd_so_step(ctypes.byref(args), _vp(ws), self.ws_bytes, stream_of(self.device)),
                   "apy_cisd_so_step")

    def tune(self, t=None, reps=2):
        """Time the candidate GEMM plans of one step on this problem's
        operands and cache the fastest per shape (process-wide)."""
        t = self.new_vector() if t is None else t
        args = _lib.CisdSoArgs(
            self.O, self.V, self.fock.data_ptr(), self.eps.data_ptr(), ctypes.pointer(self.ints.struct),
            self.hoo.data_ptr(), int(self.singles), _lib.APY_MODE_SIGMA, int(self.fov),
            self.energy_dev.data_ptr(), 1.0, 0.0, t.data_ptr(), None, t.data_ptr(), None, None, None,
            self.ring_f.data_ptr() if self.ring_f is not None else None, 0, int(self.gauss3m))
        ws = workspace(self.ws_bytes, self.device)
        _lib.check(self.lib.apy_cisd_so_tune(ctypes.byref(args), _vp(ws), self.ws_bytes, int(reps),
                                             stream_of(self.device)), "apy_cisd_so_tune")

    def sigma(self, x, out, c0=1.0, c0_dev=None, packed=False):
        """out = H_N (c0, x): the CISD sigma with reference weight c0 (or the
        device complex scalar at address c0_dev); packed vectors with
        packed=True."""
        self.step(x, None, out, mode=_lib.APY_MODE_SIGMA, c0=c0, c0_dev=c0_dev, packed=packed)

    def energy(self, t, out=None, packed=False):
        out = self.energy_dev if out is None else out
        ws = workspace(self.red_bytes, self.device)
        fn = self.lib.apy_cisd_so_energy_packed if packed else self.lib.apy_cisd_so_energy
        _lib.check(fn(self.O, self.V, int(self.singles), _vp(self.fock), _vp(self.ints["oovv"]), _vp(t), _vp(out),
                      _vp(ws), self.red_bytes, stream_of(self.device)), "apy_cisd_so_energy")
        return out

    def dots(self, xs, y, out, conj=True):
        k = len(xs)
        arr = (ctypes.c_void_p * k)(*[x.data_ptr() for x in xs])
        ws = workspace(self.red_bytes, self.device)
        _lib.check(self.lib.apy_zdot_many(k, arr, _vp(y), y.numel(), int(conj), _vp(out), _vp(ws), self.red_bytes,
                                          stream_of(self.device)), "apy_zdot_many")

    def combine(self, xs, coef, y):
        k = len(xs)
        arr = (ctypes.c_void_p * k)(*[x.data_ptr() for x in xs])
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
        _lib.check(self.lib.apy_zcombine(k, arr, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), y.numel(),
                                         _vp(y), stream_of(self.device)), "apy_zcombine")

    def diffsq(self, x, y, out):
        off = (ctypes.c_int64 * 3)(0, self.OV, self.L)
        ws = workspace(self.red_bytes, self.device)
        _lib.check(self.lib.apy_zdiffsq_segments(_vp(x), _vp(y), 2, off, _vp(out), _vp(ws), self.red_bytes,
                                                 stream_of(self.device)), "apy_zdiffsq_segments")

    # ---------------------------------------------------------------- solver
    def initial_guess(self, t):
        """t1 = F_vo^T / D_ia, t2 = <ab||ij> / D_ijab (ci_wfn.py:287, 290): one
        residual update from zero amplitudes with E = 0."""
        zero = self.new_vector()
        scratch = self.new_vector()
        self.step(zero, t, scratch, energy=self.zero_energy)
        if not self.singles:
            t[:self.OV].zero_()
        return t

    def iterations(self, diis=True, max_diis=7):
        """Generator over the reference's iterations (ci_wfn.py:302-378): each
        step is the Jacobi update (:336-337), the DIIS extrapolation
        (:340-352, utils.py:104-140), the energy (:355) and the un-conjugated
        rms of the change (:361-365).  Yields (it, E, dE, rms1, rms2) with E
        complex; the current amplitudes are self.t.
        """
        dev = self.device
        nvec = max_diis + 1
        t_cur = self.initial_guess(self.new_vector())
        t_nxt = self.new_vector()
        self.energy(t_cur)
        E = complex(self.energy_dev.cpu()[0])
        self.t = t_cur
        yield 0, E, 0j, 0j, 0j
        if diis:
            Eh = torch.zeros((nvec, self.L), dtype=CDT, device=dev)
            Th = torch.zeros((nvec, self.L), dtype=CDT, device=dev)
            Bfull = {}
            hist = []
        else:
            r = self.new_vector()
        it = 0
        while True:
            it += 1
            E_old = E
            if diis:
                slot = (it - 1) % nvec
                # drop the oldest while more than max_diis are stored (utils.py:109-112)
                while len(hist) > max_diis:
                    hist.pop(0)
                self.step(t_cur, Th[slot], Eh[slot])
                hist.append((it, slot))
                # new row of B = E^H E (utils.py:121), cached for older pairs
                self.dots([Eh[q] for _, q in hist], Eh[slot], self.scal)
                row = self.scal[:len(hist)].cpu().numpy()
                for (itm, _), b in zip(hist, row):
                    Bfull[(itm, it)] = b
                    Bfull[(it, itm)] = np.conj(b)
                for key in [k for k in Bfull if k[0] < hist[0][0] or k[1] < hist[0][0]]:
                    del Bfull[key]
                nv = len(hist)
                B = np.zeros((nv + 1, nv + 1), dtype=np.complex128)
                for a_, (ia, _) in enumerate(hist):
                    for b_, (ib, _) in enumerate(hist):
                        B[a_, b_] = Bfull[(ia, ib)]
                B[-1, :] = -1
                B[:, -1] = -1
                B[-1, -1] = 0
                A = np.zeros(nv + 1)
                A[-1] = -1
                c = np.linalg.solve(B, A)                                       # utils.py:135
                self.combine([Th[q] for _, q in hist], c[:-1], t_nxt)           # utils.py:138
            else:
                self.step(t_cur, t_nxt, r)
            self.energy(t_nxt)
            self.diffsq(t_cur, t_nxt, self.scal[1:3])
            self.scal[0:1].copy_(self.energy_dev)
            host = self.scal[:3].cpu().numpy()
            E = complex(host[0])
            rms1 = np.sqrt(host[1])
            rms2 = np.sqrt(host[2])
            t_cur, t_nxt = t_nxt, t_cur
            self.t = t_cur
            yield it, E, E_old - E, rms1, rms2

    def solve(self, params, print_level=0, max_diis=7, e_offset=0.0, control="device", graph=None):
        """The reference's iteration (ci_wfn.py:263-378 for CISD, :171-257
        for CID) until its convergence test passes or max_iterations.
        control="device" (with DIIS) keeps the DIIS solve and the
        convergence test on the device (solve_device); "host" reads the
        scalars back every iteration (solve_host).  Returns (E, t1, t2,
        iterations performed), E a Python complex."""
        if control == "device" and params['DIIS'] and max_diis + 1 <= _lib.APY_MAXVEC:
            return self.solve_device(params, print_level, max_diis, e_offset, graph=graph)
        return self.solve_host(params, print_level, max_diis, e_offset)

    def _print_header(self):
        print("\n Iter      E_elec(real)       E_elec(imaginary)        E(tot)           Delta_E(real)       "
              "Delta_E(imaginary)      RMS_T1(real)      RMS_T1(imaginary)      RMS_T2(real)      RMS_T2(imaginary)")

    @staticmethod
    def _print_row(it, E, dE, rms1, rms2, e_offset):
        print(" %02d %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f" % (
            it, E.real, E.imag, (E + e_offset).real, dE.real, dE.imag, rms1.real, rms1.imag, rms2.real, rms2.imag))

    # heavier iterations than this (seconds, measured on the first one) read the control block after each
    # iteration instead of queueing a chunk: queued iterations after convergence would cost real time
    CHUNK_MAX_ITERATION_S = 5.0e-4

    def solve_device(self, params, print_level=0, max_diis=7, e_offset=0.0, graph=None):
        """solve() with the DIIS extrapolation (apy_diis_slot), the guarded
        combination (apy_zcombine_guarded) and the convergence test
        (apy_iter_check) on the device: the (amplitude, residual) pairs live
        in a ring of max_diis + 1 slots, so each iteration's launches depend
        only on its position modulo lcm(2, slots).  Light problems queue that
        many iterations per host read (iterations queued past convergence do
        nothing that is kept); with graph=True they replay the steady-state
        chunk from a HIP graph captured in this call, which pays off only for
        long solves (at 15-35 iterations the capture costs more than the
        launches it saves, so it is off by default).  Heavy problems read the
        control block after every iteration.  The iterates are the
        reference's: same window, same solve, same test."""
        dev, lib = self.device, self.lib
        nvec = max_diis + 1
        max_it = int(params['max_iterations'])
        if print_level > 0:
            self._print_header()
        bufs = [self.initial_guess(self.new_vector()), self.new_vector()]
        self.energy(bufs[0])
        E0 = complex(self.energy_dev[0].cpu())
        if max_it < 1:
            self.t = bufs[0]
            t1, t2 = self.split(self.t)
            return E0, t1, t2, 0
        Eh = torch.zeros((nvec, self.L), dtype=CDT, device=dev)
        Th = torch.zeros((nvec, self.L), dtype=CDT, device=dev)
        B = torch.zeros((nvec, nvec), dtype=CDT, device=dev)
        coef = torch.zeros(nvec, dtype=CDT, device=dev)
        row = torch.zeros(nvec, dtype=CDT, device=dev)
        sums = torch.zeros(2, dtype=CDT, device=dev)
        ctl = torch.zeros(2, dtype=torch.int32, device=dev)
        E_prev = self.energy_dev[:1].clone()
        cap = max_it + 1
        log = torch.zeros((cap, 4), dtype=CDT, device=dev)
        e_conv, d_conv = float(params['e_convergence']), float(params['d_convergence'])
        th_ptrs = (ctypes.c_void_p * nvec)(*[Th[q].data_ptr() for q in range(nvec)])
        eh_list = [Eh[q] for q in range(nvec)]
        full = (1 << nvec) - 1

        def iteration(it):
            slot = (it - 1) % nvec
            t_cur, t_nxt = bufs[(it - 1) % 2], bufs[it % 2]
            valid = full if it >= nvec else (1 << it) - 1
            st = stream_of(dev)
            self.step(t_cur, Th[slot], Eh[slot])
            self.dots(eh_list, Eh[slot], row)                                    # utils.py:121
            _lib.check(lib.apy_diis_slot(nvec, slot, valid, _vp(row), _vp(B), _vp(coef), _vp(ctl), st),
                       "apy_diis_slot")                                        # utils.py:122-135
            _lib.check(lib.apy_zcombine_guarded(nvec, th_ptrs, _vp(coef), self.L, _vp(t_nxt), _vp(ctl), st),
                       "apy_zcombine_guarded")                                 # utils.py:138
            self.energy(t_nxt)
            self.diffsq(t_cur, t_nxt, sums)
            _lib.check(lib.apy_iter_check(_vp(self.energy_dev), _vp(sums), int(self.singles), e_conv, d_conv,
                                          max_it, _vp(E_prev), _vp(log), cap, _vp(ctl), st), "apy_iter_check")

        period = 2 * nvec // math.gcd(2, nvec)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        iteration(1)
        state = ctl.cpu().numpy()
        light = time.perf_counter() - t0 < self.CHUNK_MAX_ITERATION_S
        use_graph = bool(graph) and light
        queued, printed, g = 1, 0, None
        keep = workspace(0, dev)        # the scratch buffer a captured graph points into stays alive
        while True:
            done, status = int(state[0]), int(state[1])
            if print_level > 0 and done > printed:
                rows = log[printed + 1:done + 1].cpu().numpy()
                for k, (E, dE, r1, r2) in enumerate(rows):
                    self._print_row(printed + 1 + k, complex(E), complex(dE), complex(r1), complex(r2), e_offset)
                printed = done
            if status != 0:
                break
            if not light:
                iteration(queued + 1)
                queued += 1
            elif use_graph and queued % period == 0 and queued >= period:
                if g is None:
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g):
                        for q in range(period):
                            iteration(queued + 1 + q)
                g.replay()
                queued += period
            else:
                n = period - queued % period
                for q in range(n):
                    iteration(queued + 1 + q)
                queued += n
            state = ctl.cpu().numpy()
        del keep
        it = int(state[0])
        if int(state[1]) == 3:
            raise np.linalg.LinAlgError("singular DIIS system")
        E, dE, r1, r2 = (complex(x) for x in log[it].cpu().numpy())
        if int(state[1]) == 2 and (abs(dE) > params['e_convergence'] or r2 > params['d_convergence']):
            print("Not converged.")
        self.t = bufs[it % 2]
        self.energy(self.t)                 # queued iterations past the stop left another energy here
        t1, t2 = self.split(self.t)
        return E, t1, t2, it

    def solve_host(self, params, print_level=0, max_diis=7, e_offset=0.0):
        """solve() with the scalars of every iteration read back and the
        DIIS system solved on the host (also the form without DIIS)."""
        if print_level > 0:
            self._print_header()
        gen = self.iterations(diis=bool(params['DIIS']), max_diis=max_diis)
        _, E, _, _, _ = next(gen)
        it = 0
        dE = rms2 = 0j
        if params['max_iterations'] >= 1:
            for it, E, dE, rms1, rms2 in gen:
                if print_level > 0:
                    self._print_row(it, E, dE, rms1, rms2, e_offset)
                checks = [rms1, rms2] if self.singles else [rms2]
                if _converged(it, dE, checks, params):
                    break
                if it == params['max_iterations']:
                    if abs(dE) > params['e_convergence'] or rms2 > params['d_convergence']:
                        print("Not converged.")
                    break
        gen.close()
        t1, t2 = self.split(self.t)
        return E, t1, t2, it


def step_flops(O, V, singles=True, fold_fvv=True):
    """Real flops of the GEMMs of one apy_cisd_so_step (8 per complex
    multiply-add), per GEMM; 'ppl' is the dominant kernel.  With the F_vv
    terms folded into the ring block (fold_fvv) they cost no separate
    product.  The T1 -> T2 term through <kb||ij> (2 O flops per packed
    doubles element) is computed in the epilogue and not counted.  In the
    three-product form (SOCI gauss3m) 'ppl' and 'ring' execute 6 flops per
    complex multiply-add: see executed_flops."""
    Po, Pv, OV = npair(O), npair(V), O * V
    f = {
        'ppl': 8.0 * Po * Pv * Pv,
        'hhl': 8.0 * Po * Pv * Po,
        'ring': 8.0 * OV * OV * OV,
    }
    if not fold_fvv:
        f['fvv'] = 8.0 * O * O * V * V * V
    if singles:
        f.update({
            'q': 8.0 * Pv * O * O * V,
            's1_fvv': 8.0 * O * V * V,
            's1_foo': 8.0 * O * V * O,
            's1_ring': 8.0 * OV * OV,
            's1_fov': 8.0 * O * OV * V,
            's1_vovv': 8.0 * O * V * O * Pv,
            's1_ooov': 8.0 * O * O * O * V * V,
        })
    return f


def executed_flops(f, gauss3m=False):
    """Matrix-core flops one step executes for the per-GEMM counts f of
    step_flops: the PPL and ring products run 3 real products instead of 4
    in the three-product form."""
    tot = sum(f.values())
    if gauss3m:
        tot -= 0.25 * (f['ppl'] + f['ring'])
    return tot


def dense_flops(O, V):
    """The dense operation count of one sigma with every P(ij)/P(ab) image
    counted once: 8 [O^2V^4 + O^4V^2 + O^3V^3 + 3 O^2V^3 + 3 O^3V^2 + 3 O^2V^2]."""
    return 8.0 * (O**2 * V**4 + O**4 * V**2 + O**3 * V**3 + 3 * O**2 * V**3 + 3 * O**3 * V**2 + 3 * O**2 * V**2)
