"""Lowest root of the Hermitian spin-orbital CISD matrix by Davidson
iteration, and its energy gradient with respect to F, on the GPU.

The operator is the reference's CISD (ci_wfn.py:309-334) with the
P(ij)P(ab) f_bj t_ia coupling included (SOCI(fov_coupling=True)), so the
matrix is Hermitian and the lowest root variational; with canonical
Hartree-Fock orbitals (F_ov = 0) it is the reference's solution.  The
reference's Jacobi + DIIS iteration diverges on systems far from
diagonal dominance; Davidson converges there.

Subspace vectors live on the device as packed [c1 | c2 over (i<j, a<b) |
c0] (dense [c1 | c2 | c0] for the batched solver); inner products count each
determinant once.  Every vector operation is a library call: sigma
(apy_cisd_so_step, SIGMA mode, on the dense image from apy_ci_unpack, packed
back with apy_ci_pack), the reference row (apy_cisd_so_energy), dot
products (apy_zdot_many on packed vectors, apy_ci_dots with the c2 weight on
dense ones), linear combinations (apy_zcombine, apy_zcombine_seg) and the
preconditioner (apy_ci_precondition[_packed]).  The k x k
projected eigenproblem is solved on the host.

The gradient (adjoint) needs no iteration history: for the variational
root, dE/dF = D, the one-particle density of the normalised vector
(apy_cisd_so_opd), and dE/d<pq||rs> = G, its two-particle density
(apy_cisd_so_tpd, block layouts; apy_so_ints_grad_to_full for the whole
tensor); the left eigenvector is the conjugate of the right one.
`CISDEnergy` / `CISDEnergyFull` wrap E(F) and E(F, A) with this backward
in torch.autograd.Functions.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .kernels import stream_of, workspace

CDT = torch.complex128


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class CISDDavidson:
    """packed=True keeps the subspace as packed vectors [c1 | c2 over (i<j, a<b) | c0]
    whose sigma apy_cisd_so_step forms directly in that layout: the CI metric is
    then the Euclidean product and every vector pass moves a quarter of the bytes.
    packed=False keeps dense vectors (the layout MultiDavidson's batched sigma
    reads)."""

    def __init__(self, soci, max_space=16, packed=True, restart="sigma"):
        if restart not in ("sigma", "combine"):
            raise ValueError("restart must be 'sigma' or 'combine'")
        if not soci.fov or not soci.singles:
            raise ValueError("Davidson needs the Hermitian CISD operator: SOCI(..., singles=True, fov_coupling=True)")
        self.soci = soci
        self.lib = soci.lib
        self.dev = soci.device
        self.O, self.V = soci.O, soci.V
        self.OV, self.L = soci.OV, soci.L
        self.max_space = int(max_space)
        # at a restart the Ritz vector's sigma is a fresh product ("sigma") or sum_j y_j S_j of the stored
        # products ("combine": sigma is linear, one pass over the k products instead of a sigma)
        self.restart = restart
        if not 2 <= self.max_space <= _lib.APY_MAXVEC:
            raise ValueError(f"max_space must be in [2, {_lib.APY_MAXVEC}]")
        self.packed = bool(packed) and self.O >= 2 and self.V >= 2
        self.Lv = int(self.lib.apy_ci_packed_length(self.O, self.V)) if self.packed else self.L + 1
        self.B = torch.zeros((self.max_space, self.Lv), dtype=CDT, device=self.dev)
        self.S = torch.zeros((self.max_space, self.Lv), dtype=CDT, device=self.dev)
        self.work = torch.zeros(self.Lv, dtype=CDT, device=self.dev)
        self.part = torch.zeros(max(_lib.APY_MAXVEC, self.max_space + 1), dtype=CDT, device=self.dev)
        self.c = None
        self.E = None
        self.iterations = 0
        self.nsigma = 0

    # -------------------------------------------------------------- primitives
    def _sigma(self, q):
        """S[q] = H_N B[q]; the reference weight is read by the kernels from
        device memory."""
        self.nsigma += 1
        b, s = self.B[q], self.S[q]
        n = self.Lv - 1
        self.soci.sigma(b[:n], s[:n], c0_dev=b[n:].data_ptr(), packed=self.packed)
        self.soci.energy(b[:n], out=s[n:], packed=self.packed)

    def _dots_dev(self, xs, y, out):
        """CI-metric <x_m|y> into the device array out (len(xs) <= APY_MAXVEC), no host copy:
        apy_zdot_many on packed vectors (the metric is Euclidean there), apy_ci_dots on dense."""
        arr = (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])
        nb = int(self.lib.apy_reduce_workspace(0, len(xs)))
        ws = workspace(nb, self.dev)
        if self.packed:
            _lib.check(self.lib.apy_zdot_many(len(xs), arr, _vp(y), y.numel(), 1, _vp(out), _vp(ws), nb,
                                              stream_of(self.dev)), "apy_zdot_many")
        else:
            _lib.check(self.lib.apy_ci_dots(self.O, self.V, len(xs), arr, _vp(y), _vp(out), _vp(ws), nb,
                                            stream_of(self.dev)), "apy_ci_dots")

    def _precondition(self, theta, r, d):
        fn = self.lib.apy_ci_precondition_packed if self.packed else self.lib.apy_ci_precondition
        _lib.check(fn(self.O, self.V, _vp(self.soci.eps), theta, _vp(r), _vp(d), stream_of(self.dev)),
                   "apy_ci_precondition")

    def _gs_pass(self, basis, d, normalize):
        """d <- d - sum_j <B_j|d> B_j (normalize: then divided by its norm), all on the
        device: dots, apy_gs_coefficients, apy_zcombine_seg."""
        k = len(basis)
        self._dots_dev(basis + [d], d, self._gdots)
        _lib.check(self.lib.apy_gs_coefficients(k, _vp(self._gdots), int(normalize), _vp(self._gcoef),
                                                _vp(self._ibuf[1:2]), stream_of(self.dev)), "apy_gs_coefficients")
        xs = basis + [d]
        arr = (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])
        _lib.check(self.lib.apy_zcombine_seg(k + 1, arr, _vp(self._gcoef), 1, d.numel(), _vp(d),
                                             stream_of(self.dev)), "apy_zcombine_seg")

    def dots(self, xs, y):
        """CI-metric <x_m|y> for m < len(xs) (host complex array): one
        pass per APY_MAXVEC vectors, one host copy."""
        out = np.zeros(len(xs), dtype=np.complex128)
        M = _lib.APY_MAXVEC
        for s0 in range(0, len(xs), M):
            self._dots_dev(xs[s0:s0 + M], y, self.part[s0:])
        n = len(xs)
        out[:] = self.part[:n].cpu().numpy()
        return out

    def combine(self, xs, coef, y):
        """y = sum c_m x_m.  y may be one of the x_m when len(xs) fits one
        call (the kernel reads every input at an index before writing it)."""
        M = _lib.APY_MAXVEC
        if len(xs) > M and any(x.data_ptr() == y.data_ptr() for x in xs):
            raise ValueError("combine: output aliases an input of a multi-call combination")
        bounds = [(0, min(M, len(xs)))]
        while bounds[-1][1] < len(xs):               # later calls accumulate through y: M - 1 inputs each
            lo = bounds[-1][1]
            bounds.append((lo, min(lo + M - 1, len(xs))))
        for lo, hi in bounds:
            chunk = list(xs[lo:hi])
            c = list(coef[lo:hi])
            if lo > 0:
                chunk.append(y)
                c.append(1.0)
            arr = (ctypes.c_void_p * len(chunk))(*[x.data_ptr() for x in chunk])
            cc = np.ascontiguousarray(np.asarray(c, dtype=np.complex128)).view(np.float64)
            _lib.check(self.lib.apy_zcombine(len(chunk), arr, cc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             y.numel(), _vp(y), stream_of(self.dev)), "apy_zcombine")

    # ------------------------------------------------------------------ solve
    def iterate(self):
        """Generator over Davidson iterations.  Each step forms the Ritz pair
        of the current subspace, its residual, and extends the subspace by
        the preconditioned, twice-orthogonalised, normalised residual and its
        sigma, all on the device; one host read per iteration then brings
        back the residual norm and the new column of the projected matrix,
        and the step yields (Ritz value, residual norm) of the pair it
        started from.  finish() forms that pair's vector."""
        self._device_mode = False
        B, S = self.B, self.S
        B.zero_()
        S.zero_()
        B[0, self.Lv - 1] = 1.0                     # the reference determinant
        self._sigma(0)
        k = 1
        G = np.zeros((self.max_space, self.max_space), dtype=np.complex128)
        G[0, 0] = self.dots([B[0]], S[0])[0]
        self.x = x = torch.zeros(self.Lv, dtype=CDT, device=self.dev)
        r = self.work
        M = _lib.APY_MAXVEC
        # device scratch: [0] <r|r>, [1] Gram-Schmidt status (nn, <d|d>), [2] restart <B0|S0>,
        # [4:4+k+1] the new column; Gram-Schmidt dots and coefficients
        self._ibuf = torch.zeros(4 + M, dtype=CDT, device=self.dev)
        self._gdots = torch.zeros(M, dtype=CDT, device=self.dev)
        self._gcoef = torch.zeros(M, dtype=CDT, device=self.dev)
        while True:
            Gk = G[:k, :k]
            w, U = np.linalg.eigh(0.5 * (Gk + Gk.conj().T))
            theta, y = float(w[0]), U[:, 0]
            # residual straight from the basis, r = sum y_j S_j - theta sum y_j B_j
            self.combine([S[j] for j in range(k)] + [B[j] for j in range(k)], list(y) + list(-theta * y), r)
            self._dots_dev([r], r, self._ibuf[0:1])
            self._ritz = (k, y)
            self.E = theta
            restart = k == self.max_space
            if restart:                             # restart from the Ritz vector
                self._form_ritz()
                B[0].copy_(x)
                if self.restart == "combine":
                    self.combine([S[j] for j in range(k)], y, S[0])
                else:
                    self._sigma(0)
                self._ritz = (1, np.ones(1))
                self._dots_dev([B[0]], S[0], self._ibuf[2:3])
                k = 1
            d = B[k]
            self._precondition(theta, r, d)
            basis = [B[j] for j in range(k)]
            self._gs_pass(basis, d, False)          # Gram-Schmidt against the basis twice, the second
            self._gs_pass(basis, d, True)           # pass normalising by the Pythagoras norm
            self._sigma(k)
            self._dots_dev([B[j] for j in range(k + 1)], S[k], self._ibuf[4:5 + k])
            host = self._ibuf[:5 + k].cpu().numpy()          # the one host read of the iteration
            if not host[1].real > 0.0:
                raise RuntimeError("Davidson: the preconditioned residual lies in the subspace")
            if restart:
                G[0, 0] = host[2]
            rn = float(np.sqrt(abs(host[0])))
            col = host[4:5 + k]
            yield theta, rn
            G[:k + 1, k] = col
            G[k, :k + 1] = np.conj(col)
            k += 1

    # ------------------------------------------------- device-controlled iteration
    def _comb_dev(self, xs, coef, y):
        """y = sum_m coef[m] x_m with the coefficients in device memory (len(xs) <= APY_MAXVEC)."""
        arr = (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])
        _lib.check(self.lib.apy_zcombine_seg(len(xs), arr, _vp(coef), 1, y.numel(), _vp(y), stream_of(self.dev)),
                   "apy_zcombine_seg")

    def _device_init(self, e_conv, r_conv, cap):
        if not self.packed:
            raise ValueError("device-controlled Davidson runs on packed vectors")
        if self.max_space > _lib.APY_MAXVEC - 1:
            raise ValueError(f"device-controlled Davidson needs max_space <= {_lib.APY_MAXVEC - 1}")
        M = self.max_space
        self.G_dev = torch.zeros((M, M), dtype=CDT, device=self.dev)      # row j: column j of the projected matrix
        self.eig = torch.zeros(2 * _lib.APY_MAXVEC + 2, dtype=CDT, device=self.dev)
        self.rr = torch.zeros(1, dtype=CDT, device=self.dev)
        self.ctl = torch.tensor([0.0, float("nan"), 0.0, 0.0], dtype=torch.float64, device=self.dev)
        self.cap = int(cap)
        self.log = torch.zeros(2 * max(self.cap, 1), dtype=torch.float64, device=self.dev)
        self.one = torch.ones(1, dtype=CDT, device=self.dev)
        self._conv = (float(e_conv), float(r_conv))
        M2 = _lib.APY_MAXVEC
        self._ibuf = torch.zeros(4 + M2, dtype=CDT, device=self.dev)
        self._gdots = torch.zeros(M2, dtype=CDT, device=self.dev)
        self._gcoef = torch.zeros(M2, dtype=CDT, device=self.dev)
        self.x = torch.zeros(self.Lv, dtype=CDT, device=self.dev)
        self._graphs = {}
        self._fws_bytes = int(self.lib.apy_davidson_fused_workspace())
        self._fws = torch.empty(self._fws_bytes, dtype=torch.uint8, device=self.dev)
        self._fdots = torch.zeros(M2 + 2, dtype=CDT, device=self.dev)

    def _device_iteration(self, k):
        """One Davidson iteration at subspace size k with no host read: the
        Ritz pair from apy_heev_lowest, residual, convergence test, restart
        (k == max_space), preconditioner with theta from device memory,
        Gram-Schmidt, sigma and the new column.  Returns the new size.
        Below max_space (fused=True) the residual, preconditioner and first
        Gram-Schmidt dots are one pass over the subspace (apy_davidson_residual)
        and the first Gram-Schmidt combination carries the second pass's dots
        (apy_davidson_combine_dots)."""
        B, S, r = self.B, self.S, self.work
        st = stream_of(self.dev)
        M, MV = self.max_space, _lib.APY_MAXVEC
        _lib.check(self.lib.apy_heev_lowest(k, _vp(self.G_dev), M, _vp(self.eig), st), "apy_heev_lowest")
        if self.fused and k < M:
            d = B[k]
            sl = (ctypes.c_void_p * k)(*[S[j].data_ptr() for j in range(k)])
            bl = (ctypes.c_void_p * k)(*[B[j].data_ptr() for j in range(k)])
            fd, ws, nb = self._fdots, self._fws, self._fws_bytes
            _lib.check(self.lib.apy_davidson_residual(k, sl, bl, _vp(self.eig), _vp(self.soci.eps), self.O, self.V,
                                                      _vp(d), _vp(fd), _vp(ws), nb, st), "apy_davidson_residual")
            e_conv, r_conv = self._conv
            _lib.check(self.lib.apy_davidson_check(_vp(self.eig), _vp(fd[k + 1:k + 2]), e_conv, r_conv,
                                                   _vp(self.ctl), _vp(self.log), self.cap, st), "apy_davidson_check")
            self._ritz_dev = (k, self.eig[1:])
            _lib.check(self.lib.apy_gs_coefficients(k, _vp(fd), 0, _vp(self._gcoef), None, st), "apy_gs_coefficients")
            _lib.check(self.lib.apy_davidson_combine_dots(k, bl, _vp(self._gcoef), d.numel(), _vp(d), _vp(fd), _vp(ws),
                                                          nb, st), "apy_davidson_combine_dots")
            _lib.check(self.lib.apy_gs_coefficients(k, _vp(fd), 1, _vp(self._gcoef), _vp(self._ibuf[1:2]), st),
                       "apy_gs_coefficients")
            self._comb_dev([B[j] for j in range(k)] + [d], self._gcoef, d)
            self._sigma(k)
            self._dots_dev([B[j] for j in range(k + 1)], S[k], self.G_dev[k][0:k + 1])
            return k + 1
        self._comb_dev([S[j] for j in range(k)], self.eig[1:], r)
        self._comb_dev([B[j] for j in range(k)] + [r], self.eig[1 + MV:], r)
        self._dots_dev([r], r, self.rr)
        e_conv, r_conv = self._conv
        _lib.check(self.lib.apy_davidson_check(_vp(self.eig), _vp(self.rr), e_conv, r_conv, _vp(self.ctl),
                                               _vp(self.log), self.cap, st), "apy_davidson_check")
        if k == M:                                  # restart from the Ritz vector
            self._comb_dev([B[j] for j in range(k)], self.eig[1:], B[0])
            if self.restart == "combine":
                self._comb_dev([S[j] for j in range(k)], self.eig[1:], S[0])
            else:
                self._sigma(0)
            self._dots_dev([B[0]], S[0], self.G_dev[0][0:1])
            k = 1
            self._ritz_dev = (1, self.one)
        else:
            self._ritz_dev = (k, self.eig[1:])
        d = B[k]
        _lib.check(self.lib.apy_ci_precondition_packed_dev(self.O, self.V, _vp(self.soci.eps), _vp(self.eig), _vp(r),
                                                           _vp(d), st), "apy_ci_precondition_packed_dev")
        basis = [B[j] for j in range(k)]
        self._gs_pass(basis, d, False)
        self._gs_pass(basis, d, True)
        self._sigma(k)
        self._dots_dev([B[j] for j in range(k + 1)], S[k], self.G_dev[k][0:k + 1])
        return k + 1

    def prepare_graphs(self):
        """Capture one HIP graph per subspace size (1 .. max_space): each holds
        the whole device-controlled iteration at that size, so a solve
        replays a fixed sequence of graphs.  Capture records without running.
        The graphs point into the per-device scratch buffer that is current
        now (already sized by the sigma of the first vector); this object
        keeps a reference to it, so a later, larger request elsewhere (which
        replaces the cached buffer) cannot free memory the graphs write."""
        self._graph_ws = workspace(max(self.soci.ws_bytes, self.soci.red_bytes), self.dev)
        for k in range(1, self.max_space + 1):
            if k in self._graphs:
                continue
            torch.cuda.synchronize(self.dev)
            g = torch.cuda.CUDAGraph()
            n0 = self.nsigma
            with torch.cuda.graph(g):
                nk = self._device_iteration(k)
            if workspace(0, self.dev) is not self._graph_ws:
                raise RuntimeError("the scratch buffer grew during graph capture")
            ns, self.nsigma = self.nsigma - n0, n0          # sigmas per replay; capture ran none
            self._graphs[k] = (g, nk, self._ritz_dev, ns)

    def iterate_device(self, e_conv=0.0, r_conv=0.0, graph=False, cap=0, fused=True):
        """Generator over device-controlled Davidson iterations: the same
        iteration as iterate() with the projected eigenproblem (Jacobi,
        apy_heev_lowest), the residual coefficients and the convergence test
        (apy_davidson_check) on the GPU, so no iteration waits for the host;
        graph=True replays one captured HIP graph per subspace size.  Yields
        nothing per step (read_state() brings back the Ritz value, residual
        norm and status); finish() forms the vector of the last Ritz pair."""
        self._device_init(e_conv, r_conv, cap)
        self._device_mode = True
        self.fused = bool(fused)
        B, S = self.B, self.S
        B.zero_()
        S.zero_()
        self.G_dev.zero_()
        B[0, self.Lv - 1] = 1.0                     # the reference determinant
        self._sigma(0)
        self._dots_dev([B[0]], S[0], self.G_dev[0][0:1])
        if graph:
            self.prepare_graphs()
        k = 1
        while True:
            if graph:
                g, nk, ritz, ns = self._graphs[k]
                g.replay()
                k, self._ritz_dev = nk, ritz
                self.nsigma += ns
            else:
                k = self._device_iteration(k)
            yield None

    def read_state(self):
        """(iterations, Ritz value, residual norm, converged) of the device
        iteration: one host read."""
        c = self.ctl.cpu().numpy()
        return int(c[0]), float(c[1]), float(c[3]), bool(c[2] == 1.0)

    def solve_device(self, e_conv=1e-12, r_conv=1e-8, maxit=200, chunk=4, graph=False, fused=True):
        """solve() with device-controlled iterations, checked once per chunk of
        iterations (one host read each); iterations after convergence inside
        a chunk only extend the subspace.  Returns the correlation energy."""
        gen = self.iterate_device(e_conv, r_conv, graph=graph, cap=maxit, fused=fused)
        it = 0
        while it < maxit:
            for _ in range(min(chunk, maxit - it)):
                next(gen)
                it += 1
            if self.read_state()[3]:
                break
        gen.close()
        n, theta, rn, _ = self.read_state()
        lg = self.log[:2 * min(n, self.cap)].cpu().numpy().reshape(-1, 2)
        self.history = [(float(a), float(b)) for a, b in lg]
        self.E = theta
        self.iterations = n
        self.finish()
        return theta

    def _form_ritz(self):
        """self.x = sum_j y_j B_j for the Ritz pair of the last iteration."""
        if getattr(self, "_device_mode", False):
            k, coef = self._ritz_dev
            self._comb_dev([self.B[j] for j in range(k)], coef, self.x)
            return
        k, y = self._ritz
        self.combine([self.B[j] for j in range(k)], y, self.x)

    def finish(self):
        """Form and normalise the current Ritz vector into self.c (dense
        [c1 | c2 | c0])."""
        self._form_ritz()
        x = self.x
        nx = float(np.sqrt(abs(self.dots([x], x)[0])))
        self.combine([x], [1.0 / nx], x)
        if self.packed:
            self.c = torch.empty(self.L + 1, dtype=CDT, device=self.dev)
            _lib.check(self.lib.apy_ci_unpack(self.O, self.V, _vp(x), _vp(self.c), stream_of(self.dev)),
                       "apy_ci_unpack")
        else:
            self.c = x
        return self.c

    def solve(self, e_conv=1e-12, r_conv=1e-8, maxit=200):
        """Lowest root: stops when the Ritz value changes by less than e_conv
        and the residual norm is below r_conv (the residual reaches a floor
        near 1e-10 set by rounding in sigma; r_conv must stay above it).
        Returns the correlation energy (real); self.c holds the normalised
        vector [c1 | c2 | c0]; self.history the (Ritz value, residual norm)
        of every iteration."""
        self.history = []
        theta_old = None
        it = 0
        gen = self.iterate()
        for it, (theta, rn) in enumerate(gen, start=1):
            self.history.append((theta, rn))
            if (theta_old is not None and abs(theta - theta_old) < e_conv and rn < r_conv) or it >= maxit:
                break
            theta_old = theta
        gen.close()
        self.finish()
        self.iterations = it
        return theta

    def amplitudes(self):
        """(t1, t2) in intermediate normalisation (c / c0)."""
        c0 = complex(self.c[self.L].cpu())
        t = self.c[:self.L] / c0
        return t[:self.OV].view(self.O, self.V), t[self.OV:].view(self.O, self.O, self.V, self.V)

    def gradient_fock(self):
        """dE/dF [n, n] (apy_cisd_so_opd on the normalised root)."""
        n = self.O + self.V
        D = torch.empty((n, n), dtype=CDT, device=self.dev)
        nb = int(self.lib.apy_cisd_so_opd_workspace(self.O, self.V))
        ws = workspace(nb, self.dev)
        _lib.check(self.lib.apy_cisd_so_opd(self.O, self.V, _vp(self.c), _vp(D), _vp(ws), nb, stream_of(self.dev)),
                   "apy_cisd_so_opd")
        return D


    def gradient_integrals(self, blocks=None):
        """dE/d<pq||rs> in the block layouts of the device integrals
        (apy_cisd_so_tpd on the normalised root): an SOIntegrals holding,
        per stored entry, the derivative with respect to that entry, so
        that dE = sum over blocks of grad . d(block) for perturbations
        antisymmetric in (p,q) and (r,s)."""
        from .so_ci import ALL_BLOCKS, SOIntegrals
        G = SOIntegrals(self.O, self.V, self.dev, blocks or ALL_BLOCKS)
        nb = int(self.lib.apy_cisd_so_tpd_workspace(self.O, self.V))
        ws = workspace(nb, self.dev)
        _lib.check(self.lib.apy_cisd_so_tpd(self.O, self.V, _vp(self.c), ctypes.byref(G.struct), _vp(ws), nb,
                                            stream_of(self.dev)), "apy_cisd_so_tpd")
        return G


def full_gradient(G):
    """Whole-tensor dE/dA [n,n,n,n] (antisymmetric in each index pair) from
    the block gradients of CISDDavidson.gradient_integrals."""
    n = G.O + G.V
    out = torch.empty((n, n, n, n), dtype=CDT, device=G.device)
    _lib.check(_lib.load().apy_so_ints_grad_to_full(n, ctypes.byref(G.struct), _vp(out), stream_of(G.device)),
               "apy_so_ints_grad_to_full")
    return out


class CISDEnergy(torch.autograd.Function):
    """E(F) of the lowest Hermitian CISD root at fixed integrals, with the
    adjoint backward dE/dF = D.  E is real and F Hermitian; for a real
    loss the vector-Jacobian product is grad_output * conj(D)."""

    @staticmethod
    def forward(ctx, fock, ints, eps, e_conv=1e-12, r_conv=1e-9):
        from .so_ci import SOCI
        dav = CISDDavidson(SOCI(ints, fock.detach(), eps, singles=True, fov_coupling=True))
        E = dav.solve(e_conv, r_conv)
        ctx.dav = dav
        return torch.tensor(E, dtype=torch.float64, device=fock.device)

    @staticmethod
    def backward(ctx, gE):
        D = ctx.dav.gradient_fock()
        return gE * D.conj(), None, None, None, None


class CISDEnergyFull(torch.autograd.Function):
    """E(F, A) of the lowest Hermitian CISD root with A = <pq||rs> the whole
    antisymmetrised tensor [n,n,n,n] (device), with the adjoint backward
    dE/dF = D and dE/dA = G (two-particle density, antisymmetric in each
    index pair).  For a real loss the vector-Jacobian products are
    grad_output * conj(D) and grad_output * conj(G)."""

    @staticmethod
    def forward(ctx, fock, A, eps, O, e_conv=1e-12, r_conv=1e-9):
        from .so_ci import SOCI, SOIntegrals
        ints = SOIntegrals.from_full(A.detach(), int(O))
        dav = CISDDavidson(SOCI(ints, fock.detach(), eps, singles=True, fov_coupling=True))
        E = dav.solve(e_conv, r_conv)
        ctx.dav = dav
        return torch.tensor(E, dtype=torch.float64, device=fock.device)

    @staticmethod
    def backward(ctx, gE):
        D = ctx.dav.gradient_fock()
        G = full_gradient(ctx.dav.gradient_integrals())
        return gE * D.conj(), gE * G.conj(), None, None, None, None


def cisd_energy(fock, ints, eps, O=None):
    """Lowest Hermitian CISD root energy behind torch.autograd: `ints` is an
    SOIntegrals (gradient with respect to F only) or the whole tensor
    A = <pq||rs> [n,n,n,n] with O occupied spin orbitals (gradients with
    respect to F and A)."""
    if isinstance(ints, torch.Tensor):
        if O is None:
            raise ValueError("cisd_energy with a whole integral tensor needs O")
        return CISDEnergyFull.apply(fock, ints, eps, O)
    return CISDEnergy.apply(fock, ints, eps)


class MultiDavidson:
    """Lowest roots of several Hermitian CISD problems that share the
    integrals (the +/- field-displaced copies of one system) solved
    together: every iteration applies sigma to the new vectors of all
    unconverged instances in one apy_cisd_so_sigma_batch call, so the
    integral blocks stream through the matrix cores once per iteration for
    the whole set.  Each instance keeps its own subspace, Ritz pair and
    convergence test (those of CISDDavidson).  With one instance this is
    the per-displacement solve on the same kernels."""

    def __init__(self, ints, focks, eps, max_space=12):
        from .so_ci import SOCI
        self.ints = ints
        self.dav = [CISDDavidson(SOCI(ints, F, eps, singles=True, fov_coupling=True, fold_fvv=False), max_space,
                                 packed=False) for F in focks]
        self.lib = self.dav[0].lib
        self.dev = self.dav[0].dev
        self.O, self.V, self.L = self.dav[0].O, self.dav[0].V, self.dav[0].L
        self.ev_dominant = None

    def _sigma_batch(self, entries):
        """S[slot] = H_N B[slot] for every (instance, slot) in entries."""
        nb = len(entries)
        L = self.L
        c0 = torch.stack([d.B[q][L] for d, q in entries]).cpu().numpy()
        f = (ctypes.c_void_p * nb)(*[d.soci.fock.data_ptr() for d, _ in entries])
        h = (ctypes.c_void_p * nb)(*[d.soci.hoo.data_ptr() for d, _ in entries])
        t = (ctypes.c_void_p * nb)(*[d.B[q].data_ptr() for d, q in entries])
        o = (ctypes.c_void_p * nb)(*[d.S[q].data_ptr() for d, q in entries])
        cc = np.ascontiguousarray(c0.astype(np.complex128)).view(np.float64)
        ev = self.ev_dominant or (None, None)
        args = _lib.CisdSoSigmaBatchArgs(self.O, self.V, nb, 1, 1, ctypes.pointer(self.ints.struct), f, h,
                                         cc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), t, o, ev[0], ev[1])
        nbytes = int(self.lib.apy_cisd_so_sigma_batch_workspace(self.O, self.V, nb, 1))
        ws = workspace(nbytes, self.dev)
        _lib.check(self.lib.apy_cisd_so_sigma_batch(ctypes.byref(args), _vp(ws), nbytes, stream_of(self.dev)),
                   "apy_cisd_so_sigma_batch")
        for d, q in entries:
            d.soci.energy(d.B[q][:L], out=d.S[q][L:])

    def solve(self, e_conv=1e-12, r_conv=1e-8, maxit=200, guess=None):
        """Returns the array of correlation energies; each instance's
        normalised vector is in dav[k].c (iteration counts in dav[k].iterations).
        guess: optional start vectors [c1 | c2 | c0] per instance (e.g. the
        root of the undisplaced system for its displaced copies) in place of
        the reference determinant."""
        L = self.L
        st = []
        for k, d in enumerate(self.dav):
            d.B.zero_()
            d.S.zero_()
            if guess is not None and guess[k] is not None:
                d.B[0].copy_(guess[k])
                nrm = float(np.sqrt(abs(d.dots([d.B[0]], d.B[0])[0])))
                d.combine([d.B[0]], [1.0 / nrm], d.B[0])
            else:
                d.B[0, L] = 1.0
            d.history = []
            st.append({'k': 1, 'G': np.zeros((d.max_space, d.max_space), np.complex128), 'theta_old': None,
                       'x': torch.zeros(L + 1, dtype=CDT, device=self.dev), 'done': False})
        self._sigma_batch([(d, 0) for d in self.dav])
        for d, s in zip(self.dav, st):
            s['G'][0, 0] = d.dots([d.B[0]], d.S[0])[0]
        for it in range(1, maxit + 1):
            entries, cols = [], []
            for d, s in zip(self.dav, st):
                if s['done']:
                    continue
                k, G, x, r = s['k'], s['G'], s['x'], d.work
                Gk = G[:k, :k]
                w, U = np.linalg.eigh(0.5 * (Gk + Gk.conj().T))
                theta, y = float(w[0]), U[:, 0]
                d.combine([d.B[j] for j in range(k)], y, x)
                d.combine([d.S[j] for j in range(k)] + [x], list(y) + [-theta], r)
                rn = float(np.sqrt(abs(d.dots([r], r)[0])))
                d.history.append((theta, rn))
                d.iterations = it
                d.E = theta
                if s['theta_old'] is not None and abs(theta - s['theta_old']) < e_conv and rn < r_conv:
                    s['done'] = True
                    continue
                s['theta_old'] = theta
                if k == d.max_space:                         # restart from the Ritz vector
                    d.B[0].copy_(x)
                    k = 1
                    entries.append((d, 0))
                    s['restart'] = True
                dv = d.B[k]
                _lib.check(self.lib.apy_ci_precondition(self.O, self.V, _vp(d.soci.eps), theta, _vp(r), _vp(dv),
                                                        stream_of(self.dev)), "apy_ci_precondition")
                for _ in range(2):
                    c = d.dots([d.B[j] for j in range(k)], dv)
                    d.combine([d.B[j] for j in range(k)] + [dv], list(-c) + [1.0], dv)
                nd = float(np.sqrt(abs(d.dots([dv], dv)[0])))
                d.combine([dv], [1.0 / nd], dv)
                s['k'] = k
                entries.append((d, k))
                cols.append((d, s))
            if not entries:
                break
            self._sigma_batch(entries)
            for d, s in cols:
                k, G = s['k'], s['G']
                if s.pop('restart', False):
                    G[0, 0] = d.dots([d.B[0]], d.S[0])[0]
                col = d.dots([d.B[j] for j in range(k + 1)], d.S[k])
                G[:k + 1, k] = col
                G[k, :k + 1] = np.conj(col)
                s['k'] = k + 1
        for d, s in zip(self.dav, st):
            x = s['x']
            nx = float(np.sqrt(abs(d.dots([x], x)[0])))
            d.combine([x], [1.0 / nx], x)
            d.c = x
        return np.array([d.E for d in self.dav])


class FieldSetDavidson:
    """Lowest roots of a set of Hermitian CISD problems that differ only in a
    one-electron term, H_k = H(F0 + s_k dF_{d_k}) with shared <pq||rs> (the
    +/- field- or one-electron-displaced instances of fin_diff.py:335-370),
    solved in ONE common subspace.

    The sigma is affine in F, so H_k b = H(F0) b + s_k L(dF_d) b with L the
    Fock-linear part (apy_cisd_so_fock_sigma, O^2 V^3).  Every subspace vector
    b_j therefore carries one full sigma S_j = H(F0) b_j and one cheap
    product P_dj = L(dF_d) b_j per direction, and every instance's projected
    matrix is G0 + s_k G_d with G0 = B^H S, G_d = B^H P_d.  The root of F0 is
    solved first (CISDDavidson on packed vectors); the shared phase starts
    from it and, each iteration, adds the preconditioned residuals of the
    unconverged instances after Gram-Schmidt against the subspace and each
    other, dropping a candidate whose part outside the span is below `drop`
    of its norm: the residuals of a +/- pair are equal and opposite to first
    order in the step, so a pair adds one direction (one full sigma) where
    two separate solves would each need their own.  When the subspace is
    full it collapses onto the span of the instances' Ritz vectors (their
    sigmas and L-products by linear combination, no new sigma).  Each
    instance stops on the test of CISDDavidson.solve (Ritz value change below
    e_conv and residual norm below r_conv)."""

    def __init__(self, ints, F0, eps, directions, instances, max_space=16, base_space=12, drop=1e-2):
        from .so_ci import SOCI, npair
        self.soci = SOCI(ints, F0, eps, singles=True, fov_coupling=True)
        self.base = CISDDavidson(self.soci, base_space, packed=True)
        if not self.base.packed:
            raise ValueError("FieldSetDavidson needs O >= 2 and V >= 2 (packed vectors)")
        self.lib, self.dev = self.base.lib, self.base.dev
        self.O, self.V, self.Lv = self.base.O, self.base.V, self.base.Lv
        self.dirs = [torch.as_tensor(d, dtype=CDT, device=self.dev).contiguous() for d in directions]
        n = self.O + self.V
        for d in self.dirs:
            if d.shape != (n, n):
                raise ValueError(f"direction matrices must be [{n}][{n}]")
        self.inst = [(int(d), float(s)) for d, s in instances]
        if any(not 0 <= d < len(self.dirs) for d, _ in self.inst):
            raise ValueError("instance direction index out of range")
        self.M = int(max_space)
        if self.M < len(self.inst) + 2:
            raise ValueError("max_space must exceed the number of instances by at least 2")
        self.drop = float(drop)
        Po = npair(self.O)
        self.hoo_d = []
        for d in self.dirs:
            h = torch.zeros(Po * Po, dtype=CDT, device=self.dev)
            _lib.check(self.lib.apy_cisd_so_pair_operator(self.O, self.V, _vp(d), None, _vp(h), stream_of(self.dev)),
                       "apy_cisd_so_pair_operator")
            self.hoo_d.append(h)
        self.fs_bytes = int(self.lib.apy_cisd_so_fock_sigma_workspace(self.O, self.V))
        nd, K, M, Lv = len(self.dirs), len(self.inst), self.M, self.Lv
        # two copies of [B | S | P_0 .. P_nd-1], swapped at a collapse
        self._bufs = [torch.zeros((2 + nd, M, Lv), dtype=CDT, device=self.dev) for _ in range(2)]
        self._cur = 0
        self.R = torch.zeros((K, Lv), dtype=CDT, device=self.dev)        # residuals
        self.C = torch.zeros((K, Lv), dtype=CDT, device=self.dev)        # corrections (candidates)
        self.X = torch.zeros((K, Lv), dtype=CDT, device=self.dev)        # Ritz vectors (final)
        self._dbuf = torch.zeros(_lib.APY_MAXVEC, dtype=CDT, device=self.dev)
        self.nsigma = 0
        self.nfock = 0

    # -------------------------------------------------------------- primitives
    def _arrays(self, which=None):
        buf = self._bufs[self._cur if which is None else which]
        return buf[0], buf[1], [buf[2 + q] for q in range(len(self.dirs))]

    def _dots(self, xs, y):
        """<x_m|y> (packed vectors: Euclidean), host array; one read per 16 vectors."""
        out = np.zeros(len(xs), dtype=np.complex128)
        M = _lib.APY_MAXVEC
        for s0 in range(0, len(xs), M):
            self.base._dots_dev(xs[s0:s0 + M], y, self._dbuf)
            out[s0:s0 + M] = self._dbuf[:len(xs[s0:s0 + M])].cpu().numpy()
        return out

    def _subtract(self, xs, coef, y):
        """y -= sum_m coef_m x_m, in calls of at most APY_MAXVEC - 1 inputs plus y."""
        M = _lib.APY_MAXVEC - 1
        for s0 in range(0, len(xs), M):
            chunk = list(xs[s0:s0 + M])
            self.base.combine(chunk + [y], list(-np.asarray(coef[s0:s0 + M])) + [1.0], y)

    def _products(self, q):
        """S[q] = H(F0) B[q] (full sigma) and P_d[q] = L(dF_d) B[q] for every direction."""
        B, S, P = self._arrays()
        n = self.Lv - 1
        self.soci.sigma(B[q][:n], S[q][:n], c0_dev=B[q][n:].data_ptr(), packed=True)
        self.soci.energy(B[q][:n], out=S[q][n:], packed=True)
        self.nsigma += 1
        self._fock(q)

    def _fock(self, q):
        """P_d[q] = L(dF_d) B[q] for every direction (one apy_cisd_so_fock_sigma call)."""
        B, S, P = self._arrays()
        nd = len(self.dirs)
        f = (ctypes.c_void_p * nd)(*[d.data_ptr() for d in self.dirs])
        h = (ctypes.c_void_p * nd)(*[x.data_ptr() for x in self.hoo_d])
        o = (ctypes.c_void_p * nd)(*[Pd[q].data_ptr() for Pd in P])
        args = _lib.CisdSoFockSigmaArgs(self.O, self.V, nd, f, h, B[q].data_ptr(), o)
        ws = workspace(self.fs_bytes, self.dev)
        _lib.check(self.lib.apy_cisd_so_fock_sigma(ctypes.byref(args), _vp(ws), self.fs_bytes, stream_of(self.dev)),
                   "apy_cisd_so_fock_sigma")
        self.nfock += 1

    def _column(self, q):
        """Row/column q of G0 and every G_d from dots of the new vector."""
        B, S, P = self._arrays()
        bs = [B[j] for j in range(q + 1)]
        for G, Y in [(self.G0, S)] + list(zip(self.Gd, P)):
            col = self._dots(bs, Y[q])
            G[:q + 1, q] = col
            G[q, :q + 1] = np.conj(col)

    def _collapse(self, Ys):
        """Replace the m-vector subspace by the span of the coefficient vectors
        Ys (m x K), orthonormalised on the host; B, S, P follow by linear
        combination into the other buffer."""
        m = self.m
        U, sv, _ = np.linalg.svd(Ys, full_matrices=False)
        keep = sv > 1e-12 * sv[0]
        U = U[:, keep]
        src, dst = self._cur, 1 - self._cur
        Bs, Ss, Ps = self._arrays(src)
        Bd, Sd, Pd = self._arrays(dst)
        for c in range(U.shape[1]):
            coef = list(U[:, c])
            self.base.combine([Bs[j] for j in range(m)], coef, Bd[c])
            self.base.combine([Ss[j] for j in range(m)], coef, Sd[c])
            for Pq_s, Pq_d in zip(Ps, Pd):
                self.base.combine([Pq_s[j] for j in range(m)], coef, Pq_d[c])
        self._cur = dst
        Uh = U.conj().T
        k = U.shape[1]
        G0 = np.zeros_like(self.G0)
        G0[:k, :k] = Uh @ self.G0[:m, :m] @ U
        self.G0 = G0
        for q in range(len(self.Gd)):
            Gq = np.zeros_like(self.Gd[q])
            Gq[:k, :k] = Uh @ self.Gd[q][:m, :m] @ U
            self.Gd[q] = Gq
        self.m = k
        return U

    # ------------------------------------------------------------------ solve
    def solve(self, e_conv=1e-12, r_conv=1e-8, maxit=200, base_maxit=200, seed_subspace=False):
        """Returns the array of the instances' correlation energies; self.E0
        holds the root of F0, self.iterations the shared-phase iterations,
        self.history[k] the (Ritz value, residual norm) of instance k per
        iteration, and vector(k) its normalised packed vector.
        seed_subspace: start the shared phase from the base solve's whole
        final subspace (its sigmas are known; only the L-products are new)
        instead of its root alone."""
        self.E0 = self.base.solve(e_conv=e_conv, r_conv=r_conv, maxit=base_maxit)
        self.base_iterations = self.base.iterations
        self.nsigma = self.base.nsigma
        K, nd, M = len(self.inst), len(self.dirs), self.M
        self.G0 = np.zeros((M, M), np.complex128)
        self.Gd = [np.zeros((M, M), np.complex128) for _ in range(nd)]
        B, S, P = self._arrays()
        kb, yb = self.base._ritz
        if seed_subspace and kb + K <= M:
            # the base subspace (orthonormal, sigmas known); its span holds the root of F0
            for j in range(kb):
                B[j].copy_(self.base.B[j])
                S[j].copy_(self.base.S[j])
        else:
            # the root of F0 and its sigma (combination of the base subspace's sigmas)
            self.base.combine([self.base.B[j] for j in range(kb)], list(yb), B[0])
            self.base.combine([self.base.S[j] for j in range(kb)], list(yb), S[0])
            nx = float(np.sqrt(abs(self._dots([B[0]], B[0])[0])))
            self.base.combine([B[0]], [1.0 / nx], B[0])
            self.base.combine([S[0]], [1.0 / nx], S[0])
            kb = 1
        for j in range(kb):
            self._fock(j)
            self._column(j)
        self.m = kb
        done = [False] * K
        theta_old = [None] * K
        self.E = np.zeros(K)
        self.y = [None] * K
        self.history = [[] for _ in range(K)]
        self.iterations = 0
        for it in range(1, maxit + 1):
            self.iterations = it
            m = self.m
            B, S, P = self._arrays()
            cand = []
            for k, (d, s) in enumerate(self.inst):
                Gk = self.G0[:m, :m] + s * self.Gd[d][:m, :m]
                w, U = np.linalg.eigh(0.5 * (Gk + Gk.conj().T))
                theta, y = float(w[0]), U[:, 0]
                self.y[k] = y
                self.E[k] = theta
                if done[k]:
                    continue
                r = self.R[k]
                self.base.combine([S[j] for j in range(m)] + [P[d][j] for j in range(m)] + [B[j] for j in range(m)],
                                  list(y) + list(s * y) + list(-theta * y), r)
                rn = float(np.sqrt(abs(self._dots([r], r)[0])))
                self.history[k].append((theta, rn))
                if theta_old[k] is not None and abs(theta - theta_old[k]) < e_conv and rn < r_conv:
                    done[k] = True
                    continue
                theta_old[k] = theta
                self.base._precondition(theta, r, self.C[k])
                cand.append((rn, k))
            if not cand:
                break
            cand = [k for _, k in sorted(cand, reverse=True)]      # largest residuals first
            # Gram-Schmidt of the candidates against the subspace and each other
            accepted = []
            for k in cand:
                dv = self.C[k]
                n0 = float(np.sqrt(abs(self._dots([dv], dv)[0])))
                if not n0 > 0.0:
                    continue
                basis = [B[j] for j in range(m)] + [self.C[a] for a in accepted]
                for _ in range(2):
                    self._subtract(basis, self._dots(basis, dv), dv)
                n1 = float(np.sqrt(abs(self._dots([dv], dv)[0])))
                if n1 > self.drop * n0:
                    self.base.combine([dv], [1.0 / n1], dv)
                    accepted.append(k)
            if not accepted:
                raise RuntimeError("FieldSetDavidson: every correction lies in the subspace")
            if m + len(accepted) > M:
                Ys = np.stack(self.y, axis=1)
                self._collapse(Ys)
                m = self.m
                B, S, P = self._arrays()
                accepted = accepted[:M - m]      # the collapsed span holds at most K <= M - 2 vectors
            for a in accepted:
                B[m].copy_(self.C[a])
                self._products(m)
                self.m = m + 1
                self._column(m)
                m += 1
        self.converged = all(done)
        # final Ritz vectors
        B, S, P = self._arrays()
        for k in range(K):
            y = self.y[k]
            self.base.combine([B[j] for j in range(len(y))], list(y), self.X[k])
            nx = float(np.sqrt(abs(self._dots([self.X[k]], self.X[k])[0])))
            self.base.combine([self.X[k]], [1.0 / nx], self.X[k])
        return self.E.copy()

    def vector(self, k):
        """Instance k's normalised vector as dense [c1 | c2 | c0]."""
        c = torch.empty(self.soci.L + 1, dtype=CDT, device=self.dev)
        _lib.check(self.lib.apy_ci_unpack(self.O, self.V, _vp(self.X[k]), _vp(c), stream_of(self.dev)),
                   "apy_ci_unpack")
        return c
