"""Finite-difference atomic axial tensors with apyib's interface (aats.py).

`AAT(parameters, wfn, C0, basis0, T0, nuc_pos_C, nuc_neg_C, nuc_pos_basis,
nuc_neg_basis, nuc_pos_T, nuc_neg_T, mag_pos_C, mag_neg_C, mag_pos_basis,
mag_neg_basis, mag_pos_T, mag_neg_T, dR, dH)` with
`compute_spatial_aats(alpha, beta, normalization='full')` and
`compute_SO_aats(alpha, beta, normalization='full')` as in the reference
(aats.py:20-1055), plus `compute_aats(normalization)` returning the whole
[3N, 3] tensor from one batch of GPU work.

All overlaps (aats.py:56-115), their determinant-lemma records, the
amplitude combinations and the amplitude-determinant sums run in
libapyib_amd.so (apy_mo_overlap, apy_so_fock_from_mo, apy_aat_gamma,
apy_aat_norms, apy_aat_axpby, apy_aat_terms).  The host combines the 20
scalars of each overlap class into the reference's terms:

spatial (x d^2 per class; aats.py:713-1047), with d = det of the spatial
occupied overlap:
    0S = 2 sy N_bra        S0 = 2 sx N_ket
    0D = 2 kyy + kyyp      D0 = 2 axx + bxx
    SS = 2 s1 + 4 sx sy
    DS = 4 s3 + 4 sy axx + 2 bxx sy + 2 s2
    SD = -4 s4 + 4 sx kyy + 2 sx kyyp + 2 s5
    DD = 2 e1 + 4 axx kyy - 8 s6 + 4 s7 + 2 kyyp axx - 2 s8 - 2 s9
         + 2 bxx kyy + bxx kyyp + s10 + s11 + e2
  with the reference's branch structure: 0S/S0/0D/D0 and singles only for
  CISD (for MP2 and CID only the HF term and DD enter).
spin-orbital (x d per class, aats.py:161-516), A0/K0 = reference weight of
the bra/ket side:
    A0 K0 + A0 (sy + kyy/2) + K0 (sx + axx/2) + sx sy + s1 + s3 + sy axx/2
    - s4 + sx kyy/2 + e1/4 + axx kyy/4 - s6
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import transforms as tr
from .integrals import mixed_overlaps
from .kernels import stream_of, workspace
from .utils import device, mixed_to_basis

CDT = torch.complex128
NT = _lib.APY_AAT_NTERMS

# term scalars (indices of apy_aat_terms' output)
D, SX, SY, BXX, AXX, KYYP, KYY = 0, 1, 2, 3, 4, 5, 6
S1, S2, S3, S4, S5, S6, S7, S8, S9, S10, S11, E1, E2 = range(7, 20)

# class: (overlap key, bra kind, ket kind, sign).  kinds: 'dR' (conj of the
# normalised nuclear difference), 'dH' (normalised field difference), 't'
# (normalised unperturbed amplitudes; conjugated on the bra side)
CLASSES = (('uu', 'dR', 'dH', +1), ('up', 'dR', 't', +1), ('un', 'dR', 't', -1),
           ('pu', 't', 'dH', +1), ('nu', 't', 'dH', -1),
           ('pp', 't', 't', +1), ('pn', 't', 't', -1), ('np', 't', 't', -1), ('nn', 't', 't', +1))


def _vp(t):
    return ctypes.c_void_p(t.data_ptr() if t is not None else 0)


def _dev_amp(x, shape, dev):
    if x is None or np.isscalar(x) or (isinstance(x, np.ndarray) and x.ndim == 0):
        return None
    if isinstance(x, torch.Tensor):
        return x.to(dev, CDT).contiguous().reshape(shape)
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).to(dev).reshape(shape)


def spatial_terms(r):
    """The reference's spatial terms of one class from its scalar vector r
    (divided by d^2)."""
    sx, sy, bxx, axx, kyyp, kyy = r[SX], r[SY], r[BXX], r[AXX], r[KYYP], r[KYY]
    return {
        '0S': 2 * sy, 'S0': 2 * sx, '0D': 2 * kyy + kyyp, 'D0': 2 * axx + bxx,
        'SS': 2 * r[S1] + 4 * sx * sy,
        'DS': 4 * r[S3] + 4 * sy * axx + 2 * bxx * sy + 2 * r[S2],
        'SD': -4 * r[S4] + 4 * sx * kyy + 2 * sx * kyyp + 2 * r[S5],
        'DD': (2 * r[E1] + 4 * axx * kyy - 8 * r[S6] + 4 * r[S7] + 2 * kyyp * axx - 2 * r[S8] - 2 * r[S9]
               + 2 * bxx * kyy + bxx * kyyp + r[S10] + r[S11] + r[E2]),
    }


def so_terms(r, A0, K0):
    """<bra|ket> / d of one spin-orbital class split into the reference's
    terms (compute_SO_I_00, _0S, _0D, _S0, _D0, _SS, _DS, _SD, _DD)."""
    sx, sy, axx, kyy = r[SX], r[SY], r[AXX], r[KYY]
    return {'00': A0 * K0, '0S': A0 * sy, '0D': A0 * 0.5 * kyy, 'S0': K0 * sx, 'D0': K0 * 0.5 * axx,
            'SS': sx * sy + r[S1], 'DS': r[S3] + 0.5 * sy * axx, 'SD': -r[S4] + 0.5 * sx * kyy,
            'DD': 0.25 * r[E1] + 0.25 * axx * kyy - r[S6]}


def so_overlap(r, A0, K0):
    """<bra|ket> / d of one spin-orbital class."""
    return sum(so_terms(r, A0, K0).values())


class AAT(object):
    """The atomic axial tensor object computed by finite difference
    (aats.py:20-115)."""

    def __init__(self, parameters, wfn, unperturbed_wfn, unperturbed_basis, unperturbed_T, nuc_pos_wfn, nuc_neg_wfn,
                 nuc_pos_basis, nuc_neg_basis, nuc_pos_T, nuc_neg_T, mag_pos_wfn, mag_neg_wfn, mag_pos_basis,
                 mag_neg_basis, mag_pos_T, mag_neg_T, nuc_pert_strength, mag_pert_strength):
        self.lib = _lib.load()
        self.dev = device()
        self.parameters = parameters
        self.method = parameters['method']
        self.so = self.method.endswith('_SO')
        self.nbf = wfn.nbf
        self.ndocc = wfn.ndocc
        self.nfzc = wfn.H.basis_set.n_frozen_core() if parameters.get('freeze_core', False) else 0
        if self.so and self.nfzc:
            raise ValueError("spin-orbital AATs are defined without frozen core (aats.py:120-516)")
        self.nuc_pert_strength = nuc_pert_strength
        self.mag_pert_strength = mag_pert_strength
        self.natom = len(nuc_pos_wfn) // 3
        self.unperturbed_T = unperturbed_T
        self.nuc_pos_T, self.nuc_neg_T = nuc_pos_T, nuc_neg_T
        self.mag_pos_T, self.mag_neg_T = mag_pos_T, mag_neg_T

        nA = len(nuc_pos_wfn)
        u = (unperturbed_basis, unperturbed_wfn)
        pairs = [('uu', None, None, u, u)]
        pairs += [(k, None, b, u, (bs[b], cs[b])) for k, cs, bs in (('up', mag_pos_wfn, mag_pos_basis),
                                                                     ('un', mag_neg_wfn, mag_neg_basis))
                  for b in range(3)]
        pairs += [(k, a, None, (bs[a], cs[a]), u) for k, cs, bs in (('pu', nuc_pos_wfn, nuc_pos_basis),
                                                                     ('nu', nuc_neg_wfn, nuc_neg_basis))
                  for a in range(nA)]
        for k, (cr, br), (cm, bm) in (('pp', (nuc_pos_wfn, nuc_pos_basis), (mag_pos_wfn, mag_pos_basis)),
                                      ('pn', (nuc_pos_wfn, nuc_pos_basis), (mag_neg_wfn, mag_neg_basis)),
                                      ('np', (nuc_neg_wfn, nuc_neg_basis), (mag_pos_wfn, mag_pos_basis)),
                                      ('nn', (nuc_neg_wfn, nuc_neg_basis), (mag_neg_wfn, mag_neg_basis))):
            pairs += [(k, a, b, (br[a], cr[a]), (bm[b], cm[b])) for a in range(nA) for b in range(3)]
        self.index = {(k, a, b): q for q, (k, a, b, _, _) in enumerate(pairs)}
        self._overlaps(pairs)
        self._amplitudes()

    # ------------------------------------------------------------------ setup
    def _overlaps(self, pairs):
        """All MO overlaps of aats.py:56-115 in one batched product, then
        their determinant-lemma records."""
        dev, nbf = self.dev, self.nbf
        nov = len(pairs)
        fn = lambda b: b.functions if hasattr(b, 'functions') else b
        S_ao = mixed_overlaps([(fn(bb), fn(kb)) for _, _, _, (bb, _), (kb, _) in pairs], dev)
        if any(getattr(b, 'transform', None) is not None for _, _, _, (bb, _), (kb, _) in pairs for b in (bb, kb)):
            S_ao = [torch.from_numpy(mixed_to_basis(bb, kb, S.cpu().numpy())).to(dev)
                    for S, (_, _, _, (bb, _), (kb, _)) in zip(S_ao, pairs)]
        S_ao = torch.stack(S_ao)
        Cb = np.stack([np.asarray(c, dtype=np.complex128) for _, _, _, (_, c), _ in pairs])
        Ck = np.stack([np.asarray(c, dtype=np.complex128) for _, _, _, _, (_, c) in pairs])
        S_mo = tr.mo_overlap(tr.to_dev(Cb, dev), tr.to_dev(S_ao, dev), tr.to_dev(Ck, dev))
        if self.so:
            S_mo = torch.stack([tr.so_fock(S_mo[q]) for q in range(nov)])
            n, no, nf = 2 * nbf, 2 * self.ndocc, 0
        else:
            n, no, nf = nbf, self.ndocc, self.nfzc
        self.n, self.no, self.nf = n, no, nf
        self.oa, self.nv = no - nf, n - no
        self.S_mo = S_mo.contiguous()
        self.record = int(self.lib.apy_aat_gamma_record(no, nf, n - no))
        self.gamma = torch.empty((nov, self.record), dtype=CDT, device=dev)
        nbytes = int(self.lib.apy_aat_gamma_workspace(n, no, nov))
        ws = workspace(nbytes, dev)
        _lib.check(self.lib.apy_aat_gamma(n, no, nf, nov, _vp(self.S_mo), n * n, _vp(self.gamma), _vp(ws), nbytes,
                                          stream_of(dev)), "apy_aat_gamma")
        self.det = self.gamma[:, 0].cpu().numpy()
        if np.any(self.det == 0):
            raise np.linalg.LinAlgError("singular occupied-block overlap in the AAT overlap set")

    def _amplitudes(self):
        dev, oa, nv = self.dev, self.oa, self.nv
        s1, s2 = (oa, nv), (oa, oa, nv, nv)
        singles = self.method in ('CISD', 'CISD_SO')
        up = lambda T: ((_dev_amp(T[1], s1, dev) if singles else None), _dev_amp(T[2], s2, dev))
        self.amps = {'u': up(self.unperturbed_T),
                     'np': [up(T) for T in self.nuc_pos_T], 'nn': [up(T) for T in self.nuc_neg_T],
                     'mp': [up(T) for T in self.mag_pos_T], 'mn': [up(T) for T in self.mag_neg_T]}
        self.norm = self._norms() if self.method != 'RHF' else None

    def _norms(self):
        """Per amplitude set: |t1|^2, |t2|^2, Re conj(t2).t2^ab."""
        sets = [self.amps['u']] + self.amps['np'] + self.amps['nn'] + self.amps['mp'] + self.amps['mn']
        n = len(sets)
        t1 = (ctypes.c_void_p * n)(*[s[0].data_ptr() if s[0] is not None else None for s in sets])
        t2 = (ctypes.c_void_p * n)(*[s[1].data_ptr() for s in sets])
        out = torch.empty((n, 3), dtype=torch.float64, device=self.dev)
        nbytes = int(self.lib.apy_aat_norms_workspace(n))
        ws = workspace(nbytes, self.dev)
        _lib.check(self.lib.apy_aat_norms(self.oa, self.nv, n, t1, t2, _vp(out), _vp(ws), nbytes,
                                          stream_of(self.dev)), "apy_aat_norms")
        v = out.cpu().numpy()
        nA = len(self.amps['np'])
        keys = ['u'] + [('np', a) for a in range(nA)] + [('nn', a) for a in range(nA)] + \
            [('mp', b) for b in range(3)] + [('mn', b) for b in range(3)]
        return dict(zip(keys, v))

    def _N(self, key, normalization):
        """Normalisation factors of aats.py:134-157 (SO) and 652-669 (spatial)."""
        if self.method == 'RHF' or normalization == 'intermediate':
            return 1.0
        n1, n2, nx = self.norm[key]
        if self.so:
            x = 1.0 + 0.25 * n2 + (n1 if self.method == 'CISD_SO' else 0.0)
        else:
            x = 1.0 + 2 * n2 - nx + (2 * n1 if self.method == 'CISD' else 0.0)
        return 1.0 / np.sqrt(x)

    # ------------------------------------------------------------------ AATs
    def _element_jobs(self, alpha, beta, normalization, combos, jobs, meta):
        """Jobs of one element; combos collects the amplitude combinations."""
        N = self._N('u', normalization)
        N_np, N_nn = self._N(('np', alpha), normalization), self._N(('nn', alpha), normalization)
        N_mp, N_mn = self._N(('mp', beta), normalization), self._N(('mn', beta), normalization)
        A = self.amps
        if self.so:          # the SO form uses unnormalised differences and amplitudes (aats.py:161-516)
            cu, cp, cn, dp, dn = 1.0, 1.0, 1.0, 1.0, 1.0
        else:                # spatial: normalised amplitudes (aats.py:680-704)
            cu, cp, cn, dp, dn = N, N_np, N_nn, N_mp, N_mn
        bra = {'t': combos.get(('tb', normalization), A['u'], None, cu, 0.0, True),
               'dR': combos.get(('dR', alpha, normalization), A['np'][alpha], A['nn'][alpha], cp, -cn, True)}
        ket = {'t': combos.get(('tk', normalization), A['u'], None, cu, 0.0, False),
               'dH': combos.get(('dH', beta, normalization), A['mp'][beta], A['mn'][beta], dp, -dn, False)}
        for key, bk, kk, sign in CLASSES:
            idx = self.index[(key, alpha if key[0] in 'pn' else None, beta if key[1] in 'pn' else None)]
            jobs.append((idx, bra[bk], ket[kk]))
            meta.append((alpha, beta, key, bk, kk, sign, (N, N_np, N_nn, N_mp, N_mn)))

    TERMS = ('00', '0S', '0D', 'S0', 'D0', 'SS', 'SD', 'DS', 'DD')

    def _run_terms(self, elements, normalization):
        """Evaluate the elements [(alpha, beta)] -> {term: array of AAT
        contributions} (the reference's I_00, I_0S, ... of each element)."""
        out = {t: np.zeros(len(elements)) for t in self.TERMS}
        f = 1 / (4 * self.nuc_pert_strength * self.mag_pert_strength)
        if self.method == 'RHF':
            for q, (a, b) in enumerate(elements):
                out['00'][q] = self._i00(a, b, (1.0,) * 5)
            return out
        combos = _Combos(self)
        jobs, meta = [], []
        for a, b in elements:
            self._element_jobs(a, b, normalization, combos, jobs, meta)
        combos.run()
        R = self._terms(jobs)
        acc = {}
        for r, (a, b, key, bk, kk, sign, Ns) in zip(R, meta):
            e = acc.setdefault((a, b), {t: 0.0 for t in self.TERMS})
            for t, v in self._class_value(r, key, bk, kk, Ns).items():
                e[t] = e[t] + sign * v
        for q, (a, b) in enumerate(elements):
            Ns = meta[len(CLASSES) * q][6]
            if not self.so:              # the SO class overlaps hold the reference-reference term themselves
                acc[(a, b)]['00'] = acc[(a, b)]['00'] + self._i00_complex(a, b, Ns)
            for t in self.TERMS:
                out[t][q] = f * np.imag(acc[(a, b)][t])
        return out

    def _run(self, elements, normalization):
        """Evaluate the elements [(alpha, beta)] -> array of AATs."""
        T = self._run_terms(elements, normalization)
        return sum(T[t] for t in self.TERMS)

    def _class_value(self, r, key, bk, kk, Ns):
        """{term: contribution} of one overlap class (before the class sign)."""
        N, N_np, N_nn, N_mp, N_mn = Ns
        d = r[D]
        if self.so:
            nb = {'u': N, 'p': N_np, 'n': N_nn}[key[0]]
            nk = {'u': N, 'p': N_mp, 'n': N_mn}[key[1]]
            return {t: nb * nk * d * v for t, v in so_terms(r, 1.0 if bk == 't' else 0.0,
                                                            1.0 if kk == 't' else 0.0).items()}
        T = spatial_terms(r)
        names = ['DD']
        if self.method == 'CISD':
            names += ['SS', 'SD', 'DS']
            if kk == 't':
                names.append('D0')
                T['S0'] = T['S0'] * {'p': N_mp, 'n': N_mn}[key[1]]
                names.append('S0')
            if bk == 't':
                names.append('0D')
                T['0S'] = T['0S'] * {'p': N_np, 'n': N_nn}[key[0]]
                names.append('0S')
        return {k: d * d * T[k] for k in names}

    def _i00_complex(self, alpha, beta, Ns):
        """HF term (aats.py:672-677; SO aats.py:161-179)."""
        N, N_np, N_nn, N_mp, N_mn = Ns
        d = lambda k: self.det[self.index[(k, alpha, beta)]]
        p = (lambda x: x) if self.so else (lambda x: x * x)
        return (p(d('pp')) * N_np * N_mp - p(d('pn')) * N_np * N_mn - p(d('np')) * N_nn * N_mp
                + p(d('nn')) * N_nn * N_mn)

    def _i00(self, alpha, beta, Ns):
        return (1 / (4 * self.nuc_pert_strength * self.mag_pert_strength)) * np.imag(
            self._i00_complex(alpha, beta, Ns))

    def _terms(self, jobs):
        n = len(jobs)
        table = (_lib.AatJob * n)()
        for q, (g, (b1, b2), (k1, k2)) in enumerate(jobs):
            table[q] = _lib.AatJob(g, b1.data_ptr() if b1 is not None else None, b2.data_ptr(),
                                   k1.data_ptr() if k1 is not None else None, k2.data_ptr())
        out = torch.empty((n, NT), dtype=CDT, device=self.dev)
        nbytes = int(self.lib.apy_aat_terms_workspace(self.oa, self.nv, n))
        ws = workspace(nbytes, self.dev)
        _lib.check(self.lib.apy_aat_terms(self.oa, self.nv, _vp(self.gamma), self.record, self.gamma.shape[0], n, table,
                                          _vp(out),
                                          _vp(ws), nbytes, stream_of(self.dev)), "apy_aat_terms")
        return out.cpu().numpy()

    def compute_spatial_aats(self, alpha, beta, normalization='full'):
        """aats.py:646-1055."""
        if self.so:
            raise ValueError("compute_spatial_aats needs a spatial method (RHF, MP2, CID, CISD)")
        return float(self._run([(alpha, beta)], normalization)[0])

    def compute_SO_aats(self, alpha, beta, normalization='full'):
        """aats.py:520-554 (the overlaps are expanded to spin orbitals at
        construction for the _SO methods)."""
        if not self.so:
            raise ValueError("compute_SO_aats needs an _SO method")
        return float(self._run([(alpha, beta)], normalization)[0])

    def _so_term(self, term, alpha, beta, normalization):
        if not self.so and self.method != 'RHF':
            raise ValueError("the compute_SO_I_* terms need an _SO method")
        return float(self._run_terms([(alpha, beta)], normalization)[term][0])

    def compute_SO_I_00(self, alpha, beta, normalization='full'):
        """aats.py:161-179."""
        return self._so_term('00', alpha, beta, normalization)

    def compute_SO_I_0D(self, alpha, beta, normalization='full'):
        """aats.py:183-211."""
        return self._so_term('0D', alpha, beta, normalization)

    def compute_SO_I_D0(self, alpha, beta, normalization='full'):
        """aats.py:215-243."""
        return self._so_term('D0', alpha, beta, normalization)

    def compute_SO_I_DD(self, alpha, beta, normalization='full'):
        """aats.py:247-302."""
        return self._so_term('DD', alpha, beta, normalization)

    def compute_SO_I_0S(self, alpha, beta, normalization='full'):
        """aats.py:306-332."""
        return self._so_term('0S', alpha, beta, normalization)

    def compute_SO_I_S0(self, alpha, beta, normalization='full'):
        """aats.py:336-362."""
        return self._so_term('S0', alpha, beta, normalization)

    def compute_SO_I_SS(self, alpha, beta, normalization='full'):
        """aats.py:366-412."""
        return self._so_term('SS', alpha, beta, normalization)

    def compute_SO_I_SD(self, alpha, beta, normalization='full'):
        """aats.py:416-464."""
        return self._so_term('SD', alpha, beta, normalization)

    def compute_SO_I_DS(self, alpha, beta, normalization='full'):
        """aats.py:468-516."""
        return self._so_term('DS', alpha, beta, normalization)

    def compute_aat_terms(self, normalization='full'):
        """{term: [3N, 3]} for the whole tensor in one batch."""
        nA = 3 * self.natom
        els = [(a, b) for a in range(nA) for b in range(3)]
        return {t: v.reshape(nA, 3) for t, v in self._run_terms(els, normalization).items()}

    def compute_aats(self, normalization='full'):
        """The [3N, 3] tensor in one batch (the element map of
        parallel.py:33-47)."""
        nA = 3 * self.natom
        els = [(a, b) for a in range(nA) for b in range(3)]
        return self._run(els, normalization).reshape(nA, 3)


class _Combos:
    """Normalised amplitude combinations, formed on the device in one
    apy_aat_axpby launch and shared between the jobs that use them."""

    def __init__(self, aat):
        self.aat = aat
        self.ops = []
        self.cache = {}

    def get(self, key, A, B, ca, cb, conj):
        if key in self.cache:
            return self.cache[key]
        out = []
        for k in (0, 1):
            a = A[k]
            if a is None:
                out.append(None)
                continue
            o = torch.empty_like(a)
            b = B[k] if B is not None else None
            self.ops.append(_lib.AatAxpbyOp(a.numel(), a.data_ptr(), b.data_ptr() if b is not None else None,
                                            o.data_ptr(), float(ca), float(cb), int(conj)))
            out.append(o)
        self.cache[key] = tuple(out)
        return self.cache[key]

    def run(self):
        if not self.ops:
            return
        lib, dev = self.aat.lib, self.aat.dev
        n = len(self.ops)
        for s in range(0, n, 65535):
            chunk = self.ops[s:s + 65535]
            arr = (_lib.AatAxpbyOp * len(chunk))(*chunk)
            nbytes = int(lib.apy_aat_axpby_workspace(len(chunk)))
            ws = workspace(nbytes, dev)
            _lib.check(lib.apy_aat_axpby(len(chunk), arr, ctypes.c_void_p(ws.data_ptr()), nbytes, stream_of(dev)),
                       "apy_aat_axpby")
