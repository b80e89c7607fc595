"""Gaussian integrals over s and p Cartesian functions (host, NumPy).

Replaces the psi4 MintsHelper calls of hamiltonian.py:29-35, 49, 55 and
utils.py:371-376 for the basis sets used by the reference's tests (STO-nG on
H, C, N, O).  McMurchie-Davidson: Hermite expansion coefficients E_t^{ij}
and Hermite Coulomb integrals R_tuv built from the Boys function, evaluated
for all primitive combinations of a contracted quartet at once.

Conventions
  overlap    S_mn  = <m|n>
  kinetic    T_mn  = <m| -1/2 nabla^2 |n>
  potential  V_mn  = <m| -sum_C Z_C / |r - C| |n>
  ERI        (mn|ls) chemists' order, real
  angular    L_mn[k] = <m| (r x nabla)_k |n>  about the origin; real and
             antisymmetric (the operator r x p without the factor -i)
"""
import math
from itertools import product

import numpy as np
from scipy.special import gammainc, gamma

ANG = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
       2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]}   # Cartesian d, psi4's order


def boys(nmax, T):
    """F_n(T) for n = 0..nmax, T an array.  Upper order from the regularized
    incomplete gamma function, lower orders by downward recursion
    F_n = (2T F_{n+1} + e^{-T}) / (2n+1); Taylor series for tiny T."""
    T = np.asarray(T, dtype=np.float64)
    out = np.empty((nmax + 1,) + T.shape)
    small = T < 1e-9
    Ts = np.where(small, 1.0, T)
    a = nmax + 0.5
    top = gamma(a) * gammainc(a, Ts) / (2.0 * Ts ** a)
    series = 1.0 / (2 * nmax + 1) - T / (2 * nmax + 3)
    out[nmax] = np.where(small, series, top)
    e = np.exp(-T)
    for n in range(nmax - 1, -1, -1):
        rec = (2.0 * T * out[n + 1] + e) / (2 * n + 1)
        out[n] = np.where(small, 1.0 / (2 * n + 1) - T / (2 * n + 3), rec)
    return out


def hermite_E(i, j, Q, a, b):
    """E_t^{ij} for t = 0..i+j (arrays over primitive pairs); Q = A - B."""
    p = a + b
    q = a * b / p
    E = {(0, 0): [np.exp(-q * Q * Q)]}

    def get(ii, jj, t):
        if t < 0 or t > ii + jj:
            return 0.0
        return E[(ii, jj)][t]

    for ii in range(i + 1):
        for jj in range(j + 1):
            if (ii, jj) == (0, 0):
                continue
            if ii > 0:
                pi, pj = ii - 1, jj
                vals = [get(pi, pj, t - 1) / (2 * p) - (q * Q / a) * get(pi, pj, t) + (t + 1) * get(pi, pj, t + 1)
                        for t in range(ii + jj + 1)]
            else:
                pi, pj = ii, jj - 1
                vals = [get(pi, pj, t - 1) / (2 * p) + (q * Q / b) * get(pi, pj, t) + (t + 1) * get(pi, pj, t + 1)
                        for t in range(ii + jj + 1)]
            E[(ii, jj)] = vals
    return E[(i, j)]


def hermite_R(tmax, umax, vmax, p, X, Y, Z):
    """R_tuv^0 for t<=tmax, u<=umax, v<=vmax (arrays over primitive sets)."""
    nmax = tmax + umax + vmax
    R2 = X * X + Y * Y + Z * Z
    F = boys(nmax, p * R2)
    R = {}
    for n in range(nmax + 1):
        R[(0, 0, 0, n)] = (-2.0 * p) ** n * F[n]

    def r(t, u, v, n):
        key = (t, u, v, n)
        if key in R:
            return R[key]
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t > 0:
            val = (t - 1) * r(t - 2, u, v, n + 1) + X * r(t - 1, u, v, n + 1)
        elif u > 0:
            val = (u - 1) * r(t, u - 2, v, n + 1) + Y * r(t, u - 1, v, n + 1)
        else:
            val = (v - 1) * r(t, u, v - 2, n + 1) + Z * r(t, u, v - 1, n + 1)
        R[key] = val
        return val

    return {(t, u, v): r(t, u, v, 0) for t in range(tmax + 1) for u in range(umax + 1) for v in range(vmax + 1)}


class BasisFunction:
    """Contracted Cartesian Gaussian with unit norm."""

    def __init__(self, center, lmn, exps, coefs):
        self.center = np.asarray(center, dtype=np.float64)
        self.lmn = tuple(lmn)
        self.exps = np.asarray(exps, dtype=np.float64)
        L = sum(lmn)
        lx, ly, lz = lmn
        dfact = lambda k: float(np.prod(np.arange(k, 0, -2))) if k > 0 else 1.0
        norm = ((2.0 * self.exps / np.pi) ** 0.75 * (4.0 * self.exps) ** (L / 2.0)
                / np.sqrt(dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1)))
        c = np.asarray(coefs, dtype=np.float64) * norm
        # normalise the contraction
        a = self.exps[:, None] + self.exps[None, :]
        s = (np.pi / a) ** 1.5 * dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1) / (2.0 * a) ** L
        self.coefs = c / np.sqrt(c @ s @ c)


def _pair(f, g):
    """Primitive-pair arrays for functions f, g (flattened)."""
    a = np.repeat(f.exps, len(g.exps))
    b = np.tile(g.exps, len(f.exps))
    ca = np.repeat(f.coefs, len(g.coefs))
    cb = np.tile(g.coefs, len(f.coefs))
    p = a + b
    P = (a[:, None] * f.center + b[:, None] * g.center) / p[:, None]
    return a, b, ca * cb, p, P


def overlap(f, g):
    a, b, c, p, _ = _pair(f, g)
    Q = f.center - g.center
    val = np.ones_like(p)
    for d in range(3):
        val = val * hermite_E(f.lmn[d], g.lmn[d], Q[d], a, b)[0]
    return float(np.sum(c * val * (np.pi / p) ** 1.5))


def _overlap_1d(i, j, Q, a, b, p):
    if i < 0 or j < 0:
        return 0.0
    return hermite_E(i, j, Q, a, b)[0] * np.sqrt(np.pi / p)


def kinetic(f, g):
    a, b, c, p, _ = _pair(f, g)
    Q = f.center - g.center
    S = [[_overlap_1d(f.lmn[d], g.lmn[d] + k, Q[d], a, b, p) for k in (-2, 0, 2)] for d in range(3)]
    total = np.zeros_like(p)
    for d in range(3):
        j = g.lmn[d]
        Td = j * (j - 1) * S[d][0] - 2.0 * b * (2 * j + 1) * S[d][1] + 4.0 * b * b * S[d][2]
        term = -0.5 * Td
        for e in range(3):
            if e != d:
                term = term * S[e][1]
        total = total + term
    return float(np.sum(c * total))


def potential(f, g, charges):
    a, b, c, p, P = _pair(f, g)
    Q = f.center - g.center
    Ex = hermite_E(f.lmn[0], g.lmn[0], Q[0], a, b)
    Ey = hermite_E(f.lmn[1], g.lmn[1], Q[1], a, b)
    Ez = hermite_E(f.lmn[2], g.lmn[2], Q[2], a, b)
    tm, um, vm = f.lmn[0] + g.lmn[0], f.lmn[1] + g.lmn[1], f.lmn[2] + g.lmn[2]
    total = np.zeros_like(p)
    for Z, C in charges:
        PC = P - np.asarray(C)
        R = hermite_R(tm, um, vm, p, PC[:, 0], PC[:, 1], PC[:, 2])
        s = np.zeros_like(p)
        for t in range(tm + 1):
            for u in range(um + 1):
                for v in range(vm + 1):
                    s = s + Ex[t] * Ey[u] * Ez[v] * R[(t, u, v)]
        total = total - Z * s
    return float(np.sum(c * 2.0 * np.pi / p * total))


def eri(f, g, h, k):
    a, b, cab, p, P = _pair(f, g)
    cc, d, ccd, q, Qc = _pair(h, k)
    Qab = f.center - g.center
    Qcd = h.center - k.center
    Eab = [hermite_E(f.lmn[x], g.lmn[x], Qab[x], a, b) for x in range(3)]
    Ecd = [hermite_E(h.lmn[x], k.lmn[x], Qcd[x], cc, d) for x in range(3)]
    n1, n2 = len(p), len(q)
    pp = np.repeat(p, n2)
    qq = np.tile(q, n1)
    alpha = pp * qq / (pp + qq)
    PQ = np.repeat(P, n2, axis=0) - np.tile(Qc, (n1, 1))
    L1 = [f.lmn[x] + g.lmn[x] for x in range(3)]
    L2 = [h.lmn[x] + k.lmn[x] for x in range(3)]
    R = hermite_R(L1[0] + L2[0], L1[1] + L2[1], L1[2] + L2[2], alpha, PQ[:, 0], PQ[:, 1], PQ[:, 2])
    total = np.zeros(n1 * n2)
    for t, u, v in product(range(L1[0] + 1), range(L1[1] + 1), range(L1[2] + 1)):
        e1 = np.repeat(Eab[0][t] * Eab[1][u] * Eab[2][v], n2)
        for tau, nu, phi in product(range(L2[0] + 1), range(L2[1] + 1), range(L2[2] + 1)):
            e2 = np.tile(Ecd[0][tau] * Ecd[1][nu] * Ecd[2][phi], n1)
            sign = -1.0 if (tau + nu + phi) % 2 else 1.0
            total = total + sign * e1 * e2 * R[(t + tau, u + nu, v + phi)]
    pref = 2.0 * np.pi ** 2.5 / (pp * qq * np.sqrt(pp + qq))
    c = np.repeat(cab, n2) * np.tile(ccd, n1)
    return float(np.sum(c * pref * total))


def _multipole_1d(i, j, e, Q, a, b, p, PX, C):
    """1-D integral of x_A^i (x - C)^e x_B^j: moments about C via the
    Hermite expansion (x - C = (x - P) + (P - C))."""
    E = hermite_E(i, j, Q, a, b)
    if e == 0:
        return E[0] * np.sqrt(np.pi / p)
    # e == 1: int x_P * Lambda_t = t-dependent; <x - P> = E_1 * sqrt(pi/p)
    E1 = E[1] if len(E) > 1 else 0.0
    return (E1 + (PX - C) * E[0]) * np.sqrt(np.pi / p)


def _deriv_1d(i, j, Q, a, b, p):
    """1-D integral of x_A^i d/dx (x_B^j e^{-b x_B^2})."""
    up = _overlap_1d(i, j + 1, Q, a, b, p)
    down = _overlap_1d(i, j - 1, Q, a, b, p) if j > 0 else 0.0
    return j * down - 2.0 * b * up


def angular_momentum(f, g, origin=(0.0, 0.0, 0.0)):
    """<f| (r - O) x nabla |g> (three components)."""
    a, b, c, p, P = _pair(f, g)
    Q = f.center - g.center
    O = np.asarray(origin, dtype=np.float64)
    S = [_overlap_1d(f.lmn[x], g.lmn[x], Q[x], a, b, p) for x in range(3)]
    M = [_multipole_1d(f.lmn[x], g.lmn[x], 1, Q[x], a, b, p, P[:, x], O[x]) for x in range(3)]
    D = [_deriv_1d(f.lmn[x], g.lmn[x], Q[x], a, b, p) for x in range(3)]
    out = []
    for (y, z, x) in ((1, 2, 0), (2, 0, 1), (0, 1, 2)):
        # (r x nabla)_x = y d/dz - z d/dy
        val = M[y] * D[z] * S[x] - M[z] * D[y] * S[x]
        out.append(float(np.sum(c * val)))
    return out


def dipole(f, g, origin=(0.0, 0.0, 0.0)):
    """<f| (r - O) |g> (three components)."""
    a, b, c, p, P = _pair(f, g)
    Q = f.center - g.center
    O = np.asarray(origin, dtype=np.float64)
    S = [_overlap_1d(f.lmn[x], g.lmn[x], Q[x], a, b, p) for x in range(3)]
    out = []
    for x in range(3):
        val = _multipole_1d(f.lmn[x], g.lmn[x], 1, Q[x], a, b, p, P[:, x], O[x])
        for y in range(3):
            if y != x:
                val = val * S[y]
        out.append(float(np.sum(c * val)))
    return out


def build_one_electron(basis, charges):
    n = len(basis)
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            S[i, j] = S[j, i] = overlap(basis[i], basis[j])
            T[i, j] = T[j, i] = kinetic(basis[i], basis[j])
            V[i, j] = V[j, i] = potential(basis[i], basis[j], charges)
    return S, T, V


def build_eri(basis):
    """Full (mn|ls) tensor using the 8-fold permutational symmetry."""
    n = len(basis)
    G = np.zeros((n, n, n, n))
    for i in range(n):
        for j in range(i + 1):
            ij = i * (i + 1) // 2 + j
            for k in range(n):
                for l in range(k + 1):
                    kl = k * (k + 1) // 2 + l
                    if kl > ij:
                        continue
                    v = eri(basis[i], basis[j], basis[k], basis[l])
                    for (p, q) in ((i, j), (j, i)):
                        for (r, s) in ((k, l), (l, k)):
                            G[p, q, r, s] = v
                            G[r, s, p, q] = v
    return G


def build_mixed_overlap(bra, ket):
    return np.array([[overlap(f, g) for g in ket] for f in bra])


def build_angular_momentum(basis, origin=(0.0, 0.0, 0.0)):
    n = len(basis)
    L = np.zeros((3, n, n))
    for i in range(n):
        for j in range(n):
            L[:, i, j] = angular_momentum(basis[i], basis[j], origin)
    return L


def build_dipole(basis, origin=(0.0, 0.0, 0.0)):
    n = len(basis)
    D = np.zeros((3, n, n))
    for i in range(n):
        for j in range(i + 1):
            D[:, i, j] = D[:, j, i] = dipole(basis[i], basis[j], origin)
    return D


# ----------------------------------------------------------------------------------------------------------
# First derivatives with respect to nuclear positions (the integrals psi4's MintsHelper gives to
# analytic_aats.py:84-96: mo_oei_deriv1, mo_tei_deriv1, mo_overlap_half_deriv1)

class RawGaussian:
    """Contracted Cartesian Gaussian with the coefficients taken as given
    (no renormalisation): the terms of a differentiated basis function."""

    def __init__(self, center, lmn, exps, coefs):
        self.center = center
        self.lmn = tuple(lmn)
        self.exps = exps
        self.coefs = coefs


def center_derivative(f, d):
    """d f / d A_d for f centred at A: sum_p c_p [2 a_p x_A^(l+1) - l x_A^(l-1)]
    e^(-a_p r_A^2) along axis d, as a list of RawGaussian terms."""
    up = list(f.lmn)
    up[d] += 1
    out = [RawGaussian(f.center, up, f.exps, 2.0 * f.exps * f.coefs)]
    if f.lmn[d] > 0:
        dn = list(f.lmn)
        dn[d] -= 1
        out.append(RawGaussian(f.center, dn, f.exps, -float(f.lmn[d]) * f.coefs))
    return out


def _dsum(fn, terms, *rest):
    return sum(fn(t, *rest) for t in terms)


def build_deriv1(basis, atom_of, charges, eri_derivatives=True):
    """First-derivative AO integrals for every atom A and axis x (the ERI
    part only with eri_derivatives=True):
        S[A][x], T[A][x], V[A][x]  d/dR_Ax of S, T, V (V with the derivative
                                   of the nuclear attraction operator of A)
        half_S[A][x][m, n] = <d m/dR_Ax | n>  (bra functions only)
        ERI[A][x]                  d/dR_Ax of (mn|ls), chemists' order
    A basis function moves with its atom (d m/dR_A = d m/dA for m on A);
    the operator term uses translational invariance of <m|V_C|n>:
    d/dC = -(d/dA + d/dB)."""
    n = len(basis)
    natom = len(charges)
    atom_of = np.asarray(atom_of)
    dS = np.zeros((3, n, n))                  # <d m/dx | n> for every m (its own centre)
    dT = np.zeros((3, n, n))
    dVC = np.zeros((natom, 3, n, n))          # <d m/dx | V_C | n>
    for x in range(3):
        for m in range(n):
            terms = center_derivative(basis[m], x)
            for k in range(n):
                dS[x, m, k] = _dsum(overlap, terms, basis[k])
                dT[x, m, k] = _dsum(kinetic, terms, basis[k])
                for C, (Z, R) in enumerate(charges):
                    dVC[C, x, m, k] = _dsum(potential, terms, basis[k], [(Z, R)])
    # D[x][m,n,l,s] = (d m/dx n | l s), (l<=s) computed, (l,s) symmetric
    D = np.zeros((3, n, n, n, n)) if eri_derivatives else None
    for x in range(3 if eri_derivatives else 0):
        for m in range(n):
            terms = center_derivative(basis[m], x)
            for q in range(n):
                for l in range(n):
                    for s in range(l + 1):
                        v = sum(eri(t, basis[q], basis[l], basis[s]) for t in terms)
                        D[x, m, q, l, s] = D[x, m, q, s, l] = v
    return assemble_deriv1(dS, dT, dVC, atom_of, D)


def assemble_deriv1(dS, dT, dVC, atom_of, D=None):
    """The per-atom derivative integrals of build_deriv1 from the bra-side
    derivative integrals dS[x] = <d m/dx | n>, dT[x] = <d m/dx | T | n>,
    dVC[C, x] = <d m/dx | V_C | n> (V_C the attraction of nucleus C alone)
    and, when given, D[x] = (d m/dx n | l s)."""
    natom = dVC.shape[0]
    atom_of = np.asarray(atom_of)
    dV = dVC.sum(0)
    out = {'S': [], 'T': [], 'V': [], 'half_S': [], 'ERI': []}
    for A in range(natom):
        on = (atom_of == A).astype(float)
        mS, mT, mV, mH, mE = [], [], [], [], []
        for x in range(3):
            bra = on[:, None] * dS[x]
            mH.append(bra)
            mS.append(bra + bra.T)
            t = on[:, None] * dT[x]
            mT.append(t + t.T)
            v = on[:, None] * dV[x]
            op = dVC[A, x] + dVC[A, x].T
            mV.append(v + v.T - op)
            if D is not None:
                Dx = D[x]
                e = (on[:, None, None, None] * Dx
                     + on[None, :, None, None] * Dx.transpose(1, 0, 2, 3)
                     + on[None, None, :, None] * Dx.transpose(2, 3, 0, 1)
                     + on[None, None, None, :] * Dx.transpose(2, 3, 1, 0))
                mE.append(e)
        out['S'].append(mS)
        out['T'].append(mT)
        out['V'].append(mV)
        out['half_S'].append(mH)
        out['ERI'].append(mE)
    return out
