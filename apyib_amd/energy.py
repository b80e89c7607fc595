"""Method dispatch with apyib's interface (energy.py:17-150)."""
import numpy as np

from .ci_wfn import ci_wfn
from .hamiltonian import build_hamiltonians
from .hf_wfn import hf_wfn, solve_SCF_batch
from .integrals import prime_device_eri
from .mp2_wfn import mp2_wfn
from .spatial_ci import SpatialCISet
from .utils import compute_phase, device


def _solve(parameters, wfn, print_level):
    E, t1, t2 = 0, 0, 0
    m = parameters['method']
    if m == 'MP2':
        E, t2 = mp2_wfn(parameters, wfn).solve_MP2()
    if m == 'CID':
        E, t2 = ci_wfn(parameters, wfn).solve_CID(print_level)
    if m == 'CISD':
        E, t1, t2 = ci_wfn(parameters, wfn).solve_CISD(print_level)
    if m == 'MP2_SO':
        E, t2 = mp2_wfn(parameters, wfn).solve_MP2_SO()
    if m == 'CID_SO':
        E, t2 = ci_wfn(parameters, wfn).solve_CID_SO(print_level)
    if m == 'CISD_SO':
        E, t1, t2 = ci_wfn(parameters, wfn).solve_CISD_SO(print_level)
    return E, t1, t2


def energy(parameters, return_H=False, print_level=0):
    """energy.py:17-81.  Returns (E_list, T_list, C, basis[, H]) with
    E_list = [E_SCF, E_corr, E_nuc] and T_list = [1, t1, t2]."""
    H = build_hamiltonians([parameters], device())[0]
    wfn = hf_wfn(H)
    E_SCF, C = wfn.solve_SCF(parameters, print_level)
    basis = H.basis_set
    E_nuc = H.E_nuc
    E, t1, t2 = _solve(parameters, wfn, print_level)
    E_list = [E_SCF, E, E_nuc]
    T_list = [1, t1, t2]
    if print_level == 2:
        print("Method: ", parameters['method'])
        print("Electronic Hartree-Fock Energy: ", E_SCF)
        if parameters['method'] != 'RHF':
            print("Electronic Post-Hartree-Fock Energy: ", E)
        print("Total Energy: ", E_SCF + E + E_nuc)
        print(parameters)
    if return_H:
        return E_list, T_list, C, basis, H
    return E_list, T_list, C, basis


def phase_corrected_energy(parameters, unperturbed_basis, unperturbed_C, return_H=False, print_level=0):
    """energy.py:84-150: the MO phases are aligned to the unperturbed
    reference (compute_phase) before the correlated solve."""
    H = build_hamiltonians([parameters], device())[0]
    wfn = hf_wfn(H)
    E_SCF, C = wfn.solve_SCF(parameters, print_level)
    basis = H.basis_set
    E_nuc = H.E_nuc
    wfn.C = compute_phase(wfn.ndocc, wfn.nbf, unperturbed_basis, unperturbed_C, basis, C)
    E, t1, t2 = _solve(parameters, wfn, print_level)
    E_list = [E_SCF, E, E_nuc]
    T_list = [1, t1, t2]
    if print_level > 0:
        print("Method: ", parameters['method'])
        print("Electronic Hartree-Fock Energy: ", E_SCF)
        if parameters['method'] != 'RHF':
            print("Electronic Post-Hartree-Fock Energy: ", E)
        print("Total Energy: ", E_SCF + E + E_nuc)
    if return_H:
        return E_list, T_list, wfn.C, basis, H
    return E_list, T_list, wfn.C, basis


def phase_corrected_energies(parameter_list, unperturbed_basis, unperturbed_C, print_level=0):
    """phase_corrected_energy for a list of parameter sets (the displaced
    geometries and fields of fin_diff.py:267-372), with the Hartree-Fock
    part batched: every AO integral set (one- and two-electron) in one
    launch per basis, the SCFs
    iterated together (hf_wfn.solve_SCF_batch, one Fock build per iteration
    for all of them); the phase alignment and the correlated solve then run
    per instance.  Each result equals phase_corrected_energy's.  Returns
    [(E_list, T_list, C, basis)]."""
    Hs = build_hamiltonians(parameter_list, device())
    prime_device_eri(Hs, device())
    wfns = [hf_wfn(H) for H in Hs]
    out = []
    nbf = {w.nbf for w in wfns}
    if len(nbf) == 1:
        scf = solve_SCF_batch(wfns, parameter_list[0], print_level)
    else:
        scf = [w.solve_SCF(p, print_level) for w, p in zip(wfns, parameter_list)]
    for H, wfn, (E_SCF, C) in zip(Hs, wfns, scf):
        wfn.C = compute_phase(wfn.ndocc, wfn.nbf, unperturbed_basis, unperturbed_C, H.basis_set, C)
    corr = _solve_spatial_set(parameter_list, wfns) if print_level == 0 else None
    for q, (p, H, wfn, (E_SCF, C)) in enumerate(zip(parameter_list, Hs, wfns, scf)):
        E, t1, t2 = corr[q] if corr is not None else _solve(p, wfn, print_level)
        out.append(([E_SCF, E, H.E_nuc], [1, t1, t2], wfn.C, H.basis_set))
    return out


_SOLVER_KEYS = ('method', 'e_convergence', 'd_convergence', 'DIIS', 'max_iterations', 'freeze_core')


def _solve_spatial_set(parameter_list, wfns):
    """The CISD / CID solves of a finite-difference set as one
    spatial_ci.SpatialCISet: every instance's MO integrals and Fock matrix
    from its own ci_wfn, the Jacobi + DIIS iterations of all of them in one
    batched spin-adapted step per iteration.  Spin-adapted methods
    (ci_wfn.py:420-574, :51-167): each result equals its own ci_wfn solve
    (printing the same amplitude lines).  Spin-orbital methods (CISD_SO,
    CID_SO, ci_wfn.py:171-416): for the RHF reference the spin-orbital
    equations are solved by the spin-adapted amplitudes expanded to spin
    orbitals (apy_sp_amplitudes_to_so), so the set is solved spin-adapted
    and returned in the spin-orbital layout, equal to the per-instance
    spin-orbital solves to their convergence threshold.  None when the set
    does not qualify (other methods, one instance, differing solver settings
    or sizes)."""
    p0 = parameter_list[0]
    if p0['method'] not in ('CISD', 'CID', 'CISD_SO', 'CID_SO') or len(parameter_list) < 2:
        return None
    if any(any(p.get(k) != p0.get(k) for k in _SOLVER_KEYS) for p in parameter_list):
        return None
    cws = [ci_wfn(p, w) for p, w in zip(parameter_list, wfns)]
    shapes = {(len(c.eps_o), len(c.eps_v)) for c in cws}
    if len(shapes) != 1:
        return None
    singles = p0['method'] in ('CISD', 'CISD_SO')
    cset = SpatialCISet([c._ERI_MO for c in cws], [c._F_MO for c in cws],
                        [np.concatenate([c.eps_o, c.eps_v]) for c in cws], len(cws[0].eps_o), singles=singles)
    E, t1, t2, _ = cset.solve(p0)
    out = []
    if p0['method'].endswith('_SO'):                  # the spin-orbital solves print nothing at print_level 0
        T1, T2 = cset.spin_orbital_amplitudes()
        for q in range(len(cws)):
            t1q = T1[q].cpu().numpy() if singles else 0
            out.append((np.complex128(E[q]), t1q, T2[q].cpu().numpy()))
        return out
    for q in range(len(cws)):
        t2s = t2[q].cpu().numpy()
        print("t-Amplitude Data:")
        if singles:
            t1s = t1[q].cpu().numpy()
            print("Maximum t1: ", np.max(t1s))
            print("Maximum t2: ", np.max(t2s))
            out.append((np.complex128(E[q]), t1s, t2s))
        else:
            print("Maximum t2: ", np.max(t2s))
            out.append((np.complex128(E[q]), 0, t2s))
    return out
