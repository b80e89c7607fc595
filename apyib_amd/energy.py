"""Method dispatch with apyib's interface (energy.py:17-150)."""
from .ci_wfn import ci_wfn
from .hamiltonian import Hamiltonian
from .hf_wfn import hf_wfn
from .mp2_wfn import mp2_wfn
from .utils import compute_phase


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
    H = Hamiltonian(parameters)
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
    H = Hamiltonian(parameters)
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
