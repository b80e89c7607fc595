"""AAT tensor driver with apyib's interface (parallel.py:16-48).

compute_parallel_aats(parameters, nuc_pert_strength, mag_pert_strength,
normalization='full', num_processes=4) -> [3N, 3] array.  The reference
maps the 9N elements over a process pool; here every element is one job of
a single batched GPU evaluation (AAT.compute_aats), so num_processes is
accepted for signature compatibility and not used.
"""
from .aats import AAT
from .energy import energy
from .fin_diff import finite_difference
from .hf_wfn import hf_wfn


def compute_parallel_aats(parameters, nuc_pert_strength, mag_pert_strength, normalization='full', num_processes=4):
    E_list, T_list, C, basis, H = energy(parameters, return_H=True)
    print("Total Energy: ", E_list[0] + E_list[1] + E_list[2])
    wfn = hf_wfn(H)    # parallel.py:19-20 builds the same Hamiltonian a second time
    fd = finite_difference(parameters, basis, C)
    (nuc_pos_C, nuc_neg_C, nuc_pos_basis, nuc_neg_basis, nuc_pos_T, nuc_neg_T, mag_pos_C, mag_neg_C, mag_pos_basis,
     mag_neg_basis, mag_pos_T, mag_neg_T) = fd.compute_AAT(nuc_pert_strength, mag_pert_strength)
    aats = AAT(parameters, wfn, C, basis, T_list, nuc_pos_C, nuc_neg_C, nuc_pos_basis, nuc_neg_basis, nuc_pos_T,
               nuc_neg_T, mag_pos_C, mag_neg_C, mag_pos_basis, mag_neg_basis, mag_pos_T, mag_neg_T,
               nuc_pert_strength, mag_pert_strength)
    I = aats.compute_aats(normalization)
    print(I, "\n")
    return I
