"""ctypes binding of the native library ``libapyib_amd.so`` (C ABI in
``include/apyib_amd.h``).

The library is built in-tree by ``apyib_amd/build.py``.  There is no
fallback: if the library is missing, every GPU entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libapyib_amd.so")

APY_OK = 0
APY_E_ARG = -1
APY_E_HIP = -2
APY_E_WORKSPACE = -3
APY_E_SINGULAR = -4
APY_E_TOO_LARGE = -5

_ERR = {
    APY_E_ARG: "invalid argument",
    APY_E_HIP: "HIP runtime error",
    APY_E_WORKSPACE: "workspace too small",
    APY_E_SINGULAR: "singular linear system",
    APY_E_TOO_LARGE: "problem larger than the library's static scratch (APY_NOALLOC_SCRATCH_BYTES)",
}

APY_MK, APY_KM = 0, 1
APY_KN, APY_NK = 0, 1

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_size = ctypes.c_size_t


class GemmDesc(ctypes.Structure):
    _fields_ = [
        ("m", c_i64), ("n", c_i64), ("k", c_i64),
        ("a_layout", c_int), ("b_layout", c_int),
        ("conj_a", c_int), ("conj_b", c_int),
        ("alpha_re", c_dbl), ("alpha_im", c_dbl),
        ("beta_re", c_dbl), ("beta_im", c_dbl),
        ("a", c_vp), ("lda", c_i64), ("stride_a", c_i64),
        ("b", c_vp), ("ldb", c_i64), ("stride_b", c_i64),
        ("c", c_vp), ("ldc", c_i64), ("stride_c", c_i64),
        ("batch", c_i64),
        ("reduce_batch", c_int),
        ("split_k", c_int),
        ("tile", c_int),
        ("gauss3m", c_int),
    ]


class SoInts(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64),
        ("vvvv_p", c_vp), ("ring", c_vp), ("oooo_p", c_vp), ("ovoo_p", c_vp), ("vvvo_p", c_vp),
        ("vovv_p", c_vp), ("ooov", c_vp), ("vvoo_t", c_vp), ("oovv", c_vp),
    ]


SO_BLOCKS = ("vvvv_p", "ring", "oooo_p", "ovoo_p", "vvvo_p", "vovv_p", "ooov", "vvoo_t", "oovv")

SP_BLOCKS = ("vvvs", "vvva", "oooo", "w21", "w3", "vvoo", "l", "vvvo", "vooo", "s1", "w4", "w5", "w4t")


class SpInts(ctypes.Structure):
    _fields_ = [("o", c_i64), ("v", c_i64)] + [(b, c_vp) for b in SP_BLOCKS]


APY_GTO_MAXPRIM = 6
APY_GTO_MAXL = 3


class Gto(ctypes.Structure):
    _fields_ = [
        ("atom", ctypes.c_int32), ("nprim", ctypes.c_int32), ("nterm", ctypes.c_int32),
        ("lmn", (ctypes.c_int32 * 3) * 2), ("pad", ctypes.c_int32),
        ("exps", c_dbl * APY_GTO_MAXPRIM), ("coefs", (c_dbl * APY_GTO_MAXPRIM) * 2),
    ]


APY_OEI_OVERLAP, APY_OEI_KINETIC, APY_OEI_POTENTIAL, APY_OEI_DIPOLE, APY_OEI_ANGMOM = 1, 2, 4, 8, 16


class OeiArgs(ctypes.Structure):
    _fields_ = [
        ("nset", c_i64), ("fa", c_vp), ("na", c_i64), ("ca", c_vp), ("nca", c_i64),
        ("fb", c_vp), ("nb", c_i64), ("cb", c_vp), ("ncb", c_i64),
        ("nuclei", c_vp), ("charges", c_vp), ("nnuc", c_i64), ("origin", c_dbl * 3),
        ("which", ctypes.c_int32), ("pad", ctypes.c_int32), ("out", c_vp),
    ]


class CisdSpArgs(ctypes.Structure):
    _fields_ = [
        ("o", c_i64), ("v", c_i64), ("fock", c_vp), ("eps", c_vp), ("ints", ctypes.POINTER(SpInts)),
        ("hoo", c_vp), ("g", c_vp), ("singles", c_int), ("energy", c_vp), ("t", c_vp), ("t_new", c_vp),
        ("out", c_vp), ("resp_b", c_vp), ("resp_t", c_vp), ("resp_energy", c_vp),
        ("resp_shift_re", c_dbl), ("resp_shift_im", c_dbl), ("w21_f", c_vp), ("gauss3m", c_int),
    ]

APY_NOALLOC_SCRATCH_BYTES = 256 << 20

APY_MODE_RESIDUAL = 0
APY_MODE_SIGMA = 1
APY_MAXVEC = 16


class CisdSoArgs(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64),
        ("fock", c_vp), ("eps", c_vp),
        ("ints", ctypes.POINTER(SoInts)),
        ("hoo", c_vp),
        ("singles", c_int), ("mode", c_int), ("fov_coupling", c_int),
        ("energy", c_vp),
        ("c0_re", c_dbl), ("c0_im", c_dbl),
        ("t", c_vp), ("t_new", c_vp), ("out", c_vp),
        ("ev_begin", c_vp), ("ev_end", c_vp), ("c0_dev", c_vp),
        ("ring_f", c_vp), ("packed", c_int), ("gauss3m", c_int),
    ]


class CisdSoSigmaBatchArgs(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64),
        ("nb", c_int), ("singles", c_int), ("fov_coupling", c_int),
        ("ints", ctypes.POINTER(SoInts)),
        ("fock", ctypes.POINTER(c_vp)), ("hoo", ctypes.POINTER(c_vp)),
        ("c0", ctypes.POINTER(c_dbl)),
        ("t", ctypes.POINTER(c_vp)), ("out", ctypes.POINTER(c_vp)),
        ("ev_begin", c_vp), ("ev_end", c_vp), ("gauss3m", c_int),
    ]


class CisdSoFockSigmaArgs(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64),
        ("nf", c_int),
        ("fock", ctypes.POINTER(c_vp)), ("hoo", ctypes.POINTER(c_vp)),
        ("t", c_vp), ("out", ctypes.POINTER(c_vp)),
    ]


class Mp2Instance(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64),
        ("eps", c_vp), ("vvoo_t", c_vp), ("oovv", c_vp), ("t2", c_vp),
    ]


class Mp2NcArgs(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64),
        ("fock", c_vp), ("eps", c_vp),
        ("transpose_f", c_int), ("drv_scale", c_dbl),
        ("drv", c_vp), ("x", c_vp), ("x_new", c_vp), ("rnorm", c_vp),
    ]


class Mp2AdjointItem(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64), ("t", c_vp), ("lam", c_vp), ("grad_f", c_vp), ("ndir", c_i64),
        ("dF", ctypes.POINTER(c_vp)), ("dW", ctypes.POINTER(c_vp)), ("dV", ctypes.POINTER(c_vp)),
    ]


class Mp2NcInstance(ctypes.Structure):
    _fields_ = [
        ("O", c_i64), ("V", c_i64),
        ("fock", c_vp), ("eps", c_vp), ("drv", c_vp), ("x", c_vp), ("x_new", c_vp), ("w", c_vp),
        ("transpose_f", c_int), ("drv_scale", c_dbl),
    ]


APY_AAT_NTERMS = 20


class AatAxpbyOp(ctypes.Structure):
    _fields_ = [
        ("n", c_i64), ("a", c_vp), ("b", c_vp), ("out", c_vp),
        ("ca", c_dbl), ("cb", c_dbl), ("conj", c_int),
    ]


class AatJob(ctypes.Structure):
    _fields_ = [
        ("gamma", c_i64), ("b1", c_vp), ("b2", c_vp), ("k1", c_vp), ("k2", c_vp),
    ]


_P_I64 = ctypes.POINTER(c_i64)
_P_DBL = ctypes.POINTER(c_dbl)
_P_VP = ctypes.POINTER(c_vp)

# name -> (restype, argtypes); every symbol here is declared in include/apyib_amd.h
SIGNATURES = {
    "apy_version": (c_int, []),
    "apy_zgemm_workspace": (c_size, [ctypes.POINTER(GemmDesc)]),
    "apy_zgemm": (c_int, [ctypes.POINTER(GemmDesc), c_vp, c_size, c_vp]),
    "apy_zgemm_plan": (c_int, [ctypes.POINTER(GemmDesc), ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "apy_zgemm_tune": (c_int, [ctypes.POINTER(GemmDesc), c_vp, c_size, c_int, c_vp, ctypes.POINTER(c_int),
                               ctypes.POINTER(c_int), ctypes.POINTER(c_dbl)]),
    "apy_zgemm_tune_clear": (c_int, []),
    "apy_zgemm_tune_set": (c_int, [ctypes.POINTER(GemmDesc), c_int, c_int]),
    "apy_zgemm_tune_export": (c_int, [_P_I64, c_i64, _P_I64]),
    "apy_so_ints_sizes": (c_int, [c_i64, c_i64, _P_I64]),
    "apy_so_ints_from_full": (c_int, [c_i64, c_vp, ctypes.POINTER(SoInts), c_vp]),
    "apy_so_ints_from_mo": (c_int, [c_i64, c_vp, ctypes.POINTER(SoInts), c_vp]),
    "apy_synth_so_ints": (c_int, [ctypes.c_uint64, ctypes.c_uint64, c_dbl, c_int, c_dbl, c_int,
                                  ctypes.POINTER(SoInts), c_vp]),
    "apy_so_fock_from_mo": (c_int, [c_i64, c_vp, c_vp, c_vp]),
    "apy_cisd_so_pair_operator": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "apy_cisd_so_ring_fock": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "apy_cisd_so_workspace": (c_size, [c_i64, c_i64, c_int]),
    "apy_cisd_so_step": (c_int, [ctypes.POINTER(CisdSoArgs), c_vp, c_size, c_vp]),
    "apy_cisd_so_step_noalloc": (c_int, [ctypes.POINTER(CisdSoArgs), c_vp]),
    "apy_cisd_so_noalloc_bytes": (c_size, [ctypes.POINTER(CisdSoArgs)]),
    "apy_cisd_so_tune": (c_int, [ctypes.POINTER(CisdSoArgs), c_vp, c_size, c_int, c_vp]),
    "apy_cisd_so_sigma_batch_workspace": (c_size, [c_i64, c_i64, c_int, c_int]),
    "apy_cisd_so_sigma_batch": (c_int, [ctypes.POINTER(CisdSoSigmaBatchArgs), c_vp, c_size, c_vp]),
    "apy_cisd_so_fock_sigma_workspace": (c_size, [c_i64, c_i64]),
    "apy_cisd_so_fock_sigma": (c_int, [ctypes.POINTER(CisdSoFockSigmaArgs), c_vp, c_size, c_vp]),
    "apy_cisd_so_energy": (c_int, [c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_cisd_so_energy_packed": (c_int, [c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_zdot_many": (c_int, [c_int, _P_VP, c_vp, c_i64, c_int, c_vp, c_vp, c_size, c_vp]),
    "apy_zdiffsq_segments": (c_int, [c_vp, c_vp, c_int, _P_I64, c_vp, c_vp, c_size, c_vp]),
    "apy_zcombine": (c_int, [c_int, _P_VP, _P_DBL, c_i64, c_vp, c_vp]),
    "apy_reduce_workspace": (c_size, [c_i64, c_int]),
    "apy_eri_gto": (c_int, [c_i64, c_i64, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "apy_eri_deriv_assemble": (c_int, [c_i64, c_vp, c_int, c_vp, c_vp, c_vp]),
    "apy_oei_components": (c_int, [c_int]),
    "apy_diis_slot": (c_int, [c_int, c_int, ctypes.c_uint32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "apy_zcombine_guarded": (c_int, [c_int, _P_VP, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "apy_iter_check": (c_int, [c_vp, c_vp, c_int, c_dbl, c_dbl, c_int, c_vp, c_vp, c_int, c_vp, c_vp]),
    "apy_oei_gto": (c_int, [ctypes.POINTER(OeiArgs), c_vp]),
    "apy_reduce_seg_workspace": (c_size, [c_i64, c_int]),
    "apy_gs_coefficients": (c_int, [c_int, c_vp, c_int, c_vp, c_vp, c_vp]),
    "apy_ci_packed_length": (c_i64, [c_i64, c_i64]),
    "apy_ci_pack": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp]),
    "apy_ci_unpack": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp]),
    "apy_ci_precondition_packed": (c_int, [c_i64, c_i64, c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "apy_ci_precondition_packed_dev": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "apy_heev_lowest": (c_int, [c_int, c_vp, c_int, c_vp, c_vp]),
    "apy_davidson_check": (c_int, [c_vp, c_vp, c_dbl, c_dbl, c_vp, c_vp, c_int, c_vp]),
    "apy_davidson_fused_workspace": (c_size, []),
    "apy_peak_mfma_f64_flops": (c_dbl, [c_int, c_int]),
    "apy_peak_mfma_f64": (c_int, [c_int, c_int, c_vp, c_vp]),
    "apy_peak_copy": (c_int, [c_i64, c_vp, c_vp, c_int, c_vp]),
    "apy_peak_read": (c_int, [c_i64, c_vp, c_vp, c_int, c_vp]),
    "apy_davidson_residual": (c_int, [c_int, _P_VP, _P_VP, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_davidson_combine_dots": (c_int, [c_int, _P_VP, c_vp, c_i64, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_zdot_many_seg": (c_int, [c_int, _P_VP, c_vp, c_i64, c_i64, c_int, c_vp, c_vp, c_size, c_vp]),
    "apy_zdiffsq_batch": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_size, c_vp]),
    "apy_zcombine_seg": (c_int, [c_int, _P_VP, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "apy_mp2_so_batched_workspace": (c_size, [c_int, ctypes.POINTER(Mp2Instance)]),
    "apy_mp2_so_batched": (c_int, [c_int, ctypes.POINTER(Mp2Instance), c_vp, c_vp, c_size, c_vp]),
    "apy_mp2_so": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_mp2_spatial_workspace": (c_size, []),
    "apy_mp2_spatial": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_mp2_nc_workspace": (c_size, [c_i64, c_i64]),
    "apy_mp2_nc_step": (c_int, [ctypes.POINTER(Mp2NcArgs), c_vp, c_size, c_vp]),
    "apy_mp2_nc_density_workspace": (c_size, [c_i64, c_i64]),
    "apy_mp2_nc_densities": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_mp2_adjoint_sums_workspace": (c_size, [c_i64, ctypes.POINTER(Mp2AdjointItem)]),
    "apy_mp2_adjoint_sums": (c_int, [c_i64, ctypes.POINTER(Mp2AdjointItem), c_vp, c_vp, c_size, c_vp]),
    "apy_eri_mo_workspace": (c_size, [c_i64, c_i64]),
    "apy_eri_mo": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_fock_build": (c_int, [c_i64, c_vp, c_vp, c_vp, c_dbl, c_vp, c_vp]),
    "apy_fock_build_batch": (c_int, [c_i64, c_i64, c_vp, c_vp, c_i64, c_vp, c_dbl, c_vp, c_vp]),
    "apy_eri_so_from_mo": (c_int, [c_i64, c_vp, c_vp, c_vp]),
    "apy_mo_overlap_workspace": (c_size, [c_i64, c_i64, c_i64]),
    "apy_mo_overlap": (c_int, [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_size,
                               c_vp]),
    "apy_phase_align": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "apy_t2_packed_length": (c_i64, [c_i64, c_i64]),
    "apy_t2_pack": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp]),
    "apy_t2_unpack": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp]),
    "apy_mp2_nc_batched_workspace": (c_size, [c_int, ctypes.POINTER(Mp2NcInstance)]),
    "apy_mp2_nc_batched_setup": (c_int, [c_int, ctypes.POINTER(Mp2NcInstance), c_vp, c_size, _P_I64, c_vp]),
    "apy_mp2_nc_batched_step": (c_int, [c_int, c_i64, c_int, c_vp, c_vp, c_vp]),
    "apy_mp2_nc_batched_solve": (c_int, [c_int, c_i64, c_dbl, c_int, c_vp, c_vp, c_vp, c_vp]),
    "apy_mp2_nc_batched_energy": (c_int, [c_int, c_i64, c_int, c_dbl, c_vp, c_vp, c_vp]),
    "apy_ci_precondition": (c_int, [c_i64, c_i64, c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "apy_cisd_so_opd_workspace": (c_size, [c_i64, c_i64]),
    "apy_cisd_so_opd": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_ci_dots": (c_int, [c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_sp_ints_sizes": (c_int, [c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "apy_sp_ints_from_mo": (c_int, [c_i64, c_vp, ctypes.POINTER(SpInts), c_vp]),
    "apy_cisd_sp_prepare": (c_int, [c_i64, c_i64, c_vp, ctypes.POINTER(SpInts), c_vp, c_vp, c_vp]),
    "apy_cisd_sp_workspace": (c_size, [c_i64, c_i64, c_int]),
    "apy_cisd_sp_step": (c_int, [ctypes.POINTER(CisdSpArgs), c_vp, c_size, c_vp]),
    "apy_cisd_sp_energy_workspace": (c_size, []),
    "apy_cisd_sp_energy": (c_int, [c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_cisd_sp_batch_workspace": (c_size, [c_i64, c_i64, c_int, c_i64]),
    "apy_cisd_sp_step_batch": (c_int, [ctypes.POINTER(CisdSpArgs), c_i64, c_vp, c_vp, c_size, c_vp]),
    "apy_cisd_sp_energy_batch": (c_int, [c_i64, c_i64, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_cisd_sp_multi_workspace": (c_size, [c_i64, c_i64, c_int, c_i64]),
    "apy_cisd_sp_step_multi": (c_int, [ctypes.POINTER(CisdSpArgs), c_i64, c_vp, c_size, c_vp]),
    "apy_cisd_sp_tune": (c_int, [ctypes.POINTER(CisdSpArgs), c_i64, c_int, c_vp, c_size, c_int, c_vp]),
    "apy_cisd_sp_fold_fock": (c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "apy_sp_amplitudes_to_so": (c_int, [c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp]),
    "apy_mp2_sp_response_workspace": (c_size, [c_i64, c_i64, c_i64]),
    "apy_mp2_sp_response": (c_int, [c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_mp2_sp_densities_workspace": (c_size, [c_i64, c_i64]),
    "apy_mp2_sp_densities": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_mp2_sp_directional_workspace": (c_size, [c_i64]),
    "apy_mp2_sp_directional": (c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_size,
                                       c_vp]),
    "apy_cisd_sp_energy_multi": (c_int, [c_i64, c_i64, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "apy_cisd_so_tpd_workspace": (c_size, [c_i64, c_i64]),
    "apy_cisd_so_tpd": (c_int, [c_i64, c_i64, c_vp, ctypes.POINTER(SoInts), c_vp, c_size, c_vp]),
    "apy_so_ints_grad_to_full": (c_int, [c_i64, ctypes.POINTER(SoInts), c_vp, c_vp]),
    "apy_aat_gamma_record": (c_i64, [c_i64, c_i64, c_i64]),
    "apy_aat_gamma_workspace": (c_size, [c_i64, c_i64, c_i64]),
    "apy_aat_gamma": (c_int, [c_i64, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_size, c_vp]),
    "apy_aat_axpby_workspace": (c_size, [c_i64]),
    "apy_aat_axpby": (c_int, [c_i64, ctypes.POINTER(AatAxpbyOp), c_vp, c_size, c_vp]),
    "apy_aat_terms_workspace": (c_size, [c_i64, c_i64, c_i64]),
    "apy_aat_terms": (c_int, [c_i64, c_i64, c_vp, c_i64, c_i64, c_i64, ctypes.POINTER(AatJob), c_vp, c_vp, c_size,
                               c_vp]),
    "apy_aat_norms_workspace": (c_size, [c_i64]),
    "apy_aat_norms": (c_int, [c_i64, c_i64, c_i64, _P_VP, _P_VP, c_vp, c_vp, c_size, c_vp]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def load():
    """Load the native library (once).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError(
            f"apyib_amd native library not found at {LIB_PATH}; "
            "build it with `python -m apyib_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what):
    if status != APY_OK:
        raise NativeError(f"{what} failed: {_ERR.get(status, status)} (status {status})")
