"""The C-ABI library loads without a GPU and exports every function
include/apyib_amd.h declares; the ctypes table matches the header.  No
compute calls (CPU only)."""
import ctypes
import os
import re

import pytest

from apyib_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "apyib_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(apy_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_functions():
    names = declared_functions()
    assert "apy_zgemm" in names and "apy_cisd_so_step" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_load_binds_signatures():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.apy_version() >= 100
    sizes = (ctypes.c_int64 * 9)()
    assert lib.apy_so_ints_sizes(20, 160, sizes) == 0
    Po, Pv = 190, 12720
    assert list(sizes) == [Pv * Pv, 20 * 160 * 20 * 160, Po * Po, 20 * 160 * Po, Pv * 20 * 160,
                           160 * 20 * Pv, 20 * 20 * 20 * 160, 20 * 20 * 160 * 160, 20 * 20 * 160 * 160]
    assert lib.apy_so_ints_sizes(-1, 3, sizes) == _lib.APY_E_ARG


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(_lib.NativeError):
        _lib.load()


def test_struct_layouts_match_header_sizes():
    # int64 x3 + int x4 + double x4 + (ptr,int64,int64) x3 + int64 + int x3 (padded)
    assert ctypes.sizeof(_lib.GemmDesc) == 8 * 3 + 4 * 4 + 8 * 4 + 24 * 3 + 8 + 4 * 3 + 4
    assert ctypes.sizeof(_lib.SoInts) == 8 * 2 + 8 * 9


def test_spatial_sizes_and_struct_layouts():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    o, v = 5, 9
    ov = o * v
    sizes = (ctypes.c_int64 * 13)()
    assert lib.apy_sp_ints_sizes(o, v, sizes) == 0
    sv, av = v * (v + 1) // 2, v * (v - 1) // 2
    assert list(sizes) == [sv * sv, av * av, o ** 4, 2 * ov * ov, ov * ov, ov * ov, ov * ov, v * v * o * v,
                           v * o ** 3, ov * ov, v * o * v * v, o ** 3 * v, v * v * o * v]
    assert lib.apy_sp_ints_sizes(0, v, sizes) == _lib.APY_E_ARG
    assert ctypes.sizeof(_lib.SpInts) == 8 * 2 + 8 * 13
    # o, v, 5 pointers, int (padded), 4 pointers, 3 response pointers, 2 doubles, folded ring operator,
    # three-product flag (padded)
    assert ctypes.sizeof(_lib.CisdSpArgs) == 8 * 2 + 8 * 5 + 8 + 8 * 4 + 8 * 3 + 8 * 2 + 8 + 8


def test_argument_errors_return_before_any_device_work():
    """Null or inconsistent arguments come back as APY_E_ARG (or
    APY_E_WORKSPACE) from the host-side checks, so these calls are safe
    without a GPU."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    E = _lib.APY_E_ARG
    assert lib.apy_cisd_sp_step(None, None, 0, None) == E
    assert lib.apy_cisd_so_step(None, None, 0, None) == E
    assert lib.apy_zgemm(None, None, 0, None) == E
    assert lib.apy_sp_ints_from_mo(10, None, None, None) == E
    assert lib.apy_cisd_so_tpd(2, 3, None, None, None, 0, None) == E
    assert lib.apy_ci_dots(2, 3, 0, None, None, None, None, 0, None) == E
    assert lib.apy_ci_dots(2, 3, _lib.APY_MAXVEC + 1, None, None, None, None, 0, None) == E
    assert lib.apy_cisd_sp_energy(0, 3, 1, None, None, None, None, None, 0, None) == E
    # a spatial args struct whose integrals disagree with (o, v)
    ints = _lib.SpInts(3, 4)
    args = _lib.CisdSpArgs(3, 5, 1, 1, ctypes.pointer(ints), 1, 1, 1, 1, 1, 2, 3)
    assert lib.apy_cisd_sp_step(ctypes.byref(args), None, 0, None) == E
    # a consistent struct whose response term has no reference vector, or
    # writes over one of its own inputs
    ints = _lib.SpInts(3, 5, *([1] * len(_lib.SP_BLOCKS)))
    for rb, rt in ((7, None), (7, 2), (3, 9), (9, 3)):
        args = _lib.CisdSpArgs(3, 5, 1, 1, ctypes.pointer(ints), 1, 1, 1, 1, 1, 2, 3, rb, rt)
        assert lib.apy_cisd_sp_step(ctypes.byref(args), None, 0, None) == E


def test_gto_table_layout_and_argument_errors():
    """apy_gto is 184 bytes with the exponents at byte 40 and the
    coefficients at 88 (the header's layout); null tables are rejected
    before any device work."""
    assert ctypes.sizeof(_lib.Gto) == 184
    assert _lib.Gto.exps.offset == 40 and _lib.Gto.coefs.offset == 88
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.apy_eri_gto(1, 1, None, None, 1, None, 1, None, 1, None, 1, None, None) == _lib.APY_E_ARG
    assert lib.apy_eri_gto(0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, None) == _lib.APY_E_ARG
    assert lib.apy_eri_deriv_assemble(4, None, 0, None, None, None) == _lib.APY_E_ARG
    assert lib.apy_eri_deriv_assemble(4, 1, 0, 8, 8, None) == _lib.APY_E_ARG


def test_one_electron_args_layout_and_errors():
    """apy_oei_args is 136 bytes (header); the component count follows the
    selection bits; invalid selections, empty sets, null tables and a
    potential request without nuclei are rejected before any device work."""
    assert ctypes.sizeof(_lib.OeiArgs) == 136
    assert _lib.OeiArgs.origin.offset == 96 and _lib.OeiArgs.out.offset == 128
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert [lib.apy_oei_components(w) for w in (0, 1, 2, 4, 8, 16, 31, 32, -1)] == [0, 1, 1, 1, 3, 3, 9, 0, 0]
    good = dict(nset=1, fa=8, na=1, ca=8, nca=1, fb=8, nb=1, cb=8, ncb=1, nuclei=8, charges=8, nnuc=1, which=31,
                out=8)
    for bad in (dict(which=0), dict(which=32), dict(nset=0), dict(fa=None), dict(na=0), dict(cb=None),
                dict(ncb=0), dict(out=None), dict(nuclei=None), dict(charges=None), dict(nnuc=0)):
        a = _lib.OeiArgs(**{**good, **bad})
        assert lib.apy_oei_gto(ctypes.byref(a), None) == _lib.APY_E_ARG, bad
    assert lib.apy_oei_gto(None, None) == _lib.APY_E_ARG


def test_device_iteration_control_argument_errors():
    """apy_diis_slot, apy_zcombine_guarded and apy_iter_check reject bad
    sizes, a slot outside its window and null pointers before any device
    work."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    ok = 8
    assert lib.apy_diis_slot(0, 0, 1, ok, ok, ok, ok, None) == _lib.APY_E_ARG
    assert lib.apy_diis_slot(17, 0, 1, ok, ok, ok, ok, None) == _lib.APY_E_ARG
    assert lib.apy_diis_slot(8, 8, 0xFF, ok, ok, ok, ok, None) == _lib.APY_E_ARG
    assert lib.apy_diis_slot(8, 3, 0x07, ok, ok, ok, ok, None) == _lib.APY_E_ARG      # slot not in the window
    assert lib.apy_diis_slot(8, 0, 0x1FF, ok, ok, ok, ok, None) == _lib.APY_E_ARG     # bit beyond nslot
    assert lib.apy_diis_slot(8, 0, 0x01, None, ok, ok, ok, None) == _lib.APY_E_ARG
    assert lib.apy_diis_slot(8, 0, 0x01, ok, ok, ok, None, None) == _lib.APY_E_ARG
    ptrs = (ctypes.c_void_p * 2)(8, 8)
    assert lib.apy_zcombine_guarded(0, ptrs, ok, 4, ok, ok, None) == _lib.APY_E_ARG
    assert lib.apy_zcombine_guarded(2, ptrs, ok, 4, ok, None, None) == _lib.APY_E_ARG
    assert lib.apy_zcombine_guarded(2, ptrs, ok, 0, ok, ok, None) == _lib.APY_E_ARG
    assert lib.apy_iter_check(ok, ok, 1, 1e-12, 1e-12, 10, ok, ok, 0, ok, None) == _lib.APY_E_ARG
    assert lib.apy_iter_check(ok, None, 1, 1e-12, 1e-12, 10, ok, ok, 4, ok, None) == _lib.APY_E_ARG
    assert lib.apy_iter_check(ok, ok, 1, 1e-12, 1e-12, 10, ok, ok, 4, None, None) == _lib.APY_E_ARG


def test_model_plan_repeats_without_deadlock():
    """apy_zgemm_plan of one shape three times (the second and later calls
    read the cost model's memoized choice): the same plan each time, in a
    child process with a time limit so that a lock taken twice fails the
    test instead of hanging the suite.  Host-only (the device queries fall
    back to defaults without a GPU)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    import subprocess
    import sys
    code = (
        "import ctypes\n"
        "from apyib_amd import _lib\n"
        "lib = _lib.load()\n"
        "buf = (ctypes.c_byte * 64)()\n"
        "p = ctypes.addressof(buf) + 16 - ctypes.addressof(buf) % 16\n"
        "d = _lib.GemmDesc(64, 96, 40, _lib.APY_MK, _lib.APY_KN, 0, 0, 1.0, 0.0, 0.0, 0.0, p, 40, 0, p, 96, 0, p, 96,"
        " 0, 1, 0, 0, 0, 0)\n"
        "out = []\n"
        "for _ in range(3):\n"
        "    t, s = ctypes.c_int(), ctypes.c_int()\n"
        "    assert lib.apy_zgemm_plan(ctypes.byref(d), ctypes.byref(t), ctypes.byref(s)) == 0\n"
        "    out.append((t.value, s.value))\n"
        "assert out[0] == out[1] == out[2] and out[0][0] >= 1, out\n"
        "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
