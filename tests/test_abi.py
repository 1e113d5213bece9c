"""The C-ABI library loads and exports every symbol include/ygpu.h declares
(CPU only: no compute call is made without a GPU)."""
import ctypes
import re
from pathlib import Path

from yelmo_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "ygpu.h").read_text()
    return sorted(set(re.findall(r"\b(ygpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_python_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_default_params_mirror_yelmo_defaults():
    lib = _lib.load()
    p = _lib.SsaParamsC()
    lib.ygpu_default_params(ctypes.byref(p))
    # input/yelmo_defaults.nml:140-173 and input/yelmo_phys_const.nml:6-10
    assert p.visc_method == 1 and p.beta_method == 1 and p.beta_gl_stag == 1
    assert p.ssa_iter_max == 20 and abs(p.ssa_iter_rel - 0.7) < 1e-12
    assert p.beta_min == 100.0 and p.eps_0 == 1e-6 and p.ssa_vel_max == 1e4
    assert p.rho_ice == 910.0 and p.rho_sw == 1028.0 and p.g == 9.81
    # ssa_solver = "energy" with ssa_lis_opt_energy (input/yelmo_defaults.nml:159-161)
    assert p.ssa_solver == 1
    assert p.ssa_lis_opt.decode() == "-i cg -p jacobi -maxiter 200 -tol 1.0e-2 -initx_zeros false"
    assert ctypes.sizeof(_lib.SsaParamsC) == 4 * 10 + 8 * 15 + 4 * 2 + 256


def test_footprint_grows_with_size():
    lib = _lib.load()
    p = _lib.SsaParamsC()
    lib.ygpu_default_params(ctypes.byref(p))
    small = lib.ygpu_footprint_bytes(ctypes.byref(p), 81, 81, 11)
    big = lib.ygpu_footprint_bytes(ctypes.byref(p), 161, 161, 11)
    assert 3.5 < big / small < 4.5


def test_create_without_gpu_fails_cleanly():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.ygpu_create(0, ctypes.byref(h))
    # on a machine without a GPU the call reports a device error, it never aborts
    assert rc in (0, _lib.YGPU_ERR_DEVICE)
    if rc == 0:
        lib.ygpu_destroy(h)
