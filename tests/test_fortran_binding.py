"""The Fortran ISO_C_BINDING module compiles with amdflang and binds libygpu.so.

A small Fortran program uses yelmo_amd/fortran/ygpu_mod.f90 the way a Yelmo
build would: it fills the parameter struct through ygpu_default_params,
checks the &ydyn defaults (input/yelmo_defaults.nml:140-173) arrive intact in
the bind(C) type, sets an option string, and, without a GPU, expects
ygpu_create to return an error code instead of aborting.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

PROGRAM = r"""
program t
    use, intrinsic :: iso_c_binding
    use ygpu_mod
    implicit none
    type(ygpu_ssa_params) :: par
    type(c_ptr) :: h
    integer(c_int) :: rc
    call ygpu_default_params(par)
    if (par%visc_method /= 1 .or. par%beta_gl_stag /= 1) stop 2
    if (abs(par%ssa_iter_rel - 0.7d0) > 1d-12 .or. abs(par%rho_sw - 1028d0) > 1d-9) stop 3
    if (par%ssa_lis_opt(1) /= '-' .or. par%ssa_lis_opt(2) /= 'i') stop 4
    call ygpu_set_lis_opt(par, "-i bicgstab -p jacobi -maxiter 50 -tol 1.0e-6")
    if (par%ssa_lis_opt(14) /= 'p' .or. par%ssa_lis_opt(46) /= c_null_char) stop 5
    h = c_null_ptr
    rc = ygpu_create(0, h)
    if (rc == YGPU_OK) then
        rc = ygpu_destroy(h)
        print '(a)', 'created'
    else
        print '(a,i0)', 'no device ', rc
    end if
end program
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_fortran_module_binds_library(tmp_path):
    lib = os.path.join(ROOT, "yelmo_amd", "libygpu.so")
    if not os.path.exists(lib):
        pytest.skip("libygpu.so not built")
    src = os.path.join(ROOT, "yelmo_amd", "fortran", "ygpu_mod.f90")
    prog = tmp_path / "t.f90"
    prog.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run([FC, "-J", str(tmp_path), "-o", str(exe), src, str(prog), lib,
                    f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True, cwd=tmp_path,
                   capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout.strip()
    assert out == "created" or out.startswith("no device -2"), out
