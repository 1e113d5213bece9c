"""A Fortran program drives a solve on the GPU through yelmo_amd/fortran/ygpu_mod.f90.

The program is what a Yelmo build would do at the inner boundary
(linear_solver_matrix_solve, solver_linear.F90:151): it fills 1-based CSR
arrays of a five-point system in the reference's ordering n = (i-1) ny + j
(here a diagonally dominant advection-like operator on a 30 x 20 grid),
calls ygpu_csr_solve with an LIS option string, and checks the returned
solution against the system itself: ||b - A x|| / ||b|| below the requested
tolerance, LIS status 0, a positive iteration count, and an empty
substitution note.  A second call with "-i minres" must succeed and leave
the note that says it ran BiCGSTAB.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

pytestmark = pytest.mark.gpu

PROGRAM = r"""
program t
    use, intrinsic :: iso_c_binding
    use ygpu_mod
    implicit none
    integer, parameter :: nx = 30, ny = 20, n = nx * ny
    integer(c_int32_t) :: a_ptr(n + 1), a_idx(5 * n)
    real(c_double) :: a_val(5 * n), b(n), x(n), r(n)
    integer(c_int32_t) :: iters, status, nnz
    real(c_double) :: rel, secs, rn, bn
    type(c_ptr) :: h
    integer(c_int) :: rc
    integer :: i, j, k, row, col
    ! five-point operator: diagonal 4.2 + cell-dependent part, -1 (west),
    ! -0.9 (east), -1.1 (south), -0.8 (north); no wrap (zero rows at borders)
    k = 0
    a_ptr(1) = 1
    do i = 1, nx
        do j = 1, ny
            row = (i - 1) * ny + j
            if (j > 1) then
                k = k + 1; a_idx(k) = row - 1; a_val(k) = -1.1d0
            end if
            if (i > 1) then
                k = k + 1; a_idx(k) = row - ny; a_val(k) = -1.0d0
            end if
            k = k + 1; a_idx(k) = row; a_val(k) = 4.2d0 + 0.01d0 * mod(i * 7 + j * 3, 11)
            if (i < nx) then
                k = k + 1; a_idx(k) = row + ny; a_val(k) = -0.9d0
            end if
            if (j < ny) then
                k = k + 1; a_idx(k) = row + 1; a_val(k) = -0.8d0
            end if
            a_ptr(row + 1) = k + 1
            b(row) = sin(0.3d0 * i) + cos(0.2d0 * j)
        end do
    end do
    nnz = k
    rc = ygpu_create(0, h)
    if (rc /= YGPU_OK) stop 10
    x = 0.0d0
    rc = ygpu_csr_solve(h, int(n, c_int32_t), nnz, a_ptr, a_idx, a_val, b, x, &
                        "-i bicgstab -p jacobi -maxiter 1000 -tol 1.0e-10 -initx_zeros false" // c_null_char, &
                        iters, status, rel, secs)
    if (rc /= YGPU_OK) then
        print '(a)', ygpu_error(h)
        stop 11
    end if
    if (status /= 0 .or. iters <= 0) stop 12
    if (len(ygpu_note(h)) /= 0) stop 13
    ! residual in Fortran
    do row = 1, n
        r(row) = b(row)
        do k = a_ptr(row), a_ptr(row + 1) - 1
            col = a_idx(k)
            r(row) = r(row) - a_val(k) * x(col)
        end do
    end do
    rn = sqrt(sum(r * r))
    bn = sqrt(sum(b * b))
    print '(a,i0,a,es10.3,a,es10.3)', 'iters ', iters, ' rel ', rel, ' check ', rn / bn
    if (rn / bn > 1.0d-9) stop 14
    x = 0.0d0
    rc = ygpu_csr_solve(h, int(n, c_int32_t), nnz, a_ptr, a_idx, a_val, b, x, &
                        "-i minres -p jacobi -maxiter 1000 -tol 1.0e-10" // c_null_char, &
                        iters, status, rel, secs)
    if (rc /= YGPU_OK .or. status /= 0) stop 15
    if (ygpu_note(h) /= "-i minres solved with BiCGSTAB") stop 16
    rc = ygpu_destroy(h)
    print '(a)', 'ok'
end program
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed on this machine")
def test_fortran_program_solves_on_gpu(tmp_path):
    lib = os.path.join(ROOT, "yelmo_amd", "libygpu.so")
    assert os.path.exists(lib), "libygpu.so not built"
    src = os.path.join(ROOT, "yelmo_amd", "fortran", "ygpu_mod.f90")
    prog = tmp_path / "t.f90"
    prog.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run([FC, "-J", str(tmp_path), "-o", str(exe), src, str(prog), lib,
                    f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True, cwd=tmp_path,
                   capture_output=True, timeout=300)
    print("case fortran csr solve 30 x 20", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
