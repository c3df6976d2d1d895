"""The Fortran binding (duct_amd/fortran/duct_hip.f90) compiles with the image's
flang and a program calling every bound entry point links against
libduct_hip.so - CPU only (nothing is run)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

PROGRAM = """
program use_duct
  use iso_c_binding
  use duct_hip
  implicit none
  integer, parameter :: m = 9, nm = 2
  complex(c_double_complex) :: rhs(m, m, nm), sol(m, m, nm)
  real(c_double) :: P(m-2, m-2), Pinv(m-2, m-2), lam(m-2), DC(m, m), AM1(4, 4), U(m, m)
  integer(c_size_t) :: r(1), s(1)
  real(c_double) :: alams(1)
  rhs = 0; P = 0; Pinv = 0; lam = 0; DC = 0; AM1 = 0; U = 0; r = 0; s = 0; alams = 0
  call duct_check(duct_init(0_c_int), 'init')
  call duct_check(duct_setup('d', m, P, Pinv, lam, DC, AM1, 1_c_int, 0.0_c_double, 1_c_size_t), 'pseudomat')
  call duct_check(duct_helm3d(rhs, sol, m, nm, 0_c_int, 1.0_c_double, 1000.0_c_double, 'd', 1_c_size_t), 'helm3d')
  call duct_check(duct_helm3d_dev(r(1), s(1), m, nm, 0_c_int, 1.0_c_double, 0.0_c_double, 'n', 1_c_size_t), 'dev')
  call duct_check(duct_helm3d_batch_dev(1_c_int, r, s, 'd', alams, m, nm, 0_c_int, 1.0_c_double, 1_c_size_t), 'b')
  call duct_check(duct_int2d3d(sol, rhs, 'd', 2.0_c_double, 1_c_int, 1_c_size_t), 'int2d3d')
  call duct_check(duct_solvediag(U, 'n', 2.0_c_double, 1_c_size_t), 'solvediag')
  call duct_check(duct_set_fold(1_c_int), 'fold')
  call duct_check(duct_set_overlap(1_c_int), 'overlap')
  call duct_check(duct_helm3d_mixed_batch_dev(1_c_int, r, s, 'd', alams, m, nm, 0_c_int, 1.0_c_double, &
       3_c_int, 3_c_int, 1_c_int, c_null_ptr, 1_c_size_t), 'mixed')
  call duct_check(duct_helm3d_adjoint_batch_dev(1_c_int, r, s, 'd', alams, m, nm, 0_c_int, 1.0_c_double, &
       1_c_size_t), 'adjoint')
  call duct_check(duct_colloc_deriv_dev(r(1), s(1), m, nm, 'y', 0_c_int, 1.0_c_double, 0.0_c_double, &
       1_c_size_t), 'deriv')
  call duct_check(duct_derivx_dev(r(1), s(1), m, nm, 0_c_int, 1.0_c_double, 1.0_c_double, 0.0_c_double), 'dx')
  call duct_check(duct_advection_dev(r(1), r(1), r(1), s(1), s(1), s(1), m, 16_c_int, 48_c_int, &
       1.0_c_double, 1.0_c_double), 'advection')
  call duct_check(duct_transyz_dev(r(1), s(1), m, nm, 'y', -1_c_int, 1_c_size_t), 'transyz')
  call duct_check(duct_derivyz_dev(r(1), s(1), m, nm, 'z', 1_c_int, 1.0_c_double, 'n', 1.0_c_double, &
       1_c_size_t, 1_c_size_t), 'derivyz')
  call duct_check(duct_fourx_dev(r(1), s(1), m, m, 16_c_int, 48_c_int, 1_c_int), 'fourx')
  call duct_check(duct_rk3_substep_dev(r(1), r(1), r(1), r(1), s(1), s(1), s(1), m, 16_c_int, 48_c_int, &
       1.0_c_double, 1.0_c_double, 0.01_c_double, 0.1_c_double, 1_c_int, 0.0_c_double), 'rk3')
  call duct_check(duct_setup_temp('dndn', m, P, Pinv, lam, DC, P, Pinv, lam, DC, 4_c_size_t), 'setup_temp')
  call duct_check(duct_helm3d_temp_dev(r(1), s(1), r(1), r(1), m, nm, 0_c_int, 1.0_c_double, 2.0_c_double), 'temp')
  call duct_check(duct_setup_shen('d', m, P, P, P, P, lam, lam, lam, lam, 1_c_size_t), 'preshen')
  call duct_check(duct_helm3d_shen_dev(r(1), s(1), m, nm, 0_c_int, 1.0_c_double, 2.0_c_double, 'd', 1_c_size_t), &
       'shen')
  call duct_check(duct_shutdown(), 'shutdown')
end program use_duct
"""


@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_module_compiles_and_links(tmp_path):
    lib = os.path.join(ROOT, "duct_amd", "libduct_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libduct_hip.so not built")
    mod = os.path.join(ROOT, "duct_amd", "fortran", "duct_hip.f90")
    prog = tmp_path / "use_duct.f90"
    prog.write_text(PROGRAM)
    subprocess.run([FC, "-c", mod, "-J", str(tmp_path), "-o", str(tmp_path / "duct_hip.o")], check=True,
                   cwd=tmp_path)
    subprocess.run([FC, str(prog), str(tmp_path / "duct_hip.o"), "-I", str(tmp_path), "-L",
                    os.path.dirname(lib), "-lduct_hip", "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + os.path.dirname(lib), "-o", str(tmp_path / "use_duct")],
                   check=True, cwd=tmp_path)
    assert (tmp_path / "use_duct").exists()
