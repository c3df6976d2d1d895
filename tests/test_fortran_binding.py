"""The Fortran binding (duct_amd/fortran/duct_hip.f90) compiles with the image's
flang and a program calling every bound entry point links against
libduct_hip.so (CPU); the field-file reader, which is host code, is also run from
Fortran on a fixture.  On a GPU a Fortran program runs the drop-in solve sequence
and is checked against the oracle (test_fortran_program_solves_on_gpu)."""
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
  real(c_double) :: alams(1), rh(7)
  integer(c_int) :: ih(5), nw
  character(kind=c_char) :: wall(8)
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
  call duct_check(duct_setup_temp('dndn', m, m, P, Pinv, lam, DC, P, Pinv, lam, DC, 4_c_size_t), 'setup_temp')
  call duct_check(duct_helm3d_temp_dev(r(1), s(1), r(1), r(1), m, m, nm, 0_c_int, 1.0_c_double, 2.0_c_double), &
       'temp')
  call duct_check(duct_setup_shen('d', m, P, P, P, P, lam, lam, lam, lam, 1_c_size_t), 'preshen')
  call duct_check(duct_helm3d_shen_dev(r(1), s(1), m, nm, 0_c_int, 1.0_c_double, 2.0_c_double, 'd', 1_c_size_t), &
       'shen')
  call duct_check(duct_h5_read_header('f.h5', ih, rh, 4_c_size_t), 'h5 header')
  call duct_check(duct_h5_read_field('f.h5', 'u', rhs, m, m, 0_c_int, nm, 4_c_size_t, 1_c_size_t), 'h5 field')
  call duct_check(duct_h5_read_doubles('f.h5', 'bcy', U, 4_c_long_long, 4_c_size_t, 3_c_size_t), 'h5 bcy')
  call duct_check(duct_h5_read_chars('f.h5', 'boundy1', wall, 8_c_int, nw, 4_c_size_t, 7_c_size_t), 'h5 wall')
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


READER = """
program read_uvwp
  use iso_c_binding
  use duct_hip
  implicit none
  character(len=512) :: fn
  integer(c_int) :: ih(5), nw
  real(c_double) :: rh(7), bcy(9, 2, 0:1)
  character(kind=c_char) :: wall(4)
  complex(c_double_complex), allocatable :: a(:, :, :)
  call get_command_argument(1, fn)
  ! the whole blank-padded variable and its full length, as a Fortran caller passes it
  call duct_check(duct_h5_read_header(fn, ih, rh, int(len(fn), c_size_t)), 'header')
  allocate(a(ih(3), ih(4), 0:ih(2)))
  call duct_check(duct_h5_read_field(fn, 'w', a, ih(3), ih(4), 0_c_int, ih(2) + 1_c_int, &
                                     int(len(fn), c_size_t), 1_c_size_t), 'read w')
  call duct_check(duct_h5_read_doubles(fn, 'bcy', bcy, 36_c_long_long, int(len(fn), c_size_t), 3_c_size_t), 'bcy')
  call duct_check(duct_h5_read_chars(fn, 'boundz2', wall, 4_c_int, nw, int(len(fn), c_size_t), 7_c_size_t), 'wall')
  write(*, '(5i6)') ih
  write(*, '(3es24.16)') rh(1), rh(3), rh(4)
  write(*, '(6es24.16)') a(1, 1, 0), a(3, 2, 1), a(ih(3), ih(4), ih(2))
  write(*, '(2es24.16)') bcy(5, 2, 1), bcy(9, 1, 0)
  write(*, '(i3, 1x, a1)') nw, wall(1)
end program read_uvwp
"""


@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_reads_a_field_file(tmp_path):
    """read_4fields_fou_hdf5's replacement called from Fortran on the temperature fixture: the field
    w lands in a(my, mz, 0:mx1) in the reference's order (y fastest, then z, then the mode)."""
    lib = os.path.join(ROOT, "duct_amd", "libduct_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libduct_hip.so not built")
    mod = os.path.join(ROOT, "duct_amd", "fortran", "duct_hip.f90")
    prog = tmp_path / "read_uvwp.f90"
    prog.write_text(READER)
    subprocess.run([FC, "-c", mod, "-J", str(tmp_path), "-o", str(tmp_path / "duct_hip.o")], check=True,
                   cwd=tmp_path)
    subprocess.run([FC, str(prog), str(tmp_path / "duct_hip.o"), "-I", str(tmp_path), "-L",
                    os.path.dirname(lib), "-lduct_hip", "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + os.path.dirname(lib), "-o", str(tmp_path / "read_uvwp")],
                   check=True, cwd=tmp_path)
    fixture = os.path.join(ROOT, "tests", "golden", "uvwpt_m9.h5")
    out = subprocess.run([str(tmp_path / "read_uvwp"), fixture], check=True, capture_output=True, text=True,
                         timeout=60).stdout.split("\n")
    assert [int(x) for x in out[0].split()] == [6, 1, 9, 9, 13]
    assert [float(x) for x in out[1].split()] == [1.0, 12.5, 0.00125]

    def w(k, z, y):       # the fixture's field 2 ("w"), mode k, 0-based z and y; mode 0 is real
        v = 3.0 + k * 0.5 + z * 0.03125 + y * 0.0009765625
        return [v, -(v + 0.25) if k else 0.0]
    assert [float(x) for x in out[2].split()] == w(0, 0, 0) + w(1, 1, 2) + w(1, 8, 8)
    # bcy(z=5, wall 2, mode 1) and bcy(z=9, wall 1, mode 0), 1-based z
    assert [float(x) for x in out[3].split()] == [100.0 + 8.0 + 4.0 + 4 * 0.125, 100.0 + 8 * 0.125]
    assert out[4].split() == ["1", "n"]


SOLVER = """
program solve_cpu33
  ! the reference's sequence (main1.f90:149-150 device init, pseudomat.f90:188-221 and 265-298
  ! uploads, helm3d.f90:48-73 solve) through the bind(C) module, on host arrays
  use iso_c_binding
  use duct_hip
  implicit none
  character(len=512) :: fin, fout
  integer(c_int) :: m, nm, iofz(2)
  real(c_double) :: wanted(2)
  real(c_double), allocatable :: P(:, :, :), Pinv(:, :, :), lam(:, :), DC(:, :), AM1(:, :, :)
  complex(c_double_complex), allocatable :: rhs(:, :, :, :), sol(:, :, :, :)
  character(kind=c_char) :: bnd(2)
  real(c_double) :: alam(2)
  integer :: b
  bnd = ['d', 'n']
  alam = [1000.0_c_double, 0.0_c_double]
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) m, nm
  allocate(P(m-2, m-2, 2), Pinv(m-2, m-2, 2), lam(m-2, 2), DC(m, m), AM1(4, 4, 2))
  allocate(rhs(m, m, nm, 2), sol(m, m, nm, 2))
  read(10) DC
  do b = 1, 2
    read(10) P(:, :, b), Pinv(:, :, b), lam(:, b), AM1(:, :, b), iofz(b), wanted(b)
  end do
  read(10) rhs
  close(10)
  call duct_check(duct_init(0_c_int), 'init')
  do b = 1, 2
    call duct_check(duct_setup(bnd(b), m, P(:, :, b), Pinv(:, :, b), lam(:, b), DC, AM1(:, :, b), iofz(b), &
                               wanted(b), 1_c_size_t), 'setup')
  end do
  do b = 1, 2
    call duct_check(duct_helm3d(rhs(:, :, :, b), sol(:, :, :, b), m, nm, 0_c_int, 1.0_c_double, alam(b), &
                                bnd(b), 1_c_size_t), 'helm3d')
  end do
  call duct_check(duct_shutdown(), 'shutdown')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) sol
  close(11)
  write(*, '(a)') 'solved'
end program solve_cpu33
"""


@pytest.mark.gpu
@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_program_solves_on_gpu(tmp_path, oracle_mod):
    """A flang-built program runs the drop-in sequence on the GPU: duct_init, duct_setup for 'd' and
    'n' with the oracle's own bases (what the reference's pseudomat computes and hands over), then
    duct_helm3d on the 33^2 x 16 CPU workload's right-hand sides (u, Dirichlet, alambda 1000, and p,
    Neumann, alambda 0).  Both solutions match the oracle's helm3d to the FP64 bar."""
    import numpy as np
    from duct_amd import workloads
    lib = os.path.join(ROOT, "duct_amd", "libduct_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libduct_hip.so not built")
    mod = os.path.join(ROOT, "duct_amd", "fortran", "duct_hip.f90")
    prog = tmp_path / "solve_cpu33.f90"
    prog.write_text(SOLVER)
    subprocess.run([FC, "-c", mod, "-J", str(tmp_path), "-o", str(tmp_path / "duct_hip.o")], check=True,
                   cwd=tmp_path)
    subprocess.run([FC, str(prog), str(tmp_path / "duct_hip.o"), "-I", str(tmp_path), "-L",
                    os.path.dirname(lib), "-lduct_hip", "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + os.path.dirname(lib), "-o", str(tmp_path / "solve_cpu33")],
                   check=True, cwd=tmp_path)
    cfg = workloads.CPU_33
    m, nm = cfg.m, cfg.nmodes
    ob = oracle_mod.OracleBasis(oracle_mod.pseudomat(m))
    b = ob.b
    f = [workloads.rhs_field(cfg, 0), workloads.rhs_field(cfg, 3)]     # u ('d') and p ('n')
    assert workloads.BOUNDS[0] == "d" and workloads.BOUNDS[3] == "n"

    def fa(a):
        return np.asarray(a, dtype=np.float64).tobytes(order="F")
    blob = [np.array([m, nm], dtype=np.int32).tobytes(), fa(b.DC)]
    for P, Pi, lam, am1, io, wa in ((b.PD, b.PDM1, b.ED, np.zeros((4, 4)), 0, 0.0),
                                    (b.PN, b.PNM1, b.EN, b.AM1, b.iofzero, b.wanted)):
        blob += [fa(P), fa(Pi), fa(lam), fa(am1), np.array([io], dtype=np.int32).tobytes(), fa([wa])]
    blob.append(np.ascontiguousarray(np.stack(f)).tobytes())          # Fortran rhs(m, m, nm, 2)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(b"".join(blob))
    print("running the Fortran drop-in program, 33^2 x 16, d and n", flush=True)
    out = subprocess.run([str(tmp_path / "solve_cpu33"), str(fin), str(fout)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = np.frombuffer(fout.read_bytes(), dtype=np.complex128).reshape(2, nm, m, m)
    for i, (bound, alam) in enumerate((("d", 1000.0), ("n", 0.0))):
        ref, _ = oracle_mod.helm3d(ob, f[i], bound, alam, alp=1.0, ib=0)
        err = float(np.max(np.abs(got[i] - ref)) / np.max(np.abs(ref)))
        print("Fortran drop-in %s: max rel err vs oracle %.2e" % (bound, err), flush=True)
        assert err <= 1e-12, (bound, err)
