! Fortran binding of libduct_hip.so (include/duct_hip.h) for duct's Fortran
! callers: pseudomat.f90 hands its eigenbases to duct_setup, and helm3d.f90 /
! int2d3d.f90 / solvediag.f90 call duct_helm3d / duct_int2d3d / duct_solvediag
! instead of their CPU loops or the cuBLAS path.  Every entry returns an
! integer status; duct_check prints the library's message and stops, the way
! the reference stops on a bad boundary character (solvediag.f90:324-326).
module duct_hip
  use iso_c_binding
  implicit none
  private
  public :: duct_init, duct_shutdown, duct_setup, duct_helm3d, duct_helm3d_dev, &
            duct_helm3d_batch_dev, duct_int2d3d, duct_solvediag, duct_check, duct_set_fold, &
            duct_set_overlap, duct_set_fused_quarters, &
            duct_helm3d_mixed_batch_dev, duct_helm3d_adjoint_batch_dev, duct_colloc_deriv_dev, &
            duct_derivx_dev, duct_advection_dev, duct_transyz_dev, duct_derivyz_dev, duct_fourx_dev, &
            duct_rk3_substep_dev, duct_setup_temp, duct_helm3d_temp_dev, &
            duct_setup_shen, duct_helm3d_shen_dev, duct_h5_read_header, duct_h5_read_field, &
            duct_h5_read_doubles, duct_h5_read_chars

  interface
    integer(c_int) function duct_init(device) bind(C, name="duct_init_")
      import :: c_int
      integer(c_int), intent(in) :: device
    end function

    integer(c_int) function duct_shutdown() bind(C, name="duct_shutdown_")
      import :: c_int
    end function

    integer(c_int) function duct_last_error(buf, len) bind(C, name="duct_last_error_")
      import :: c_int, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), intent(in) :: len
    end function

    ! pseudomat.f90:188-221 / :265-298 (after dgeev, dgetrf/dgetri)
    integer(c_int) function duct_setup(bound, m, P, Pinv, lam, DC, AM1, iofzero, wanted, bound_len) &
        bind(C, name="duct_setup_")
      import :: c_int, c_double, c_char, c_size_t
      character(kind=c_char), intent(in) :: bound
      integer(c_int), intent(in) :: m, iofzero
      real(c_double), intent(in) :: P(*), Pinv(*), lam(*), DC(*), AM1(*)
      real(c_double), intent(in) :: wanted
      integer(c_size_t), value :: bound_len
    end function

    ! helm3d.f90:2 on host arrays: rhs, sol complex(8) (my, mz, nmodes)
    integer(c_int) function duct_helm3d(rhs, sol, m, nmodes, ib, alp, alambda, bound, bound_len) &
        bind(C, name="duct_helm3d_")
      import :: c_int, c_double, c_double_complex, c_char, c_size_t
      complex(c_double_complex), intent(inout) :: rhs(*)
      complex(c_double_complex), intent(out) :: sol(*)
      integer(c_int), intent(in) :: m, nmodes, ib
      real(c_double), intent(in) :: alp, alambda
      character(kind=c_char), intent(in) :: bound
      integer(c_size_t), value :: bound_len
    end function

    ! helm3d on device-resident fields (addresses held in integer(8), as the
    ! reference's gpudevice module holds its devptr_t values, mod.f90:442-481)
    integer(c_int) function duct_helm3d_dev(rhs, sol, m, nmodes, ib, alp, alambda, bound, bound_len) &
        bind(C, name="duct_helm3d_dev_")
      import :: c_int, c_double, c_size_t, c_char
      integer(c_size_t), intent(in) :: rhs, sol
      integer(c_int), intent(in) :: m, nmodes, ib
      real(c_double), intent(in) :: alp, alambda
      character(kind=c_char), intent(in) :: bound
      integer(c_size_t), value :: bound_len
    end function

    integer(c_int) function duct_helm3d_batch_dev(nfields, rhs, sol, bounds, alambda, m, nmodes, ib, alp, &
        bounds_len) bind(C, name="duct_helm3d_batch_dev_")
      import :: c_int, c_double, c_size_t, c_char
      integer(c_int), intent(in) :: nfields, m, nmodes, ib
      integer(c_size_t), intent(in) :: rhs(*), sol(*)
      character(kind=c_char), intent(in) :: bounds(*)
      real(c_double), intent(in) :: alambda(*), alp
      integer(c_size_t), value :: bounds_len
    end function

    ! int2d3d.f90:2: one complex plane
    integer(c_int) function duct_int2d3d(sol, rhs, bound, xlambda, imode, bound_len) &
        bind(C, name="duct_int2d3d_")
      import :: c_int, c_double, c_double_complex, c_char, c_size_t
      complex(c_double_complex), intent(out) :: sol(*)
      complex(c_double_complex), intent(in) :: rhs(*)
      character(kind=c_char), intent(in) :: bound
      real(c_double), intent(in) :: xlambda
      integer(c_int), intent(in) :: imode
      integer(c_size_t), value :: bound_len
    end function

    ! solvediag.f90:2: one real plane U(my,mz), in place
    integer(c_int) function duct_solvediag(U, bound, xlambda, bound_len) bind(C, name="duct_solvediag_")
      import :: c_int, c_double, c_char, c_size_t
      real(c_double), intent(inout) :: U(*)
      character(kind=c_char), intent(in) :: bound
      real(c_double), intent(in) :: xlambda
      integer(c_size_t), value :: bound_len
    end function

    ! even/odd folded eigenbasis for the FP64 solves: 0 off (default), 1 Dirichlet, 2 both
    integer(c_int) function duct_set_fold(mode) bind(C, name="duct_set_fold_")
      import :: c_int
      integer(c_int), intent(in) :: mode
    end function

    ! Neumann group of a batched solve on a second stream (1, default) or in order (0)
    integer(c_int) function duct_set_overlap(on) bind(C, name="duct_set_overlap_")
      import :: c_int
      integer(c_int), intent(in) :: on
    end function

    ! folded passes 2 and 3 on LDS-resident quarters (1, default) or as two planar launches (0)
    integer(c_int) function duct_set_fused_quarters(on) bind(C, name="duct_set_fused_quarters_")
      import :: c_int
      integer(c_int), intent(in) :: on
    end function

    ! Field files (host code, no device needed). Strings are passed with their lengths;
    ! trailing blanks are ignored, so a blank-padded character variable and len() will do.
    ! readheader_hdf5 (readwrite_hdf5.f90:798-850): ints = mgalx, mx1, my, mz, nparams;
    ! reals = alp, aspect, time, fnu, cfl, dtfixed, cmassflow
    integer(c_int) function duct_h5_read_header(filename, ints, reals, filename_len) &
        bind(C, name="duct_h5_read_header_")
      import :: c_int, c_double, c_char, c_size_t
      character(kind=c_char), intent(in) :: filename(*)
      integer(c_int), intent(out) :: ints(5)
      real(c_double), intent(out) :: reals(7)
      integer(c_size_t), value :: filename_len
    end function

    ! one field of read_4fields_fou_hdf5 / read_5fields_fou_hdf5 (readwrite_hdf5.f90:2-264)
    ! into a(my, mz, ib:ib+nmodes-1)
    integer(c_int) function duct_h5_read_field(filename, name, a, my, mz, ib, nmodes, filename_len, name_len) &
        bind(C, name="duct_h5_read_field_")
      import :: c_int, c_double_complex, c_char, c_size_t
      character(kind=c_char), intent(in) :: filename(*), name(*)
      complex(c_double_complex), intent(out) :: a(*)
      integer(c_int), intent(in) :: my, mz, ib, nmodes
      integer(c_size_t), value :: filename_len, name_len
    end function

    ! a numeric dataset (header scalars, bcy, bcz) of exactly count elements
    integer(c_int) function duct_h5_read_doubles(filename, name, out, count, filename_len, name_len) &
        bind(C, name="duct_h5_read_doubles_")
      import :: c_int, c_double, c_char, c_size_t, c_long_long
      character(kind=c_char), intent(in) :: filename(*), name(*)
      real(c_double), intent(out) :: out(*)
      integer(c_long_long), intent(in) :: count
      integer(c_size_t), value :: filename_len, name_len
    end function

    ! a character dataset (boundy1 ... boundz2)
    integer(c_int) function duct_h5_read_chars(filename, name, out, capacity, length, filename_len, name_len) &
        bind(C, name="duct_h5_read_chars_")
      import :: c_int, c_char, c_size_t
      character(kind=c_char), intent(in) :: filename(*), name(*)
      character(kind=c_char), intent(out) :: out(*)
      integer(c_int), intent(in) :: capacity
      integer(c_int), intent(out) :: length
      integer(c_size_t), value :: filename_len, name_len
    end function

    ! mixed-precision refined variant of duct_helm3d_batch_dev
    integer(c_int) function duct_helm3d_mixed_batch_dev(nfields, rhs, sol, bounds, alambda, m, nmodes, ib, &
        alp, scheme0, scheme1, niter, stats, bounds_len) bind(C, name="duct_helm3d_mixed_batch_dev_")
      import :: c_int, c_double, c_size_t, c_char, c_ptr
      integer(c_int), intent(in) :: nfields, m, nmodes, ib, scheme0, scheme1, niter
      integer(c_size_t), intent(in) :: rhs(*), sol(*)
      character(kind=c_char), intent(in) :: bounds(*)
      real(c_double), intent(in) :: alambda(*), alp
      type(c_ptr), value :: stats
      integer(c_size_t), value :: bounds_len
    end function

    ! transposed solve (the backward of helm3d)
    integer(c_int) function duct_helm3d_adjoint_batch_dev(nfields, solbar, rhsbar, bounds, alambda, m, nmodes, &
        ib, alp, bounds_len) bind(C, name="duct_helm3d_adjoint_batch_dev_")
      import :: c_int, c_double, c_size_t, c_char
      integer(c_int), intent(in) :: nfields, m, nmodes, ib
      integer(c_size_t), intent(in) :: solbar(*), rhsbar(*)
      character(kind=c_char), intent(in) :: bounds(*)
      real(c_double), intent(in) :: alambda(*), alp
      integer(c_size_t), value :: bounds_len
    end function

    ! collocation derivative products along y or z, and d/dx in Fourier space
    integer(c_int) function duct_colloc_deriv_dev(in, out, m, nmodes, cdir, trans, scal, beta, cdir_len) &
        bind(C, name="duct_colloc_deriv_dev_")
      import :: c_int, c_double, c_size_t, c_char
      integer(c_size_t), intent(in) :: in, out
      integer(c_int), intent(in) :: m, nmodes, trans
      character(kind=c_char), intent(in) :: cdir
      real(c_double), intent(in) :: scal, beta
      integer(c_size_t), value :: cdir_len
    end function

    integer(c_int) function duct_derivx_dev(in, out, m, nmodes, ib, alp, scal, beta) &
        bind(C, name="duct_derivx_dev_")
      import :: c_int, c_double, c_size_t
      integer(c_size_t), intent(in) :: in, out
      integer(c_int), intent(in) :: m, nmodes, ib
      real(c_double), intent(in) :: alp, scal, beta
    end function

    ! Shen-Legendre Galerkin solve (helm3d_shen): preshen uploads and the solve
    integer(c_int) function duct_setup_shen(bound, m, Ay, Az, By, Bz, md, mu, sd, lam, bound_len) &
        bind(C, name="duct_setup_shen_")
      import :: c_int, c_double, c_size_t, c_char
      character(kind=c_char), intent(in) :: bound(*)
      integer(c_int), intent(in) :: m
      real(c_double), intent(in) :: Ay(*), Az(*), By(*), Bz(*), md(*), mu(*), sd(*), lam(*)
      integer(c_size_t), value :: bound_len
    end function
    integer(c_int) function duct_helm3d_shen_dev(rhs, sol, m, nmodes, ib, alp, alambda, bound, bound_len) &
        bind(C, name="duct_helm3d_shen_dev_")
      import :: c_int, c_size_t, c_double, c_char
      integer(c_size_t), intent(in) :: rhs, sol
      integer(c_int), intent(in) :: m, nmodes, ib
      real(c_double), intent(in) :: alp, alambda
      character(kind=c_char), intent(in) :: bound(*)
      integer(c_size_t), value :: bound_len
    end function

    ! mixed-boundary (temperature) solve: pseudomat_temp uploads and helm3d_temp
    ! (my may differ from mz: PY, PYinv my-2 square, PZ, PZinv mz-2 square; DY my x my, DZ mz x mz)
    integer(c_int) function duct_setup_temp(bounds, my, mz, PY, PYinv, ETY, DY, PZ, PZinv, ETZ, DZ, bounds_len) &
        bind(C, name="duct_setup_temp_")
      import :: c_int, c_double, c_size_t, c_char
      character(kind=c_char), intent(in) :: bounds(*)
      integer(c_int), intent(in) :: my, mz
      real(c_double), intent(in) :: PY(*), PYinv(*), ETY(*), DY(*), PZ(*), PZinv(*), ETZ(*), DZ(*)
      integer(c_size_t), value :: bounds_len
    end function
    ! rhs, sol complex(8) (my, mz, nmodes); bcy real(8) (mz, 2, nmodes), bcz (my, 2, nmodes)
    integer(c_int) function duct_helm3d_temp_dev(rhs, sol, bcy, bcz, my, mz, nmodes, ib, alp, alambda) &
        bind(C, name="duct_helm3d_temp_dev_")
      import :: c_int, c_size_t, c_double
      integer(c_size_t), intent(in) :: rhs, sol, bcy, bcz
      integer(c_int), intent(in) :: my, mz, nmodes, ib
      real(c_double), intent(in) :: alp, alambda
    end function

    ! one RK3 substep of u, v, w, p (main1.f90:384-655)
    integer(c_int) function duct_rk3_substep_dev(u, v, w, p, adv1, adv2, adv3, m, nmodes, mgalx, alp, aspect, &
                                                 fnu, dt, krk, dpdx) bind(C, name="duct_rk3_substep_dev_")
      import :: c_int, c_size_t, c_double
      integer(c_size_t), intent(in) :: u, v, w, p, adv1, adv2, adv3
      integer(c_int), intent(in) :: m, nmodes, mgalx, krk
      real(c_double), intent(in) :: alp, aspect, fnu, dt, dpdx
    end function

    ! fourx (fourier.F:1-115) on the z-cut layouts
    integer(c_int) function duct_fourx_dev(fou, phys, my, nz, nmodes, mgalx, iopt) bind(C, name="duct_fourx_dev_")
      import :: c_int, c_size_t
      integer(c_size_t), intent(in) :: fou, phys
      integer(c_int), intent(in) :: my, nz, nmodes, mgalx, iopt
    end function

    ! transyz (chebyshev.f90:458-540) and derivyz (:1-295) on device fields
    integer(c_int) function duct_transyz_dev(in, out, m, nmodes, cdir, iopt, cdir_len) &
        bind(C, name="duct_transyz_dev_")
      import :: c_int, c_size_t, c_char
      integer(c_size_t), intent(in) :: in, out
      integer(c_int), intent(in) :: m, nmodes, iopt
      character(kind=c_char), intent(in) :: cdir
      integer(c_size_t), value :: cdir_len
    end function

    integer(c_int) function duct_derivyz_dev(u, du, m, nmodes, cdir, ider, scal, what, aspect, cdir_len, &
        what_len) bind(C, name="duct_derivyz_dev_")
      import :: c_int, c_double, c_size_t, c_char
      integer(c_size_t), intent(in) :: u, du
      integer(c_int), intent(in) :: m, nmodes, ider
      character(kind=c_char), intent(in) :: cdir, what
      real(c_double), intent(in) :: scal, aspect
      integer(c_size_t), value :: cdir_len, what_len
    end function

    ! advection3d (operators.F:9-276) on device fields
    integer(c_int) function duct_advection_dev(u, v, w, adv1, adv2, adv3, m, nmodes, mgalx, alp, aspect) &
        bind(C, name="duct_advection_dev_")
      import :: c_int, c_double, c_size_t
      integer(c_size_t), intent(in) :: u, v, w, adv1, adv2, adv3
      integer(c_int), intent(in) :: m, nmodes, mgalx
      real(c_double), intent(in) :: alp, aspect
    end function
  end interface

contains

  subroutine duct_check(status, where)
    integer(c_int), intent(in) :: status
    character(*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, st
    if (status == 0) return
    st = duct_last_error(buf, 512_c_int)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    write(6,*) 'in ', where, ': libduct_hip status ', status, ': ', trim(msg)
    stop
  end subroutine duct_check

end module duct_hip
