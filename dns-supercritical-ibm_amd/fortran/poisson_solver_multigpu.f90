! Drop-in replacement of the reference's module poisson_solver_multigpu
! (src/poisson_solver_mgpu_dtdma/poisson_solver_multigpu.f90) for one GPU.
!
! The subroutines keep the reference's names and dummy lists, so main.f90's
!   use poisson_solver_multigpu, only: init_spec_poisson_multigpu, run_spec_poisson_multigpu, &
!                                      get_lo_n_bounds_1d, planinfo_hipfft
! and its calls compile unchanged; they reach the MI355X library libdnsproj.so
! through iso_c_binding (include/dnsproj.h).  The handle of the C solver is
! kept in plan_fwd_x%plan_0 (the reference's hipFFT plan slot).
!
! Differences a caller can see:
!   * one rank only: mpi_divs_y = mpi_divs_z = 1, else error stop;
!   * the work arrays (poi_C, spec_x, spec_z, P_z, ...), the other plans and
!     the diezDecomp transpose objects are accepted and not touched; the
!     library owns its device workspace (diezDecomp types are class(*) here,
!     so this module does not depend on diezDecomp);
!   * a_x, a_z, poi_is_00x, poi_is_00z are filled as the reference's FFTW
!     branch fills them (:132-155), for callers that read them.
! Errors from the library end in error stop, as the reference's checks do.
module poisson_solver_multigpu
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only: dp => real64
  implicit none
  private
  public :: planinfo_hipfft, init_spec_poisson_multigpu, run_spec_poisson_multigpu, get_lo_n_bounds_1d

  type planinfo_hipfft
    type(C_PTR)  ::  plan_0 = c_null_ptr, plan_1 = c_null_ptr
    integer      ::  nreps_0, nreps_1, batch_0, batch_1, dj_in, dj_out
    logical      ::  is_fwd
    integer      ::  batch_max = 2**16
  end type

  interface
    function c_dnsp_poisson_init(nx, nz, ny, inv_dx, inv_dz, a, b, c, a00, b00, c00, precision, h) &
        bind(C, name="dnsp_poisson_init") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, nz, ny, precision
      real(c_double), value :: inv_dx, inv_dz
      real(c_double) :: a(*), b(*), c(*), a00(*), b00(*), c00(*)
      type(c_ptr) :: h
      integer(c_int) :: rc
    end function
    function c_dnsp_poisson_set_form(h, refinements) bind(C, name="dnsp_poisson_set_form") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: refinements
      integer(c_int) :: rc
    end function
    function c_dnsp_poisson_run(h, P, stream) bind(C, name="dnsp_poisson_run") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: h, P, stream
      integer(c_int) :: rc
    end function
    function c_dnsp_poisson_run_host(h, P) bind(C, name="dnsp_poisson_run_host") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: P(*)
      integer(c_int) :: rc
    end function
    function c_dnsp_poisson_min_pivot(h, v) bind(C, name="dnsp_poisson_min_pivot") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: v
      integer(c_int) :: rc
    end function
    function c_dnsp_stream_synchronize(stream) bind(C, name="dnsp_stream_synchronize") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: stream
      integer(c_int) :: rc
    end function
    function c_dnsp_get_lo_n_bounds_1d(divs, n, lo, nb) bind(C, name="dnsp_get_lo_n_bounds_1d") result(rc)
      import :: c_int
      integer(c_int), value :: divs, n
      integer(c_int) :: lo(*), nb(*)
      integer(c_int) :: rc
    end function
    function c_dnsp_last_error() bind(C, name="dnsp_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
  end interface

contains

  subroutine dnsp_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c_dnsp_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(6,'(A,A,A,I0,A,A)') 'ERROR: ', where, ' (dnsproj status ', rc, '): ', trim(text); flush(6)
    error stop 'dnsproj call failed'
  end subroutine

  subroutine init_spec_poisson_multigpu(inv_poi_B, poi_A, poi_C, poi_slabs_tips, &
                                        ts_band_a, ts_band_b, ts_cp, a_x, a_z, poi_is_00x, poi_is_00z, &
                                        P_band_a, P_band_b, P_band_c, P_band_a_00, P_band_b_00, P_band_c_00, P, &
                                        spec_x, spec_z, P_z, &
                                        nx_glob, nz_glob, n_fft_x, n_fft_z, mpi_divs_z, mpi_divs_y, mpi_pos_z, mpi_pos_y, &
                                        nx_transp, nz_transp_py, nz_loc, ny_loc, inv_dx, inv_dz, &
                                        plan_fwd_x, plan_bwd_x, plan_fwd_z, plan_bwd_z, utils_red_scalar, &
                                        tr_poi_yz, buffer_transp, lo_xt_blocks_mpi)
    integer     ::  nx_glob, nz_glob, n_fft_x, n_fft_z, mpi_divs_z, mpi_divs_y, mpi_pos_z, mpi_pos_y, &
                    nx_transp, nz_transp_py, nz_loc, ny_loc, lo_xt_blocks_mpi(0:), poi_is_00x(0:), poi_is_00z(0:)
    real(dp)    ::  inv_dx, inv_dz, &
                    inv_poi_B(0:,0:,0:), poi_A(0:,0:,0:), poi_C(0:,0:,0:), poi_slabs_tips(0:,0:,0:), &
                    ts_band_a(0:,0:,0:), ts_band_b(0:,0:,0:), ts_cp(0:,0:,0:), a_x(0:), a_z(0:), &
                    P_band_a(0:), P_band_b(0:), P_band_c(0:), P_band_a_00(0:), P_band_b_00(0:), P_band_c_00(0:), &
                    P(0:,0:,0:), buffer_transp(0:), spec_x(0:,0:,0:), spec_z(0:,0:,0:), P_z(0:,0:,0:)
    real(dp), contiguous :: utils_red_scalar(0:)
    class(*)              ::  tr_poi_yz
    type(planinfo_hipfft) ::  plan_fwd_x, plan_bwd_x, plan_fwd_z, plan_bwd_z
    real(c_double), allocatable :: ba(:), bb(:), bc(:), ba0(:), bb0(:), bc0(:)
    real(c_double) :: min_abs_div_bacp
    integer :: i, ii, form_len, form_stat, prec
    real(dp) :: const_pi
    character(len=16) :: form_env

    if (mpi_divs_y /= 1 .or. mpi_divs_z /= 1) error stop 'poisson_solver_multigpu (dnsproj): one rank only'
    if (ny_loc < 2) error stop 'ny_loc  <2'
    allocate(ba(ny_loc), bb(ny_loc), bc(ny_loc), ba0(ny_loc), bb0(ny_loc), bc0(ny_loc))
    ba = P_band_a(0:ny_loc-1);    bb = P_band_b(0:ny_loc-1);    bc = P_band_c(0:ny_loc-1)
    ba0 = P_band_a_00(0:ny_loc-1); bb0 = P_band_b_00(0:ny_loc-1); bc0 = P_band_c_00(0:ny_loc-1)
    ! DNSP_POISSON_FORM=mixed selects the mixed-precision solve (FP32 transforms,
    ! one FP64 residual refinement); default: FP64 as the reference
    prec = 0
    call get_environment_variable('DNSP_POISSON_FORM', form_env, form_len, form_stat)
    if (form_stat == 0 .and. trim(form_env) == 'mixed') prec = 1
    call dnsp_check(c_dnsp_poisson_init(int(nx_glob, c_int), int(nz_glob, c_int), int(ny_loc, c_int), &
                                        real(inv_dx, c_double), real(inv_dz, c_double), &
                                        ba, bb, bc, ba0, bb0, bc0, int(prec, c_int), plan_fwd_x%plan_0), &
                    'init_spec_poisson_multigpu')
    if (prec == 1) call dnsp_check(c_dnsp_poisson_set_form(plan_fwd_x%plan_0, 1_c_int), &
                                   'init_spec_poisson_multigpu: DNSP_POISSON_FORM')
    plan_bwd_x%plan_0 = plan_fwd_x%plan_0
    plan_fwd_z%plan_0 = plan_fwd_x%plan_0
    plan_bwd_z%plan_0 = plan_fwd_x%plan_0

    ! a_x, a_z, poi_is_00x/z as the FFTW branch fills them (:136-154)
    const_pi = 4.d0*datan(1.d0)
    do i = 0, size(a_x) - 1
      ii = hc_wavenumber(i + lo_xt_blocks_mpi(mpi_pos_z), nx_glob)
      a_x(i) = -4*(inv_dx**2)*(sin(ii*const_pi/nx_glob)**2)
      poi_is_00x(i) = merge(1, 0, ii == 0)
    end do
    do i = 0, size(a_z) - 1
      ii = hc_wavenumber(i, nz_glob)
      a_z(i) = -4*(inv_dz**2)*(sin(ii*const_pi/nz_glob)**2)
      poi_is_00z(i) = merge(1, 0, ii == 0)
    end do

    call dnsp_check(c_dnsp_poisson_min_pivot(plan_fwd_x%plan_0, min_abs_div_bacp), 'min pivot')
    write(6,'(A52,E14.6)') 'INFO: init_spec_poisson_multigpu, min_abs_div_bacp: ', min_abs_div_bacp; flush(6)
  end subroutine

  integer function hc_wavenumber(i, n)
    integer, intent(in) :: i, n
    if (i <= n/2) then
      hc_wavenumber = i
    else
      hc_wavenumber = max(1, (n+1)/2-1 - (i-(n/2 + 1)))
    end if
  end function

  subroutine run_spec_poisson_multigpu(inv_poi_B, poi_A, poi_C, poi_slabs_tips, &
                                       ts_band_a, ts_band_b, ts_cp, a_x, a_z, poi_is_00x, poi_is_00z, &
                                       P_band_a, P_band_b, P_band_c, P_band_a_00, P_band_b_00, P_band_c_00, P, &
                                       spec_x, spec_z, P_z, trispec, &
                                       nx_glob, nz_glob, n_fft_x, n_fft_z, mpi_divs_z, mpi_divs_y, mpi_pos_z, mpi_pos_y, &
                                       nx_transp, nz_transp_py, nx_loc, nz_loc, ny_loc, &
                                       plan_fwd_x, plan_bwd_x, plan_fwd_z, plan_bwd_z, &
                                       tr_poi_xy, tr_poi_yx, tr_poi_yz, tr_poi_zy, buffer_transp)
    integer     ::  nx_glob, nz_glob, n_fft_x, n_fft_z, mpi_divs_z, mpi_divs_y, mpi_pos_z, mpi_pos_y, &
                    nx_transp, nz_transp_py, nx_loc, nz_loc, ny_loc, poi_is_00x(0:), poi_is_00z(0:)
    real(dp)    ::  inv_poi_B(0:,0:,0:), poi_A(0:,0:,0:), poi_C(0:,0:,0:), &
                    ts_band_a(0:,0:,0:), ts_band_b(0:,0:,0:), ts_cp(0:,0:,0:), a_x(0:), a_z(0:), &
                    P_band_a(0:), P_band_b(0:), P_band_c(0:), P_band_a_00(0:), P_band_b_00(0:), P_band_c_00(0:), &
                    buffer_transp(0:), &
                    spec_x(0:,0:,0:), spec_z(0:,0:,0:), P_z(0:,0:,0:), poi_slabs_tips(0:,0:,0:), trispec(0:,0:,0:)
    real(dp), contiguous, target :: P(0:,0:,0:)
    class(*)              ::  tr_poi_xy, tr_poi_yx, tr_poi_yz, tr_poi_zy
    type(planinfo_hipfft) ::  plan_fwd_x, plan_bwd_x, plan_fwd_z, plan_bwd_z

    if (.not. c_associated(plan_fwd_x%plan_0)) error stop 'run_spec_poisson_multigpu: init not called'
    if (size(P,1) /= nx_glob .or. size(P,2) /= nz_loc .or. size(P,3) /= ny_loc) &
      error stop 'run_spec_poisson_multigpu: P shape'
#if defined(_OPENACC)
    ! device-resident P (OpenACC data region): solve in place on the device
    !$acc wait
    !$acc host_data use_device(P)
    call dnsp_check(c_dnsp_poisson_run(plan_fwd_x%plan_0, c_loc(P), c_null_ptr), 'run_spec_poisson_multigpu')
    !$acc end host_data
    call dnsp_check(c_dnsp_stream_synchronize(c_null_ptr), 'run_spec_poisson_multigpu sync')
#else
    ! host P: copy in, solve, copy out
    call dnsp_check(c_dnsp_poisson_run_host(plan_fwd_x%plan_0, P), 'run_spec_poisson_multigpu')
#endif
  end subroutine

  subroutine get_lo_n_bounds_1d(mpi_divs, n_glob, lo_blocks_mpi, n_blocks_mpi)
    integer :: mpi_divs, n_glob, lo_blocks_mpi(0:), n_blocks_mpi(0:)
    integer(c_int), allocatable :: lo(:), nb(:)
    allocate(lo(mpi_divs), nb(mpi_divs))
    call dnsp_check(c_dnsp_get_lo_n_bounds_1d(int(mpi_divs, c_int), int(n_glob, c_int), lo, nb), &
                    'get_lo_n_bounds_1d')
    lo_blocks_mpi(0:mpi_divs-1) = lo
    n_blocks_mpi(0:mpi_divs-1) = nb
    if (sum(n_blocks_mpi(0:mpi_divs-1)).ne.n_glob) error stop 'sum(n_blocks_mpi).ne.n_glob'
  end subroutine

end module poisson_solver_multigpu
