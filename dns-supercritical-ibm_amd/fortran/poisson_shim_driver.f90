! Test driver of the Fortran drop-in module, following the reference test's
! flow (tests/test_poisson_solver_and_ibm_rand/main_poi_test.f90:177-220,
! sizes of alloc_poisson_solver_multigpu.f90:3-54) on one rank with host
! arrays.
!   poisson_shim_driver <input> <output> <eigen_output>
! input : int32 nx, nz, ny; float64 inv_dx, inv_dz; bands a, b, c, a00, b00,
!         c00 (ny each); RHS (nx*nz*ny, x fastest)
! output: the reference's binary array format (int32 nx, nz, ny, then the
!         solution x fastest); eigen_output: a_x(0:n_fft_x-1), a_z(0:n_fft_z-1)
program poisson_shim_driver
  use, intrinsic :: iso_fortran_env, only: dp => real64, i4 => int32
  use poisson_solver_multigpu, only: init_spec_poisson_multigpu, run_spec_poisson_multigpu, &
                                     get_lo_n_bounds_1d, planinfo_hipfft
  implicit none
  character(len=1024) :: fin, fout, feig
  integer(i4) :: hdr(3)
  integer :: nx_glob, nz_glob, ny_glob, n_fft_x, n_fft_z, mpi_divs_z, mpi_divs_y, mpi_pos_z, mpi_pos_y, &
             nx_loc, nz_loc, ny_loc, nx_transp, nz_transp_py
  real(dp) :: inv_dx, inv_dz
  real(dp), allocatable :: P(:,:,:), spec_x(:,:,:), spec_z(:,:,:), P_z(:,:,:), poi_C(:,:,:), &
                           inv_poi_B(:,:,:), poi_A(:,:,:), poi_slabs_tips(:,:,:), ts_band_a(:,:,:), &
                           ts_band_b(:,:,:), ts_cp(:,:,:), trispec(:,:,:), a_x(:), a_z(:), buffer_transp(:), &
                           utils_red_scalar(:), bands(:,:)
  integer, allocatable :: poi_is_00x(:), poi_is_00z(:), lo_xt_blocks_mpi(:), nxt_blocks_mpi(:)
  type(planinfo_hipfft) :: plan_fwd_x, plan_bwd_x, plan_fwd_z, plan_bwd_z
  integer :: tr_poi_xy, tr_poi_yx, tr_poi_yz, tr_poi_zy   ! stand for the transpose objects (unused on one rank)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, feig)
  open(19, file=trim(fin), action='read', status='old', form='unformatted', access='stream')
  read(19) hdr
  nx_glob = hdr(1); nz_glob = hdr(2); ny_glob = hdr(3)
  read(19) inv_dx, inv_dz
  allocate(bands(0:ny_glob-1, 0:5))
  read(19) bands
  mpi_divs_z = 1; mpi_divs_y = 1; mpi_pos_z = 0; mpi_pos_y = 0
  n_fft_x = 2*(nx_glob/2 + 1)
  n_fft_z = 2*(nz_glob/2 + 1)
  allocate(lo_xt_blocks_mpi(0:0), nxt_blocks_mpi(0:0))
  call get_lo_n_bounds_1d(mpi_divs_z, n_fft_x, lo_xt_blocks_mpi, nxt_blocks_mpi)
  nx_loc = nx_glob; nz_loc = nz_glob; ny_loc = ny_glob
  nx_transp = nxt_blocks_mpi(0); nz_transp_py = n_fft_z
  allocate(P(0:nx_glob-1, 0:nz_loc-1, 0:ny_loc-1), spec_x(0:n_fft_x-1, 0:nz_loc-1, 0:ny_loc-1), &
           P_z(0:nz_glob-1, 0:nx_transp-1, 0:ny_loc-1), spec_z(0:n_fft_z-1, 0:nx_transp-1, 0:ny_loc-1), &
           a_x(0:nx_transp-1), a_z(0:n_fft_z-1), poi_is_00x(0:nx_transp-1), poi_is_00z(0:n_fft_z-1), &
           poi_C(0:n_fft_z-1, 0:nx_transp-1, 0:ny_loc-1))
  ! multi-rank work arrays: empty on one rank
  allocate(inv_poi_B(0:-1,0:-1,0:-1), poi_A(0:-1,0:-1,0:-1), poi_slabs_tips(0:-1,0:-1,0:-1), &
           ts_band_a(0:-1,0:-1,0:-1), ts_band_b(0:-1,0:-1,0:-1), ts_cp(0:-1,0:-1,0:-1), trispec(0:-1,0:-1,0:-1), &
           buffer_transp(0:0), utils_red_scalar(0:0))
  P = 0; a_x = 0; a_z = 0
  tr_poi_xy = 0; tr_poi_yx = 0; tr_poi_yz = 0; tr_poi_zy = 0

  call init_spec_poisson_multigpu(inv_poi_B, poi_A, poi_C, poi_slabs_tips, &
                                  ts_band_a, ts_band_b, ts_cp, a_x, a_z, poi_is_00x, poi_is_00z, &
                                  bands(:,0), bands(:,1), bands(:,2), bands(:,3), bands(:,4), bands(:,5), P, &
                                  spec_x, spec_z, P_z, &
                                  nx_glob, nz_glob, n_fft_x, n_fft_z, mpi_divs_z, mpi_divs_y, mpi_pos_z, mpi_pos_y, &
                                  nx_transp, nz_transp_py, nz_loc, ny_loc, inv_dx, inv_dz, &
                                  plan_fwd_x, plan_bwd_x, plan_fwd_z, plan_bwd_z, utils_red_scalar, &
                                  tr_poi_yz, buffer_transp, lo_xt_blocks_mpi)
  read(19) P
  close(19)
  call run_spec_poisson_multigpu(inv_poi_B, poi_A, poi_C, poi_slabs_tips, &
                                 ts_band_a, ts_band_b, ts_cp, a_x, a_z, poi_is_00x, poi_is_00z, &
                                 bands(:,0), bands(:,1), bands(:,2), bands(:,3), bands(:,4), bands(:,5), P, &
                                 spec_x, spec_z, P_z, trispec, &
                                 nx_glob, nz_glob, n_fft_x, n_fft_z, mpi_divs_z, mpi_divs_y, mpi_pos_z, mpi_pos_y, &
                                 nx_transp, nz_transp_py, nx_loc, nz_loc, ny_loc, &
                                 plan_fwd_x, plan_bwd_x, plan_fwd_z, plan_bwd_z, &
                                 tr_poi_xy, tr_poi_yx, tr_poi_yz, tr_poi_zy, buffer_transp)
  open(20, file=trim(fout), action='write', status='replace', form='unformatted', access='stream')
  write(20) int(nx_glob, i4), int(nz_glob, i4), int(ny_glob, i4)
  write(20) P
  close(20)
  open(21, file=trim(feig), action='write', status='replace', form='unformatted', access='stream')
  write(21) a_x, a_z
  close(21)
end program
