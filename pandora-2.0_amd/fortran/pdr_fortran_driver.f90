!---------------------------------------------------------------------------------
! program pdr_fortran_driver
!> A Fortran host that drives the MI355X step through the pdr_rk3_amd shim the
!! way the reference's control loop drives g_RK3Calc (src/pg_control.f90:177-194).
!! Odd steps call g_RK3Calc; even steps call f_RK3Stage1/2/3 and g_RK3UpdateB
!! one by one, as g_RK3Calc does internally (src/pg_rk3.f90:82-129).
!!
!! usage: pdr_fortran_driver N NSTEPS FTYPE SEED OUT_RESTART
!!   FTYPE 1: decaying isotropic, FIXED dt = 0.005, nu = 0.01, spectrum k0 = 4
!!   FTYPE 2: shear BF_ROT, DECIDE, F_REMESH_T = 3, S = 20, nu = 0.005, k0 = 6
program pdr_fortran_driver
  use iso_c_binding
  use pdr_rk3_amd
  implicit none
  type(pdr_params_t)  :: p
  integer             :: ierr, tstep, nsteps, n, ftype
  integer(c_int64_t)  :: seed
  real(kind=c_double) :: simtime, tstart, dt, kp
  character(len=256)  :: arg, outpath

  call get_command_argument(1, arg); read(arg, *) n
  call get_command_argument(2, arg); read(arg, *) nsteps
  call get_command_argument(3, arg); read(arg, *) ftype
  call get_command_argument(4, arg); read(arg, *) seed
  call get_command_argument(5, outpath)

  p%nodes_x = n
  p%nodes_y = n
  p%nodes_z = n
  p%f_type = ftype
  if ( ftype == 1 ) then
    p%rk3_timestep = 2
    p%rk3_dt = 0.005d0
    p%f_nu = 0.01d0
    kp = 4.0d0
  else
    p%rk3_timestep = 1
    p%f_remesh_t = 3
    p%f_deform_s = 20.0d0
    p%f_nu = 0.005d0
    kp = 6.0d0
  end if

  ierr = 0
  call pdr_Configure(p, ierr)
  call g_RK3Init(ierr)
  call check('g_RK3Init')
  call pdr_InitField(seed, 2.0d0/3.0d0, kp, ierr)
  call check('pdr_InitField')

  simtime = 0.0d0
  dt = 99999.0d0
  TIMELOOP: do tstep = 1, nsteps
    if ( mod(tstep, 2) == 1 ) then
      call g_RK3Calc(tstep, simtime, ierr)
      call check('g_RK3Calc')
    else
      tstart = simtime
      call f_RK3Stage1(tstep, simtime, dt, ierr)
      call check('f_RK3Stage1')
      simtime = tstart + (1.0d0/3.0d0) * dt
      call g_RK3UpdateB(tstep, dt, simtime, ierr)
      call f_RK3Stage2(tstep, simtime, dt, ierr)
      call check('f_RK3Stage2')
      simtime = tstart + (3.0d0/4.0d0) * dt
      call g_RK3UpdateB(tstep, dt, simtime, ierr)
      call f_RK3Stage3(tstep, simtime, dt, ierr)
      call check('f_RK3Stage3')
      simtime = tstart + dt
      call g_RK3UpdateB(tstep, dt, simtime, ierr)
    end if
  end do TIMELOOP

  call pdr_RestartWrite(trim(outpath), .false., ierr)
  call check('pdr_RestartWrite')
  write(*,'(a,es25.17)') 'simtime ', simtime
  call g_RK3Finalise(ierr)

contains

  subroutine check ( what )
    character(len=*), intent(in) :: what
    if ( ierr /= 0 ) then
      write(*,'(a,a,a,i0,a,a)') 'ERROR in ', what, ': code ', ierr, ' ', trim(pdr_ErrorMessage())
      stop 1
    end if
  end subroutine check

end program pdr_fortran_driver
