!---------------------------------------------------------------------------------
! program pdr_fortran_driver
!> A Fortran host that drives the MI355X step through the pdr_rk3_amd shim the
!! way the reference's control loop drives g_RK3Calc (src/pg_control.f90:177-194).
!! Odd steps call g_RK3Calc; even steps call f_RK3Stage1/2/3 and g_RK3UpdateB
!! one by one, as g_RK3Calc does internally (src/pg_rk3.f90:82-129).
!!
!! usage: pdr_fortran_driver N NSTEPS FTYPE SEED OUT_RESTART [NP MODE]
!!   FTYPE 1: decaying isotropic, FIXED dt = 0.005, nu = 0.01, spectrum k0 = 4
!!   FTYPE 2: shear BF_ROT, DECIDE, F_REMESH_T = 3, S = 20, nu = 0.005, k0 = 6
!!   NP tracers at seeded random positions (seed SEED + 100), MODE:
!!     1  advanced inside the stages (f_RK3Stage1..3 / g_RK3Calc)
!!     2  advanced by this program after each f_RK3StageN, as f_RK3RealSpace
!!        does (src/pf_rk3.f90:1145-1176): p_ControlRK3 on odd steps,
!!        p_RK3SolveParticle[BF] on even steps, then p_ControlFinishTimeStep;
!!        every step then goes through the stage calls
!!   positions before and after are written to OUT_RESTART.x0 / .x1
!!   (3*NP float64, particle_xp(1:3, np) order)
program pdr_fortran_driver
  use iso_c_binding
  use pdr_rk3_amd
  implicit none
  type(pdr_params_t)  :: p
  integer             :: ierr, tstep, nsteps, n, ftype, mode, stage
  integer(c_int64_t)  :: seed, np
  real(kind=c_double) :: simtime, tstart, dt, kp
  real(kind=c_double), allocatable :: xp(:,:)
  character(len=256)  :: arg, outpath
  real(kind=c_double), parameter :: cstage(3) = (/ 1.0d0/3.0d0, 3.0d0/4.0d0, 1.0d0 /)

  call get_command_argument(1, arg); read(arg, *) n
  call get_command_argument(2, arg); read(arg, *) nsteps
  call get_command_argument(3, arg); read(arg, *) ftype
  call get_command_argument(4, arg); read(arg, *) seed
  call get_command_argument(5, outpath)
  np = 0
  mode = 0
  if ( command_argument_count() >= 7 ) then
    call get_command_argument(6, arg); read(arg, *) np
    call get_command_argument(7, arg); read(arg, *) mode
  end if

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
  if ( np > 0 ) then
    allocate ( xp(3, np) )
    call pdr_InitParticles(seed + 100_c_int64_t, np, ierr)
    call check('pdr_InitParticles')
    call pdr_GetParticles(xp, ierr)
    call check('pdr_GetParticles')
    call write_positions(trim(outpath)//'.x0')
    if ( mode == 2 ) then
      call pdr_SetTracersInStage(.false., ierr)
      call check('pdr_SetTracersInStage')
    end if
  end if

  simtime = 0.0d0
  dt = 99999.0d0
  TIMELOOP: do tstep = 1, nsteps
    if ( mod(tstep, 2) == 1 .and. mode /= 2 ) then
      call g_RK3Calc(tstep, simtime, ierr)
      call check('g_RK3Calc')
    else
      tstart = simtime
      do stage = 1, 3
        select case ( stage )
          case ( 1 )
            call f_RK3Stage1(tstep, simtime, dt, ierr)
          case ( 2 )
            call f_RK3Stage2(tstep, simtime, dt, ierr)
          case default
            call f_RK3Stage3(tstep, simtime, dt, ierr)
        end select
        call check('f_RK3Stage')
        if ( mode == 2 ) call advance_tracers(stage)
        simtime = tstart + cstage(stage) * dt
        call g_RK3UpdateB(tstep, dt, simtime, ierr)
      end do
    end if
  end do TIMELOOP

  call pdr_RestartWrite(trim(outpath), .false., ierr)
  call check('pdr_RestartWrite')
  if ( np > 0 ) then
    call pdr_GetParticles(xp, ierr)
    call check('pdr_GetParticles')
    call write_positions(trim(outpath)//'.x1')
  end if
  write(*,'(a,es25.17)') 'simtime ', simtime
  call g_RK3Finalise(ierr)

contains

  !> the particle calls of f_RK3RealSpace (src/pf_rk3.f90:1145-1176)
  subroutine advance_tracers ( st )
    integer, intent(in) :: st
    if ( mod(tstep, 2) == 1 ) then
      call p_ControlRK3(st, tstep, dt, ierr)
      call check('p_ControlRK3')
    else if ( ftype == 1 ) then
      call p_RK3SolveParticle(int(np), st, dt, ierr)
      call check('p_RK3SolveParticle')
    else
      call p_RK3SolveParticleBF(int(np), st, dt, ierr)
      call check('p_RK3SolveParticleBF')
    end if
    call p_ControlFinishTimeStep(tstep, ierr)
    call check('p_ControlFinishTimeStep')
  end subroutine advance_tracers

  subroutine write_positions ( path )
    character(len=*), intent(in) :: path
    integer :: u
    open(newunit=u, file=path, access='stream', form='unformatted', status='replace')
    write(u) xp
    close(u)
  end subroutine write_positions

  subroutine check ( what )
    character(len=*), intent(in) :: what
    if ( ierr /= 0 ) then
      write(*,'(a,a,a,i0,a,a)') 'ERROR in ', what, ': code ', ierr, ' ', trim(pdr_ErrorMessage())
      stop 1
    end if
  end subroutine check

end program pdr_fortran_driver
