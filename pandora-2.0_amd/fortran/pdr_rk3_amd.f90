!---------------------------------------------------------------------------------
! module pdr_rk3_amd
!> ISO_C_BINDING shim that keeps the reference's time-step interfaces
!! (g_RK3Init / g_RK3Calc / g_RK3UpdateB in src/pg_rk3.f90:46-385 and
!! f_RK3Stage1/2/3 in src/pf_rk3.f90:40-200) and runs them on the MI355X
!! library declared in include/pandora_amd.h, together with the tracer entry
!! points p_ControlRK3 / p_RK3SolveParticle[BF|R] / p_ControlFinishTimeStep
!! (src/pp_control.f90:518, 1440; src/pp_rk3.f90:57, 169, 307), which advance
!! the tracers on the device when the caller drives them itself
!! (pdr_SetTracersInStage(.false.)).  The Fortran caller keeps its
!! TIMELOOP (src/pg_control.f90:177-194) and its pf_arrays-layout arrays
!! arr(0:1, 0:NKK-1, 0:NKJ-1, 0:NKI-1); the whole box lives on one GPU.
!!
!! Differences from the reference module: run parameters are handed over once
!! with pdr_Configure (the reference reads module globals of g_parameters), and
!! state moves between host arrays and the device with pdr_PutVelocity /
!! pdr_GetVelocity (the reference keeps PETSc Vecs).  ierr is 0 on success and
!! the library's negative error code otherwise; pdr_ErrorMessage returns the text.
module pdr_rk3_amd
  use iso_c_binding
  implicit none
  private

  !> Mirror of struct pdr_params (include/pandora_amd.h)
  type, bind(C), public :: pdr_params_t
    integer(c_int32_t) :: nodes_x = 0, nodes_y = 0, nodes_z = 0
    integer(c_int32_t) :: f_type = 1
    integer(c_int32_t) :: ns_deformed = 1   ! 1 BF_ROT, 2 R_CONV, 3 R_ROT
    integer(c_int32_t) :: dealiasing = 1    ! 1 two-thirds, 2 spherical, 3 LES
    integer(c_int32_t) :: rk3_timestep = 2
    integer(c_int32_t) :: f_remesh_t = 0
    integer(c_int32_t) :: b110 = 1
    integer(c_int32_t) :: p_interp_order = 4
    integer(c_int32_t) :: precision = 0
    integer(c_int32_t) :: force_stop = 0
    real(c_double)     :: rk3_dt = 0.01d0
    real(c_double)     :: rk3_cfl = 0.75d0
    real(c_double)     :: f_nu = 0.0d0
    real(c_double)     :: f_deform_s = 0.0d0
    real(c_double)     :: force_tf = 0.0d0
    real(c_double)     :: force_kf = 0.0d0
    real(c_double)     :: force_sigma = 0.0d0
    integer(c_int64_t) :: force_seed = 0
    real(c_double)     :: les_cutoff = 1.0d0
    integer(c_int32_t) :: b220 = 1, b330 = 1
  end type pdr_params_t

  integer(c_int32_t), parameter, public :: PDR_FP64 = 0, PDR_FP32_MIXED = 1

  type(c_ptr), save         :: ctx = c_null_ptr
  type(pdr_params_t), save  :: params

  public :: pdr_Configure, g_RK3Init, g_RK3Calc, g_RK3UpdateB, g_RK3Finalise
  public :: f_RK3Stage1, f_RK3Stage2, f_RK3Stage3, f_RK3Remesh
  public :: pdr_PutVelocity, pdr_GetVelocity, pdr_PutParticles, pdr_GetParticles
  public :: pdr_InitField, pdr_RestartWrite, pdr_RestartRead, pdr_RestartReadScaled, pdr_Spectrum, pdr_ErrorMessage
  public :: pdr_GetFrame, pdr_PutForcing, pdr_GetForcing, pdr_Stats, pdr_RsStats
  public :: pdr_ParticleStatsConfig, pdr_ParticleStats, pdr_GetParticlesAux
  public :: pdr_InitParticles, pdr_SetTracersInStage
  public :: p_ControlRK3, p_ControlFinishTimeStep
  public :: p_RK3SolveParticle, p_RK3SolveParticleBF, p_RK3SolveParticleR
  public :: pdr_RsCorrelation, pdr_ParticleRestartWrite, pdr_ParticleRestartRead

  interface
    integer(c_int) function c_pdr_create(out, p) bind(C, name="pdr_create")
      import :: c_int, c_ptr, pdr_params_t
      type(c_ptr), intent(out) :: out
      type(pdr_params_t), intent(in) :: p
    end function
    integer(c_int) function c_pdr_destroy(c) bind(C, name="pdr_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function c_pdr_step(c, tstep, simtime) bind(C, name="pdr_step")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), value :: tstep
      real(c_double), intent(inout) :: simtime
    end function
    integer(c_int) function c_pdr_stage(c, stage, tstep, time, dt) bind(C, name="pdr_stage")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), value :: stage, tstep
      real(c_double), value :: time
      real(c_double), intent(inout) :: dt
    end function
    integer(c_int) function c_pdr_update_b(c, tstep, simtime) bind(C, name="pdr_update_b")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), value :: tstep
      real(c_double), value :: simtime
    end function
    integer(c_int) function c_pdr_remesh(c, tstep, done) bind(C, name="pdr_remesh")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: c
      integer(c_int32_t), value :: tstep
      integer(c_int32_t), intent(out) :: done
    end function
    integer(c_int) function c_pdr_upload_state(c, u1, u2, u3) bind(C, name="pdr_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: u1(*), u2(*), u3(*)
    end function
    integer(c_int) function c_pdr_download_state(c, u1, u2, u3) bind(C, name="pdr_download_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: u1(*), u2(*), u3(*)
    end function
    integer(c_int) function c_pdr_particles_upload(c, xp, np) bind(C, name="pdr_particles_upload")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: c
      real(c_double), intent(in) :: xp(*)
      integer(c_int64_t), value :: np
    end function
    integer(c_int) function c_pdr_particles_download(c, xp, uprk) bind(C, name="pdr_particles_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: xp(*)
      type(c_ptr), value :: uprk
    end function
    integer(c_int) function c_pdr_get_frame(c, bmat, binv) bind(C, name="pdr_get_frame")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: bmat(9), binv(9)
    end function
    integer(c_int) function c_pdr_set_forcing(c, nk, f1, f2, f3) bind(C, name="pdr_set_forcing")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: c
      integer(c_int32_t), value :: nk
      real(c_double), intent(in) :: f1(*), f2(*), f3(*)
    end function
    integer(c_int) function c_pdr_get_forcing(c, nk, f1, f2, f3) bind(C, name="pdr_get_forcing")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: c
      integer(c_int32_t), value :: nk
      real(c_double), intent(out) :: f1(*), f2(*), f3(*)
    end function
    integer(c_int) function c_pdr_init_field(c, seed, u0, kp) bind(C, name="pdr_init_field")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: c
      integer(c_int64_t), value :: seed
      real(c_double), value :: u0, kp
    end function
    integer(c_int) function c_pdr_restart_write(c, path, i64) bind(C, name="pdr_restart_write")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: i64
    end function
    integer(c_int) function c_pdr_restart_read(c, path) bind(C, name="pdr_restart_read")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function c_pdr_restart_read_scaled(c, path, scale) bind(C, name="pdr_restart_read_scaled")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: scale
    end function
    integer(c_int) function c_pdr_spectrum(c, ek, dk, dsk, tot) bind(C, name="pdr_spectrum")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: ek(*), dk(*), dsk(*), tot(6)
    end function
    integer(c_int) function c_pdr_stats(c, q, spec) bind(C, name="pdr_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: q(33), spec(*)
    end function
    integer(c_int) function c_pdr_rs_stats(c, q) bind(C, name="pdr_rs_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: q(25)
    end function
    integer(c_int) function c_pdr_rs_correlation(c, r1i, lint11) bind(C, name="pdr_rs_correlation")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: r1i(*), lint11
    end function
    integer(c_int) function c_pdr_particles_restart_write(c, path, i64) bind(C, name="pdr_particles_restart_write")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: i64
    end function
    integer(c_int) function c_pdr_particles_restart_read(c, path, np) bind(C, name="pdr_particles_restart_read")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: np
    end function
    integer(c_int) function c_pdr_particle_stats_config(c, enable, keep, release) &
        bind(C, name="pdr_particle_stats_config")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: c
      integer(c_int32_t), value :: enable, keep, release
    end function
    integer(c_int) function c_pdr_particle_stats(c, tstep, q) bind(C, name="pdr_particle_stats")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), intent(out) :: tstep
      real(c_double), intent(out) :: q(30)
    end function
    integer(c_int) function c_pdr_particles_download_aux(c, up, xp0) bind(C, name="pdr_particles_download_aux")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, up, xp0
    end function
    integer(c_int) function c_pdr_init_particles(c, seed, np) bind(C, name="pdr_init_particles")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: c
      integer(c_int64_t), value :: seed, np
    end function
    integer(c_int) function c_pdr_set_tracers_in_stage(c, on) bind(C, name="pdr_set_tracers_in_stage")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: c
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function c_pdr_particles_control_rk3(c, stage, tstep, deltat) &
        bind(C, name="pdr_particles_control_rk3")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), value :: stage, tstep
      real(c_double), value :: deltat
    end function
    integer(c_int) function c_pdr_particles_solve(c, form, npin, stage, deltat) bind(C, name="pdr_particles_solve")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), value :: form
      integer(c_int64_t), value :: npin
      integer(c_int32_t), value :: stage
      real(c_double), value :: deltat
    end function
    integer(c_int) function c_pdr_particles_finish_step(c, tstep) bind(C, name="pdr_particles_finish_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: c
      integer(c_int32_t), value :: tstep
    end function
    type(c_ptr) function c_pdr_last_error() bind(C, name="pdr_last_error")
      import :: c_ptr
    end function
  end interface

contains

  !> Hand the run parameters (the g_parameters subset the step uses) to the shim.
  subroutine pdr_Configure ( p, ierr )
    type(pdr_params_t), intent(in) :: p
    integer, intent(inout)         :: ierr
    params = p
    ierr = 0
  end subroutine pdr_Configure

  !> g_RK3Init (src/pg_rk3.f90:46-70): checks the parameters (CFL < 1 as
  !! g_RK3Check) and creates the resident device state.
  subroutine g_RK3Init ( ierr )
    integer, intent(inout) :: ierr
    if ( c_associated(ctx) ) ierr = int(c_pdr_destroy(ctx))
    ierr = int(c_pdr_create(ctx, params))
  end subroutine g_RK3Init

  !> g_RK3Calc (src/pg_rk3.f90:82-129): three stages with the shear-frame update.
  subroutine g_RK3Calc ( tstep, simtime, ierr )
    integer, intent(in)               :: tstep
    real(kind=c_double), intent(inout) :: simtime
    integer, intent(inout)            :: ierr
    ierr = int(c_pdr_step(ctx, int(tstep, c_int32_t), simtime))
  end subroutine g_RK3Calc

  !> g_RK3UpdateB (src/pg_rk3.f90:280-385); rk3dt is unused, as in the reference
  subroutine g_RK3UpdateB ( tstep, rk3dt, simtime, ierr )
    integer, intent(in)             :: tstep
    real(kind=c_double), intent(in) :: rk3dt
    real(kind=c_double), intent(in) :: simtime
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_update_b(ctx, int(tstep, c_int32_t), simtime))
  end subroutine g_RK3UpdateB

  !> f_RK3Stage1 (src/pf_rk3.f90:40-115): remesh check, stage 1, sets dt
  subroutine f_RK3Stage1 ( tstep, time, dt, ierr )
    integer, intent(in)                :: tstep
    real(kind=c_double), intent(in)    :: time
    real(kind=c_double), intent(inout) :: dt
    integer, intent(inout)             :: ierr
    ierr = int(c_pdr_stage(ctx, 1_c_int32_t, int(tstep, c_int32_t), time, dt))
  end subroutine f_RK3Stage1

  !> f_RK3Stage2 (src/pf_rk3.f90:121-155)
  subroutine f_RK3Stage2 ( tstep, time, dt, ierr )
    integer, intent(in)                :: tstep
    real(kind=c_double), intent(in)    :: time
    real(kind=c_double), intent(inout) :: dt
    integer, intent(inout)             :: ierr
    ierr = int(c_pdr_stage(ctx, 2_c_int32_t, int(tstep, c_int32_t), time, dt))
  end subroutine f_RK3Stage2

  !> f_RK3Stage3 (src/pf_rk3.f90:161-200)
  subroutine f_RK3Stage3 ( tstep, time, dt, ierr )
    integer, intent(in)                :: tstep
    real(kind=c_double), intent(in)    :: time
    real(kind=c_double), intent(inout) :: dt
    integer, intent(inout)             :: ierr
    ierr = int(c_pdr_stage(ctx, 3_c_int32_t, int(tstep, c_int32_t), time, dt))
  end subroutine f_RK3Stage3

  !> f_RK3Remesh (src/pf_rk3.f90:3958-4049)
  subroutine f_RK3Remesh ( tstep, ierr )
    integer, intent(in)    :: tstep
    integer, intent(inout) :: ierr
    integer(c_int32_t)     :: done
    ierr = int(c_pdr_remesh(ctx, int(tstep, c_int32_t), done))
  end subroutine f_RK3Remesh

  !> Release the device state (ProgramFinalise, src/pg_control.f90:200-259)
  subroutine g_RK3Finalise ( ierr )
    integer, intent(inout) :: ierr
    ierr = 0
    if ( c_associated(ctx) ) ierr = int(c_pdr_destroy(ctx))
    ctx = c_null_ptr
  end subroutine g_RK3Finalise

  !> Copy u1_3w, u2_3w, u3_3w (pf_arrays layout, FP64) to the device
  subroutine pdr_PutVelocity ( u1, u2, u3, ierr )
    real(kind=c_double), intent(in) :: u1(0:,0:,0:,0:), u2(0:,0:,0:,0:), u3(0:,0:,0:,0:)
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_upload_state(ctx, u1, u2, u3))
  end subroutine pdr_PutVelocity

  !> Copy the velocity back into pf_arrays-layout host arrays
  subroutine pdr_GetVelocity ( u1, u2, u3, ierr )
    real(kind=c_double), intent(out) :: u1(0:,0:,0:,0:), u2(0:,0:,0:,0:), u3(0:,0:,0:,0:)
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_download_state(ctx, u1, u2, u3))
  end subroutine pdr_GetVelocity

  !> Tracer positions particle_xp(1:3, np) (src/pp_arrays.f90:25-34)
  subroutine pdr_PutParticles ( xp, np, ierr )
    real(kind=c_double), intent(in) :: xp(:,:)
    integer(c_int64_t), intent(in)  :: np
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_particles_upload(ctx, xp, np))
  end subroutine pdr_PutParticles

  subroutine pdr_GetParticles ( xp, ierr )
    real(kind=c_double), intent(out) :: xp(:,:)
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_particles_download(ctx, xp, c_null_ptr))
  end subroutine pdr_GetParticles

  !> bmat, binv (src/pg_domain.f90:78-87), Fortran column-major
  subroutine pdr_GetFrame ( bmat, binv, ierr )
    real(kind=c_double), intent(out) :: bmat(3,3), binv(3,3)
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_get_frame(ctx, bmat, binv))
  end subroutine pdr_GetFrame

  !> OU forcing arrays fluid_f1..3(0:1,-NK:NK,-NK:NK,0:NK) (src/pf_force.f90:27-33)
  subroutine pdr_PutForcing ( nk, f1, f2, f3, ierr )
    integer, intent(in)             :: nk
    real(kind=c_double), intent(in) :: f1(0:1,-nk:nk,-nk:nk,0:nk), f2(0:1,-nk:nk,-nk:nk,0:nk), &
                                       f3(0:1,-nk:nk,-nk:nk,0:nk)
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_set_forcing(ctx, int(nk, c_int32_t), f1, f2, f3))
  end subroutine pdr_PutForcing

  subroutine pdr_GetForcing ( nk, f1, f2, f3, ierr )
    integer, intent(in)              :: nk
    real(kind=c_double), intent(out) :: f1(0:1,-nk:nk,-nk:nk,0:nk), f2(0:1,-nk:nk,-nk:nk,0:nk), &
                                        f3(0:1,-nk:nk,-nk:nk,0:nk)
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_get_forcing(ctx, int(nk, c_int32_t), f1, f2, f3))
  end subroutine pdr_GetForcing

  !> Seeded Rogallo initial field, type-10 spectrum (f_InitialiseFluid)
  subroutine pdr_InitField ( seed, u0, kp, ierr )
    integer(c_int64_t), intent(in)  :: seed
    real(kind=c_double), intent(in) :: u0, kp
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_init_field(ctx, seed, u0, kp))
  end subroutine pdr_InitField

  !> g_RestartPetscFluid(WRITE_FILE) (src/pg_restart.f90:97-150)
  subroutine pdr_RestartWrite ( path, int64header, ierr )
    character(len=*), intent(in) :: path
    logical, intent(in)          :: int64header
    integer, intent(inout)       :: ierr
    integer(c_int32_t)           :: flag
    flag = 0
    if ( int64header ) flag = 1
    ierr = int(c_pdr_restart_write(ctx, trim(path)//c_null_char, flag))
  end subroutine pdr_RestartWrite

  !> g_RestartPetscFluid(READ_FILE) (src/pg_restart.f90:57-96)
  subroutine pdr_RestartRead ( path, ierr )
    character(len=*), intent(in) :: path
    integer, intent(inout)       :: ierr
    ierr = int(c_pdr_restart_read(ctx, trim(path)//c_null_char))
  end subroutine pdr_RestartRead

  !> g_RestartPetscFluid(READ_FILE) with F_RESTART_SCALE > 1 and f_InitialiseCopyRestart
  subroutine pdr_RestartReadScaled ( path, scale, ierr )
    character(len=*), intent(in) :: path
    integer, intent(in)          :: scale
    integer, intent(inout)       :: ierr
    ierr = int(c_pdr_restart_read_scaled(ctx, trim(path)//c_null_char, int(scale, c_int32_t)))
  end subroutine pdr_RestartReadScaled

  !> g_RestartPetscParticle(WRITE_FILE) (src/pg_restart.f90:264-318) after
  !! p_ControlPrepareWriteRestart: xp, up, dp, status[, xp0] records
  subroutine pdr_ParticleRestartWrite ( path, int64header, ierr )
    character(len=*), intent(in) :: path
    logical, intent(in)          :: int64header
    integer, intent(inout)       :: ierr
    integer(c_int32_t)           :: flag
    flag = 0
    if ( int64header ) flag = 1
    ierr = int(c_pdr_particles_restart_write(ctx, trim(path)//c_null_char, flag))
  end subroutine pdr_ParticleRestartWrite

  !> Reads a particle restart file of np particles (positions, particle_up,
  !! particle_xp0) back onto the device
  subroutine pdr_ParticleRestartRead ( path, np, ierr )
    character(len=*), intent(in) :: path
    integer, intent(in)          :: np
    integer, intent(inout)       :: ierr
    ierr = int(c_pdr_particles_restart_read(ctx, trim(path)//c_null_char, int(np, c_int64_t)))
  end subroutine pdr_ParticleRestartRead

  !> E(k), D(k) shell sums and totals (f_FluidstatsFsPrimaryQuants)
  subroutine pdr_Spectrum ( ek, dk, dsk, tot, ierr )
    real(kind=c_double), intent(out) :: ek(:), dk(:), dsk(:), tot(6)
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_spectrum(ctx, ek, dk, dsk, tot))
  end subroutine pdr_Spectrum

  !> f_FluidstatsFs on the device: q(1:33) in the order of pdr_stats_t (tke,
  !! fs_u1u1..fs_u2u3, f_ken, f_dis, f_disskew, fn_ken.., fn_int.., tn_tur,
  !! ln_tur, ln_tay, rn_tay, ln_int.., tn_edy, ln_kol, tn_kol, disskew,
  !! C_epsilon); spec(kkmax, 7) = spec_ken, spec_ken1..3, spec_dis,
  !! spec_disskew, dens_mod
  subroutine pdr_Stats ( q, spec, ierr )
    real(kind=c_double), intent(out) :: q(33), spec(:,:)
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_stats(ctx, q, spec))
  end subroutine pdr_Stats

  !> f_FluidstatsRs2D + f_FluidstatsRsCollect on the device: q(1:25) in the
  !! order of pdr_rs_stats_t (rs_u1u1..rs_u2u3, rs_skew1..3, rs_flat1..3,
  !! rs_rmsvort1..3, rs_skewvort1..3, rs_flatvort1..3, omegamax1..3, ucfl)
  subroutine pdr_RsStats ( q, ierr )
    real(kind=c_double), intent(out) :: q(25)
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_rs_stats(ctx, q))
  end subroutine pdr_RsStats

  !> The TS_CORR part of f_FluidstatsRs2D / f_FluidstatsRsCollect: r1i(0:NODES_X/2)
  !! and rs_lint11 of the current state
  subroutine pdr_RsCorrelation ( r1i, lint11, ierr )
    real(kind=c_double), intent(out) :: r1i(0:), lint11
    integer, intent(inout)           :: ierr
    ierr = int(c_pdr_rs_correlation(ctx, r1i, lint11))
  end subroutine pdr_RsCorrelation

  !> Tracer statistics of f_RK3Stage1 (p_StatsPrimary, p_StatsDispersion,
  !! p_StatsRelativeVelocity): track = P_TRACK_PART statistics on/off,
  !! keep = P_KEEP_INITIAL_POSITION, release = P_TS_RELEASE
  subroutine pdr_ParticleStatsConfig ( track, keep, release, ierr )
    integer, intent(in)    :: track, keep, release
    integer, intent(inout) :: ierr
    ierr = int(c_pdr_particle_stats_config(ctx, int(track, c_int32_t), int(keep, c_int32_t), &
                                           int(release, c_int32_t)))
  end subroutine pdr_ParticleStatsConfig

  !> The last p_StatsWrite record: q(1:30) = u1u1, u2u2, u3u3, u1u2, u1u3, u2u3,
  !! tke, ptke, vrel1..3, vrelabs, xrel1..3, dfp1..3, x1x2, x1x3, x2x3, vt(3),
  !! vs(3), Rep, Stk, Ste; tstep = its time step (-1 if none yet)
  subroutine pdr_ParticleStats ( tstep, q, ierr )
    integer, intent(out)             :: tstep
    real(kind=c_double), intent(out) :: q(30)
    integer, intent(inout)           :: ierr
    integer(c_int32_t)               :: t
    ierr = int(c_pdr_particle_stats(ctx, t, q))
    tstep = int(t)
  end subroutine pdr_ParticleStats

  !> particle_up(1:3,np) and particle_xp0(1:3,np) from the device
  subroutine pdr_GetParticlesAux ( up, xp0, ierr )
    real(kind=c_double), intent(out), target :: up(:,:), xp0(:,:)
    integer, intent(inout)                   :: ierr
    ierr = int(c_pdr_particles_download_aux(ctx, c_loc(up), c_loc(xp0)))
  end subroutine pdr_GetParticlesAux

  !> Uniform random tracer positions (P_INIT_RANDOM, src/pp_control.f90:657-686)
  subroutine pdr_InitParticles ( seed, np, ierr )
    integer(c_int64_t), intent(in) :: seed, np
    integer, intent(inout)         :: ierr
    ierr = int(c_pdr_init_particles(ctx, seed, np))
  end subroutine pdr_InitParticles

  !> .true. (default): f_RK3Stage1..3 and g_RK3Calc advance the tracers; .false.:
  !! the caller does, with p_ControlRK3 or p_RK3SolveParticle[BF|R] and then
  !! p_ControlFinishTimeStep after each f_RK3StageN, before g_RK3UpdateB
  subroutine pdr_SetTracersInStage ( on, ierr )
    logical, intent(in)    :: on
    integer, intent(inout) :: ierr
    integer(c_int32_t)     :: flag
    flag = 0
    if ( on ) flag = 1
    ierr = int(c_pdr_set_tracers_in_stage(ctx, flag))
  end subroutine pdr_SetTracersInStage

  !> p_ControlRK3 (src/pp_control.f90:518-580): release handling at
  !! P_TS_RELEASE, then the frame's p_RK3SolveParticle[BF|R] for every tracer
  subroutine p_ControlRK3 ( stage, tstep, deltat, ierr )
    integer, intent(in)             :: stage
    integer, intent(in)             :: tstep
    real(kind=c_double), intent(in) :: deltat
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_particles_control_rk3(ctx, int(stage, c_int32_t), int(tstep, c_int32_t), deltat))
  end subroutine p_ControlRK3

  !> p_ControlFinishTimeStep (src/pp_control.f90:1440-1450): periodic
  !! correction of the positions (no ranks to redistribute to)
  subroutine p_ControlFinishTimeStep ( tstep, ierr )
    integer, intent(in)    :: tstep
    integer, intent(inout) :: ierr
    ierr = int(c_pdr_particles_finish_step(ctx, int(tstep, c_int32_t)))
  end subroutine p_ControlFinishTimeStep

  !> p_RK3SolveParticle (src/pp_rk3.f90:57-165), isotropic runs
  subroutine p_RK3SolveParticle ( npin, stage, deltat, ierr )
    integer, intent(in)             :: npin
    integer, intent(in)             :: stage
    real(kind=c_double), intent(in) :: deltat
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_particles_solve(ctx, 0_c_int32_t, int(npin, c_int64_t), int(stage, c_int32_t), deltat))
  end subroutine p_RK3SolveParticle

  !> p_RK3SolveParticleBF (src/pp_rk3.f90:169-303), NS_BF_ROT shear runs
  subroutine p_RK3SolveParticleBF ( npin, stage, deltat, ierr )
    integer, intent(in)             :: npin
    integer, intent(in)             :: stage
    real(kind=c_double), intent(in) :: deltat
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_particles_solve(ctx, 1_c_int32_t, int(npin, c_int64_t), int(stage, c_int32_t), deltat))
  end subroutine p_RK3SolveParticleBF

  !> p_RK3SolveParticleR (src/pp_rk3.f90:307-403), NS_R_ROT / NS_R_CONV shear runs
  subroutine p_RK3SolveParticleR ( npin, stage, deltat, ierr )
    integer, intent(in)             :: npin
    integer, intent(in)             :: stage
    real(kind=c_double), intent(in) :: deltat
    integer, intent(inout)          :: ierr
    ierr = int(c_pdr_particles_solve(ctx, 2_c_int32_t, int(npin, c_int64_t), int(stage, c_int32_t), deltat))
  end subroutine p_RK3SolveParticleR

  !> Text of the last library error
  function pdr_ErrorMessage () result ( msg )
    character(len=256)          :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(c_pdr_last_error(), p, [256])
    do i = 1, 256
      if ( p(i) == c_null_char ) exit
      msg(i:i) = p(i)
    end do
  end function pdr_ErrorMessage

end module pdr_rk3_amd
