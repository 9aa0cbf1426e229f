/*
 * pandora_amd.h -- C-ABI of the MI355X-native PANDORA 2.0 pseudo-spectral RK3 step.
 *
 * The whole periodic (or Rogallo shear-deforming) box lives on one GPU.  The
 * Fortran host keeps its own drivers and calls these entry points through
 * ISO_C_BINDING (pandora-2.0_amd/fortran/pdr_rk3_amd.f90 is the shim that keeps
 * the reference's g_RK3Calc / g_RK3UpdateB / f_RK3Stage1..3 / f_RK3Remesh /
 * p_ControlRK3 / p_RK3SolveParticle[BF|R] / p_ControlFinishTimeStep signatures).
 *
 * Conventions
 *  - Every function returns PDR_OK (0) or a negative PDR_ERR_* code; the text
 *    of the last error of the calling thread is returned by pdr_last_error().
 *    The Fortran shim maps the code onto the reference's `ierr` argument.
 *  - Host arrays are borrowed for the duration of a call; the context owns all
 *    device memory.  Calls are synchronous with respect to host buffers.
 *  - Wave-space host arrays use the reference's pf_arrays layout: Fortran
 *    arr(0:1, 0:NKK-1, 0:NKJ-1, 0:NKI-1) of real(C_DOUBLE), i.e. (re, im)
 *    fastest, then k, then j, then i (src/pf_arrays.f90:267-280, 761-810),
 *    with k,j = 0..K positive and K+1..2K holding -K..-1, i = 0..KI.
 *  - 3x3 tensors bmat/binv/amat are Fortran column-major (src/pg_domain.f90:78-87).
 *  - Particle arrays are the reference's particle_xp(1:3, np) (src/pp_arrays.f90:25-34).
 *  - One context per host thread; a context is not thread-safe (neither is the
 *    reference's module state).
 */
#ifndef PANDORA_AMD_H
#define PANDORA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDR_OK               0
#define PDR_ERR_ARG         -1   /* invalid argument or parameter combination  */
#define PDR_ERR_HIP         -2   /* HIP runtime error (message has details)    */
#define PDR_ERR_NOMEM       -3   /* device or host allocation failed           */
#define PDR_ERR_UNSUPPORTED -4   /* option the reference has but this build lacks */
#define PDR_ERR_IO          -5   /* restart file could not be read / written   */
#define PDR_ERR_STATE       -6   /* call out of order (e.g. no particles set)  */
#define PDR_ERR_NONFINITE   -7   /* a NaN/Inf was detected in the state        */

#define PDR_FP64        0        /* every field and transform in FP64 (parity form) */
#define PDR_FP32_MIXED  1        /* FFTs, products and state in FP32; frame, dt,
                                    viscous factors and tracer positions in FP64 */

/* Reference parameter values (src/pg_parameters.f90) */
#define PDR_F_ISOTROPIC           1
#define PDR_F_SHEAR_DNS           2
#define PDR_F_SHEAR_RDT_VISCOUS   3
#define PDR_F_SHEAR_RDT_INVISCID  4
#define PDR_F_ISOTROPIC_DEFORMED  5
#define PDR_NS_BF_ROT             1
#define PDR_NS_R_CONV             2
#define PDR_NS_R_ROT              3
#define PDR_DEALIASING_TWOTHIRD   1
#define PDR_DEALIASING_SPHERICAL  2
#define PDR_DEALIASING_LES        3
#define PDR_RK3_TIMESTEP_DECIDE   1
#define PDR_RK3_TIMESTEP_FIXED    2
#define PDR_RK3_TIMESTEP_CFL      3

typedef struct pdr_ctx pdr_ctx;

/* Run parameters: the subset of the reference's namelist globals
 * (src/pg_parameters.f90:24-300, read in src/pg_control.f90:385-436)
 * that the time step uses. */
typedef struct pdr_params {
  int32_t nodes_x, nodes_y, nodes_z; /* NODES_X/Y/Z: 2^a * {1,3,5} (radix 2/3/5 FFTs)  */
  int32_t f_type;                    /* F_TYPE: 1 isotropic, 2 shear DNS, 3/4 shear RDT,
                                        5 isotropic on a deformed grid                   */
  int32_t ns_deformed;               /* NS_DEFORMED: 1 BF_ROT, 2 R_CONV, 3 R_ROT        */
  int32_t dealiasing;                /* DEALIASING: 1 cubic 2/3 rule by storage; 2 also
                                        spherical |k_lab| <= KMAX; 3 LES |k_lab| <=
                                        LES_CUTOFF KMAX (KMAX = min(KI, KJ, KK))         */
  int32_t rk3_timestep;              /* RK3_TIMESTEP: 1 DECIDE, 2 FIXED, 3 CFL          */
  int32_t f_remesh_t;                /* F_REMESH_T (shear, DECIDE)                      */
  int32_t b110;                      /* B110: box stretch in x1 (<= 0 taken as 1; other
                                        than 1 only with F_TYPE 5)                       */
  int32_t p_interp_order;            /* P_INTERP_ORDER: 1..8 (reference default 4)      */
  int32_t precision;                 /* PDR_FP64 or PDR_FP32_MIXED                      */
  int32_t force_stop;                /* FORCE_STOP: OU forcing while tstep < FORCE_STOP */
  double  rk3_dt;                    /* RK3_DT  (FIXED)                                 */
  double  rk3_cfl;                   /* RK3_CFL (CFL, DECIDE-isotropic)                 */
  double  f_nu;                      /* F_NU                                            */
  double  f_deform_s;                /* F_DEFORM_S: amat(1,3) = dU1/dx3                 */
  double  force_tf;                  /* FORCE_TF: OU time scale; dt cap (DECIDE-iso)    */
  double  force_kf;                  /* FORCE_KF: forced modes 0 < |k| <= FORCE_KF      */
  double  force_sigma;               /* FORCE_SIGMA: OU amplitude                       */
  int64_t force_seed;                /* seed of the forcing draws (counter-based RNG)   */
  double  les_cutoff;                /* LES_CUTOFF (DEALIASING = 3; reference default 1) */
  int32_t b220, b330;                /* B220, B330: box stretch in x2, x3 (as B110)     */
} pdr_params;

/* ---- library / context -------------------------------------------------- */

/* Version string of this library. No reference counterpart. */
const char* pdr_version(void);

/* Text of the last error on this thread ("" if none). No reference counterpart
 * (the reference prints and calls MPI_Abort). */
const char* pdr_last_error(void);

/* Create a context: validates parameters, computes the domain (g_DomainCalc,
 * src/pg_domain.f90:147-243), allocates the resident device state and work
 * buffers (f_ArraysInit, src/pf_arrays.f90:165) and plans the transforms
 * (f_FftwInit, src/pf_fftw.f90:98-315).  State and RK registers start at zero,
 * bmat/binv/amat at their g_DomainCalc values, dt at 99999.0 (src/pg_rk3.f90:92). */
int pdr_create(pdr_ctx** out, const pdr_params* params);

/* Release all device and host memory of a context (ProgramFinalise,
 * src/pg_control.f90:200-259). */
int pdr_destroy(pdr_ctx* ctx);

/* Sizes: out[0..7] = NODES_KI, NODES_KJ, NODES_KK, Mw (modes per wave field),
 * M1 (modes per 1r2w field), kkmax (spectrum bins), device bytes held,
 * number of tracers.  (src/pg_domain.f90:220-228, src/pf_fluidstats.f90:343-374) */
int pdr_sizes(const pdr_ctx* ctx, int64_t out[8]);

/* ---- state ---------------------------------------------------------------- */

/* Copy the velocity u1_3w, u2_3w, u3_3w in (pf_arrays layout, FP64) to the device
 * (the DMDAVecGetArray/VecLoad path of src/pg_restart.f90:39-97). */
int pdr_upload_state(pdr_ctx* ctx, const double* u1, const double* u2, const double* u3);
/* Copy the velocity back to the host (pf_arrays layout, FP64). */
int pdr_download_state(pdr_ctx* ctx, double* u1, double* u2, double* u3);
/* Low-storage RK registers rk3_3w_1..3 (src/pf_arrays.f90:777-781). */
int pdr_upload_rk3(pdr_ctx* ctx, const double* r1, const double* r2, const double* r3);
int pdr_download_rk3(pdr_ctx* ctx, double* r1, double* r2, double* r3);

/* Shear-frame tensors bmat, binv (Fortran column-major, FP64; g_domain module
 * variables, src/pg_domain.f90:78-87). */
int pdr_set_frame(pdr_ctx* ctx, const double bmat[9], const double binv[9]);
int pdr_get_frame(const pdr_ctx* ctx, double bmat[9], double binv[9]);

/* Ornstein-Uhlenbeck forcing state fluid_f1..3 (src/pf_force.f90:27-33): three
 * arrays in the reference's order f(0:1, -NK:NK, -NK:NK, 0:NK) (re/im, k, j, i;
 * Fortran column-major), each 2 (2NK+1)^2 (NK+1) doubles, NK = aint(FORCE_KF).
 * The state starts at zero (f_ForceAlloc, src/pf_force.f90:275-303) and is
 * advanced on the device once per stage (f_ForceCalc, src/pf_force.f90:106-268,
 * called from f_RK3SolveNavierStokes, src/pf_rk3.f90:2063) while tstep <
 * FORCE_STOP; it is applied in the isotropic FormF (src/pf_rk3.f90:2421-2467).
 * nk must equal the context's NK.  Contexts without forcing (FORCE_STOP <= 0
 * or F_TYPE != 1) return PDR_ERR_STATE. */
int pdr_set_forcing(pdr_ctx* ctx, int32_t nk, const double* f1, const double* f2, const double* f3);
int pdr_get_forcing(pdr_ctx* ctx, int32_t nk, double* f1, double* f2, double* f3);

/* sim_dt, the SAVEd step width of g_RK3Calc (src/pg_rk3.f90:92). */
int pdr_set_dt(pdr_ctx* ctx, double dt);
int pdr_get_dt(const pdr_ctx* ctx, double* dt);

/* ---- time stepping ------------------------------------------------------- */

/* One Runge-Kutta stage: f_RK3Stage1/2/3(tstep, time, dt, ierr)
 * (src/pf_rk3.f90:40, 121, 161).  Stage 1 also does the shear remesh check
 * (f_RK3Remesh, src/pf_rk3.f90:3958) and sets *dt (f_RK3Tstep, 3852-3948).
 * Tracers, if set, are advanced inside the stage as f_RK3RealSpace does with
 * p_ControlRK3 + p_ControlFinishTimeStep (src/pf_rk3.f90:1145-1176), unless
 * pdr_set_tracers_in_stage(ctx, 0) leaves that to the caller.  The
 * state-space statistics calls of the reference stage are not made (see
 * pdr_spectrum). */
int pdr_stage(pdr_ctx* ctx, int32_t stage, int32_t tstep, double time, double* dt);

/* Remesh check and remesh of the Rogallo frame: f_RK3Remesh(tstep, ierr)
 * (src/pf_rk3.f90:3958-4049) with f_RK3RemeshWaveSpace (4056-4178) and
 * p_ControlRemesh (src/pp_control.f90:1658-1743).  pdr_stage(1, ...) calls it
 * first; *done (may be NULL) is set to 1 if this tstep remeshed. */
int pdr_remesh(pdr_ctx* ctx, int32_t tstep, int32_t* done);

/* Spherical / LES truncation of the state: f_ArraysDealiasSpherical and
 * f_ArraysLESCutoff (src/pf_arrays.f90:1865-1987): u1..u3 of every mode whose
 * laboratory wavenumber |bmat^T k| exceeds KMAX (DEALIASING = 2) or
 * LES_CUTOFF * KMAX (DEALIASING = 3) are set to zero; no-op for DEALIASING = 1.
 * pdr_stage applies it where f_RK3PrepareRealSpace and f_RK3SolveNavierStokes
 * do (src/pf_rk3.f90:269-277, 2101-2109), pdr_init_field and pdr_restart_read
 * where f_InitialiseInit does (src/pf_initialise.f90:75-83). */
int pdr_dealias(pdr_ctx* ctx);

/* Shear-frame update after a stage: g_RK3UpdateB(tstep, dt, simtime)
 * (src/pg_rk3.f90:280-385).  FP64 on the host. */
int pdr_update_b(pdr_ctx* ctx, int32_t tstep, double simtime);

/* One full time step: g_RK3Calc(tstep, simtime, ierr) (src/pg_rk3.f90:82-129). */
int pdr_step(pdr_ctx* ctx, int32_t tstep, double* simtime);

/* nsteps consecutive g_RK3Calc calls (tstep0 .. tstep0+nsteps-1) without a host
 * round trip between steps: the body of TIMELOOP (src/pg_control.f90:177-194)
 * minus its reporting. */
int pdr_run(pdr_ctx* ctx, int32_t tstep0, int32_t nsteps, double* simtime);

/* Wait for all queued device work of the context. */
int pdr_synchronize(pdr_ctx* ctx);

/* Checks that the velocity state is finite (no NaN/Inf); returns PDR_OK or
 * PDR_ERR_NONFINITE.  Replaces the reference's absence of such a check
 * (failure detection for long runs). */
int pdr_check_finite(pdr_ctx* ctx);

/* ---- tracers ------------------------------------------------------------- */

/* Set np passive tracers at positions xp(1:3, np) (moving frame), zeroing
 * particle_up_rk3: p_ControlInitialiseParticles' result (src/pp_control.f90:589-727)
 * handed over by the host.  np = 0 removes them. */
int pdr_particles_upload(pdr_ctx* ctx, const double* xp, int64_t np);
/* Read positions xp(1:3,np) and, if up_rk3 != NULL, particle_up_rk3(1:3,np). */
int pdr_particles_download(pdr_ctx* ctx, double* xp, double* up_rk3);
/* particle_up(1:3,np) (the laboratory-frame fluid velocity at the tracer from
 * the last stage: uf, or binv uf with shear, src/pp_rk3.f90:124, 268) and
 * particle_xp0(1:3,np) (P_KEEP_INITIAL_POSITION); either pointer may be NULL.
 * The reference keeps both in p_arrays (src/pp_arrays.f90:25-42) and its
 * restart; these move them in and out.  particle_up starts at zero on upload. */
int pdr_particles_download_aux(pdr_ctx* ctx, double* up, double* xp0);
/* Re-sort the device slots of the tracers by the (z, y) grid row of their
 * position every `every` steps (stage 1; 0 = never; default 10, and once at
 * upload), so that the interpolation gathers of neighbouring lanes share cache
 * lines.  No reference counterpart: the reference's particle order is its MPI
 * ownership order.  Results do not depend on it; every transfer to and from the
 * host uses the caller's numbering. */
int pdr_set_particle_sort(pdr_ctx* ctx, int32_t every);
int pdr_particles_upload_aux(pdr_ctx* ctx, const double* up, const double* xp0);

/* Who advances the tracers.  on != 0 (default): pdr_stage and pdr_step do, as
 * f_RK3RealSpace calls p_ControlRK3 and p_ControlFinishTimeStep inside the
 * fluid stage.  on = 0: pdr_stage only writes the stage's real-space velocity
 * for them (u1_3r..u3_3r, src/pf_rk3.f90:1108-1115) and the caller advances
 * them with the three entry points below, after pdr_stage and before
 * pdr_update_b (the interpolation uses the frame of the stage either way).
 * pdr_step and pdr_run do not advance tracers in that mode.  No reference
 * counterpart (the reference's particle driver is called from inside
 * f_RK3RealSpace). */
int pdr_set_tracers_in_stage(pdr_ctx* ctx, int32_t on);

/* p_ControlRK3(stage, tstep, deltat, ierr) (src/pp_control.f90:518-580): at
 * the release step (stage 1, tstep = P_TS_RELEASE) particle_up_rk3 = 0 and
 * particle_xp0 = particle_xp (p_ControlInitPartVelocity[BF|R]); then the
 * interpolation of the stage's velocity at every tracer, particle_up and the
 * RK3 update with the frame's branch: p_RK3SolveParticle (isotropic),
 * p_RK3SolveParticleBF (NS_BF_ROT) or p_RK3SolveParticleR (NS_R_ROT,
 * NS_R_CONV).  Positions are left unwrapped until pdr_particles_finish_step.
 * Needs pdr_set_tracers_in_stage(ctx, 0) and the tracer field of this stage and
 * tstep (PDR_ERR_STATE otherwise). */
int pdr_particles_control_rk3(pdr_ctx* ctx, int32_t stage, int32_t tstep, double deltat);

/* p_RK3SolveParticle / p_RK3SolveParticleBF / p_RK3SolveParticleR(npin, stage,
 * deltat, ierr) (src/pp_rk3.f90:57-165, 169-303, 307-403): the advance alone
 * (interpolation, particle_up, particle_up_rk3 and position update; no release
 * handling, no wrap).  form = 0, 1, 2 selects the routine and must be the
 * run's frame branch (PDR_ERR_ARG otherwise); npin must equal the number of
 * tracers set: one box, so one owner of every tracer. */
int pdr_particles_solve(pdr_ctx* ctx, int32_t form, int64_t npin, int32_t stage, double deltat);

/* p_ControlFinishTimeStep(tstep, ierr) (src/pp_control.f90:1440-1450): the
 * periodic correction of every position (p_MPICorrectPosition,
 * src/pp_mpi.f90:1011-1057; particle_xp0 moves with it, src/pp_control.f90:
 * 1574-1585); the MPI redistribution has nothing to do on one box.  After a
 * stage-1 advance with tracer statistics on it also takes
 * p_StatsRelativeVelocity, as f_RK3Stage1 does next (src/pf_rk3.f90:88-97). */
int pdr_particles_finish_step(pdr_ctx* ctx, int32_t tstep);

/* Tracer statistics: p_StatsPrimary, p_StatsDispersion and
 * p_StatsRelativeVelocity (src/pp_stats.f90:119-390), called by f_RK3Stage1
 * (src/pf_rk3.f90:62-99).  enable = P_TRACK_PART statistics; keep_initial =
 * P_KEEP_INITIAL_POSITION (particle_xp0 = particle_xp now, at upload and at the
 * release step; it follows the periodic wraps and the remesh of its particle,
 * src/pp_control.f90:1540-1588, src/pp_mpi.f90:1095-1129); ts_release =
 * P_TS_RELEASE: particles move from tstep = ts_release on (particle_up_rk3 and
 * particle_xp0 reset at that step, src/pp_control.f90:378-510) and the
 * statistics are taken in stage 1 of every tstep > ts_release, on the device
 * without a host synchronisation.  Resets the saved displacement used for dfp. */
int pdr_particle_stats_config(pdr_ctx* ctx, int32_t enable, int32_t keep_initial, int32_t ts_release);

/* The last record, in the order of p_StatsWrite (src/pp_stats.f90:379-390).
 * Tracers have dp = 0, so the particle kinetic energy, particle Reynolds number
 * and Stokes numbers are 0; without gravity vs (upgrav) is 0.  vt is the mean
 * particle_up (p_RK3AbsoluteVelocity), xrel the mean squared laboratory-frame
 * displacement binv (x - x0) and dfp = (xrel - xrel_previous) / (2 dt). */
typedef struct pdr_pstats_t {
  double rs_u1u1, rs_u2u2, rs_u3u3, rs_u1u2, rs_u1u3, rs_u2u3, tke, pken;
  double vrel1, vrel2, vrel3, vrel_abs;
  double xrel1, xrel2, xrel3, dfp1, dfp2, dfp3, x1x2, x1x3, x2x3;
  double vt1, vt2, vt3, vs1, vs2, vs3, rep, stk, ste;
} pdr_pstats_t;
/* *tstep = the record's tstep (-1 if none yet). */
int pdr_particle_stats(pdr_ctx* ctx, int32_t* tstep, pdr_pstats_t* out);

/* ---- initial conditions, restart, statistics ----------------------------- */

/* Seeded initial field: Rogallo alpha/beta solenoidal construction of
 * f_InitialiseFluid (src/pf_initialise.f90:97-512) with the spectrum of
 * f_InitialiseSpec type 10, Ahmed/Elghobashi (src/pf_initialise.f90:1081-1083):
 * E(k) = 16 sqrt(2/pi) (3/2) U0 k^4/kp^5 exp(-2k^2/kp^2).  The gfortran
 * intrinsic RNG is replaced by a counter-based generator, so fields are
 * reproducible from (seed) on any machine but are not the reference's stream.
 * Hermitian symmetry on the i = 0 plane is enforced as in the reference.  A
 * stretched box (B110 B220 B330 > 1, F_TYPE 5) returns PDR_ERR_UNSUPPORTED: the
 * reference's stretched variant of the spectrum (src/pf_initialise.f90:67-71,
 * 550-632) is not built, such a run starts from an uploaded state. */
int pdr_init_field(pdr_ctx* ctx, uint64_t seed, double init_u0, double init_kp);

/* Host-only form of pdr_init_field: fills u1..u3 (pf_arrays layout for an
 * nx*ny*nz grid, 2*NKI*NKJ*NKK doubles each) without a context or a GPU, for
 * hosts that build the initial state in their own arrays (f_InitialiseFluid,
 * src/pf_initialise.f90:97-512). */
int pdr_init_field_host(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double init_u0,
                        double init_kp, double* u1, double* u2, double* u3);

/* Uniform random tracer positions in [-dx/2, L-dx/2)^3 (P_INIT_RANDOM,
 * src/pp_control.f90:657-686) from a counter-based generator. */
int pdr_init_particles(pdr_ctx* ctx, uint64_t seed, int64_t np);

/* PETSc-binary restart of u1_3w, u2_3w, u3_3w: g_RestartPetscFluid
 * (src/pg_restart.f90:39-194).  Three records of {int32 1211214 (VEC_FILE_CLASSID),
 * PetscInt length, big-endian float64 data in DMDA natural order}; int64_header
 * selects 64-bit PetscInt (--with-64-bit-indices builds). */
int pdr_restart_write(pdr_ctx* ctx, const char* path, int32_t int64_header);
/* Reads either header width (detected from the file size). */
int pdr_restart_read(pdr_ctx* ctx, const char* path);

/* Host-only restart I/O of three n-double records (n = 2*NKI*NKJ*NKK), same
 * format as pdr_restart_write/read; no context or GPU needed. */
int pdr_restart_write_host(const char* path, int64_t n, const double* u1, const double* u2,
                           const double* u3, int32_t int64_header);
int pdr_restart_read_host(const char* path, int64_t n, double* u1, double* u2, double* u3);

/* Particle restart: g_RestartPetscParticle(WRITE_FILE) (src/pg_restart.f90:264-318)
 * of the arrays p_ControlPrepareWriteRestart gathers (src/pp_control.f90:1219-1430):
 * PETSc Vec records of np doubles, in order particle_xp(1,:), (2,:), (3,:),
 * particle_up(1..3,:), the diameter (0 for tracers), particle_status(1,:) (the
 * particle number 1..np, src/pp_control.f90:646, 690) and, with
 * P_KEEP_INITIAL_POSITION, particle_xp0(1..3,:): 8 or 11 records, the
 * particles in the caller's numbering.  PDR_ERR_STATE without particles. */
int pdr_particles_restart_write(pdr_ctx* ctx, const char* path, int32_t int64_header);
/* Reads such a file of np particles (either header width, 8 or 11 records),
 * placing each entry by its particle number: positions, particle_up and, when
 * the file has them and the context keeps initial positions, particle_xp0;
 * particle_up_rk3 starts at zero.  No reference counterpart: the reference's
 * read branch is commented out (src/pg_restart.f90:220-262). */
int pdr_particles_restart_read(pdr_ctx* ctx, const char* path, int64_t np);
/* Host-only forms (no context or GPU): xp, up, xp0 as (3, np) arrays; xp0 may
 * be NULL (8 records); *has_xp0 (may be NULL) tells whether the file had them. */
int pdr_particles_restart_write_host(const char* path, int64_t np, const double* xp, const double* up,
                                     const double* xp0, int32_t int64_header);
int pdr_particles_restart_read_host(const char* path, int64_t np, double* xp, double* up, double* xp0,
                                    int32_t* has_xp0);

/* Scaled restart (F_RESTART_SCALE > 1): g_RestartPetscFluid of a file written on
 * the grid (nx, ny, nz) / scale, then f_InitialiseCopyRestart
 * (src/pf_initialise.f90:1184-1299, src/pf_arrays.f90:1047-1312): each coarse
 * mode keeps its signed wavenumber on the (nx, ny, nz) grid, every other mode is
 * zero.  The reference copies the negative-j half only on ranks whose j range
 * starts above the positive half (IFJISPOS), so its result depends on the
 * decomposition; this copies both halves, as a decomposition with a rank
 * boundary there does.  Host-only form: fills the fine u1..u3. */
int pdr_restart_read_scaled_host(const char* path, int32_t nx, int32_t ny, int32_t nz, int32_t scale,
                                 double* u1, double* u2, double* u3);
/* Device form: the copy above into the context's state, then the truncation
 * of pdr_restart_read. */
int pdr_restart_read_scaled(pdr_ctx* ctx, const char* path, int32_t scale);

/* Wave-space statistics of the current velocity, computed on the device:
 * f_FluidstatsFs (src/pf_fluidstats.f90:1293-1361).  Modes are taken to the
 * laboratory frame as in f_FluidstatsFsPrimaryQuantsBFRot (1586-1792; the
 * identity for isotropic runs), binned by |k_lab| in [kk-1/2, kk+1/2),
 * kk = 1..kkmax, and summed in FP64 with compensated accumulation; the mode
 * density (696-774, 1977-2049), the normalised spectra and totals (1547-1575)
 * and the secondary quantities (f_FluidstatsFsSecondaryQuants, 2491-2558)
 * follow on the host.  Field names are the reference's module variables. */
typedef struct pdr_stats_t {
  double tke;                                   /* 1/2 (u1u1 + u2u2 + u3u3)      */
  double u1u1, u2u2, u3u3, u1u2, u1u3, u2u3;    /* fs_u1u1 ... Reynolds sums     */
  double f_ken, f_dis, f_disskew;               /* un-normalised totals          */
  double fn_ken, fn_ken1, fn_ken2, fn_ken3;     /* normalised kinetic energies   */
  double fn_dis, fn_disskew;                    /* normalised dissipation (eps)  */
  double fn_int, fn_int1, fn_int2, fn_int3;     /* r.m.s. velocities u'          */
  double tn_tur, ln_tur;                        /* t_eps, l_eps                  */
  double ln_tay, rn_tay;                        /* Taylor scale, Re_lambda       */
  double ln_int, ln_int1, ln_int2, ln_int3;     /* integral length scales        */
  double tn_edy, ln_kol, tn_kol;                /* eddy turnover, eta, tau_eta   */
  double disskew, c_epsilon;                    /* dissipation skewness, C_eps   */
} pdr_stats_t;

/* spec (may be NULL): 7 * kkmax doubles, the normalised spectra spec_ken,
 * spec_ken1..3, spec_dis, spec_disskew and dens_mod, bin kk at index kk-1. */
int pdr_stats(pdr_ctx* ctx, pdr_stats_t* out, double* spec);

/* pdr_stats without any device allocation.  No counterpart in the reference:
 * an entry point for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The workgroup partials live in a fixed 16 MiB device array of the library
 * (2^21 doubles per GPU, held until the process ends) instead of the
 * context's lazily allocated buffer; boxes whose statistics rows do not fit
 * are refused with PDR_ERR_UNSUPPORTED.  spec (7 * kkmax doubles, as above) is
 * required: it is also the working buffer.  Results equal pdr_stats' to
 * rounding (the partials may be summed over fewer workgroups).  All calls on
 * one GPU share the array: calls that overlap on one GPU race. */
int pdr_stats_noalloc(pdr_ctx* ctx, pdr_stats_t* out, double* spec);

/* Streamwise 1-D spectra of f_FluidstatsFs2D + f_FluidstatsFsCollect
 * (src/pf_fluidstats.f90:1154-1283) for the current state: spec_1i[i],
 * spec_2i[i], i = 0..NKI-1, the sums over the (y, z) points of 2 |u_lab,c|^2 of
 * the y- and z-transformed velocity divided by NODES_X (times binv(1,1)^3 with
 * shear), as the reference forms them in stage 1 from the 1r1w arrays
 * (src/pf_rk3.f90:1504-1590).  Computed from the wave state in the statistics
 * pass (Parseval: 2 NY NZ / NX^2 sum_{j,k} |u_lab(i,j,k)|^2).  h (may be NULL)
 * receives the f_FluidstatsHomogeneous record (src/pf_fluidstats.f90:
 * 2568-2936) the reference writes for shear runs; its ln_int_21, ln_int_12 and
 * ln_int_13 are never assigned, so lstar = ln_int_11 / 0 (infinite) there and
 * here.  Either spectrum pointer may be NULL. */
typedef struct pdr_homog_t {
  double sstar, kstar, lstar;                   /* S* = S tn_edy, K*, L*          */
  double ln_int1, ln_int2, ln_int3;
  double ln_int_11, ln_int_12, ln_int_13;       /* pi E11(0) / (2 u1u1), 0, 0     */
  double fn_ken1, fn_ken2, fn_ken3;
  double fn_int1, fn_int2, fn_int3;
} pdr_homog_t;
int pdr_spectra_1d(pdr_ctx* ctx, double* spec_1i, double* spec_2i, pdr_homog_t* h);

/* Real-space statistics of the current velocity: f_FluidstatsRs2D and
 * f_FluidstatsRsCollect (src/pf_fluidstats.f90:827-951, 962-1146), which the
 * reference computes on the real-space planes of stage 1 (src/pf_rk3.f90:
 * 1075-1078, 1124-1127).  Velocities are laboratory-frame (binv u; u for
 * isotropic runs), vorticities those of the real-space planes.  Moments are
 * normalised by NX NY NZ as there, including rs_rmsvort1, which the reference
 * leaves unnormalised (src/pf_fluidstats.f90:1068) and this entry point does
 * likewise.  One extra pass (z, y and x transforms of the state); the two-point
 * correlation rs_lint11 of TS_CORR steps is not computed. */
typedef struct pdr_rs_stats_t {
  double rs_u1u1, rs_u2u2, rs_u3u3, rs_u1u2, rs_u1u3, rs_u2u3;
  double rs_skew1, rs_skew2, rs_skew3, rs_flat1, rs_flat2, rs_flat3;
  double rs_rmsvort1, rs_rmsvort2, rs_rmsvort3;
  double rs_skewvort1, rs_skewvort2, rs_skewvort3;
  double rs_flatvort1, rs_flatvort2, rs_flatvort3;
  double omegamax1, omegamax2, omegamax3;
  double ucfl;                                  /* max(|u1|+|u2|+|u3|), lab frame */
} pdr_rs_stats_t;
int pdr_rs_stats(pdr_ctx* ctx, pdr_rs_stats_t* out);

/* The TS_CORR two-point correlation of f_FluidstatsRs2D / f_FluidstatsRsCollect
 * (src/pf_fluidstats.f90:933-945, 1092-1106) of the current state:
 * r1i[NX/2 + 1] = sum over grid points of u1_lab(x) u1(x + s) / (rs_u1u1 NX NY NZ)
 * (laboratory u1 at x, stored u1 at x + s, x periodic) and lint11 = dx_lab
 * times that sum over s = 0 .. (NX-1)/2 + 1.  Evaluated from per-i
 * correlation sums of the wave-space statistics pass (an exact trigonometric
 * sum on the grid), with no inverse transform. */
int pdr_rs_correlation(pdr_ctx* ctx, double* r1i, double* lint11);

/* Statistics of stage 1 inside the step, as f_RK3Stage1 takes them
 * (src/pf_rk3.f90:58-60 and f_RK3RealSpace, 1073-1078, 1124-1143): with
 * every > 0, stage 1 of each tstep with (tstep - 1) mod every = 0 runs the
 * wave-space statistics pass on the state after the remesh check
 * (f_FluidstatsFs((tstep-1))) and accumulates the real-space moments in its own
 * x pass (f_FluidstatsRs2D / RsCollect on the stage's planes), so no extra
 * transform is made; the record stays on the device until read.  every = 0
 * turns it off.  The record buffer is allocated here, never during a step.
 * The 1-D spectra come from the same wave-space pass (Parseval, see
 * pdr_spectra_1d): for DEALIASING 2 and 3 they are those of the state before
 * the stage's entry truncation, which the reference applies first. */
int pdr_set_stage_stats(pdr_ctx* ctx, int32_t every);

/* The last stage-1 record: *tstep = its statistics step (the reference's
 * tstep - 1; -1 if none yet), then the quantities of pdr_stats, pdr_rs_stats
 * and pdr_spectra_1d for that state.  Any output pointer may be NULL. */
int pdr_stage_stats(pdr_ctx* ctx, int32_t* tstep, pdr_stats_t* fs, double* spec, pdr_rs_stats_t* rs,
                    double* spec_1i, double* spec_2i, pdr_homog_t* h);
/* The same correlation from the stage-1 record (as pdr_rs_correlation of the
 * state the record's stage 1 started from); *tstep as in pdr_stage_stats. */
int pdr_stage_correlation(pdr_ctx* ctx, int32_t* tstep, double* r1i, double* lint11);

/* Shell sums before mode-density normalisation from the same device pass:
 * ek, dk, dsk have kkmax entries (energy, 2 nu k^2 E, 4 nu^2 k^4 E);
 * tot[6] = {f_ken, f_dis, f_disskew, u1u1, u2u2, u3u3}. */
int pdr_spectrum(pdr_ctx* ctx, double* ek, double* dk, double* dsk, double tot[6]);

/* ---- batched ensembles ---------------------------------------------------- */

/* Many independent realisations of uneven sizes advanced together: each RK3
 * pass is one launch over the tiles of every member (per-member work
 * descriptors) and the 15 launches of a step are captured once into a HIP graph
 * that pdr_ensemble_run replays.  No reference counterpart (the reference runs
 * one box per job); each member runs the device code pdr_step runs for it, with
 * the same arguments (compiled into other kernels: equal to rounding).
 * Members are contexts from pdr_create, not owned by the ensemble, all of one
 * precision, cubic with 32..256 points (a specialised length), isotropic with
 * RK3_TIMESTEP = FIXED, unforced and without tracers; others give
 * PDR_ERR_UNSUPPORTED.  Their buffers must not be reallocated while the
 * ensemble exists (state uploads and statistics are fine). */
typedef struct pdr_ensemble pdr_ensemble;
int pdr_ensemble_create(pdr_ensemble** out, pdr_ctx* const* members, int32_t count);
/* nsteps g_RK3Calc steps of every member; simtime[k] of member k advances by
 * its RK3_DT per step (tstep0 is accepted for symmetry with pdr_run). */
int pdr_ensemble_run(pdr_ensemble* e, int32_t tstep0, int32_t nsteps, double* simtime);
/* Fields per workgroup of the inverse / forward y passes (1, 2, 3, 6 / 1, 3;
 * 0 = each member's own choice); rebuilds the graph. */
int pdr_ensemble_set_split(pdr_ensemble* e, int32_t fpb_y, int32_t fpb_yf);
/* Times `steps` steps for every split, keeps the fastest and restores the
 * members' state and RK registers; returns the choice in *fpb_y, *fpb_yf. */
int pdr_ensemble_autotune(pdr_ensemble* e, int32_t steps, int32_t* fpb_y, int32_t* fpb_yf);
int pdr_ensemble_destroy(pdr_ensemble* e);

/* ---- diagnostics ---------------------------------------------------------- */

/* Copy 1r2w intermediate field f (0..5 = u1..u6_1r2w after the z pass of
 * f_RK3PrepareRealSpace, src/pf_rk3.f90:205-948) to the host as
 * (0:1, 0:NKJ-1, 0:NZ-1, 0:NKI-1) FP64.  Runs pass 1 of a stage on the
 * current state; the state is not modified. */
int pdr_debug_prepare(pdr_ctx* ctx, int32_t field, double* out);

/* Which kernel forms this context runs (no reference counterpart; for tests
 * that must know which code a box size exercises): out[0] 1 = length-
 * specialised kernels, 0 = runtime-radix kernels; out[1] z planes per x-pencil
 * chunk (NZ = whole box); out[2] 1 = two chunks in flight on two streams;
 * out[3] 1 = fused z-forward + FormF kernel; with the specialised kernels
 * out[4] x rows per x-pass workgroup, out[5] 1 = x tile twiddles read from the
 * global table, out[6] 1 = register z inverse (velocity tile loaded once),
 * out[7] lines per z-inverse tile (0 for out[4..7] on the runtime-radix path). */
int pdr_kernel_forms(const pdr_ctx* ctx, int32_t out[8]);

/* Measured device copy bandwidth (read + write bytes per second, GB/s) of a
 * plain 16-byte-vector copy of `bytes` bytes, timed over `iters` launches.
 * No reference counterpart: the roofline reference for the kernels. */
int pdr_copy_bandwidth(int64_t bytes, int32_t iters, double* gbps);

/* HBM ceiling probe: kind 0 copy, 1 read, 2 write, 3 triad (d = a + s b) over
 * `bytes`-byte buffers, each thread keeping eight 16-byte loads in flight and
 * storing non-temporally; *gbps = bytes moved (reads + writes) per second over
 * `iters` launches.  No reference counterpart: the measured HBM ceiling that
 * kernel fractions are quoted against beside the nominal 8 TB/s. */
int pdr_hbm_probe(int32_t kind, int64_t bytes, int32_t iters, double* gbps);

/* Per-kernel timing with HIP events on the context's stream.  enable != 0
 * starts collection; pdr_profile_get returns, for up to n kernel classes,
 * names (comma-separated into names[namelen]), launch counts and total ms. */
int pdr_profile_enable(pdr_ctx* ctx, int32_t enable);
int pdr_profile_get(pdr_ctx* ctx, int32_t n, char* names, int32_t namelen,
                    int64_t* counts, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* PANDORA_AMD_H */
