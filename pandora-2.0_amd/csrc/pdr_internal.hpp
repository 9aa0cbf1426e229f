// pdr_internal.hpp -- shared declarations of the MI355X RK3-step library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pandora_amd.h"
#include "pdr_fft.hpp"

// Fortran-order 3x3 tensor access a(m, n), 1-based, as the reference's bmat/binv/amat
#define T9(a, m, n) ((a)[((m)-1) + 3 * ((n)-1)])

namespace pdr {

// Domain sizes (g_DomainCalc, src/pg_domain.f90:220-228).
struct Geo {
  int nx, ny, nz;
  int ki, kj, kk;     // KI_MAX, KJ_MAX, KK_MAX
  int nki, nkj, nkk;  // NODES_KI, NODES_KJ, NODES_KK
  int ldj;            // j stride of the 1r2w buffer (NKJ padded to 16)
  int ldi;            // i stride of the x-pencil buffer (NKI padded to 16)
  long long mw;       // modes per wave field      NKI*NKJ*NKK
  long long m1;       // elements per 1r2w field   NKI*NZ*ldj
  long long m2;       // elements per x-pencil field of the chunk buffer  zc*NY*ldi
  int zc;             // z planes per chunk of the x-pencil buffer (NZ when unchunked)
  long long nr;       // real points               NX*NY*NZ
};

// Per-stage operator coefficients, evaluated on the host in FP64 from the
// current frame (bmat, binv, amat and their existence masks), so that one
// kernel serves the isotropic and the BF_ROT shear forms:
//   vorticity   w_c = i * sum_n (sum_m G[c][n][m] k_m) u_n   (f_RK3VorticityBFRot,
//               src/pf_rk3.f90:3786-3843; the isotropic i k x u is the identity frame)
//   product     NL  = w x (V u), V = masked binv              (src/pf_rk3.f90:1597-1711)
//   FormF       nlf = Bm NL; visc = -nu (k.M k) u; p1 = M k; p4 = k.M k;
//               p3 = 2 sum_c (sum_m k_m P[m][c]) u_c; hom = 2 H u
//               (src/pf_rk3.f90:2864-3209, 2323-2487)
struct Coef {
  double G[27];  // [c][n][m]
  double V[9];   // [r][m] masked binv
  float Vf[9];   // V rounded to FP32 (the FP32-mixed x pass)
  double Bm[9];  // [i][c] masked bmat
  double M[9];   // [l][n] sum_j bmat(l,j) bmat(n,j)
  double P[9];   // [m][c]
  double H[9];   // [i][c]
  double nu;
  int use_nl;    // F_TYPE 1, 2
  int use_visc;  // F_TYPE 1, 2, 3
  int has_hom;   // shear
  int form;      // 0: isotropic / BF_ROT tensors; 1: NS_R_ROT (nonlinear term of the
                 // unmapped velocity, pressure along M k = bmat^T k, continuity
                 // fix through Bm and V)
  int fdef;      // F_TYPE 5: forcing read at (i, j, k) / fdiv where fdiv divides them,
  int fdiv[3];   // projected along the laboratory wavenumber fbm[m] k_m (1 and 1.0
  double fbm[3]; // for the isotropic form); finv = 1 / fdiv
  double finv[3];
};

// B110, B220, B330 (m = 1, 2, 3); values <= 0 stand for the reference default 1
inline int box_stretch(const pdr_params& p, int m) {
  const int b = m == 1 ? p.b110 : (m == 2 ? p.b220 : p.b330);
  return b > 0 ? b : 1;
}

// One member of a batched ensemble (pdr_ens.hip): everything a batched pass
// kernel needs to run that member's tiles, read from device memory by the
// workgroups whose block index falls in [start[p], start[p] + blocks[p]).
struct EnsMember {
  Geo g;
  Coef cf;
  int n;                  // line length (cubic member)
  int nt[5];              // tiles per line group of pass p (zinv, yinv, xpass, yfwd, zfwd+FormF)
  int fpb_y, fpb_yf;      // fields per workgroup of the y passes
  unsigned start[5];      // first block of the member in pass p
  double dt;              // RK3_DT
  double norm;            // 1 / (NX NY NZ)
  const void* u[3];
  void* rk[3];
  void* w1;
  void* w2;
  unsigned long long* umax;
  const void* twz;
  const void* twy;
  const void* twx;
};

// Tracer frame transforms (p_RK3Mov2Lab / p_RK3Lab2Mov, src/pp_rk3.f90:989-1043).
struct TracerFrame {
  double binv[9];  // masked, [m][n]
  double bmat[9];  // masked, [m][n]
  int bf;          // 0: isotropic (no transform)
  int order;       // P_INTERP_ORDER
  double dx, dy, dz, lx, ly, lz;
};

enum KernelClass {
  K_ZINV = 0,
  K_YINV,
  K_XPASS,
  K_YFWD,
  K_ZFWD,
  K_FORMF,
  K_ZFF,  // fused z-forward + FormF
  K_TRACER,
  K_REMESH,
  K_MISC,
  K_STATS,  // statistics of stage 1 inside the step (wave-space pass)
  K_NCLASS
};

struct Tables {
  void* twz = nullptr;
  void* twy = nullptr;
  void* twx = nullptr;
};

}  // namespace pdr

struct pdr_ctx {
  pdr_params p;
  pdr::Geo g;
  int prec = 0;     // 0 FP64, 1 FP32-mixed
  size_t csz = 16;  // bytes per complex element
  size_t rsz = 8;   // bytes per real element
  // resident state
  void* u[3] = {nullptr, nullptr, nullptr};
  void* rk[3] = {nullptr, nullptr, nullptr};
  // work buffers
  void* w1 = nullptr;  // 6 fields x m1 (1r2w layout)
  void* w2 = nullptr;  // 6 fields x m2 (x-pencil layout, zc planes of z)
  void *w1b = nullptr, *w1c = nullptr, *w2b = nullptr, *w2c = nullptr;  // NS_R_CONV: du/dy, du/dz sweeps
  void* nl = nullptr;  // NL [c][i][j][k] written by pass 5 (aliases w1 fields 3..5)
  int fpb_y = 6;       // fields per workgroup of the chunked inverse / forward y passes
  int fpb_yf = 3;
  void* ur = nullptr;  // real-space velocity for tracers, [z][y][x][3]
  double* xp = nullptr;
  double* uprk = nullptr;
  double* up = nullptr;   // particle_up (3, np): laboratory-frame fluid velocity of the last stage
  double* xp0 = nullptr;  // particle_xp0 (3, np) with P_KEEP_INITIAL_POSITION, else null
  int64_t np = 0;
  // tracer statistics (p_StatsPrimary / p_StatsRelativeVelocity)
  int pst_on = 0, keep_x0 = 0, ts_release = 0;
  double xrel_prev[3] = {0.0, 0.0, 0.0};
  double pst[30] = {};
  int pst_tstep = -1;
  double* pst_buf = nullptr;  // device partial sums
  // tracer order: slot s holds the tracer uploaded as number pid[s]; the slots
  // are re-sorted by (z, y) grid row every sort_every steps (0: never)
  uint32_t* pid = nullptr;
  int sort_every = 10;
  int since_sort = 0;
  // tracers advanced inside pdr_stage (default) or by the caller through the
  // p_ControlRK3 / p_RK3SolveParticle* entry points (the stage then only
  // writes the tracer field)
  bool tracers_in_stage = true;
  int ur_tstep = -1, ur_stage = 0;  // the step and stage whose x pass wrote the tracer field
  pdr::TracerFrame tf_ur{};           // the frame of that stage
  int adv_stage = 0, adv_tstep = -1;  // the last caller-driven tracer advance
  void* sort_buf = nullptr;   // keys, indices and the radix-sort workspace
  size_t sort_bytes = 0;
  double* spare = nullptr;    // 3 * np doubles for the reordering
  unsigned long long* umax = nullptr;  // CFL reduction (bits of a non-negative double)
  double* red = nullptr;               // spectrum reduction scratch
  void* stat_buf = nullptr;            // statistics partials (allocated on first use)
  size_t stat_bytes = 0;
  void* rs_buf = nullptr;              // real-space moment partials (allocated on first use)
  size_t rs_bytes = 0;
  // statistics of stage 1 inside the step (pdr_set_stage_stats): every sst_every
  // steps, into a record buffer allocated by the set call
  int sst_every = 0;
  void* sst_buf = nullptr;
  size_t sst_bytes = 0;
  int sst_tstep = -1;    // the record's statistics step (tstep - 1, as the reference passes it)
  double sst_b11 = 1.0;  // binv(1,1) of the record's frame
  double* force = nullptr;             // OU forcing state, [i][j+NK][k+NK][c] complex FP64
  int force_nk = 0;                    // NK = aint(FORCE_KF)
  bool forcing = false;                // F_TYPE 1 with FORCE_STOP > 0
  uint64_t force_calls = 0;            // f_ForceCalc calls so far (draw stream index)
  pdr::Tables tab;
  pdr::FftPlan pz, py, px;
  // frame (FP64 on the host)
  double bmat[9], binv[9], amat[9];
  int bval[9], aval[9];
  double dt = 99999.0;
  double ucfl = 0.0;
  // launch shapes
  bool spec = false;  // length-specialised kernels (pdr_kernels2.hip)
  bool fuse_zf = true;  // fused z-forward + FormF (specialised path; PDR_FUSE_ZFORMF=0 disables)
  int tj_z = 1, ti_y = 1;
  size_t lds_max = 65536;
  // workgroups of the statistics pass that the device holds at once (found on
  // first use: 0 not yet, -1 unknown)
  mutable int stats_resident = 0;
  hipStream_t stream = nullptr;
  // the tracer advance runs on its own stream beside the forward passes of the
  // stage: ev_ur orders it after the x pass that writes the tracer field,
  // ev_tr orders the next stage after it
  hipStream_t tstream = nullptr;
  hipEvent_t ev_ur = nullptr;
  hipEvent_t ev_tr = nullptr;
  // two z chunks in flight (nz >= 512 unchunked, or PDR_ZSTREAMS=2): odd chunks run their
  // y/x/y passes on zstream with their own x-pencil buffer w2z, beside the even ones
  bool zs2 = false;
  hipStream_t zstream = nullptr;
  void* w2z = nullptr;
  hipEvent_t ev_za = nullptr, ev_zb = nullptr;
  size_t dev_bytes = 0;
  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double prof_ms[pdr::K_NCLASS] = {0};
  long long prof_cnt[pdr::K_NCLASS] = {0};
};

namespace pdr {

// launchers (pdr_kernels.hip); return hipSuccess or the launch error
hipError_t launch_zinv(pdr_ctx* c, const Coef& cf);
hipError_t launch_yinv(pdr_ctx* c, int z0, int nzc);
hipError_t launch_xpass(pdr_ctx* c, const Coef& cf, bool tracers, bool cfl, int z0, int nzc,
                        double* rsp = nullptr);
hipError_t launch_yfwd(pdr_ctx* c, int z0, int nzc);
hipError_t launch_zfwd(pdr_ctx* c);
hipError_t launch_formf(pdr_ctx* c, const Coef& cf, double beta, double gdt);
hipError_t launch_zfwd_formf(pdr_ctx* c, const Coef& cf, double beta, double gdt);
hipError_t launch_force(pdr_ctx* c, int32_t tstep, double dt);
// the tracer advance (interpolation, RK3 update, particle_up); wrap: also the
// periodic correction of p_ControlFinishTimeStep (moving particle_xp0 with it)
hipError_t launch_tracers(pdr_ctx* c, const TracerFrame& tf, double beta, double gdt, hipStream_t s,
                          bool wrap = true);
hipError_t launch_tracers_wrap(pdr_ctx* c, const TracerFrame& tf, hipStream_t s);
hipError_t launch_remesh(pdr_ctx* c, int bfrot);
hipError_t launch_remesh_particles(pdr_ctx* c, double lx, double dx);
// batched ensemble passes (pdr_kernels2.hip)
bool ens_size_supported(int n);
int ens_tiles(int prec, int n, const Geo& g, int fpb_y, int fpb_yf, int nt[5], unsigned blocks[5]);
size_t ens_lds(int prec, int n, int pass);
hipError_t ens_prepare(int prec, size_t lds_max);
hipError_t launch_ens_pass(int prec, int pass, const EnsMember* mem, int nm, unsigned nblk, size_t lds,
                           double beta, double gamma, hipStream_t s);
// tracer order (pdr_sort.hip)
int particles_sort(pdr_ctx* c);
int particles_to_host(pdr_ctx* c, const double* dev, double* host);
int particles_from_host(pdr_ctx* c, const double* host, double* dev);
void particles_free_sort(pdr_ctx* c);
// tracer statistics (pdr_pstats.hip)
int pstats_alloc(pdr_ctx* c);
int pstats_reset(pdr_ctx* c);
hipError_t launch_pstats_primary(pdr_ctx* c, double dt);
hipError_t launch_pstats_rel(pdr_ctx* c, int tstep, hipStream_t s);
int pstats_read(pdr_ctx* c, int32_t* tstep, double* rec);
hipError_t launch_finite(pdr_ctx* c, int* flag);
// statistics of stage 1 inside the step (pdr_stats.hip)
bool stage_stats_due(const pdr_ctx* c, int stage, int tstep);
hipError_t stage_stats_wave(pdr_ctx* c, int tstep);
double* stage_stats_rs_partials(pdr_ctx* c);
int kkmax_of(const pdr_ctx* c);  // spectrum bins of the box (f_FluidstatsAlloc)
hipError_t stage_stats_rs_reduce(pdr_ctx* c);
void stage_stats_free(pdr_ctx* c);
hipError_t launch_copy(const void* a, void* b, long long bytes, hipStream_t s);
// HBM probes: kind 0 copy, 1 read, 2 write (fill with 1.0f), 3 triad
hipError_t launch_probe(int kind, const void* a, const void* b, void* d, long long bytes, float* sink,
                        hipStream_t s);

// profiling hooks (pdr_api.hip)
void prof_begin(pdr_ctx* c, int cls, hipStream_t s = nullptr);
void prof_end(pdr_ctx* c, int cls, hipStream_t s = nullptr);

void fft_plan(FftPlan* p, int n);

// length-specialised kernels (pdr_kernels2.hip)
bool spec_shape_ok(int nx, int ny, int nz);
int spec_forms(int prec, int nx, int ny, int nz, int out[5]);
bool spec_supported(int n);
size_t spec_lds_need(int prec, int nx, int ny, int nz);
hipError_t launch2_zinv(pdr_ctx* c, const Coef& cf);
hipError_t launch2_yinv(pdr_ctx* c, int z0, int nzc);
hipError_t launch2_xpass(pdr_ctx* c, const Coef& cf, bool tracers, bool cfl, int z0, int nzc, double* rsp);
int xpass_tiles(const pdr_ctx* c);  // x-pass workgroups per z plane (specialised path)
Coef make_coef(const pdr_ctx* c);
// Rogallo shear run, NS_R_ROT or NS_R_CONV (state = laboratory velocity on the
// sheared grid; the two share FormF, remesh, tracer and statistics forms)
inline bool rogallo_form(const pdr_ctx* c) {
  return c->p.f_type != PDR_F_ISOTROPIC && (c->p.ns_deformed == PDR_NS_R_ROT || c->p.ns_deformed == PDR_NS_R_CONV);
}
inline bool conv_form(const pdr_ctx* c) {
  return c->p.f_type != PDR_F_ISOTROPIC && c->p.ns_deformed == PDR_NS_R_CONV;
}
// LDS of the convective x pass: two buffers of 6 complex lines of NX + 1
inline size_t conv_xpass_lds(int nx, size_t csz) { return 12 * size_t(nx + 1) * csz; }
Coef conv_coef(const Coef& cf, int d);  // G selecting field 3 + c = d u_c / d x_d
hipError_t launch_conv_front(pdr_ctx* c, const Coef& cf, bool tr, bool cfl, double* rsp);
hipError_t launch_dealias(pdr_ctx* c);  // no-op for DEALIASING = 1
hipError_t launch2_yfwd(pdr_ctx* c, int z0, int nzc);
hipError_t launch2_zfwd(pdr_ctx* c);
hipError_t launch2_zfwd_formf(pdr_ctx* c, const Coef& cf, double beta, double gdt);

// records the text returned by pdr_last_error() and returns code
int set_error(int code, const std::string& msg);

}  // namespace pdr
