// pdr_api.hip -- C-ABI of the MI355X RK3 step: context, state transfer, stage
// orchestration, shear-frame clock and remesh decision (all FP64 on the host).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pdr_internal.hpp"

using namespace pdr;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int pdr::set_error(int code, const std::string& msg) { return fail(code, msg); }

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(PDR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)


// Reference RK3 coefficients (src/pg_constants.f90:56-62, src/pf_rk3.f90:3478-3488)
static void rk_coeffs(int stage, double* beta, double* gamma) {
  if (stage == 1) {
    *beta = 0.0;
    *gamma = 1.0 / 3.0;
  } else if (stage == 2) {
    *beta = -(5.0 / 9.0);
    *gamma = 15.0 / 16.0;
  } else {
    *beta = -(153.0 / 128.0);
    *gamma = 8.0 / 15.0;
  }
}

// float32 pi widened to double (LONGONEPI is assigned from a default-real
// literal, src/pg_constants.f90:126)
static const double kTwoPiRef = 2.0 * double(3.14159265358979323846f);

namespace pdr {

void fft_plan(FftPlan* p, int n) {
  p->n = n;
  p->npass = 0;
  int m = n;
  while (m % 8 == 0 && p->npass < kMaxPasses) { p->radix[p->npass++] = 8; m /= 8; }
  if (m % 4 == 0) { p->radix[p->npass++] = 4; m /= 4; }
  if (m % 2 == 0) { p->radix[p->npass++] = 2; m /= 2; }
  while (m % 3 == 0 && p->npass < kMaxPasses) { p->radix[p->npass++] = 3; m /= 3; }
  while (m % 5 == 0 && p->npass < kMaxPasses) { p->radix[p->npass++] = 5; m /= 5; }
}

void prof_begin(pdr_ctx* c, int cls, hipStream_t s) {
  if (!c->prof) return;
  hipEvent_t a, b;
  if (c->ev_pool.size() >= 2) {
    a = c->ev_pool.back();
    c->ev_pool.pop_back();
    b = c->ev_pool.back();
    c->ev_pool.pop_back();
  } else {
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) return;
  }
  (void)hipEventRecord(a, s ? s : c->stream);
  c->ev_pending.push_back({cls, {a, b}});
}

void prof_end(pdr_ctx* c, int cls, hipStream_t s) {
  if (!c->prof || c->ev_pending.empty()) return;
  auto& last = c->ev_pending.back();
  if (last.first != cls) return;
  (void)hipEventRecord(last.second.second, s ? s : c->stream);
}

}  // namespace pdr

static bool radix235(int n) {
  if (n < 2) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

// g_DomainCalc frame initialisation (src/pg_domain.f90:177-216)
static void frame_init(pdr_ctx* c) {
  for (int a = 0; a < 9; ++a) {
    c->bmat[a] = c->binv[a] = c->amat[a] = 0.0;
    c->bval[a] = c->aval[a] = 0;
  }
  T9(c->binv, 1, 1) = double(box_stretch(c->p, 1));
  T9(c->bval, 1, 1) = 1;
  T9(c->binv, 2, 2) = double(box_stretch(c->p, 2));
  T9(c->bval, 2, 2) = 1;
  T9(c->binv, 3, 3) = double(box_stretch(c->p, 3));
  T9(c->bval, 3, 3) = 1;
  // the shear entries exist in every form but the deformed isotropic one
  if (c->p.f_type != PDR_F_ISOTROPIC && c->p.f_type != PDR_F_ISOTROPIC_DEFORMED) {
    T9(c->bval, 1, 3) = 1;
    T9(c->amat, 1, 3) = c->p.f_deform_s;
    T9(c->aval, 1, 3) = 1;
  }
  T9(c->bmat, 1, 1) = 1.0 / T9(c->binv, 1, 1);
  T9(c->bmat, 2, 2) = 1.0 / T9(c->binv, 2, 2);
  T9(c->bmat, 3, 3) = 1.0 / T9(c->binv, 3, 3);
}

// The forcing of the deformed isotropic form, f_RK3BFRotForcing and the identical
// f_RK3RRotForcing (src/pf_rk3.f90:3256-3349, and its R_ROT twin): forced modes are
// those whose indices the stretches divide, read at the quotient, made solenoidal
// along the laboratory wavenumber bmat(m,m) k_m.
static void deformed_forcing(const pdr_ctx* c, Coef& f) {
  const bool def = c->p.f_type == PDR_F_ISOTROPIC_DEFORMED;
  f.fdef = def ? 1 : 0;
  for (int m = 1; m <= 3; ++m) {
    f.fdiv[m - 1] = def ? box_stretch(c->p, m) : 1;
    f.finv[m - 1] = 1.0 / double(f.fdiv[m - 1]);
    f.fbm[m - 1] = def ? T9(c->bmat, m, m) : 1.0;  // the isotropic form: unit stretch
  }
}

// Operator tensors from the current frame (see Coef).
static Coef make_coef_fp64(const pdr_ctx* c) {
  Coef f;
  std::memset(&f, 0, sizeof(f));
  const double* B = c->bmat;
  const double* BI = c->binv;
  const double* A = c->amat;
  const int* bv = c->bval;
  const int* av = c->aval;
  auto G = [&](int cc, int n, int m) -> double& { return f.G[((cc - 1) * 3 + (n - 1)) * 3 + (m - 1)]; };
  if (rogallo_form(c)) {
    // NS_R_ROT: omega = curl' u with d'_j = i sum_m bmat(m,j) k_m
    // (f_RK3DerivativeRRot, src/pf_rk3.f90:3749-3780; accumulation 416-421,
    // 533-537, 645-649); the nonlinear term is the isotropic omega x u of the
    // unmapped velocity (1088-1089), FormF is f_RK3FormFRRot (2680-2856)
    for (int m = 1; m <= 3; ++m) {
      G(1, 3, m) += T9(B, m, 2);
      G(1, 2, m) -= T9(B, m, 3);
      G(2, 1, m) += T9(B, m, 3);
      G(2, 3, m) -= T9(B, m, 1);
      G(3, 2, m) += T9(B, m, 1);
      G(3, 1, m) -= T9(B, m, 2);
    }
    for (int r = 1; r <= 3; ++r)
      for (int m = 1; m <= 3; ++m) {
        f.V[3 * (r - 1) + (m - 1)] = T9(bv, r, m) ? T9(BI, r, m) : 0.0;
        f.Bm[3 * (r - 1) + (m - 1)] = T9(bv, r, m) ? T9(B, r, m) : 0.0;
        f.M[3 * (r - 1) + (m - 1)] = T9(bv, m, r) ? T9(B, m, r) : 0.0;  // p1 = bmat^T k
        // amat / 2: P enters as 2 sum_m p1_m P[m][c] u_c, H as 2 H u
        f.P[3 * (r - 1) + (m - 1)] = T9(av, r, m) ? 0.5 * T9(A, r, m) : 0.0;
        f.H[3 * (r - 1) + (m - 1)] = T9(av, r, m) ? 0.5 * T9(A, r, m) : 0.0;
      }
    f.nu = c->p.f_nu;
    f.use_nl = (c->p.f_type == PDR_F_SHEAR_DNS || c->p.f_type == PDR_F_ISOTROPIC_DEFORMED) ? 1 : 0;
    f.use_visc = (c->p.f_type != PDR_F_SHEAR_RDT_INVISCID) ? 1 : 0;
    // the deformed isotropic branch of f_RK3FormFRRot adds no source term (2820-2828)
    f.has_hom = (c->p.f_type != PDR_F_ISOTROPIC_DEFORMED) ? 1 : 0;
    f.form = 1;
    deformed_forcing(c, f);
    return f;
  }
  for (int n = 1; n <= 3; ++n)
    for (int m = 1; m <= 3; ++m) {
      if (T9(bv, m, 1)) {
        if (T9(bv, 3, n)) G(2, n, m) -= T9(B, m, 1) * T9(BI, 3, n);
        if (T9(bv, 2, n)) G(3, n, m) += T9(B, m, 1) * T9(BI, 2, n);
      }
      if (T9(bv, m, 2)) {
        if (T9(bv, 3, n)) G(1, n, m) += T9(B, m, 2) * T9(BI, 3, n);
        if (T9(bv, 1, n)) G(3, n, m) -= T9(B, m, 2) * T9(BI, 1, n);
      }
      if (T9(bv, m, 3)) {
        if (T9(bv, 2, n)) G(1, n, m) -= T9(B, m, 3) * T9(BI, 2, n);
        if (T9(bv, 1, n)) G(2, n, m) += T9(B, m, 3) * T9(BI, 1, n);
      }
    }
  for (int r = 1; r <= 3; ++r)
    for (int m = 1; m <= 3; ++m) {
      f.V[3 * (r - 1) + (m - 1)] = T9(bv, r, m) ? T9(BI, r, m) : 0.0;
      f.Bm[3 * (r - 1) + (m - 1)] = T9(bv, r, m) ? T9(B, r, m) : 0.0;
    }
  for (int l = 1; l <= 3; ++l)
    for (int n = 1; n <= 3; ++n) {
      double s = 0.0;
      for (int j = 1; j <= 3; ++j)
        if (T9(bv, l, j) && T9(bv, n, j)) s += T9(B, l, j) * T9(B, n, j);
      f.M[3 * (l - 1) + (n - 1)] = s;
    }
  for (int m = 1; m <= 3; ++m)
    for (int cc = 1; cc <= 3; ++cc) {
      double s = 0.0, h = 0.0;
      for (int r = 1; r <= 3; ++r)
        for (int s2 = 1; s2 <= 3; ++s2) {
          if (T9(bv, m, r) && T9(av, r, s2) && T9(bv, s2, cc)) s += T9(B, m, r) * T9(A, r, s2) * T9(BI, s2, cc);
        }
      f.P[3 * (m - 1) + (cc - 1)] = s;
      for (int r = 1; r <= 3; ++r)
        for (int j = 1; j <= 3; ++j)
          if (T9(bv, m, r) && T9(av, r, j) && T9(bv, j, cc)) h += T9(B, m, r) * T9(A, r, j) * T9(BI, j, cc);
      f.H[3 * (m - 1) + (cc - 1)] = h;
    }
  f.nu = c->p.f_nu;
  f.use_nl = (c->p.f_type == PDR_F_ISOTROPIC || c->p.f_type == PDR_F_SHEAR_DNS ||
              c->p.f_type == PDR_F_ISOTROPIC_DEFORMED) ? 1 : 0;
  f.use_visc = (c->p.f_type != PDR_F_SHEAR_RDT_INVISCID) ? 1 : 0;
  // the deformed isotropic form adds no homogeneous source term (its amat is zero)
  f.has_hom = (c->p.f_type != PDR_F_ISOTROPIC && c->p.f_type != PDR_F_ISOTROPIC_DEFORMED) ? 1 : 0;
  deformed_forcing(c, f);
  return f;
}

Coef pdr::make_coef(const pdr_ctx* c) {
  Coef f = make_coef_fp64(c);
  for (int q = 0; q < 9; ++q) f.Vf[q] = float(f.V[q]);
  return f;
}

static TracerFrame make_tracer_frame(const pdr_ctx* c) {
  TracerFrame t;
  std::memset(&t, 0, sizeof(t));
  for (int m = 1; m <= 3; ++m)
    for (int n = 1; n <= 3; ++n) {
      t.binv[3 * (m - 1) + (n - 1)] = T9(c->bval, m, n) ? T9(c->binv, m, n) : 0.0;
      t.bmat[3 * (m - 1) + (n - 1)] = T9(c->bval, m, n) ? T9(c->bmat, m, n) : 0.0;
    }
  // BF_ROT tracers move in the moving frame (p_RK3SolveParticleBF); R_ROT ones
  // take the isotropic update (p_RK3SolveParticleR, src/pp_rk3.f90:307-403)
  t.bf = (c->p.f_type != PDR_F_ISOTROPIC && !rogallo_form(c)) ? 1 : 0;
  t.order = c->p.p_interp_order;
  t.lx = kTwoPiRef;
  t.ly = kTwoPiRef;
  t.lz = kTwoPiRef;
  t.dx = t.lx / double(c->g.nx);
  t.dy = t.ly / double(c->g.ny);
  t.dz = t.lz / double(c->g.nz);
  return t;
}

template <typename T>
static void make_twiddles(int n, std::vector<unsigned char>& out) {
  out.resize(sizeof(cx<T>) * size_t(n));
  cx<T>* w = reinterpret_cast<cx<T>*>(out.data());
  const long double pil = 3.14159265358979323846264338327950288L;
  for (int m = 0; m < n; ++m) {
    const long double a = -2.0L * pil * (long double)m / (long double)n;
    w[m].x = T(double(cosl(a)));
    w[m].y = T(double(sinl(a)));
  }
}

static int upload_tw(pdr_ctx* c, int n, void** dst) {
  std::vector<unsigned char> h;
  if (c->prec)
    make_twiddles<float>(n, h);
  else
    make_twiddles<double>(n, h);
  HIPCHK(hipMalloc(dst, h.size()));
  c->dev_bytes += h.size();
  HIPCHK(hipMemcpy(*dst, h.data(), h.size(), hipMemcpyHostToDevice));
  return PDR_OK;
}

static int dalloc(pdr_ctx* c, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // clear it: later launch checks read hipGetLastError()
    *p = nullptr;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    (void)hipGetLastError();
    return fail(PDR_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " +
                                   hipGetErrorString(e) + " (context holds " + std::to_string(c->dev_bytes) +
                                   " bytes; device free " + std::to_string(free_b) + " of " +
                                   std::to_string(total_b) + ", x-pencil chunk " + std::to_string(c->g.zc) +
                                   " planes)");
  }
  c->dev_bytes += bytes;
  return PDR_OK;
}

extern "C" {

const char* pdr_version(void) { return "pandora-amd 0.1 (gfx950, HIP)"; }

const char* pdr_last_error(void) { return g_err.c_str(); }

int pdr_destroy(pdr_ctx* c) {
  if (!c) return PDR_OK;
  // every stream of the context, before any buffer is freed
  for (hipStream_t s : {c->stream, c->zstream, c->tstream})
    if (s) (void)hipStreamSynchronize(s);
  for (int q = 0; q < 3; ++q) {
    if (c->u[q]) (void)hipFree(c->u[q]);
    if (c->rk[q]) (void)hipFree(c->rk[q]);
  }
  if (c->w1) (void)hipFree(c->w1);
  if (c->w2) (void)hipFree(c->w2);
  if (c->w2z) (void)hipFree(c->w2z);
  for (void* q : {c->w1b, c->w1c, c->w2b, c->w2c})
    if (q) (void)hipFree(q);
  if (c->ur) (void)hipFree(c->ur);
  if (c->xp) (void)hipFree(c->xp);
  if (c->uprk) (void)hipFree(c->uprk);
  if (c->up) (void)hipFree(c->up);
  if (c->xp0) (void)hipFree(c->xp0);
  if (c->pid) (void)hipFree(c->pid);
  particles_free_sort(c);
  if (c->pst_buf) (void)hipFree(c->pst_buf);
  if (c->umax) (void)hipFree(c->umax);
  if (c->red) (void)hipFree(c->red);
  if (c->force) (void)hipFree(c->force);
  if (c->stat_buf) (void)hipFree(c->stat_buf);
  if (c->rs_buf) (void)hipFree(c->rs_buf);
  stage_stats_free(c);
  if (c->tab.twz) (void)hipFree(c->tab.twz);
  if (c->tab.twy) (void)hipFree(c->tab.twy);
  if (c->tab.twx) (void)hipFree(c->tab.twx);
  for (auto& e : c->ev_pending) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  for (auto& e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev_ur) (void)hipEventDestroy(c->ev_ur);
  if (c->ev_tr) (void)hipEventDestroy(c->ev_tr);
  if (c->ev_za) (void)hipEventDestroy(c->ev_za);
  if (c->ev_zb) (void)hipEventDestroy(c->ev_zb);
  if (c->zstream) (void)hipStreamDestroy(c->zstream);
  if (c->tstream) (void)hipStreamDestroy(c->tstream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PDR_OK;
}

// bytes of the forcing state: 3 components x (NK+1)(2NK+1)^2 complex FP64
static size_t force_bytes(const pdr_ctx* c) {
  if (!c->forcing) return 0;
  const size_t w = size_t(2 * c->force_nk + 1);
  return 3 * size_t(c->force_nk + 1) * w * w * 16;
}

// x-pencil chunk (MB) used when the whole-box buffers do not fit
constexpr double kZChunkMB = 2048.0;

int pdr_create(pdr_ctx** out, const pdr_params* p) {
  if (!out || !p) return fail(PDR_ERR_ARG, "pdr_create: null argument");
  *out = nullptr;
  const int nx = p->nodes_x, ny = p->nodes_y, nz = p->nodes_z;
  if (!radix235(nx) || !radix235(ny) || !radix235(nz) || nx % 2 || nx < 8 || ny < 8 || nz < 8)
    return fail(PDR_ERR_UNSUPPORTED, "grid sizes must be even, >= 8 and of the form 2^a 3^b 5^c");
  if (p->f_type < 1 || p->f_type > 5)
    return fail(PDR_ERR_UNSUPPORTED, "F_TYPE must be 1 (isotropic), 2..4 (shear DNS/RDT) or 5 (deformed isotropic)");
  if (p->f_type != PDR_F_ISOTROPIC && (p->ns_deformed < PDR_NS_BF_ROT || p->ns_deformed > PDR_NS_R_ROT))
    return fail(PDR_ERR_UNSUPPORTED, "NS_DEFORMED must be 1 (BF_ROT), 2 (R_CONV) or 3 (R_ROT)");
  if (p->dealiasing < PDR_DEALIASING_TWOTHIRD || p->dealiasing > PDR_DEALIASING_LES)
    return fail(PDR_ERR_UNSUPPORTED, "DEALIASING must be 1 (2/3 rule), 2 (spherical) or 3 (LES)");
  if (p->dealiasing == PDR_DEALIASING_LES && !(p->les_cutoff > 0.0))
    return fail(PDR_ERR_ARG, "DEALIASING = 3 needs LES_CUTOFF > 0");
  if (p->rk3_timestep < 1 || p->rk3_timestep > 3) return fail(PDR_ERR_ARG, "RK3_TIMESTEP must be 1, 2 or 3");
  if (p->rk3_cfl >= 1.0) return fail(PDR_ERR_ARG, "STOP: Courant number is greater than 1.0 (g_RK3Check)");
  if (p->f_type != PDR_F_ISOTROPIC && p->rk3_timestep == PDR_RK3_TIMESTEP_DECIDE &&
      (p->f_remesh_t <= 0 || p->f_deform_s == 0.0))
    return fail(PDR_ERR_ARG, "shear with RK3_TIMESTEP=DECIDE needs F_REMESH_T > 0 and F_DEFORM_S != 0");
  if (p->p_interp_order < 1 || p->p_interp_order > 8)
    return fail(PDR_ERR_UNSUPPORTED, "P_INTERP_ORDER must be 1..8");
  if (p->precision != PDR_FP64 && p->precision != PDR_FP32_MIXED)
    return fail(PDR_ERR_ARG, "precision must be PDR_FP64 or PDR_FP32_MIXED");
  if ((p->b110 > 1 || p->b220 > 1 || p->b330 > 1) && p->f_type == PDR_F_ISOTROPIC)
    return fail(PDR_ERR_UNSUPPORTED, "box stretches B110, B220, B330 need a deformed frame (F_TYPE 2..5)");
  // f_ForceCalc runs in every stage; its arrays enter FormF in the isotropic and
  // the deformed isotropic forms (src/pf_rk3.f90:2063, 2421-2467, 3256-3349)
  const bool forcing = (p->f_type == PDR_F_ISOTROPIC || p->f_type == PDR_F_ISOTROPIC_DEFORMED) && p->force_stop > 0;
  if (forcing && !(p->force_tf > 0.0 && p->force_kf > 0.0 && p->force_kf <= 64.0 && p->force_sigma >= 0.0))
    return fail(PDR_ERR_ARG, "forcing needs FORCE_TF > 0, 0 < FORCE_KF <= 64 and FORCE_SIGMA >= 0");

  pdr_ctx* c = new pdr_ctx();
  c->p = *p;
  c->prec = p->precision;
  c->csz = c->prec ? 8 : 16;
  c->rsz = c->prec ? 4 : 8;
  Geo& g = c->g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.ki = nx / 3;
  g.kj = ny / 3;
  g.kk = nz / 3;
  g.nki = g.ki + 1;
  g.nkj = 2 * g.kj + 1;
  g.nkk = 2 * g.kk + 1;
  g.ldj = (g.nkj + 15) / 16 * 16;
  g.ldi = (g.nki + 15) / 16 * 16;
  g.mw = (long long)g.nki * g.nkj * g.nkk;
  g.m1 = (long long)g.nki * g.nz * g.ldj;
  g.zc = g.nz;
  g.m2 = (long long)g.nz * g.ny * g.ldi;
  g.nr = (long long)nx * ny * nz;
  frame_init(c);
  c->dt = 99999.0;
  // f_ForceAlloc: FORCE_NK = aint(FORCE_KF) (src/pf_force.f90:289)
  c->forcing = forcing;
  c->force_nk = forcing ? int(std::trunc(p->force_kf)) : 0;

  int dev = 0;
  hipError_t he = hipGetDevice(&dev);
  if (he != hipSuccess) {
    delete c;
    return fail(PDR_ERR_HIP, std::string("no HIP device: ") + hipGetErrorString(he));
  }
  int lds = 65536;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
  c->lds_max = size_t(lds);
  // keep two workgroups per CU (160 KB of LDS per CU on gfx950)
  const size_t target = std::min<size_t>(c->lds_max, 81920);
  auto pick = [&](int len, int nlines) {
    int t = 16;
    while (t > 1 && (2 * size_t(t) * (len + 1) * c->csz > target || t > nlines)) t /= 2;
    return t;
  };
  c->tj_z = pick(nz, g.nkj);
  c->ti_y = pick(ny, g.nki);
  if (6 * size_t(nx + 1) * c->csz > c->lds_max || 2 * size_t(nz + 1) * c->csz > c->lds_max ||
      2 * size_t(ny + 1) * c->csz > c->lds_max || (conv_form(c) && conv_xpass_lds(nx, c->csz) > c->lds_max)) {
    delete c;
    return fail(PDR_ERR_UNSUPPORTED, "grid too large for the LDS of this device");
  }
  {
    const char* env = std::getenv("PDR_GENERIC_KERNELS");
    const bool generic = env && env[0] && env[0] != '0';
    // NS_R_CONV runs on the generic kernels (its x pass takes twelve rows)
    c->spec = !generic && !conv_form(c) && spec_supported(nx) && spec_supported(ny) && spec_supported(nz) &&
              spec_shape_ok(nx, ny, nz) &&
              spec_lds_need(c->prec, nx, ny, nz) <= c->lds_max;
  }
  // The x-pencil buffer may hold a chunk of zc planes instead of the whole box
  // (specialised kernels only): the y/x passes then run chunk by chunk.  It
  // costs time (measured at 512^3 FP32: 73 steps/s whole, 66 with 192 MB
  // chunks, 55 with 64 MB), so it is used only when the whole-box buffers
  // would not fit in device memory, or when PDR_ZCHUNK_MB asks for it.
  if (c->spec) {
    const char* env = std::getenv("PDR_ZCHUNK_MB");
    double mb = env && env[0] ? std::atof(env) : 0.0;
    const double plane = 6.0 * double(ny) * double(g.ldi) * double(c->csz);
    if (!(env && env[0])) {
      // free device memory; if the query fails, the device's total memory
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        int dev = 0;
        hipDeviceProp_t pr;
        free_b = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess)
                     ? size_t(pr.totalGlobalMem)
                     : 0;
        (void)hipGetLastError();
      }
      const double need = double(g.mw) * 6.0 * double(c->csz) + 6.0 * double(g.m1) * double(c->csz) +
                          plane * double(nz) + 3.0 * double(g.nr) * double(c->rsz);
      if (free_b == 0 || need > 0.9 * double(free_b)) mb = kZChunkMB;
    }
    int zc = mb > 0.0 ? int(mb * 1048576.0 / plane) : nz;
    zc = std::max(1, std::min(zc, nz));
    // Two chunks in flight: the even chunks' y/x/y passes on the main stream,
    // the odd ones' on a second stream with their own x-pencil buffer, so that
    // the HBM-bound y passes of one run beside the issue-bound x pass of the
    // other.  Medians of same-process runs at 512^3 FP32 with 1M tracers show
    // no gain (83.34 against 83.47 steps/s, spread 0.4 %), so it is on by
    // default only where it was measured faster: FP64 from 1024 planes up
    // (+3 %) when the whole-box buffer fits; not over chunks cut for memory
    // (1536^3 FP64: 0.750 against 0.771 steps/s).  PDR_ZSTREAMS=2 forces it,
    // 1 turns it off; the same buffer bytes either way.
    const char* zs = std::getenv("PDR_ZSTREAMS");
    c->zs2 = (zs && zs[0] ? zs[0] == '2' : (c->prec == 0 && nz >= 1024 && zc == nz)) && nz > 1;
    if (c->zs2) zc = std::min(zc, (nz + 1) / 2);
    g.zc = zc;
    g.m2 = (long long)zc * g.ny * g.ldi;
    // split the fields of the y passes over workgroups while a chunk is small
    const int tiles = (g.nki + 15) / 16;
    c->fpb_y = 6;
    while (c->fpb_y > 1 && tiles * zc * (6 / c->fpb_y) < 1024) c->fpb_y = (c->fpb_y == 6) ? 3 : (c->fpb_y == 3 ? 2 : 1);
    c->fpb_yf = 3;
    while (c->fpb_yf > 1 && tiles * zc * (3 / c->fpb_yf) < 1024) c->fpb_yf = 1;
  }
  {
    // fused z-forward + FormF holds 3 NL values per mode in registers; above
    // 512 points per line the tile has too many modes per thread for that
    const char* env = std::getenv("PDR_FUSE_ZFORMF");
    c->fuse_zf = !(env && env[0] == '0') && nz <= 512;
  }
  fft_plan(&c->pz, nz);
  fft_plan(&c->py, ny);
  fft_plan(&c->px, nx);

  int rc = PDR_OK;
  he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&c->tstream, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_ur, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_tr, hipEventDisableTiming);
  if (he == hipSuccess && c->zs2) he = hipStreamCreateWithFlags(&c->zstream, hipStreamNonBlocking);
  if (he == hipSuccess && c->zs2) he = hipEventCreateWithFlags(&c->ev_za, hipEventDisableTiming);
  if (he == hipSuccess && c->zs2) he = hipEventCreateWithFlags(&c->ev_zb, hipEventDisableTiming);
  if (he != hipSuccess) {
    pdr_destroy(c);
    return fail(PDR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  const size_t wbytes = size_t(g.mw) * c->csz;
  for (int q = 0; q < 3 && rc == PDR_OK; ++q) {
    rc = dalloc(c, &c->u[q], wbytes);
    if (rc == PDR_OK) rc = dalloc(c, &c->rk[q], wbytes);
  }
  if (rc == PDR_OK) rc = dalloc(c, &c->w1, 6 * size_t(g.m1) * c->csz);
  if (rc == PDR_OK) rc = dalloc(c, &c->w2, 6 * size_t(g.m2) * c->csz);
  if (rc == PDR_OK && c->zs2) rc = dalloc(c, &c->w2z, 6 * size_t(g.m2) * c->csz);
  if (rc == PDR_OK) c->nl = static_cast<char*>(c->w1) + 3 * size_t(g.m1) * c->csz;
  if (rc == PDR_OK && conv_form(c)) {
    for (void** q : {&c->w1b, &c->w1c})
      if (rc == PDR_OK) rc = dalloc(c, q, 6 * size_t(g.m1) * c->csz);
    for (void** q : {&c->w2b, &c->w2c})
      if (rc == PDR_OK) rc = dalloc(c, q, 6 * size_t(g.m2) * c->csz);
  }
  // 256 bytes owned by the context: the CFL maximum at offset 0, the
  // non-finite flag of pdr_check_finite at offset 64
  if (rc == PDR_OK) rc = dalloc(c, (void**)&c->umax, 256);
  const size_t fbytes = force_bytes(c);
  if (rc == PDR_OK && fbytes) {
    rc = dalloc(c, (void**)&c->force, fbytes);
    if (rc == PDR_OK && hipMemsetAsync(c->force, 0, fbytes, c->stream) != hipSuccess)
      rc = fail(PDR_ERR_HIP, "zeroing forcing failed");
  }
  if (rc == PDR_OK) rc = upload_tw(c, nz, &c->tab.twz);
  if (rc == PDR_OK) rc = upload_tw(c, ny, &c->tab.twy);
  if (rc == PDR_OK) rc = upload_tw(c, nx, &c->tab.twx);
  if (rc == PDR_OK) {
    for (int q = 0; q < 3; ++q) {
      if (hipMemsetAsync(c->u[q], 0, wbytes, c->stream) != hipSuccess ||
          hipMemsetAsync(c->rk[q], 0, wbytes, c->stream) != hipSuccess)
        rc = fail(PDR_ERR_HIP, "zeroing state failed");
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(PDR_ERR_HIP, "stream sync failed");
  }
  if (rc != PDR_OK) {
    std::string keep = g_err;
    pdr_destroy(c);
    g_err = keep;
    return rc;
  }
  *out = c;
  return PDR_OK;
}

int pdr_sizes(const pdr_ctx* c, int64_t out[8]) {
  if (!c || !out) return fail(PDR_ERR_ARG, "pdr_sizes: null argument");
  const Geo& g = c->g;
  out[0] = g.nki;
  out[1] = g.nkj;
  out[2] = g.nkk;
  out[3] = g.mw;
  out[4] = g.m1;
  out[5] = kkmax_of(c);
  out[6] = (int64_t)c->dev_bytes;
  out[7] = c->np;
  return PDR_OK;
}

int pdr_kernel_forms(const pdr_ctx* c, int32_t out[8]) {
  if (!c || !out) return fail(PDR_ERR_ARG, "pdr_kernel_forms: null argument");
  int f[5] = {0, 0, 0, 0, 0};
  if (c->spec) (void)spec_forms(c->prec, c->g.nx, c->g.ny, c->g.nz, f);
  out[0] = c->spec ? 1 : 0;
  out[1] = c->g.zc;
  out[2] = c->zs2 ? 1 : 0;
  out[3] = (c->spec && c->fuse_zf) ? 1 : 0;
  for (int q = 0; q < 4; ++q) out[4 + q] = f[q];
  return PDR_OK;
}

static int put_field(pdr_ctx* c, void* dst, const double* src, long long n) {
  if (!src) return fail(PDR_ERR_ARG, "null host array");
  if (c->prec == PDR_FP64) {
    HIPCHK(hipMemcpy(dst, src, size_t(n) * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(static_cast<size_t>(n));
    for (long long q = 0; q < n; ++q) tmp[size_t(q)] = float(src[q]);
    HIPCHK(hipMemcpy(dst, tmp.data(), size_t(n) * sizeof(float), hipMemcpyHostToDevice));
  }
  return PDR_OK;
}

static int get_field(pdr_ctx* c, double* dst, const void* src, long long n) {
  if (!dst) return fail(PDR_ERR_ARG, "null host array");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->prec == PDR_FP64) {
    HIPCHK(hipMemcpy(dst, src, size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(static_cast<size_t>(n));
    HIPCHK(hipMemcpy(tmp.data(), src, size_t(n) * sizeof(float), hipMemcpyDeviceToHost));
    for (long long q = 0; q < n; ++q) dst[q] = double(tmp[size_t(q)]);
  }
  return PDR_OK;
}

int pdr_upload_state(pdr_ctx* c, const double* a, const double* b, const double* d) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  const long long n = 2 * c->g.mw;
  int rc = put_field(c, c->u[0], a, n);
  if (rc == PDR_OK) rc = put_field(c, c->u[1], b, n);
  if (rc == PDR_OK) rc = put_field(c, c->u[2], d, n);
  return rc;
}

int pdr_download_state(pdr_ctx* c, double* a, double* b, double* d) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  const long long n = 2 * c->g.mw;
  int rc = get_field(c, a, c->u[0], n);
  if (rc == PDR_OK) rc = get_field(c, b, c->u[1], n);
  if (rc == PDR_OK) rc = get_field(c, d, c->u[2], n);
  return rc;
}

int pdr_upload_rk3(pdr_ctx* c, const double* a, const double* b, const double* d) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  const long long n = 2 * c->g.mw;
  int rc = put_field(c, c->rk[0], a, n);
  if (rc == PDR_OK) rc = put_field(c, c->rk[1], b, n);
  if (rc == PDR_OK) rc = put_field(c, c->rk[2], d, n);
  return rc;
}

int pdr_download_rk3(pdr_ctx* c, double* a, double* b, double* d) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  const long long n = 2 * c->g.mw;
  int rc = get_field(c, a, c->rk[0], n);
  if (rc == PDR_OK) rc = get_field(c, b, c->rk[1], n);
  if (rc == PDR_OK) rc = get_field(c, d, c->rk[2], n);
  return rc;
}

int pdr_set_frame(pdr_ctx* c, const double bmat[9], const double binv[9]) {
  if (!c || !bmat || !binv) return fail(PDR_ERR_ARG, "pdr_set_frame: null argument");
  std::memcpy(c->bmat, bmat, sizeof(c->bmat));
  std::memcpy(c->binv, binv, sizeof(c->binv));
  return PDR_OK;
}

int pdr_get_frame(const pdr_ctx* c, double bmat[9], double binv[9]) {
  if (!c || !bmat || !binv) return fail(PDR_ERR_ARG, "pdr_get_frame: null argument");
  std::memcpy(bmat, c->bmat, sizeof(c->bmat));
  std::memcpy(binv, c->binv, sizeof(c->binv));
  return PDR_OK;
}

int pdr_set_dt(pdr_ctx* c, double dt) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  c->dt = dt;
  return PDR_OK;
}

int pdr_get_dt(const pdr_ctx* c, double* dt) {
  if (!c || !dt) return fail(PDR_ERR_ARG, "null argument");
  *dt = c->dt;
  return PDR_OK;
}

int pdr_synchronize(pdr_ctx* c) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return PDR_OK;
}

// f_RK3Remesh (src/pf_rk3.f90:3958-4049) + f_RK3RemeshWaveSpace frame updates (4056-4178)
static int maybe_remesh(pdr_ctx* c, int tstep, int32_t* done) {
  if (done) *done = 0;
  if (c->p.f_type == PDR_F_ISOTROPIC || c->p.rk3_timestep != PDR_RK3_TIMESTEP_DECIDE) return PDR_OK;
  const int b110 = box_stretch(c->p, 1);
  const int a = tstep - (b110 * c->p.f_remesh_t) - 1;
  const int b = 2 * b110 * c->p.f_remesh_t;
  if (a % b != 0) return PDR_OK;  // Fortran mod truncates toward zero like C %
  const TracerFrame tf = make_tracer_frame(c);
  HIPCHK(launch_remesh_particles(c, tf.lx, tf.dx));
  T9(c->binv, 1, 3) = (double(b110) * 0.5) / T9(c->bmat, 3, 3);
  T9(c->bmat, 1, 3) = -T9(c->bmat, 1, 1) * (-(double(b110) * 0.5));
  HIPCHK(launch_remesh(c, c->p.ns_deformed == PDR_NS_BF_ROT ? 1 : 0));
  T9(c->binv, 1, 3) = (-(double(b110) * 0.5)) / T9(c->bmat, 3, 3);
  if (done) *done = 1;
  return PDR_OK;
}

// p_ControlInitPartVelocity[BF|R] at the release (src/pp_control.f90:378-510):
// particle_up_rk3 = 0, particle_xp0 = particle_xp (particle_up is set by the advance)
static hipError_t tracers_release(pdr_ctx* c, hipStream_t s) {
  hipError_t e = hipMemsetAsync(c->uprk, 0, 24 * size_t(c->np), s);
  if (e == hipSuccess && c->xp0) e = hipMemcpyAsync(c->xp0, c->xp, 24 * size_t(c->np), hipMemcpyDeviceToDevice, s);
  return e;
}

int pdr_stage(pdr_ctx* c, int32_t stage, int32_t tstep, double time, double* dt) {
  (void)time;
  if (!c || !dt) return fail(PDR_ERR_ARG, "pdr_stage: null argument");
  if (stage < 1 || stage > 3) return fail(PDR_ERR_ARG, "stage must be 1, 2 or 3");
  if (stage == 1) {
    int rc = maybe_remesh(c, tstep, nullptr);
    if (rc != PDR_OK) return rc;
  }
  // f_FluidstatsFs((tstep-1)) after the remesh check (src/pf_rk3.f90:58-60); the
  // real-space moments then ride on this stage's x pass (1073-1078)
  const bool sst = stage_stats_due(c, stage, tstep);
  double* rsp = sst ? stage_stats_rs_partials(c) : nullptr;
  if (sst) {
    prof_begin(c, K_STATS);
    HIPCHK(stage_stats_wave(c, tstep));
    prof_end(c, K_STATS);
  }
  // f_RK3Stage1: p_StatsPrimary after the remesh, p_StatsRelativeVelocity after
  // the particle advance (src/pf_rk3.f90:62-99), for tstep > P_TS_RELEASE
  if (stage == 1 && c->np > 1 && c->sort_every > 0 && ++c->since_sort >= c->sort_every) {
    int rc = particles_sort(c);
    if (rc != PDR_OK) return rc;
  }
  const bool pstats = stage == 1 && c->pst_on && c->np > 0 && tstep > c->ts_release;
  if (pstats) HIPCHK(launch_pstats_primary(c, *dt));
  // spherical / LES truncation on entry to f_RK3PrepareRealSpace (src/pf_rk3.f90:269-277)
  HIPCHK(launch_dealias(c));
  const Coef cf = make_coef(c);
  // particles move from tstep = P_TS_RELEASE on (src/pf_rk3.f90:1146): the x pass
  // writes their field, and the stage advances them unless the caller does
  const bool field = c->np > 0 && tstep >= c->ts_release;
  const bool tracers = field && c->tracers_in_stage;
  const bool need_cfl =
      stage == 1 && (c->p.rk3_timestep == PDR_RK3_TIMESTEP_CFL ||
                     (c->p.rk3_timestep == PDR_RK3_TIMESTEP_DECIDE &&
                      (c->p.f_type == PDR_F_ISOTROPIC || c->p.f_type == PDR_F_ISOTROPIC_DEFORMED)));
  if (conv_form(c)) {
    if (need_cfl) HIPCHK(hipMemsetAsync(c->umax, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(launch_conv_front(c, cf, field, need_cfl, rsp));
    HIPCHK(launch_yfwd(c, 0, c->g.nz));
  } else {
    HIPCHK(launch_zinv(c, cf));
    if (need_cfl) HIPCHK(hipMemsetAsync(c->umax, 0, sizeof(unsigned long long), c->stream));
    // chunks touch disjoint planes of w1 and the tracer field; with two in
    // flight the odd ones run on zstream with their own x-pencil buffer
    const bool two = c->zs2 && c->g.zc < c->g.nz;
    if (two) {
      HIPCHK(hipEventRecord(c->ev_za, c->stream));
      HIPCHK(hipStreamWaitEvent(c->zstream, c->ev_za, 0));
    }
    int chunk = 0;
    hipError_t he = hipSuccess;
    for (int z0 = 0; z0 < c->g.nz && he == hipSuccess; z0 += c->g.zc, ++chunk) {
      const int nzc = std::min(c->g.zc, c->g.nz - z0);
      const bool alt = two && (chunk & 1);
      if (alt) {
        std::swap(c->stream, c->zstream);
        std::swap(c->w2, c->w2z);
      }
      he = launch_yinv(c, z0, nzc);
      if (he == hipSuccess) he = launch_xpass(c, cf, field, need_cfl, z0, nzc, rsp);
      if (he == hipSuccess) he = launch_yfwd(c, z0, nzc);
      if (alt) {
        std::swap(c->stream, c->zstream);
        std::swap(c->w2, c->w2z);
      }
    }
    // join the second stream also when a launch failed, so that nothing freed
    // later is still in use there
    if (two) {
      hipError_t je = hipEventRecord(c->ev_zb, c->zstream);
      if (je == hipSuccess) je = hipStreamWaitEvent(c->stream, c->ev_zb, 0);
      if (he == hipSuccess) he = je;
    }
    HIPCHK(he);
  }
  if (sst) HIPCHK(stage_stats_rs_reduce(c));  // f_FluidstatsRsCollect's sums (src/pf_rk3.f90:1124-1127)
  if (field) {
    c->ur_tstep = tstep;
    c->ur_stage = stage;
    c->tf_ur = make_tracer_frame(c);
  }
  if (stage == 1) {
    // f_RK3Tstep (src/pf_rk3.f90:3852-3948)
    if (need_cfl) {
      unsigned long long bits = 0;
      HIPCHK(hipMemcpyAsync(&bits, c->umax, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      double u;
      std::memcpy(&u, &bits, sizeof(u));
      c->ucfl = u;
    }
    const double dx = kTwoPiRef / double(c->g.nx);
    switch (c->p.rk3_timestep) {
      case PDR_RK3_TIMESTEP_FIXED: *dt = c->p.rk3_dt; break;
      case PDR_RK3_TIMESTEP_CFL: *dt = (c->ucfl == 0.0) ? double(0.02f) : c->p.rk3_cfl * dx / c->ucfl; break;
      default:
        if (c->p.f_type == PDR_F_ISOTROPIC || c->p.f_type == PDR_F_ISOTROPIC_DEFORMED) {
          // the deformed form takes the smallest moving-frame spacing (3905-3919)
          const double h = c->p.f_type == PDR_F_ISOTROPIC
                               ? dx
                               : std::min(dx, std::min(kTwoPiRef / double(c->g.ny), kTwoPiRef / double(c->g.nz)));
          if (c->ucfl == 0.0) {
            *dt = double(0.02f);
          } else {
            *dt = c->p.rk3_cfl * h / c->ucfl;
            if (*dt > c->p.force_tf) *dt = c->p.force_tf;
          }
        } else {
          *dt = ((0.5 / c->p.f_deform_s) / double(c->p.f_remesh_t));
        }
    }
  }
  double beta, gamma;
  rk_coeffs(stage, &beta, &gamma);
  const double gdt = gamma * (*dt);
  if (tracers) {
    // the advance reads the tracer field the x pass wrote and nothing the forward
    // passes touch: it runs on the tracer stream beside them
    HIPCHK(hipEventRecord(c->ev_ur, c->stream));
    HIPCHK(hipStreamWaitEvent(c->tstream, c->ev_ur, 0));
    if (stage == 1 && tstep == c->ts_release) HIPCHK(tracers_release(c, c->tstream));
    HIPCHK(launch_tracers(c, c->tf_ur, beta, gdt, c->tstream));
    if (pstats) HIPCHK(launch_pstats_rel(c, tstep, c->tstream));
    HIPCHK(hipEventRecord(c->ev_tr, c->tstream));
  } else if (pstats && !field) {
    HIPCHK(launch_pstats_rel(c, tstep, c->stream));
  }
  // with the advance left to the caller, p_StatsRelativeVelocity follows its
  // pdr_particles_finish_step
  if (c->forcing) HIPCHK(launch_force(c, tstep, *dt));  // f_ForceCalc (src/pf_rk3.f90:2063)
  if (c->spec && c->fuse_zf) {
    HIPCHK(launch_zfwd_formf(c, cf, beta, gdt));
  } else {
    HIPCHK(launch_zfwd(c));
    HIPCHK(launch_formf(c, cf, beta, gdt));
  }
  // and at the end of f_RK3SolveNavierStokes (src/pf_rk3.f90:2101-2109)
  HIPCHK(launch_dealias(c));
  if (tracers) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_tr, 0));
  return PDR_OK;
}

int pdr_set_tracers_in_stage(pdr_ctx* c, int32_t on) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->tracers_in_stage = on != 0;
  c->adv_stage = 0;
  return PDR_OK;
}

// 0: p_RK3SolveParticle (isotropic), 1: p_RK3SolveParticleBF, 2: p_RK3SolveParticleR
// (the branch p_ControlRK3 takes, src/pp_control.f90:556-574)
static int tracer_form(const pdr_ctx* c) {
  if (c->p.f_type == PDR_F_ISOTROPIC) return 0;
  return rogallo_form(c) ? 2 : 1;
}

static int caller_advance_check(pdr_ctx* c, int32_t stage) {
  if (stage < 1 || stage > 3) return fail(PDR_ERR_ARG, "stage must be 1, 2 or 3");
  if (c->np <= 0) return fail(PDR_ERR_STATE, "no particles set");
  if (c->tracers_in_stage)
    return fail(PDR_ERR_STATE, "pdr_stage advances the tracers; call pdr_set_tracers_in_stage(ctx, 0) first");
  if (c->ur_stage != stage)
    return fail(PDR_ERR_STATE, "the tracer field is not that of this stage: call pdr_stage for it first");
  return PDR_OK;
}

int pdr_particles_control_rk3(pdr_ctx* c, int32_t stage, int32_t tstep, double deltat) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  int rc = caller_advance_check(c, stage);
  if (rc != PDR_OK) return rc;
  if (c->ur_tstep != tstep) return fail(PDR_ERR_STATE, "the tracer field is not that of this tstep");
  // p_ControlInitPartVelocity[BF|R] at the release step (src/pp_control.f90:540-551)
  if (stage == 1 && tstep == c->ts_release) HIPCHK(tracers_release(c, c->stream));
  double beta, gamma;
  rk_coeffs(stage, &beta, &gamma);
  HIPCHK(launch_tracers(c, c->tf_ur, beta, gamma * deltat, c->stream, false));
  c->adv_stage = stage;
  c->adv_tstep = tstep;
  return PDR_OK;
}

int pdr_particles_solve(pdr_ctx* c, int32_t form, int64_t npin, int32_t stage, double deltat) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  if (form < 0 || form > 2) return fail(PDR_ERR_ARG, "form must be 0 (isotropic), 1 (BF) or 2 (R)");
  if (form != tracer_form(c))
    return fail(PDR_ERR_ARG, "this run's frame takes p_RK3SolveParticle" +
                                 std::string(tracer_form(c) == 0 ? "" : (tracer_form(c) == 1 ? "BF" : "R")));
  int rc = caller_advance_check(c, stage);
  if (rc != PDR_OK) return rc;
  if (npin != c->np) return fail(PDR_ERR_ARG, "npin must equal the number of tracers set (one box, one owner)");
  double beta, gamma;
  rk_coeffs(stage, &beta, &gamma);
  HIPCHK(launch_tracers(c, c->tf_ur, beta, gamma * deltat, c->stream, false));
  c->adv_stage = stage;
  c->adv_tstep = c->ur_tstep;
  return PDR_OK;
}

int pdr_particles_finish_step(pdr_ctx* c, int32_t tstep) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  if (c->np <= 0) return fail(PDR_ERR_STATE, "no particles set");
  HIPCHK(launch_tracers_wrap(c, make_tracer_frame(c), c->stream));
  // f_RK3Stage1: p_StatsRelativeVelocity after the stage-1 advance (src/pf_rk3.f90:88-97)
  if (c->adv_stage == 1 && c->pst_on && tstep > c->ts_release) HIPCHK(launch_pstats_rel(c, tstep, c->stream));
  c->adv_stage = 0;
  return PDR_OK;
}

int pdr_dealias(pdr_ctx* c) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  HIPCHK(launch_dealias(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PDR_OK;
}

int pdr_remesh(pdr_ctx* c, int32_t tstep, int32_t* done) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  return maybe_remesh(c, tstep, done);
}

// g_RK3UpdateB (src/pg_rk3.f90:280-385)
int pdr_update_b(pdr_ctx* c, int32_t tstep, double simtime) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  if (c->p.f_type == PDR_F_ISOTROPIC) return PDR_OK;
  const int b110 = box_stretch(c->p, 1);
  double localtime;
  if (c->p.rk3_timestep == PDR_RK3_TIMESTEP_DECIDE) {
    if (tstep <= (b110 * c->p.f_remesh_t)) {
      localtime = simtime;
    } else {
      const double a = double(tstep - (b110 * c->p.f_remesh_t) - 1);
      const double b = double(2 * (b110 * c->p.f_remesh_t));
      const int localtstep = int(a / b);
      localtime = simtime - (double(localtstep + 1) / c->p.f_deform_s * double(b110));
    }
  } else {
    localtime = simtime;
  }
  T9(c->bmat, 1, 3) = -T9(c->bmat, 1, 1) * c->p.f_deform_s * localtime;
  T9(c->binv, 1, 3) = c->p.f_deform_s * localtime / T9(c->bmat, 3, 3);
  return PDR_OK;
}

// g_RK3Calc (src/pg_rk3.f90:82-129)
int pdr_step(pdr_ctx* c, int32_t tstep, double* simtime) {
  if (!c || !simtime) return fail(PDR_ERR_ARG, "pdr_step: null argument");
  const double tstart = *simtime;
  int rc = pdr_stage(c, 1, tstep, *simtime, &c->dt);
  if (rc != PDR_OK) return rc;
  *simtime = tstart + (1.0 / 3.0) * c->dt;
  pdr_update_b(c, tstep, *simtime);
  rc = pdr_stage(c, 2, tstep, *simtime, &c->dt);
  if (rc != PDR_OK) return rc;
  *simtime = tstart + (3.0 / 4.0) * c->dt;
  pdr_update_b(c, tstep, *simtime);
  rc = pdr_stage(c, 3, tstep, *simtime, &c->dt);
  if (rc != PDR_OK) return rc;
  *simtime = tstart + c->dt;
  pdr_update_b(c, tstep, *simtime);
  return PDR_OK;
}

int pdr_run(pdr_ctx* c, int32_t tstep0, int32_t nsteps, double* simtime) {
  if (!c || !simtime || nsteps < 0) return fail(PDR_ERR_ARG, "pdr_run: bad argument");
  for (int32_t s = 0; s < nsteps; ++s) {
    int rc = pdr_step(c, tstep0 + s, simtime);
    if (rc != PDR_OK) return rc;
  }
  return PDR_OK;
}

int pdr_check_finite(pdr_ctx* c) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  // the flag lives in the context's own small buffer: no allocation per call
  int* flag = reinterpret_cast<int*>(reinterpret_cast<char*>(c->umax) + 64);
  int h = 0;
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), c->stream);
  if (e == hipSuccess) e = launch_finite(c, flag);
  if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(PDR_ERR_HIP, std::string("finite check: ") + hipGetErrorString(e));
  if (h) return fail(PDR_ERR_NONFINITE, "velocity state contains NaN or Inf");
  return PDR_OK;
}

int pdr_particles_upload(pdr_ctx* c, const double* xp, int64_t np) {
  if (!c || np < 0 || (np > 0 && !xp)) return fail(PDR_ERR_ARG, "pdr_particles_upload: bad argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  particles_free_sort(c);
  double** arrays[4] = {&c->xp, &c->uprk, &c->up, &c->xp0};
  for (double** a : arrays)
    if (*a) {
      (void)hipFree(*a);
      c->dev_bytes -= 24 * size_t(c->np);
      *a = nullptr;
    }
  if (c->pid) {
    (void)hipFree(c->pid);
    c->dev_bytes -= 4 * size_t(c->np);
    c->pid = nullptr;
  }
  c->np = 0;
  if (np == 0) return PDR_OK;
  if (np > int64_t(0xFFFFFFFF)) return fail(PDR_ERR_ARG, "pdr_particles_upload: at most 2^32 - 1 particles");
  int rc = dalloc(c, (void**)&c->xp, 24 * size_t(np));
  if (rc == PDR_OK) rc = dalloc(c, (void**)&c->uprk, 24 * size_t(np));
  if (rc == PDR_OK) rc = dalloc(c, (void**)&c->up, 24 * size_t(np));
  if (rc == PDR_OK && c->keep_x0) rc = dalloc(c, (void**)&c->xp0, 24 * size_t(np));
  if (rc == PDR_OK) rc = dalloc(c, (void**)&c->pid, 4 * size_t(np));
  if (rc == PDR_OK && !c->ur) rc = dalloc(c, &c->ur, 3 * size_t(c->g.nr) * c->rsz);
  if (rc != PDR_OK) return rc;
  std::vector<uint32_t> ident(static_cast<size_t>(np));
  for (size_t s = 0; s < ident.size(); ++s) ident[s] = uint32_t(s);
  HIPCHK(hipMemcpy(c->pid, ident.data(), 4 * size_t(np), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->xp, xp, 24 * size_t(np), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->uprk, 0, 24 * size_t(np)));
  HIPCHK(hipMemset(c->up, 0, 24 * size_t(np)));
  // particle_xp0 = particle_xp at placement (src/pp_control.f90:405-407)
  if (c->xp0) HIPCHK(hipMemcpy(c->xp0, xp, 24 * size_t(np), hipMemcpyHostToDevice));
  c->np = np;
  c->since_sort = 0;
  if (c->sort_every > 0) {
    rc = particles_sort(c);
    if (rc != PDR_OK) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return PDR_OK;
}

int pdr_set_particle_sort(pdr_ctx* c, int32_t every) {
  if (!c || every < 0) return fail(PDR_ERR_ARG, "pdr_set_particle_sort: bad argument");
  c->sort_every = every;
  c->since_sort = 0;
  return PDR_OK;
}

int pdr_particles_download_aux(pdr_ctx* c, double* up, double* xp0) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  if (c->np == 0) return fail(PDR_ERR_STATE, "no particles set");
  if (xp0 && !c->xp0) return fail(PDR_ERR_STATE, "initial positions are not kept (pdr_particle_stats_config)");
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = up ? particles_to_host(c, c->up, up) : PDR_OK;
  if (rc == PDR_OK && xp0) rc = particles_to_host(c, c->xp0, xp0);
  return rc;
}

int pdr_particles_upload_aux(pdr_ctx* c, const double* up, const double* xp0) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  if (c->np == 0) return fail(PDR_ERR_STATE, "no particles set");
  if (xp0 && !c->xp0) return fail(PDR_ERR_STATE, "initial positions are not kept (pdr_particle_stats_config)");
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = up ? particles_from_host(c, up, c->up) : PDR_OK;
  if (rc == PDR_OK && xp0) rc = particles_from_host(c, xp0, c->xp0);
  return rc;
}

int pdr_particle_stats_config(pdr_ctx* c, int32_t enable, int32_t keep_initial, int32_t ts_release) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (keep_initial && !c->xp0 && c->np > 0) {
    int rc = dalloc(c, (void**)&c->xp0, 24 * size_t(c->np));
    if (rc != PDR_OK) return rc;
    HIPCHK(hipMemcpy(c->xp0, c->xp, 24 * size_t(c->np), hipMemcpyDeviceToDevice));
  }
  if (!keep_initial && c->xp0) {
    (void)hipFree(c->xp0);
    c->dev_bytes -= 24 * size_t(c->np);
    c->xp0 = nullptr;
  }
  c->keep_x0 = keep_initial ? 1 : 0;
  c->ts_release = ts_release;
  c->pst_on = enable ? 1 : 0;
  if (enable) {
    int rc = pstats_alloc(c);
    if (rc != PDR_OK) return rc;
  }
  return pstats_reset(c);
}

int pdr_particle_stats(pdr_ctx* c, int32_t* tstep, pdr_pstats_t* out) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  static_assert(sizeof(pdr_pstats_t) == 30 * sizeof(double), "pdr_pstats_t is 30 doubles");
  return pstats_read(c, tstep, reinterpret_cast<double*>(out));
}

int pdr_particles_download(pdr_ctx* c, double* xp, double* uprk) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  if (c->np == 0) return fail(PDR_ERR_STATE, "no particles set");
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = xp ? particles_to_host(c, c->xp, xp) : PDR_OK;
  if (rc == PDR_OK && uprk) rc = particles_to_host(c, c->uprk, uprk);
  return rc;
}

int pdr_debug_prepare(pdr_ctx* c, int32_t field, double* out) {
  if (!c || !out || field < 0 || field > 5) return fail(PDR_ERR_ARG, "pdr_debug_prepare: bad argument");
  // NS_R_CONV: fields 3..5 are du/dx (the reference's u4..u6)
  const Coef cf = conv_form(c) ? conv_coef(make_coef(c), 0) : make_coef(c);
  HIPCHK(launch_zinv(c, cf));
  const Geo& g = c->g;
  const size_t off = size_t(field) * size_t(g.m1) * c->csz;
  std::vector<double> pad(size_t(2 * g.m1));
  int rc = get_field(c, pad.data(), (const char*)c->w1 + off, 2 * g.m1);
  if (rc != PDR_OK) return rc;
  // drop the padding of the j stride: (0:1, 0:NKJ-1, 0:NZ-1, 0:NKI-1)
  for (long long line = 0; line < (long long)g.nki * g.nz; ++line)
    std::memcpy(out + 2 * line * g.nkj, pad.data() + 2 * line * g.ldj, sizeof(double) * 2 * size_t(g.nkj));
  return PDR_OK;
}

int pdr_copy_bandwidth(int64_t bytes, int32_t iters, double* gbps) {
  if (bytes < 16 || iters < 1 || !gbps) return fail(PDR_ERR_ARG, "pdr_copy_bandwidth: bad argument");
  bytes &= ~int64_t(15);
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, size_t(bytes));
  if (e == hipSuccess) e = hipMalloc(&b, size_t(bytes));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, size_t(bytes), s);
  if (e == hipSuccess) e = launch_copy(a, b, bytes, s);  // warm-up
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int q = 0; q < iters && e == hipSuccess; ++q) e = launch_copy(a, b, bytes, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess) return fail(PDR_ERR_HIP, std::string("copy bandwidth: ") + hipGetErrorString(e));
  *gbps = 2.0 * double(bytes) * iters / (double(ms) * 1e-3) / 1e9;
  return PDR_OK;
}

int pdr_hbm_probe(int32_t kind, int64_t bytes, int32_t iters, double* gbps) {
  if (kind < 0 || kind > 3 || bytes < 16 || iters < 1 || !gbps) return fail(PDR_ERR_ARG, "pdr_hbm_probe: bad argument");
  bytes &= ~int64_t(15);
  // streams per launch: copy reads and writes, read and write one pass, triad reads two and writes one
  const double passes[4] = {2.0, 1.0, 1.0, 3.0};
  void *a = nullptr, *b = nullptr, *d = nullptr;
  float* sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, size_t(bytes));
  if (e == hipSuccess && kind == 3) e = hipMalloc(&b, size_t(bytes));
  if (e == hipSuccess && kind != 1) e = hipMalloc(&d, size_t(bytes));
  if (e == hipSuccess) e = hipMalloc((void**)&sink, 256 * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  // inputs hold 1.0f (the read probe's sum is then never -1)
  if (e == hipSuccess) e = launch_probe(2, nullptr, nullptr, a, bytes, sink, s);
  if (e == hipSuccess && b) e = launch_probe(2, nullptr, nullptr, b, bytes, sink, s);
  if (e == hipSuccess) e = launch_probe(kind, a, b, d, bytes, sink, s);  // warm-up
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int q = 0; q < iters && e == hipSuccess; ++q) e = launch_probe(kind, a, b, d, bytes, sink, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  for (void* q : {a, b, d, (void*)sink})
    if (q) (void)hipFree(q);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(PDR_ERR_HIP, std::string("hbm probe: ") + hipGetErrorString(e));
  }
  *gbps = passes[kind] * double(bytes) * iters / (double(ms) * 1e-3) / 1e9;
  return PDR_OK;
}

int pdr_profile_enable(pdr_ctx* c, int32_t enable) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  c->prof = enable != 0;
  if (c->prof) {
    for (int k = 0; k < K_NCLASS; ++k) {
      c->prof_ms[k] = 0.0;
      c->prof_cnt[k] = 0;
    }
  }
  return PDR_OK;
}

int pdr_profile_get(pdr_ctx* c, int32_t n, char* names, int32_t namelen, int64_t* counts, double* total_ms) {
  if (!c) return fail(PDR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& e : c->ev_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess) {
      c->prof_ms[e.first] += ms;
      c->prof_cnt[e.first] += 1;
    }
    c->ev_pool.push_back(e.second.first);
    c->ev_pool.push_back(e.second.second);
  }
  c->ev_pending.clear();
  static const char* kNames[K_NCLASS] = {"zinv",      "yinv",    "xpass",  "yfwd", "zfwd", "formf",
                                         "zfwdformf", "tracers", "remesh", "misc", "stats"};
  std::string all;
  for (int k = 0; k < K_NCLASS && k < n; ++k) {
    if (counts) counts[k] = c->prof_cnt[k];
    if (total_ms) total_ms[k] = c->prof_ms[k];
    if (k) all += ",";
    all += kNames[k];
  }
  if (names && namelen > 0) {
    std::strncpy(names, all.c_str(), size_t(namelen) - 1);
    names[namelen - 1] = 0;
  }
  return PDR_OK;
}


// Forcing state in the reference's array order f(0:1, -NK:NK (k), -NK:NK (j), 0:NK (i))
// per component; the device holds [i][j+NK][k+NK][c].
static int force_check(pdr_ctx* c, int32_t nk, const void* a, const void* b, const void* d) {
  if (!c || !a || !b || !d) return fail(PDR_ERR_ARG, "forcing: null argument");
  if (!c->forcing) return fail(PDR_ERR_STATE, "forcing is off (F_TYPE != 1 or FORCE_STOP <= 0)");
  if (nk != c->force_nk) return fail(PDR_ERR_ARG, "forcing: nk does not match aint(FORCE_KF)");
  return PDR_OK;
}

int pdr_set_forcing(pdr_ctx* c, int32_t nk, const double* f1, const double* f2, const double* f3) {
  int rc = force_check(c, nk, f1, f2, f3);
  if (rc != PDR_OK) return rc;
  const int w = 2 * nk + 1;
  std::vector<double> dev(force_bytes(c) / 8);
  const double* src[3] = {f1, f2, f3};
  for (int i = 0; i <= nk; ++i)
    for (int j = 0; j < w; ++j)
      for (int k = 0; k < w; ++k)
        for (int q = 0; q < 3; ++q) {
          const size_t h = 2 * ((size_t(i) * w + j) * w + k);
          const size_t d = 2 * (((size_t(i) * w + j) * w + k) * 3 + q);
          dev[d] = src[q][h];
          dev[d + 1] = src[q][h + 1];
        }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->force, dev.data(), dev.size() * 8, hipMemcpyHostToDevice));
  return PDR_OK;
}

int pdr_get_forcing(pdr_ctx* c, int32_t nk, double* f1, double* f2, double* f3) {
  int rc = force_check(c, nk, f1, f2, f3);
  if (rc != PDR_OK) return rc;
  const int w = 2 * nk + 1;
  std::vector<double> dev(force_bytes(c) / 8);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(dev.data(), c->force, dev.size() * 8, hipMemcpyDeviceToHost));
  double* dst[3] = {f1, f2, f3};
  for (int i = 0; i <= nk; ++i)
    for (int j = 0; j < w; ++j)
      for (int k = 0; k < w; ++k)
        for (int q = 0; q < 3; ++q) {
          const size_t h = 2 * ((size_t(i) * w + j) * w + k);
          const size_t d = 2 * (((size_t(i) * w + j) * w + k) * 3 + q);
          dst[q][h] = dev[d];
          dst[q][h + 1] = dev[d + 1];
        }
  return PDR_OK;
}

}  // extern "C"
