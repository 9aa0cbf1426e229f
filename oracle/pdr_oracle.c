/*
 * pdr_oracle.c -- FP64 CPU restatement of PANDORA 2.0's pseudo-spectral RK3 step.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product never
 * links or calls it.
 *
 * Each routine below follows the reference operation for operation (citations
 * are file:line in the reference repository).  The FFT is this file's own
 * recursive mixed-radix Cooley-Tukey transform with twiddles evaluated in long
 * double; it follows the FFTW conventions the reference uses (sign -1 forward,
 * +1 backward, unnormalised, c2r ignores the imaginary part of the DC term) and
 * is checked against brute-force DFTs in tests/.
 *
 * Parity pins (tests/test_oracle.py): reference known answer for the nonlinear
 * term (src/ptf_rk3.f90:1030-1198), brute-force DFTs and 3-D direct sums,
 * single-mode viscous decay under the Williamson scheme, inviscid energy
 * conservation of the truncated nonlinear term, S=0 shear == isotropic,
 * shear energy budget, remesh map enumeration, Neville polynomial reproduction,
 * uniform-flow tracer advection.
 */
#include "pdr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Constants: src/pg_constants.f90:26-141.  LONGONEPI is assigned from a      */
/* default-real (float32) literal, so ONEPI/TWOPI are float32 pi widened.     */
/* ------------------------------------------------------------------------- */
#define ONEPI_REF ((double)3.14159265358979323846f)
#define TWOPI_REF (2.0 * ONEPI_REF)
static const double THIRD = 1.0 / 3.0;
static const double FIVE_OVER_NINE = 5.0 / 9.0;
static const double FIFTEEN_OVER_SIXTEEN = 15.0 / 16.0;
static const double ONEFIVETHREE_OVER_ONETWOEIGHT = 153.0 / 128.0;
static const double EIGHT_OVER_FIFTEEN = 8.0 / 15.0;

/* ------------------------------------------------------------------------- */
/* FFT: recursive mixed-radix decimation in time                              */
/* ------------------------------------------------------------------------- */
typedef struct fftp {
  int n;
  int nf;
  int f[48];
  double* tw; /* tw[2m], tw[2m+1] = cos, sin of -2 pi m / n */
} fftp;

static void fftp_init(fftp* p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
  for (int q = 3; m > 1; q += 2) {
    while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
  }
  p->tw = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
  const long double pil = 3.14159265358979323846264338327950288L;
  for (int k = 0; k < n; k++) {
    long double a = -2.0L * pil * (long double)k / (long double)n;
    p->tw[2 * k] = (double)cosl(a);
    p->tw[2 * k + 1] = (double)sinl(a);
  }
}

static void fftp_free(fftp* p) { free(p->tw); p->tw = NULL; }

static void fft_rec(const fftp* p, const double* in, long is, double* out, int n, int fi,
                    int tws, int sign) {
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  const int r = p->f[fi];
  const int m = n / r;
  for (int q = 0; q < r; q++)
    fft_rec(p, in + 2 * q * is, is * r, out + 2 * (long)q * m, m, fi + 1, tws * r, sign);
  const int N = p->n;
  const double sg = (sign > 0) ? -1.0 : 1.0; /* conj of table for sign +1 */
  double t[2 * 64];
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < r; q++) {
      double yr = out[2 * (q * m + k)], yi = out[2 * (q * m + k) + 1];
      long e = ((long)q * k * tws) % N;
      double wr = p->tw[2 * e], wi = sg * p->tw[2 * e + 1];
      t[2 * q] = yr * wr - yi * wi;
      t[2 * q + 1] = yr * wi + yi * wr;
    }
    for (int s = 0; s < r; s++) {
      double ar = 0.0, ai = 0.0;
      for (int q = 0; q < r; q++) {
        long e = ((long)((q * s) % r) * (N / r)) % N;
        double wr = p->tw[2 * e], wi = sg * p->tw[2 * e + 1];
        ar += t[2 * q] * wr - t[2 * q + 1] * wi;
        ai += t[2 * q] * wi + t[2 * q + 1] * wr;
      }
      out[2 * (k + s * m)] = ar;
      out[2 * (k + s * m) + 1] = ai;
    }
  }
}

/* out must not alias in */
static void fft_exec(const fftp* p, const double* in, double* out, int sign) {
  fft_rec(p, in, 1, out, p->n, 0, 1, sign);
}

void orc_fft(int n, int sign, const double* in, double* out) {
  fftp p;
  fftp_init(&p, n);
  fft_exec(&p, in, out, sign);
  fftp_free(&p);
}

/* ------------------------------------------------------------------------- */
/* Context                                                                    */
/* ------------------------------------------------------------------------- */
struct orc {
  orc_params p;
  int nx, ny, nz;
  int ki, kj, kk;       /* KI_MAX, KJ_MAX, KK_MAX   src/pg_domain.f90:220-228 */
  int nki, nkj, nkk;    /* NODES_KI, NODES_KJ, NODES_KK */
  long mw, m1;
  int kkmax;
  double bmat[9], binv[9], amat[9];
  int bval[9], aval[9];
  double dt;            /* sim_dt, SAVEd in g_RK3Calc (src/pg_rk3.f90:92) */
  double ucfl;
  double dx, dy, dz, lx, ly, lz;
  double dx_lab;        /* dx_laboratoryframe = TWOPI B110 / NODES_X */
  double* u[3];
  double* rk[3];
  double* w[12];        /* u1..u6_1r2w (u1..u12 for R_CONV) */
  int nf;               /* inverse-transformed fields: 6, or 12 for R_CONV */
  double* ur[3];        /* real-space velocity for tracers */
  long np;
  double* xp;
  double* up;           /* particle_up: lab-frame fluid velocity of the last stage */
  double* uprk;
  double* xp0;          /* particle_xp0 (P_KEEP_INITIAL_POSITION) or NULL */
  int pst_on;           /* tracer statistics at stage 1 (P_TRACK_PART) */
  int keep_x0;          /* P_KEEP_INITIAL_POSITION */
  int ts_release;       /* P_TS_RELEASE */
  double xrel_prev[3];  /* p_stats' saved xrel (zero at start) */
  double pst[30];       /* last p_StatsWrite record */
  int pst_tstep;        /* its tstep, -1 before the first */
  fftp fz, fy, fx;
  int nthreads;
  int forcing;          /* F_TYPE 1 and FORCE_STOP > 0 */
  int force_nk;         /* FORCE_NK = aint(FORCE_KF), src/pf_force.f90:289 */
  double* force;        /* [c][i][j+NK][k+NK][re,im] */
  uint64_t force_calls; /* f_ForceCalc calls that drew (draw stream index) */
};

#define T9(a, m, n) ((a)[((m) - 1) + 3 * ((n) - 1)])

static int nthr(const orc* o) {
#ifdef _OPENMP
  return o->nthreads;
#else
  (void)o;
  return 1;
#endif
}

orc* orc_create(const orc_params* p) {
  /* NS_DEFORMED: 1 BF_ROT, 2 R_CONV, 3 R_ROT */
  if (p->f_type != 1 && (p->ns_deformed < 1 || p->ns_deformed > 3)) return NULL;
  orc* o = (orc*)calloc(1, sizeof(orc));
  o->p = *p;
  o->pst_tstep = -1;
  o->nx = p->nodes_x; o->ny = p->nodes_y; o->nz = p->nodes_z;
  /* src/pg_domain.f90:220-228 */
  o->ki = o->nx / 3; o->kj = o->ny / 3; o->kk = o->nz / 3;
  o->nki = o->ki + 1; o->nkj = 2 * o->kj + 1; o->nkk = 2 * o->kk + 1;
  o->mw = (long)o->nki * o->nkj * o->nkk;
  o->m1 = (long)o->nki * o->nkj * o->nz;
  const int b110 = p->b110 > 0 ? p->b110 : 1;
  const int b220 = p->b220 > 0 ? p->b220 : 1;
  const int b330 = p->b330 > 0 ? p->b330 : 1;
  /* src/pf_fluidstats.f90:343-374 (k_min = 1): k_max over the stretched index
   * ranges outside the isotropic form */
  if (p->f_type == 1)
    o->kkmax = (int)lround(sqrt((double)(o->ki * o->ki + o->kj * o->kj + o->kk * o->kk)));
  else
    o->kkmax = (int)lround(sqrt(((double)(o->ki * o->ki) / (double)(b110 * b110) +
                                 (double)(o->kj * o->kj) / (double)(b220 * b220)) +
                                (double)(o->kk * o->kk) / (double)(b330 * b330)));
  /* src/pg_domain.f90:168-216 */
  o->lx = TWOPI_REF; o->ly = TWOPI_REF; o->lz = TWOPI_REF;
  o->dx = o->lx / (double)o->nx; o->dy = o->ly / (double)o->ny; o->dz = o->lz / (double)o->nz;
  o->dx_lab = (TWOPI_REF * (double)b110) / (double)o->nx;
  for (int a = 0; a < 9; a++) { o->bmat[a] = o->binv[a] = o->amat[a] = 0.0; o->bval[a] = o->aval[a] = 0; }
  T9(o->binv, 1, 1) = (double)b110; T9(o->bval, 1, 1) = 1;
  T9(o->binv, 2, 2) = (double)b220; T9(o->bval, 2, 2) = 1;
  T9(o->binv, 3, 3) = (double)b330; T9(o->bval, 3, 3) = 1;
  if (p->f_type != 1 && p->f_type != 5) {
    T9(o->bval, 1, 3) = 1;
    T9(o->amat, 1, 3) = p->f_deform_s;
    T9(o->aval, 1, 3) = 1;
  }
  T9(o->bmat, 1, 1) = 1.0 / T9(o->binv, 1, 1);
  T9(o->bmat, 2, 2) = 1.0 / T9(o->binv, 2, 2);
  T9(o->bmat, 3, 3) = 1.0 / T9(o->binv, 3, 3);
  o->dt = 99999.0;
  for (int c = 0; c < 3; c++) {
    o->u[c] = (double*)calloc((size_t)(2 * o->mw), sizeof(double));
    o->rk[c] = (double*)calloc((size_t)(2 * o->mw), sizeof(double));
    o->ur[c] = NULL;
  }
  /* R_CONV transforms u and its nine raw-wavenumber gradients (src/pf_arrays.f90:788-964) */
  o->nf = (p->f_type != 1 && p->ns_deformed == 2) ? 12 : 6;
  for (int f = 0; f < o->nf; f++) o->w[f] = (double*)calloc((size_t)(2 * o->m1), sizeof(double));
  fftp_init(&o->fz, o->nz);
  fftp_init(&o->fy, o->ny);
  fftp_init(&o->fx, o->nx);
#ifdef _OPENMP
  o->nthreads = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#else
  o->nthreads = 1;
#endif
  o->forcing = ((p->f_type == 1 || p->f_type == 5) && p->force_stop > 0);
  o->force_nk = o->forcing ? (int)trunc(p->force_kf) : 0;
  if (o->forcing) {
    const long w = 2 * o->force_nk + 1;
    o->force = (double*)calloc((size_t)(3 * (o->force_nk + 1) * w * w * 2), sizeof(double));
  }
  return o;
}

/* ------------------------------------------------------------------------- */
/* Forcing: f_ForceCalc, src/pf_force.f90:106-268                             */
/* ------------------------------------------------------------------------- */
static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static double u01(uint64_t seed, uint64_t idx) {
  const uint64_t h = splitmix64(splitmix64(seed) ^ (idx * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
/* g_RandomNormal (src/pg_random.f90:112-146), mean 0, variance 1, on
 * counter-based uniforms: attempt a reads uniforms key*64 + 2a, 2a + 1. */
void orc_normal_pair(uint64_t seed, uint64_t key, double* r1, double* r2) {
  *r1 = 0.0; *r2 = 0.0;
  for (int a = 0; a < 32; a++) {
    const double x1 = 2.0 * u01(seed, key * 64u + (uint64_t)(2 * a)) - 1.0;
    const double x2 = 2.0 * u01(seed, key * 64u + (uint64_t)(2 * a + 1)) - 1.0;
    const double w = x1 * x1 + x2 * x2;
    if (w < 1.0 && w > 0.0) {
      const double f = sqrt((-2.0 * log(w)) / w);
      *r1 = x1 * f; *r2 = x2 * f;
      return;
    }
  }
}

int orc_force_nk(const orc* o) { return o->forcing ? o->force_nk : -1; }

static double* fat(const orc* o, int c, int i, int j, int k) {
  const long w = 2 * o->force_nk + 1;
  return o->force + 2 * ((((long)c * (o->force_nk + 1) + i) * w + (j + o->force_nk)) * w + (k + o->force_nk));
}

void orc_set_forcing(orc* o, const double* f1, const double* f2, const double* f3) {
  const double* f[3] = {f1, f2, f3};
  const int nk = o->force_nk;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i <= nk; i++)
      for (int j = -nk; j <= nk; j++)
        for (int k = -nk; k <= nk; k++) {
          const long h = 2 * (((long)i * (2 * nk + 1) + (j + nk)) * (2 * nk + 1) + (k + nk));
          double* d = fat(o, c, i, j, k);
          d[0] = f[c][h]; d[1] = f[c][h + 1];
        }
}

void orc_get_forcing(const orc* o, double* f1, double* f2, double* f3) {
  double* f[3] = {f1, f2, f3};
  const int nk = o->force_nk;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i <= nk; i++)
      for (int j = -nk; j <= nk; j++)
        for (int k = -nk; k <= nk; k++) {
          const long h = 2 * (((long)i * (2 * nk + 1) + (j + nk)) * (2 * nk + 1) + (k + nk));
          const double* d = fat(o, c, i, j, k);
          f[c][h] = d[0]; f[c][h + 1] = d[1];
        }
}

/* one OU update of the three components of mode (i, j, k) with draw keys key..key+2 */
static void ou_mode(orc* o, int i, int j, int k, double om, double sd, uint64_t key) {
  for (int c = 0; c < 3; c++) {
    double re, im;
    orc_normal_pair((uint64_t)o->p.force_seed, key + (uint64_t)c, &re, &im);
    double* p = fat(o, c, i, j, k);
    p[0] = p[0] * om + re * sd;
    p[1] = p[1] * om + im * sd;
  }
}
static void conj_copy(orc* o, int i, int j, int k, int si, int sj, int sk) {
  for (int c = 0; c < 3; c++) {
    const double* s = fat(o, c, si, sj, sk);
    double* d = fat(o, c, i, j, k);
    d[0] = s[0]; d[1] = -s[1];
  }
}

void orc_force_calc(orc* o, int tstep, double dt) {
  if (!o->forcing) return;
  const int nk = o->force_nk;
  const long w = 2 * nk + 1;
  if (tstep >= o->p.force_stop) {  /* IFDECAY */
    memset(o->force, 0, sizeof(double) * (size_t)(3 * (nk + 1) * w * w * 2));
    return;
  }
  const double a = dt / o->p.force_tf;
  const double om = 1.0 - a;
  const double sd = sqrt(2.0 * (o->p.force_sigma * o->p.force_sigma) * a);
  const uint64_t key0 = o->force_calls << 24;
  /* i = 0 plane, conjugate symmetric (src/pf_force.f90:143-221) */
  for (int j = 0; j <= nk; j++)
    for (int k = 0; k <= nk; k++) {
      const uint64_t t = (uint64_t)(j * (nk + 1) + k);
      ou_mode(o, 0, j, k, om, sd, key0 + 6 * t);
      conj_copy(o, 0, -j, -k, 0, j, k);
      if (j > 0) {
        ou_mode(o, 0, -j, k, om, sd, key0 + 6 * t + 3);
        conj_copy(o, 0, j, -k, 0, -j, k);
      }
    }
  /* i >= 1 (src/pf_force.f90:223-258) */
  const uint64_t n0 = (uint64_t)((nk + 1) * (nk + 1));
  uint64_t u = 0;
  for (int i = 1; i <= nk; i++)
    for (int j = -nk; j <= nk; j++)
      for (int k = -nk; k <= nk; k++, u++) ou_mode(o, i, j, k, om, sd, key0 + 6 * n0 + 3 * u);
  o->force_calls++;
}

void orc_destroy(orc* o) {
  free(o->force);
  if (!o) return;
  for (int c = 0; c < 3; c++) { free(o->u[c]); free(o->rk[c]); free(o->ur[c]); }
  for (int f = 0; f < o->nf; f++) free(o->w[f]);
  free(o->xp); free(o->up); free(o->uprk); free(o->xp0);
  fftp_free(&o->fz); fftp_free(&o->fy); fftp_free(&o->fx);
  free(o);
}

void orc_sizes(const orc* o, int64_t* s) {
  s[0] = o->nki; s[1] = o->nkj; s[2] = o->nkk; s[3] = o->mw; s[4] = o->m1; s[5] = o->kkmax;
}

void orc_set_state(orc* o, const double* a, const double* b, const double* c) {
  size_t n = sizeof(double) * 2 * (size_t)o->mw;
  memcpy(o->u[0], a, n); memcpy(o->u[1], b, n); memcpy(o->u[2], c, n);
}
void orc_get_state(const orc* o, double* a, double* b, double* c) {
  size_t n = sizeof(double) * 2 * (size_t)o->mw;
  memcpy(a, o->u[0], n); memcpy(b, o->u[1], n); memcpy(c, o->u[2], n);
}
void orc_set_rk(orc* o, const double* a, const double* b, const double* c) {
  size_t n = sizeof(double) * 2 * (size_t)o->mw;
  memcpy(o->rk[0], a, n); memcpy(o->rk[1], b, n); memcpy(o->rk[2], c, n);
}
void orc_get_rk(const orc* o, double* a, double* b, double* c) {
  size_t n = sizeof(double) * 2 * (size_t)o->mw;
  memcpy(a, o->rk[0], n); memcpy(b, o->rk[1], n); memcpy(c, o->rk[2], n);
}
void orc_set_frame(orc* o, const double* b, const double* bi) {
  memcpy(o->bmat, b, 9 * sizeof(double)); memcpy(o->binv, bi, 9 * sizeof(double));
}
void orc_get_frame(const orc* o, double* b, double* bi) {
  memcpy(b, o->bmat, 9 * sizeof(double)); memcpy(bi, o->binv, 9 * sizeof(double));
}
void orc_set_dt(orc* o, double dt) { o->dt = dt; }
double orc_get_dt(const orc* o) { return o->dt; }
double orc_get_ucfl(const orc* o) { return o->ucfl; }
void orc_get_1r2w(const orc* o, int f, double* out) {
  memcpy(out, o->w[f], sizeof(double) * 2 * (size_t)o->m1);
}

/* f_ArraysWavenumber, src/pf_arrays.f90:1825-1834 (K0 = 1) */
static inline double wvn(int k, int kmax) {
  return (k <= kmax) ? (double)k : (double)(k - (2 * kmax + 1));
}

/* ------------------------------------------------------------------------- */
/* f_RK3Derivative (src/pf_rk3.f90:3735-3743) and vorticity forms            */
/* ------------------------------------------------------------------------- */
static inline void deriv(const double* u, double k, double* d) {
  d[1] = k * u[0];
  d[0] = -k * u[1];
}

/* Rogallo forms (NS_DEFORMED = 2 R_CONV, 3 R_ROT) of a sheared run */
static inline int rogallo(const orc* o) { return o->p.f_type != 1 && o->p.ns_deformed != 1; }

/* f_RK3DerivativeRRot (src/pf_rk3.f90:3749-3780): derivative along the sheared
 * direction j, sum over m of bmat(m,j) k_m; the existence test there has an empty
 * body, so every m contributes (absent entries of bmat are zero). */
static inline void deriv_rrot(const orc* o, const double* uk, const double* k, int j, double* d) {
  d[0] = 0.0;
  d[1] = 0.0;
  for (int m = 1; m <= 3; m++) {
    d[1] = d[1] + (T9(o->bmat, m, j) * k[m - 1] * uk[0]);
    d[0] = d[0] - (T9(o->bmat, m, j) * k[m - 1] * uk[1]);
  }
}

/* f_RK3VorticityBFRot, src/pf_rk3.f90:3786-3843.  om is (0:1,1:3) row major [c][2]. */
static void vort_bfrot(const orc* o, const double* un, const double* k, int n, double om[3][2]) {
  for (int c = 0; c < 3; c++) om[c][0] = om[c][1] = 0.0;
  const double* B = o->bmat;
  const double* BI = o->binv;
  const int* bv = o->bval;
  for (int m = 1; m <= 3; m++) {
    if (T9(bv, m, 1)) {
      if (T9(bv, 3, n)) {
        om[1][1] = om[1][1] - (T9(B, m, 1) * T9(BI, 3, n) * k[m - 1] * un[0]);
        om[1][0] = om[1][0] + (T9(B, m, 1) * T9(BI, 3, n) * k[m - 1] * un[1]);
      }
      if (T9(bv, 2, n)) {
        om[2][1] = om[2][1] + (T9(B, m, 1) * T9(BI, 2, n) * k[m - 1] * un[0]);
        om[2][0] = om[2][0] - (T9(B, m, 1) * T9(BI, 2, n) * k[m - 1] * un[1]);
      }
    }
    if (T9(bv, m, 2)) {
      if (T9(bv, 3, n)) {
        om[0][1] = om[0][1] + (T9(B, m, 2) * T9(BI, 3, n) * k[m - 1] * un[0]);
        om[0][0] = om[0][0] - (T9(B, m, 2) * T9(BI, 3, n) * k[m - 1] * un[1]);
      }
      if (T9(bv, 1, n)) {
        om[2][1] = om[2][1] - (T9(B, m, 2) * T9(BI, 1, n) * k[m - 1] * un[0]);
        om[2][0] = om[2][0] + (T9(B, m, 2) * T9(BI, 1, n) * k[m - 1] * un[1]);
      }
    }
    if (T9(bv, m, 3)) {
      if (T9(bv, 2, n)) {
        om[0][1] = om[0][1] - (T9(B, m, 3) * T9(BI, 2, n) * k[m - 1] * un[0]);
        om[0][0] = om[0][0] + (T9(B, m, 3) * T9(BI, 2, n) * k[m - 1] * un[1]);
      }
      if (T9(bv, 1, n)) {
        om[1][1] = om[1][1] + (T9(B, m, 3) * T9(BI, 1, n) * k[m - 1] * un[0]);
        om[1][0] = om[1][0] - (T9(B, m, 3) * T9(BI, 1, n) * k[m - 1] * un[1]);
      }
    }
  }
}

/* Vorticity of one mode, accumulated component by component as in
 * f_RK3PrepareVelocity (src/pf_rk3.f90:336-723). */
static void mode_vorticity(const orc* o, const double* u1, const double* u2, const double* u3,
                           double ki_, double kj_, double kk_, double om[3][2]) {
  if (o->p.f_type == 1) {
    double dudy[2], dudz[2], dvdx[2], dvdz[2], dwdx[2], dwdy[2];
    deriv(u1, kj_, dudy); deriv(u1, kk_, dudz);
    om[0][0] = 0.0; om[0][1] = 0.0;
    om[1][0] = dudz[0]; om[1][1] = dudz[1];
    om[2][0] = -dudy[0]; om[2][1] = -dudy[1];
    deriv(u2, ki_, dvdx); deriv(u2, kk_, dvdz);
    om[0][0] = om[0][0] - dvdz[0]; om[0][1] = om[0][1] - dvdz[1];
    om[2][0] = om[2][0] + dvdx[0]; om[2][1] = om[2][1] + dvdx[1];
    deriv(u3, ki_, dwdx); deriv(u3, kj_, dwdy);
    om[0][0] = om[0][0] + dwdy[0]; om[0][1] = om[0][1] + dwdy[1];
    om[1][0] = om[1][0] - dwdx[0]; om[1][1] = om[1][1] - dwdx[1];
  } else if (rogallo(o)) {
    /* NS_R_ROT branches of f_RK3PrepareVelocity (src/pf_rk3.f90:416-421, 533-537, 645-649) */
    const double kv[3] = {ki_, kj_, kk_};
    double dudy[2], dudz[2], dvdx[2], dvdz[2], dwdx[2], dwdy[2];
    deriv_rrot(o, u1, kv, 2, dudy); deriv_rrot(o, u1, kv, 3, dudz);
    om[0][0] = 0.0; om[0][1] = 0.0;
    om[1][0] = dudz[0]; om[1][1] = dudz[1];
    om[2][0] = -dudy[0]; om[2][1] = -dudy[1];
    deriv_rrot(o, u2, kv, 1, dvdx); deriv_rrot(o, u2, kv, 3, dvdz);
    om[0][0] = om[0][0] - dvdz[0]; om[0][1] = om[0][1] - dvdz[1];
    om[2][0] = om[2][0] + dvdx[0]; om[2][1] = om[2][1] + dvdx[1];
    deriv_rrot(o, u3, kv, 1, dwdx); deriv_rrot(o, u3, kv, 2, dwdy);
    om[0][0] = om[0][0] + dwdy[0]; om[0][1] = om[0][1] + dwdy[1];
    om[1][0] = om[1][0] - dwdx[0]; om[1][1] = om[1][1] - dwdx[1];
  } else {
    const double kv[3] = {ki_, kj_, kk_};
    double t[3][2];
    vort_bfrot(o, u1, kv, 1, t);
    for (int c = 0; c < 3; c++) { om[c][0] = t[c][0]; om[c][1] = t[c][1]; }
    vort_bfrot(o, u2, kv, 2, t);
    for (int c = 0; c < 3; c++) { om[c][0] = om[c][0] + t[c][0]; om[c][1] = om[c][1] + t[c][1]; }
    vort_bfrot(o, u3, kv, 3, t);
    for (int c = 0; c < 3; c++) { om[c][0] = om[c][0] + t[c][0]; om[c][1] = om[c][1] + t[c][1]; }
  }
}

/* ------------------------------------------------------------------------- */
/* f_RK3PrepareRealSpace: vorticity, zero pad in k, z FFT (sign +1), store  */
/* in the 1r2w arrays.  src/pf_rk3.f90:205-948, src/pf_fftw.f90:586-624     */
/* ------------------------------------------------------------------------- */
void orc_prepare_realspace(orc* o) {
  const int nz = o->nz, nkj = o->nkj, nkk = o->nkk, K = o->kk;
#pragma omp parallel num_threads(nthr(o))
  {
    double* in = (double*)malloc(sizeof(double) * 2 * nz);
    double* out = (double*)malloc(sizeof(double) * 2 * nz);
    double* om = (double*)malloc(sizeof(double) * 2 * (size_t)nkk * 9);
    const int conv = (o->nf == 12);
#pragma omp for schedule(dynamic)
    for (int i = 0; i < o->nki; i++) {
      const double wi = wvn(i, o->ki);
      for (int j = 0; j < nkj; j++) {
        const double wj = wvn(j, o->kj);
        const long base = ((long)i * nkj + j) * nkk;
        for (int k = 0; k < nkk; k++) {
          const double wk = wvn(k, K);
          if (conv) {
            /* NS_R_CONV branches of f_RK3PrepareVelocity (src/pf_rk3.f90:422-427, 538-543,
             * 650-656): u(4 + 3d + c) = d u_c / d x_d with the raw wavenumbers */
            const double kd[3] = {wi, wj, wk};
            for (int c = 0; c < 3; c++)
              for (int d = 0; d < 3; d++) {
                double g[2];
                deriv(&o->u[c][2 * (base + k)], kd[d], g);
                om[2 * ((3 * d + c) * nkk + k)] = g[0];
                om[2 * ((3 * d + c) * nkk + k) + 1] = g[1];
              }
            continue;
          }
          double t[3][2];
          mode_vorticity(o, &o->u[0][2 * (base + k)], &o->u[1][2 * (base + k)],
                         &o->u[2][2 * (base + k)], wi, wj, wk, t);
          for (int c = 0; c < 3; c++) { om[2 * (c * nkk + k)] = t[c][0]; om[2 * (c * nkk + k) + 1] = t[c][1]; }
        }
        for (int f = 0; f < o->nf; f++) {
          const double* src = (f < 3) ? &o->u[f][2 * base] : &om[2 * (f - 3) * nkk];
          /* zero-padded buffer: K+1 positive, NZ-2K-1 zeros, K negative */
          for (int z = 0; z < 2 * nz; z++) in[z] = 0.0;
          for (int k = 0; k <= K; k++) { in[2 * k] = src[2 * k]; in[2 * k + 1] = src[2 * k + 1]; }
          for (int k = K + 1; k < nkk; k++) {
            int z = nz - nkk + k;
            in[2 * z] = src[2 * k]; in[2 * z + 1] = src[2 * k + 1];
          }
          fft_exec(&o->fz, in, out, +1);
          double* dst = o->w[f];
          for (int z = 0; z < nz; z++) {
            long d = (((long)i * nz + z) * nkj + j) * 2;
            dst[d] = out[2 * z]; dst[d + 1] = out[2 * z + 1];
          }
        }
      }
    }
    free(in); free(out); free(om);
  }
}

/* ------------------------------------------------------------------------- */
/* Nonlinear products in real space                                          */
/* ------------------------------------------------------------------------- */
/* f_RK3ComputeNonlinearIsotropic, src/pf_rk3.f90:1597-1628 */
static inline void nl_iso(const double* u, const double* w, double* n) {
  n[0] = u[2] * w[1] - u[1] * w[2];
  n[1] = u[0] * w[2] - u[2] * w[0];
  n[2] = u[1] * w[0] - u[0] * w[1];
}
/* f_RK3ComputeNonlinearBFRot, src/pf_rk3.f90:1634-1711 */
static inline void nl_bfrot(const orc* o, const double* u, const double* om, double* n) {
  double n1 = 0.0, n2 = 0.0, n3 = 0.0;
  const double* BI = o->binv;
  const int* bv = o->bval;
  for (int m = 1; m <= 3; m++) {
    if (T9(bv, 1, m)) {
      n3 = n3 - (om[1] * T9(BI, 1, m) * u[m - 1]);
      n2 = n2 + (om[2] * T9(BI, 1, m) * u[m - 1]);
    }
    if (T9(bv, 2, m)) {
      n1 = n1 - (om[2] * T9(BI, 2, m) * u[m - 1]);
      n3 = n3 + (om[0] * T9(BI, 2, m) * u[m - 1]);
    }
    if (T9(bv, 3, m)) {
      n1 = n1 + (om[1] * T9(BI, 3, m) * u[m - 1]);
      n2 = n2 - (om[0] * T9(BI, 3, m) * u[m - 1]);
    }
  }
  n[0] = n1; n[1] = n2; n[2] = n3;
}

/* f_RK3ComputeNonlinearRConv, src/pf_rk3.f90:1717-1774: NL_i = sum_j sum_k
 * bmat(k,j) du_i/dx_k u_j over the existing entries; g[3k + i] = du_i/dx_k */
static void nl_rconv(const orc* o, const double* u, const double* g, double* n) {
  for (int i = 1; i <= 3; i++) {
    double a = 0.0;
    for (int j = 1; j <= 3; j++)
      for (int k = 1; k <= 3; k++)
        if (T9(o->bval, k, j)) a = a + (T9(o->bmat, k, j) * g[3 * (k - 1) + (i - 1)] * u[j - 1]);
    n[i - 1] = a;
  }
}

/* ------------------------------------------------------------------------- */
/* f_RK3RealSpace: per z-plane y/x inverse transforms, nonlinear term,       */
/* forward x/y transforms with truncation.                                   */
/* src/pf_rk3.f90:960-1237, 1245-1498, 1780-1940; src/pf_fftw.f90:632-834   */
/* ------------------------------------------------------------------------- */
/* u1..u6_2r of plane z from the 1r2w arrays: y transforms for each i
 * (src/pf_rk3.f90:1437-1498, pf_fftw.f90:673-759), zero pad i to NX/2+1 and c2r
 * in x (imaginary part of the DC term ignored).  r holds 6 planes [f][y][x];
 * in/out hold 2 max(NX,NY) doubles, pc 2 NY NKI. */
static void plane_fields(const orc* o, int z, double* in, double* out, double* pc, double* r) {
  const int nx = o->nx, ny = o->ny, nz = o->nz, nki = o->nki, nkj = o->nkj, KJ = o->kj;
  const long plane = (long)nx * ny;
  for (int f = 0; f < o->nf; f++) {
    /* y transforms for each i (src/pf_rk3.f90:1437-1498, pf_fftw.f90:673-759) */
    for (int i = 0; i < nki; i++) {
      for (int y = 0; y < 2 * ny; y++) in[y] = 0.0;
      const double* src = &o->w[f][(((long)i * nz + z) * nkj) * 2];
      for (int j = 0; j <= KJ; j++) { in[2 * j] = src[2 * j]; in[2 * j + 1] = src[2 * j + 1]; }
      for (int j = KJ + 1; j < nkj; j++) {
        int y = ny - nkj + j;
        in[2 * y] = src[2 * j]; in[2 * y + 1] = src[2 * j + 1];
      }
      fft_exec(&o->fy, in, out, +1);
      for (int y = 0; y < ny; y++) {
        pc[2 * ((long)y * nki + i)] = out[2 * y];
        pc[2 * ((long)y * nki + i) + 1] = out[2 * y + 1];
      }
    }
    /* zero pad i to NX/2+1, c2r in x (imaginary part of the DC term ignored) */
    for (int y = 0; y < ny; y++) {
      for (int x = 0; x < 2 * nx; x++) in[x] = 0.0;
      for (int i = 0; i < nki; i++) {
        double re = pc[2 * ((long)y * nki + i)], im = pc[2 * ((long)y * nki + i) + 1];
        if (i == 0) { in[0] = re; in[1] = 0.0; continue; }
        in[2 * i] = re; in[2 * i + 1] = im;
        in[2 * (nx - i)] = re; in[2 * (nx - i) + 1] = -im;
      }
      fft_exec(&o->fx, in, out, +1);
      double* rr = &r[f * plane + (long)y * nx];
      for (int x = 0; x < nx; x++) rr[x] = out[2 * x];
    }
  }
}

/* u_lab(m) = sum_n binv(m,n) u(n) over the existing entries
 * (f_FluidstatsRs2D, src/pf_fluidstats.f90:853-866) */
static void lab_velocity(const orc* o, double u1, double u2, double u3, double* ul) {
  const double um[3] = {u1, u2, u3};
  if (o->p.f_type == 1) { ul[0] = u1; ul[1] = u2; ul[2] = u3; return; }
  for (int m = 1; m <= 3; m++) {
    ul[m - 1] = 0.0;
    for (int n = 1; n <= 3; n++)
      if (T9(o->bval, m, n)) ul[m - 1] = ul[m - 1] + T9(o->binv, m, n) * um[n - 1];
  }
}

void orc_realspace(orc* o, int stage, int tstep, double* dt) {
  (void)tstep;
  const int nx = o->nx, ny = o->ny, nz = o->nz, nki = o->nki, nkj = o->nkj, KJ = o->kj;
  const long plane = (long)nx * ny;
  const int want_tracers = (o->np > 0);
  if (want_tracers) {
    for (int c = 0; c < 3; c++)
      if (!o->ur[c]) o->ur[c] = (double*)malloc(sizeof(double) * (size_t)plane * nz);
  }
  double umax_all = 0.0;
#pragma omp parallel num_threads(nthr(o)) reduction(max : umax_all)
  {
    double* in = (double*)malloc(sizeof(double) * 2 * (size_t)(ny > nx ? ny : nx));
    double* out = (double*)malloc(sizeof(double) * 2 * (size_t)(ny > nx ? ny : nx));
    double* pc = (double*)malloc(sizeof(double) * 2 * (size_t)ny * nki);   /* ji plane */
    double* r = (double*)malloc(sizeof(double) * (size_t)o->nf * (size_t)plane);  /* u1..u6(12)_2r */
#pragma omp for schedule(dynamic)
    for (int z = 0; z < nz; z++) {
      plane_fields(o, z, in, out, pc, r);
      /* ucfl (src/pf_fluidstats.f90:914-917): computed at stage 1 from the
       * laboratory-frame velocity u_lab = binv u_mov (853-866; identity for
       * isotropic runs) */
      if (stage == 1) {
        for (long q = 0; q < plane; q++) {
          double ul[3];
          lab_velocity(o, r[q], r[plane + q], r[2 * plane + q], ul);
          double uu = fabs(ul[0]) + fabs(ul[1]) + fabs(ul[2]);
          if (uu > umax_all) umax_all = uu;
        }
      }
      /* nonlinear term */
      if (o->nf == 12) {
        for (long q = 0; q < plane; q++) {
          double uu[3] = {r[q], r[plane + q], r[2 * plane + q]};
          double g[9];
          for (int a = 0; a < 9; a++) g[a] = r[(3 + a) * plane + q];
          double nn[3];
          nl_rconv(o, uu, g, nn);
          r[3 * plane + q] = nn[0]; r[4 * plane + q] = nn[1]; r[5 * plane + q] = nn[2];
        }
      } else
      for (long q = 0; q < plane; q++) {
        double uu[3] = {r[q], r[plane + q], r[2 * plane + q]};
        double ww[3] = {r[3 * plane + q], r[4 * plane + q], r[5 * plane + q]};
        double nn[3];
        if (o->p.f_type == 1 || rogallo(o)) nl_iso(uu, ww, nn); else nl_bfrot(o, uu, ww, nn);  /* 1080-1093 */
        r[3 * plane + q] = nn[0]; r[4 * plane + q] = nn[1]; r[5 * plane + q] = nn[2];
      }
      if (want_tracers) {
        for (int c = 0; c < 3; c++) memcpy(&o->ur[c][(long)z * plane], &r[c * plane], sizeof(double) * plane);
      }
      /* forward: r2c in x, keep i <= KI, y FFT (sign -1), truncate j */
      for (int f = 3; f < 6; f++) {
        for (int y = 0; y < ny; y++) {
          const double* rr = &r[f * plane + (long)y * nx];
          for (int x = 0; x < nx; x++) { in[2 * x] = rr[x]; in[2 * x + 1] = 0.0; }
          fft_exec(&o->fx, in, out, -1);
          for (int i = 0; i < nki; i++) {
            pc[2 * ((long)y * nki + i)] = out[2 * i];
            pc[2 * ((long)y * nki + i) + 1] = out[2 * i + 1];
          }
        }
        for (int i = 0; i < nki; i++) {
          for (int y = 0; y < ny; y++) {
            in[2 * y] = pc[2 * ((long)y * nki + i)];
            in[2 * y + 1] = pc[2 * ((long)y * nki + i) + 1];
          }
          fft_exec(&o->fy, in, out, -1);
          double* dst = &o->w[f][(((long)i * nz + z) * nkj) * 2];
          for (int j = 0; j <= KJ; j++) { dst[2 * j] = out[2 * j]; dst[2 * j + 1] = out[2 * j + 1]; }
          for (int j = KJ + 1; j < nkj; j++) {
            int y = ny - nkj + j;
            dst[2 * j] = out[2 * y]; dst[2 * j + 1] = out[2 * y + 1];
          }
        }
      }
    }
    free(in); free(out); free(pc); free(r);
  }
  if (stage == 1) {
    o->ucfl = umax_all;
    /* f_RK3Tstep, src/pf_rk3.f90:3852-3948 */
    switch (o->p.rk3_timestep) {
      case 2: *dt = o->p.rk3_dt; break;
      case 3: *dt = (o->ucfl == 0.0) ? (double)0.02f : o->p.rk3_cfl * o->dx / o->ucfl; break;
      default:
        if (o->p.f_type == 1 || o->p.f_type == 5) {
          /* the deformed form: min(DX, DY, DZ) of the moving frame (3905-3919) */
          const double h = (o->p.f_type == 1) ? o->dx : fmin(o->dx, fmin(o->dy, o->dz));
          if (o->ucfl == 0.0) *dt = (double)0.02f;
          else {
            *dt = o->p.rk3_cfl * h / o->ucfl;
            if (*dt > o->p.force_tf) *dt = o->p.force_tf;
          }
        } else {
          *dt = ((0.5 / o->p.f_deform_s) / (double)o->p.f_remesh_t);
        }
        break;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Real-space statistics: f_FluidstatsRs2D + f_FluidstatsRsCollect            */
/* (src/pf_fluidstats.f90:827-951, 962-1146) of the current state.            */
/* raw[25]: sums v1v1 v2v2 v3v3 v1v2 v1v3 v2v3, v^3 (3), v^4 (3), w^2 (3),     */
/* w^3 (3), w^4 (3); maxima |v1|+|v2|+|v3|, |w1|, |w2|, |w3|.                  */
/* ------------------------------------------------------------------------- */
void orc_rs_raw(orc* o, double* raw) {
  orc_prepare_realspace(o);
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const long plane = (long)nx * ny;
  /* one row of 25 per z plane, planes in parallel, rows combined in z order */
  double* zr = (double*)calloc(25 * (size_t)nz, sizeof(double));
#pragma omp parallel num_threads(nthr(o))
  {
    double* in = (double*)malloc(sizeof(double) * 2 * (size_t)(ny > nx ? ny : nx));
    double* out = (double*)malloc(sizeof(double) * 2 * (size_t)(ny > nx ? ny : nx));
    double* pc = (double*)malloc(sizeof(double) * 2 * (size_t)ny * o->nki);
    double* r = (double*)malloc(sizeof(double) * (size_t)o->nf * (size_t)plane);
#pragma omp for schedule(dynamic)
    for (int z = 0; z < nz; z++) {
      double* rz = &zr[25 * (size_t)z];
      plane_fields(o, z, in, out, pc, r);
      for (long q = 0; q < plane; q++) {
        double v[3];
        lab_velocity(o, r[q], r[plane + q], r[2 * plane + q], v);
        const double w[3] = {r[3 * plane + q], r[4 * plane + q], r[5 * plane + q]};
        rz[0] += v[0] * v[0]; rz[1] += v[1] * v[1]; rz[2] += v[2] * v[2];
        rz[3] += v[0] * v[1]; rz[4] += v[0] * v[2]; rz[5] += v[1] * v[2];
        for (int c = 0; c < 3; c++) {
          rz[6 + c] += v[c] * v[c] * v[c];
          rz[9 + c] += v[c] * v[c] * v[c] * v[c];
          rz[12 + c] += w[c] * w[c];
          rz[15 + c] += w[c] * w[c] * w[c];
          rz[18 + c] += w[c] * w[c] * w[c] * w[c];
          if (fabs(w[c]) > rz[22 + c]) rz[22 + c] = fabs(w[c]);
        }
        const double u = fabs(v[0]) + fabs(v[1]) + fabs(v[2]);
        if (u > rz[21]) rz[21] = u;
      }
    }
    free(in); free(out); free(pc); free(r);
  }
  for (int q = 0; q < 25; q++) raw[q] = 0.0;
  for (int z = 0; z < nz; z++) {
    const double* rz = &zr[25 * (size_t)z];
    for (int q = 0; q < 21; q++) raw[q] += rz[q];
    for (int q = 21; q < 25; q++) if (rz[q] > raw[q]) raw[q] = rz[q];
  }
  free(zr);
}

/* Longitudinal two-point correlation of f_FluidstatsRs2D on TS_CORR steps
 * (src/pf_fluidstats.f90:933-945) and its normalisation in f_FluidstatsRsCollect
 * (1092-1106), on one box (x_min_2r = 0, x_width_2r = NX): for every grid point,
 * the laboratory-frame u1 there times the stored u1 at x + s (x periodic), for
 * s = 0 .. (NX-1)/2 + 1, summed into r1i(s) and, each product times dx, into
 * lint11; both divided by rs_u1u1 * NX NY NZ.  r1i holds NX/2 + 1 entries (the
 * reference's array extent: for odd NX its loop's last separation enters lint11
 * only).  Planes in parallel, each plane's sums combined in z order. */
void orc_rs_correlation(orc* o, double* r1i, double* lint11) {
  orc_prepare_realspace(o);
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const long plane = (long)nx * ny;
  const int smax = (nx - 1) / 2 + 1;
  const int w = smax + 3; /* separations 0..smax, lint11, u1u1 */
  double* zr = (double*)calloc((size_t)w * (size_t)nz, sizeof(double));
#pragma omp parallel num_threads(nthr(o))
  {
    double* in = (double*)malloc(sizeof(double) * 2 * (size_t)(ny > nx ? ny : nx));
    double* out = (double*)malloc(sizeof(double) * 2 * (size_t)(ny > nx ? ny : nx));
    double* pc = (double*)malloc(sizeof(double) * 2 * (size_t)ny * o->nki);
    double* r = (double*)malloc(sizeof(double) * (size_t)o->nf * (size_t)plane);
#pragma omp for schedule(dynamic)
    for (int z = 0; z < nz; z++) {
      double* rz = &zr[(size_t)w * z];
      plane_fields(o, z, in, out, pc, r);
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          const long q = (long)y * nx + x;
          double v[3];
          lab_velocity(o, r[q], r[plane + q], r[2 * plane + q], v);
          rz[smax + 2] += v[0] * v[0];
          for (int s = 0; s <= smax; s++) {
            int xx = x + s;
            if (xx >= nx) xx -= nx;
            const double c = v[0] * r[(long)y * nx + xx];
            rz[s] += c;
            rz[smax + 1] += c * o->dx_lab;
          }
        }
    }
    free(in); free(out); free(pc); free(r);
  }
  double* t = (double*)calloc((size_t)w, sizeof(double));
  for (int z = 0; z < nz; z++)
    for (int q = 0; q < w; q++) t[q] += zr[(size_t)w * z + q];
  const double n3 = ((double)nx * (double)ny) * (double)nz;
  const double u1u1 = t[smax + 2] / n3;
  for (int s = 0; s <= nx / 2; s++) r1i[s] = t[s] / (u1u1 * n3);
  *lint11 = t[smax + 1] / (u1u1 * n3);
  free(t);
  free(zr);
}

/* f_FluidstatsRsCollect normalisation (src/pf_fluidstats.f90:992-1116),
 * including its rs_rmsvort1 that is not divided by the number of points.
 * out[25]: rs_u1u1 rs_u2u2 rs_u3u3 rs_u1u2 rs_u1u3 rs_u2u3, rs_skew1..3,
 * rs_flat1..3, rs_rmsvort1..3, rs_skewvort1..3, rs_flatvort1..3,
 * omegamax1..3, ucfl */
void orc_rs_finish(int nx, int ny, int nz, const double* raw, double* out) {
  const double n3 = ((double)nx * (double)ny) * (double)nz;
  for (int q = 0; q < 6; q++) out[q] = raw[q] / n3;
  for (int c = 0; c < 3; c++) {
    out[6 + c] = (raw[6 + c] / n3) / pow(out[c], 1.5);
    out[9 + c] = (raw[9 + c] / n3) / (out[c] * out[c]);
  }
  double rms[3], sk[3], fl[3];
  rms[0] = raw[12];
  rms[1] = raw[13] / n3;
  rms[2] = raw[14] / n3;
  for (int c = 0; c < 3; c++) { sk[c] = raw[15 + c] / n3; fl[c] = raw[18 + c] / n3; }
  for (int c = 0; c < 3; c++) {
    rms[c] = pow(rms[c], 0.5);
    out[12 + c] = rms[c];
    out[15 + c] = sk[c] / (rms[c] * rms[c] * rms[c]);
    out[18 + c] = fl[c] / ((rms[c] * rms[c]) * (rms[c] * rms[c]));
  }
  out[21] = raw[22]; out[22] = raw[23]; out[23] = raw[24];
  out[24] = raw[21];
}

void orc_rs_stats(orc* o, double* out) {
  double raw[25];
  orc_rs_raw(o, raw);
  orc_rs_finish(o->nx, o->ny, o->nz, raw, out);
}

/* f_FluidstatsFs2D + f_FluidstatsFsCollect (src/pf_fluidstats.f90:1154-1283) on
 * the 1r1w arrays that f_RK3TransformjiFs2RsGetSpectrum fills (src/pf_rk3.f90:
 * 1504-1590): per z slice, the y-transformed velocity over (i, y) divided by
 * NODES_X, taken to the laboratory frame with shear; spec_1i(i) += 2 |u1|^2
 * (times binv(1,1)^3 with shear), y inner, z outer.  Overwrites the 1r2w arrays. */
void orc_spectra_1d(orc* o, double* s1, double* s2) {
  orc_prepare_realspace(o);
  const int ny = o->ny, nz = o->nz, nki = o->nki, nkj = o->nkj, KJ = o->kj;
  const int shear = (o->p.f_type != 1);
  const double b11 = T9(o->binv, 1, 1);
  const double b3 = (b11 * b11) * b11;
  const double nxd = (double)o->nx;
  double* in = (double*)malloc(sizeof(double) * 2 * (size_t)ny);
  double* out = (double*)malloc(sizeof(double) * 2 * (size_t)ny);
  double* pc = (double*)malloc(sizeof(double) * 3 * 2 * (size_t)ny * nki); /* [c][y][i] */
  for (int i = 0; i < nki; i++) { s1[i] = 0.0; s2[i] = 0.0; }
  for (int z = 0; z < nz; z++) {
    for (int c = 0; c < 3; c++)
      for (int i = 0; i < nki; i++) {
        for (int y = 0; y < 2 * ny; y++) in[y] = 0.0;
        const double* src = &o->w[c][(((long)i * nz + z) * nkj) * 2];
        for (int j = 0; j <= KJ; j++) { in[2 * j] = src[2 * j]; in[2 * j + 1] = src[2 * j + 1]; }
        for (int j = KJ + 1; j < nkj; j++) {
          int y = ny - nkj + j;
          in[2 * y] = src[2 * j]; in[2 * y + 1] = src[2 * j + 1];
        }
        fft_exec(&o->fy, in, out, +1);
        for (int y = 0; y < ny; y++) {
          double* d = &pc[2 * (((long)c * ny + y) * nki + i)];
          d[0] = out[2 * y] / nxd;
          d[1] = out[2 * y + 1] / nxd;
        }
      }
    for (int y = 0; y < ny; y++)
      for (int i = 0; i < nki; i++) {
        const double* a = &pc[2 * (((long)0 * ny + y) * nki + i)];
        const double* b = &pc[2 * (((long)1 * ny + y) * nki + i)];
        const double* c3 = &pc[2 * (((long)2 * ny + y) * nki + i)];
        double ur[3], ui[3];
        if (shear) {
          lab_velocity(o, a[0], b[0], c3[0], ur);
          lab_velocity(o, a[1], b[1], c3[1], ui);
        } else {
          ur[0] = a[0]; ur[1] = b[0]; ui[0] = a[1]; ui[1] = b[1];
        }
        double p11 = ur[0] * ur[0] + ui[0] * ui[0];
        double p22 = ur[1] * ur[1] + ui[1] * ui[1];
        if (shear) { p11 = p11 * b3; p22 = p22 * b3; }
        s1[i] = s1[i] + (2.0 * p11);
        s2[i] = s2[i] + (2.0 * p22);
      }
  }
  free(in); free(out); free(pc);
}

/* f_FluidstatsHomogeneous (src/pf_fluidstats.f90:2568-2936) for the current
 * state: out[15] = S*, K*, L*, ln_int1..3, ln_int_11..13, fn_ken1..3,
 * fn_int1..3.  ln_int_21, ln_int_12 and ln_int_13 are never assigned there
 * (zero), so L* = ln_int_11 / 0. */
void orc_homogeneous(orc* o, double* out) {
  double* spec = (double*)malloc(sizeof(double) * 7 * (size_t)o->kkmax);
  double q[33];
  orc_stats(o, spec, q);
  double* s1 = (double*)malloc(sizeof(double) * 2 * (size_t)o->nki);
  orc_spectra_1d(o, s1, s1 + o->nki);
  const double onepi = (double)3.14159265358979323846f;
  const double zero = 0.0;
  const double l11 = onepi * s1[0] / (2.0 * q[1]);
  out[0] = q[28] * o->p.f_deform_s;
  out[1] = 2.0 * q[1] / (q[2] + q[3]);
  out[2] = l11 / (2.0 * zero);
  out[3] = q[25]; out[4] = q[26]; out[5] = q[27];
  out[6] = l11; out[7] = 0.0; out[8] = 0.0;
  out[9] = q[11]; out[10] = q[12]; out[11] = q[13];
  out[12] = q[17]; out[13] = q[18]; out[14] = q[19];
  free(spec); free(s1);
}

/* ------------------------------------------------------------------------- */
/* z forward transform of u4..u6_1r2w for one i slice, x 1/(NX NY NZ),       */
/* truncated to the stored k modes.  src/pf_rk3.f90:2148-2316,              */
/* src/pf_fftw.f90:295-297, 841-879.  nl[c] is [j][k] complex.               */
/* ------------------------------------------------------------------------- */
static void kforward_slice(const orc* o, int i, double* in, double* out, double* nl[3]) {
  const int nz = o->nz, nkj = o->nkj, nkk = o->nkk, K = o->kk;
  const double norm = 1.0 / ((double)o->nx * (double)o->ny * (double)o->nz);
  for (int c = 0; c < 3; c++) {
    const double* src = o->w[3 + c];
    for (int j = 0; j < nkj; j++) {
      for (int z = 0; z < nz; z++) {
        long s = (((long)i * nz + z) * nkj + j) * 2;
        in[2 * z] = src[s]; in[2 * z + 1] = src[s + 1];
      }
      fft_exec(&o->fz, in, out, -1);
      for (int z = 0; z < 2 * nz; z++) out[z] = out[z] * norm;
      double* d = &nl[c][(long)j * nkk * 2];
      for (int k = 0; k <= K; k++) { d[2 * k] = out[2 * k]; d[2 * k + 1] = out[2 * k + 1]; }
      for (int k = K + 1; k < nkk; k++) {
        int z = nz - nkk + k;
        d[2 * k] = out[2 * z]; d[2 * k + 1] = out[2 * z + 1];
      }
    }
  }
}

void orc_nl_forward(orc* o, double* a, double* b, double* c) {
  double* dst[3] = {a, b, c};
  const long slab = (long)o->nkj * o->nkk;
#pragma omp parallel num_threads(nthr(o))
  {
    double* in = (double*)malloc(sizeof(double) * 2 * o->nz);
    double* out = (double*)malloc(sizeof(double) * 2 * o->nz);
#pragma omp for
    for (int i = 0; i < o->nki; i++) {
      double* nl[3] = {&dst[0][i * slab * 2], &dst[1][i * slab * 2], &dst[2][i * slab * 2]};
      kforward_slice(o, i, in, out, nl);
    }
    free(in); free(out);
  }
}

/* ------------------------------------------------------------------------- */
/* FormF terms for the BF_ROT shear form: src/pf_rk3.f90:2864-3209           */
/* ------------------------------------------------------------------------- */
static void bf_nonlinear(const orc* o, const double nl[3][2], double f[3][2]) {
  const double* B = o->bmat;
  const int* bv = o->bval;
  for (int i = 1; i <= 3; i++) {
    f[i - 1][0] = 0.0; f[i - 1][1] = 0.0;
    for (int c = 1; c <= 3; c++) {
      if (T9(bv, i, c)) {
        f[i - 1][0] = f[i - 1][0] + T9(B, i, c) * nl[c - 1][0];
        f[i - 1][1] = f[i - 1][1] + T9(B, i, c) * nl[c - 1][1];
      }
    }
  }
}

static void bf_viscous(const orc* o, const double u[3][2], const double* k, double v[3][2]) {
  const double* B = o->bmat;
  const int* bv = o->bval;
  const double nu = o->p.f_nu;
  for (int c = 0; c < 3; c++) v[c][0] = v[c][1] = 0.0;
  for (int l = 1; l <= 3; l++)
    for (int j = 1; j <= 3; j++)
      if (T9(bv, l, j))
        for (int n = 1; n <= 3; n++)
          if (T9(bv, n, j)) {
            double coef = nu * T9(B, l, j) * T9(B, n, j) * k[l - 1] * k[n - 1];
            for (int c = 0; c < 3; c++) {
              v[c][0] = v[c][0] - coef * u[c][0];
              v[c][1] = v[c][1] - coef * u[c][1];
            }
          }
}

static void bf_pressure(const orc* o, const double u[3][2], const double* k, const double nl[3][2],
                        double pf[3][2]) {
  const double* B = o->bmat;
  const double* BI = o->binv;
  const double* A = o->amat;
  const int* bv = o->bval;
  const int* av = o->aval;
  double p1[3] = {0, 0, 0}, p2[2] = {0, 0}, p3[2] = {0, 0}, p4 = 0.0;
  for (int m = 1; m <= 3; m++) {
    p2[0] = p2[0] + (k[m - 1] * nl[m - 1][0]);
    p2[1] = p2[1] + (k[m - 1] * nl[m - 1][1]);
    for (int r = 1; r <= 3; r++) {
      if (!T9(bv, m, r)) continue;
      for (int s = 1; s <= 3; s++) {
        if (!T9(av, r, s)) continue;
        for (int kk = 1; kk <= 3; kk++) {
          if (!T9(bv, s, kk)) continue;
          double cf = k[m - 1] * T9(B, m, r) * T9(A, r, s) * T9(BI, s, kk);
          p3[0] = p3[0] + (2.0 * (cf * u[kk - 1][0]));
          p3[1] = p3[1] + (2.0 * (cf * u[kk - 1][1]));
        }
      }
      for (int kk = 1; kk <= 3; kk++)
        if (T9(bv, kk, r)) p4 = p4 + (k[m - 1] * T9(B, m, r) * T9(B, kk, r) * k[kk - 1]);
    }
  }
  for (int i = 1; i <= 3; i++)
    for (int p = 1; p <= 3; p++)
      if (T9(bv, i, p))
        for (int j = 1; j <= 3; j++)
          if (T9(bv, j, p)) p1[i - 1] = p1[i - 1] + T9(B, i, p) * T9(B, j, p) * k[j - 1];
  if (p4 > 0.0) {
    for (int c = 0; c < 3; c++) {
      pf[c][0] = (p1[c] / p4) * (p2[0] + p3[0]);
      pf[c][1] = (p1[c] / p4) * (p2[1] + p3[1]);
    }
  } else {
    for (int c = 0; c < 3; c++) pf[c][0] = pf[c][1] = 0.0;
  }
}

static void bf_source(const orc* o, const double u[3][2], double h[3][2]) {
  const double* B = o->bmat;
  const double* BI = o->binv;
  const double* A = o->amat;
  const int* bv = o->bval;
  const int* av = o->aval;
  for (int c = 0; c < 3; c++) h[c][0] = h[c][1] = 0.0;
  for (int i = 1; i <= 3; i++)
    for (int r = 1; r <= 3; r++)
      if (T9(bv, i, r))
        for (int j = 1; j <= 3; j++)
          if (T9(av, r, j))
            for (int c = 1; c <= 3; c++)
              if (T9(bv, j, c)) {
                double cf = T9(B, i, r) * T9(A, r, j) * T9(BI, j, c);
                h[i - 1][0] = h[i - 1][0] + (2.0 * (cf * u[c - 1][0]));
                h[i - 1][1] = h[i - 1][1] + (2.0 * (cf * u[c - 1][1]));
              }
}

/* ------------------------------------------------------------------------- */
/* FormF terms for the R_ROT shear form: src/pf_rk3.f90:2945-3249            */
/* ------------------------------------------------------------------------- */
/* f_RK3RRotViscous (2945-2979): -nu bmat(l,j) bmat(n,j) k_l k_n u, term by term */
static void rr_viscous(const orc* o, const double u[3][2], const double* k, double v[3][2]) {
  const double* B = o->bmat;
  const int* bv = o->bval;
  const double nu = o->p.f_nu;
  for (int c = 0; c < 3; c++) v[c][0] = v[c][1] = 0.0;
  for (int l = 1; l <= 3; l++)
    for (int j = 1; j <= 3; j++)
      if (T9(bv, l, j))
        for (int n = 1; n <= 3; n++)
          if (T9(bv, n, j))
            for (int c = 0; c < 3; c++)
              for (int r = 0; r < 2; r++)
                v[c][r] = v[c][r] - nu * T9(B, l, j) * T9(B, n, j) * k[l - 1] * k[n - 1] * u[c][r];
}

/* f_RK3RRotPressure (3078-3156): (p1_c / p4) (p2 + p3) with p1 = bmat^T k,
 * p2 = k . bmat nl, p3 = k . bmat amat u (no factor two), p4 = |bmat^T k|^2 */
static void rr_pressure(const orc* o, const double u[3][2], const double* k, const double nl[3][2],
                        double pf[3][2]) {
  const double* B = o->bmat;
  const double* A = o->amat;
  const int* bv = o->bval;
  const int* av = o->aval;
  double p1[3] = {0, 0, 0}, p2[2] = {0, 0}, p3[2] = {0, 0}, p4 = 0.0;
  for (int m = 1; m <= 3; m++)
    for (int r = 1; r <= 3; r++) {
      if (!T9(bv, m, r)) continue;
      for (int q = 0; q < 2; q++) p2[q] = p2[q] + (k[m - 1] * T9(B, m, r) * nl[r - 1][q]);
      for (int kk = 1; kk <= 3; kk++)
        if (T9(av, r, kk))
          for (int q = 0; q < 2; q++) {
            const double t = k[m - 1] * T9(B, m, r) * T9(A, r, kk) * u[kk - 1][q];
            p3[q] = p3[q] + (o->p.rrot_p3_two ? 2.0 * t : t);
          }
      for (int kk = 1; kk <= 3; kk++)
        if (T9(bv, kk, r)) p4 = p4 + (k[m - 1] * T9(B, m, r) * k[kk - 1] * T9(B, kk, r));
    }
  for (int i = 1; i <= 3; i++)
    for (int p = 1; p <= 3; p++)
      if (T9(bv, p, i)) p1[i - 1] = p1[i - 1] + k[p - 1] * T9(B, p, i);
  for (int c = 0; c < 3; c++)
    for (int q = 0; q < 2; q++) pf[c][q] = (p4 > 0.0) ? (p1[c] / p4) * (p2[q] + p3[q]) : 0.0;
}

/* f_RK3RRotSource (3218-3249): amat u over the existing entries */
static void rr_source(const orc* o, const double u[3][2], double h[3][2]) {
  for (int i = 1; i <= 3; i++)
    for (int q = 0; q < 2; q++) {
      h[i - 1][q] = 0.0;
      for (int j = 1; j <= 3; j++)
        if (T9(o->aval, i, j)) h[i - 1][q] = h[i - 1][q] + (T9(o->amat, i, j) * u[j - 1][q]);
    }
}

/* f_RK3ContinuityFixRRot (src/pf_rk3.f90:3620-3693): map to the moving frame
 * with bmat, remove k (k . u) / |k|^2, map back with binv */
static void continuity_rrot(const orc* o, double* u[3], long e, const double* kv, double kmag2) {
  double um[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int m = 1; m <= 3; m++)
    for (int n = 1; n <= 3; n++)
      if (T9(o->bval, m, n))
        for (int q = 0; q < 2; q++) um[m - 1][q] = um[m - 1][q] + (T9(o->bmat, m, n) * u[n - 1][e + q]);
  if (!(kmag2 > 0.0)) return;
  for (int q = 0; q < 2; q++) {
    const double kd = kv[0] * um[0][q] + kv[1] * um[1][q] + kv[2] * um[2][q];
    for (int c = 0; c < 3; c++) um[c][q] = um[c][q] - (kv[c] * kd / kmag2);
  }
  for (int m = 1; m <= 3; m++)
    for (int q = 0; q < 2; q++) {
      double a = 0.0;
      for (int n = 1; n <= 3; n++)
        if (T9(o->bval, m, n)) a = a + (T9(o->binv, m, n) * um[n - 1][q]);
      u[m - 1][e + q] = a;
    }
}

/* f_RK3RungeKutta (src/pf_rk3.f90:3457-3498) */
static inline void rk_only(double* rk[3], double* u[3], long e, const double tot[3][2], double beta,
                           double gdt) {
  for (int c = 0; c < 3; c++) {
    rk[c][e] = (beta * rk[c][e]) + tot[c][0];
    rk[c][e + 1] = (beta * rk[c][e + 1]) + tot[c][1];
  }
  for (int c = 0; c < 3; c++) {
    u[c][e] = u[c][e] + (gdt * rk[c][e]);
    u[c][e + 1] = u[c][e + 1] + (gdt * rk[c][e + 1]);
  }
}

/* f_RK3RungeKutta + f_RK3ContinuityFixIsotropic (3504-3531) */
static inline void rk_update(double* rk[3], double* u[3], long e, const double tot[3][2], double beta,
                             double gdt, const double* kv, double kmag2) {
  rk_only(rk, u, e, tot, beta, gdt);
  double kd0 = kv[0] * u[0][e] + kv[1] * u[1][e] + kv[2] * u[2][e];
  double kd1 = kv[0] * u[0][e + 1] + kv[1] * u[1][e + 1] + kv[2] * u[2][e + 1];
  for (int c = 0; c < 3; c++) {
    u[c][e] = u[c][e] - (kv[c] * kd0 / kmag2);
    u[c][e + 1] = u[c][e + 1] - (kv[c] * kd1 / kmag2);
  }
}

/* f_RK3BFRotForcing (src/pf_rk3.f90:3256-3349), the forcing of the deformed
 * isotropic form: a mode (i, j, k) whose indices B110, B220, B330 divide takes the
 * forcing arrays at (i/B110, j/B220, k/B330) if its laboratory wavenumber
 * wvnlab(m) = bmat(m,m) k_m has 0 < |wvnlab| <= FORCE_KF, made solenoidal along
 * wvnlab. */
static void bf_forcing(const orc* o, int i, int j, int k, const double* kv, double f[3][2]) {
  for (int c = 0; c < 3; c++) f[c][0] = f[c][1] = 0.0;
  if (!o->forcing) return;
  const int b1 = (int)lround(T9(o->binv, 1, 1)), b2 = (int)lround(T9(o->binv, 2, 2)), b3 = (int)lround(T9(o->binv, 3, 3));
  if (i % b1 != 0 || j % b2 != 0 || k % b3 != 0) return;
  double wl[3] = {0.0, 0.0, 0.0};
  for (int m = 0; m < 3; m++) wl[m] = wl[m] + T9(o->bmat, m + 1, m + 1) * kv[m];
  const double kmag2 = ((wl[0] * wl[0]) + (wl[1] * wl[1])) + (wl[2] * wl[2]);
  if (!(sqrt(kmag2) <= o->p.force_kf) || !(sqrt(kmag2) > 0.0)) return;
  const int il = i / b1, jl = j / b2, kl = k / b3, nk = o->force_nk;
  if (il > nk || jl < -nk || jl > nk || kl < -nk || kl > nk) return;
  const double* ff[3] = {fat(o, 0, il, jl, kl), fat(o, 1, il, jl, kl), fat(o, 2, il, jl, kl)};
  for (int r = 0; r < 2; r++) {
    const double kb = (wl[0] * ff[0][r] + wl[1] * ff[1][r]) + wl[2] * ff[2][r];
    for (int c = 0; c < 3; c++) f[c][r] = ff[c][r] - ((wl[c] / kmag2) * kb);
  }
}

/* f_RK3SolveNavierStokes, src/pf_rk3.f90:1996-2142 with f_RK3FormFIsotropic
 * (2323-2487) and f_RK3FormFBFRot (2494-2673).  The isotropic form adds the
 * OU forcing, the deformed isotropic one (F_TYPE 5) f_RK3BFRotForcing's; the
 * shear forms carry none; the two-way coupling term is zero. */
static void solve_ns(orc* o, double dt, int stage) {
  double beta = 0.0, gamma = THIRD;
  if (stage == 2) { beta = -FIVE_OVER_NINE; gamma = FIFTEEN_OVER_SIXTEEN; }
  if (stage == 3) { beta = -ONEFIVETHREE_OVER_ONETWOEIGHT; gamma = EIGHT_OVER_FIFTEEN; }
  const double gdt = gamma * dt;
  const int nkj = o->nkj, nkk = o->nkk;
  const long slab = (long)nkj * nkk;
#pragma omp parallel num_threads(nthr(o))
  {
    double* in = (double*)malloc(sizeof(double) * 2 * o->nz);
    double* out = (double*)malloc(sizeof(double) * 2 * o->nz);
    double* nlbuf = (double*)malloc(sizeof(double) * 2 * 3 * slab);
    double* nl[3] = {nlbuf, nlbuf + 2 * slab, nlbuf + 4 * slab};
#pragma omp for schedule(dynamic)
    for (int i = 0; i < o->nki; i++) {
      kforward_slice(o, i, in, out, nl);
      const double wi = wvn(i, o->ki);
      for (int j = 0; j < nkj; j++) {
        const double wj = wvn(j, o->kj);
        for (int k = 0; k < nkk; k++) {
          const double wk = wvn(k, o->kk);
          const double kmag2 = (wi * wi) + (wj * wj) + (wk * wk);
          if (!(kmag2 > 0.0)) continue;
          const long e = (((long)i * nkj + j) * nkk + k) * 2;
          const long q = ((long)j * nkk + k) * 2;
          const double kv[3] = {wi, wj, wk};
          double tot[3][2];
          double uu[3][2], nn[3][2];
          for (int c = 0; c < 3; c++) {
            uu[c][0] = o->u[c][e]; uu[c][1] = o->u[c][e + 1];
            nn[c][0] = nl[c][q]; nn[c][1] = nl[c][q + 1];
          }
          if (o->p.f_type == 1) {
            double kdn0 = (wi * nn[0][0]) + (wj * nn[1][0]) + (wk * nn[2][0]);
            double kdn1 = (wi * nn[0][1]) + (wj * nn[1][1]) + (wk * nn[2][1]);
            /* forcing on 0 < |k| <= FORCE_KF (src/pf_rk3.f90:2421-2467) */
            double fo[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            const int nk = o->force_nk;
            const int jj = (int)wj, kq = (int)wk;
            if (o->forcing && i <= nk && jj >= -nk && jj <= nk && kq >= -nk && kq <= nk &&
                sqrt(kmag2) <= o->p.force_kf) {
              const double* f[3] = {fat(o, 0, i, jj, kq), fat(o, 1, i, jj, kq), fat(o, 2, i, jj, kq)};
              const double kb0 = (wi * f[0][0] + wj * f[1][0]) + wk * f[2][0];
              const double kb1 = (wi * f[0][1] + wj * f[1][1]) + wk * f[2][1];
              for (int c = 0; c < 3; c++) {
                fo[c][0] = f[c][0] - ((kv[c] / kmag2) * kb0);
                fo[c][1] = f[c][1] - ((kv[c] / kmag2) * kb1);
              }
            }
            for (int c = 0; c < 3; c++) {
              double visc0 = -(o->p.f_nu * kmag2 * uu[c][0]);
              double visc1 = -(o->p.f_nu * kmag2 * uu[c][1]);
              double pr0 = (kv[c] / kmag2) * kdn0;
              double pr1 = (kv[c] / kmag2) * kdn1;
              tot[c][0] = (((visc0 - nn[c][0]) + pr0) + fo[c][0]) + 0.0;
              tot[c][1] = (((visc1 - nn[c][1]) + pr1) + fo[c][1]) + 0.0;
            }
          } else if (rogallo(o)) {
            /* f_RK3FormFRRot (src/pf_rk3.f90:2680-2856); the nonlinear term enters
             * only for F_SHEAR_DNS, unmapped */
            double nlf[3][2], pf[3][2], hom[3][2], visc[3][2];
            for (int c = 0; c < 3; c++)
              for (int r = 0; r < 2; r++) nlf[c][r] = (o->p.f_type == 2 || o->p.f_type == 5) ? nn[c][r] : 0.0;
            rr_pressure(o, uu, kv, nlf, pf);
            rr_source(o, uu, hom);
            if (o->p.f_type == 2 || o->p.f_type == 3 || o->p.f_type == 5) rr_viscous(o, uu, kv, visc);
            /* f_RK3RRotForcing is f_RK3BFRotForcing's twin (the deformed branch, 2820-2828) */
            double fo[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            if (o->p.f_type == 5) bf_forcing(o, i, (int)wj, (int)wk, kv, fo);
            for (int c = 0; c < 3; c++)
              for (int r = 0; r < 2; r++) {
                if (o->p.f_type == 2) tot[c][r] = (((visc[c][r] - nlf[c][r]) + pf[c][r]) + 0.0) - hom[c][r];
                else if (o->p.f_type == 3) tot[c][r] = ((visc[c][r] + pf[c][r]) + 0.0) - hom[c][r];
                else if (o->p.f_type == 5) tot[c][r] = (((visc[c][r] - nlf[c][r]) + pf[c][r]) + 0.0) + fo[c][r];
                else tot[c][r] = pf[c][r] - hom[c][r];
              }
            rk_only(o->rk, o->u, e, tot, beta, gdt);
            continuity_rrot(o, o->u, e, kv, kmag2);
            continue;
          } else {
            double nlf[3][2], pf[3][2], hom[3][2], visc[3][2];
            if (o->p.f_type == 2 || o->p.f_type == 5) bf_nonlinear(o, nn, nlf);
            else for (int c = 0; c < 3; c++) nlf[c][0] = nlf[c][1] = 0.0;
            bf_pressure(o, uu, kv, nlf, pf);
            bf_source(o, uu, hom);
            if (o->p.f_type == 2 || o->p.f_type == 3 || o->p.f_type == 5) bf_viscous(o, uu, kv, visc);
            double fo[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            if (o->p.f_type == 5) bf_forcing(o, i, (int)wj, (int)wk, kv, fo);
            for (int c = 0; c < 3; c++)
              for (int r = 0; r < 2; r++) {
                if (o->p.f_type == 2) tot[c][r] = (((visc[c][r] - nlf[c][r]) + pf[c][r]) + 0.0) - hom[c][r];
                else if (o->p.f_type == 3) tot[c][r] = ((visc[c][r] + pf[c][r]) + 0.0) - hom[c][r];
                else if (o->p.f_type == 5) tot[c][r] = (((visc[c][r] - nlf[c][r]) + pf[c][r]) + 0.0) + fo[c][r];
                else tot[c][r] = pf[c][r] - hom[c][r];
              }
          }
          rk_update(o->rk, o->u, e, tot, beta, gdt, kv, kmag2);
        }
      }
    }
    free(in); free(out); free(nlbuf);
  }
}

/* ------------------------------------------------------------------------- */
/* Remesh: src/pf_rk3.f90:3958-4178, src/pp_mpi.f90:1062-1092                */
/* ------------------------------------------------------------------------- */
int orc_remesh_target(int kmax, int i, int k) {
  const int nodes = 2 * kmax + 1;
  int actualk = (k > kmax) ? k - nodes : k;
  int knew = actualk - i;
  if (knew >= -kmax) {
    if (knew < 0) knew += nodes;
    return knew;
  }
  return -1;
}

static void wrap_one(const orc* o, double* x);

int orc_remesh(orc* o, int tstep) {
  if (o->p.f_type == 1 || o->p.rk3_timestep != 1) return 0;
  const int b110 = o->p.b110 > 0 ? o->p.b110 : 1;
  const int a = tstep - (b110 * o->p.f_remesh_t) - 1;
  const int b = 2 * b110 * o->p.f_remesh_t;
  if ((a % b) != 0) return 0; /* Fortran mod truncates like C % */
  /* particles first: p_ControlRemeshInitial (x0_1 += x0_3, shifted as x1 + x3
   * wraps, src/pp_mpi.f90:1095-1129), then x1 += x3, wrap x1 */
  for (long q = 0; q < o->np && o->xp0; q++) {
    const double* x = &o->xp[3 * q];
    double* x0 = &o->xp0[3 * q];
    const double x1 = x[0] + x[2];
    const double x1min = -0.5 * o->dx, x1max = x1min + o->lx;
    x0[0] = x0[0] + x0[2];
    if (x1 < x1min) x0[0] = x0[0] + o->lx;
    if (x1 >= x1max) x0[0] = x0[0] - o->lx;
  }
  for (long q = 0; q < o->np; q++) {
    double* x = &o->xp[3 * q];
    x[0] = x[0] + x[2];
    const double x1min = -0.5 * o->dx, x1max = x1min + o->lx;
    if (x[0] < x1min) x[0] = x[0] + o->lx;
    if (x[0] >= x1max) x[0] = x[0] - o->lx;
  }
  T9(o->binv, 1, 3) = ((double)b110 * 0.5) / T9(o->bmat, 3, 3);
  T9(o->bmat, 1, 3) = -T9(o->bmat, 1, 1) * (-((double)b110 * 0.5));
  const int nkk = o->nkk, nkj = o->nkj, K = o->kk;
  double* buf = (double*)malloc(sizeof(double) * 6 * nkk);
  for (int i = 0; i < o->nki; i++)
    for (int j = 0; j < nkj; j++) {
      for (int q = 0; q < 6 * nkk; q++) buf[q] = 0.0;
      const long base = (((long)i * nkj + j) * nkk) * 2;
      for (int k = 0; k < nkk; k++) {
        int kn = orc_remesh_target(K, i, k);
        if (kn < 0) continue;
        for (int c = 0; c < 3; c++) {
          buf[2 * (c * nkk + kn)] = o->u[c][base + 2 * k];
          buf[2 * (c * nkk + kn) + 1] = o->u[c][base + 2 * k + 1];
        }
      }
      for (int k = 0; k < nkk; k++)
        for (int r = 0; r < 2; r++) {
          double b1 = buf[2 * k + r], b2 = buf[2 * (nkk + k) + r], b3 = buf[2 * (2 * nkk + k) + r];
          o->u[0][base + 2 * k + r] = (o->p.ns_deformed == 1) ? b1 + b3 : b1;
          o->u[1][base + 2 * k + r] = b2;
          o->u[2][base + 2 * k + r] = b3;
        }
    }
  free(buf);
  T9(o->binv, 1, 3) = (-((double)b110 * 0.5)) / T9(o->bmat, 3, 3);
  return 1;
}

/* ------------------------------------------------------------------------- */
/* Tracers                                                                    */
/* ------------------------------------------------------------------------- */
void orc_set_particles(orc* o, const double* xp, int64_t np) {
  free(o->xp); free(o->up); free(o->uprk); free(o->xp0);
  o->np = np;
  o->xp = (double*)malloc(sizeof(double) * 3 * (size_t)(np > 0 ? np : 1));
  o->up = (double*)calloc(3 * (size_t)(np > 0 ? np : 1), sizeof(double));
  o->uprk = (double*)calloc(3 * (size_t)(np > 0 ? np : 1), sizeof(double));
  memcpy(o->xp, xp, sizeof(double) * 3 * (size_t)np);
  o->xp0 = NULL;
  if (o->keep_x0) {  /* particle_xp0 = particle_xp at placement, src/pp_control.f90:405-407 */
    o->xp0 = (double*)malloc(sizeof(double) * 3 * (size_t)(np > 0 ? np : 1));
    memcpy(o->xp0, xp, sizeof(double) * 3 * (size_t)np);
  }
}

void orc_get_particles_aux(const orc* o, double* up, double* xp0) {
  if (up) memcpy(up, o->up, sizeof(double) * 3 * (size_t)o->np);
  if (xp0 && o->xp0) memcpy(xp0, o->xp0, sizeof(double) * 3 * (size_t)o->np);
}

void orc_set_particles_aux(orc* o, const double* up, const double* xp0) {
  if (up) memcpy(o->up, up, sizeof(double) * 3 * (size_t)o->np);
  if (xp0 && o->xp0) memcpy(o->xp0, xp0, sizeof(double) * 3 * (size_t)o->np);
}

void orc_particle_stats_config(orc* o, int enable, int keep_initial, int ts_release) {
  o->pst_on = enable;
  o->ts_release = ts_release;
  if (keep_initial && !o->xp0 && o->np > 0) {
    o->xp0 = (double*)malloc(sizeof(double) * 3 * (size_t)o->np);
    memcpy(o->xp0, o->xp, sizeof(double) * 3 * (size_t)o->np);
  }
  if (!keep_initial) { free(o->xp0); o->xp0 = NULL; }
  o->keep_x0 = keep_initial;
  for (int c = 0; c < 3; c++) o->xrel_prev[c] = 0.0;
  for (int a = 0; a < 30; a++) o->pst[a] = 0.0;
  o->pst_tstep = -1;
}

int orc_particle_stats(const orc* o, double* rec) {
  if (rec) memcpy(rec, o->pst, sizeof(o->pst));
  return o->pst_tstep;
}

void orc_get_particles(const orc* o, double* xp, double* uprk) {
  if (xp) memcpy(xp, o->xp, sizeof(double) * 3 * (size_t)o->np);
  if (uprk) memcpy(uprk, o->uprk, sizeof(double) * 3 * (size_t)o->np);
}

void orc_get_tracer_field(const orc* o, double* a, double* b, double* c) {
  size_t n = sizeof(double) * (size_t)o->nx * o->ny * o->nz;
  if (o->ur[0]) { memcpy(a, o->ur[0], n); memcpy(b, o->ur[1], n); memcpy(c, o->ur[2], n); }
}

static inline long pmod(long a, long n) { long r = a % n; return r < 0 ? r + n : r; }

/* p_InterpFindMolecule + p_InterpNeville3D, src/pp_interp.f90:406-560, with
 * LX_HALO from g_DomainInitHalo (src/pg_domain.f90:360-384). */
static void interp_one(const orc* o, const double* f[3], const double* xp, double* uf) {
  const int P = o->p.p_interp_order > 0 ? o->p.p_interp_order : 4;
  const int nhalo = P / 2;
  double lxh = (double)nhalo * o->dx, lyh = (double)nhalo * o->dy, lzh = (double)nhalo * o->dz;
  double xd1, xd2, xd3;
  if (P % 2 == 0) {
    lxh = lxh - (0.5 * o->dx); lyh = lyh - (0.5 * o->dy); lzh = lzh - (0.5 * o->dz);
    xd1 = xp[0] + lxh + o->dx + (0.5 * o->dx);
    xd2 = xp[1] + lyh + o->dy + (0.5 * o->dy);
    xd3 = xp[2] + lzh + o->dz + (0.5 * o->dz);
  } else {
    xd1 = xp[0] + lxh + (0.5 * o->dx);
    xd2 = xp[1] + lyh + (0.5 * o->dy);
    xd3 = xp[2] + lzh + (0.5 * o->dz);
  }
  const int i0 = (int)(xd1 / o->dx) - nhalo - nhalo;
  const int j0 = (int)(xd2 / o->dy) - nhalo - nhalo;
  const int k0 = (int)(xd3 / o->dz) - nhalo - nhalo;
  double ix[16], iy[16], iz[16];
  for (int m = 0; m < P; m++) {
    ix[m] = (double)(i0 + m) * o->dx;
    iy[m] = (double)(j0 + m) * o->dy;
    iz[m] = (double)(k0 + m) * o->dz;
  }
  /* v[c][x][z][y] as interpolationvalues(c, x, z, y) */
  double v[3][8][8][8];
  for (int a = 0; a < P; a++)
    for (int c2 = 0; c2 < P; c2++)
      for (int b = 0; b < P; b++) {
        long gx = pmod(i0 + a, o->nx), gz = pmod(k0 + c2, o->nz), gy = pmod(j0 + b, o->ny);
        long idx = (gz * o->ny + gy) * o->nx + gx;
        for (int c = 0; c < 3; c++) v[c][a][c2][b] = f[c][idx];
      }
  for (int col = 1; col <= P - 1; col++)
    for (int nv = 1; nv <= P - col; nv++) {
      double d1 = xp[1] - iy[nv + col - 1];
      double d2 = iy[nv - 1] - xp[1];
      double d3 = iy[nv - 1] - iy[nv + col - 1];
      for (int zi = 0; zi < P; zi++)
        for (int xi = 0; xi < P; xi++)
          for (int c = 0; c < 3; c++)
            v[c][xi][zi][nv - 1] = ((d1 * v[c][xi][zi][nv - 1]) + (d2 * v[c][xi][zi][nv])) / d3;
    }
  for (int col = 1; col <= P - 1; col++)
    for (int nv = 1; nv <= P - col; nv++) {
      double d1 = xp[2] - iz[nv + col - 1];
      double d2 = iz[nv - 1] - xp[2];
      double d3 = iz[nv - 1] - iz[nv + col - 1];
      for (int xi = 0; xi < P; xi++)
        for (int c = 0; c < 3; c++)
          v[c][xi][nv - 1][0] = ((d1 * v[c][xi][nv - 1][0]) + (d2 * v[c][xi][nv][0])) / d3;
    }
  for (int col = 1; col <= P - 1; col++)
    for (int nv = 1; nv <= P - col; nv++) {
      double d1 = xp[0] - ix[nv + col - 1];
      double d2 = ix[nv - 1] - xp[0];
      double d3 = ix[nv - 1] - ix[nv + col - 1];
      for (int c = 0; c < 3; c++)
        v[c][nv - 1][0][0] = ((d1 * v[c][nv - 1][0][0]) + (d2 * v[c][nv][0][0])) / d3;
    }
  for (int c = 0; c < 3; c++) uf[c] = v[c][0][0][0];
}

void orc_interpolate(const orc* o, const double* f1, const double* f2, const double* f3,
                     const double* xp, int64_t np, double* uf) {
  const double* f[3] = {f1, f2, f3};
#pragma omp parallel for num_threads(nthr(o))
  for (long q = 0; q < np; q++) interp_one(o, f, &xp[3 * q], &uf[3 * q]);
}

/* p_MPICorrectPosition, src/pp_mpi.f90:1011-1057 (x3 upper test is '>') */
static void wrap_one(const orc* o, double* x) {
  const double x1min = -0.5 * o->dx, x1max = x1min + o->lx;
  const double x2min = -0.5 * o->dy, x2max = x2min + o->ly;
  const double x3min = -0.5 * o->dz, x3max = x3min + o->lz;
  if (x[0] < x1min) x[0] = x[0] + o->lx;
  if (x[0] >= x1max) x[0] = x[0] - o->lx;
  if (x[1] < x2min) x[1] = x[1] + o->ly;
  if (x[1] >= x2max) x[1] = x[1] - o->ly;
  if (x[2] < x3min) x[2] = x[2] + o->lz;
  if (x[2] > x3max) x[2] = x[2] - o->lz;
}

void orc_correct_position(const orc* o, double* xp, int64_t np) {
  for (long q = 0; q < np; q++) wrap_one(o, &xp[3 * q]);
}

/* p_RK3SolveParticle / p_RK3SolveParticleBF tracer branch (dp = 0),
 * src/pp_rk3.f90:57-165, 169-303, 989-1043; then p_ControlFinishTimeStep wrap. */
static void tracers_stage(orc* o, int stage, double dt) {
  double beta = 0.0, gamma = THIRD;
  if (stage == 2) { beta = -FIVE_OVER_NINE; gamma = FIFTEEN_OVER_SIXTEEN; }
  if (stage == 3) { beta = -ONEFIVETHREE_OVER_ONETWOEIGHT; gamma = EIGHT_OVER_FIFTEEN; }
  const double* f[3] = {o->ur[0], o->ur[1], o->ur[2]};
  const int bf = (o->p.f_type != 1) && !rogallo(o);  /* R forms: p_RK3SolveParticleR (307-403) */
#pragma omp parallel for num_threads(nthr(o))
  for (long q = 0; q < o->np; q++) {
    double* x = &o->xp[3 * q];
    double* up = &o->up[3 * q];
    double* ur = &o->uprk[3 * q];
    double uf[3];
    interp_one(o, f, x, uf);
    if (!bf) {
      for (int c = 0; c < 3; c++) up[c] = uf[c];
      for (int c = 0; c < 3; c++) ur[c] = (beta * ur[c]) + up[c];
      for (int c = 0; c < 3; c++) x[c] = x[c] + (gamma * dt * ur[c]);
    } else {
      double xl[3] = {0, 0, 0}, ul[3] = {0, 0, 0}, xm[3] = {0, 0, 0};
      for (int m = 1; m <= 3; m++)
        for (int n = 1; n <= 3; n++)
          if (T9(o->bval, m, n)) {
            xl[m - 1] = xl[m - 1] + T9(o->binv, m, n) * x[n - 1];
            ul[m - 1] = ul[m - 1] + T9(o->binv, m, n) * uf[n - 1];
          }
      for (int c = 0; c < 3; c++) up[c] = ul[c];
      for (int c = 0; c < 3; c++) ur[c] = (beta * ur[c]) + up[c];
      for (int c = 0; c < 3; c++) xl[c] = xl[c] + (gamma * dt * ur[c]);
      for (int m = 1; m <= 3; m++)
        for (int n = 1; n <= 3; n++)
          if (T9(o->bval, m, n)) xm[m - 1] = xm[m - 1] + T9(o->bmat, m, n) * xl[n - 1];
      for (int c = 0; c < 3; c++) x[c] = xm[c];
    }
    const double xu[3] = {x[0], x[1], x[2]};
    wrap_one(o, x);
    /* p_ControlUpdateInitial (src/pp_control.f90:1574-1585): xp0 + xp_corrected - xp,
     * evaluated left to right on every call */
    if (o->xp0)
      for (int c = 0; c < 3; c++) o->xp0[3 * q + c] = (o->xp0[3 * q + c] + x[c]) - xu[c];
  }
}

/* p_StatsPrimary + p_StatsDispersion (src/pp_stats.f90:119-233, 303-370) for
 * tracers (dp = 0: particle kinetic energy, response time, Stokes numbers are
 * zero), with the mean velocity of p_RK3AbsoluteVelocity[BF] (src/pp_rk3.f90:
 * 602-710) that p_StatsRelativeVelocity reports as vt.  Sequential sums in
 * particle order, as on one rank. */
static void pstats_primary(orc* o, double dt) {
  double s[6] = {0, 0, 0, 0, 0, 0}, va[3] = {0, 0, 0};
  double lx[3] = {0, 0, 0}, lx2[3] = {0, 0, 0}, l12 = 0, l13 = 0, l23 = 0;
  for (long q = 0; q < o->np; q++) {
    const double* u = &o->up[3 * q];
    s[0] = s[0] + u[0] * u[0];
    s[1] = s[1] + u[1] * u[1];
    s[2] = s[2] + u[2] * u[2];
    s[3] = s[3] + u[0] * u[1];
    s[4] = s[4] + u[0] * u[2];
    s[5] = s[5] + u[1] * u[2];
    if (rogallo(o)) {
      /* p_RK3AbsoluteVelocityR (src/pp_rk3.f90:720-773): binv particle_up */
      for (int m = 1; m <= 3; m++) {
        double a = 0.0;
        for (int n = 1; n <= 3; n++)
          if (T9(o->bval, m, n)) a = a + T9(o->binv, m, n) * u[n - 1];
        va[m - 1] = va[m - 1] + a;
      }
    } else {
      for (int c = 0; c < 3; c++) va[c] = va[c] + u[c];
    }
    if (o->xp0) {
      double xm[3], xl[3] = {0, 0, 0};
      for (int c = 0; c < 3; c++) xm[c] = o->xp[3 * q + c] - o->xp0[3 * q + c];
      if (o->p.f_type == 1) {
        for (int c = 0; c < 3; c++) xl[c] = xm[c];
      } else {
        for (int m = 1; m <= 3; m++)
          for (int n = 1; n <= 3; n++)
            if (T9(o->bval, m, n)) xl[m - 1] = xl[m - 1] + T9(o->binv, m, n) * xm[n - 1];
      }
      for (int c = 0; c < 3; c++) { lx[c] = lx[c] + xl[c]; lx2[c] = lx2[c] + (xl[c] * xl[c]); }
      l12 = l12 + (xl[0] * xl[1]);
      l13 = l13 + (xl[0] * xl[2]);
      l23 = l23 + (xl[1] * xl[2]);
    }
  }
  const double npg = (double)o->np;
  double* r = o->pst;
  for (int a = 0; a < 30; a++) r[a] = 0.0;
  for (int a = 0; a < 6; a++) r[a] = s[a] / npg;
  r[6] = 0.5 * (r[0] + r[1] + r[2]);
  r[7] = 0.0;                                         /* pken: dp = 0 */
  for (int c = 0; c < 3; c++) r[21 + c] = va[c] / npg; /* vt = vabs */
  if (o->xp0) {
    double xrel[3];
    for (int c = 0; c < 3; c++) xrel[c] = lx2[c] / npg;
    r[18] = l12 / npg;
    r[19] = l13 / npg;
    r[20] = l23 / npg;
    for (int c = 0; c < 3; c++) {
      r[12 + c] = xrel[c];
      r[15 + c] = 0.5 * (xrel[c] - o->xrel_prev[c]) / dt;
      o->xrel_prev[c] = xrel[c];
    }
  }
}

/* p_RK3RelativeVelocity for tracers (uf = 0, src/pp_rk3.f90:407-464) summed over
 * the stage-1 velocities, and p_StatsRelativeVelocity (src/pp_stats.f90:240-296). */
static void pstats_relvel(orc* o, int tstep) {
  double ab = 0, v[3] = {0, 0, 0};
  for (long q = 0; q < o->np; q++) {
    const double* u = &o->up[3 * q];
    const double a = (u[0] * u[0]) + (u[1] * u[1]) + (u[2] * u[2]);
    ab = ab + a;
    for (int c = 0; c < 3; c++) v[c] = v[c] + u[c] * u[c];
  }
  const double npg = (double)o->np;
  for (int c = 0; c < 3; c++) o->pst[8 + c] = v[c] / npg;
  o->pst[11] = sqrt(ab / npg);
  o->pst_tstep = tstep;
}

/* f_ArraysDealiasSpherical / f_ArraysLESCutoff (src/pf_arrays.f90:1865-1987):
 * zero u1..u3 where |k_lab| > cutoff, k_lab(n) = sum_m bmat(m,n) k(m) over the
 * existing entries, cutoff = KMAX = min(KI, KJ, KK) (src/pg_domain.f90:231) or
 * LES_CUTOFF * KMAX */
void orc_dealias(orc* o) {
  if (o->p.dealiasing != 2 && o->p.dealiasing != 3) return;
  const int km_i = o->ki < o->kj ? (o->ki < o->kk ? o->ki : o->kk) : (o->kj < o->kk ? o->kj : o->kk);
  const double kmax = (double)km_i;
  const double cutoff = (o->p.dealiasing == 3) ? o->p.les_cutoff * kmax : kmax;
  for (int i = 0; i < o->nki; i++)
    for (int j = 0; j < o->nkj; j++)
      for (int k = 0; k < o->nkk; k++) {
        const double kv[3] = {wvn(i, o->ki), wvn(j, o->kj), wvn(k, o->kk)};
        double kl[3] = {0.0, 0.0, 0.0};
        for (int m = 1; m <= 3; m++)
          for (int n = 1; n <= 3; n++)
            if (T9(o->bval, m, n)) kl[n - 1] = kl[n - 1] + T9(o->bmat, m, n) * kv[m - 1];
        const double kmag = sqrt((kl[0] * kl[0]) + (kl[1] * kl[1]) + (kl[2] * kl[2]));
        if (kmag > cutoff) {
          const long e = (((long)i * o->nkj + j) * o->nkk + k) * 2;
          for (int c = 0; c < 3; c++) o->u[c][e] = o->u[c][e + 1] = 0.0;
        }
      }
}

/* ------------------------------------------------------------------------- */
/* Stage / step drivers                                                       */
/* ------------------------------------------------------------------------- */
void orc_stage(orc* o, int stage, int tstep, double time, double* dt) {
  (void)time;
  if (stage == 1) orc_remesh(o, tstep);
  /* f_RK3Stage1: p_StatsPrimary before the stage, p_StatsRelativeVelocity after
   * it (src/pf_rk3.f90:62-99), when tstep > P_TS_RELEASE */
  const int pst = stage == 1 && o->pst_on && o->np > 0 && tstep > o->ts_release;
  if (pst) pstats_primary(o, *dt);
  orc_dealias(o);  /* entry of f_RK3PrepareRealSpace, src/pf_rk3.f90:269-277 */
  orc_prepare_realspace(o);
  orc_realspace(o, stage, tstep, dt);
  /* particles move from tstep = P_TS_RELEASE on (src/pf_rk3.f90:1146); at the
   * release, p_ControlInitPartVelocity[BF] (src/pp_control.f90:378-510) resets
   * particle_up_rk3 and particle_xp0 (particle_up is reassigned by the solve) */
  if (o->np > 0 && tstep >= o->ts_release) {
    if (stage == 1 && tstep == o->ts_release) {
      memset(o->uprk, 0, sizeof(double) * 3 * (size_t)o->np);
      if (o->xp0) memcpy(o->xp0, o->xp, sizeof(double) * 3 * (size_t)o->np);
    }
    tracers_stage(o, stage, *dt);
  }
  if (pst) pstats_relvel(o, tstep);
  orc_force_calc(o, tstep, *dt);  /* src/pf_rk3.f90:2063 */
  solve_ns(o, *dt, stage);
  orc_dealias(o);  /* end of f_RK3SolveNavierStokes, src/pf_rk3.f90:2101-2109 */
}

/* g_RK3UpdateB, src/pg_rk3.f90:280-385 */
void orc_update_b(orc* o, int tstep, double simtime) {
  if (o->p.f_type == 1) return;
  const int b110 = o->p.b110 > 0 ? o->p.b110 : 1;
  double localtime;
  if (o->p.rk3_timestep == 1) {
    if (tstep <= (b110 * o->p.f_remesh_t)) {
      localtime = simtime;
    } else {
      double a = (double)(tstep - (b110 * o->p.f_remesh_t) - 1);
      double b = (double)(2 * (b110 * o->p.f_remesh_t));
      int localtstep = (int)(a / b);
      localtime = simtime - ((double)(localtstep + 1) / o->p.f_deform_s * (double)b110);
    }
  } else {
    localtime = simtime;
  }
  T9(o->bmat, 1, 3) = -T9(o->bmat, 1, 1) * o->p.f_deform_s * localtime;
  T9(o->binv, 1, 3) = o->p.f_deform_s * localtime / T9(o->bmat, 3, 3);
}

void orc_step(orc* o, int tstep, double* simtime) {
  const double tstart = *simtime;
  orc_stage(o, 1, tstep, *simtime, &o->dt);
  *simtime = tstart + (1.0 / 3.0) * o->dt;
  orc_update_b(o, tstep, *simtime);
  orc_stage(o, 2, tstep, *simtime, &o->dt);
  *simtime = tstart + (3.0 / 4.0) * o->dt;
  orc_update_b(o, tstep, *simtime);
  orc_stage(o, 3, tstep, *simtime, &o->dt);
  *simtime = tstart + o->dt;
  orc_update_b(o, tstep, *simtime);
}

/* ------------------------------------------------------------------------- */
/* Wave-space statistics: src/pf_fluidstats.f90:1423-1576                     */
/* ------------------------------------------------------------------------- */
/* ------------------------------------------------------------------------- */
/* Wave-space statistics: f_FluidstatsFs (src/pf_fluidstats.f90:1293-1361)    */
/* ------------------------------------------------------------------------- */
/* lab-frame wavenumber and velocity of a mode: k_lab(n) = sum_m bmat(m,n) k(m),
 * u_lab(m) = sum_n binv(m,n) u(n) over existing entries
 * (f_FluidstatsFsPrimaryQuantsBFRot, src/pf_fluidstats.f90:1633-1642); the
 * isotropic frame is the identity, which reproduces f_FluidstatsFsPrimaryQuants. */
static void lab_frame(const orc* o, const double* kv, const double um[3][2], double* kl, double ul[3][2]) {
  for (int n = 0; n < 3; n++) { kl[n] = 0.0; ul[n][0] = ul[n][1] = 0.0; }
  for (int m = 1; m <= 3; m++)
    for (int n = 1; n <= 3; n++)
      if (T9(o->bval, m, n)) {
        kl[n - 1] = kl[n - 1] + T9(o->bmat, m, n) * kv[m - 1];
        ul[m - 1][0] = ul[m - 1][0] + T9(o->binv, m, n) * um[n - 1][0];
        ul[m - 1][1] = ul[m - 1][1] + T9(o->binv, m, n) * um[n - 1][1];
      }
  /* f_FluidstatsFsPrimaryQuantsRRot (src/pf_fluidstats.f90:1801-1970): the
   * Rogallo forms hold laboratory components, ulab = umov */
  if (rogallo(o))
    for (int n = 0; n < 3; n++) { ul[n][0] = um[n][0]; ul[n][1] = um[n][1]; }
}

/* bin kk in 1..kkmax with kk - 1/2 <= |k| < kk + 1/2, or 0 (src/pf_fluidstats.f90:1530-1545) */
static int shell_of(double km, int kkmax) {
  const int kc = (int)floor(km + 0.5);
  for (int kk = kc - 1; kk <= kc + 1; kk++) {
    if (kk < 1 || kk > kkmax) continue;
    const double up_ = (double)kk + 0.5, lo_ = (double)kk - 0.5;
    if (km >= lo_ && km < up_) return kk;
  }
  return 0;
}

/* raw[7*kkmax]: shell sums ken, ken1, ken2, ken3, dis, disskew, then the mode
 * count (1 for i = 0, 2 otherwise); uu[6]: u1u1 u2u2 u3u3 u1u2 u1u3 u2u3 */
void orc_stats_raw(const orc* o, double* raw, double* uu) {
  const int kkmax = o->kkmax;
  for (int q = 0; q < 7 * kkmax; q++) raw[q] = 0.0;
  for (int q = 0; q < 6; q++) uu[q] = 0.0;
  const double nu = o->p.f_nu;
  for (int i = 0; i < o->nki; i++) {
    const double wi = wvn(i, o->ki);
    for (int j = 0; j < o->nkj; j++) {
      const double wj = wvn(j, o->kj);
      for (int k = 0; k < o->nkk; k++) {
        const double kv[3] = {wi, wj, wvn(k, o->kk)};
        const long e = (((long)i * o->nkj + j) * o->nkk + k) * 2;
        double um[3][2], ul[3][2], kl[3];
        for (int c = 0; c < 3; c++) { um[c][0] = o->u[c][e]; um[c][1] = o->u[c][e + 1]; }
        lab_frame(o, kv, um, kl, ul);
        const double km = sqrt((kl[0] * kl[0]) + (kl[1] * kl[1]) + (kl[2] * kl[2]));
        double ph[6];
        ph[0] = ul[0][0] * ul[0][0] + ul[0][1] * ul[0][1];
        ph[1] = ul[1][0] * ul[1][0] + ul[1][1] * ul[1][1];
        ph[2] = ul[2][0] * ul[2][0] + ul[2][1] * ul[2][1];
        ph[3] = ul[0][0] * ul[1][0] + ul[0][1] * ul[1][1];
        ph[4] = ul[0][0] * ul[2][0] + ul[0][1] * ul[2][1];
        ph[5] = ul[1][0] * ul[2][0] + ul[1][1] * ul[2][1];
        if (i != 0) for (int q = 0; q < 6; q++) ph[q] = ph[q] * 2.0;
        for (int q = 0; q < 6; q++) uu[q] += ph[q];
        const double e_ = 0.5 * (ph[0] + ph[1] + ph[2]);
        const double d_ = 2.0 * km * km * nu * e_;
        const double ds = 4.0 * km * km * km * km * nu * nu * e_;
        const int kk = shell_of(km, kkmax);
        if (kk) {
          double* r = raw + (kk - 1);
          r[0] += e_;
          r[kkmax] += 0.5 * ph[0];
          r[2 * kkmax] += 0.5 * ph[1];
          r[3 * kkmax] += 0.5 * ph[2];
          r[4 * kkmax] += d_;
          r[5 * kkmax] += ds;
          r[6 * kkmax] += (i != 0) ? 2.0 : 1.0;
        }
      }
    }
  }
}

/* From the raw sums: mode density (f_FluidstatsModeDensity[BFRot],
 * src/pf_fluidstats.f90:696-774, 1977-2049), the normalised spectra and totals
 * (1547-1575) and the secondary quantities (f_FluidstatsFsSecondaryQuants,
 * 2491-2558).  spec[7*kkmax] = spec_ken, ken1, ken2, ken3, dis, disskew,
 * dens_mod; q[33] in the order of pdr_stats_t. */
void orc_stats_finish(int kkmax, double nu, const double* raw, const double* uu, double* spec, double* q) {
  double f_ken = 0, f_dis = 0, f_ds = 0, fn[6] = {0, 0, 0, 0, 0, 0};
  for (int kk = 1; kk <= kkmax; kk++) {
    const double up_ = (double)kk + 0.5, lo_ = (double)kk - 0.5;
    const double dens = raw[6 * kkmax + kk - 1] / (4.0 / 3.0 * ONEPI_REF * (pow(up_, 3) - pow(lo_, 3)));
    spec[6 * kkmax + kk - 1] = dens;
    for (int a = 0; a < 6; a++) spec[a * kkmax + kk - 1] = 0.0;
    if (dens != 0.0) {
      f_ken = f_ken + raw[kk - 1];
      f_dis = f_dis + raw[4 * kkmax + kk - 1];
      f_ds = f_ds + raw[5 * kkmax + kk - 1];
      for (int a = 0; a < 6; a++) {
        spec[a * kkmax + kk - 1] = raw[a * kkmax + kk - 1] / dens;
        fn[a] = fn[a] + spec[a * kkmax + kk - 1];
      }
    }
  }
  q[0] = 0.5 * (uu[0] + uu[1] + uu[2]);
  for (int a = 0; a < 6; a++) q[1 + a] = uu[a];
  q[7] = f_ken; q[8] = f_dis; q[9] = f_ds;
  for (int a = 0; a < 6; a++) q[10 + a] = fn[a];
  const double fn_ken = fn[0], fn_dis = fn[4], fn_ds = fn[5];
  const double fint = sqrt(2.0 * fn_ken / 3.0);
  q[16] = fint;
  q[17] = sqrt(2.0 * fn[1]);
  q[18] = sqrt(2.0 * fn[2]);
  q[19] = sqrt(2.0 * fn[3]);
  q[20] = fn_ken / fn_dis;
  q[21] = pow(fn_ken, 3.0 / 2.0) / fn_dis;
  q[22] = sqrt((15.0 * nu * (fint * fint)) / fn_dis);
  q[23] = q[22] * fint / nu;
  double li[4] = {0, 0, 0, 0};
  for (int kk = 1; kk <= kkmax; kk++) {
    const double f = ONEPI_REF / 2.0 / (fint * fint);
    li[0] = li[0] + (f * spec[kk - 1] / (double)kk);
    li[1] = li[1] + (f * spec[kkmax + kk - 1] / (double)kk);
    li[2] = li[2] + (f * spec[2 * kkmax + kk - 1] / (double)kk);
    li[3] = li[3] + (f * spec[3 * kkmax + kk - 1] / (double)kk);
  }
  for (int a = 0; a < 4; a++) q[24 + a] = li[a];
  q[28] = li[0] / fint;
  q[29] = pow(nu * nu * nu / fn_dis, 1.0 / 4.0);
  q[30] = sqrt(nu / fn_dis);
  q[31] = fn_ds * pow(15.0 / fn_dis, 3.0 / 2.0) * sqrt(nu) / 35.0;
  q[32] = fn_dis * li[0] / (fint * fint * fint);
}

int orc_kkmax(const orc* o) { return o->kkmax; }

void orc_stats(const orc* o, double* spec, double* q) {
  const int kkmax = o->kkmax;
  double* raw = (double*)malloc(sizeof(double) * 7 * (size_t)kkmax);
  double uu[6];
  orc_stats_raw(o, raw, uu);
  orc_stats_finish(kkmax, o->p.f_nu, raw, uu, spec, q);
  free(raw);
}

/* shell sums before normalisation (ek, dk, dsk) and {f_ken, f_dis, f_disskew, u1u1, u2u2, u3u3} */
void orc_spectrum(const orc* o, double* ek, double* dk, double* dsk, double* tot) {
  const int kkmax = o->kkmax;
  double* raw = (double*)malloc(sizeof(double) * 7 * (size_t)kkmax);
  double uu[6];
  orc_stats_raw(o, raw, uu);
  double fk = 0, fd = 0, fs = 0;
  for (int q = 0; q < kkmax; q++) {
    ek[q] = raw[q]; dk[q] = raw[4 * kkmax + q]; dsk[q] = raw[5 * kkmax + q];
    if (raw[6 * kkmax + q] != 0.0) { fk += ek[q]; fd += dk[q]; fs += dsk[q]; }
  }
  tot[0] = fk; tot[1] = fd; tot[2] = fs; tot[3] = uu[0]; tot[4] = uu[1]; tot[5] = uu[2];
  free(raw);
}
