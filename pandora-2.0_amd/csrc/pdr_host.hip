// pdr_host.hip -- host-side pieces of the C-ABI that sit either side of the
// time step: seeded initial conditions and PETSc-binary restart I/O.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pdr_internal.hpp"
#include "pdr_rng.hpp"

namespace {

using pdr::u01;

// float32 pi widened to double: ONEPI, TWOPI, FOURPI of src/pg_constants.f90:126-131
const double kOnePi = double(3.14159265358979323846f);
const double kTwoPi = 2.0 * kOnePi;
const double kFourPi = 4.0 * kOnePi;
const int32_t kVecFileClassId = 1211214;  // PETSc VEC_FILE_CLASSID

inline double wvn(int k, int kmax) { return (k <= kmax) ? double(k) : double(k - (2 * kmax + 1)); }

struct Mode {
  double t[3][2];
};

// One Rogallo mode (src/pf_initialise.f90:161-219) with the type-10 spectrum
// E(k) = 16 (2/pi)^0.5 (3/2) U0 k^4/kp^5 exp(-2k^2/kp^2) (src/pf_initialise.f90:1081-1083).
Mode rogallo(double ki, double kj, double kk, double kmag, double k2d, double r1, double r2, double r3,
             double u0, double kp) {
  Mode m;
  std::memset(&m, 0, sizeof(m));
  if (!(kmag > 0.0) || !(k2d > 0.0)) return m;
  const double ek = 16.0 * std::sqrt(2.0 / kOnePi) * 1.5 * u0 *
                    ((kmag * kmag * kmag * kmag) / (kp * kp * kp * kp * kp)) *
                    std::exp(-(2.0 * kmag * kmag) / (kp * kp));
  const double factor = std::sqrt(ek / (kFourPi * kmag * kmag));
  const double al[2] = {factor * std::cos(r1) * std::cos(r3), factor * std::sin(r1) * std::cos(r3)};
  const double be[2] = {factor * std::cos(r2) * std::sin(r3), factor * std::sin(r2) * std::sin(r3)};
  const double s2 = std::sqrt(k2d);
  for (int q = 0; q < 2; ++q) {
    m.t[0][q] = (al[q] * kmag * kj + be[q] * ki * kk) / (kmag * s2);
    m.t[1][q] = (be[q] * kj * kk - al[q] * kmag * ki) / (kmag * s2);
    m.t[2][q] = -(be[q] * s2) / kmag;
  }
  return m;
}

void put(double** u, long long e, const Mode& m, bool conj) {
  for (int c = 0; c < 3; ++c) {
    u[c][2 * e] = m.t[c][0];
    u[c][2 * e + 1] = conj ? -m.t[c][1] : m.t[c][1];
  }
}

void be_store(unsigned char* p, uint64_t v, int n) {
  for (int q = 0; q < n; ++q) p[q] = (unsigned char)((v >> (8 * (n - 1 - q))) & 0xff);
}
uint64_t be_load(const unsigned char* p, int n) {
  uint64_t v = 0;
  for (int q = 0; q < n; ++q) v = (v << 8) | p[q];
  return v;
}

// Neumaier compensated summation

}  // namespace

extern "C" {

int pdr_init_field_host(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double init_u0,
                        double init_kp, double* u1, double* u2, double* u3) {
  if (nx < 2 || ny < 2 || nz < 2 || !u1 || !u2 || !u3 || !(init_kp > 0.0))
    return pdr::set_error(PDR_ERR_ARG, "pdr_init_field_host: bad argument");
  const int kj = ny / 3, kk = nz / 3;
  const int nki = nx / 3 + 1, nkj = 2 * kj + 1, nkk = 2 * kk + 1;
  const long long mw = (long long)nki * nkj * nkk;
  double* u[3] = {u1, u2, u3};
  for (int c = 0; c < 3; ++c) std::memset(u[c], 0, sizeof(double) * size_t(2 * mw));
  auto idx = [&](int i, int j, int k) { return ((long long)i * nkj + j) * nkk + k; };
  auto draw = [&](long long id, double* r) {
    for (int q = 0; q < 3; ++q) r[q] = kTwoPi * u01(seed, uint64_t(id) * 3 + uint64_t(q));
  };
  // i = 0 plane with conjugate symmetry u(0,-j,-k) = conj u(0,j,k) (src/pf_initialise.f90:150-411)
  for (int j = 0; j <= kj; ++j)
    for (int k = 0; k <= kk; ++k) {
      const double wj = wvn(j, kj), wk = wvn(k, kk);
      const double kmag = std::sqrt((wj * wj) + (wk * wk));
      const double k2d = wj * wj;
      double r[3];
      draw(2 * idx(0, j, k), r);
      const Mode m = rogallo(0.0, wj, wk, kmag, k2d, r[0], r[1], r[2], init_u0, init_kp);
      const int jj = (j != 0) ? nkj - j : j;
      const int kc = (k != 0) ? nkk - k : k;
      put(u, idx(0, j, k), m, false);
      put(u, idx(0, jj, kc), m, true);
      if (k != 0) {
        draw(2 * idx(0, j, k) + 1, r);
        const Mode m2 = rogallo(0.0, wvn(jj, kj), wk, kmag, k2d, r[0], r[1], r[2], init_u0, init_kp);
        put(u, idx(0, jj, k), m2, false);
        put(u, idx(0, j, kc), m2, true);
      }
    }
  // i >= 1: no symmetry to enforce (src/pf_initialise.f90:416-505)
  for (int i = 1; i < nki; ++i)
    for (int j = 0; j < nkj; ++j)
      for (int k = 0; k < nkk; ++k) {
        const double wi = double(i), wj = wvn(j, kj), wk = wvn(k, kk);
        const double kmag = std::sqrt((wi * wi) + (wj * wj) + (wk * wk));
        const double k2d = (wi * wi) + (wj * wj);
        double r[3];
        draw(2 * idx(i, j, k), r);
        put(u, idx(i, j, k), rogallo(wi, wj, wk, kmag, k2d, r[0], r[1], r[2], init_u0, init_kp), false);
      }
  return PDR_OK;
}

int pdr_init_field(pdr_ctx* ctx, uint64_t seed, double init_u0, double init_kp) {
  if (!ctx) return pdr::set_error(PDR_ERR_ARG, "pdr_init_field: null context");
  // a stretched box takes f_InitialiseSpec's deformed variant (src/pf_initialise.f90:
  // 67-71, 550-632), which is not built: such a run starts from an uploaded state
  if (pdr::box_stretch(ctx->p, 1) * pdr::box_stretch(ctx->p, 2) * pdr::box_stretch(ctx->p, 3) != 1)
    return pdr::set_error(PDR_ERR_UNSUPPORTED, "pdr_init_field: the stretched-box initial field is not built; upload a state");
  const pdr::Geo& g = ctx->g;
  std::vector<double> u[3];
  for (int c = 0; c < 3; ++c) u[c].assign(size_t(2 * g.mw), 0.0);
  int rc = pdr_init_field_host(g.nx, g.ny, g.nz, seed, init_u0, init_kp, u[0].data(), u[1].data(), u[2].data());
  if (rc != PDR_OK) return rc;
  rc = pdr_upload_state(ctx, u[0].data(), u[1].data(), u[2].data());
  if (rc != PDR_OK) return rc;
  for (int c = 0; c < 3; ++c) std::fill(u[c].begin(), u[c].end(), 0.0);
  rc = pdr_upload_rk3(ctx, u[0].data(), u[1].data(), u[2].data());
  // f_InitialiseInit truncates the new field (src/pf_initialise.f90:75-83)
  return rc == PDR_OK ? pdr_dealias(ctx) : rc;
}

int pdr_init_particles(pdr_ctx* ctx, uint64_t seed, int64_t np) {
  if (!ctx || np < 0) return pdr::set_error(PDR_ERR_ARG, "pdr_init_particles: bad argument");
  const long double lx = kTwoPi, ly = kTwoPi, lz = kTwoPi;
  const long double dx = (long double)(kTwoPi / double(ctx->g.nx));
  const long double dy = (long double)(kTwoPi / double(ctx->g.ny));
  const long double dz = (long double)(kTwoPi / double(ctx->g.nz));
  std::vector<double> xp(size_t(3 * np));
  // P_INIT_RANDOM: xp = rn * L - dx/2 evaluated in long double (src/pp_control.f90:657-686)
  for (int64_t q = 0; q < np; ++q) {
    const uint64_t b = uint64_t(q) * 3;
    xp[size_t(3 * q)] = double((long double)u01(seed, b) * lx - dx * 0.5L);
    xp[size_t(3 * q + 1)] = double((long double)u01(seed, b + 1) * ly - dy * 0.5L);
    xp[size_t(3 * q + 2)] = double((long double)u01(seed, b + 2) * lz - dz * 0.5L);
  }
  return pdr_particles_upload(ctx, xp.data(), np);
}

int pdr_restart_write_host(const char* path, int64_t n, const double* u1, const double* u2,
                           const double* u3, int32_t int64_header) {
  if (!path || n <= 0 || !u1 || !u2 || !u3) return pdr::set_error(PDR_ERR_ARG, "pdr_restart_write_host: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return pdr::set_error(PDR_ERR_IO, std::string("cannot open for writing: ") + path);
  const int hl = int64_header ? 8 : 4;
  const double* u[3] = {u1, u2, u3};
  std::vector<unsigned char> buf(size_t(n) * 8);
  bool ok = true;
  for (int c = 0; c < 3 && ok; ++c) {
    unsigned char h[12];
    be_store(h, uint64_t(uint32_t(kVecFileClassId)), 4);
    be_store(h + 4, uint64_t(n), hl);
    ok = std::fwrite(h, 1, size_t(4 + hl), f) == size_t(4 + hl);
    for (int64_t q = 0; q < n && ok; ++q) {
      uint64_t bits;
      std::memcpy(&bits, &u[c][q], 8);
      be_store(&buf[size_t(8 * q)], bits, 8);
    }
    if (ok) ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? PDR_OK : pdr::set_error(PDR_ERR_IO, std::string("write failed: ") + path);
}

int pdr_restart_read_host(const char* path, int64_t n, double* u1, double* u2, double* u3) {
  if (!path || n <= 0 || !u1 || !u2 || !u3) return pdr::set_error(PDR_ERR_ARG, "pdr_restart_read_host: bad argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return pdr::set_error(PDR_ERR_IO, std::string("cannot open: ") + path);
  std::fseek(f, 0, SEEK_END);
  const long long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  int hl = 0;
  if (size == 3 * (8 + 8 * n)) hl = 4;
  if (size == 3 * (12 + 8 * n)) hl = 8;
  if (!hl) {
    std::fclose(f);
    return pdr::set_error(PDR_ERR_IO, "restart file size does not match three PETSc Vec records of this grid");
  }
  double* u[3] = {u1, u2, u3};
  std::vector<unsigned char> buf(size_t(n) * 8);
  bool ok = true;
  for (int c = 0; c < 3 && ok; ++c) {
    unsigned char h[12];
    ok = std::fread(h, 1, size_t(4 + hl), f) == size_t(4 + hl);
    if (ok) ok = int32_t(be_load(h, 4)) == kVecFileClassId && (int64_t)be_load(h + 4, hl) == n;
    if (ok) ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    if (!ok) break;
    for (int64_t q = 0; q < n; ++q) {
      const uint64_t bits = be_load(&buf[size_t(8 * q)], 8);
      std::memcpy(&u[c][q], &bits, 8);
    }
  }
  std::fclose(f);
  if (!ok) return pdr::set_error(PDR_ERR_IO, "bad PETSc Vec header or short read in restart file");
  return PDR_OK;
}

int pdr_restart_write(pdr_ctx* ctx, const char* path, int32_t int64_header) {
  if (!ctx || !path) return pdr::set_error(PDR_ERR_ARG, "pdr_restart_write: null argument");
  const long long n = 2 * ctx->g.mw;
  std::vector<double> u[3];
  for (int c = 0; c < 3; ++c) u[c].resize(static_cast<size_t>(n));
  int rc = pdr_download_state(ctx, u[0].data(), u[1].data(), u[2].data());
  if (rc != PDR_OK) return rc;
  return pdr_restart_write_host(path, n, u[0].data(), u[1].data(), u[2].data(), int64_header);
}

// Particle restart records (g_RestartPetscParticle, src/pg_restart.f90:264-318):
// xp1, xp2, xp3, up1, up2, up3, dp, status[, xp01, xp02, xp03], each a PETSc Vec
// record {int32 VEC_FILE_CLASSID, PetscInt np, np big-endian doubles}.
int pdr_particles_restart_write_host(const char* path, int64_t np, const double* xp, const double* up,
                                     const double* xp0, int32_t int64_header) {
  if (!path || np <= 0 || !xp || !up)
    return pdr::set_error(PDR_ERR_ARG, "pdr_particles_restart_write_host: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return pdr::set_error(PDR_ERR_IO, std::string("cannot open for writing: ") + path);
  const int hl = int64_header ? 8 : 4;
  const int nrec = xp0 ? 11 : 8;
  std::vector<unsigned char> buf(size_t(np) * 8);
  bool ok = true;
  for (int r = 0; r < nrec && ok; ++r) {
    unsigned char h[12];
    be_store(h, uint64_t(uint32_t(kVecFileClassId)), 4);
    be_store(h + 4, uint64_t(np), hl);
    ok = std::fwrite(h, 1, size_t(4 + hl), f) == size_t(4 + hl);
    for (int64_t q = 0; q < np && ok; ++q) {
      double v;
      if (r < 3)
        v = xp[3 * q + r];
      else if (r < 6)
        v = up[3 * q + (r - 3)];
      else if (r == 6)
        v = 0.0;  // particle diameter: tracers
      else if (r == 7)
        v = double(q + 1);  // particle_status(1,:): the particle number
      else
        v = xp0[3 * q + (r - 8)];
      uint64_t bits;
      std::memcpy(&bits, &v, 8);
      be_store(&buf[size_t(8 * q)], bits, 8);
    }
    if (ok) ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? PDR_OK : pdr::set_error(PDR_ERR_IO, std::string("write failed: ") + path);
}

int pdr_particles_restart_read_host(const char* path, int64_t np, double* xp, double* up, double* xp0,
                                    int32_t* has_xp0) {
  if (!path || np <= 0 || !xp || !up) return pdr::set_error(PDR_ERR_ARG, "pdr_particles_restart_read_host: bad argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return pdr::set_error(PDR_ERR_IO, std::string("cannot open: ") + path);
  std::fseek(f, 0, SEEK_END);
  const long long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  int hl = 0, nrec = 0;
  for (int r : {8, 11})
    for (int w : {4, 8})
      if (size == (long long)r * (4 + w + 8 * np)) {
        hl = w;
        nrec = r;
      }
  if (!hl) {
    std::fclose(f);
    return pdr::set_error(PDR_ERR_IO, "particle restart size does not match 8 or 11 PETSc Vec records of np particles");
  }
  std::vector<double> rec(size_t(nrec) * size_t(np));
  std::vector<unsigned char> buf(size_t(np) * 8);
  bool ok = true;
  for (int r = 0; r < nrec && ok; ++r) {
    unsigned char h[12];
    ok = std::fread(h, 1, size_t(4 + hl), f) == size_t(4 + hl);
    if (ok) ok = int32_t(be_load(h, 4)) == kVecFileClassId && (int64_t)be_load(h + 4, hl) == np;
    if (ok) ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    for (int64_t q = 0; q < np && ok; ++q) {
      const uint64_t bits = be_load(&buf[size_t(8 * q)], 8);
      std::memcpy(&rec[size_t(r) * size_t(np) + size_t(q)], &bits, 8);
    }
  }
  std::fclose(f);
  if (!ok) return pdr::set_error(PDR_ERR_IO, "bad PETSc Vec header or short read in particle restart file");
  // place every entry by its particle number (status record); it must be a permutation of 1..np
  std::vector<char> seen(size_t(np), 0);
  const double* st = &rec[size_t(7) * size_t(np)];
  for (int64_t q = 0; q < np; ++q) {
    const double v = st[q];
    const int64_t id = int64_t(v) - 1;
    if (!(v == std::floor(v)) || id < 0 || id >= np || seen[size_t(id)])
      return pdr::set_error(PDR_ERR_IO, "particle restart: status record is not a numbering 1..np");
    seen[size_t(id)] = 1;
    for (int c = 0; c < 3; ++c) {
      xp[3 * id + c] = rec[size_t(c) * size_t(np) + size_t(q)];
      up[3 * id + c] = rec[size_t(3 + c) * size_t(np) + size_t(q)];
      if (xp0 && nrec == 11) xp0[3 * id + c] = rec[size_t(8 + c) * size_t(np) + size_t(q)];
    }
  }
  if (has_xp0) *has_xp0 = nrec == 11 ? 1 : 0;
  return PDR_OK;
}

int pdr_particles_restart_write(pdr_ctx* ctx, const char* path, int32_t int64_header) {
  if (!ctx || !path) return pdr::set_error(PDR_ERR_ARG, "pdr_particles_restart_write: null argument");
  if (ctx->np <= 0) return pdr::set_error(PDR_ERR_STATE, "pdr_particles_restart_write: no particles set");
  const size_t n3 = size_t(3) * size_t(ctx->np);
  std::vector<double> xp(n3), up(n3), xp0(ctx->xp0 ? n3 : 0);
  int rc = pdr_particles_download(ctx, xp.data(), nullptr);
  if (rc == PDR_OK) rc = pdr_particles_download_aux(ctx, up.data(), ctx->xp0 ? xp0.data() : nullptr);
  if (rc != PDR_OK) return rc;
  return pdr_particles_restart_write_host(path, ctx->np, xp.data(), up.data(), ctx->xp0 ? xp0.data() : nullptr,
                                          int64_header);
}

int pdr_particles_restart_read(pdr_ctx* ctx, const char* path, int64_t np) {
  if (!ctx || !path || np <= 0) return pdr::set_error(PDR_ERR_ARG, "pdr_particles_restart_read: bad argument");
  const size_t n3 = size_t(3) * size_t(np);
  std::vector<double> xp(n3), up(n3), xp0(n3);
  int32_t has0 = 0;
  int rc = pdr_particles_restart_read_host(path, np, xp.data(), up.data(), xp0.data(), &has0);
  if (rc == PDR_OK) rc = pdr_particles_upload(ctx, xp.data(), np);
  if (rc == PDR_OK) rc = pdr_particles_upload_aux(ctx, up.data(), (has0 && ctx->xp0) ? xp0.data() : nullptr);
  return rc;
}

int pdr_restart_read_scaled_host(const char* path, int32_t nx, int32_t ny, int32_t nz, int32_t scale,
                                 double* u1, double* u2, double* u3) {
  if (!path || scale < 1 || nx < 1 || ny < 1 || nz < 1 || !u1 || !u2 || !u3)
    return pdr::set_error(PDR_ERR_ARG, "pdr_restart_read_scaled_host: bad argument");
  // fine (this grid) and coarse (file) compact sizes, src/pg_domain.f90:220-228 and
  // f_ArraysVectorCreateRestart (src/pf_arrays.f90:1067-1073)
  const int kif = nx / 3, kjf = ny / 3, kkf = nz / 3;
  const int nkjf = 2 * kjf + 1, nkkf = 2 * kkf + 1;
  const int kic = (nx / scale) / 3, kjc = (ny / scale) / 3, kkc = (nz / scale) / 3;
  const int nkic = kic + 1, nkjc = 2 * kjc + 1, nkkc = 2 * kkc + 1;
  if (kic < 1 || kjc < 1 || kkc < 1)
    return pdr::set_error(PDR_ERR_ARG, "pdr_restart_read_scaled_host: coarse grid too small");
  const int64_t nc = 2LL * nkic * nkjc * nkkc;
  std::vector<double> c[3];
  for (auto& v : c) v.resize(size_t(nc));
  int rc = pdr_restart_read_host(path, nc, c[0].data(), c[1].data(), c[2].data());
  if (rc != PDR_OK) return rc;
  double* f[3] = {u1, u2, u3};
  const size_t nf = size_t(2) * size_t(kif + 1) * nkjf * nkkf;
  for (auto* p : f) std::fill(p, p + nf, 0.0);
  // signed wavenumber -> compact index: s >= 0 at s, s < 0 at n + s
  for (int i = 0; i <= kic; ++i)
    for (int sj = -kjc; sj <= kjc; ++sj)
      for (int sk = -kkc; sk <= kkc; ++sk) {
        const size_t jc = size_t(sj >= 0 ? sj : nkjc + sj), kc = size_t(sk >= 0 ? sk : nkkc + sk);
        const size_t jf = size_t(sj >= 0 ? sj : nkjf + sj), kf = size_t(sk >= 0 ? sk : nkkf + sk);
        const size_t ec = 2 * ((size_t(i) * nkjc + jc) * nkkc + kc);
        const size_t ef = 2 * ((size_t(i) * nkjf + jf) * nkkf + kf);
        for (int q = 0; q < 3; ++q) {
          f[q][ef] = c[q][ec];
          f[q][ef + 1] = c[q][ec + 1];
        }
      }
  return PDR_OK;
}

int pdr_restart_read_scaled(pdr_ctx* ctx, const char* path, int32_t scale) {
  if (!ctx || !path) return pdr::set_error(PDR_ERR_ARG, "pdr_restart_read_scaled: null argument");
  const pdr::Geo& g = ctx->g;
  std::vector<double> u[3];
  for (auto& v : u) v.resize(size_t(2 * g.mw));
  int rc = pdr_restart_read_scaled_host(path, g.nx, g.ny, g.nz, scale, u[0].data(), u[1].data(), u[2].data());
  if (rc != PDR_OK) return rc;
  rc = pdr_upload_state(ctx, u[0].data(), u[1].data(), u[2].data());
  return rc == PDR_OK ? pdr_dealias(ctx) : rc;
}

int pdr_restart_read(pdr_ctx* ctx, const char* path) {
  if (!ctx || !path) return pdr::set_error(PDR_ERR_ARG, "pdr_restart_read: null argument");
  const long long n = 2 * ctx->g.mw;
  std::vector<double> u[3];
  for (int c = 0; c < 3; ++c) u[c].resize(static_cast<size_t>(n));
  int rc = pdr_restart_read_host(path, n, u[0].data(), u[1].data(), u[2].data());
  if (rc != PDR_OK) return rc;
  rc = pdr_upload_state(ctx, u[0].data(), u[1].data(), u[2].data());
  return rc == PDR_OK ? pdr_dealias(ctx) : rc;  // as after a restart in f_InitialiseInit
}

}  // extern "C"
