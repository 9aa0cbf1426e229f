// pdr_pstats.hip -- tracer statistics on the device: p_StatsPrimary,
// p_StatsDispersion and p_StatsRelativeVelocity (src/pp_stats.f90:119-390) for
// tracers (dp = 0), with the mean particle velocity of p_RK3AbsoluteVelocity[BF|R]
// (src/pp_rk3.f90:602-773) that the record reports as vt.
//
// Stage 1 of a step (src/pf_rk3.f90:62-99) runs two reductions over the
// particles: before the particle advance, the second moments and the mean of
// particle_up (the velocity of the previous stage) and, with
// P_KEEP_INITIAL_POSITION, the moments of the laboratory-frame displacement
// binv (x - x0); after the advance, the mean squares of the new particle_up.
// Each is a grid-stride pass with FP64 per-thread sums, a fixed-order LDS tree
// per workgroup, and a Neumaier sum over the workgroup partials; a one-thread
// kernel then forms the record (including the diffusivity from the saved
// previous displacement), so nothing synchronises with the host inside a step.
#include <cmath>
#include <vector>

#include "pdr_internal.hpp"

namespace pdr {
namespace {

constexpr int kPsThreads = 256;
constexpr int kPsBlocks = 1024;
constexpr int kPrimW = 18;  // uu(6), sum u(3), sum xl(3), sum xl^2(3), x1x2, x1x3, x2x3
constexpr int kRelW = 4;    // sum |u|^2, sum u_c^2 (3)
constexpr int kRecW = 30;   // p_StatsWrite record (src/pp_stats.f90:379-390)

// Layout of the device buffer (doubles).
struct PsLayout {
  size_t prim_part, prim_sum, rel_part, rel_sum, rec, xprev, tstep, total;
};

PsLayout ps_layout() {
  PsLayout L;
  L.prim_part = 0;
  L.prim_sum = L.prim_part + size_t(kPsBlocks) * kPrimW;
  L.rel_part = L.prim_sum + kPrimW;
  L.rel_sum = L.rel_part + size_t(kPsBlocks) * kRelW;
  L.rec = L.rel_sum + kRelW;
  L.xprev = L.rec + kRecW;
  L.tstep = L.xprev + 3;
  L.total = L.tstep + 1;
  return L;
}

struct PsFrame {
  double bi[9];  // binv over the existing entries, [m][n]; identity for isotropic runs
  int vt_binv;   // R_ROT: vt is the mean of binv particle_up (p_RK3AbsoluteVelocityR)
};

template <int W>
__device__ __forceinline__ void ps_block_out(double (&s)[W], double* sh, double* __restrict__ out) {
  for (int q = 0; q < W; ++q) sh[q * kPsThreads + threadIdx.x] = s[q];
  __syncthreads();
  for (int w = kPsThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int q = 0; q < W; ++q) sh[q * kPsThreads + threadIdx.x] += sh[q * kPsThreads + threadIdx.x + w];
    __syncthreads();
  }
  if (int(threadIdx.x) < W) out[threadIdx.x] = sh[threadIdx.x * kPsThreads];
}

// Primary sums of the particles before the stage-1 advance.
__global__ void __launch_bounds__(kPsThreads)
    k_pstats_primary(const double* __restrict__ up, const double* __restrict__ xp, const double* __restrict__ xp0,
                     long long np, PsFrame f, double* __restrict__ part) {
  __shared__ double sh[kPrimW * kPsThreads];
  double s[kPrimW];
  for (int q = 0; q < kPrimW; ++q) s[q] = 0.0;
  for (long long p = (long long)blockIdx.x * kPsThreads + threadIdx.x; p < np;
       p += (long long)gridDim.x * kPsThreads) {
    const double u1 = up[3 * p], u2 = up[3 * p + 1], u3 = up[3 * p + 2];
    s[0] += u1 * u1;
    s[1] += u2 * u2;
    s[2] += u3 * u3;
    s[3] += u1 * u2;
    s[4] += u1 * u3;
    s[5] += u2 * u3;
    if (f.vt_binv) {
      const double u[3] = {u1, u2, u3};
      for (int m = 0; m < 3; ++m) {
        double a = 0.0;
        for (int n = 0; n < 3; ++n) a = a + f.bi[3 * m + n] * u[n];
        s[6 + m] += a;
      }
    } else {
      s[6] += u1;
      s[7] += u2;
      s[8] += u3;
    }
    if (xp0) {
      double xm[3], xl[3];
      for (int c = 0; c < 3; ++c) xm[c] = xp[3 * p + c] - xp0[3 * p + c];
      for (int m = 0; m < 3; ++m) {
        double a = 0.0;
        for (int n = 0; n < 3; ++n) a = a + f.bi[3 * m + n] * xm[n];
        xl[m] = a;
      }
      for (int c = 0; c < 3; ++c) {
        s[9 + c] += xl[c];
        s[12 + c] += xl[c] * xl[c];
      }
      s[15] += xl[0] * xl[1];
      s[16] += xl[0] * xl[2];
      s[17] += xl[1] * xl[2];
    }
  }
  ps_block_out<kPrimW>(s, sh, part + (size_t)blockIdx.x * kPrimW);
}

// Relative-velocity sums after the stage-1 advance (uf = 0 for tracers).
__global__ void __launch_bounds__(kPsThreads)
    k_pstats_rel(const double* __restrict__ up, long long np, double* __restrict__ part) {
  __shared__ double sh[kRelW * kPsThreads];
  double s[kRelW] = {0.0, 0.0, 0.0, 0.0};
  for (long long p = (long long)blockIdx.x * kPsThreads + threadIdx.x; p < np;
       p += (long long)gridDim.x * kPsThreads) {
    const double u1 = up[3 * p], u2 = up[3 * p + 1], u3 = up[3 * p + 2];
    s[0] += ((u1 * u1) + (u2 * u2)) + (u3 * u3);
    s[1] += u1 * u1;
    s[2] += u2 * u2;
    s[3] += u3 * u3;
  }
  ps_block_out<kRelW>(s, sh, part + (size_t)blockIdx.x * kRelW);
}

// out[q] = Neumaier sum over the nblk workgroup partials of column q
__global__ void k_ps_reduce(const double* __restrict__ part, int nblk, int width, double* __restrict__ out) {
  const int q = int(threadIdx.x);
  if (q >= width) return;
  double s = 0.0, c = 0.0;
  for (int b = 0; b < nblk; ++b) {
    const double v = part[(size_t)b * width + q];
    const double t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  out[q] = s + c;
}

// p_StatsPrimary's finishing statements (src/pp_stats.f90:178-231)
__global__ void k_ps_finish_primary(double* __restrict__ buf, size_t sum_off, size_t rec_off, size_t xprev_off,
                                    double npg, double dt, int disp) {
  const double* g = buf + sum_off;
  double* r = buf + rec_off;
  double* xprev = buf + xprev_off;
  for (int q = 0; q < kRecW; ++q) r[q] = 0.0;
  for (int q = 0; q < 6; ++q) r[q] = g[q] / npg;
  r[6] = 0.5 * ((r[0] + r[1]) + r[2]);
  for (int c = 0; c < 3; ++c) r[21 + c] = g[6 + c] / npg;  // vt = vabs
  if (disp) {
    for (int c = 0; c < 3; ++c) {
      const double xrel = g[12 + c] / npg;
      r[12 + c] = xrel;
      r[15 + c] = 0.5 * (xrel - xprev[c]) / dt;
      xprev[c] = xrel;
    }
    r[18] = g[15] / npg;
    r[19] = g[16] / npg;
    r[20] = g[17] / npg;
  }
}

// p_StatsRelativeVelocity (src/pp_stats.f90:240-296)
__global__ void k_ps_finish_rel(double* __restrict__ buf, size_t sum_off, size_t rec_off, size_t tstep_off,
                                double npg, double tstep) {
  const double* g = buf + sum_off;
  double* r = buf + rec_off;
  for (int c = 0; c < 3; ++c) r[8 + c] = g[1 + c] / npg;
  r[11] = sqrt(g[0] / npg);
  buf[tstep_off] = tstep;
}

int ps_blocks(long long np) {
  const long long want = (np + 4LL * kPsThreads - 1) / (4LL * kPsThreads);
  return int(want < 1 ? 1 : (want < kPsBlocks ? want : kPsBlocks));
}

}  // namespace

int pstats_alloc(pdr_ctx* c) {
  if (c->pst_buf) return PDR_OK;
  const PsLayout L = ps_layout();
  if (hipMalloc(&c->pst_buf, L.total * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    c->pst_buf = nullptr;
    return set_error(PDR_ERR_NOMEM, "tracer statistics: device allocation failed");
  }
  c->dev_bytes += L.total * sizeof(double);
  return pstats_reset(c);
}

int pstats_reset(pdr_ctx* c) {
  if (!c->pst_buf) return PDR_OK;
  const PsLayout L = ps_layout();
  std::vector<double> h(L.total - L.rec, 0.0);
  h.back() = -1.0;  // no record yet
  if (hipMemcpyAsync(c->pst_buf + L.rec, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return set_error(PDR_ERR_HIP, "tracer statistics: reset failed");
  return PDR_OK;
}

hipError_t launch_pstats_primary(pdr_ctx* c, double dt) {
  const PsLayout L = ps_layout();
  PsFrame f;
  const bool iso = c->p.f_type == PDR_F_ISOTROPIC;
  for (int m = 1; m <= 3; ++m)
    for (int n = 1; n <= 3; ++n)
      f.bi[3 * (m - 1) + (n - 1)] = iso ? (m == n ? 1.0 : 0.0) : (T9(c->bval, m, n) ? T9(c->binv, m, n) : 0.0);
  f.vt_binv = rogallo_form(c) ? 1 : 0;
  const int nb = ps_blocks(c->np);
  k_pstats_primary<<<dim3(unsigned(nb)), kPsThreads, 0, c->stream>>>(c->up, c->xp, c->xp0, c->np, f,
                                                                     c->pst_buf + L.prim_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_ps_reduce<<<1, 64, 0, c->stream>>>(c->pst_buf + L.prim_part, nb, kPrimW, c->pst_buf + L.prim_sum);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_ps_finish_primary<<<1, 1, 0, c->stream>>>(c->pst_buf, L.prim_sum, L.rec, L.xprev, double(c->np), dt,
                                             c->xp0 ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_pstats_rel(pdr_ctx* c, int tstep, hipStream_t s) {
  const PsLayout L = ps_layout();
  const int nb = ps_blocks(c->np);
  k_pstats_rel<<<dim3(unsigned(nb)), kPsThreads, 0, s>>>(c->up, c->np, c->pst_buf + L.rel_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_ps_reduce<<<1, 64, 0, s>>>(c->pst_buf + L.rel_part, nb, kRelW, c->pst_buf + L.rel_sum);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_ps_finish_rel<<<1, 1, 0, s>>>(c->pst_buf, L.rel_sum, L.rec, L.tstep, double(c->np), double(tstep));
  return hipGetLastError();
}

int pstats_read(pdr_ctx* c, int32_t* tstep, double* rec) {
  const PsLayout L = ps_layout();
  double h[kRecW + 4];
  if (!c->pst_buf) {
    if (tstep) *tstep = -1;
    if (rec)
      for (int q = 0; q < kRecW; ++q) rec[q] = 0.0;
    return PDR_OK;
  }
  if (hipMemcpyAsync(h, c->pst_buf + L.rec, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return set_error(PDR_ERR_HIP, "tracer statistics: copy failed");
  if (tstep) *tstep = int32_t(h[kRecW + 3]);
  if (rec)
    for (int q = 0; q < kRecW; ++q) rec[q] = h[q];
  return PDR_OK;
}

}  // namespace pdr
