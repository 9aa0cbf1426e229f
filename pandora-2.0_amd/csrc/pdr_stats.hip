// pdr_stats.hip -- wave-space statistics of the velocity on the device:
// f_FluidstatsFs (src/pf_fluidstats.f90:1293-1361) with the primary quantities
// of f_FluidstatsFsPrimaryQuants / ...BFRot (1423-1576, 1586-1792), the mode
// density f_FluidstatsModeDensity / ...BFRot (696-774, 1977-2049) and the
// secondary quantities f_FluidstatsFsSecondaryQuants (2491-2558).
//
// One pass over the state: every mode is taken to the laboratory frame
// (k_lab(n) = sum_m bmat(m,n) k(m), u_lab(m) = sum_n binv(m,n) u(n) over the
// existing entries; the identity for isotropic runs), binned by |k_lab| into
// [kk - 1/2, kk + 1/2), and its energy, component energies, dissipation and
// dissipation-skewness terms and its mode count (1 on i = 0, 2 elsewhere) are
// added to a per-workgroup LDS histogram.  A thread sums the terms of
// consecutive modes of one shell in registers (Neumaier) and adds each such run
// to the histogram.  The LDS adds are FP64 atomics that return the old value,
// so each thread recovers the exact rounding error of its add (TwoSum) and adds
// it to a second histogram: the per-workgroup sums are compensated.  Workgroup
// partials are then reduced in workgroup order with Neumaier summation, and the
// host finishes the kkmax-sized arrays in FP64.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "pdr_internal.hpp"
#include "pdr_rsstats.hpp"

namespace pdr {

void spectra_finish(const pdr_ctx* ctx, const std::vector<double>& raw, const double uu[6],
                    const std::vector<double>& s12, double b11, double* spec_1i, double* spec_2i,
                    pdr_homog_t* h);
void corr_finish(const pdr_ctx* ctx, const std::vector<double>& s12, double* r1i, double* lint11);

// spectrum bins (f_FluidstatsAlloc, src/pf_fluidstats.f90:348-374): kkmax =
// nint(k_max) with k_max the largest laboratory |k| of the box, each index range
// divided by its stretch outside the isotropic form
int kkmax_of(const pdr_ctx* c) {
  const Geo& g = c->g;
  if (c->p.f_type == PDR_F_ISOTROPIC) return int(std::llround(std::sqrt(double(g.ki * g.ki + g.kj * g.kj + g.kk * g.kk))));
  const double b1 = box_stretch(c->p, 1), b2 = box_stretch(c->p, 2), b3 = box_stretch(c->p, 3);
  const double k2 = (double(g.ki * g.ki) / (b1 * b1) + double(g.kj * g.kj) / (b2 * b2)) + double(g.kk * g.kk) / (b3 * b3);
  return int(std::llround(std::sqrt(k2)));
}

namespace {

constexpr int kStatThreads = 256;
constexpr int kStatBlocks = 1024;
#ifndef PDR_STATS_RUN
#define PDR_STATS_RUN 4
#endif
constexpr int kStatRun = PDR_STATS_RUN;

struct StatsFrame {
  double bm[9];  // masked bmat, [m][n]
  double bi[9];  // masked binv, [m][n]
  double nu;
  int kkmax;
  int nki;
  int lab_state;  // NS_R_ROT: the state holds laboratory velocities
                  // (f_FluidstatsFsPrimaryQuantsRRot, src/pf_fluidstats.f90:1801-1970)
};

// Width of the per-workgroup histogram: 13 shell arrays of kkmax, then the
// streamwise bins over i: hi and lo of sum |u1_lab|^2, of |u2_lab|^2, and of the
// real and imaginary parts of conj(u1_lab) u1 (the two-point correlation).
__host__ __device__ inline int stats_hist(int kkmax, int nki) { return 13 * kkmax + 8 * nki; }

__device__ __forceinline__ double wvn_s(int k, int kmax) {
  return (k <= kmax) ? double(k) : double(k - (2 * kmax + 1));
}

__device__ __forceinline__ void two_sum_add(double* hi, double* lo, double v) {
  const double old = atomicAdd(hi, v);
  const double t = old + v;  // the value the atomic stored
  const double bv = t - old;
  const double err = (old - (t - bv)) + (v - bv);
  if (err != 0.0) atomicAdd(lo, err);
}

// LDS: hist[13][kkmax] = hi sums of 6 quantities, their lo parts, mode counts;
// the streamwise bins s1 (hi, lo), s2 (hi, lo), cre (hi, lo), cim (hi, lo) of
// nki each; then 12 doubles per wave for the Reynolds-stress sums.
// Partial of workgroup b: part[b][stats_hist + 12].
template <typename T>
__global__ void __launch_bounds__(kStatThreads)
    k_stats(Geo g, StatsFrame sf, const cx<T>* __restrict__ u0, const cx<T>* __restrict__ u1,
            const cx<T>* __restrict__ u2, double* __restrict__ part) {
  // no contraction into FMAs: each mode's |k_lab| and terms are then rounded
  // exactly as in a plain FP64 evaluation, so shell membership near a bin
  // edge cannot depend on the compiler
#pragma clang fp contract(off)
  extern __shared__ double sh[];
  const int kkmax = sf.kkmax;
  const int nh = stats_hist(kkmax, sf.nki);
  double* s1 = sh + 13 * kkmax;
  double* red = sh + nh;
  for (int q = threadIdx.x; q < nh; q += kStatThreads) sh[q] = 0.0;
  __syncthreads();
  double uu[6] = {0, 0, 0, 0, 0, 0}, uc[6] = {0, 0, 0, 0, 0, 0};
  // Each workgroup takes a contiguous range of modes (the lanes of a wave still
  // read consecutive modes), so a thread's modes stay in one i plane for long
  // stretches: the per-i bins (streamwise spectra and the correlation sums) are
  // summed in registers (Neumaier) and added to the LDS bins only when the plane
  // changes, instead of one contended LDS atomic per mode.
  const long long per = (g.mw + gridDim.x - 1) / gridDim.x;
  const long long e_beg = (long long)blockIdx.x * per;
  const long long e_end = e_beg + per < g.mw ? e_beg + per : g.mw;
  int icur = -1;
  double pi_s[4] = {0, 0, 0, 0}, pi_c[4] = {0, 0, 0, 0};
  auto flush_i = [&]() {
    if (icur >= 0) {
#pragma unroll
      for (int a = 0; a < 4; ++a) two_sum_add(&s1[2 * a * sf.nki + icur], &s1[(2 * a + 1) * sf.nki + icur], pi_s[a] + pi_c[a]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      pi_s[a] = 0.0;
      pi_c[a] = 0.0;
    }
  };
  // Shell bins: a thread takes runs of kStatRun consecutive modes (consecutive
  // k3, so |k| moves slowly and neighbouring modes mostly share a shell); it sums
  // a shell's terms in registers (Neumaier) while the shell stays the same and
  // adds them to the LDS bins when it changes, instead of seven contended LDS
  // atomics per mode.
  int kcur = 0;
  double bs[7], bc[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) bs[q] = bc[q] = 0.0;
  auto flush_k = [&]() {
    if (kcur) {
      const int o = kcur - 1;
#pragma unroll
      for (int q = 0; q < 6; ++q) two_sum_add(&sh[q * kkmax + o], &sh[(6 + q) * kkmax + o], bs[q] + bc[q]);
      atomicAdd(&sh[12 * kkmax + o], bs[6]);  // mode count: integers, exact
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) bs[q] = bc[q] = 0.0;
  };
  auto mode = [&](int i, int j, int k, const cx<T>& a, const cx<T>& b, const cx<T>& c) {
    if (i != icur) {
      flush_i();
      icur = i;
    }
    const double kv[3] = {double(i), wvn_s(j, g.kj), wvn_s(k, g.kk)};
    const double um[3][2] = {{double(a.x), double(a.y)}, {double(b.x), double(b.y)}, {double(c.x), double(c.y)}};
    double kl[3] = {0, 0, 0}, ul[3][2] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        kl[n] = kl[n] + sf.bm[3 * m + n] * kv[m];
        ul[m][0] = ul[m][0] + sf.bi[3 * m + n] * um[n][0];
        ul[m][1] = ul[m][1] + sf.bi[3 * m + n] * um[n][1];
      }
    const double km = sqrt((kl[0] * kl[0]) + (kl[1] * kl[1]) + (kl[2] * kl[2]));
    // per i, over all stored modes: |u1_lab|^2 and |u2_lab|^2 (streamwise spectra;
    // f_FluidstatsFs2D maps with binv in every sheared form) and the two-point
    // correlation of the laboratory u1 with the stored u1 along x, conj(u1_lab) u1
    // (Wiener-Khinchin form of the real-space sum of f_FluidstatsRs2D,
    // src/pf_fluidstats.f90:933-945)
    {
      const double pv[4] = {ul[0][0] * ul[0][0] + ul[0][1] * ul[0][1], ul[1][0] * ul[1][0] + ul[1][1] * ul[1][1],
                            ul[0][0] * um[0][0] + ul[0][1] * um[0][1], ul[0][0] * um[0][1] - ul[0][1] * um[0][0]};
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double t = pi_s[a] + pv[a];
        pi_c[a] += (fabs(pi_s[a]) >= fabs(pv[a])) ? (pi_s[a] - t) + pv[a] : (pv[a] - t) + pi_s[a];
        pi_s[a] = t;
      }
    }
    const double(*v)[2] = sf.lab_state ? um : ul;
    double ph[6];
    ph[0] = v[0][0] * v[0][0] + v[0][1] * v[0][1];
    ph[1] = v[1][0] * v[1][0] + v[1][1] * v[1][1];
    ph[2] = v[2][0] * v[2][0] + v[2][1] * v[2][1];
    ph[3] = v[0][0] * v[1][0] + v[0][1] * v[1][1];
    ph[4] = v[0][0] * v[2][0] + v[0][1] * v[2][1];
    ph[5] = v[1][0] * v[2][0] + v[1][1] * v[2][1];
    if (i != 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) ph[q] = ph[q] * 2.0;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {  // Neumaier per thread
      const double t = uu[q] + ph[q];
      uc[q] += (fabs(uu[q]) >= fabs(ph[q])) ? (uu[q] - t) + ph[q] : (ph[q] - t) + uu[q];
      uu[q] = t;
    }
    const double ek = 0.5 * (ph[0] + ph[1] + ph[2]);
    const double dk = 2.0 * km * km * sf.nu * ek;
    const double dsk = 4.0 * km * km * km * km * sf.nu * sf.nu * ek;
    int kk = 0;
    const int kc = int(floor(km + 0.5));
    for (int kb = kc - 1; kb <= kc + 1; ++kb)
      if (kb >= 1 && kb <= kkmax && km >= double(kb) - 0.5 && km < double(kb) + 0.5) kk = kb;
    if (kk != kcur) {
      flush_k();
      kcur = kk;
    }
    if (kk) {
      const double tv[6] = {ek, 0.5 * ph[0], 0.5 * ph[1], 0.5 * ph[2], dk, dsk};
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const double t = bs[q] + tv[q];
        bc[q] += (fabs(bs[q]) >= fabs(tv[q])) ? (bs[q] - t) + tv[q] : (tv[q] - t) + bs[q];
        bs[q] = t;
      }
      bs[6] += (i != 0) ? 2.0 : 1.0;
    }
  };
  for (long long eb = e_beg + (long long)threadIdx.x * kStatRun; eb < e_end; eb += (long long)kStatThreads * kStatRun) {
    // the run's loads first, all in flight together; (i, j, k) from one
    // division, then stepped
    cx<T> ra[kStatRun], rb[kStatRun], rc[kStatRun];
#pragma unroll
    for (int r = 0; r < kStatRun; ++r) {
      const bool ok = eb + r < e_end;
      ra[r] = ok ? u0[eb + r] : cx<T>{T(0), T(0)};
      rb[r] = ok ? u1[eb + r] : cx<T>{T(0), T(0)};
      rc[r] = ok ? u2[eb + r] : cx<T>{T(0), T(0)};
    }
    const long long ij = eb / g.nkk;
    int k = int(eb - ij * g.nkk);
    int j = int(ij % g.nkj);
    int i = int(ij / g.nkj);
#pragma unroll
    for (int r = 0; r < kStatRun; ++r) {
      if (eb + r < e_end) mode(i, j, k, ra[r], rb[r], rc[r]);
      if (++k == g.nkk) {
        k = 0;
        if (++j == g.nkj) {
          j = 0;
          ++i;
        }
      }
    }
  }
  flush_i();
  flush_k();
  // the 12 Reynolds-stress sums: over each wave (butterfly, every lane ends with
  // the same sums), then over the waves in wave order
#pragma unroll
  for (int q = 0; q < 6; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      uu[q] += __shfl_xor(uu[q], off);
      uc[q] += __shfl_xor(uc[q], off);
    }
  }
  const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      red[q * (kStatThreads / 64) + wave] = uu[q];
      red[(6 + q) * (kStatThreads / 64) + wave] = uc[q];
    }
  }
  __syncthreads();
  double* out = part + (size_t)blockIdx.x * size_t(nh + 12);
  for (int q = threadIdx.x; q < nh; q += kStatThreads) out[q] = sh[q];
  if (threadIdx.x < 12) {
    double r = red[threadIdx.x * (kStatThreads / 64)];
    for (int w = 1; w < kStatThreads / 64; ++w) r += red[threadIdx.x * (kStatThreads / 64) + w];
    out[nh + threadIdx.x] = r;
  }
}

// out[q] = Neumaier sum over workgroups b = 0, 1, ... of part[b][q]
__global__ void __launch_bounds__(kStatThreads)
    k_stats_reduce(const double* __restrict__ part, int nblk, int width, double* __restrict__ out) {
  const int q = int(blockIdx.x * kStatThreads + threadIdx.x);
  if (q >= width) return;
  double s = 0.0, c = 0.0;
  // unrolled so that eight loads are in flight ahead of the dependent sums
#pragma unroll 8
  for (int b = 0; b < nblk; ++b) {
    const double v = part[(size_t)b * width + q];
    const double t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  out[q] = s + c;
}

template <typename T>
hipError_t stats_t(pdr_ctx* c, const StatsFrame& sf, int nblk, double* part, double* out) {
  const int width = stats_hist(sf.kkmax, sf.nki) + 12;
  const size_t lds = size_t(width - 12 + 12 * (kStatThreads / 64)) * sizeof(double);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_stats<T>, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  k_stats<T><<<dim3(unsigned(nblk)), kStatThreads, lds, c->stream>>>(
      c->g, sf, (const cx<T>*)c->u[0], (const cx<T>*)c->u[1], (const cx<T>*)c->u[2], part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_stats_reduce<<<dim3(unsigned((width + kStatThreads - 1) / kStatThreads)), kStatThreads, 0, c->stream>>>(
      part, nblk, width, out);
  return hipGetLastError();
}


StatsFrame make_stats_frame(const pdr_ctx* c) {
  StatsFrame sf;
  for (int m = 1; m <= 3; ++m)
    for (int n = 1; n <= 3; ++n) {
      sf.bm[3 * (m - 1) + (n - 1)] = T9(c->bval, m, n) ? T9(c->bmat, m, n) : 0.0;
      sf.bi[3 * (m - 1) + (n - 1)] = T9(c->bval, m, n) ? T9(c->binv, m, n) : 0.0;
    }
  sf.nu = c->p.f_nu;
  sf.kkmax = kkmax_of(c);
  sf.nki = c->g.nki;
  sf.lab_state = rogallo_form(c) ? 1 : 0;
  return sf;
}

bool stats_fits_lds(const pdr_ctx* c) {
  const size_t lds = size_t(stats_hist(kkmax_of(c), c->g.nki) + 12 * (kStatThreads / 64)) * sizeof(double);
  return lds <= c->lds_max;
}

// Workgroups of the statistics pass that the device runs at once (its LDS
// histogram admits three per CU at 512^3), or -1 if the runtime cannot say.
static int stats_resident_of(const pdr_ctx* c) {
  if (!stats_fits_lds(c)) return -1;
  const size_t lds = size_t(stats_hist(kkmax_of(c), c->g.nki) + 12 * (kStatThreads / 64)) * sizeof(double);
  const void* f = c->prec ? (const void*)k_stats<float> : (const void*)k_stats<double>;
  int per_cu = 0, dev = 0, ncu = 0;
  bool ok = lds <= 65536 || hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) == hipSuccess;
  ok = ok && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kStatThreads, lds) == hipSuccess && per_cu > 0;
  ok = ok && hipGetDevice(&dev) == hipSuccess &&
       hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0;
  if (!ok) {
    (void)hipGetLastError();  // leave no error for the next launch check to report
    return -1;
  }
  return per_cu * ncu;
}

// Workgroups of the statistics pass and the width of one partial row.  Every
// workgroup takes an equal range of modes, so the pass runs in one round: at
// most as many workgroups as the device holds at once (1024 at three per CU on
// 256 CUs would leave a second round a third full).
int stats_nblk(const pdr_ctx* c) {
  const long long want = (c->g.mw + 4LL * kStatThreads - 1) / (4LL * kStatThreads);
  if (c->stats_resident == 0) c->stats_resident = stats_resident_of(c);
  const long long cap = c->stats_resident > 0 ? std::min<long long>(kStatBlocks, c->stats_resident) : kStatBlocks;
  return int(want < cap ? (want < 1 ? 1 : want) : cap);
}
int stats_width(const pdr_ctx* c) { return stats_hist(kkmax_of(c), c->g.nki) + 12; }

// the reduced row h[stats_width] of the pass -> raw[7 * kkmax] (ken, ken1, ken2,
// ken3, dis, disskew, count), uu[6] and, on request, the streamwise bins
void stats_unpack(const pdr_ctx* c, const double* h, std::vector<double>& raw, double uu[6],
                  std::vector<double>* s12) {
  const int kkmax = kkmax_of(c);
  const int nki = c->g.nki;
  raw.assign(size_t(7) * kkmax, 0.0);
  for (int a = 0; a < 6; ++a)
    for (int q = 0; q < kkmax; ++q)
      raw[size_t(a * kkmax + q)] = h[size_t(a * kkmax + q)] + h[size_t((6 + a) * kkmax + q)];
  for (int q = 0; q < kkmax; ++q) raw[size_t(6 * kkmax + q)] = h[size_t(12 * kkmax + q)];
  const size_t nh = size_t(stats_hist(kkmax, nki));
  for (int a = 0; a < 6; ++a) uu[a] = h[nh + a] + h[nh + 6 + a];
  if (s12) {  // [s1 over i, s2 over i, Re and Im of the correlation sums over i]
    s12->assign(size_t(4) * nki, 0.0);
    const size_t o = size_t(13) * kkmax;
    for (int i = 0; i < nki; ++i)
      for (int a = 0; a < 4; ++a)
        (*s12)[size_t(a * nki + i)] = h[o + size_t(2 * a * nki + i)] + h[o + size_t((2 * a + 1) * nki + i)];
  }
}

// raw[7 * kkmax] (ken, ken1, ken2, ken3, dis, disskew, count) and uu[6]
int stats_raw(pdr_ctx* c, std::vector<double>& raw, double uu[6], std::vector<double>* s12 = nullptr) {
  if (!stats_fits_lds(c)) return set_error(PDR_ERR_UNSUPPORTED, "statistics: spectrum too long for the LDS");
  const StatsFrame sf = make_stats_frame(c);
  const int nblk = stats_nblk(c);
  const int width = stats_width(c);
  const size_t need = (size_t(nblk) + 1) * size_t(width) * sizeof(double);
  if (c->stat_bytes < need) {
    if (c->stat_buf) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(c->stat_buf);
      c->dev_bytes -= c->stat_bytes;
      c->stat_buf = nullptr;
      c->stat_bytes = 0;
    }
    if (hipMalloc(&c->stat_buf, need) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(PDR_ERR_NOMEM, "statistics: device allocation failed");
    }
    c->stat_bytes = need;
    c->dev_bytes += need;
  }
  double* part = (double*)c->stat_buf;
  double* out = part + size_t(nblk) * width;
  hipError_t e = c->prec ? stats_t<float>(c, sf, nblk, part, out) : stats_t<double>(c, sf, nblk, part, out);
  if (e != hipSuccess) return set_error(PDR_ERR_HIP, std::string("statistics kernels: ") + hipGetErrorString(e));
  std::vector<double> h(static_cast<size_t>(width));
  e = hipMemcpyAsync(h.data(), out, size_t(width) * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_error(PDR_ERR_HIP, std::string("statistics copy: ") + hipGetErrorString(e));
  stats_unpack(c, h.data(), raw, uu, s12);
  return PDR_OK;
}

// workgroup partials of the real-space moments: one per x-pass workgroup
long long rs_nblk(const pdr_ctx* c) {
  return c->spec ? (long long)xpass_tiles(c) * c->g.nz : (long long)c->g.ny * c->g.nz;
}

// Column q of the real-space partials: a sum (q < kRsSums, Neumaier per thread,
// then a fixed-order tree) or a maximum, over nblk workgroup partials.
__global__ void __launch_bounds__(kStatThreads)
    k_rs_reduce(const double* __restrict__ part, long long nblk, double* __restrict__ out) {
  __shared__ double sh[kStatThreads];
  const int q = int(blockIdx.x);
  const bool is_max = q >= kRsSums;
  double s = 0.0, c = 0.0;
  for (long long b = threadIdx.x; b < nblk; b += kStatThreads) {
    const double v = part[b * kRsWidth + q];
    if (is_max) {
      s = v > s ? v : s;
    } else {
      const double t = s + v;
      c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
      s = t;
    }
  }
  sh[threadIdx.x] = is_max ? s : s + c;
  __syncthreads();
  for (int w = kStatThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      const double a = sh[threadIdx.x], b = sh[threadIdx.x + w];
      sh[threadIdx.x] = is_max ? (b > a ? b : a) : a + b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = sh[0];
}

// First level for many partials: workgroup w folds rows [w per, (w + 1) per)
// into one row.  A thread takes whole rows (25 consecutive doubles, so a wave's
// loads are contiguous), Neumaier per column, then a fixed-order tree over the
// workgroup; the maxima by maximum.  k_rs_reduce then reduces the folded rows.
constexpr int kRsFold = 512;
__global__ void __launch_bounds__(kStatThreads)
    k_rs_fold(const double* __restrict__ part, long long nblk, long long per, double* __restrict__ out) {
  __shared__ double sh[kRsWidth * kStatThreads];
  const long long b0 = (long long)blockIdx.x * per;
  const long long b1 = b0 + per < nblk ? b0 + per : nblk;
  double s[kRsWidth], c[kRsSums];
#pragma unroll
  for (int q = 0; q < kRsWidth; ++q) s[q] = 0.0;
#pragma unroll
  for (int q = 0; q < kRsSums; ++q) c[q] = 0.0;
  for (long long b = b0 + threadIdx.x; b < b1; b += kStatThreads) {
    const double* row = part + b * kRsWidth;
#pragma unroll
    for (int q = 0; q < kRsWidth; ++q) {
      const double v = row[q];
      if (q >= kRsSums) {
        s[q] = v > s[q] ? v : s[q];
      } else {
        const double t = s[q] + v;
        c[q] += (fabs(s[q]) >= fabs(v)) ? (s[q] - t) + v : (v - t) + s[q];
        s[q] = t;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kRsWidth; ++q) sh[q * kStatThreads + threadIdx.x] = q < kRsSums ? s[q] + c[q] : s[q];
  __syncthreads();
  for (int w = kStatThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
#pragma unroll
      for (int q = 0; q < kRsWidth; ++q) {
        const double a = sh[q * kStatThreads + threadIdx.x], b = sh[q * kStatThreads + threadIdx.x + w];
        sh[q * kStatThreads + threadIdx.x] = q >= kRsSums ? (b > a ? b : a) : a + b;
      }
    }
    __syncthreads();
  }
  if (int(threadIdx.x) < kRsWidth) out[(long long)blockIdx.x * kRsWidth + threadIdx.x] = sh[threadIdx.x * kStatThreads];
}

// Real-space moments: the nblk partials (fold: room for kRsFold rows) -> out.
hipError_t rs_reduce(const double* part, long long nblk, double* fold, double* out, hipStream_t st) {
  if (nblk > 4 * kRsFold) {
    const long long per = (nblk + kRsFold - 1) / kRsFold;
    const long long nf = (nblk + per - 1) / per;
    k_rs_fold<<<dim3(unsigned(nf)), kStatThreads, 0, st>>>(part, nblk, per, fold);
    part = fold;
    nblk = nf;
  }
  k_rs_reduce<<<dim3(unsigned(kRsWidth)), kStatThreads, 0, st>>>(part, nblk, out);
  return hipGetLastError();
}

// float32 pi widened to double (ONEPI, src/pg_constants.f90:126)
const double kOnePiRef = double(3.14159265358979323846f);

}  // namespace
}  // namespace pdr

extern "C" {

}  // extern "C"

namespace pdr {
namespace {

// Mode-density normalisation and secondary quantities of f_FluidstatsFs, in
// place: sp[7 * kkmax] holds the raw shell sums (ken, ken1, ken2, ken3, dis,
// disskew, count) on entry and the normalised spectra and mode density on exit.
void stats_finish_inplace(const pdr_ctx* ctx, double* sp, const double uu[6], pdr_stats_t* q) {
  const int kkmax = pdr::kkmax_of(ctx);
  const double nu = ctx->p.f_nu;
  // normalisation by the mode density (src/pf_fluidstats.f90:1547-1575, 767-769)
  double f_ken = 0, f_dis = 0, f_ds = 0, fn[6] = {0, 0, 0, 0, 0, 0};
  for (int kk = 1; kk <= kkmax; ++kk) {
    const double up = double(kk) + 0.5, lo = double(kk) - 0.5;
    const double dens = sp[size_t(6 * kkmax + kk - 1)] / (4.0 / 3.0 * pdr::kOnePiRef * (up * up * up - lo * lo * lo));
    sp[size_t(6 * kkmax + kk - 1)] = dens;
    if (dens != 0.0) {
      f_ken = f_ken + sp[size_t(kk - 1)];
      f_dis = f_dis + sp[size_t(4 * kkmax + kk - 1)];
      f_ds = f_ds + sp[size_t(5 * kkmax + kk - 1)];
      for (int a = 0; a < 6; ++a) {
        sp[size_t(a * kkmax + kk - 1)] = sp[size_t(a * kkmax + kk - 1)] / dens;
        fn[a] = fn[a] + sp[size_t(a * kkmax + kk - 1)];
      }
    } else {
      for (int a = 0; a < 6; ++a) sp[size_t(a * kkmax + kk - 1)] = 0.0;
    }
  }
  q->tke = 0.5 * (uu[0] + uu[1] + uu[2]);
  q->u1u1 = uu[0];
  q->u2u2 = uu[1];
  q->u3u3 = uu[2];
  q->u1u2 = uu[3];
  q->u1u3 = uu[4];
  q->u2u3 = uu[5];
  q->f_ken = f_ken;
  q->f_dis = f_dis;
  q->f_disskew = f_ds;
  q->fn_ken = fn[0];
  q->fn_ken1 = fn[1];
  q->fn_ken2 = fn[2];
  q->fn_ken3 = fn[3];
  q->fn_dis = fn[4];
  q->fn_disskew = fn[5];
  // f_FluidstatsFsSecondaryQuants (src/pf_fluidstats.f90:2511-2556)
  const double fint = std::sqrt(2.0 * fn[0] / 3.0);
  q->fn_int = fint;
  q->fn_int1 = std::sqrt(2.0 * fn[1]);
  q->fn_int2 = std::sqrt(2.0 * fn[2]);
  q->fn_int3 = std::sqrt(2.0 * fn[3]);
  q->tn_tur = fn[0] / fn[4];
  q->ln_tur = std::pow(fn[0], 3.0 / 2.0) / fn[4];
  q->ln_tay = std::sqrt((15.0 * nu * (fint * fint)) / fn[4]);
  q->rn_tay = q->ln_tay * fint / nu;
  double li[4] = {0, 0, 0, 0};
  for (int kk = 1; kk <= kkmax; ++kk) {
    const double f = pdr::kOnePiRef / 2.0 / (fint * fint);
    for (int a = 0; a < 4; ++a) li[a] = li[a] + (f * sp[size_t(a * kkmax + kk - 1)] / double(kk));
  }
  q->ln_int = li[0];
  q->ln_int1 = li[1];
  q->ln_int2 = li[2];
  q->ln_int3 = li[3];
  q->tn_edy = li[0] / fint;
  q->ln_kol = std::pow(nu * nu * nu / fn[4], 1.0 / 4.0);
  q->tn_kol = std::sqrt(nu / fn[4]);
  q->disskew = fn[5] * std::pow(15.0 / fn[4], 3.0 / 2.0) * std::sqrt(nu) / 35.0;
  q->c_epsilon = fn[4] * li[0] / (fint * fint * fint);
}

void stats_finish(const pdr_ctx* ctx, const std::vector<double>& raw, const double uu[6], pdr_stats_t* q,
                  double* spec) {
  std::vector<double> sp(raw);
  stats_finish_inplace(ctx, sp.data(), uu, q);
  if (spec) std::copy(sp.begin(), sp.end(), spec);
}

}  // namespace
}  // namespace pdr

extern "C" {

int pdr_stats(pdr_ctx* ctx, pdr_stats_t* q, double* spec) {
  if (!ctx || !q) return pdr::set_error(PDR_ERR_ARG, "pdr_stats: null argument");
  std::vector<double> raw;
  double uu[6];
  int rc = pdr::stats_raw(ctx, raw, uu);
  if (rc != PDR_OK) return rc;
  pdr::stats_finish(ctx, raw, uu, q, spec);
  return PDR_OK;
}

// f_FluidstatsFs2D + f_FluidstatsFsCollect (src/pf_fluidstats.f90:1154-1283).
// The reference sums 2 |u_lab(i, y, z) / NX|^2 over the (y, z) points of the
// y- and z-transformed velocity (src/pf_rk3.f90:1504-1590); by Parseval for
// those unnormalised transforms this equals 2 NY NZ / NX^2 sum_{j,k} |u_lab(i,j,k)|^2,
// which the statistics pass accumulates per i from the wave state.
int pdr_spectra_1d(pdr_ctx* ctx, double* spec_1i, double* spec_2i, pdr_homog_t* h) {
  if (!ctx) return pdr::set_error(PDR_ERR_ARG, "pdr_spectra_1d: null context");
  std::vector<double> raw, s12;
  double uu[6];
  int rc = pdr::stats_raw(ctx, raw, uu, &s12);
  if (rc != PDR_OK) return rc;
  pdr::spectra_finish(ctx, raw, uu, s12, T9(ctx->binv, 1, 1), spec_1i, spec_2i, h);
  return PDR_OK;
}

// The TS_CORR two-point correlation of f_FluidstatsRs2D / RsCollect
// (src/pf_fluidstats.f90:933-945, 1092-1106) of the current state, from the
// per-i correlation bins of the statistics pass.
int pdr_rs_correlation(pdr_ctx* ctx, double* r1i, double* lint11) {
  if (!ctx || !r1i || !lint11) return pdr::set_error(PDR_ERR_ARG, "pdr_rs_correlation: null argument");
  std::vector<double> raw, s12;
  double uu[6];
  int rc = pdr::stats_raw(ctx, raw, uu, &s12);
  if (rc != PDR_OK) return rc;
  pdr::corr_finish(ctx, s12, r1i, lint11);
  return PDR_OK;
}

}  // extern "C"

namespace pdr {
namespace {

// rs_raw[kRsWidth] (the reduced real-space partials) -> the pdr_rs_stats_t record
// of f_FluidstatsRsCollect (src/pf_fluidstats.f90:992-1116); rs_rmsvort1 is
// not divided by the number of points, as there
void rs_finish(const pdr_ctx* c, const double* raw, pdr_rs_stats_t* out) {
  const pdr::Geo& g = c->g;
  const double n3 = (double(g.nx) * double(g.ny)) * double(g.nz);
  double uu[6], sk[3], fl[3], rms[3], skv[3], flv[3];
  for (int q = 0; q < 6; ++q) uu[q] = raw[q] / n3;
  for (int q = 0; q < 3; ++q) {
    sk[q] = (raw[6 + q] / n3) / std::pow(uu[q], 1.5);
    fl[q] = (raw[9 + q] / n3) / (uu[q] * uu[q]);
    rms[q] = (q == 0) ? raw[12] : raw[12 + q] / n3;
    rms[q] = std::pow(rms[q], 0.5);
    skv[q] = (raw[15 + q] / n3) / (rms[q] * rms[q] * rms[q]);
    flv[q] = (raw[18 + q] / n3) / ((rms[q] * rms[q]) * (rms[q] * rms[q]));
  }
  out->rs_u1u1 = uu[0];
  out->rs_u2u2 = uu[1];
  out->rs_u3u3 = uu[2];
  out->rs_u1u2 = uu[3];
  out->rs_u1u3 = uu[4];
  out->rs_u2u3 = uu[5];
  out->rs_skew1 = sk[0];
  out->rs_skew2 = sk[1];
  out->rs_skew3 = sk[2];
  out->rs_flat1 = fl[0];
  out->rs_flat2 = fl[1];
  out->rs_flat3 = fl[2];
  out->rs_rmsvort1 = rms[0];
  out->rs_rmsvort2 = rms[1];
  out->rs_rmsvort3 = rms[2];
  out->rs_skewvort1 = skv[0];
  out->rs_skewvort2 = skv[1];
  out->rs_skewvort3 = skv[2];
  out->rs_flatvort1 = flv[0];
  out->rs_flatvort2 = flv[1];
  out->rs_flatvort3 = flv[2];
  out->omegamax1 = raw[22];
  out->omegamax2 = raw[23];
  out->omegamax3 = raw[24];
  out->ucfl = raw[21];
}

}  // namespace

// streamwise 1-D spectra and the homogeneous-shear record from the unpacked
// statistics pass; b11 = binv(1,1) of the frame the pass ran in
void spectra_finish(const pdr_ctx* ctx, const std::vector<double>& raw, const double uu[6],
                    const std::vector<double>& s12, double b11, double* spec_1i, double* spec_2i,
                    pdr_homog_t* h) {
  const pdr::Geo& g = ctx->g;
  const bool shear = ctx->p.f_type != PDR_F_ISOTROPIC;
  const double f = 2.0 * (double(g.ny) * double(g.nz)) / (double(g.nx) * double(g.nx)) *
                   (shear ? (b11 * b11) * b11 : 1.0);
  std::vector<double> s1(size_t(g.nki)), s2(size_t(g.nki));
  for (int i = 0; i < g.nki; ++i) {
    s1[size_t(i)] = f * s12[size_t(i)];
    s2[size_t(i)] = f * s12[size_t(g.nki + i)];
  }
  if (spec_1i) std::copy(s1.begin(), s1.end(), spec_1i);
  if (spec_2i) std::copy(s2.begin(), s2.end(), spec_2i);
  if (h) {
    // f_FluidstatsHomogeneous (src/pf_fluidstats.f90:2568-2936); ln_int_21,
    // ln_int_12 and ln_int_13 are never assigned there, so they are zero and
    // L* = ln_int_11 / 0
    pdr_stats_t q;
    pdr::stats_finish(ctx, raw, uu, &q, nullptr);
    const double zero = 0.0;
    const double l11 = pdr::kOnePiRef * s1[0] / (2.0 * q.u1u1);
    h->sstar = q.tn_edy * ctx->p.f_deform_s;
    h->kstar = 2.0 * q.u1u1 / (q.u2u2 + q.u3u3);
    h->lstar = l11 / (2.0 * zero);
    h->ln_int1 = q.ln_int1;
    h->ln_int2 = q.ln_int2;
    h->ln_int3 = q.ln_int3;
    h->ln_int_11 = l11;
    h->ln_int_12 = 0.0;
    h->ln_int_13 = 0.0;
    h->fn_ken1 = q.fn_ken1;
    h->fn_ken2 = q.fn_ken2;
    h->fn_ken3 = q.fn_ken3;
    h->fn_int1 = q.fn_int1;
    h->fn_int2 = q.fn_int2;
    h->fn_int3 = q.fn_int3;
  }
}

// r1i(s) for s = 0..NX/2 and lint11 from the per-i sums C(i) = sum over (j, k)
// of conj(u1_lab) u1 and S(i) = sum |u1_lab|^2 of the statistics pass.  On the
// grid the real-space sums of f_FluidstatsRs2D are exact trigonometric sums:
// sum_x u1_lab(x) u1(x + s) = N^3 (Re C(0) + 2 sum_{i>=1} Re(C(i) e^{2 pi i i s / NX}))
// and sum_x u1_lab^2 = N^3 (S(0) + 2 sum_{i>=1} S(i)), so r1i(s) is their ratio
// and lint11 = dx_lab times the sum of the ratio over the reference loop's
// separations s = 0 .. (NX-1)/2 + 1 (for odd NX the last one is beyond the
// r1i array and enters lint11 only).
void corr_finish(const pdr_ctx* ctx, const std::vector<double>& s12, double* r1i, double* lint11) {
  const int nx = ctx->g.nx, nki = ctx->g.nki;
  const double* S = s12.data();
  const double* cre = S + 2 * nki;
  const double* cim = S + 3 * nki;
  double den = S[0];
  for (int i = 1; i < nki; ++i) den += 2.0 * S[i];
  const int smax = (nx - 1) / 2 + 1;
  const double twopi = 2.0 * 3.14159265358979323846;
  double lsum = 0.0;
  for (int s = 0; s <= smax; ++s) {
    double num = cre[0];
    for (int i = 1; i < nki; ++i) {
      const double th = twopi * double((static_cast<long long>(i) * s) % nx) / double(nx);
      num += 2.0 * (cre[i] * std::cos(th) - cim[i] * std::sin(th));
    }
    const double r = num / den;
    if (s <= nx / 2) r1i[s] = r;
    lsum += r;
  }
  // dx_laboratoryframe = TWOPI B110 / NODES_X (src/pg_domain.f90:178-183)
  *lint11 = lsum * ((2.0 * kOnePiRef * double(box_stretch(ctx->p, 1))) / double(nx));
}

}  // namespace pdr

extern "C" {

int pdr_rs_stats(pdr_ctx* ctx, pdr_rs_stats_t* out) {
  if (!ctx || !out) return pdr::set_error(PDR_ERR_ARG, "pdr_rs_stats: null argument");
  pdr_ctx* c = ctx;
  const pdr::Geo& g = c->g;
  const long long nblk = pdr::rs_nblk(c);
  const size_t need = size_t(nblk + 1 + pdr::kRsFold) * pdr::kRsWidth * sizeof(double);
  if (c->rs_bytes < need) {
    if (c->rs_buf) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(c->rs_buf);
      c->dev_bytes -= c->rs_bytes;
      c->rs_buf = nullptr;
      c->rs_bytes = 0;
    }
    if (hipMalloc(&c->rs_buf, need) != hipSuccess) {
      (void)hipGetLastError();
      return pdr::set_error(PDR_ERR_NOMEM, "pdr_rs_stats: device allocation failed");
    }
    c->rs_bytes = need;
    c->dev_bytes += need;
  }
  double* part = (double*)c->rs_buf;
  double* dres = part + size_t(nblk) * pdr::kRsWidth;
  // the real-space planes of the current state: pass 1, then the y and x
  // passes with the moments accumulated in the x pass (its outputs are scratch)
  const pdr::Coef cf = pdr::make_coef(c);
  hipError_t e = hipSuccess;
  if (pdr::conv_form(c)) {
    e = pdr::launch_conv_front(c, cf, false, false, part);
  } else {
    e = pdr::launch_zinv(c, cf);
    for (int z0 = 0; e == hipSuccess && z0 < g.nz; z0 += g.zc) {
      const int nzc = std::min(g.zc, g.nz - z0);
      e = pdr::launch_yinv(c, z0, nzc);
      if (e == hipSuccess) e = pdr::launch_xpass(c, cf, false, false, z0, nzc, part);
    }
  }
  if (e == hipSuccess) e = pdr::rs_reduce(part, nblk, dres + pdr::kRsWidth, dres, c->stream);
  double raw[pdr::kRsWidth];
  if (e == hipSuccess) e = hipMemcpyAsync(raw, dres, sizeof(raw), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return pdr::set_error(PDR_ERR_HIP, std::string("pdr_rs_stats: ") + hipGetErrorString(e));
  pdr::rs_finish(c, raw, out);
  return PDR_OK;
}

int pdr_spectrum(pdr_ctx* ctx, double* ek, double* dk, double* dsk, double tot[6]) {
  if (!ctx || !ek || !dk || !dsk || !tot) return pdr::set_error(PDR_ERR_ARG, "pdr_spectrum: null argument");
  std::vector<double> raw;
  double uu[6];
  int rc = pdr::stats_raw(ctx, raw, uu);
  if (rc != PDR_OK) return rc;
  const int kkmax = pdr::kkmax_of(ctx);
  double fk = 0, fd = 0, fs = 0;
  for (int q = 0; q < kkmax; ++q) {
    ek[q] = raw[size_t(q)];
    dk[q] = raw[size_t(4 * kkmax + q)];
    dsk[q] = raw[size_t(5 * kkmax + q)];
    if (raw[size_t(6 * kkmax + q)] != 0.0) {
      fk += ek[q];
      fd += dk[q];
      fs += dsk[q];
    }
  }
  tot[0] = fk;
  tot[1] = fd;
  tot[2] = fs;
  tot[3] = uu[0];
  tot[4] = uu[1];
  tot[5] = uu[2];
  return PDR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Statistics of stage 1 inside the step (f_RK3Stage1 and f_RK3RealSpace,
// src/pf_rk3.f90:58-60, 1073-1078, 1124-1143): the wave-space pass of the state
// after the remesh check, and the real-space moments accumulated by the stage's
// own x pass.  Both reduce into a record buffer the context owns; nothing is
// allocated or copied to the host during the step.
// Record buffer: [wave partials nblk x width][wave row width][real-space
// partials rs_nblk x kRsWidth][real-space row kRsWidth][kRsFold x kRsWidth of
// folded real-space partials].
// ---------------------------------------------------------------------------
namespace pdr {

static size_t sst_need(const pdr_ctx* c) {
  return (size_t(stats_nblk(c)) + 1) * size_t(stats_width(c)) * sizeof(double) +
         (size_t(rs_nblk(c)) + 1 + kRsFold) * size_t(kRsWidth) * sizeof(double);
}

static double* sst_wave_part(pdr_ctx* c) { return (double*)c->sst_buf; }
static double* sst_wave_row(pdr_ctx* c) { return sst_wave_part(c) + size_t(stats_nblk(c)) * stats_width(c); }
double* stage_stats_rs_partials(pdr_ctx* c) { return sst_wave_row(c) + stats_width(c); }
static double* sst_rs_row(pdr_ctx* c) { return stage_stats_rs_partials(c) + size_t(rs_nblk(c)) * kRsWidth; }

bool stage_stats_due(const pdr_ctx* c, int stage, int tstep) {
  return stage == 1 && c->sst_every > 0 && c->sst_buf && (tstep - 1) % c->sst_every == 0;
}

hipError_t stage_stats_wave(pdr_ctx* c, int tstep) {
  const StatsFrame sf = make_stats_frame(c);
  c->sst_tstep = tstep - 1;
  c->sst_b11 = T9(c->binv, 1, 1);
  return c->prec ? stats_t<float>(c, sf, stats_nblk(c), sst_wave_part(c), sst_wave_row(c))
                 : stats_t<double>(c, sf, stats_nblk(c), sst_wave_part(c), sst_wave_row(c));
}

hipError_t stage_stats_rs_reduce(pdr_ctx* c) {
  return rs_reduce(stage_stats_rs_partials(c), rs_nblk(c), sst_rs_row(c) + kRsWidth, sst_rs_row(c), c->stream);
}

void stage_stats_free(pdr_ctx* c) {
  if (c->sst_buf) {
    (void)hipFree(c->sst_buf);
    c->dev_bytes -= c->sst_bytes;
  }
  c->sst_buf = nullptr;
  c->sst_bytes = 0;
}

}  // namespace pdr

extern "C" {

int pdr_set_stage_stats(pdr_ctx* c, int32_t every) {
  if (!c || every < 0) return pdr::set_error(PDR_ERR_ARG, "pdr_set_stage_stats: bad argument");
  if (every > 0 && !pdr::stats_fits_lds(c))
    return pdr::set_error(PDR_ERR_UNSUPPORTED, "statistics: spectrum too long for the LDS");
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return pdr::set_error(PDR_ERR_HIP, "pdr_set_stage_stats: stream synchronisation failed");
  c->sst_every = every;
  c->sst_tstep = -1;
  if (every == 0) {
    pdr::stage_stats_free(c);
    return PDR_OK;
  }
  const size_t need = pdr::sst_need(c);
  if (c->sst_bytes < need) {
    pdr::stage_stats_free(c);
    if (hipMalloc(&c->sst_buf, need) != hipSuccess) {
      (void)hipGetLastError();
      c->sst_buf = nullptr;
      c->sst_every = 0;
      return pdr::set_error(PDR_ERR_NOMEM, "pdr_set_stage_stats: device allocation failed");
    }
    c->sst_bytes = need;
    c->dev_bytes += need;
  }
  return PDR_OK;
}

int pdr_stage_stats(pdr_ctx* c, int32_t* tstep, pdr_stats_t* fs, double* spec, pdr_rs_stats_t* rs,
                    double* spec_1i, double* spec_2i, pdr_homog_t* h) {
  if (!c) return pdr::set_error(PDR_ERR_ARG, "null context");
  if (tstep) *tstep = c->sst_tstep;
  if (c->sst_tstep < 0) return PDR_OK;
  const int width = pdr::stats_width(c);
  std::vector<double> row(size_t(width) + pdr::kRsWidth);
  hipError_t e = hipMemcpyAsync(row.data(), pdr::sst_wave_row(c), size_t(width) * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(row.data() + width, pdr::sst_rs_row(c), pdr::kRsWidth * sizeof(double),
                       hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return pdr::set_error(PDR_ERR_HIP, std::string("pdr_stage_stats: ") + hipGetErrorString(e));
  std::vector<double> raw, s12;
  double uu[6];
  pdr::stats_unpack(c, row.data(), raw, uu, &s12);
  if (fs || spec) {
    pdr_stats_t q;
    pdr::stats_finish(c, raw, uu, &q, spec);
    if (fs) *fs = q;
  }
  if (spec_1i || spec_2i || h) pdr::spectra_finish(c, raw, uu, s12, c->sst_b11, spec_1i, spec_2i, h);
  if (rs) pdr::rs_finish(c, row.data() + width, rs);
  return PDR_OK;
}

// The TS_CORR two-point correlation of the stage-1 record (the state that
// stage 1 of the record's step started from), as pdr_rs_correlation.
int pdr_stage_correlation(pdr_ctx* c, int32_t* tstep, double* r1i, double* lint11) {
  if (!c || !r1i || !lint11) return pdr::set_error(PDR_ERR_ARG, "pdr_stage_correlation: null argument");
  if (tstep) *tstep = c->sst_tstep;
  if (c->sst_tstep < 0) return PDR_OK;
  const int width = pdr::stats_width(c);
  std::vector<double> row(static_cast<size_t>(width));
  hipError_t e = hipMemcpyAsync(row.data(), pdr::sst_wave_row(c), size_t(width) * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return pdr::set_error(PDR_ERR_HIP, std::string("pdr_stage_correlation: ") + hipGetErrorString(e));
  std::vector<double> raw, s12;
  double uu[6];
  pdr::stats_unpack(c, row.data(), raw, uu, &s12);
  pdr::corr_finish(c, s12, r1i, lint11);
  return PDR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free wave-space statistics (pdr_stats_noalloc).  The workgroup
// partials, the reduced row and the folded shell sums live in `scratch`, a
// fixed device array of kScratchDoubles doubles (16 MiB) that every GPU the
// library runs on holds until the process ends.  Its device address is looked
// up once per device with hipGetSymbolAddress (with that device current) and
// kept in g_scratch; the kernels take that address, never the array's name.
// Every call on a device shares that device's array and leaves the last
// call's contents there: each call writes every element it later reads
// (the partial rows, then the row, then the folded sums), so nothing needs
// clearing, and calls that overlap on one device race.
// ---------------------------------------------------------------------------
namespace pdr {
constexpr size_t kScratchDoubles = size_t(2) << 20;
}  // namespace pdr

static __device__ __align__(256) double scratch[pdr::kScratchDoubles];

namespace pdr {
namespace {

constexpr int kScratchMaxDevices = 64;
double* g_scratch[kScratchMaxDevices] = {};
std::mutex g_scratch_mu;

// this device's copy of scratch (the current device is the context's)
hipError_t scratch_ptr(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kScratchMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

// the reduced row h[stats_hist + 12] -> raw[7 * kkmax] (hi + lo of the six shell
// sums, then the mode counts) and uu[6] at raw + 7 * kkmax: stats_unpack's
// additions on the device
__global__ void __launch_bounds__(kStatThreads)
    k_stats_fold(const double* __restrict__ h, int kkmax, int nki, double* __restrict__ raw) {
  const int q = int(blockIdx.x * kStatThreads + threadIdx.x);
  if (q < 6 * kkmax) {
    raw[q] = h[q] + h[6 * kkmax + q];
  } else if (q < 7 * kkmax) {
    raw[q] = h[6 * kkmax + q];
  } else if (q < 7 * kkmax + 6) {
    const int nh = stats_hist(kkmax, nki);
    const int a = q - 7 * kkmax;
    raw[q] = h[nh + a] + h[nh + 6 + a];
  }
}

}  // namespace
}  // namespace pdr

extern "C" {

int pdr_stats_noalloc(pdr_ctx* ctx, pdr_stats_t* q, double* spec) {
  if (!ctx || !q || !spec)
    return pdr::set_error(PDR_ERR_ARG, "pdr_stats_noalloc: null argument (spec is required: it is the working buffer)");
  if (!pdr::stats_fits_lds(ctx)) return pdr::set_error(PDR_ERR_UNSUPPORTED, "statistics: spectrum too long for the LDS");
  const int kkmax = pdr::kkmax_of(ctx);
  const long long width = pdr::stats_width(ctx);
  const long long tail = 7LL * kkmax + 6;
  // rows of workgroup partials that fit beside the reduced row and the folded sums
  const long long rows = (static_cast<long long>(pdr::kScratchDoubles) - width - tail) / width;
  if (rows < 1)
    return pdr::set_error(PDR_ERR_UNSUPPORTED,
                          "pdr_stats_noalloc: the statistics of this box need " + std::to_string(2 * width + tail) +
                              " doubles of scratch, more than its maximum of " +
                              std::to_string(pdr::kScratchDoubles));
  const int nblk = int(std::min<long long>(pdr::stats_nblk(ctx), rows));
  double* base = nullptr;
  hipError_t e = pdr::scratch_ptr(&base);
  if (e != hipSuccess) return pdr::set_error(PDR_ERR_HIP, std::string("pdr_stats_noalloc: scratch address: ") + hipGetErrorString(e));
  double* part = base;
  double* row = part + size_t(nblk) * size_t(width);
  double* fold = row + width;
  const pdr::StatsFrame sf = pdr::make_stats_frame(ctx);
  e = ctx->prec ? pdr::stats_t<float>(ctx, sf, nblk, part, row) : pdr::stats_t<double>(ctx, sf, nblk, part, row);
  if (e == hipSuccess) {
    pdr::k_stats_fold<<<dim3(unsigned((tail + pdr::kStatThreads - 1) / pdr::kStatThreads)), pdr::kStatThreads, 0,
                        ctx->stream>>>(row, kkmax, ctx->g.nki, fold);
    e = hipGetLastError();
  }
  double uu[6];
  if (e == hipSuccess)
    e = hipMemcpyAsync(spec, fold, size_t(7) * kkmax * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(uu, fold + size_t(7) * kkmax, sizeof(uu), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return pdr::set_error(PDR_ERR_HIP, std::string("pdr_stats_noalloc: ") + hipGetErrorString(e));
  pdr::stats_finish_inplace(ctx, spec, uu, q);
  return PDR_OK;
}

}  // extern "C"
