// pdr_kernels2.hip -- length-specialised pass kernels (the default path for
// grid sizes in PDR_SIZES).  Same work split and data layouts as pdr_kernels.hip
// (pass 1 k_zinv ... pass 5 k_zfwd); see that file for the reference routines
// each pass replaces.  Differences: compile-time transform length and tile
// height, in-place LDS transforms, twiddles in LDS, 128-byte tile chunks on
// 16-element-padded strides, XCD-aware tile order, several x rows per block.
#include <type_traits>

#include "pdr_fft2.hpp"
#include "pdr_internal.hpp"
#include "pdr_rsstats.hpp"
#include "pdr_formf.hpp"

namespace pdr {

namespace {

__device__ __forceinline__ double wvn2(int k, int kmax) {
  return (k <= kmax) ? double(k) : double(k - (2 * kmax + 1));
}

template <typename T>
__device__ __forceinline__ cx<T> czero() {
  return mk<T>(T(0), T(0));
}

// Tile shapes.  Lines per tile of the y/z kernels and rows per block of the x
// kernel; the build may override the FP32 / FP64 values for lengths <= 512
// (PDR_LINES_F32, PDR_LINES_F64, PDR_XROWS_F32, PDR_XROWS_F64) and the FP32
// heights of longer lines (PDR_LINES_MID_F32 and PDR_LINES_MID_FWD_F32 for
// 512 < n <= 1024, PDR_LINES_LONG_F32 and PDR_LINES_LONG_FWD_F32 above) to tune.
// Longer lines keep at least 8 lines in the inverse kernels, whose global
// accesses are runs of one element of each line (64-byte runs at 8 FP32 lines;
// 1.45 against 1.26 steps/s at 1536^3 with 4), while the forward kernels of
// lines above 1024 run faster with the two workgroups per CU that 4 lines allow.
#ifndef PDR_LINES_F32
#define PDR_LINES_F32 16
#endif
#ifndef PDR_LINES_F64
#define PDR_LINES_F64 8
#endif
#ifndef PDR_XROWS_F32
#define PDR_XROWS_F32 2
#endif
#ifndef PDR_XROWS_F64
#define PDR_XROWS_F64 2
#endif
#ifndef PDR_LINES_MID_F32
#define PDR_LINES_MID_F32 8
#endif
#ifndef PDR_LINES_LONG_F32
#define PDR_LINES_LONG_F32 8
#endif
template <typename T>
constexpr int tile_lines(int n) {
  return sizeof(T) == 4 ? (n <= 512 ? PDR_LINES_F32 : (n <= 1024 ? PDR_LINES_MID_F32 : PDR_LINES_LONG_F32))
                        : (n <= 512 ? PDR_LINES_F64 : (n <= 1024 ? 4 : 2));
}
// forward y/z kernels (3 fields per tile) may use their own FP32 tile height
#ifndef PDR_LINES_FWD_F32
#define PDR_LINES_FWD_F32 8
#endif
#ifndef PDR_LINES_MID_FWD_F32
#define PDR_LINES_MID_FWD_F32 8
#endif
#ifndef PDR_LINES_LONG_FWD_F32
#define PDR_LINES_LONG_FWD_F32 4
#endif
template <typename T>
constexpr int tile_lines_fwd(int n) {
  return sizeof(T) != 4 ? tile_lines<T>(n)
                        : (n <= 512 ? PDR_LINES_FWD_F32 : (n <= 1024 ? PDR_LINES_MID_FWD_F32 : PDR_LINES_LONG_FWD_F32));
}
// y passes of short lines: a workgroup takes PP whole z planes (all NKI
// i-lines of each) instead of a tile of i-lines of one plane, so that it
// moves about three thousand points and no tile is left part-empty
// (PP = 0: tiles of tile_lines i-lines within one plane).
template <typename T>
constexpr int y_planes(int n) {
  return n <= 32 ? 8 : (n <= 48 ? 4 : (n <= 64 ? 2 : (n <= 96 ? 1 : 0)));
}
template <typename T>
constexpr int y_lines(int n) {
  return y_planes<T>(n) ? y_planes<T>(n) * (n / 3 + 1) : tile_lines<T>(n);
}
template <typename T>
constexpr int y_lines_fwd(int n) {
  return y_planes<T>(n) ? y_planes<T>(n) * (n / 3 + 1) : tile_lines_fwd<T>(n);
}

// Short lines take more rows per workgroup, so that a workgroup still moves
// about a thousand points (2 rows of 512; 16 rows of 64).
template <typename T>
constexpr int xrows(int n) {
  return n <= 64 ? 16 : (n <= 128 ? 8 : (n <= 256 ? 4 : (sizeof(T) == 4 ? (n <= 512 ? PDR_XROWS_F32 : 2)
                                                                        : (n <= 512 ? PDR_XROWS_F64 : 2))));
}

// LDS bytes of a workgroup, and the waves per SIMD that the 160 KB of LDS of a
// CU admit (each 256-thread workgroup puts one wave on each of the 4 SIMDs).
// The kernels declare that occupancy, capped at two, to the register
// allocator: it then keeps within 256 registers per lane (without the hint it
// takes more, and one wave per SIMD), and still uses fewer where it can.
constexpr int kLdsPerCu = 160 * 1024;
template <typename T>
constexpr int lds_lines_bytes(int n) {
  return (n + tile_lines<T>(n) * lds_ld<T>(n)) * int(sizeof(cx<T>));
}
template <typename T>
constexpr int lds_fwd_bytes(int n) {
  return (n + tile_lines_fwd<T>(n) * lds_ld<T>(n)) * int(sizeof(cx<T>));
}
template <typename T>
constexpr int lds_y_bytes(int n) {
  return (n + y_lines<T>(n) * lds_ld<T>(n)) * int(sizeof(cx<T>));
}
template <typename T>
constexpr int lds_yf_bytes(int n) {
  return (n + y_lines_fwd<T>(n) * lds_ld<T>(n)) * int(sizeof(cx<T>));
}
// The x tile of lines longer than PDR_X_TWG_MIN points reads its twiddles from
// the global table (through the caches) instead of staging them in LDS: that
// frees N complex of LDS, so two 2-row FP32 tiles of 1536 points fit one CU
// and the FP64 one fits at all.
#ifndef PDR_X_TWG_MIN
#define PDR_X_TWG_MIN 1024
#endif
template <typename T>
constexpr bool x_tw_global(int n) { return n > PDR_X_TWG_MIN; }
template <typename T>
constexpr int lds_x_bytes(int n) {
  return ((x_tw_global<T>(n) ? 0 : n) + 3 * xrows<T>(n) * lds_ld<T>(n)) * int(sizeof(cx<T>)) +
         (kBlock2 / 64) * int(sizeof(T));
}
// Streaming (non-temporal) stores per kernel, bit mask: 1 zinv, 2 yinv,
// 4 xpass, 8 yfwd, 16 zfwd.  Every pass writes data that only the next pass
// reads, long after the caches have turned over: streaming stores keep the
// caches for the re-read inputs (measured 73.1 -> 75.4 steps/s at 512^3).
#ifndef PDR_NT_WRITES
#define PDR_NT_WRITES 31
#endif
// Streaming loads of single-use inputs, bit mask: 1 zinv (register form),
// 2 yinv, 4 xpass, 8 yfwd, 16 zfwd.
#ifndef PDR_NT_READS
#define PDR_NT_READS 0
#endif
constexpr bool kNrZinv = (PDR_NT_READS & 1) != 0;
constexpr bool kNrYinv = (PDR_NT_READS & 2) != 0;
constexpr bool kNrXpass = (PDR_NT_READS & 4) != 0;
constexpr bool kNrYfwd = (PDR_NT_READS & 8) != 0;
constexpr bool kNrZfwd = (PDR_NT_READS & 16) != 0;
constexpr bool kNtZinv = (PDR_NT_WRITES & 1) != 0;
constexpr bool kNtYinv = (PDR_NT_WRITES & 2) != 0;
constexpr bool kNtXpass = (PDR_NT_WRITES & 4) != 0;
constexpr bool kNtYfwd = (PDR_NT_WRITES & 8) != 0;
constexpr bool kNtZfwd = (PDR_NT_WRITES & 16) != 0;

[[maybe_unused]] constexpr int waves_for(int lds) {
  return kLdsPerCu / lds < 2 ? 1 : 2;
}
#define PDR_LINE_WAVES __attribute__((amdgpu_waves_per_eu(waves_for(lds_lines_bytes<T>(N)))))
#define PDR_FWD_WAVES __attribute__((amdgpu_waves_per_eu(waves_for(lds_fwd_bytes<T>(N)))))
#define PDR_X_WAVES __attribute__((amdgpu_waves_per_eu(waves_for(lds_x_bytes<T>(N)))))
#define PDR_Y_WAVES __attribute__((amdgpu_waves_per_eu(waves_for(lds_y_bytes<T>(N)))))
#define PDR_YF_WAVES __attribute__((amdgpu_waves_per_eu(waves_for(lds_yf_bytes<T>(N)))))

// Line kernels issue all global loads of a field (QE elements per thread)
// before the LDS stores of that field, so that one HBM latency is exposed per
// field rather than one per element.  (Loading field f+1 during the transform
// of field f was measured slower: it needs more registers than the occupancy
// leaves and spills.)

// ---------------------------------------------------------------------------
// pass 1: vorticity + inverse z FFT, (i, j-tile) per block
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__(kBlock2) PDR_LINE_WAVES
    k2_zinv(Geo g, Coef cf, const cx<T>* __restrict__ u0, const cx<T>* __restrict__ u1,
            const cx<T>* __restrict__ u2, cx<T>* __restrict__ w1, const cx<T>* __restrict__ twg, int ntile) {
  constexpr int L = tile_lines<T>(N);
  constexpr int LD = lds_ld<T>(N);
  constexpr int K = N / 3;
  constexpr int NKK = 2 * K + 1;
  constexpr int QE = (L * N + kBlock2 - 1) / kBlock2;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
  cx<T>* buf = tw + N;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = int(bid % unsigned(ntile));
  const int i = int(bid / unsigned(ntile));
  const int j0 = tile * L;
  const int nkj = g.nkj;
  load_tw<T, N>(tw, twg);
  const double wi = double(i);
  bool need[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const double* G = &cf.G[(c * 3 + n) * 3];
      need[c][n] = (G[0] != 0.0) || (G[1] != 0.0) || (G[2] != 0.0);
    }
  const long long base = ((long long)i * nkj + j0) * NKK;
  const cx<T>* uu[3] = {u0 + base, u1 + base, u2 + base};
  // velocity fields: QE elements per thread, loaded in one batch
  cx<T> pre[QE];
  auto load_u = [&](int f) {
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + kBlock2 * q;
      const int line = idx / N;
      const int z = idx - line * N;
      const int k = (z <= K) ? z : ((z >= N - K) ? z - (N - NKK) : -1);
      pre[q] = ((L * N % kBlock2 == 0 || idx < L * N) && j0 + line < nkj && k >= 0) ? uu[f][line * NKK + k]
                                                                                    : czero<T>();
    }
  };
  // vorticity fields need two or three components per element: computed and
  // stored to LDS in chunks of VC elements to bound the registers in flight
  constexpr int VC0 = sizeof(T) == 4 ? 8 : 4;
  constexpr int VC = QE < VC0 ? QE : VC0;
  auto fill_w = [&](int c) {
#pragma unroll 1
    for (int q0 = 0; q0 < QE; q0 += VC) {
      cx<T> a[VC][3];
#pragma unroll
      for (int qq = 0; qq < VC; ++qq) {
        const int idx = int(threadIdx.x) + kBlock2 * (q0 + qq);
        const int line = idx / N;
        const int z = idx - line * N;
        const int k = (z <= K) ? z : ((z >= N - K) ? z - (N - NKK) : -1);
        const bool ok = (QE % VC == 0 || q0 + qq < QE) && (L * N % kBlock2 == 0 || idx < L * N) &&
                        j0 + line < nkj && k >= 0;
#pragma unroll
        for (int n = 0; n < 3; ++n) a[qq][n] = (ok && need[c][n]) ? uu[n][line * NKK + k] : czero<T>();
      }
#pragma unroll
      for (int qq = 0; qq < VC; ++qq) {
        const int idx = int(threadIdx.x) + kBlock2 * (q0 + qq);
        if ((QE % VC == 0 || q0 + qq < QE) && (L * N % kBlock2 == 0 || idx < L * N)) {
          const int line = idx / N;
          const int z = idx - line * N;
          const int k = (z <= K) ? z : ((z >= N - K) ? z - (N - NKK) : 0);
          const double kv[3] = {wi, wvn2(j0 + line, g.kj), wvn2(k, K)};
          cx<T> acc = czero<T>();
#pragma unroll
          for (int n = 0; n < 3; ++n) {
            if (need[c][n]) {
              const double* G = &cf.G[(c * 3 + n) * 3];
              acc = add(acc, scl(a[qq][n], T(G[0] * kv[0] + G[1] * kv[1] + G[2] * kv[2])));
            }
          }
          buf[line * LD + lds_off<T>(z)] = mk<T>(-acc.y, acc.x);
        }
      }
    }
  };
  for (int f = 0; f < 6; ++f) {
    if (f < 3) {
      load_u(f);
#pragma unroll
      for (int q = 0; q < QE; ++q) {
        const int idx = int(threadIdx.x) + kBlock2 * q;
        if (L * N % kBlock2 == 0 || idx < L * N) {
          const int line = idx / N;
          buf[line * LD + lds_off<T>(idx - line * N)] = pre[q];
        }
      }
    } else {
      fill_w(f - 3);
    }
    __syncthreads();
    fft_ip<T, N, L, LD, +1>(buf, tw);
    cx<T>* dst = w1 + (long long)f * g.m1 + (long long)i * N * g.ldj + j0;
    for (int idx = threadIdx.x; idx < L * N; idx += kBlock2) {
      const int z = idx / L;
      const int line = idx - z * L;
      if (j0 + line < nkj) put<kNtZinv>(&dst[z * g.ldj + line], buf[line * LD + lds_off<T>(z)]);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pass 1, register form: the three velocity fields of a tile are loaded once
// and kept in registers; the vorticity fields are formed from them, so the
// velocity is read from HBM exactly once.  Narrower tiles (PDR_ZINV_LINES)
// keep the registers per thread within two waves per SIMD.
// ---------------------------------------------------------------------------
#ifndef PDR_ZINV_LINES_F32
#define PDR_ZINV_LINES_F32 8
#endif
#ifndef PDR_ZINV_LINES_F64
#define PDR_ZINV_LINES_F64 4
#endif
#ifndef PDR_ZINV_REG
#define PDR_ZINV_REG 1
#endif
template <typename T>
constexpr int zinv_lines(int n) {
  return n <= 512 ? (sizeof(T) == 4 ? PDR_ZINV_LINES_F32 : PDR_ZINV_LINES_F64) : tile_lines<T>(n);
}
template <typename T>
constexpr int lds_zinv_bytes(int n) {
  return (n + zinv_lines<T>(n) * lds_ld<T>(n)) * int(sizeof(cx<T>));
}
#define PDR_ZINV_WAVES __attribute__((amdgpu_waves_per_eu(waves_for(lds_zinv_bytes<T>(N)))))
// The single-box register z inverse runs 512-thread workgroups over twice the
// lines (16 FP32 / 8 FP64): each thread keeps the same velocity elements as
// with 256 threads and 8 / 4 lines, and occupancy stays two waves per SIMD, but
// a tile's store of one z row is a whole 128-byte line of the 1r2w array
// instead of half of one (which the neighbouring tile wrote at another time).
// Batched ensembles keep the 256-thread form.
#ifndef PDR_ZINV_THREADS
#define PDR_ZINV_THREADS 512
#endif
template <typename T, int ZT>
constexpr int zinv_lines_t(int n) {
  return zinv_lines<T>(n) * (ZT / kBlock2);
}
template <typename T, int ZT>
constexpr int lds_zinv_bytes_t(int n) {
  return (n + zinv_lines_t<T, ZT>(n) * lds_ld<T>(n)) * int(sizeof(cx<T>));
}
#define PDR_ZINV_WAVES_T \
  __attribute__((amdgpu_waves_per_eu(PDR_ZINV_THREADS > kBlock2 ? PDR_ZINV_THREADS / kBlock2 : waves_for(lds_zinv_bytes<T>(N)))))

template <typename T, int N, int ZT = kBlock2>
__device__ __forceinline__ void k2_zinv_r_body(unsigned bid, const Geo& g, const Coef& cf, const cx<T>* __restrict__ u0, const cx<T>* __restrict__ u1,
              const cx<T>* __restrict__ u2, cx<T>* __restrict__ w1, const cx<T>* __restrict__ twg, int ntile) {
  constexpr int L = zinv_lines_t<T, ZT>(N);
  constexpr int LD = lds_ld<T>(N);
  constexpr int K = N / 3;
  constexpr int NKK = 2 * K + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
  cx<T>* buf = tw + N;
  const int tile = int(bid % unsigned(ntile));
  const int i = int(bid / unsigned(ntile));
  const int j0 = tile * L;
  const int nkj = g.nkj;
  load_tw<T, N, ZT>(tw, twg);
  const long long base = ((long long)i * nkj + j0) * NKK;
  const cx<T>* uu[3] = {u0 + base, u1 + base, u2 + base};
  // A thread keeps the stored modes e = tid + ZT q of the tile (line e / NKK,
  // mode k = e % NKK): the tile's velocity is one contiguous run of L NKK modes
  // per component, and of the truncated band of each line only the positions
  // the first FFT pass reads are written (zeros).
  constexpr int EM = L * NKK;
  constexpr int QM = (EM + ZT - 1) / ZT;
  cx<T> ua[QM], ub[QM], uc[QM];  // u1, u2, u3 of this thread's modes
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int e = int(threadIdx.x) + ZT * q;
    const int line = e / NKK;
    const bool ok = (EM % ZT == 0 || e < EM) && j0 + line < nkj;
    ua[q] = ok ? get<kNrZinv>(&uu[0][e]) : czero<T>();
    ub[q] = ok ? get<kNrZinv>(&uu[1][e]) : czero<T>();
    uc[q] = ok ? get<kNrZinv>(&uu[2][e]) : czero<T>();
  }
  const double wi = double(i);
#pragma unroll
  for (int f = 0; f < 6; ++f) {  // unrolled: every field selection is static
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int e = int(threadIdx.x) + ZT * q;
      if (EM % ZT == 0 || e < EM) {
        const int line = e / NKK;
        const int k = e - line * NKK;
        const int z = (k <= K) ? k : k + (N - NKK);
        const cx<T> un[3] = {ua[q], ub[q], uc[q]};
        cx<T> v;
        if (f < 3) {
          v = un[f];
        } else {
          // w_c = i sum_n (G[c][n] . k) u_n, the same operations as k2_zinv
          const int c = f - 3;
          const double kv[3] = {wi, wvn2(j0 + line, g.kj), wvn2(k, K)};
          cx<T> acc = czero<T>();
#pragma unroll
          for (int n = 0; n < 3; ++n) {
            const double* G = &cf.G[(c * 3 + n) * 3];
            if ((G[0] != 0.0) || (G[1] != 0.0) || (G[2] != 0.0))
              acc = add(acc, scl(un[n], T(G[0] * kv[0] + G[1] * kv[1] + G[2] * kv[2])));
          }
          v = mk<T>(-acc.y, acc.x);
        }
        buf[line * LD + lds_off<T>(z)] = v;
      }
    }
    zero_band_edges<T, N, L, LD, K, ZT>(buf);
    __syncthreads();
    fft_ip<T, N, L, LD, +1, true, ZT, K>(buf, tw);
    cx<T>* dst = w1 + (long long)f * g.m1 + (long long)i * N * g.ldj + j0;
    for (int idx = threadIdx.x; idx < L * N; idx += ZT) {
      const int z = idx / L;
      const int line = idx - z * L;
      if (j0 + line < nkj) put<kNtZinv>(&dst[z * g.ldj + line], buf[line * LD + lds_off<T>(z)]);
    }
    __syncthreads();
  }
}

template <typename T, int N>
__global__ void __launch_bounds__(PDR_ZINV_THREADS) PDR_ZINV_WAVES_T
    k2_zinv_r(Geo g, Coef cf, const cx<T>* __restrict__ u0, const cx<T>* __restrict__ u1,
              const cx<T>* __restrict__ u2, cx<T>* __restrict__ w1, const cx<T>* __restrict__ twg, int ntile) {
  k2_zinv_r_body<T, N, PDR_ZINV_THREADS>(xcd_remap(blockIdx.x, gridDim.x), g, cf, u0, u1, u2, w1, twg, ntile);
}

// ---------------------------------------------------------------------------
// pass 2: inverse y FFT, (z, i-tile) per block
// ---------------------------------------------------------------------------
template <typename T, int N, int YT = kBlock2>
__device__ __forceinline__ void k2_yinv_body(unsigned bid, const Geo& g, const cx<T>* __restrict__ w1, cx<T>* __restrict__ w2, const cx<T>* __restrict__ twg,
            int ntile, int z0, int nzc, int fpb) {
  constexpr int PP = y_planes<T>(N);
  constexpr int L = y_lines<T>(N);
  constexpr int LD = lds_ld<T>(N);
  constexpr int K = N / 3;
  constexpr int NKJ = 2 * K + 1;
  constexpr int QE = (L * N + YT - 1) / YT;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
  cx<T>* buf = tw + N;
  const int tile = int(bid % unsigned(ntile));
  const int zg = int(bid / unsigned(ntile));
  const int zl = PP ? 0 : zg % nzc;   // plane within the chunk (PP: tile = group of PP planes)
  const int fg = PP ? zg : zg / nzc;  // field group
  const int z = z0 + zl;
  const int i0 = PP ? 0 : tile * L;
  const int nki = g.nki;
  const int lstride = g.nz * g.ldj;
  load_tw<T, N, YT>(tw, twg);
  // Tiles of i-lines: a thread takes the stored modes e = tid + YT q of the tile
  // (line e / NKJ, mode j = e % NKJ); of the truncated band of each line only
  // the positions the first FFT pass reads are written (zeros).  Whole-plane
  // tiles fill every position the first pass reads.
  constexpr int EM = L * NKJ;
  constexpr int QM = PP > 0 ? QE : (EM + YT - 1) / YT;
  cx<T> pre[QM];
  auto load = [&](int f) {
    const cx<T>* src = w1 + (long long)f * g.m1 + ((long long)i0 * g.nz + z) * g.ldj;
    if constexpr (PP == 0) {
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int e = int(threadIdx.x) + YT * q;
        const int line = e / NKJ;
        cx<T> v = czero<T>();
        if ((EM % YT == 0 || e < EM) && i0 + line < nki) v = get<kNrYinv>(&src[line * lstride + (e - line * NKJ)]);
        pre[q] = v;
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + YT * q;
      const int line = idx / N;
      const int y = idx - line * N;
      const int j = (y <= K) ? y : ((y >= N - K) ? y - (N - NKJ) : -1);
      cx<T> v = czero<T>();
      if constexpr (PP > 0) {
        const int p = line / (K + 1);
        const int i = line - p * (K + 1);
        const int zz = tile * PP + p;
        if ((L * N % YT == 0 || idx < L * N) && zz < nzc && j >= 0)
          v = get<kNrYinv>(&src[((long long)i * g.nz + zz) * g.ldj + j]);
      } else {
        if ((L * N % YT == 0 || idx < L * N) && i0 + line < nki && j >= 0)
          v = get<kNrYinv>(&src[line * lstride + j]);
      }
      pre[q] = v;
    }
  };
  for (int f = fg * fpb; f < (fg + 1) * fpb; ++f) {
    load(f);
    if constexpr (PP == 0) {
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int e = int(threadIdx.x) + YT * q;
        if (EM % YT == 0 || e < EM) {
          const int line = e / NKJ;
          const int j = e - line * NKJ;
          buf[line * LD + lds_off<T>((j <= K) ? j : j + (N - NKJ))] = pre[q];
        }
      }
      zero_band_edges<T, N, L, LD, K, YT>(buf);
    } else {
      // whole-plane tiles: every position but those of the skipped band blocks
      // (the band's other positions hold the zeros of the load)
#pragma unroll
      for (int q = 0; q < QE; ++q) {
        const int idx = int(threadIdx.x) + YT * q;
        if (L * N % YT == 0 || idx < L * N) {
          const int line = idx / N;
          const int y = idx - line * N;
          if (!in_skipped_block<N, K>(y)) buf[line * LD + lds_off<T>(y)] = pre[q];
        }
      }
    }
    __syncthreads();
    fft_ip<T, N, L, LD, +1, true, YT, K>(buf, tw);
    cx<T>* dst = w2 + (long long)f * g.m2 + (long long)zl * N * g.ldi + i0;
    for (int idx = threadIdx.x; idx < L * N; idx += YT) {
      const int y = idx / L;
      const int line = idx - y * L;
      if constexpr (PP > 0) {
        const int p = line / (K + 1);
        const int zz = tile * PP + p;
        if (zz < nzc)
          put<kNtYinv>(&dst[((long long)zz * N + y) * g.ldi + (line - p * (K + 1))], buf[line * LD + lds_off<T>(y)]);
      } else {
        if (i0 + line < nki) put<kNtYinv>(&dst[y * g.ldi + line], buf[line * LD + lds_off<T>(y)]);
      }
    }
    __syncthreads();
  }
}

// The single-box inverse y pass of tiles of i-lines runs 512-thread workgroups
// over the same tile: half the elements per thread, so the registers of a wave
// fall below half of what 256 threads need and the CU holds twice the waves
// (the tile's LDS then bounds it) to keep more loads in flight.  Whole-plane
// tiles of short lines and the batched ensembles keep 256 threads.
#ifndef PDR_YINV_THREADS
#define PDR_YINV_THREADS 512
#endif
template <typename T, int N>
constexpr int yinv_threads() {
  return y_planes<T>(N) ? kBlock2 : PDR_YINV_THREADS;
}
// waves per SIMD declared to the register allocator: as for the other passes
// (at most two 256-thread workgroups), or the two waves per SIMD of each
// 512-thread workgroup times the workgroups the LDS admits, at most four
template <typename T, int N>
constexpr int yinv_waves() {
  constexpr int wg = kLdsPerCu / lds_y_bytes<T>(N) < 1 ? 1 : kLdsPerCu / lds_y_bytes<T>(N);
  return yinv_threads<T, N>() == kBlock2 ? waves_for(lds_y_bytes<T>(N))
                                          : (wg * (yinv_threads<T, N>() / kBlock2) > 4 ? 4 : wg * (yinv_threads<T, N>() / kBlock2));
}

template <typename T, int N>
__global__ void __launch_bounds__((yinv_threads<T, N>())) __attribute__((amdgpu_waves_per_eu(yinv_waves<T, N>())))
    k2_yinv(Geo g, const cx<T>* __restrict__ w1, cx<T>* __restrict__ w2, const cx<T>* __restrict__ twg,
            int ntile, int z0, int nzc, int fpb) {
  k2_yinv_body<T, N, yinv_threads<T, N>()>(xcd_remap(blockIdx.x, gridDim.x), g, w1, w2, twg, ntile, z0, nzc, fpb);
}

// ---------------------------------------------------------------------------
// pass 3: RB rows (z, y0..y0+RB-1) per block: c2r, product, r2c
// ---------------------------------------------------------------------------
template <typename T, int N, bool RS, int XT = kBlock2>
__device__ __forceinline__ void k2_xpass_body(unsigned bid, const Geo& g, const Coef& cf, cx<T>* __restrict__ w2, T* __restrict__ ur, int want_tr, int want_cfl,
             unsigned long long* __restrict__ umax, const cx<T>* __restrict__ twg, int ntile, int z0,
             double* __restrict__ rsp) {
  constexpr int RB = xrows<T>(N);
  constexpr int LD = lds_ld<T>(N);
  constexpr int KI = N / 3;
  constexpr int NKI = KI + 1;
  constexpr int IT = (RB * N + XT - 1) / XT;  // product items per thread
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ T red[XT / 64];  // the CFL maximum of each wave
  constexpr bool TWG = x_tw_global<T>(N);
  cx<T>* const lds0 = reinterpret_cast<cx<T>*>(smem);
  const cx<T>* const tw = TWG ? twg : lds0;
  cx<T>* const buf = TWG ? lds0 : lds0 + N;
  const int tile = int(bid % unsigned(ntile));
  const int zl = int(bid / unsigned(ntile));  // plane within the chunk
  const int z = z0 + zl;
  const int y0 = tile * RB;
  const int ny = g.ny;
  if constexpr (!TWG) load_tw<T, N, XT>(lds0, twg);
  cx<T>* rowp = w2 + ((long long)zl * ny + y0) * g.ldi;
  // fill: the 3 * RB lines pack pairs of real inputs (p = 0: u1 + i u2,
  // 1: u3 + i w1, 2: w2 + i w3).  Each stored mode m' = 0..KI of a pair (a, b) is
  // loaded once and written to both m' and its Hermitian image N - m'
  // (conj(a) + i conj(b)); c2r uses only the real part of the DC term; the band
  // KI < m < N - KI is zero, and only its positions that the first FFT pass
  // reads are written.  All loads of a thread precede its LDS stores.
  if constexpr (NKI >= 128) {
    // long lines: thread m takes mode m of every line, so the global row bases
    // are workgroup-uniform, the LDS line offsets compile-time constants and
    // the per-element index arithmetic is one address add
    constexpr int NL = 3 * RB;
    const cx<T>* fa[NL];
#pragma unroll
    for (int line = 0; line < NL; ++line) fa[line] = rowp + (long long)(2 * (line % 3)) * g.m2 + (line / 3) * g.ldi;
    for (int m0 = 0; m0 < NKI; m0 += XT) {
      const int m = m0 + int(threadIdx.x);
      if (m < NKI) {
        cx<T> A[NL], B[NL];
#pragma unroll
        for (int line = 0; line < NL; ++line) {
          A[line] = czero<T>();
          B[line] = czero<T>();
          if (y0 + line / 3 < ny) {
            A[line] = get<kNrXpass>(&fa[line][m]);
            B[line] = get<kNrXpass>(&fa[line][g.m2 + m]);
          }
        }
        const int om = lds_off<T>(m), on = lds_off<T>(N - m);
#pragma unroll
        for (int line = 0; line < NL; ++line) {
          cx<T>* row = buf + line * LD;
          if (m == 0) {
            row[0] = mk<T>(A[line].x, B[line].x);
          } else {
            row[om] = mk<T>(A[line].x - B[line].y, A[line].y + B[line].x);
            row[on] = mk<T>(A[line].x + B[line].y, B[line].x - A[line].y);
          }
        }
      }
    }
  } else {
    constexpr int EL = 3 * RB * NKI;
    constexpr int QL = (EL + XT - 1) / XT;
    cx<T> A[QL], B[QL];
#pragma unroll
    for (int q = 0; q < QL; ++q) {
      const int idx = int(threadIdx.x) + XT * q;
      const int line = idx / NKI;
      const int m = idx - line * NKI;
      const int r = line / 3;
      const int p = line - 3 * r;
      A[q] = czero<T>();
      B[q] = czero<T>();
      if ((EL % XT == 0 || idx < EL) && y0 + r < ny) {
        const cx<T>* fa = rowp + (long long)(2 * p) * g.m2 + r * g.ldi;
        A[q] = get<kNrXpass>(&fa[m]);
        B[q] = get<kNrXpass>(&fa[g.m2 + m]);
      }
    }
#pragma unroll
    for (int q = 0; q < QL; ++q) {
      const int idx = int(threadIdx.x) + XT * q;
      if (EL % XT == 0 || idx < EL) {
        const int line = idx / NKI;
        const int m = idx - line * NKI;
        cx<T>* row = buf + line * LD;
        if (m == 0) {
          row[0] = mk<T>(A[q].x, B[q].x);
        } else {
          row[lds_off<T>(m)] = mk<T>(A[q].x - B[q].y, A[q].y + B[q].x);
          row[lds_off<T>(N - m)] = mk<T>(A[q].x + B[q].y, B[q].x - A[q].y);
        }
      }
    }
  }
  zero_band_edges<T, N, 3 * RB, LD, KI, XT>(buf);
  __syncthreads();
  fft_ip<T, N, 3 * RB, LD, +1, !TWG, XT, KI>(buf, tw);
  T V[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    if constexpr (sizeof(T) == 4)
      V[q] = cf.Vf[q];
    else
      V[q] = T(cf.V[q]);
  }
  T nl[IT][3];
  T lmax = T(0);
  T* urow = ur + ((long long)z * ny + y0) * N * 3;
  RsAccT<T> acc;
  if (RS) rs_init(acc);
  // the product, with the frame's form a compile-time choice (no per-point select)
  auto product = [&](auto form_c) {
    constexpr bool FORM = decltype(form_c)::value;
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int idx = int(threadIdx.x) + XT * q;
    if ((RB * N) % XT == 0 || idx < RB * N) {
      const int r = idx / N;
      const int x = idx - r * N;
      const cx<T> a = buf[(3 * r) * LD + lds_off<T>(x)], b = buf[(3 * r + 1) * LD + lds_off<T>(x)],
                  c = buf[(3 * r + 2) * LD + lds_off<T>(x)];
      const T u1 = a.x, u2 = a.y, u3 = b.x, o1 = b.y, o2 = c.x, o3 = c.y;
      const T v1 = V[0] * u1 + V[1] * u2 + V[2] * u3;
      const T v2 = V[3] * u1 + V[4] * u2 + V[5] * u3;
      const T v3 = V[6] * u1 + V[7] * u2 + V[8] * u3;
      // omega x (binv u) (BF_ROT), or omega x u (isotropic with binv = I, R_ROT)
      const T w1 = FORM ? u1 : v1, w2 = FORM ? u2 : v2, w3 = FORM ? u3 : v3;
      nl[q][0] = o2 * w3 - o3 * w2;
      nl[q][1] = o3 * w1 - o1 * w3;
      nl[q][2] = o1 * w2 - o2 * w1;
      if (want_tr && y0 + r < ny) {
        T* d = urow + idx * 3;
        put1<kNtXpass>(d, u1);
        put1<kNtXpass>(d + 1, u2);
        put1<kNtXpass>(d + 2, u3);
      }
      if (want_cfl) {
        const T s = fabs(v1) + fabs(v2) + fabs(v3);  // lab-frame velocity (src/pf_fluidstats.f90:853-866, 914)
        lmax = s > lmax ? s : lmax;
      }
      if (RS && y0 + r < ny) rs_add(acc, v1, v2, v3, o1, o2, o3);
    }
  }
  };
  if (cf.form)
    product(std::true_type{});
  else
    product(std::false_type{});
  if constexpr (RS) {
    // partial of this workgroup at its place in the whole box: (z, tile); where
    // the x tile can hold the threads' partials (the product has read it), they
    // are reduced through it
    double* const rso = rsp + ((long long)z * ntile + tile) * kRsWidth;
    __shared__ double rs_lds[kRsWidth * 4];
    if constexpr (kRsWidth * (XT + 1) * sizeof(T) <= 3 * RB * LD * sizeof(cx<T>)) {
      __syncthreads();
      rs_block_out_lds<XT>(acc, reinterpret_cast<T*>(buf), rs_lds, rso);
    } else {
      static_assert(kRsWidth * (XT / 64) <= kRsWidth * 4, "rs_block_out's LDS");
      rs_block_out(acc, rs_lds, rso);
    }
  }
  __syncthreads();
  // forward lines: r -> n1 + i n2 of row r; RB + r/2 -> n3 of rows (r, r+1)
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int idx = int(threadIdx.x) + XT * q;
    if ((RB * N) % XT == 0 || idx < RB * N) {
      const int r = idx / N;
      const int x = idx - r * N;
      buf[r * LD + lds_off<T>(x)] = mk<T>(nl[q][0], nl[q][1]);
      cx<T>* p3 = &buf[(RB + r / 2) * LD + lds_off<T>(x)];
      if (r & 1)
        p3->y = nl[q][2];
      else
        p3->x = nl[q][2];
    }
  }
  __syncthreads();
  fft_ip<T, N, 3 * RB / 2, LD, -1, !TWG, XT>(buf, tw);
  // unpack the pairs: row r of n1 + i n2 and half of line RB + r/2 give the
  // stored modes m = 0..KI of the three forward fields
  auto unpack = [&](int r, int m, int om, int omr) {
    const cx<T> Z = buf[r * LD + om], Zr = buf[r * LD + omr];
    const cx<T> W = buf[(RB + r / 2) * LD + om], Wr = buf[(RB + r / 2) * LD + omr];
    cx<T>* d0 = rowp + r * g.ldi + m;
    put<kNtXpass>(d0, mk<T>(T(0.5) * (Z.x + Zr.x), T(0.5) * (Z.y - Zr.y)));
    put<kNtXpass>(d0 + g.m2, mk<T>(T(0.5) * (Z.y + Zr.y), T(-0.5) * (Z.x - Zr.x)));
    if (r & 1)
      put<kNtXpass>(d0 + 2 * g.m2, mk<T>(T(0.5) * (W.y + Wr.y), T(-0.5) * (W.x - Wr.x)));
    else
      put<kNtXpass>(d0 + 2 * g.m2, mk<T>(T(0.5) * (W.x + Wr.x), T(0.5) * (W.y - Wr.y)));
  };
  if constexpr (NKI >= 128) {
    for (int m0 = 0; m0 < NKI; m0 += XT) {
      const int m = m0 + int(threadIdx.x);
      if (m < NKI) {
        const int om = lds_off<T>(m), omr = lds_off<T>((m == 0) ? 0 : N - m);
#pragma unroll
        for (int r = 0; r < RB; ++r)
          if (y0 + r < ny) unpack(r, m, om, omr);
      }
    }
  } else {
    for (int idx = threadIdx.x; idx < RB * NKI; idx += XT) {
      const int r = idx / NKI;
      const int m = idx - r * NKI;
      if (y0 + r >= ny) continue;
      const int mr = (m == 0) ? 0 : N - m;
      unpack(r, m, lds_off<T>(m), lds_off<T>(mr));
    }
  }
  if (want_cfl) {
    // maximum over the wave, then over the XT / 64 waves of the workgroup
    T m = lmax;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T o = __shfl_xor(m, off);
      m = o > m ? o : m;
    }
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      T r = red[0];
      for (int w = 1; w < XT / 64; ++w) r = red[w] > r ? red[w] : r;
      atomicMax(umax, (unsigned long long)__double_as_longlong(double(r)));
    }
  }
}

template <typename T, int N, bool RS>
__global__ void __launch_bounds__(kBlock2) PDR_X_WAVES
    k2_xpass(Geo g, Coef cf, cx<T>* __restrict__ w2, T* __restrict__ ur, int want_tr, int want_cfl,
             unsigned long long* __restrict__ umax, const cx<T>* __restrict__ twg, int ntile, int z0,
             double* __restrict__ rsp) {
  k2_xpass_body<T, N, RS>(xcd_remap(blockIdx.x, gridDim.x), g, cf, w2, ur, want_tr, want_cfl, umax, twg, ntile, z0,
                          rsp);
}

// ---------------------------------------------------------------------------
// pass 4: forward y FFT + truncation, (z, i-tile) per block
// ---------------------------------------------------------------------------
template <typename T, int N>
__device__ __forceinline__ void k2_yfwd_body(unsigned bid, const Geo& g, const cx<T>* __restrict__ w2, cx<T>* __restrict__ w1, const cx<T>* __restrict__ twg,
            int ntile, int z0, int nzc, int fpb) {
  constexpr int PP = y_planes<T>(N);
  constexpr int L = y_lines_fwd<T>(N);
  constexpr int LD = lds_ld<T>(N);
  constexpr int K = N / 3;
  constexpr int NKJ = 2 * K + 1;
  constexpr int QE = (L * N + kBlock2 - 1) / kBlock2;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
  cx<T>* buf = tw + N;
  const int tile = int(bid % unsigned(ntile));
  const int zg = int(bid / unsigned(ntile));
  const int zl = PP ? 0 : zg % nzc;   // PP: tile = group of PP planes
  const int fg = PP ? zg : zg / nzc;
  const int z = z0 + zl;
  const int i0 = PP ? 0 : tile * L;
  const int nki = g.nki;
  const int lstride = g.nz * g.ldj;
  load_tw<T, N>(tw, twg);
  cx<T> pre[QE];
  auto load = [&](int f) {
    const cx<T>* src = w2 + (long long)f * g.m2 + (long long)zl * N * g.ldi + i0;
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + kBlock2 * q;
      const int y = idx / L;
      const int line = idx - y * L;
      if constexpr (PP > 0) {
        const int p = line / (K + 1);
        const int zz = tile * PP + p;
        pre[q] = ((L * N % kBlock2 == 0 || idx < L * N) && zz < nzc)
                     ? get<kNrYfwd>(&src[((long long)zz * N + y) * g.ldi + (line - p * (K + 1))])
                     : czero<T>();
      } else {
        pre[q] = ((L * N % kBlock2 == 0 || idx < L * N) && i0 + line < nki) ? get<kNrYfwd>(&src[y * g.ldi + line])
                                                                          : czero<T>();
      }
    }
  };
  for (int f = fg * fpb; f < (fg + 1) * fpb; ++f) {
    load(f);
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + kBlock2 * q;
      if (L * N % kBlock2 == 0 || idx < L * N) {
        const int y = idx / L;
        buf[(idx - y * L) * LD + lds_off<T>(y)] = pre[q];
      }
    }
    __syncthreads();
    fft_ip<T, N, L, LD, -1>(buf, tw);
    cx<T>* dst = w1 + (long long)f * g.m1 + ((long long)i0 * g.nz + z) * g.ldj;
    for (int idx = threadIdx.x; idx < L * NKJ; idx += kBlock2) {
      const int line = idx / NKJ;
      const int j = idx - line * NKJ;
      const int yy = (j <= K) ? j : N - NKJ + j;
      if constexpr (PP > 0) {
        const int p = line / (K + 1);
        const int zz = tile * PP + p;
        if (zz < nzc)
          put<kNtYfwd>(&dst[((long long)(line - p * (K + 1)) * g.nz + zz) * g.ldj + j], buf[line * LD + lds_off<T>(yy)]);
      } else {
        if (i0 + line < nki) put<kNtYfwd>(&dst[line * lstride + j], buf[line * LD + lds_off<T>(yy)]);
      }
    }
    __syncthreads();
  }
}

template <typename T, int N>
__global__ void __launch_bounds__(kBlock2) PDR_YF_WAVES
    k2_yfwd(Geo g, const cx<T>* __restrict__ w2, cx<T>* __restrict__ w1, const cx<T>* __restrict__ twg,
            int ntile, int z0, int nzc, int fpb) {
  k2_yfwd_body<T, N>(xcd_remap(blockIdx.x, gridDim.x), g, w2, w1, twg, ntile, z0, nzc, fpb);
}

// ---------------------------------------------------------------------------
// pass 5: forward z FFT, normalisation, truncation -> NL [c][i][j][k]
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__(kBlock2) PDR_FWD_WAVES
    k2_zfwd(Geo g, const cx<T>* __restrict__ w1, cx<T>* __restrict__ nl, const cx<T>* __restrict__ twg,
            int ntile, T norm) {
  constexpr int L = tile_lines_fwd<T>(N);
  constexpr int LD = lds_ld<T>(N);
  constexpr int K = N / 3;
  constexpr int NKK = 2 * K + 1;
  constexpr int QE = (L * N + kBlock2 - 1) / kBlock2;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
  cx<T>* buf = tw + N;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = int(bid % unsigned(ntile));
  const int i = int(bid / unsigned(ntile));
  const int j0 = tile * L;
  const int nkj = g.nkj;
  load_tw<T, N>(tw, twg);
  cx<T> pre[QE];
  auto load = [&](int f) {
    const cx<T>* src = w1 + (long long)f * g.m1 + (long long)i * N * g.ldj + j0;
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + kBlock2 * q;
      const int z = idx / L;
      const int line = idx - z * L;
      pre[q] = ((L * N % kBlock2 == 0 || idx < L * N) && j0 + line < nkj) ? get<kNrZfwd>(&src[z * g.ldj + line]) : czero<T>();
    }
  };
  for (int f = 0; f < 3; ++f) {
    load(f);
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + kBlock2 * q;
      if (L * N % kBlock2 == 0 || idx < L * N) {
        const int z = idx / L;
        buf[(idx - z * L) * LD + lds_off<T>(z)] = pre[q];
      }
    }
    __syncthreads();
    fft_ip<T, N, L, LD, -1>(buf, tw);
    cx<T>* dst = nl + (long long)f * g.mw + ((long long)i * nkj + j0) * NKK;
    const int nl_lines = (nkj - j0) < L ? (nkj - j0) : L;
    for (int idx = threadIdx.x; idx < nl_lines * NKK; idx += kBlock2) {
      const int line = idx / NKK;
      const int k = idx - line * NKK;
      const int zz = (k <= K) ? k : N - NKK + k;
      put<kNtZfwd>(&dst[idx], scl(buf[line * LD + lds_off<T>(zz)], norm));
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pass 5 + FormF: forward z FFT of the three NL fields of an (i, j-tile), then
// FormF / RK3 / continuity fix of the tile's modes straight from registers:
// NL never goes through HBM (f_RK3SolveNavierStokes, src/pf_rk3.f90:2067-2130,
// does the same per i slice).  The modes of a tile, (i, j0..j0+L-1, all k), are
// contiguous in the wave arrays.
// ---------------------------------------------------------------------------
#ifndef PDR_ZFF_PF
#define PDR_ZFF_PF 1
#endif
template <typename T, int N, int FT = kBlock2, int FRC = 2>
__device__ __forceinline__ void k2_zfwd_formf_body(unsigned bid, const Geo& g, const Coef& cf, const cx<T>* __restrict__ w1, cx<T>* __restrict__ u0, cx<T>* __restrict__ u1,
                  cx<T>* __restrict__ u2, cx<T>* __restrict__ r0, cx<T>* __restrict__ r1, cx<T>* __restrict__ r2,
                  const double2* __restrict__ frc, int fnk, double fkf, double beta, double gdt,
                  const cx<T>* __restrict__ twg, int ntile, T norm) {
  constexpr int L = tile_lines_fwd<T>(N);
  constexpr int LD = lds_ld<T>(N);
  constexpr int K = N / 3;
  constexpr int NKK = 2 * K + 1;
  constexpr int QE = (L * N + FT - 1) / FT;
  constexpr int QF = (L * NKK + FT - 1) / FT;  // modes per thread
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
  cx<T>* buf = tw + N;
  const int tile = int(bid % unsigned(ntile));
  const int i = int(bid / unsigned(ntile));
  const int j0 = tile * L;
  const int nkj = g.nkj;
  const int nl_lines = (nkj - j0) < L ? (nkj - j0) : L;
  load_tw<T, N, FT>(tw, twg);
  cx<T> pre[QE];
  cx<T> nlv[3][QF];
  auto load = [&](int f) {
    const cx<T>* src = w1 + (long long)f * g.m1 + (long long)i * N * g.ldj + j0;
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + FT * q;
      const int z = idx / L;
      const int line = idx - z * L;
      pre[q] = ((L * N % FT == 0 || idx < L * N) && j0 + line < nkj) ? src[z * g.ldj + line] : czero<T>();
    }
  };
  // FP32 lines of 256..512 points: the next field's loads are issued before
  // this field's transform, whose work then hides their latency (the registers
  // of one field's elements fit beside the transform's)
  constexpr bool PF = PDR_ZFF_PF && sizeof(T) == 4 && N >= 256 && N <= 512;
  if constexpr (PF) load(0);
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if constexpr (!PF) load(f);
#pragma unroll
    for (int q = 0; q < QE; ++q) {
      const int idx = int(threadIdx.x) + FT * q;
      if (L * N % FT == 0 || idx < L * N) {
        const int z = idx / L;
        buf[(idx - z * L) * LD + lds_off<T>(z)] = pre[q];
      }
    }
    __syncthreads();
    if constexpr (PF) {
      if (f < 2) load(f + 1);
    }
    fft_ip<T, N, L, LD, -1, true, FT>(buf, tw);
#pragma unroll
    for (int q = 0; q < QF; ++q) {
      const int idx = int(threadIdx.x) + FT * q;
      const int line = idx / NKK;
      const int k = idx - line * NKK;
      const int zz = (k <= K) ? k : N - NKK + k;
      nlv[f][q] = (idx < nl_lines * NKK) ? scl(buf[line * LD + lds_off<T>(zz)], norm) : czero<T>();
    }
    __syncthreads();
  }
  const long long e0 = ((long long)i * nkj + j0) * NKK;
#pragma unroll
  for (int q = 0; q < QF; ++q) {
    const int idx = int(threadIdx.x) + FT * q;
    if (idx < nl_lines * NKK) {
      const int line = idx / NKK;
      formf_mode<T, FRC>(g, cf, i, j0 + line, idx - line * NKK, e0 + idx, nlv[0][q], nlv[1][q], nlv[2][q], u0, u1, u2,
                    r0, r1, r2, frc, fnk, fkf, beta, gdt);
    }
  }
}

// FRC: 0 unforced, 1 isotropic forcing, 2 deformed isotropic forcing
template <typename T, int N, int FRC>
__global__ void __launch_bounds__(kBlock2) PDR_FWD_WAVES
    k2_zfwd_formf(Geo g, Coef cf, const cx<T>* __restrict__ w1, cx<T>* __restrict__ u0, cx<T>* __restrict__ u1,
                  cx<T>* __restrict__ u2, cx<T>* __restrict__ r0, cx<T>* __restrict__ r1, cx<T>* __restrict__ r2,
                  const double2* __restrict__ frc, int fnk, double fkf, double beta, double gdt,
                  const cx<T>* __restrict__ twg, int ntile, T norm) {
  k2_zfwd_formf_body<T, N, kBlock2, FRC>(xcd_remap(blockIdx.x, gridDim.x), g, cf, w1, u0, u1, u2, r0, r1, r2, frc,
                                         fnk, fkf, beta, gdt, twg, ntile, norm);
}

template <typename T, int N>
size_t lds_lines() {
  return size_t(lds_lines_bytes<T>(N));
}
template <typename T, int N>
size_t lds_fwd() {
  return size_t(lds_fwd_bytes<T>(N));
}
template <typename T, int N>
size_t lds_y() {
  return size_t(lds_y_bytes<T>(N) > lds_yf_bytes<T>(N) ? lds_y_bytes<T>(N) : lds_yf_bytes<T>(N));
}
template <typename T, int N>
size_t lds_x() {
  return size_t(lds_x_bytes<T>(N) - (kBlock2 / 64) * int(sizeof(T)));  // dynamic part
}

hipError_t big_lds(const void* fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
}

template <typename T, int N>
hipError_t zinv_n(pdr_ctx* c, const Coef& cf) {
  const Geo& g = c->g;
  if constexpr (PDR_ZINV_REG && N <= 512) {  // longer lines hold too many values per thread
    constexpr int L = zinv_lines_t<T, PDR_ZINV_THREADS>(N);
    const int nt = (g.nkj + L - 1) / L;
    const size_t lds = size_t(lds_zinv_bytes_t<T, PDR_ZINV_THREADS>(N));
    hipError_t e = big_lds((const void*)k2_zinv_r<T, N>, lds);
    if (e != hipSuccess) return e;
    k2_zinv_r<T, N><<<dim3(unsigned(nt * g.nki)), PDR_ZINV_THREADS, lds, c->stream>>>(
        g, cf, (const cx<T>*)c->u[0], (const cx<T>*)c->u[1], (const cx<T>*)c->u[2], (cx<T>*)c->w1,
        (const cx<T>*)c->tab.twz, nt);
    return hipGetLastError();
  } else {
    constexpr int L = tile_lines<T>(N);
    const int nt = (g.nkj + L - 1) / L;
    const size_t lds = lds_lines<T, N>();
    hipError_t e = big_lds((const void*)k2_zinv<T, N>, lds);
    if (e != hipSuccess) return e;
    k2_zinv<T, N><<<dim3(unsigned(nt * g.nki)), kBlock2, lds, c->stream>>>(
        g, cf, (const cx<T>*)c->u[0], (const cx<T>*)c->u[1], (const cx<T>*)c->u[2], (cx<T>*)c->w1,
        (const cx<T>*)c->tab.twz, nt);
    return hipGetLastError();
  }
}

template <typename T, int N>
hipError_t yinv_n(pdr_ctx* c, int z0, int nzc) {
  const Geo& g = c->g;
  constexpr int PP = y_planes<T>(N);
  constexpr int L = y_lines<T>(N);
  const int nt = PP ? (nzc + PP - 1) / PP : (g.nki + L - 1) / L;
  const size_t lds = size_t(lds_y_bytes<T>(N));
  hipError_t e = big_lds((const void*)k2_yinv<T, N>, lds);
  if (e != hipSuccess) return e;
  k2_yinv<T, N><<<dim3(unsigned(nt * (PP ? 1 : nzc) * (6 / c->fpb_y))), yinv_threads<T, N>(), lds, c->stream>>>(
      g, (const cx<T>*)c->w1, (cx<T>*)c->w2, (const cx<T>*)c->tab.twy, nt, z0, nzc, c->fpb_y);
  return hipGetLastError();
}

template <typename T, int N>
hipError_t xpass_n(pdr_ctx* c, const Coef& cf, bool tr, bool cfl, int z0, int nzc, double* rsp) {
  const Geo& g = c->g;
  constexpr int RB = xrows<T>(N);
  const int nt = (g.ny + RB - 1) / RB;
  const size_t lds = lds_x<T, N>();
  if (rsp) {
    hipError_t e = big_lds((const void*)k2_xpass<T, N, true>, lds);
    if (e != hipSuccess) return e;
    k2_xpass<T, N, true><<<dim3(unsigned(nt * nzc)), kBlock2, lds, c->stream>>>(
        g, cf, (cx<T>*)c->w2, (T*)c->ur, tr ? 1 : 0, cfl ? 1 : 0, c->umax, (const cx<T>*)c->tab.twx, nt, z0, rsp);
    return hipGetLastError();
  }
  hipError_t e = big_lds((const void*)k2_xpass<T, N, false>, lds);
  if (e != hipSuccess) return e;
  k2_xpass<T, N, false><<<dim3(unsigned(nt * nzc)), kBlock2, lds, c->stream>>>(
      g, cf, (cx<T>*)c->w2, (T*)c->ur, tr ? 1 : 0, cfl ? 1 : 0, c->umax, (const cx<T>*)c->tab.twx, nt, z0, rsp);
  return hipGetLastError();
}

template <typename T, int N>
hipError_t yfwd_n(pdr_ctx* c, int z0, int nzc) {
  const Geo& g = c->g;
  constexpr int PP = y_planes<T>(N);
  constexpr int L = y_lines_fwd<T>(N);
  const int nt = PP ? (nzc + PP - 1) / PP : (g.nki + L - 1) / L;
  const size_t lds = size_t(lds_yf_bytes<T>(N));
  hipError_t e = big_lds((const void*)k2_yfwd<T, N>, lds);
  if (e != hipSuccess) return e;
  k2_yfwd<T, N><<<dim3(unsigned(nt * (PP ? 1 : nzc) * (3 / c->fpb_yf))), kBlock2, lds, c->stream>>>(
      g, (const cx<T>*)c->w2, (cx<T>*)c->w1, (const cx<T>*)c->tab.twy, nt, z0, nzc, c->fpb_yf);
  return hipGetLastError();
}

template <typename T, int N>
hipError_t zfwd_n(pdr_ctx* c) {
  const Geo& g = c->g;
  constexpr int L = tile_lines_fwd<T>(N);
  const int nt = (g.nkj + L - 1) / L;
  const size_t lds = lds_fwd<T, N>();
  hipError_t e = big_lds((const void*)k2_zfwd<T, N>, lds);
  if (e != hipSuccess) return e;
  const double norm = 1.0 / (double(g.nx) * double(g.ny) * double(g.nz));
  k2_zfwd<T, N><<<dim3(unsigned(nt * g.nki)), kBlock2, lds, c->stream>>>(
      g, (const cx<T>*)c->w1, (cx<T>*)c->nl, (const cx<T>*)c->tab.twz, nt, T(norm));
  return hipGetLastError();
}

template <typename T, int N>
hipError_t zfwd_formf_n(pdr_ctx* c, const Coef& cf, double beta, double gdt) {
  if constexpr (N > 512) {  // not used above 512 points per line (see pdr_create)
    return hipErrorNotSupported;
  } else {
  const Geo& g = c->g;
  constexpr int L = tile_lines_fwd<T>(N);
  const int nt = (g.nkj + L - 1) / L;
  const size_t lds = lds_fwd<T, N>();
  const double norm = 1.0 / (double(g.nx) * double(g.ny) * double(g.nz));
  auto launch = [&](auto frc_c) -> hipError_t {
    constexpr int F = decltype(frc_c)::value;
    hipError_t e = big_lds((const void*)k2_zfwd_formf<T, N, F>, lds);
    if (e != hipSuccess) return e;
    k2_zfwd_formf<T, N, F><<<dim3(unsigned(nt * g.nki)), kBlock2, lds, c->stream>>>(
        g, cf, (const cx<T>*)c->w1, (cx<T>*)c->u[0], (cx<T>*)c->u[1], (cx<T>*)c->u[2], (cx<T>*)c->rk[0],
        (cx<T>*)c->rk[1], (cx<T>*)c->rk[2], F ? (const double2*)c->force : nullptr, F ? c->force_nk : 0,
        F ? c->p.force_kf : 0.0, beta, gdt, (const cx<T>*)c->tab.twz, nt, T(norm));
    return hipGetLastError();
  };
  if (!c->forcing) return launch(std::integral_constant<int, 0>{});
  if (cf.fdef) return launch(std::integral_constant<int, 2>{});
  return launch(std::integral_constant<int, 1>{});
  }
}

// ---------------------------------------------------------------------------
// Batched ensemble passes: one launch runs a pass for every member.  Block b
// (after the XCD remap, so that a member's tiles share one XCD's L2) belongs to
// the last member whose start is <= b; the member's descriptor supplies the
// geometry, coefficients, buffers and tile counts, and a switch on its line
// length selects the compile-time body that the single-box kernel also uses,
// so each member computes what its own launches would (up to the rounding of
// multiply-adds that the compiler contracts differently in another kernel).
// ---------------------------------------------------------------------------
#define PDR_ENS_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

// member of block b of a pass: binary search of the members' first blocks
__device__ __forceinline__ int ens_member(const EnsMember* __restrict__ mem, int nm, int pass, unsigned b) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].start[pass] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <typename T>
__global__ void __launch_bounds__(kBlock2) k_ens_zinv(const EnsMember* __restrict__ mem, int nm) {
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const EnsMember& e = mem[ens_member(mem, nm, 0, b)];
  const unsigned lb = b - e.start[0];
  switch (e.n) {
#define PDR_EC(n)                                                                                                \
  case n:                                                                                                        \
    k2_zinv_r_body<T, n>(lb, e.g, e.cf, (const cx<T>*)e.u[0], (const cx<T>*)e.u[1], (const cx<T>*)e.u[2],        \
                         (cx<T>*)e.w1, (const cx<T>*)e.twz, e.nt[0]);                                            \
    break;
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    default:
      break;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock2) k_ens_yinv(const EnsMember* __restrict__ mem, int nm) {
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const EnsMember& e = mem[ens_member(mem, nm, 1, b)];
  const unsigned lb = b - e.start[1];
  switch (e.n) {
#define PDR_EC(n)                                                                                                \
  case n:                                                                                                        \
    k2_yinv_body<T, n>(lb, e.g, (const cx<T>*)e.w1, (cx<T>*)e.w2, (const cx<T>*)e.twy, e.nt[1], 0, e.g.nz,      \
                       e.fpb_y);                                                                                 \
    break;
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    default:
      break;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock2) k_ens_xpass(const EnsMember* __restrict__ mem, int nm) {
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const EnsMember& e = mem[ens_member(mem, nm, 2, b)];
  const unsigned lb = b - e.start[2];
  switch (e.n) {
#define PDR_EC(n)                                                                                                \
  case n:                                                                                                        \
    k2_xpass_body<T, n, false>(lb, e.g, e.cf, (cx<T>*)e.w2, (T*)nullptr, 0, 0, e.umax, (const cx<T>*)e.twx,      \
                               e.nt[2], 0, nullptr);                                                             \
    break;
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    default:
      break;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock2) k_ens_yfwd(const EnsMember* __restrict__ mem, int nm) {
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const EnsMember& e = mem[ens_member(mem, nm, 3, b)];
  const unsigned lb = b - e.start[3];
  switch (e.n) {
#define PDR_EC(n)                                                                                                \
  case n:                                                                                                        \
    k2_yfwd_body<T, n>(lb, e.g, (const cx<T>*)e.w2, (cx<T>*)e.w1, (const cx<T>*)e.twy, e.nt[3], 0, e.g.nz,      \
                       e.fpb_yf);                                                                                \
    break;
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    default:
      break;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock2) k_ens_zfwd_formf(const EnsMember* __restrict__ mem, int nm, double beta,
                                                          double gamma) {
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const EnsMember& e = mem[ens_member(mem, nm, 4, b)];
  const unsigned lb = b - e.start[4];
  const double gdt = gamma * e.dt;
  switch (e.n) {
#define PDR_EC(n)                                                                                                \
  case n:                                                                                                        \
    k2_zfwd_formf_body<T, n, kBlock2, 0>(lb, e.g, e.cf, (const cx<T>*)e.w1, (cx<T>*)e.u[0], (cx<T>*)e.u[1],  \
                             (cx<T>*)e.u[2], (cx<T>*)e.rk[0], (cx<T>*)e.rk[1], (cx<T>*)e.rk[2], nullptr, 0, 0.0, \
                             beta, gdt,                                                                          \
                             (const cx<T>*)e.twz, e.nt[4], T(e.norm));                                           \
    break;
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    default:
      break;
  }
}

template <typename T, int N>
int ens_tiles_n(const Geo& g, int fpb_y, int fpb_yf, int nt[5], unsigned blocks[5]) {
  nt[0] = (g.nkj + zinv_lines<T>(N) - 1) / zinv_lines<T>(N);
  constexpr int PP = y_planes<T>(N);
  nt[1] = PP ? (g.nz + PP - 1) / PP : (g.nki + y_lines<T>(N) - 1) / y_lines<T>(N);
  nt[2] = (g.ny + xrows<T>(N) - 1) / xrows<T>(N);
  nt[3] = PP ? (g.nz + PP - 1) / PP : (g.nki + y_lines_fwd<T>(N) - 1) / y_lines_fwd<T>(N);
  nt[4] = (g.nkj + tile_lines_fwd<T>(N) - 1) / tile_lines_fwd<T>(N);
  blocks[0] = unsigned(nt[0] * g.nki);
  blocks[1] = unsigned(nt[1] * (PP ? 1 : g.nz) * (6 / fpb_y));
  blocks[2] = unsigned(nt[2] * g.nz);
  blocks[3] = unsigned(nt[3] * (PP ? 1 : g.nz) * (3 / fpb_yf));
  blocks[4] = unsigned(nt[4] * g.nki);
  return 0;
}

template <typename T, int N>
size_t ens_lds_n(int pass) {
  switch (pass) {
    case 0: return size_t(lds_zinv_bytes<T>(N));
    case 1: return size_t(lds_y_bytes<T>(N));
    case 2: return lds_x<T, N>();
    case 3: return size_t(lds_yf_bytes<T>(N));
    default: return lds_fwd<T, N>();
  }
}

}  // namespace

bool ens_size_supported(int n) {
  switch (n) {
#define PDR_EC(n) case n:
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    return true;
    default:
      return false;
  }
}

int ens_tiles(int prec, int n, const Geo& g, int fpb_y, int fpb_yf, int nt[5], unsigned blocks[5]) {
  switch (n) {
#define PDR_EC(n)                                                                                     \
  case n:                                                                                             \
    return prec ? ens_tiles_n<float, n>(g, fpb_y, fpb_yf, nt, blocks)                                 \
                : ens_tiles_n<double, n>(g, fpb_y, fpb_yf, nt, blocks);
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    default:
      return -1;
  }
}

size_t ens_lds(int prec, int n, int pass) {
  switch (n) {
#define PDR_EC(n) \
  case n:         \
    return prec ? ens_lds_n<float, n>(pass) : ens_lds_n<double, n>(pass);
    PDR_ENS_SIZES(PDR_EC)
#undef PDR_EC
    default:
      return 0;
  }
}

template <typename T>
static hipError_t ens_prepare_t(size_t lds_max) {
  const void* fns[5] = {(const void*)k_ens_zinv<T>, (const void*)k_ens_yinv<T>, (const void*)k_ens_xpass<T>,
                        (const void*)k_ens_yfwd<T>, (const void*)k_ens_zfwd_formf<T>};
  for (const void* f : fns) {
    hipError_t e = big_lds(f, lds_max);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t ens_prepare(int prec, size_t lds_max) {
  return prec ? ens_prepare_t<float>(lds_max) : ens_prepare_t<double>(lds_max);
}

template <typename T>
static void ens_launch_t(int pass, const EnsMember* mem, int nm, unsigned nblk, size_t lds, double beta,
                         double gamma, hipStream_t s) {
  const dim3 grid(nblk);
  switch (pass) {
    case 0: k_ens_zinv<T><<<grid, kBlock2, lds, s>>>(mem, nm); break;
    case 1: k_ens_yinv<T><<<grid, kBlock2, lds, s>>>(mem, nm); break;
    case 2: k_ens_xpass<T><<<grid, kBlock2, lds, s>>>(mem, nm); break;
    case 3: k_ens_yfwd<T><<<grid, kBlock2, lds, s>>>(mem, nm); break;
    default: k_ens_zfwd_formf<T><<<grid, kBlock2, lds, s>>>(mem, nm, beta, gamma); break;
  }
}

hipError_t launch_ens_pass(int prec, int pass, const EnsMember* mem, int nm, unsigned nblk, size_t lds,
                           double beta, double gamma, hipStream_t s) {
  if (prec)
    ens_launch_t<float>(pass, mem, nm, nblk, lds, beta, gamma, s);
  else
    ens_launch_t<double>(pass, mem, nm, nblk, lds, beta, gamma, s);
  return hipGetLastError();
}

namespace {
}  // namespace

#define PDR_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536)

int xpass_tiles(const pdr_ctx* c) {
  switch (c->g.nx) {
#define PDR_XT(n) \
  case n:         \
    return c->prec ? (c->g.ny + xrows<float>(n) - 1) / xrows<float>(n) : (c->g.ny + xrows<double>(n) - 1) / xrows<double>(n);
    PDR_SIZES(PDR_XT)
#undef PDR_XT
    default:
      return c->g.ny;
  }
}

// Compile-time forms the specialised kernels take for this box: out[0] x rows
// per x-pass workgroup, out[1] 1 if the x tile reads its twiddles from the
// global table, out[2] 1 for the register z inverse (velocity tile loaded once),
// out[3] lines per z-inverse tile, out[4] lines per inverse y tile (or whole
// planes x NKI for short lines).  Returns -1 for lengths without a specialised
// kernel.
int spec_forms(int prec, int nx, int ny, int nz, int out[5]) {
  int ok = 0;
  auto fx = [&](auto tag, auto nn) {
    using T = decltype(tag);
    constexpr int n = decltype(nn)::value;
    out[0] = xrows<T>(n);
    out[1] = x_tw_global<T>(n) ? 1 : 0;
    ++ok;
  };
  auto fz = [&](auto tag, auto nn) {
    using T = decltype(tag);
    constexpr int n = decltype(nn)::value;
    out[2] = (PDR_ZINV_REG && n <= 512) ? 1 : 0;
    out[3] = (PDR_ZINV_REG && n <= 512) ? zinv_lines_t<T, PDR_ZINV_THREADS>(n) : tile_lines<T>(n);
    ++ok;
  };
  auto fy = [&](auto tag, auto nn) {
    using T = decltype(tag);
    constexpr int n = decltype(nn)::value;
    out[4] = y_lines<T>(n);
    ++ok;
  };
  auto dispatch = [&](int n, auto f) {
    switch (n) {
#define PDR_SF(n)                                                    \
  case n:                                                            \
    if (prec)                                                        \
      f(float(0), std::integral_constant<int, n>{});                 \
    else                                                             \
      f(double(0), std::integral_constant<int, n>{});                \
    break;
      PDR_SIZES(PDR_SF)
#undef PDR_SF
      default:
        break;
    }
  };
  dispatch(nx, fx);
  dispatch(nz, fz);
  dispatch(ny, fy);
  return ok == 3 ? 0 : -1;
}

// The whole-plane tiles of the y passes of short lines take the i-lines per
// plane from the compile-time length of the pass: NKI = NY/3 + 1 must then be
// the box's NX/3 + 1.
bool spec_shape_ok(int nx, int ny, int nz) {
  (void)nz;
  return y_planes<float>(ny) == 0 || nx / 3 == ny / 3;
}

bool spec_supported(int n) {
  switch (n) {
#define PDR_CASE(n) case n:
    PDR_SIZES(PDR_CASE)
#undef PDR_CASE
    return true;
    default:
      return false;
  }
}

size_t spec_lds_need(int prec, int nx, int ny, int nz) {
  size_t m = 0;
  auto upd = [&](size_t v) { m = v > m ? v : m; };
#define PDR_LDS(T, n, fn)            \
  switch (n) {                       \
    case 32: upd(fn<T, 32>()); break;     \
    case 48: upd(fn<T, 48>()); break;     \
    case 64: upd(fn<T, 64>()); break;     \
    case 96: upd(fn<T, 96>()); break;     \
    case 128: upd(fn<T, 128>()); break;   \
    case 192: upd(fn<T, 192>()); break;   \
    case 256: upd(fn<T, 256>()); break;   \
    case 384: upd(fn<T, 384>()); break;   \
    case 512: upd(fn<T, 512>()); break;   \
    case 768: upd(fn<T, 768>()); break;   \
    case 1024: upd(fn<T, 1024>()); break; \
    case 1536: upd(fn<T, 1536>()); break; \
    default: break;                  \
  }
  if (prec) {
    PDR_LDS(float, nz, lds_lines)
    PDR_LDS(float, ny, lds_lines)
    PDR_LDS(float, nz, lds_fwd)
    PDR_LDS(float, ny, lds_fwd)
    PDR_LDS(float, ny, lds_y)
    PDR_LDS(float, nx, lds_x)
  } else {
    PDR_LDS(double, nz, lds_lines)
    PDR_LDS(double, ny, lds_lines)
    PDR_LDS(double, nz, lds_fwd)
    PDR_LDS(double, ny, lds_fwd)
    PDR_LDS(double, ny, lds_y)
    PDR_LDS(double, nx, lds_x)
  }
#undef PDR_LDS
  return m;
}

#define PDR_DISPATCH(n, call_f, call_d) \
  case n:                               \
    return c->prec ? call_f : call_d;

hipError_t launch2_zinv(pdr_ctx* c, const Coef& cf) {
  switch (c->g.nz) {
#define PDR_CASE(n) PDR_DISPATCH(n, (zinv_n<float, n>(c, cf)), (zinv_n<double, n>(c, cf)))
    PDR_SIZES(PDR_CASE)
#undef PDR_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch2_yinv(pdr_ctx* c, int z0, int nzc) {
  switch (c->g.ny) {
#define PDR_CASE(n) PDR_DISPATCH(n, (yinv_n<float, n>(c, z0, nzc)), (yinv_n<double, n>(c, z0, nzc)))
    PDR_SIZES(PDR_CASE)
#undef PDR_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch2_xpass(pdr_ctx* c, const Coef& cf, bool tr, bool cfl, int z0, int nzc, double* rsp) {
  switch (c->g.nx) {
#define PDR_CASE(n) \
  PDR_DISPATCH(n, (xpass_n<float, n>(c, cf, tr, cfl, z0, nzc, rsp)), (xpass_n<double, n>(c, cf, tr, cfl, z0, nzc, rsp)))
    PDR_SIZES(PDR_CASE)
#undef PDR_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch2_yfwd(pdr_ctx* c, int z0, int nzc) {
  switch (c->g.ny) {
#define PDR_CASE(n) PDR_DISPATCH(n, (yfwd_n<float, n>(c, z0, nzc)), (yfwd_n<double, n>(c, z0, nzc)))
    PDR_SIZES(PDR_CASE)
#undef PDR_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch2_zfwd(pdr_ctx* c) {
  switch (c->g.nz) {
#define PDR_CASE(n) PDR_DISPATCH(n, (zfwd_n<float, n>(c)), (zfwd_n<double, n>(c)))
    PDR_SIZES(PDR_CASE)
#undef PDR_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch2_zfwd_formf(pdr_ctx* c, const Coef& cf, double beta, double gdt) {
  switch (c->g.nz) {
#define PDR_CASE(n) \
  PDR_DISPATCH(n, (zfwd_formf_n<float, n>(c, cf, beta, gdt)), (zfwd_formf_n<double, n>(c, cf, beta, gdt)))
    PDR_SIZES(PDR_CASE)
#undef PDR_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace pdr
