// pdr_fft2.hpp -- compile-time specialised batched FFTs in LDS for CDNA4.
//
// Same Stockham algorithm and conventions as pdr_fft.hpp, but the length N, the
// number of lines and the radix plan are template constants, so every index
// division is folded, the pass loop is unrolled and the butterflies of a pass are
// held in registers: each pass reads its inputs into registers, waits at a
// barrier, then writes in place (one LDS buffer instead of two).  Twiddles are
// staged in LDS once per workgroup.
#pragma once
#include "pdr_fft.hpp"

namespace pdr {

constexpr int kBlock2 = 256;

// Radix plan: 8s, then one 4 or 2, then 3s, then 5s.
constexpr int plan_radix(int n, int p) {
  int m = n, q = 0;
  while (m % 8 == 0) {
    if (q == p) return 8;
    m /= 8;
    ++q;
  }
  if (m % 4 == 0) {
    if (q == p) return 4;
    m /= 4;
    ++q;
  }
  if (m % 2 == 0) {
    if (q == p) return 2;
    m /= 2;
    ++q;
  }
  while (m % 3 == 0) {
    if (q == p) return 3;
    m /= 3;
    ++q;
  }
  while (m % 5 == 0) {
    if (q == p) return 5;
    m /= 5;
    ++q;
  }
  return 0;
}

constexpr int plan_npass(int n) {
  int p = 0;
  while (plan_radix(n, p) != 0) ++p;
  return p;
}

// LDS layout of a tile of lines: line l, element o at l * LD + o + (o >> PS),
// i.e. one pad element per 16 elements inside a line (PS = 4; 128 bytes at FP32,
// 256 at FP64, measured faster at 512 points than one per 8, 32 or 64 at FP32)
// and a line stride LD = N + (N >> PS) + 1.  Without the pad the stride-R writes
// of the first Stockham passes (NS = 1, 8) fall on a few banks (8-way conflicts
// for radix 8); with it every access pattern of these kernels stays within ~1.5x
// of the conflict-free rate on 32- and 64-bank LDS, and the odd line stride
// keeps the line-fastest (transposing) accesses spread as well.
template <typename T>
constexpr int lds_pad_shift() {
#ifndef PDR_LDS_PAD_F32
#define PDR_LDS_PAD_F32 4
#endif
  return sizeof(T) == 4 ? PDR_LDS_PAD_F32 : 4;
}
template <typename T>
constexpr int lds_ld(int n) {
  return n + (n >> lds_pad_shift<T>()) + 1;
}
template <typename T>
__device__ __forceinline__ int lds_off(int o) {
  return o + (o >> lds_pad_shift<T>());
}

// One radix-R Stockham pass over LINES lines of length N held in LDS (line
// stride LD, padded layout above).  The per-element offsets o(r) = B + r*S are
// turned into a base plus compile-time constants wherever the pad positions
// allow it (all passes of power-of-two lengths), which keeps the address
// registers of the butterflies few.  Twiddles: PPT = true reads the per-pass
// table of load_tw (LDS), false the plain table exp(-2 pi i m / N) (global).
// ZB > 0 declares the band ZB < o < N - ZB of every line zero on entry (the
// modes a 2/3 truncation drops): the inputs r of a butterfly whose whole block
// r NBL .. r NBL + NBL - 1 lies in the band are zero at compile time and not
// read, so the caller writes zeros only to the band positions of the blocks
// that hold stored modes as well (zero_band_edges).
template <typename T, int N, int LINES, int LD, int R, int NS, int SIGN, bool PPT, int NT = kBlock2, int ZB = 0>
__device__ __forceinline__ void pass_ip(cx<T>* __restrict__ buf, const cx<T>* __restrict__ tw) {
  constexpr int NBL = N / R;
  constexpr int TOTAL = NBL * LINES;
  constexpr int NB = (TOTAL + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);
  constexpr int PS = lds_pad_shift<T>();
  constexpr int P2 = 1 << PS;
  cx<T> v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int t = int(threadIdx.x) + NT * b;
    if (TOTAL % NT == 0 || t < TOTAL) {
      const int line = t / NBL;
      const int j = t - line * NBL;
      static_assert(ZB == 0 || NS == 1, "the zero band is known only on entry to the first pass");
      if constexpr (NBL % P2 == 0) {
        const cx<T>* s = buf + line * LD + lds_off<T>(j);
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[b][r] = (ZB > 0 && r * NBL > ZB && r * NBL + NBL - 1 < N - ZB) ? mk<T>(T(0), T(0))
                                                                           : s[r * (NBL + (NBL >> PS))];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[b][r] = (ZB > 0 && r * NBL > ZB && r * NBL + NBL - 1 < N - ZB) ? mk<T>(T(0), T(0))
                                                                           : buf[line * LD + lds_off<T>(j + r * NBL)];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int t = int(threadIdx.x) + NT * b;
    if (TOTAL % NT == 0 || t < TOTAL) {
      const int line = t / NBL;
      const int j = t - line * NBL;
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> w = PPT ? tw[(NS - 1) + k * (R - 1) + (r - 1)] : tw[k * r * TWS];
          if (SIGN > 0) w.y = -w.y;
          v[b][r] = mul(v[b][r], w);
        }
      }
      Dft<T, R, SIGN>::run(v[b]);
      const int A = (j - k) * R;  // a multiple of NS * R
      if constexpr (NS % P2 == 0) {
        cx<T>* d = buf + line * LD + lds_off<T>(A + k);
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * (NS + (NS >> PS))] = v[b][r];
      } else if constexpr (P2 % NS == 0 && (NS * R) % P2 == 0) {
        cx<T>* d = buf + line * LD + lds_off<T>(A) + k;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * NS + ((r * NS) >> PS)] = v[b][r];
      } else if constexpr (P2 % (NS * R) == 0) {
        cx<T>* d = buf + line * LD + lds_off<T>(A) + k;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * NS] = v[b][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) buf[line * LD + lds_off<T>(A + k + r * NS)] = v[b][r];
      }
    }
  }
  __syncthreads();
}

template <typename T, int N, int LINES, int LD, int SIGN, bool PPT = true, int NT = kBlock2, int ZB = 0, int P = 0,
          int NS = 1>
__device__ __forceinline__ void fft_ip(cx<T>* __restrict__ buf, const cx<T>* __restrict__ tw) {
  if constexpr (P < plan_npass(N)) {
    constexpr int R = plan_radix(N, P);
    pass_ip<T, N, LINES, LD, R, NS, SIGN, PPT, NT, P == 0 ? ZB : 0>(buf, tw);
    fft_ip<T, N, LINES, LD, SIGN, PPT, NT, ZB, P + 1, NS * R>(buf, tw);
  }
}

// Zeros into the band positions ZB < o < N - ZB of LINES lines that the first
// pass of fft_ip<..., ZB> reads: those in the first-pass blocks (N / R0
// positions each) that also hold stored modes, at most two runs per line
// ([A, E1] and [S2, B]).  The caller synchronises before the transform.
template <typename T, int N, int LINES, int LD, int ZB, int NT = kBlock2>
__device__ __forceinline__ void zero_band_edges(cx<T>* __restrict__ buf) {
  constexpr int NBL = N / plan_radix(N, 0);
  constexpr int A = ZB + 1, B = N - ZB - 1;  // the band, inclusive
  constexpr int E1 = (A % NBL == 0) ? A - 1 : ((A / NBL + 1) * NBL - 1 < B ? (A / NBL + 1) * NBL - 1 : B);
  constexpr int S2 = ((B + 1) % NBL == 0) ? B + 1 : ((B / NBL) * NBL > E1 ? (B / NBL) * NBL : E1 + 1);
  constexpr int C1 = E1 - A + 1, C = C1 + (B - S2 + 1);
  static_assert(ZB > 0 && A <= B && C1 >= 0 && C >= C1, "a band inside the line");
  // one (line, position) per thread: short lines have few band positions (3 at
  // 32 and 48 points) and many lines per tile, which a thread per position
  // would write one after another
  if constexpr (C > 0) {
    for (int t = threadIdx.x; t < LINES * C; t += NT) {
      const int line = t / C, p = t - line * C;
      buf[line * LD + lds_off<T>(p < C1 ? A + p : S2 + (p - C1))] = mk<T>(T(0), T(0));
    }
  }
}

// Whether position o of a line lies in a first-pass block wholly inside the
// band ZB < o < N - ZB (a position fft_ip<..., ZB> never reads).
template <int N, int ZB>
__device__ __forceinline__ bool in_skipped_block(int o) {
  constexpr int NBL = N / plan_radix(N, 0);
  const int lo = (o / NBL) * NBL;
  return lo > ZB && lo + NBL - 1 < N - ZB;
}

// Global store / load of one complex element; NT selects the non-temporal
// (streaming) form, which does not keep the line in the caches.
template <bool NT, typename T>
__device__ __forceinline__ void put(cx<T>* p, cx<T> v) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
  } else {
    *p = v;
  }
}
template <bool NT, typename T>
__device__ __forceinline__ cx<T> get(const cx<T>* p) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return mk<T>(w.x, w.y);
  } else {
    return *p;
  }
}
template <bool NT, typename T>
__device__ __forceinline__ void put1(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// Per-pass twiddle table in LDS, built from the plain table src[m] =
// exp(-2 pi i m / N) (the caller synchronises).  Pass p (NS > 1) keeps its
// NS (R - 1) twiddles w^(k r N / (NS R)) at NS - 1 + k (R - 1) + (r - 1): the
// lanes of a butterfly wave (consecutive k) then read at stride R - 1, which is
// odd for radix 8, 4 and 2 (no bank conflicts), where the plain table's
// stride k r N / (NS R) puts many lanes on one bank.  The passes' tables tile
// [0, N - 1) exactly (sum of NS (R - 1) over the passes is N - 1).
template <typename T, int N, int NT = kBlock2, int P = 0, int NS = 1>
__device__ __forceinline__ void load_tw(cx<T>* __restrict__ dst, const cx<T>* __restrict__ src) {
  if constexpr (P < plan_npass(N)) {
    constexpr int R = plan_radix(N, P);
    if constexpr (NS > 1) {
      constexpr int TWS = N / (NS * R);
      for (int e = threadIdx.x; e < NS * (R - 1); e += NT) {
        const int k = e / (R - 1);
        const int r = e - k * (R - 1) + 1;
        dst[(NS - 1) + e] = src[k * r * TWS];
      }
    }
    load_tw<T, N, NT, P + 1, NS * R>(dst, src);
  }
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2);
// renumber them so that every XCD walks a contiguous range of tiles.
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblk) {
  constexpr unsigned kXcd = 8;
  const unsigned full = (nblk / kXcd) * kXcd;
  if (bid >= full) return bid;
  return (bid % kXcd) * (full / kXcd) + bid / kXcd;
}

}  // namespace pdr
