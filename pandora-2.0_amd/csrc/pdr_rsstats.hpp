// pdr_rsstats.hpp -- real-space moments accumulated inside the x pass, where the
// velocity and vorticity of every grid point are in registers:
// f_FluidstatsRs2D (src/pf_fluidstats.f90:827-951).  Per point, with the
// laboratory-frame velocity v (binv u; u itself for isotropic runs) and the
// vorticity w of the real-space planes:
//   s[0..5]   v1v1 v2v2 v3v3 v1v2 v1v3 v2v3
//   s[6..8]   v^3,   s[9..11]  v^4
//   s[12..14] w^2,   s[15..17] w^3,  s[18..20] w^4
//   m[0] = max(|v1| + |v2| + |v3|) (ucfl),  m[1..3] = max |w_c|
// Each workgroup writes its 25 partials; a reduction kernel and the host finish.
#pragma once
#include <hip/hip_runtime.h>

namespace pdr {

constexpr int kRsSums = 21;
constexpr int kRsWidth = 25;

// A thread's partials are of type A: double, or float in the FP32-mixed x pass,
// where a thread sums the few points of its own (4 at 512 points per line, from
// FP32 velocities) and every sum across threads and workgroups is FP64.  The
// float partials halve the accumulator's registers (25 instead of 50): the
// FP32 x pass with the moments then needs 92 VGPRs at 512 points instead of
// 130, five waves per SIMD (as many as its LDS admits) instead of three.
template <typename A = double>
struct RsAccT {
  A s[kRsSums];
  A m[4];
};
using RsAcc = RsAccT<double>;

template <typename A>
__device__ __forceinline__ void rs_init(RsAccT<A>& a) {
#pragma unroll
  for (int q = 0; q < kRsSums; ++q) a.s[q] = A(0);
#pragma unroll
  for (int q = 0; q < 4; ++q) a.m[q] = A(0);
}

template <typename A>
__device__ __forceinline__ void rs_add(RsAccT<A>& a, A v1, A v2, A v3, A w1, A w2, A w3) {
  a.s[0] += v1 * v1;
  a.s[1] += v2 * v2;
  a.s[2] += v3 * v3;
  a.s[3] += v1 * v2;
  a.s[4] += v1 * v3;
  a.s[5] += v2 * v3;
  const A v[3] = {v1, v2, v3}, w[3] = {w1, w2, w3};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const A v2c = v[c] * v[c], w2c = w[c] * w[c];
    a.s[6 + c] += v2c * v[c];
    a.s[9 + c] += v2c * v2c;
    a.s[12 + c] += w2c;
    a.s[15 + c] += w2c * w[c];
    a.s[18 + c] += w2c * w2c;
    const A aw = fabs(w[c]);
    a.m[1 + c] = aw > a.m[1 + c] ? aw : a.m[1 + c];
  }
  const A u = fabs(v1) + fabs(v2) + fabs(v3);
  a.m[0] = u > a.m[0] ? u : a.m[0];
}

// Workgroup reduction (sums and maxima) into out[kRsWidth]; lds holds
// kRsWidth * (blockDim.x / 64) doubles.  Every thread of the workgroup calls it.
// The 21 sums go through a reduce-scatter inside the wave: padded to 32 slots,
// in each round of offsets 32, 16, 8, 4, 2 a lane keeps one half of its slots
// and adds its partner's copy of that half (the lane with the offset bit set
// keeps the upper half), so lane l ends with slot l >> 1 summed over 32 lanes and
// one more exchange (offset 1) completes it: 32 shuffles instead of 21 x 6.
// The 4 maxima take a plain butterfly.
template <typename A>
__device__ __forceinline__ void rs_block_out(RsAccT<A>& a, double* lds, double* __restrict__ out) {
  const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
  const int nw = int(blockDim.x >> 6);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const A o = __shfl_xor(a.m[q], off);
      a.m[q] = o > a.m[q] ? o : a.m[q];
    }
  }
  double v[32];
#pragma unroll
  for (int q = 0; q < 32; ++q) v[q] = q < kRsSums ? double(a.s[q]) : 0.0;
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    const int off = 32 >> r, len = 16 >> r;
    const bool up = (lane & off) != 0;
#pragma unroll
    for (int k = 0; k < len; ++k) {
      const double keep = up ? v[len + k] : v[k];
      const double send = up ? v[k] : v[len + k];
      v[k] = keep + __shfl_xor(send, off);
    }
  }
  const double sum = v[0] + __shfl_xor(v[0], 1);
  __syncthreads();  // lds may alias memory the caller used before
  if (!(lane & 1) && (lane >> 1) < kRsSums) lds[wave * kRsWidth + (lane >> 1)] = sum;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) lds[wave * kRsWidth + kRsSums + q] = double(a.m[q]);
  }
  __syncthreads();
  if (int(threadIdx.x) < kRsWidth) {
    const int q = int(threadIdx.x);
    double r = lds[q];
    for (int w = 1; w < nw; ++w) {
      const double v2 = lds[w * kRsWidth + q];
      r = (q < kRsSums) ? r + v2 : (v2 > r ? v2 : r);
    }
    out[q] = r;
  }
}

// Workgroup reduction through LDS, for a caller whose LDS tile is free at this
// point (every thread has finished its reads of it): each thread stores its 25
// partials as A (slot-major rows of NT + 1 values, so that the lanes reading
// different slot rows fall on different banks), then 4 threads per slot each sum
// every fourth value of the row in FP64 (maxima by maximum) and one thread per
// slot combines the 4 in order.  About a third of the instructions of the
// shuffle reduction above.  scratch holds kRsWidth * (NT + 1) values of A,
// done kRsWidth * 4 doubles.  Ends past a barrier, with the tile free again.
template <int NT, typename A>
__device__ __forceinline__ void rs_block_out_lds(const RsAccT<A>& a, A* scratch, double* done,
                                                 double* __restrict__ out) {
  static_assert(NT % 4 == 0 && kRsWidth * 4 <= NT, "one group of 4 threads per slot");
  constexpr int RL = NT + 1;
  const int t = int(threadIdx.x);
#pragma unroll
  for (int q = 0; q < kRsSums; ++q) scratch[q * RL + t] = a.s[q];
#pragma unroll
  for (int q = 0; q < 4; ++q) scratch[(kRsSums + q) * RL + t] = a.m[q];
  __syncthreads();
  if (t < kRsWidth * 4) {
    const int q = t >> 2, p = t & 3;
    const A* row = scratch + q * RL + p;
    double r = 0.0;
    if (q < kRsSums) {
#pragma unroll 8
      for (int k = 0; k < NT / 4; ++k) r += double(row[4 * k]);
    } else {
#pragma unroll 8
      for (int k = 0; k < NT / 4; ++k) r = double(row[4 * k]) > r ? double(row[4 * k]) : r;
    }
    done[t] = r;
  }
  __syncthreads();
  if (t < kRsWidth) {
    double r = done[4 * t];
#pragma unroll
    for (int p = 1; p < 4; ++p) r = (t < kRsSums) ? r + done[4 * t + p] : (done[4 * t + p] > r ? done[4 * t + p] : r);
    out[t] = r;
  }
  __syncthreads();
}

}  // namespace pdr
