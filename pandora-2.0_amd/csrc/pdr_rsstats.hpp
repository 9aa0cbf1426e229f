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

struct RsAcc {
  double s[kRsSums];
  double m[4];
};

__device__ __forceinline__ void rs_init(RsAcc& a) {
#pragma unroll
  for (int q = 0; q < kRsSums; ++q) a.s[q] = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) a.m[q] = 0.0;
}

__device__ __forceinline__ void rs_add(RsAcc& a, double v1, double v2, double v3, double w1, double w2,
                                       double w3) {
  a.s[0] += v1 * v1;
  a.s[1] += v2 * v2;
  a.s[2] += v3 * v3;
  a.s[3] += v1 * v2;
  a.s[4] += v1 * v3;
  a.s[5] += v2 * v3;
  const double v[3] = {v1, v2, v3}, w[3] = {w1, w2, w3};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v2c = v[c] * v[c], w2c = w[c] * w[c];
    a.s[6 + c] += v2c * v[c];
    a.s[9 + c] += v2c * v2c;
    a.s[12 + c] += w2c;
    a.s[15 + c] += w2c * w[c];
    a.s[18 + c] += w2c * w2c;
    const double aw = fabs(w[c]);
    a.m[1 + c] = aw > a.m[1 + c] ? aw : a.m[1 + c];
  }
  const double u = fabs(v1) + fabs(v2) + fabs(v3);
  a.m[0] = u > a.m[0] ? u : a.m[0];
}

// Workgroup reduction (sums and maxima) into out[kRsWidth]; lds holds
// kRsWidth * (blockDim.x / 64) doubles.  Every thread of the workgroup calls it.
__device__ __forceinline__ void rs_block_out(RsAcc& a, double* lds, double* __restrict__ out) {
  const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
  const int nw = int(blockDim.x >> 6);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < kRsSums; ++q) a.s[q] += __shfl_xor(a.s[q], off);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double o = __shfl_xor(a.m[q], off);
      a.m[q] = o > a.m[q] ? o : a.m[q];
    }
  }
  __syncthreads();  // lds may alias memory the caller used before
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kRsSums; ++q) lds[wave * kRsWidth + q] = a.s[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) lds[wave * kRsWidth + kRsSums + q] = a.m[q];
  }
  __syncthreads();
  if (int(threadIdx.x) < kRsWidth) {
    const int q = int(threadIdx.x);
    double r = lds[q];
    for (int w = 1; w < nw; ++w) {
      const double v = lds[w * kRsWidth + q];
      r = (q < kRsSums) ? r + v : (v > r ? v : r);
    }
    out[q] = r;
  }
}

}  // namespace pdr
