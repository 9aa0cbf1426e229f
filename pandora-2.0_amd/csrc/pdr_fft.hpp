// pdr_fft.hpp -- batched 1-D mixed-radix (2/3/4/5/8) FFTs on LDS for CDNA4.
//
// A workgroup transforms `nlines` lines of length n held in LDS.  Each pass is a
// Stockham autosort step (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC'08): butterfly j of a radix-R pass reads
// R elements at stride n/R, multiplies by the twiddles exp(sign*2*pi*i*(j mod Ns)*r
// /(Ns*R)), does an R-point DFT in registers and writes at expand(j, Ns, R) + r*Ns.
// Passes ping-pong between two LDS buffers, one barrier per pass.
//
// Conventions follow FFTW as used by the reference (src/pf_fftw.f90:31, 199-291):
// sign -1 forward, +1 backward, no normalisation.  Twiddles come from a per-length
// table of exp(-2*pi*i*m/n) evaluated in long double on the host and rounded to
// the working precision (FP64 twiddles rounded to FP32 in the mixed form).
#pragma once
#include <hip/hip_runtime.h>

namespace pdr {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cx<T> add(cx<T> a, cx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> sub(cx<T> a, cx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> mul(cx<T> a, cx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cx<T> scl(cx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// multiply by sign*i
template <int SIGN, typename T>
__device__ __forceinline__ cx<T> rot(cx<T> a) {
  return (SIGN > 0) ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

constexpr int kMaxPasses = 16;

struct FftPlan {
  int n;
  int npass;
  int radix[kMaxPasses];
};

template <typename T, int R, int SIGN>
struct Dft;

template <typename T, int SIGN>
struct Dft<T, 2, SIGN> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
  }
};

template <typename T, int SIGN>
struct Dft<T, 3, SIGN> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    const T c = T(0.86602540378443864676);  // sin(2 pi / 3)
    cx<T> t1 = add(v[1], v[2]);
    cx<T> t2 = sub(v[0], scl(t1, T(0.5)));
    cx<T> t3 = scl(sub(v[1], v[2]), c);
    v[0] = add(v[0], t1);
    v[1] = add(t2, rot<SIGN>(t3));
    v[2] = sub(t2, rot<SIGN>(t3));
  }
};

template <typename T, int SIGN>
struct Dft<T, 4, SIGN> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> t0 = add(v[0], v[2]);
    cx<T> t1 = sub(v[0], v[2]);
    cx<T> t2 = add(v[1], v[3]);
    cx<T> t3 = rot<SIGN>(sub(v[1], v[3]));
    v[0] = add(t0, t2);
    v[2] = sub(t0, t2);
    v[1] = add(t1, t3);
    v[3] = sub(t1, t3);
  }
};

template <typename T, int SIGN>
struct Dft<T, 5, SIGN> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    const T c1 = T(0.30901699437494742410);   // cos(2 pi/5)
    const T c2 = T(-0.80901699437494742410);  // cos(4 pi/5)
    const T s1 = T(0.95105651629515357212);   // sin(2 pi/5)
    const T s2 = T(0.58778525229247312917);   // sin(4 pi/5)
    cx<T> a = v[0];
    cx<T> t1 = add(v[1], v[4]);
    cx<T> t2 = add(v[2], v[3]);
    cx<T> t3 = sub(v[1], v[4]);
    cx<T> t4 = sub(v[2], v[3]);
    cx<T> r1 = add(a, add(scl(t1, c1), scl(t2, c2)));
    cx<T> r2 = add(a, add(scl(t1, c2), scl(t2, c1)));
    cx<T> i1 = rot<SIGN>(add(scl(t3, s1), scl(t4, s2)));
    cx<T> i2 = rot<SIGN>(sub(scl(t3, s2), scl(t4, s1)));
    v[0] = add(a, add(t1, t2));
    v[1] = add(r1, i1);
    v[4] = sub(r1, i1);
    v[2] = add(r2, i2);
    v[3] = sub(r2, i2);
  }
};

template <typename T, int SIGN>
struct Dft<T, 8, SIGN> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    const T h = T(0.70710678118654752440);  // 1/sqrt(2)
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, SIGN>::run(e);
    Dft<T, 4, SIGN>::run(o);
    // w8^k = exp(SIGN*2*pi*i*k/8)
    cx<T> o1 = mk<T>(h * (o[1].x - T(SIGN) * o[1].y), h * (o[1].y + T(SIGN) * o[1].x));
    cx<T> o2 = rot<SIGN>(o[2]);
    cx<T> o3 = mk<T>(h * (-o[3].x - T(SIGN) * o[3].y), h * (-o[3].y + T(SIGN) * o[3].x));
    v[0] = add(e[0], o[0]);
    v[4] = sub(e[0], o[0]);
    v[1] = add(e[1], o1);
    v[5] = sub(e[1], o1);
    v[2] = add(e[2], o2);
    v[6] = sub(e[2], o2);
    v[3] = add(e[3], o3);
    v[7] = sub(e[3], o3);
  }
};

// One Stockham pass over `nlines` lines (line stride ld) from src to dst.
template <typename T, int R, int SIGN>
__device__ __forceinline__ void stockham_pass(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst,
                                              int n, int ns, int nlines, int ld,
                                              const cx<T>* __restrict__ tw) {
  const int nb = n / R;
  const int total = nb * nlines;
  const int twstep = n / (ns * R);
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / nb;
    const int j = t - line * nb;
    const cx<T>* s = src + line * ld;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * nb];
    const int k = j % ns;
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = tw[k * r * twstep];
        if (SIGN > 0) w.y = -w.y;
        v[r] = mul(v[r], w);
      }
    }
    Dft<T, R, SIGN>::run(v);
    cx<T>* d = dst + line * ld;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[base + r * ns] = v[r];
  }
}

// Transforms nlines lines in buffer a (scratch b).  The caller has written `a`
// and issued __syncthreads().  Returns the buffer that holds the result; all
// threads have passed a barrier after the last pass.
template <typename T, int SIGN>
__device__ cx<T>* fft_lines(cx<T>* a, cx<T>* b, const FftPlan& p, int nlines, int ld,
                            const cx<T>* __restrict__ tw) {
  int ns = 1;
  for (int q = 0; q < p.npass; ++q) {
    const int R = p.radix[q];
    switch (R) {
      case 8: stockham_pass<T, 8, SIGN>(a, b, p.n, ns, nlines, ld, tw); break;
      case 4: stockham_pass<T, 4, SIGN>(a, b, p.n, ns, nlines, ld, tw); break;
      case 2: stockham_pass<T, 2, SIGN>(a, b, p.n, ns, nlines, ld, tw); break;
      case 3: stockham_pass<T, 3, SIGN>(a, b, p.n, ns, nlines, ld, tw); break;
      default: stockham_pass<T, 5, SIGN>(a, b, p.n, ns, nlines, ld, tw); break;
    }
    __syncthreads();
    cx<T>* t = a;
    a = b;
    b = t;
    ns *= R;
  }
  return a;
}

}  // namespace pdr
