// pdr_kernels.hip -- HIP kernels of the MI355X RK3 step (gfx950).
//
// One RK3 stage is five LDS-tiled passes over HBM-resident data plus two
// elementwise kernels:
//   1 k_zinv   per (i, j-tile): vorticity + zero-pad in k + inverse z FFT of
//              u1..u3, w1..w3 -> 1r2w layout [f][i][z][j]
//              (f_RK3PrepareRealSpace, src/pf_rk3.f90:205-948; pf_fftw.f90:586-624)
//   2 k_yinv   per (z, i-tile): zero-pad in j + inverse y FFT -> x-pencils [f][z][y][i]
//              (f_RK3TransformjiFs2Rs, src/pf_rk3.f90:1437-1498; pf_fftw.f90:673-759)
//   3 k_xpass  per (z, y) row: zero-pad in i, c2r (two rows per complex FFT),
//              nonlinear product, r2c, truncation to i <= KI; optionally writes the
//              real-space velocity for the tracers and the CFL maximum
//              (src/pf_rk3.f90:1069-1124, 1597-1711, 1780-1940; pf_fftw.f90:766-834)
//   4 k_yfwd   per (z, i-tile): forward y FFT + truncation in j -> 1r2w layout
//   5 k_zfwd   per (i, j-tile): forward z FFT, x 1/(NX NY NZ), truncation in k
//              (f_RK3TransformkRs2Fs, src/pf_rk3.f90:2242-2316; pf_fftw.f90:841-879)
//   k_formf    per mode: FormF + low-storage RK3 update + continuity fix
//              (src/pf_rk3.f90:2323-2673, 3457-3531)
//   k_tracers_col  per tracer (one lane per molecule column): Neville interpolation + RK3 + wrap
//              (src/pp_interp.f90:313-560, src/pp_rk3.f90:57-303, src/pp_mpi.f90:1011-1057)
//   k_remesh   per (i, j) line: integer k-index gather of the Rogallo remesh
//              (src/pf_rk3.f90:4056-4178)
// All kernels are HBM-bandwidth bound; none is a dense contraction.
#include "pdr_internal.hpp"
#include "pdr_rng.hpp"
#include "pdr_rsstats.hpp"
#include "pdr_formf.hpp"
#include "pdr_fft2.hpp"

namespace pdr {

__device__ __forceinline__ double wvn(int k, int kmax) {
  return (k <= kmax) ? double(k) : double(k - (2 * kmax + 1));
}

template <typename T>
__device__ __forceinline__ cx<T> zero_c() {
  return mk<T>(T(0), T(0));
}

template <typename T>
__device__ __forceinline__ cx<T> conj_c(cx<T> a) {
  return mk<T>(a.x, -a.y);
}

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// pass 1: vorticity + inverse z FFT
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_zinv(Geo g, Coef cf, const cx<T>* __restrict__ u0, const cx<T>* __restrict__ u1,
           const cx<T>* __restrict__ u2, cx<T>* __restrict__ w1, FftPlan pz,
           const cx<T>* __restrict__ tw, int tj, int fsel) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nz = g.nz, K = g.kk, nkk = g.nkk, nkj = g.nkj;
  const int ld = nz + 1;
  cx<T>* A = reinterpret_cast<cx<T>*>(smem);
  cx<T>* B = A + tj * ld;
  const int i = blockIdx.y;
  const int j0 = blockIdx.x * tj;
  const double wi = wvn(i, g.ki);
  const int f_begin = (fsel < 0) ? 0 : fsel;
  const int f_end = (fsel < 0) ? 6 : fsel + 1;
  for (int f = f_begin; f < f_end; ++f) {
    for (int idx = threadIdx.x; idx < tj * nz; idx += blockDim.x) {
      const int line = idx / nz;
      const int z = idx - line * nz;
      const int j = j0 + line;
      const int k = (z <= K) ? z : ((z >= nz - K) ? z - (nz - nkk) : -1);
      cx<T> v = zero_c<T>();
      if (j < nkj && k >= 0) {
        const long long e = ((long long)i * nkj + j) * nkk + k;
        if (f == 0) {
          v = u0[e];
        } else if (f == 1) {
          v = u1[e];
        } else if (f == 2) {
          v = u2[e];
        } else {
          const double kv[3] = {wi, wvn(j, g.kj), wvn(k, K)};
          const int c = f - 3;
          const cx<T> un[3] = {u0[e], u1[e], u2[e]};
          cx<T> acc = zero_c<T>();
#pragma unroll
          for (int n = 0; n < 3; ++n) {
            const double* G = &cf.G[(c * 3 + n) * 3];
            const T co = T(G[0] * kv[0] + G[1] * kv[1] + G[2] * kv[2]);
            acc = add(acc, scl(un[n], co));
          }
          v = mk<T>(-acc.y, acc.x);  // i * acc
        }
      }
      A[line * ld + z] = v;
    }
    __syncthreads();
    cx<T>* R = fft_lines<T, +1>(A, B, pz, tj, ld, tw);
    cx<T>* dst = w1 + (long long)f * g.m1 + (long long)i * nz * g.ldj;
    for (int idx = threadIdx.x; idx < tj * nz; idx += blockDim.x) {
      const int z = idx / tj;
      const int line = idx - z * tj;
      const int j = j0 + line;
      if (j < nkj) dst[(long long)z * g.ldj + j] = R[line * ld + z];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pass 2: inverse y FFT, 1r2w [f][i][z][j] -> x-pencils [f][z][y][i]
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_yinv(Geo g, const cx<T>* __restrict__ w1, cx<T>* __restrict__ w2, FftPlan py,
           const cx<T>* __restrict__ tw, int ti) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ny = g.ny, nz = g.nz, KJ = g.kj, nkj = g.nkj, nki = g.nki;
  const int ld = ny + 1;
  cx<T>* A = reinterpret_cast<cx<T>*>(smem);
  cx<T>* B = A + ti * ld;
  const int z = blockIdx.y;
  const int i0 = blockIdx.x * ti;
  for (int f = 0; f < 6; ++f) {
    const cx<T>* src = w1 + (long long)f * g.m1;
    for (int idx = threadIdx.x; idx < ti * ny; idx += blockDim.x) {
      const int line = idx / ny;
      const int y = idx - line * ny;
      const int i = i0 + line;
      const int j = (y <= KJ) ? y : ((y >= ny - KJ) ? y - (ny - nkj) : -1);
      cx<T> v = zero_c<T>();
      if (i < nki && j >= 0) v = src[((long long)i * nz + z) * g.ldj + j];
      A[line * ld + y] = v;
    }
    __syncthreads();
    cx<T>* R = fft_lines<T, +1>(A, B, py, ti, ld, tw);
    cx<T>* dst = w2 + (long long)f * g.m2 + (long long)z * ny * g.ldi;
    for (int idx = threadIdx.x; idx < ti * ny; idx += blockDim.x) {
      const int y = idx / ti;
      const int line = idx - y * ti;
      const int i = i0 + line;
      if (i < nki) dst[(long long)y * g.ldi + i] = R[line * ld + y];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pass 3: x c2r, nonlinear product, x r2c for one (z, y) row.
// Two real rows share one complex FFT: z = a + i b.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_xpass(Geo g, Coef cf, cx<T>* __restrict__ w2, T* __restrict__ ur, int want_tr, int want_cfl,
            unsigned long long* __restrict__ umax, FftPlan px, const cx<T>* __restrict__ tw,
            double* __restrict__ rsp) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ T red[kThreads];
  const int nx = g.nx, KI = g.ki, nki = g.nki;
  const int ld = nx + 1;
  cx<T>* A = reinterpret_cast<cx<T>*>(smem);
  cx<T>* B = A + 3 * ld;
  const int y = blockIdx.x;
  const int z = blockIdx.y;
  const long long row = ((long long)z * g.ny + y) * g.ldi;
  for (int idx = threadIdx.x; idx < 3 * nx; idx += blockDim.x) {
    const int p = idx / nx;
    const int m = idx - p * nx;
    cx<T> X1 = zero_c<T>(), X2 = zero_c<T>();
    const cx<T>* fa = w2 + (long long)(2 * p) * g.m2 + row;
    const cx<T>* fb = w2 + (long long)(2 * p + 1) * g.m2 + row;
    if (m <= KI) {
      X1 = fa[m];
      X2 = fb[m];
      if (m == 0) {  // c2r uses only the real part of the DC term
        X1.y = T(0);
        X2.y = T(0);
      }
    } else if (m >= nx - KI) {
      X1 = conj_c(fa[nx - m]);
      X2 = conj_c(fb[nx - m]);
    }
    A[p * ld + m] = mk<T>(X1.x - X2.y, X1.y + X2.x);
  }
  __syncthreads();
  cx<T>* R = fft_lines<T, +1>(A, B, px, 3, ld, tw);
  cx<T>* O = (R == A) ? B : A;
  T lmax = T(0);
  RsAcc acc;
  if (rsp) rs_init(acc);
  const T V[9] = {T(cf.V[0]), T(cf.V[1]), T(cf.V[2]), T(cf.V[3]), T(cf.V[4]),
                  T(cf.V[5]), T(cf.V[6]), T(cf.V[7]), T(cf.V[8])};
  const long long rr = ((long long)z * g.ny + y) * nx;
  for (int x = threadIdx.x; x < nx; x += blockDim.x) {
    const cx<T> a = R[x], b = R[ld + x], c = R[2 * ld + x];
    const T u1 = a.x, u2 = a.y, u3 = b.x;
    const T o1 = b.y, o2 = c.x, o3 = c.y;
    const T v1 = V[0] * u1 + V[1] * u2 + V[2] * u3;
    const T v2 = V[3] * u1 + V[4] * u2 + V[5] * u3;
    const T v3 = V[6] * u1 + V[7] * u2 + V[8] * u3;
    // omega x (binv u) (BF_ROT), or omega x u (isotropic with binv = I, R_ROT)
    const T w1 = cf.form ? u1 : v1, w2 = cf.form ? u2 : v2, w3 = cf.form ? u3 : v3;
    const T n1 = o2 * w3 - o3 * w2;
    const T n2 = o3 * w1 - o1 * w3;
    const T n3 = o1 * w2 - o2 * w1;
    if (want_tr) {
      T* d = ur + (rr + x) * 3;
      d[0] = u1;
      d[1] = u2;
      d[2] = u3;
    }
    if (want_cfl) {
      const T s = fabs(v1) + fabs(v2) + fabs(v3);  // lab-frame velocity (src/pf_fluidstats.f90:853-866, 914)
      lmax = s > lmax ? s : lmax;
    }
    if (rsp) rs_add(acc, double(v1), double(v2), double(v3), double(o1), double(o2), double(o3));
    O[x] = mk<T>(n1, n2);
    O[ld + x] = mk<T>(n3, T(0));
  }
  __syncthreads();
  cx<T>* Q = fft_lines<T, -1>(O, R, px, 2, ld, tw);
  cx<T>* d0 = w2 + row;
  cx<T>* d1 = w2 + g.m2 + row;
  cx<T>* d2 = w2 + 2 * g.m2 + row;
  for (int m = threadIdx.x; m < nki; m += blockDim.x) {
    const int mr = (m == 0) ? 0 : nx - m;
    const cx<T> Z = Q[m], Zc = conj_c(Q[mr]);
    const cx<T> s = add(Z, Zc), df = sub(Z, Zc);
    d0[m] = mk<T>(T(0.5) * s.x, T(0.5) * s.y);
    d1[m] = mk<T>(T(0.5) * df.y, T(-0.5) * df.x);  // (Z - conj Z') / (2i)
    const cx<T> W = Q[ld + m], Wc = conj_c(Q[ld + mr]);
    const cx<T> s3 = add(W, Wc);
    d2[m] = mk<T>(T(0.5) * s3.x, T(0.5) * s3.y);
  }
  if (want_cfl) {
    red[threadIdx.x] = lmax;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + s] ? red[threadIdx.x] : red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const double v = double(red[0]);
      atomicMax(umax, (unsigned long long)__double_as_longlong(v));
    }
  }
  if (rsp) {
    __shared__ double rs_lds[kRsWidth * (kThreads / 64)];
    rs_block_out(acc, rs_lds, rsp + ((long long)z * g.ny + y) * kRsWidth);
  }
}

// ---------------------------------------------------------------------------
// pass 3, NS_R_CONV: x c2r of u and its nine gradients (12 real rows in 6
// complex lines: wa holds u and du/dx, fields 3..5 of wb and wc hold du/dy and
// du/dz), the convective product NL_i = sum_j sum_k bmat(k,j) du_i/dx_k u_j
// (f_RK3ComputeNonlinearRConv, src/pf_rk3.f90:1717-1774), x r2c of NL into
// fields 0..2 of wa.  The tracer field, the CFL maximum and the real-space
// moments take u, binv u and fields 3..5 (du/dx), as the reference's
// f_FluidstatsRs2D reads u4..u6 (src/pf_fluidstats.f90:846-917).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_xpass_conv(Geo g, Coef cf, cx<T>* __restrict__ wa, const cx<T>* __restrict__ wb,
                 const cx<T>* __restrict__ wc, T* __restrict__ ur, int want_tr, int want_cfl,
                 unsigned long long* __restrict__ umax, FftPlan px, const cx<T>* __restrict__ tw,
                 double* __restrict__ rsp) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ T red[kThreads];
  const int nx = g.nx, KI = g.ki, nki = g.nki;
  const int ld = nx + 1;
  cx<T>* A = reinterpret_cast<cx<T>*>(smem);
  cx<T>* B = A + 6 * ld;
  const int y = blockIdx.x;
  const int z = blockIdx.y;
  const long long row = ((long long)z * g.ny + y) * g.ldi;
  auto field = [&](int q) -> const cx<T>* {  // real row q of the 12
    return (q < 6 ? wa + (long long)q * g.m2 : (q < 9 ? wb + (long long)(q - 3) * g.m2 : wc + (long long)(q - 6) * g.m2)) +
           row;
  };
  for (int idx = threadIdx.x; idx < 6 * nx; idx += blockDim.x) {
    const int p = idx / nx;
    const int m = idx - p * nx;
    cx<T> X1 = zero_c<T>(), X2 = zero_c<T>();
    const cx<T>* fa = field(2 * p);
    const cx<T>* fb = field(2 * p + 1);
    if (m <= KI) {
      X1 = fa[m];
      X2 = fb[m];
      if (m == 0) {  // c2r uses only the real part of the DC term
        X1.y = T(0);
        X2.y = T(0);
      }
    } else if (m >= nx - KI) {
      X1 = conj_c(fa[nx - m]);
      X2 = conj_c(fb[nx - m]);
    }
    A[p * ld + m] = mk<T>(X1.x - X2.y, X1.y + X2.x);
  }
  __syncthreads();
  cx<T>* R = fft_lines<T, +1>(A, B, px, 6, ld, tw);
  cx<T>* O = (R == A) ? B : A;
  T lmax = T(0);
  RsAcc acc;
  if (rsp) rs_init(acc);
  const T V[9] = {T(cf.V[0]), T(cf.V[1]), T(cf.V[2]), T(cf.V[3]), T(cf.V[4]),
                  T(cf.V[5]), T(cf.V[6]), T(cf.V[7]), T(cf.V[8])};
  const T Bk[9] = {T(cf.Bm[0]), T(cf.Bm[1]), T(cf.Bm[2]), T(cf.Bm[3]), T(cf.Bm[4]),
                   T(cf.Bm[5]), T(cf.Bm[6]), T(cf.Bm[7]), T(cf.Bm[8])};  // [k][j]
  const long long rr = ((long long)z * g.ny + y) * nx;
  for (int x = threadIdx.x; x < nx; x += blockDim.x) {
    T f[12];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const cx<T> v = R[p * ld + x];
      f[2 * p] = v.x;
      f[2 * p + 1] = v.y;
    }
    const T u[3] = {f[0], f[1], f[2]};
    T nl[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T a = T(0);
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) a = a + (Bk[3 * k + j] * f[3 + 3 * k + i] * u[j]);
      nl[i] = a;
    }
    const T v1 = V[0] * u[0] + V[1] * u[1] + V[2] * u[2];
    const T v2 = V[3] * u[0] + V[4] * u[1] + V[5] * u[2];
    const T v3 = V[6] * u[0] + V[7] * u[1] + V[8] * u[2];
    if (want_tr) {
      T* d = ur + (rr + x) * 3;
      d[0] = u[0];
      d[1] = u[1];
      d[2] = u[2];
    }
    if (want_cfl) {
      const T sm = fabs(v1) + fabs(v2) + fabs(v3);
      lmax = sm > lmax ? sm : lmax;
    }
    if (rsp) rs_add(acc, double(v1), double(v2), double(v3), double(f[3]), double(f[4]), double(f[5]));
    O[x] = mk<T>(nl[0], nl[1]);
    O[ld + x] = mk<T>(nl[2], T(0));
  }
  __syncthreads();
  cx<T>* Q = fft_lines<T, -1>(O, R, px, 2, ld, tw);
  cx<T>* d0 = wa + row;
  cx<T>* d1 = wa + g.m2 + row;
  cx<T>* d2 = wa + 2 * g.m2 + row;
  for (int m = threadIdx.x; m < nki; m += blockDim.x) {
    const int mr = (m == 0) ? 0 : nx - m;
    const cx<T> Z = Q[m], Zc = conj_c(Q[mr]);
    const cx<T> s = add(Z, Zc), df = sub(Z, Zc);
    d0[m] = mk<T>(T(0.5) * s.x, T(0.5) * s.y);
    d1[m] = mk<T>(T(0.5) * df.y, T(-0.5) * df.x);
    const cx<T> W = Q[ld + m], Wc = conj_c(Q[ld + mr]);
    const cx<T> s3 = add(W, Wc);
    d2[m] = mk<T>(T(0.5) * s3.x, T(0.5) * s3.y);
  }
  if (want_cfl) {
    red[threadIdx.x] = lmax;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + s] ? red[threadIdx.x] : red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(umax, (unsigned long long)__double_as_longlong(double(red[0])));
  }
  if (rsp) {
    __shared__ double rs_lds[kRsWidth * (kThreads / 64)];
    rs_block_out(acc, rs_lds, rsp + ((long long)z * g.ny + y) * kRsWidth);
  }
}

// ---------------------------------------------------------------------------
// pass 4: forward y FFT + truncation, x-pencils -> 1r2w (NL fields 0..2)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_yfwd(Geo g, const cx<T>* __restrict__ w2, cx<T>* __restrict__ w1, FftPlan py,
           const cx<T>* __restrict__ tw, int ti) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ny = g.ny, nz = g.nz, nkj = g.nkj, nki = g.nki;
  const int ld = ny + 1;
  cx<T>* A = reinterpret_cast<cx<T>*>(smem);
  cx<T>* B = A + ti * ld;
  const int z = blockIdx.y;
  const int i0 = blockIdx.x * ti;
  for (int f = 0; f < 3; ++f) {
    const cx<T>* src = w2 + (long long)f * g.m2 + (long long)z * ny * g.ldi;
    for (int idx = threadIdx.x; idx < ti * ny; idx += blockDim.x) {
      const int y = idx / ti;
      const int line = idx - y * ti;
      const int i = i0 + line;
      A[line * ld + y] = (i < nki) ? src[(long long)y * g.ldi + i] : zero_c<T>();
    }
    __syncthreads();
    cx<T>* R = fft_lines<T, -1>(A, B, py, ti, ld, tw);
    cx<T>* dst = w1 + (long long)f * g.m1;
    for (int idx = threadIdx.x; idx < ti * nkj; idx += blockDim.x) {
      const int line = idx / nkj;
      const int j = idx - line * nkj;
      const int i = i0 + line;
      if (i < nki) {
        const int yy = (j <= g.kj) ? j : ny - nkj + j;
        dst[((long long)i * nz + z) * g.ldj + j] = R[line * ld + yy];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pass 5: forward z FFT, normalisation, truncation -> NL [c][i][j][k]
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_zfwd(Geo g, const cx<T>* __restrict__ w1, cx<T>* __restrict__ nl, FftPlan pz,
           const cx<T>* __restrict__ tw, int tj, T norm) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nz = g.nz, nkj = g.nkj, nkk = g.nkk;
  const int ld = nz + 1;
  cx<T>* A = reinterpret_cast<cx<T>*>(smem);
  cx<T>* B = A + tj * ld;
  const int i = blockIdx.y;
  const int j0 = blockIdx.x * tj;
  for (int f = 0; f < 3; ++f) {
    const cx<T>* src = w1 + (long long)f * g.m1 + (long long)i * nz * g.ldj;
    for (int idx = threadIdx.x; idx < tj * nz; idx += blockDim.x) {
      const int z = idx / tj;
      const int line = idx - z * tj;
      const int j = j0 + line;
      A[line * ld + z] = (j < nkj) ? src[(long long)z * g.ldj + j] : zero_c<T>();
    }
    __syncthreads();
    cx<T>* R = fft_lines<T, -1>(A, B, pz, tj, ld, tw);
    cx<T>* dst = nl + (long long)f * g.mw;
    for (int idx = threadIdx.x; idx < tj * nkk; idx += blockDim.x) {
      const int line = idx / nkk;
      const int k = idx - line * nkk;
      const int j = j0 + line;
      if (j < nkj) {
        const int zz = (k <= g.kk) ? k : nz - nkk + k;
        dst[((long long)i * nkj + j) * nkk + k] = scl(R[line * ld + zz], norm);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// FormF + RK3 + continuity, one thread per mode
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_formf(Geo g, Coef cf, cx<T>* __restrict__ u0, cx<T>* __restrict__ u1, cx<T>* __restrict__ u2,
            cx<T>* __restrict__ r0, cx<T>* __restrict__ r1, cx<T>* __restrict__ r2,
            const cx<T>* __restrict__ nl, const double2* __restrict__ frc, int fnk, double fkf, double beta,
            double gdt) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.mw) return;
  const int k = int(e % g.nkk);
  const long long ij = e / g.nkk;
  const int j = int(ij % g.nkj);
  const int i = int(ij / g.nkj);
  formf_mode<T>(g, cf, i, j, k, e, nl[e], nl[g.mw + e], nl[2 * g.mw + e], u0, u1, u2, r0, r1, r2, frc, fnk, fkf,
                beta, gdt);
}

// ---------------------------------------------------------------------------
// Ornstein-Uhlenbeck forcing update, f_ForceCalc (src/pf_force.f90:106-268).
// F is [i][j+NK][k+NK][c] complex, i = 0..NK, j, k = -NK..NK.  Thread t < (NK+1)^2
// runs one (j, k) iteration of the i = 0 loop (j, k = 0..NK) with its
// conjugate-symmetric copies, in the reference's statement order (for k = 0
// the (0, -j, 0) update reads the copy just written, as there); the other
// threads update one i >= 1 mode each.  The iterations touch disjoint modes.
// Normal pairs come from normal_pair(seed, call * 2^24 + slot); draws of one
// call never repeat another call's.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
    k_force_ou(double2* __restrict__ F, int nk, double om, double sd, unsigned long long seed,
               unsigned long long call) {
  const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
  const int w = 2 * nk + 1;
  const int n0 = (nk + 1) * (nk + 1);
  const int n1 = nk * w * w;
  if (t >= n0 + n1) return;
  auto at = [&](int i, int j, int k) { return F + ((i * w + j + nk) * w + k + nk) * 3; };
  auto ou = [&](double2* p, unsigned long long key) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double re, im;
      pdr::normal_pair(seed, key + unsigned(c), &re, &im);
      p[c].x = p[c].x * om + re * sd;
      p[c].y = p[c].y * om + im * sd;
    }
  };
  const unsigned long long key0 = call << 24;
  if (t < n0) {
    const int j = t / (nk + 1), k = t - (t / (nk + 1)) * (nk + 1);
    double2* p = at(0, j, k);
    ou(p, key0 + 6ull * unsigned(t));
    double2* m = at(0, -j, -k);
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] = make_double2(p[c].x, -p[c].y);
    if (j > 0) {
      double2* q = at(0, -j, k);
      ou(q, key0 + 6ull * unsigned(t) + 3ull);
      double2* r = at(0, j, -k);
#pragma unroll
      for (int c = 0; c < 3; ++c) r[c] = make_double2(q[c].x, -q[c].y);
    }
  } else {
    const int u = t - n0;
    const int i = 1 + u / (w * w);
    const int rem = u - (i - 1) * w * w;
    ou(at(i, rem / w - nk, rem % w - nk), key0 + 6ull * unsigned(n0) + 3ull * unsigned(u));
  }
}

__global__ void __launch_bounds__(kThreads) k_zero_d2(double2* __restrict__ p, int n) {
  const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (t < n) p[t] = make_double2(0.0, 0.0);
}

// ---------------------------------------------------------------------------
// Spherical / LES truncation (f_ArraysDealiasSpherical, f_ArraysLESCutoff,
// src/pf_arrays.f90:1865-1987): zero u1..u3 where |bmat^T k| > cutoff.  The
// decision needs only the indices and the frame, so the kernel reads nothing
// and writes the masked modes only.
// ---------------------------------------------------------------------------
struct MaskFrame {
  double bm[9];  // bmat over the existing entries, [m][n]
  double cutoff;
};

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_dealias(Geo g, MaskFrame mf, cx<T>* __restrict__ u0, cx<T>* __restrict__ u1, cx<T>* __restrict__ u2) {
#pragma clang fp contract(off)
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < g.mw; e += stride) {
    const int k = int(e % g.nkk);
    const long long ij = e / g.nkk;
    const int j = int(ij % g.nkj);
    const int i = int(ij / g.nkj);
    const double kv[3] = {wvn(i, g.ki), wvn(j, g.kj), wvn(k, g.kk)};
    double kl[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 3; ++n) kl[n] = kl[n] + mf.bm[3 * m + n] * kv[m];
    const double km = sqrt((kl[0] * kl[0]) + (kl[1] * kl[1]) + (kl[2] * kl[2]));
    if (km > mf.cutoff) {
      u0[e] = zero_c<T>();
      u1[e] = zero_c<T>();
      u2[e] = zero_c<T>();
    }
  }
}

// ---------------------------------------------------------------------------
// Tracers
// ---------------------------------------------------------------------------
__device__ __forceinline__ long long pmod(long long a, long long n) {
  long long r = a % n;
  return r < 0 ? r + n : r;
}

// p_MPICorrectPosition (src/pp_mpi.f90:1011-1057): one shift per axis into
// [-dx/2, L - dx/2); the x3 upper test is '>', so x3 = x3max stays.
__device__ __forceinline__ void wrap_position(const TracerFrame& tf, double x[3]) {
  const double x1min = -0.5 * tf.dx, x1max = x1min + tf.lx;
  const double x2min = -0.5 * tf.dy, x2max = x2min + tf.ly;
  const double x3min = -0.5 * tf.dz, x3max = x3min + tf.lz;
  if (x[0] < x1min) x[0] = x[0] + tf.lx;
  if (x[0] >= x1max) x[0] = x[0] - tf.lx;
  if (x[1] < x2min) x[1] = x[1] + tf.ly;
  if (x[1] >= x2max) x[1] = x[1] - tf.ly;
  if (x[2] < x3min) x[2] = x[2] + tf.lz;
  if (x[2] > x3max) x[2] = x[2] - tf.lz;
}

// p_ControlFinishTimeStep alone (src/pp_control.f90:1440-1450 ->
// p_ControlUpdateCount, 1454-1536): the periodic correction of every position,
// with particle_xp0 moved by the same shift (p_ControlUpdateInitial,
// src/pp_control.f90:1574-1585).  One thread per tracer.
__global__ void __launch_bounds__(kThreads)
    k_tracers_wrap(TracerFrame tf, double* __restrict__ xpa, double* __restrict__ xp0, long long np) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  double x[3] = {xpa[3 * q], xpa[3 * q + 1], xpa[3 * q + 2]};
  const double xu[3] = {x[0], x[1], x[2]};
  wrap_position(tf, x);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    xpa[3 * q + m] = x[m];
    if (xp0) xp0[3 * q + m] = (xp0[3 * q + m] + x[m]) - xu[m];
  }
}

// ---------------------------------------------------------------------------
// Tracers, one lane per x column of the molecule: a group of G lanes
// (G = tracer_group(P), at least 4) serves one tracer; lane a < P gathers the
// P x P points (y, z) of column x_a, all three components per point (one
// 3-element load each), and applies the y then z Neville passes to them; the
// group exchanges the P column values with __shfl and every lane runs the x
// pass, the RK3 update and the wrap; lane c < 3 stores component c.  The
// operation sequence per column and component is that of the reference
// (y for each (x, z), then z for each x, then x).
// ---------------------------------------------------------------------------
// Lanes per tracer: P rounded up to a power of two, and at least 4, since lane
// c < 3 stores component c of the results.
constexpr int tracer_group(int p) { return p <= 4 ? 4 : 8; }

// Workgroups take consecutive tracer slots; with the slots sorted by grid row,
// the workgroups of one XCD (renumbered by xcd_remap) then walk one contiguous
// slab of rows, and the stencil rows that neighbouring workgroups share are
// fetched into that XCD's L2 once instead of into several.
#ifndef PDR_TR_XCD
#define PDR_TR_XCD 1
#endif
#ifndef PDR_TR_RCP
#define PDR_TR_RCP 1
#endif
// Four waves per SIMD (at most 128 registers) for the orders up to 4: the gather
// is latency-bound, and without the hint the FP64 form takes 134 registers and
// three waves.
#ifndef PDR_TR_WAVES
#define PDR_TR_WAVES 4
#endif
template <typename T, int P>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(P <= 4 ? PDR_TR_WAVES : 1)))
    k_tracers_col(Geo g, TracerFrame tf, const T* __restrict__ ur, double* __restrict__ xpa,
                  double* __restrict__ uprk, double* __restrict__ upa, double* __restrict__ xp0,
                  long long np, double beta, double gdt, int wrap) {
  constexpr int G = tracer_group(P);
  const unsigned blk = PDR_TR_XCD ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const long long gt = (long long)blk * blockDim.x + threadIdx.x;
  const long long q = gt / G;
  const int a = int(gt % G);
  const bool live = q < np;
  const long long qq = live ? q : np - 1;  // np >= 1: dead lanes read a valid tracer
  double x[3] = {xpa[3 * qq], xpa[3 * qq + 1], xpa[3 * qq + 2]};
  const int nhalo = P / 2;
  // g_DomainInitHalo (src/pg_domain.f90:360-384), p_InterpFindMolecule (src/pp_interp.f90:406-467):
  // even orders shift the halo length by half a cell and the search by one and a half,
  // odd orders only the search by half a cell (the molecule is centred on the nearest node)
  double lxh = double(nhalo) * tf.dx, lyh = double(nhalo) * tf.dy, lzh = double(nhalo) * tf.dz;
  double xd1, xd2, xd3;
  if constexpr (P % 2 == 0) {
    lxh = lxh - (0.5 * tf.dx);
    lyh = lyh - (0.5 * tf.dy);
    lzh = lzh - (0.5 * tf.dz);
    xd1 = x[0] + lxh + tf.dx + (0.5 * tf.dx);
    xd2 = x[1] + lyh + tf.dy + (0.5 * tf.dy);
    xd3 = x[2] + lzh + tf.dz + (0.5 * tf.dz);
  } else {
    xd1 = x[0] + lxh + (0.5 * tf.dx);
    xd2 = x[1] + lyh + (0.5 * tf.dy);
    xd3 = x[2] + lzh + (0.5 * tf.dz);
  }
  const int i0 = int(xd1 / tf.dx) - nhalo - nhalo;
  const int j0 = int(xd2 / tf.dy) - nhalo - nhalo;
  const int k0 = int(xd3 / tf.dz) - nhalo - nhalo;
  double iy[P], iz[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    iy[m] = double(j0 + m) * tf.dy;
    iz[m] = double(k0 + m) * tf.dz;
  }
  // Each tableau entry is ((d1 v) + (d2 v')) / d3 with d3 a node difference that
  // depends only on (col, nv) and the axis: the P (P - 1) / 2 reciprocals of an
  // axis are formed once and the entries multiply by them (one FP64 division per
  // (col, nv) instead of one per row, component and entry; the entries then
  // differ from the quotients by at most an ulp of rounding).
  auto recips = [&](const double* nd, double (&rd)[P][P]) {
#pragma unroll
    for (int col = 1; col <= P - 1; ++col)
#pragma unroll
      for (int nv = 1; nv <= P - col; ++nv) rd[col][nv - 1] = 1.0 / (nd[nv - 1] - nd[nv + col - 1]);
  };
  double vc[3] = {0.0, 0.0, 0.0};  // this lane's column after the y and z passes
  if (a < P) {
    double ry[P][P];
    if (PDR_TR_RCP) recips(iy, ry);
    const long long gx3 = 3 * pmod(i0 + a, g.nx);
    double vz[3][P];  // [component][z]
#pragma unroll
    for (int qz = 0; qz < P; ++qz) {
      const long long zoff = pmod(k0 + qz, g.nz) * g.ny;
      T w[P][3];  // [y][component]
#pragma unroll
      for (int b = 0; b < P; ++b) {
        const T* pt = ur + (zoff + pmod(j0 + b, g.ny)) * g.nx * 3 + gx3;
        w[b][0] = pt[0];
        w[b][1] = pt[1];
        w[b][2] = pt[2];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double v[P];
#pragma unroll
        for (int b = 0; b < P; ++b) v[b] = double(w[b][c]);
#pragma unroll
        for (int col = 1; col <= P - 1; ++col)
#pragma unroll
          for (int nv = 1; nv <= P - col; ++nv) {
            const double d1 = x[1] - iy[nv + col - 1];
            const double d2 = iy[nv - 1] - x[1];
            const double d3 = iy[nv - 1] - iy[nv + col - 1];
            v[nv - 1] = PDR_TR_RCP ? ((d1 * v[nv - 1]) + (d2 * v[nv])) * ry[col][nv - 1]
                                   : ((d1 * v[nv - 1]) + (d2 * v[nv])) / d3;
          }
        vz[c][qz] = v[0];
      }
    }
    double rz[P][P];
    if (PDR_TR_RCP) recips(iz, rz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int col = 1; col <= P - 1; ++col)
#pragma unroll
        for (int nv = 1; nv <= P - col; ++nv) {
          const double d1 = x[2] - iz[nv + col - 1];
          const double d2 = iz[nv - 1] - x[2];
          const double d3 = iz[nv - 1] - iz[nv + col - 1];
          vz[c][nv - 1] = PDR_TR_RCP ? ((d1 * vz[c][nv - 1]) + (d2 * vz[c][nv])) * rz[col][nv - 1]
                                     : ((d1 * vz[c][nv - 1]) + (d2 * vz[c][nv])) / d3;
        }
      vc[c] = vz[c][0];
    }
  }
  // x pass over the P columns of the group
  const int lane0 = int(threadIdx.x & 63u) - a;  // first lane of the group within the wave
  double uf[3];
  double ix[P], rx[P][P];
#pragma unroll
  for (int m = 0; m < P; ++m) ix[m] = double(i0 + m) * tf.dx;
  if (PDR_TR_RCP) recips(ix, rx);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double vx[P];
#pragma unroll
    for (int m = 0; m < P; ++m) vx[m] = __shfl(vc[c], lane0 + m);
#pragma unroll
    for (int col = 1; col <= P - 1; ++col)
#pragma unroll
      for (int nv = 1; nv <= P - col; ++nv) {
        const double xa = ix[nv + col - 1], xb = ix[nv - 1];
        const double d1 = x[0] - xa;
        const double d2 = xb - x[0];
        const double d3 = xb - xa;
        vx[nv - 1] = PDR_TR_RCP ? ((d1 * vx[nv - 1]) + (d2 * vx[nv])) * rx[col][nv - 1]
                                : ((d1 * vx[nv - 1]) + (d2 * vx[nv])) / d3;
      }
    uf[c] = vx[0];
  }
  double r[3] = {uprk[3 * qq], uprk[3 * qq + 1], uprk[3 * qq + 2]};
  double up[3];  // particle_up: the fluid velocity, laboratory frame
  if (!tf.bf) {
    // p_RK3SolveParticle tracer branch (src/pp_rk3.f90:103-134)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      up[m] = uf[m];
      r[m] = (beta * r[m]) + uf[m];
      x[m] = x[m] + (gdt * r[m]);
    }
  } else {
    // p_RK3SolveParticleBF tracer branch (src/pp_rk3.f90:223-288)
    double xl[3] = {0.0, 0.0, 0.0}, ul[3] = {0.0, 0.0, 0.0}, xm[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        xl[m] = xl[m] + tf.binv[3 * m + n] * x[n];
        ul[m] = ul[m] + tf.binv[3 * m + n] * uf[n];
      }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      up[m] = ul[m];
      r[m] = (beta * r[m]) + ul[m];
      xl[m] = xl[m] + (gdt * r[m]);
    }
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 3; ++n) xm[m] = xm[m] + tf.bmat[3 * m + n] * xl[n];
    x[0] = xm[0];
    x[1] = xm[1];
    x[2] = xm[2];
  }
  const double xu[3] = {x[0], x[1], x[2]};
  if (wrap) wrap_position(tf, x);  // p_ControlFinishTimeStep, fused into the advance
  if (live && a < 3) {
    const double xa = (a == 0) ? x[0] : ((a == 1) ? x[1] : x[2]);
    xpa[3 * q + a] = xa;
    uprk[3 * q + a] = (a == 0) ? r[0] : ((a == 1) ? r[1] : r[2]);
    upa[3 * q + a] = (a == 0) ? up[0] : ((a == 1) ? up[1] : up[2]);
    if (xp0 && wrap) {
      // p_ControlUpdateInitial (src/pp_control.f90:1574-1585): xp0 + x_wrapped - x,
      // left to right, on every stage
      const double xb = (a == 0) ? xu[0] : ((a == 1) ? xu[1] : xu[2]);
      xp0[3 * q + a] = (xp0[3 * q + a] + xa) - xb;
    }
  }
}

// ---------------------------------------------------------------------------
// Remesh: per (i, j) line new[s] = old[s + i] for s + i <= K (signed k indices),
// BF_ROT: u1 <- u1' + u3'.  In place; one workgroup per line.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_remesh(Geo g, cx<T>* __restrict__ u0, cx<T>* __restrict__ u1, cx<T>* __restrict__ u2, int bfrot) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  const int nkk = g.nkk, K = g.kk;
  const int j = blockIdx.x, i = blockIdx.y;
  const long long base = ((long long)i * g.nkj + j) * nkk;
  for (int k = threadIdx.x; k < nkk; k += blockDim.x) {
    buf[k] = u0[base + k];
    buf[nkk + k] = u1[base + k];
    buf[2 * nkk + k] = u2[base + k];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nkk; k += blockDim.x) {
    const int s = (k <= K) ? k : k - nkk;  // new signed index
    const int a = s + i;                   // old signed index
    cx<T> b1 = zero_c<T>(), b2 = zero_c<T>(), b3 = zero_c<T>();
    if (a <= K) {
      const int ko = (a >= 0) ? a : a + nkk;
      b1 = buf[ko];
      b2 = buf[nkk + ko];
      b3 = buf[2 * nkk + ko];
    }
    u0[base + k] = bfrot ? add(b1, b3) : b1;
    u1[base + k] = b2;
    u2[base + k] = b3;
  }
}

// p_MPIRemeshPosition (src/pp_mpi.f90:1062-1092)
__global__ void k_remesh_particles(double* __restrict__ xp, double* __restrict__ xp0, long long np, double lx,
                                   double dx) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const double xs = xp[3 * q] + xp[3 * q + 2];
  const double x1min = -0.5 * dx, x1max = x1min + lx;
  if (xp0) {
    // p_MPIRemeshInitial (src/pp_mpi.f90:1095-1129): x0_1 += x0_3, shifted as x1 + x3 wraps
    double x01 = xp0[3 * q] + xp0[3 * q + 2];
    if (xs < x1min) x01 = x01 + lx;
    if (xs >= x1max) x01 = x01 - lx;
    xp0[3 * q] = x01;
  }
  double x1 = xs;
  if (x1 < x1min) x1 = x1 + lx;
  if (x1 >= x1max) x1 = x1 - lx;
  xp[3 * q] = x1;
}

template <typename T>
__global__ void k_finite(const T* __restrict__ a, long long n, int* flag) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    if (!isfinite(a[q])) {
      atomicOr(flag, 1);
      return;
    }
  }
}

// Plain streaming copy (16-byte vectors): the measured HBM reference for the
// roofline and the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters.
__global__ void __launch_bounds__(kThreads) k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    b[q] = a[q];
}

hipError_t launch_copy(const void* a, void* b, long long bytes, hipStream_t s) {
  const long long n = bytes / 16;
  k_copy<<<dim3(8192), kThreads, 0, s>>>((const float4*)a, (float4*)b, n);
  return hipGetLastError();
}

// HBM ceiling probes: every thread owns U 16-byte vectors spaced one workgroup
// apart (coalesced), issues all U loads before any store, and writes with
// non-temporal stores, so each wave keeps U * 1 KiB of reads in flight.
// KIND 0: copy d = a; 1: read (sum of a, kept live through a never-taken
// store); 2: write (fill d); 3: triad d = a + s b.
template <int KIND, int U>
__global__ void __launch_bounds__(kThreads)
    k_probe(const float4* __restrict__ a, const float4* __restrict__ b, float4* __restrict__ d, long long n, float s,
            float* __restrict__ sink) {
  const long long base = (long long)blockIdx.x * (kThreads * U) + threadIdx.x;
  typedef float v4 __attribute__((ext_vector_type(4)));
  if constexpr (KIND == 2) {
    const v4 f = {s, s, s, s};
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long q = base + (long long)u * kThreads;
      if (q < n) __builtin_nontemporal_store(f, reinterpret_cast<v4*>(d) + q);
    }
  } else {
    v4 va[U], vb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long q = base + (long long)u * kThreads;
      const v4 z = {0.f, 0.f, 0.f, 0.f};
      va[u] = q < n ? reinterpret_cast<const v4*>(a)[q] : z;
      if constexpr (KIND == 3) vb[u] = q < n ? reinterpret_cast<const v4*>(b)[q] : z;
    }
    if constexpr (KIND == 1) {
      float acc = 0.f;
#pragma unroll
      for (int u = 0; u < U; ++u) acc += (va[u].x + va[u].y) + (va[u].z + va[u].w);
      if (acc == -1.0f) sink[threadIdx.x] = acc;  // never true for the probe's data (all ones)
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long q = base + (long long)u * kThreads;
        v4 v = va[u];
        if constexpr (KIND == 3) v = va[u] + s * vb[u];
        if (q < n) __builtin_nontemporal_store(v, reinterpret_cast<v4*>(d) + q);
      }
    }
  }
}

hipError_t launch_probe(int kind, const void* a, const void* b, void* d, long long bytes, float* sink,
                        hipStream_t s) {
  constexpr int U = 8;
  const long long n = bytes / 16;
  const dim3 grid(unsigned((n + kThreads * U - 1) / (kThreads * U)));
  const float4* fa = (const float4*)a;
  const float4* fb = (const float4*)b;
  float4* fd = (float4*)d;
  switch (kind) {
    case 0: k_probe<0, U><<<grid, kThreads, 0, s>>>(fa, fb, fd, n, 1.0f, sink); break;
    case 1: k_probe<1, U><<<grid, kThreads, 0, s>>>(fa, fb, fd, n, 1.0f, sink); break;
    case 2: k_probe<2, U><<<grid, kThreads, 0, s>>>(fa, fb, fd, n, 1.0f, sink); break;
    default: k_probe<3, U><<<grid, kThreads, 0, s>>>(fa, fb, fd, n, 1.0f, sink); break;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <typename T>
static hipError_t set_lds(const void* fn, size_t bytes) {
  if (bytes > 65536) return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
  return hipSuccess;
}

template <typename T>
static hipError_t zinv_t(pdr_ctx* c, const Coef& cf, int fsel, void* w1) {
  const Geo& g = c->g;
  const int tj = c->tj_z;
  const size_t lds = 2 * size_t(tj) * (g.nz + 1) * sizeof(cx<T>);
  hipError_t e = set_lds<T>((const void*)k_zinv<T>, lds);
  if (e != hipSuccess) return e;
  dim3 grid((g.nkj + tj - 1) / tj, g.nki);
  k_zinv<T><<<grid, kThreads, lds, c->stream>>>(g, cf, (const cx<T>*)c->u[0], (const cx<T>*)c->u[1],
                                                (const cx<T>*)c->u[2], (cx<T>*)w1, c->pz,
                                                (const cx<T>*)c->tab.twz, tj, fsel);
  return hipGetLastError();
}

hipError_t launch_zinv(pdr_ctx* c, const Coef& cf) {
  prof_begin(c, K_ZINV);
  hipError_t e = c->spec   ? launch2_zinv(c, cf)
                 : c->prec ? zinv_t<float>(c, cf, -1, c->w1)
                           : zinv_t<double>(c, cf, -1, c->w1);
  prof_end(c, K_ZINV);
  return e;
}

template <typename T>
static hipError_t yinv_t(pdr_ctx* c, const void* w1, void* w2) {
  const Geo& g = c->g;
  const int ti = c->ti_y;
  const size_t lds = 2 * size_t(ti) * (g.ny + 1) * sizeof(cx<T>);
  hipError_t e = set_lds<T>((const void*)k_yinv<T>, lds);
  if (e != hipSuccess) return e;
  dim3 grid((g.nki + ti - 1) / ti, g.nz);
  k_yinv<T><<<grid, kThreads, lds, c->stream>>>(g, (const cx<T>*)w1, (cx<T>*)w2, c->py,
                                                (const cx<T>*)c->tab.twy, ti);
  return hipGetLastError();
}

// The generic y/x kernels work on the whole box (the buffer is not chunked
// for them: zc = NZ); the specialised ones take a chunk of z planes.
hipError_t launch_yinv(pdr_ctx* c, int z0, int nzc) {
  prof_begin(c, K_YINV);
  hipError_t e = c->spec   ? launch2_yinv(c, z0, nzc)
                 : c->prec ? yinv_t<float>(c, c->w1, c->w2)
                           : yinv_t<double>(c, c->w1, c->w2);
  prof_end(c, K_YINV);
  return e;
}

template <typename T>
static hipError_t xpass_t(pdr_ctx* c, const Coef& cf, bool tr, bool cfl, double* rsp) {
  const Geo& g = c->g;
  const size_t lds = 6 * size_t(g.nx + 1) * sizeof(cx<T>);
  hipError_t e = set_lds<T>((const void*)k_xpass<T>, lds);
  if (e != hipSuccess) return e;
  dim3 grid(g.ny, g.nz);
  k_xpass<T><<<grid, kThreads, lds, c->stream>>>(g, cf, (cx<T>*)c->w2, (T*)c->ur, tr ? 1 : 0, cfl ? 1 : 0,
                                                 c->umax, c->px, (const cx<T>*)c->tab.twx, rsp);
  return hipGetLastError();
}

template <typename T>
static hipError_t xpass_conv_t(pdr_ctx* c, const Coef& cf, bool tr, bool cfl, double* rsp) {
  const Geo& g = c->g;
  const size_t lds = conv_xpass_lds(g.nx, c->csz);
  hipError_t e = set_lds<T>((const void*)k_xpass_conv<T>, lds);
  if (e != hipSuccess) return e;
  dim3 grid(g.ny, g.nz);
  k_xpass_conv<T><<<grid, kThreads, lds, c->stream>>>(g, cf, (cx<T>*)c->w2, (const cx<T>*)c->w2b,
                                                      (const cx<T>*)c->w2c, (T*)c->ur, tr ? 1 : 0, cfl ? 1 : 0,
                                                      c->umax, c->px, (const cx<T>*)c->tab.twx, rsp);
  return hipGetLastError();
}

Coef conv_coef(const Coef& cf, int d) {
  Coef r = cf;
  for (int q = 0; q < 27; ++q) r.G[q] = 0.0;
  for (int cc = 0; cc < 3; ++cc) r.G[(cc * 3 + cc) * 3 + d] = 1.0;  // field 3 + c = i k_d u_c
  return r;
}

// NS_R_CONV front of a stage (generic kernels, whole box): three inverse z/y
// sweeps give u with du/dx, du/dy and du/dz (fields 3..5 of w, wb, wc), then
// the convective x pass leaves NL in fields 0..2 of the x-pencil buffer.
hipError_t launch_conv_front(pdr_ctx* c, const Coef& cf, bool tr, bool cfl, double* rsp) {
  void* w1s[3] = {c->w1, c->w1b, c->w1c};
  void* w2s[3] = {c->w2, c->w2b, c->w2c};
  hipError_t e = hipSuccess;
  prof_begin(c, K_ZINV);
  for (int d = 0; d < 3 && e == hipSuccess; ++d) {
    const Coef cd = conv_coef(cf, d);
    e = c->prec ? zinv_t<float>(c, cd, -1, w1s[d]) : zinv_t<double>(c, cd, -1, w1s[d]);
  }
  prof_end(c, K_ZINV);
  if (e != hipSuccess) return e;
  prof_begin(c, K_YINV);
  for (int d = 0; d < 3 && e == hipSuccess; ++d)
    e = c->prec ? yinv_t<float>(c, w1s[d], w2s[d]) : yinv_t<double>(c, w1s[d], w2s[d]);
  prof_end(c, K_YINV);
  if (e != hipSuccess) return e;
  prof_begin(c, K_XPASS);
  e = c->prec ? xpass_conv_t<float>(c, cf, tr, cfl, rsp) : xpass_conv_t<double>(c, cf, tr, cfl, rsp);
  prof_end(c, K_XPASS);
  return e;
}

hipError_t launch_xpass(pdr_ctx* c, const Coef& cf, bool tr, bool cfl, int z0, int nzc, double* rsp) {
  prof_begin(c, K_XPASS);
  hipError_t e = c->spec ? launch2_xpass(c, cf, tr, cfl, z0, nzc, rsp)
                 : c->prec ? xpass_t<float>(c, cf, tr, cfl, rsp) : xpass_t<double>(c, cf, tr, cfl, rsp);
  prof_end(c, K_XPASS);
  return e;
}

template <typename T>
static hipError_t yfwd_t(pdr_ctx* c) {
  const Geo& g = c->g;
  const int ti = c->ti_y;
  const size_t lds = 2 * size_t(ti) * (g.ny + 1) * sizeof(cx<T>);
  hipError_t e = set_lds<T>((const void*)k_yfwd<T>, lds);
  if (e != hipSuccess) return e;
  dim3 grid((g.nki + ti - 1) / ti, g.nz);
  k_yfwd<T><<<grid, kThreads, lds, c->stream>>>(g, (const cx<T>*)c->w2, (cx<T>*)c->w1, c->py,
                                                (const cx<T>*)c->tab.twy, ti);
  return hipGetLastError();
}

hipError_t launch_yfwd(pdr_ctx* c, int z0, int nzc) {
  prof_begin(c, K_YFWD);
  hipError_t e = c->spec ? launch2_yfwd(c, z0, nzc) : c->prec ? yfwd_t<float>(c) : yfwd_t<double>(c);
  prof_end(c, K_YFWD);
  return e;
}

template <typename T>
static hipError_t zfwd_t(pdr_ctx* c) {
  const Geo& g = c->g;
  const int tj = c->tj_z;
  const size_t lds = 2 * size_t(tj) * (g.nz + 1) * sizeof(cx<T>);
  hipError_t e = set_lds<T>((const void*)k_zfwd<T>, lds);
  if (e != hipSuccess) return e;
  dim3 grid((g.nkj + tj - 1) / tj, g.nki);
  const double norm = 1.0 / (double(g.nx) * double(g.ny) * double(g.nz));
  k_zfwd<T><<<grid, kThreads, lds, c->stream>>>(g, (const cx<T>*)c->w1, (cx<T>*)c->nl, c->pz,
                                                (const cx<T>*)c->tab.twz, tj, T(norm));
  return hipGetLastError();
}

hipError_t launch_zfwd(pdr_ctx* c) {
  prof_begin(c, K_ZFWD);
  hipError_t e = c->spec ? launch2_zfwd(c) : c->prec ? zfwd_t<float>(c) : zfwd_t<double>(c);
  prof_end(c, K_ZFWD);
  return e;
}

template <typename T>
static hipError_t formf_t(pdr_ctx* c, const Coef& cf, double beta, double gdt) {
  const Geo& g = c->g;
  const long long nb = (g.mw + kThreads - 1) / kThreads;
  k_formf<T><<<dim3((unsigned)nb), kThreads, 0, c->stream>>>(
      g, cf, (cx<T>*)c->u[0], (cx<T>*)c->u[1], (cx<T>*)c->u[2], (cx<T>*)c->rk[0], (cx<T>*)c->rk[1],
      (cx<T>*)c->rk[2], (const cx<T>*)c->nl, (const double2*)c->force, c->forcing ? c->force_nk : 0,
      c->p.force_kf, beta, gdt);
  return hipGetLastError();
}

hipError_t launch_force(pdr_ctx* c, int32_t tstep, double dt) {
  const int nk = c->force_nk, w = 2 * nk + 1;
  prof_begin(c, K_MISC);
  if (tstep >= c->p.force_stop) {
    // IFDECAY branch: the forcing arrays are zeroed, no draws
    const int n = (nk + 1) * w * w * 3;
    k_zero_d2<<<dim3(unsigned((n + kThreads - 1) / kThreads)), kThreads, 0, c->stream>>>((double2*)c->force, n);
  } else {
    const double a = dt / c->p.force_tf;  // dt_over_TF
    const double sd = sqrt(2.0 * (c->p.force_sigma * c->p.force_sigma) * a);
    const int nt = (nk + 1) * (nk + 1) + nk * w * w;
    k_force_ou<<<dim3(unsigned((nt + kThreads - 1) / kThreads)), kThreads, 0, c->stream>>>(
        (double2*)c->force, nk, 1.0 - a, sd, (unsigned long long)c->p.force_seed,
        (unsigned long long)c->force_calls);
    ++c->force_calls;
  }
  prof_end(c, K_MISC);
  return hipGetLastError();
}

hipError_t launch_zfwd_formf(pdr_ctx* c, const Coef& cf, double beta, double gdt) {
  prof_begin(c, K_ZFF);
  hipError_t e = launch2_zfwd_formf(c, cf, beta, gdt);
  prof_end(c, K_ZFF);
  return e;
}

hipError_t launch_formf(pdr_ctx* c, const Coef& cf, double beta, double gdt) {
  prof_begin(c, K_FORMF);
  hipError_t e = c->prec ? formf_t<float>(c, cf, beta, gdt) : formf_t<double>(c, cf, beta, gdt);
  prof_end(c, K_FORMF);
  return e;
}

template <typename T>
static hipError_t tracers_t(pdr_ctx* c, const TracerFrame& tf, double beta, double gdt, hipStream_t s,
                            int wrap) {
  const T* ur = (const T*)c->ur;
  auto go = [&](auto kern, int order) {
    const long long nb = (tracer_group(order) * c->np + kThreads - 1) / kThreads;
    kern<<<dim3((unsigned)nb), kThreads, 0, s>>>(c->g, tf, ur, c->xp, c->uprk, c->up, c->xp0, c->np,
                                                         beta, gdt, wrap);
  };
  switch (tf.order) {
    case 1: go(k_tracers_col<T, 1>, 1); break;
    case 2: go(k_tracers_col<T, 2>, 2); break;
    case 3: go(k_tracers_col<T, 3>, 3); break;
    case 5: go(k_tracers_col<T, 5>, 5); break;
    case 6: go(k_tracers_col<T, 6>, 6); break;
    case 7: go(k_tracers_col<T, 7>, 7); break;
    case 8: go(k_tracers_col<T, 8>, 8); break;
    default: go(k_tracers_col<T, 4>, 4); break;
  }
  return hipGetLastError();
}

hipError_t launch_tracers(pdr_ctx* c, const TracerFrame& tf, double beta, double gdt, hipStream_t s,
                          bool wrap) {
  if (c->np <= 0) return hipSuccess;
  prof_begin(c, K_TRACER, s);
  hipError_t e = c->prec ? tracers_t<float>(c, tf, beta, gdt, s, wrap ? 1 : 0)
                         : tracers_t<double>(c, tf, beta, gdt, s, wrap ? 1 : 0);
  prof_end(c, K_TRACER, s);
  return e;
}

hipError_t launch_tracers_wrap(pdr_ctx* c, const TracerFrame& tf, hipStream_t s) {
  if (c->np <= 0) return hipSuccess;
  const long long nb = (c->np + kThreads - 1) / kThreads;
  k_tracers_wrap<<<dim3((unsigned)nb), kThreads, 0, s>>>(tf, c->xp, c->xp0, c->np);
  return hipGetLastError();
}

template <typename T>
static hipError_t remesh_t(pdr_ctx* c, int bfrot) {
  const Geo& g = c->g;
  const size_t lds = 3 * size_t(g.nkk) * sizeof(cx<T>);
  hipError_t e = set_lds<T>((const void*)k_remesh<T>, lds);
  if (e != hipSuccess) return e;
  k_remesh<T><<<dim3(g.nkj, g.nki), kThreads, lds, c->stream>>>(g, (cx<T>*)c->u[0], (cx<T>*)c->u[1],
                                                              (cx<T>*)c->u[2], bfrot);
  return hipGetLastError();
}

hipError_t launch_remesh(pdr_ctx* c, int bfrot) {
  prof_begin(c, K_REMESH);
  hipError_t e = c->prec ? remesh_t<float>(c, bfrot) : remesh_t<double>(c, bfrot);
  prof_end(c, K_REMESH);
  return e;
}

hipError_t launch_remesh_particles(pdr_ctx* c, double lx, double dx) {
  if (c->np <= 0) return hipSuccess;
  const long long nb = (c->np + kThreads - 1) / kThreads;
  k_remesh_particles<<<dim3((unsigned)nb), kThreads, 0, c->stream>>>(c->xp, c->xp0, c->np, lx, dx);
  return hipGetLastError();
}

hipError_t launch_dealias(pdr_ctx* c) {
  if (c->p.dealiasing == PDR_DEALIASING_TWOTHIRD) return hipSuccess;
  const Geo& g = c->g;
  MaskFrame mf;
  for (int m = 1; m <= 3; ++m)
    for (int n = 1; n <= 3; ++n) mf.bm[3 * (m - 1) + (n - 1)] = T9(c->bval, m, n) ? T9(c->bmat, m, n) : 0.0;
  // KMAX = min(KI_MAX, KJ_MAX, KK_MAX) (src/pg_domain.f90:231)
  const double kmax = double(std::min(g.ki, std::min(g.kj, g.kk)));
  mf.cutoff = c->p.dealiasing == PDR_DEALIASING_LES ? c->p.les_cutoff * kmax : kmax;
  const long long want = (g.mw + kThreads - 1) / kThreads;
  const unsigned nb = unsigned(want < 65536 ? want : 65536);
  if (c->prec)
    k_dealias<float><<<nb, kThreads, 0, c->stream>>>(g, mf, (cx<float>*)c->u[0], (cx<float>*)c->u[1],
                                                     (cx<float>*)c->u[2]);
  else
    k_dealias<double><<<nb, kThreads, 0, c->stream>>>(g, mf, (cx<double>*)c->u[0], (cx<double>*)c->u[1],
                                                      (cx<double>*)c->u[2]);
  return hipGetLastError();
}

hipError_t launch_finite(pdr_ctx* c, int* flag) {
  const long long n = 2 * c->g.mw;
  for (int q = 0; q < 3; ++q) {
    if (c->prec)
      k_finite<float><<<dim3(1024), kThreads, 0, c->stream>>>((const float*)c->u[q], n, flag);
    else
      k_finite<double><<<dim3(1024), kThreads, 0, c->stream>>>((const double*)c->u[q], n, flag);
  }
  return hipGetLastError();
}

}  // namespace pdr
