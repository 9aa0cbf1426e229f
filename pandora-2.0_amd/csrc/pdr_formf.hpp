// pdr_formf.hpp -- FormF, low-storage RK3 update and continuity fix of one
// wave mode, shared by the stand-alone FormF kernel and the fused z-forward +
// FormF kernel.  Generic over the isotropic and BF_ROT shear forms through
// the host-built operator tensors of Coef:
//   f_RK3FormFIsotropic / f_RK3FormFBFRot (src/pf_rk3.f90:2323-2673) with
//   f_RK3BFRot{Nonlinear,Viscous,Pressure,Source} (2864-3209), the isotropic
//   forcing (2421-2467) and that of the deformed isotropic form
//   (f_RK3BFRotForcing, 3256-3349), f_RK3RungeKutta (3457-3498) and
//   f_RK3ContinuityFixIsotropic (3504-3531).
#pragma once
#include "pdr_internal.hpp"

namespace pdr {

__device__ __forceinline__ double wvn_m(int k, int kmax) {
  return (k <= kmax) ? double(k) : double(k - (2 * kmax + 1));
}

// Mode (i, j, k) at element e of the wave arrays; a, b, c are its three
// components of the nonlinear term NL (already normalised and truncated).
// FRC: 0 leaves the forcing code out (callers without forcing arrays), so that
// it costs the fused kernels' unforced runs no registers; 1 the isotropic
// forcing; 2 that of the deformed isotropic form (cf.fdef).
template <typename T, int FRC = 2>
__device__ __forceinline__ void formf_mode(const Geo& g, const Coef& cf, int i, int j, int k, long long e, cx<T> a,
                                           cx<T> b, cx<T> c, cx<T>* __restrict__ u0, cx<T>* __restrict__ u1,
                                           cx<T>* __restrict__ u2, cx<T>* __restrict__ r0, cx<T>* __restrict__ r1,
                                           cx<T>* __restrict__ r2, const double2* __restrict__ frc, int fnk,
                                           double fkf, double beta, double gdt) {
  const double kv[3] = {wvn_m(i, g.ki), wvn_m(j, g.kj), wvn_m(k, g.kk)};
  const double kmag2 = (kv[0] * kv[0]) + (kv[1] * kv[1]) + (kv[2] * kv[2]);
  if (!(kmag2 > 0.0)) return;  // the k = 0 mode is never updated (src/pf_rk3.f90:2385)
  cx<T> u[3] = {u0[e], u1[e], u2[e]};
  cx<T> r[3] = {r0[e], r1[e], r2[e]};
  const bool rg = cf.form != 0;  // NS_R_ROT (f_RK3FormFRRot, src/pf_rk3.f90:2680-2856)
  cx<T> n[3];
  if (cf.use_nl && rg) {
    n[0] = a;
    n[1] = b;
    n[2] = c;
  } else if (cf.use_nl) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double* Bm = &cf.Bm[3 * q];
      n[q] = add(add(scl(a, T(Bm[0])), scl(b, T(Bm[1]))), scl(c, T(Bm[2])));
    }
  } else {
    n[0] = n[1] = n[2] = mk<T>(T(0), T(0));
  }
  // p1 = M k; p4 = k . M k (BF_ROT, M = bmat bmat^T) or |bmat^T k|^2 (R_ROT, M = bmat^T)
  double p1[3];
#pragma unroll
  for (int l = 0; l < 3; ++l) p1[l] = cf.M[3 * l] * kv[0] + cf.M[3 * l + 1] * kv[1] + cf.M[3 * l + 2] * kv[2];
  const double p4 = rg ? (p1[0] * p1[0] + p1[1] * p1[1]) + p1[2] * p1[2] : kv[0] * p1[0] + kv[1] * p1[1] + kv[2] * p1[2];
  // p2 = kp . nlf ; p3 = 2 sum_c (sum_m kp_m P[m][c]) u_c, along kp = k (BF_ROT)
  // or bmat^T k (R_ROT: f_RK3RRotPressure, 3078-3156)
  const double kp[3] = {rg ? p1[0] : kv[0], rg ? p1[1] : kv[1], rg ? p1[2] : kv[2]};
  cx<T> p23 = add(add(scl(n[0], T(kp[0])), scl(n[1], T(kp[1]))), scl(n[2], T(kp[2])));
  if (cf.has_hom) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double q = 2.0 * (kp[0] * cf.P[c] + kp[1] * cf.P[3 + c] + kp[2] * cf.P[6 + c]);
      p23 = add(p23, scl(u[c], T(q)));
    }
  }
  const T visc = cf.use_visc ? T(-cf.nu * p4) : T(0);
  cx<T> tot[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T pc = (p4 > 0.0) ? T(p1[c] / p4) : T(0);
    cx<T> t = sub(scl(u[c], visc), n[c]);
    t = add(t, scl(p23, pc));
    if (cf.has_hom) {
      const double* H = &cf.H[3 * c];
      const cx<T> h = add(add(scl(u[0], T(2.0 * H[0])), scl(u[1], T(2.0 * H[1]))), scl(u[2], T(2.0 * H[2])));
      t = sub(t, h);
    }
    tot[c] = t;
  }
  // forcing on 0 < |k| <= FORCE_KF, made solenoidal (src/pf_rk3.f90:2421-2467);
  // frc is [i][j+NK][k+NK][c] (isotropic forms only, fnk = 0 otherwise).  The
  // deformed isotropic form (f_RK3BFRotForcing, 3256-3349, and its R_ROT twin)
  // forces only modes whose indices the stretches fdiv divide, reads the arrays
  // at the quotient and projects along the laboratory wavenumber fbm[m] k_m; the
  // isotropic form is its unit-stretch case (fbm = 1: the same wavenumber)
  if (FRC != 0 && fnk > 0) {
    int il = i, jl = int(kv[1]), kl = int(kv[2]);
    bool hit = true;
    if (FRC == 2 && cf.fdef) {
      // quotients by rounding the product with the reciprocal (exact for the
      // divisible indices, which are the only ones kept): no integer division
      const int qi = int(rint(double(il) * cf.finv[0])), qj = int(rint(double(jl) * cf.finv[1])),
                qk = int(rint(double(kl) * cf.finv[2]));
      hit = qi * cf.fdiv[0] == il && qj * cf.fdiv[1] == jl && qk * cf.fdiv[2] == kl;
      il = qi;
      jl = qj;
      kl = qk;
    }
    const double wl[3] = {FRC == 2 ? cf.fbm[0] * kv[0] : kv[0], FRC == 2 ? cf.fbm[1] * kv[1] : kv[1],
                          FRC == 2 ? cf.fbm[2] * kv[2] : kv[2]};
    const double kl2 = (wl[0] * wl[0] + wl[1] * wl[1]) + wl[2] * wl[2];
    if (hit && il <= fnk && jl >= -fnk && jl <= fnk && kl >= -fnk && kl <= fnk && sqrt(kl2) <= fkf && kl2 > 0.0) {
      const int w = 2 * fnk + 1;
      const double2* f = frc + ((long long)(il * w + jl + fnk) * w + kl + fnk) * 3;
      const double kbr = (wl[0] * f[0].x + wl[1] * f[1].x) + wl[2] * f[2].x;
      const double kbi = (wl[0] * f[0].y + wl[1] * f[1].y) + wl[2] * f[2].y;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double q = wl[c] / kl2;
        tot[c] = add(tot[c], mk<T>(T(f[c].x - q * kbr), T(f[c].y - q * kbi)));
      }
    }
  }
  const T tb = T(beta), tg = T(gdt);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    r[c] = add(scl(r[c], tb), tot[c]);
    u[c] = add(u[c], scl(r[c], tg));
  }
  if (rg) {
    // f_RK3ContinuityFixRRot (src/pf_rk3.f90:3620-3693): um = bmat u,
    // um -= k (k.um) / |k|^2, u = binv um
    cx<T> um[3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
      um[m] = add(add(scl(u[0], T(cf.Bm[3 * m])), scl(u[1], T(cf.Bm[3 * m + 1]))), scl(u[2], T(cf.Bm[3 * m + 2])));
    const cx<T> kd = add(add(scl(um[0], T(kv[0])), scl(um[1], T(kv[1]))), scl(um[2], T(kv[2])));
#pragma unroll
    for (int m = 0; m < 3; ++m) um[m] = sub(um[m], scl(kd, T(kv[m] / kmag2)));
#pragma unroll
    for (int m = 0; m < 3; ++m)
      u[m] = add(add(scl(um[0], T(cf.V[3 * m])), scl(um[1], T(cf.V[3 * m + 1]))), scl(um[2], T(cf.V[3 * m + 2])));
  } else {
    // continuity fix u -= k (k.u) / |k|^2 (src/pf_rk3.f90:3504-3531)
    const cx<T> kd = add(add(scl(u[0], T(kv[0])), scl(u[1], T(kv[1]))), scl(u[2], T(kv[2])));
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c] = sub(u[c], scl(kd, T(kv[c] / kmag2)));
  }
  u0[e] = u[0];
  u1[e] = u[1];
  u2[e] = u[2];
  r0[e] = r[0];
  r1[e] = r[1];
  r2[e] = r[2];
}

}  // namespace pdr
