// Device evaluation of the infinite-depth wave kernel (the Green function's wave part and its
// derivatives), FP64.
//
//   G = 1/r + 1/r' + K Ft(X, Y) + i K Gi(X, Y),  X = K R,  Y = K |z + zeta|,
//   Gi = -2 pi exp(-Y) J0(X)                       (green_inf.rs:1-12, Newman 1985 eqs. 1-3)
//
// Inside [0, XT) x [0, YT) the identity form of the reference (green_inf.rs:340-349):
//   Ft = -2 e^-Y J0 (log(Y + R0) + gamma - log 2) - pi e^-Y Yt(X) - 2 R0 V(X, Y)
// with J0, J1, Yt, Yt' from 1D tiles and V, V_X from 2D tiles (tables built by green_tables.hip).
// Beyond: Hankel and Struve asymptotics and the partial-integration series of Havelock's integral
// (green_inf.rs:350-362; Abramowitz & Stegun 9.2.5-10 and 12.1.29; Newman 1985 eq. 7).
#pragma once

#include "hbg_internal.h"

namespace hbg {

struct WaveK {
  double F, FX, I, IX;  // Ft, dFt/dX, Gi, dGi/dX; Ft_Y = -Ft - 2/R0 and Gi_Y = -Gi
};

// A table pointer that loads through the constant address space: where the tile index is the
// same on every lane of the wavefront (fin_kernel_u below), such loads become scalar loads (one
// fetch serves the 64 lanes); elsewhere they are ordinary vector loads.
typedef __attribute__((address_space(4))) const double cdouble;

// Tile u of a table whose tiles are `stride` doubles apart, as a pointer that the compiler cannot
// tie to any lane's own index (the empty asm pins it to scalar registers; without it, inside
// `if (key == u)` the compiler substitutes the lane's key and the loads become vector loads).
__device__ __forceinline__ cdouble* uniform_tile(const double* base, int u, int stride) {
  cdouble* c = (cdouble*)(base + (size_t)u * stride);
  asm volatile("" : "+s"(c));
  return c;
}

// Powers t^0 .. t^(M - 1) of a tile variable, for the power-basis evaluations below
template <int M>
__device__ __forceinline__ void powers(double t, double* pw) {
  pw[0] = 1.0;
  pw[1] = t;
#pragma unroll
  for (int a = 2; a < M; ++a) pw[a] = pw[a - 1] * t;
}

// sum_a c[a stride] t^a with the powers given: with c in scalar registers each term is one
// fused multiply-add whose scalar operand is the coefficient (Horner's fma(p, t, c) would need
// the coefficient copied into vector registers first)
template <int M, int STRIDE>
__device__ __forceinline__ double poly_u(cdouble* c, const double* pw) {
  double r = c[0];
#pragma unroll
  for (int a = 1; a < M; ++a) r = fma(c[a * STRIDE], pw[a], r);
  return r;
}

// Runs BODY once per distinct value of KEY among the lanes where ACTIVE holds: U is the
// round's key, the same on every lane, and MINE tells whether this lane's key is U. With the
// lanes' keys mostly equal the loop runs once or twice, and the tile loads inside are scalar.
#define HBG_UNIFORM_ROUNDS(KEY, ACTIVE, U, MINE, BODY)                                          \
  {                                                                                             \
    unsigned long long todo_ = __ballot(ACTIVE);                                                \
    while (todo_) {                                                                             \
      const int U = __builtin_amdgcn_readlane((KEY), (int)__builtin_ctzll(todo_));             \
      const bool MINE = (ACTIVE) && (KEY) == U;                                                 \
      BODY                                                                                      \
      todo_ &= ~__ballot(MINE);                                                                 \
    }                                                                                           \
  }

// The 1D tile of x in [0, XT): index and tile variable t = 2 (x - i) - 1
__device__ __forceinline__ int tile1d(double X, double& t) {
  int i = (int)X;
  if (i >= XT) i = XT - 1;
  t = 2.0 * (X - (double)i) - 1.0;
  return i;
}

template <class P>
__device__ __forceinline__ void table1d_eval(P c, double t, double& j0, double& j1, double& yt, double& dyt) {
  double r0 = c[4 * D1], r1 = c[4 * D1 + 1], r2 = c[4 * D1 + 2], r3 = c[4 * D1 + 3];
#pragma unroll
  for (int j = D1 - 1; j >= 0; --j) {
    r0 = fma(r0, t, c[4 * j]);
    r1 = fma(r1, t, c[4 * j + 1]);
    r2 = fma(r2, t, c[4 * j + 2]);
    r3 = fma(r3, t, c[4 * j + 3]);
  }
  j0 = r0;
  j1 = r1;
  yt = r2;
  dyt = r3;
}

// table1d_eval and table2d_eval for a wavefront-uniform tile, in the power basis
__device__ __forceinline__ void table1d_eval_u(cdouble* c, double t, double& j0, double& j1, double& yt, double& dyt) {
  double pw[M1];
  powers<M1>(t, pw);
  j0 = poly_u<M1, 4>(c, pw);
  j1 = poly_u<M1, 4>(c + 1, pw);
  yt = poly_u<M1, 4>(c + 2, pw);
  dyt = poly_u<M1, 4>(c + 3, pw);
}

__device__ __forceinline__ void table1d(const double* __restrict__ t1d, double X, double& j0, double& j1,
                                        double& yt, double& dyt) {
  double t;
  const int i = tile1d(X, t);
  table1d_eval(t1d + (size_t)i * TILE1, t, j0, j1, yt, dyt);
}

// The 2D tile of (X, Y) in [0, XT) x [0, YT): index iy XT + ix and the tile variables
__device__ __forceinline__ int tile2d(double X, double Y, double& tx, double& ty) {
  int ix = (int)X, iy = (int)Y;
  if (ix >= XT) ix = XT - 1;
  if (iy >= YT) iy = YT - 1;
  tx = 2.0 * (X - (double)ix) - 1.0;
  ty = 2.0 * (Y - (double)iy) - 1.0;
  return iy * XT + ix;
}

template <class P>
__device__ __forceinline__ void table2d_eval(P c, double tx, double ty, double& V, double& VX) {
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int b = D2; b >= 0; --b) {
    const int o0 = b * M2, o1 = M2 * M2 + b * M2;
    double r0 = c[o0 + D2], r1 = c[o1 + D2];
#pragma unroll
    for (int a = D2 - 1; a >= 0; --a) {
      r0 = fma(r0, tx, c[o0 + a]);
      r1 = fma(r1, tx, c[o1 + a]);
    }
    s0 = fma(s0, ty, r0);
    s1 = fma(s1, ty, r1);
  }
  V = s0;
  VX = s1;
}

__device__ __forceinline__ void table2d_eval_u(cdouble* c, double tx, double ty, double& V, double& VX) {
  double pw[M2];
  powers<M2>(tx, pw);
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int b = D2; b >= 0; --b) {
    s0 = fma(s0, ty, poly_u<M2, 1>(c + b * M2, pw));
    s1 = fma(s1, ty, poly_u<M2, 1>(c + M2 * M2 + b * M2, pw));
  }
  V = s0;
  VX = s1;
}

__device__ __forceinline__ void table2d(const double* __restrict__ t2d, double X, double Y, double& V, double& VX) {
  double tx, ty;
  const int i = tile2d(X, Y, tx, ty);
  table2d_eval(t2d + (size_t)i * TILE2, tx, ty, V, VX);
}

// Hankel asymptotic expansions of J0, Y0, J1, Y1 for x >= XT (A&S 9.2.5-10):
// P(nu) ~ sum (-1)^k a_2k / x^2k, Q(nu) ~ sum (-1)^k a_2k+1 / x^2k+1,
// a_k(nu) = prod_{m=1..k} (4 nu^2 - (2m - 1)^2) / (k! 8^k).
__device__ __forceinline__ void hankel01(double x, double& j0, double& y0, double& j1, double& y1) {
  const double inv = 1.0 / x;
  double P[2], Q[2];
#pragma unroll
  for (int nu = 0; nu < 2; ++nu) {
    const double mu = 4.0 * nu * nu;
    double a = 1.0, pw = 1.0, p = 0.0, q = 0.0;
#pragma unroll
    for (int k = 0; k < 20; ++k) {
      // term a_k / x^k with a_k by its recurrence
      const double term = a * pw;
      const int r = k & 3;  // k = 0: +P, 1: +Q, 2: -P, 3: -Q
      if (r == 0) p += term;
      if (r == 1) q += term;
      if (r == 2) p -= term;
      if (r == 3) q -= term;
      const double m = 2.0 * k + 1.0;
      a *= (mu - m * m) / (8.0 * (k + 1));
      pw *= inv;
    }
    P[nu] = p;
    Q[nu] = q;
  }
  const double chi = x - 0.25 * PI;
  const double sq = sqrt(2.0 / (PI * x));
  double sn, cs;
  sincos(chi, &sn, &cs);
  j0 = sq * (P[0] * cs - Q[0] * sn);
  y0 = sq * (P[0] * sn + Q[0] * cs);
  j1 = sq * (P[1] * sn + Q[1] * cs);
  y1 = sq * (-P[1] * cs + Q[1] * sn);
}

// H0 - Y0 and H1 - Y1 for x >= XT (A&S 12.1.29 with nu = 0, 1):
// H0 - Y0 ~ (2/pi) sum_k (-1)^k ((2k-1)!!)^2 / x^(2k+1),
// H1 - Y1 ~ (2/pi) [1 + sum_{k>=1} (-1)^(k-1) (2k-1)!! (2k-3)!! / x^(2k)].
__device__ __forceinline__ void struve_minus_y01(double x, double& s0, double& s1) {
  const double inv2 = 1.0 / (x * x);
  double c0 = 1.0, c1 = 1.0, pw = 1.0 / x, pw1 = inv2;
  double a0 = pw, a1 = 1.0;
#pragma unroll
  for (int k = 1; k < 14; ++k) {
    const double m = 2.0 * k - 1.0;
    c0 *= -m * m;  // (-1)^k ((2k-1)!!)^2
    pw *= inv2;
    a0 += c0 * pw;
    if (k >= 2) c1 *= -(m * (m - 2.0));  // (-1)^(k-1) (2k-1)!! (2k-3)!!
    a1 += c1 * pw1;
    pw1 *= inv2;
  }
  s0 = (2.0 / PI) * a0;
  s1 = (2.0 / PI) * a1;
}

// sum_m (-1)^m m! a_m(t0) and its X derivative, a_m the Taylor coefficients of
// f(t) = (X^2 + t^2)^(-1/2) at t0 (recurrence from (X^2 + t^2) f' + t f = 0), stopped when the
// bound m! / A^((m+1)/2) of the next term falls below 1e-16.
__device__ __forceinline__ void taylor_sum(double X, double t0, double& P, double& PX) {
  const double A = X * X + t0 * t0;
  const double ia = 1.0 / A;
  const double ir = sqrt(ia);
  double a = ir, b = -X * ir * ia, am1 = 0.0, bm1 = 0.0;
  double fact = 1.0, size = ir;
  P = a;
  PX = b;
  for (int n = 0; n < 24; ++n) {
    const double an = a, bn = b, nf = (double)n, k = 2.0 * n + 1.0;
    const double inv = 1.0 / (nf + 1.0);
    a = -(k * t0 * an + nf * am1) * ia * inv;
    b = -(k * t0 * bn + nf * bm1) * ia * inv - a * 2.0 * X * ia;
    am1 = an;
    bm1 = bn;
    fact *= nf + 1.0;
    size *= (nf + 1.0) * ir;
    const double s = (n & 1) ? fact : -fact;
    P += s * a;
    PX += s * b;
    if (size < 1e-16) break;
  }
}

// The wave kernel from the table values it needs: (j0, j1, yt, dyt) of the 1D tile of X when
// X < XT, and (V, VX) of the 2D tile of (X, Y) when also Y < YT (the others are not read).
__device__ __forceinline__ WaveK wave_kernel_tab(double X, double Y, double j0, double j1, double yt, double dyt,
                                                 double V, double VX) {
  WaveK w;
  const double R0 = sqrt(X * X + Y * Y);
  const double eY = exp(-Y);
  if (X < (double)XT && Y < (double)YT) {
    const double lg = log(Y + R0) + (GAMMA - LOG2);
    w.F = -2.0 * eY * j0 * lg - PI * eY * yt - 2.0 * R0 * V;
    w.FX = 2.0 * eY * j1 * lg - 2.0 * eY * j0 * X / (R0 * (Y + R0)) - PI * eY * dyt - 2.0 * (X / R0) * V -
           2.0 * R0 * VX;
  } else if (Y >= (double)YT) {
    // the exp(-Y) terms are below 1e-20 and dropped
    if (X >= (double)XT) {
      double y0, y1;
      hankel01(X, j0, y0, j1, y1);
    }
    double P, PX;
    taylor_sum(X, Y, P, PX);
    w.F = -2.0 * P;
    w.FX = -2.0 * PX;
  } else {
    double y0, y1, s0, s1;
    hankel01(X, j0, y0, j1, y1);
    struve_minus_y01(X, s0, s1);
    const double h0py0 = s0 + 2.0 * y0;                 // H0 + Y0
    const double dh0py0 = 2.0 / PI - s1 - 2.0 * y1;     // (H0 + Y0)' = 2/pi - H1 - Y1
    double P, PX, P0, PX0;
    taylor_sum(X, Y, P, PX);
    taylor_sum(X, 0.0, P0, PX0);
    P -= eY * P0;
    PX -= eY * PX0;
    w.F = -PI * eY * h0py0 - 2.0 * P;
    w.FX = -PI * eY * dh0py0 - 2.0 * PX;
  }
  w.I = -2.0 * PI * eY * j0;
  w.IX = 2.0 * PI * eY * j1;
  return w;
}

// wave_kernel for a wavefront whose lanes take different points (`valid` false: no point, the
// lane only joins the rounds), the table reads wavefront-uniform (HBG_UNIFORM_ROUNDS, scalar
// loads, power basis).
__device__ __forceinline__ WaveK wave_kernel_u(const WaveTables& T, bool valid, double X, double Y) {
  double j0 = 0.0, j1 = 0.0, yt = 0.0, dyt = 0.0, V = 0.0, VX = 0.0;
  {
    double t;
    const int i1 = tile1d(X, t);
    const bool act = valid && X < (double)XT;
    HBG_UNIFORM_ROUNDS(i1, act, u, mine, {
      cdouble* c = uniform_tile(T.t1d, u, TILE1);
      if (mine) table1d_eval_u(c, t, j0, j1, yt, dyt);
    })
  }
  {
    double tx, ty;
    const int i2 = tile2d(X, Y, tx, ty);
    const bool act = valid && X < (double)XT && Y < (double)YT;
    HBG_UNIFORM_ROUNDS(i2, act, u, mine, {
      cdouble* c = uniform_tile(T.t2d, u, TILE2);
      if (mine) table2d_eval_u(c, tx, ty, V, VX);
    })
  }
  return wave_kernel_tab(X, Y, j0, j1, yt, dyt, V, VX);
}

__device__ __forceinline__ WaveK wave_kernel(const WaveTables& T, double X, double Y) {
  double j0 = 0.0, j1 = 0.0, yt = 0.0, dyt = 0.0, V = 0.0, VX = 0.0;
  if (X < (double)XT) {
    table1d(T.t1d, X, j0, j1, yt, dyt);
    if (Y < (double)YT) table2d(T.t2d, X, Y, V, VX);
  }
  return wave_kernel_tab(X, Y, j0, j1, yt, dyt, V, VX);
}

}  // namespace hbg
