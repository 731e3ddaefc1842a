// Device evaluation of the finite-depth wave kernel, FP64 (green_fin.rs:1-20, 1007-1095).
//
// Depth h, K = w^2/g, H = K h, field point depth z, source depth zeta, horizontal distance R,
// v1 = |z - zeta|, v2 = z + zeta + 2h, v3 = -(z + zeta), v4 = z - zeta + 2h, v5 = -z + zeta + 2h,
// v6 = z + zeta + 4h, r_p = sqrt(R^2 + v_p^2). The regular part Wr = Re G - sum_p 1/r_p and
// Wi = Im G with their derivatives in R, zeta (source) and z (field point):
//   R >= h/2: John's eigenfunction series,
//     Re G = -A0 E Y0(k0 R) + sum_n c_n cos k_n(z+h) cos k_n(zeta+h) K0(k_n R), stopped at k_n R >= 18;
//   R < h/2: Newman's decomposition,
//     Wr = K Ft(X, V3) + (1/h) [Phi1(R/h, v1/h) + Phi2(R/h, v2/h)],
//     Phi1 = H [Ft(X, 2H - V1) + Ft(X, 2H + V1)] + Q(xb, vb1),  Phi2 = H Ft(X, 2H + V2) + Q(xb, vb2),
//     Q(xb, vb) = PV int_0^inf (k + H)^2 cosh(k vb) e^{-3k} J0(k xb) / [(k sinh k - H cosh k)(k - H)] dk,
//   both: Im G = -A0 E J0(k0 R), E = sum_{p=3..6} e^{-k0 v_p},
//         A0 = 2 pi k0 / (1 - e^{-4 k0 h} + 4 k0 h e^{-2 k0 h}), c_n = 4 (k_n^2 + K^2)/(k_n^2 h + K^2 h - K).
// Phi1 and Phi2 come from per-frequency tiles (8 x 8 over [0, 1/2] x [0, 1] and 8 x 16 over
// [0, 1/2] x [0, 2], degree FD, value and both derivatives fitted separately) that
// green_fin.hip builds on the device from the contour integral of Q; K0 and K1 from unit tiles
// (degree 17) of [1/2, 18).
#pragma once

#include "green_inf_dev.h"
#include "hbg_internal.h"

namespace hbg {

constexpr int FIN_ROOTS = 24;       // evanescent roots kept (green_fin.rs:113)
constexpr double SERIES_STOP = 18.0;  // stop of the series at k_n R >= 18 (green_fin.rs:349)
constexpr int FTS = 8, FTW = 8, FTV = 16;  // tiles of Phi1 (FTS x FTW) and Phi2 (FTS x FTV)
constexpr int FD = 10;                  // degree of the Phi tiles
constexpr int FM = FD + 1;
constexpr int FTILE = 3 * FM * FM;      // doubles per tile: [value, d/dxb, d/dvb][vb power][xb power]
constexpr int KD = 17;                  // degree of the K0, K1 tiles
constexpr int KTILES = 18;              // [1/2, 1), [1, 2), ..., [17, 18)

// Per-frequency constants of the finite-depth kernel.
struct FinFreq {
  double h, K, H, x0, k0, A0;
  double kn[FIN_ROOTS], cn[FIN_ROOTS];
};

// Per-frequency device data: Phi tiles and, per panel, the propagating-mode factors
// (e^{k0 z}, e^{-k0 (z + h)}, e^{-k0 (h - z)}, e^{-k0 (z + 2h)}) and cos, sin k_n (z + h).
struct FinTables {
  const double* phi1;   // [FTS][FTW][FTILE]
  const double* phi2;   // [FTS][FTV][FTILE]
  const double* ktab;   // [KTILES][KD + 1][2] monomial coefficients of K0, K1
  const double4* pf;    // [N] propagating factors
  const double2* pcs;   // [N][FIN_ROOTS] (cos, sin) k_n (z + h)
};

struct FinK {
  double Wr, WrR, WrZ, Wrz, Wi, WiR, WiZ, Wiz;
};

// The tile of x in [1/2, 18) of the K0, K1 table and its tile variable
__device__ __forceinline__ int ktile(double x, double& t) {
  int i;
  if (x < 1.0) {
    i = 0;
    t = 4.0 * (x - 0.5) - 1.0;
  } else {
    i = (int)x;
    if (i > KTILES - 1) i = KTILES - 1;
    t = 2.0 * (x - (double)i) - 1.0;
  }
  return i;
}

constexpr int KTILE = (KD + 1) * 2;  // doubles per K tile

template <class P>
__device__ __forceinline__ void ktile_eval(P c, double t, double& g0, double& g1) {
  double a0 = c[2 * KD], a1 = c[2 * KD + 1];
#pragma unroll
  for (int j = KD - 1; j >= 0; --j) {
    a0 = fma(a0, t, c[2 * j]);
    a1 = fma(a1, t, c[2 * j + 1]);
  }
  g0 = a0;
  g1 = a1;
}

__device__ __forceinline__ void ktile_eval_u(cdouble* c, double t, double& g0, double& g1) {
  double pw[KD + 1];
  powers<KD + 1>(t, pw);
  g0 = poly_u<KD + 1, 2>(c, pw);
  g1 = poly_u<KD + 1, 2>(c + 1, pw);
}

__device__ __forceinline__ void k01(const double* __restrict__ kt, double x, double& g0, double& g1) {
  double t;
  const int i = ktile(x, t);
  ktile_eval(kt + (size_t)i * KTILE, t, g0, g1);
}

// The Phi tile of (xb, vb) (ny tiles over vb in [0, ymax]): index ix ny + iy and tile variables
__device__ __forceinline__ int phi_index(int ny, double ymax, double xb, double vb, double& tx, double& ty) {
  const double ux = xb * (FTS / 0.5), uy = vb * (ny / ymax);
  int ix = (int)ux, iy = (int)uy;
  if (ix > FTS - 1) ix = FTS - 1;
  if (iy > ny - 1) iy = ny - 1;
  if (ix < 0) ix = 0;
  if (iy < 0) iy = 0;
  tx = 2.0 * (ux - (double)ix) - 1.0;
  ty = 2.0 * (uy - (double)iy) - 1.0;
  return ix * ny + iy;
}

// (Phi, dPhi/dxb, dPhi/dvb) from the tile's coefficients
template <class P>
__device__ __forceinline__ void phi_eval(P c, double tx, double ty, double* out) {
#pragma unroll 1
  for (int q = 0; q < 3; ++q) {
    const int oq = q * FM * FM;
    double s = 0.0;
#pragma unroll 1
    for (int b = FD; b >= 0; --b) {
      const int ob = oq + b * FM;
      double p = c[ob + FD];
#pragma unroll
      for (int a = FD - 1; a >= 0; --a) p = fma(p, tx, c[ob + a]);
      s = fma(s, ty, p);
    }
    if (q == 0) out[0] = s;
    if (q == 1) out[1] = s;
    if (q == 2) out[2] = s;
  }
}

// phi_eval for a wavefront-uniform tile, rows in the power basis of xb
__device__ __forceinline__ void phi_eval_u(cdouble* c, double tx, double ty, double* out) {
  double pw[FM];
  powers<FM>(tx, pw);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double s = 0.0;
#pragma unroll
    for (int b = FD; b >= 0; --b) s = fma(s, ty, poly_u<FM, 1>(c + q * FM * FM + b * FM, pw));
    out[q] = s;
  }
}

// (Phi, dPhi/dxb, dPhi/dvb) from the tile of (xb, vb): which 1 over vb in [0, 1], 2 over [0, 2]
__device__ __forceinline__ void phi_tile(const double* __restrict__ tab, int ny, double ymax, double xb, double vb,
                                         double* out) {
  double tx, ty;
  const int i = phi_index(ny, ymax, xb, vb, tx, ty);
  phi_eval(tab + (size_t)i * FTILE, tx, ty, out);
}

// Y0, Y1 at x < XT from the 1D tile's (j0, j1, Yt, Yt')
__device__ __forceinline__ void bessel_y01(double x, double j0, double j1, double yt, double dyt, double& y0,
                                           double& y1) {
  const double L = log(0.5 * x) + GAMMA;
  y0 = yt + 2.0 / PI * L * j0;
  y1 = -dyt - 2.0 / PI * (j0 / x - L * j1);
}

// The table-based Bessel functions of the infinite-depth module: (J0, Y0, J1, Y1) at x > 0.
__device__ __forceinline__ void bessel_jy01(const WaveTables& T, double x, double& j0, double& y0, double& j1,
                                            double& y1) {
  if (x >= (double)XT) {
    hankel01(x, j0, y0, j1, y1);
    return;
  }
  double yt, dyt;
  table1d(T.t1d, x, j0, j1, yt, dyt);
  bessel_y01(x, j0, j1, yt, dyt, y0, y1);
}

// The finite-depth kernel of the pair (field point k at depth z, source panel n at depth zeta).
__device__ __forceinline__ FinK fin_kernel(const FinFreq* __restrict__ fp, const FinTables& F, const WaveTables& T,
                                           int k, int n, double R, double z, double zeta) {
  FinK o;
  const FinFreq& f = *fp;
  const double h = f.h, K = f.K, k0 = f.k0;
  const double d1 = z - zeta;
  const double sg = d1 > 0.0 ? 1.0 : (d1 < 0.0 ? -1.0 : 0.0);
  const double v1 = fabs(d1), v2 = z + zeta + 2.0 * h, v3 = -(z + zeta);
  const double4 pz = F.pf[k], pq = F.pf[n];
  const double e3 = pz.x * pq.x, e4 = pz.y * pq.z, e5 = pz.z * pq.y, e6 = pz.w * pq.w;
  const double E = e3 + e4 + e5 + e6, Ez = k0 * (e3 + e4 - e5 - e6), Ef = k0 * (e3 - e4 + e5 - e6);
  const double x = k0 * R;
  double j0, j1, y0, y1;
  if (R >= 0.5 * h) {
    bessel_jy01(T, x, j0, y0, j1, y1);
    double re = -f.A0 * E * y0, reR = f.A0 * E * k0 * y1, reZ = -f.A0 * Ez * y0, rez = -f.A0 * Ef * y0;
    const double2* cz = F.pcs + (size_t)k * FIN_ROOTS;
    const double2* cq = F.pcs + (size_t)n * FIN_ROOTS;
    for (int m = 0; m < FIN_ROOTS; ++m) {
      const double kk = f.kn[m], xm = kk * R;
      if (xm >= SERIES_STOP) break;
      double K0, K1;
      k01(F.ktab, xm, K0, K1);
      const double2 a = cz[m], b = cq[m];
      const double c = f.cn[m];
      re += c * a.x * b.x * K0;
      reR -= c * a.x * b.x * kk * K1;
      reZ -= c * a.x * kk * b.y * K0;
      rez -= c * kk * a.y * b.x * K0;
    }
    const double v[6] = {v1, v2, v3, d1 + 2.0 * h, -d1 + 2.0 * h, z + zeta + 4.0 * h};
    const double dvz[6] = {-sg, 1.0, -1.0, -1.0, 1.0, 1.0}, dvf[6] = {sg, 1.0, -1.0, 1.0, -1.0, 1.0};
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const double r2 = R * R + v[p] * v[p], ir = rsqrt(r2), ir3 = ir * ir * ir;
      re -= ir;
      reR += R * ir3;
      reZ += v[p] * dvz[p] * ir3;
      rez += v[p] * dvf[p] * ir3;
    }
    o.Wr = re;
    o.WrR = reR;
    o.WrZ = reZ;
    o.Wrz = rez;
  } else {
    if (x >= (double)XT) {
      hankel01(x, j0, y0, j1, y1);
    } else {
      double yt, dyt;
      table1d(T.t1d, x, j0, j1, yt, dyt);
    }
    const double ih = 1.0 / h, ih2 = ih * ih;
    double f1[3], f2[3];
    phi_tile(F.phi1, FTW, 1.0, R * ih, v1 * ih, f1);
    phi_tile(F.phi2, FTV, 2.0, R * ih, v2 * ih, f2);
    WaveK w = wave_kernel(T, K * R, K * v3);
    const double X = K * R, Y = K * v3;
    double FY = -w.F - 2.0 / sqrt(X * X + Y * Y);
    if (R == 0.0 && v3 == 0.0) {  // coincident points on z = 0: the limit of Ft + 2 psi (green_fin.rs:816-822)
      w.F = -2.0 * (GAMMA - LOG2 + log(K));
      w.FX = 0.0;
      FY = 0.0;
    }
    o.Wr = ih * (f1[0] + f2[0]) + K * w.F;
    o.WrR = ih2 * (f1[1] + f2[1]) + K * K * w.FX;
    const double A1z = -sg * ih2 * f1[2];
    const double Bz = ih2 * f2[2] - K * K * FY;
    o.WrZ = A1z + Bz;
    o.Wrz = -A1z + Bz;
  }
  o.Wi = -f.A0 * E * j0;
  o.WiR = f.A0 * E * k0 * j1;
  o.WiZ = -f.A0 * Ez * j0;
  o.Wiz = -f.A0 * Ef * j0;
  return o;
}

// fin_kernel for a wavefront whose lanes take different pairs (`valid` false: no pair, the lane
// only joins the wavefront-wide rounds), with every table read made wavefront-uniform: each
// tile lookup runs once per distinct tile among the lanes (HBG_UNIFORM_ROUNDS), its
// coefficients fetched by scalar loads and summed in the power basis of the tile variable (the
// per-lane fin_kernel uses Horner's rule: the two differ by rounding only). It pays when the
// lanes' pairs share tiles, as after the assembly's sort of the pairs by tile.
__device__ __forceinline__ FinK fin_kernel_u(const FinFreq* __restrict__ fp, const FinTables& F, const WaveTables& T,
                                             bool valid, int k, int n, double R, double z, double zeta) {
  FinK o;
  const FinFreq& f = *fp;
  const double h = f.h, K = f.K, k0 = f.k0;
  const double d1 = z - zeta;
  const double sg = d1 > 0.0 ? 1.0 : (d1 < 0.0 ? -1.0 : 0.0);
  const double v1 = fabs(d1), v2 = z + zeta + 2.0 * h, v3 = -(z + zeta);
  const double4 pz = F.pf[k], pq = F.pf[n];
  const double e3 = pz.x * pq.x, e4 = pz.y * pq.z, e5 = pz.z * pq.y, e6 = pz.w * pq.w;
  const double E = e3 + e4 + e5 + e6, Ez = k0 * (e3 + e4 - e5 - e6), Ef = k0 * (e3 - e4 + e5 - e6);
  const double x = k0 * R;
  const bool series = valid && R >= 0.5 * h, near = valid && !(R >= 0.5 * h);
  // the 1D tile at x = k0 R: J0, J1 for both forms, Y0, Y1 for the series
  double j0 = 0.0, j1 = 0.0, yt = 0.0, dyt = 0.0, y0 = 0.0, y1 = 0.0;
  {
    double t;
    const int i1 = tile1d(x, t);
    const bool act = valid && x < (double)XT;
    HBG_UNIFORM_ROUNDS(i1, act, u, mine, {
      cdouble* c = uniform_tile(T.t1d, u, TILE1);
      if (mine) table1d_eval_u(c, t, j0, j1, yt, dyt);
    })
  }
  if (valid && x >= (double)XT) hankel01(x, j0, y0, j1, y1);
  double Wr = 0.0, WrR = 0.0, WrZ = 0.0, Wrz = 0.0;
  if (__ballot(series)) {
    if (series && x < (double)XT) bessel_y01(x, j0, j1, yt, dyt, y0, y1);
    double re = -f.A0 * E * y0, reR = f.A0 * E * k0 * y1, reZ = -f.A0 * Ez * y0, rez = -f.A0 * Ef * y0;
    const double2* cz = F.pcs + (size_t)k * FIN_ROOTS;
    const double2* cq = F.pcs + (size_t)n * FIN_ROOTS;
    // the roots k_n increase, so a lane leaves the series at its first k_n R >= 18 as in fin_kernel
    bool on = series;
    for (int m = 0; m < FIN_ROOTS; ++m) {
      const double kk = f.kn[m], xm = kk * R;
      on = on && xm < SERIES_STOP;
      if (!__ballot(on)) break;
      double t, K0 = 0.0, K1 = 0.0;
      const int it = ktile(xm, t);
      HBG_UNIFORM_ROUNDS(it, on, u, mine, {
        cdouble* c = uniform_tile(F.ktab, u, KTILE);
        if (mine) ktile_eval_u(c, t, K0, K1);
      })
      if (on) {
        const double2 a = cz[m], b = cq[m];
        const double c = f.cn[m];
        re += c * a.x * b.x * K0;
        reR -= c * a.x * b.x * kk * K1;
        reZ -= c * a.x * kk * b.y * K0;
        rez -= c * kk * a.y * b.x * K0;
      }
    }
    if (series) {
      const double v[6] = {v1, v2, v3, d1 + 2.0 * h, -d1 + 2.0 * h, z + zeta + 4.0 * h};
      const double dvz[6] = {-sg, 1.0, -1.0, -1.0, 1.0, 1.0}, dvf[6] = {sg, 1.0, -1.0, 1.0, -1.0, 1.0};
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        const double r2 = R * R + v[p] * v[p], ir = rsqrt(r2), ir3 = ir * ir * ir;
        re -= ir;
        reR += R * ir3;
        reZ += v[p] * dvz[p] * ir3;
        rez += v[p] * dvf[p] * ir3;
      }
      Wr = re;
      WrR = reR;
      WrZ = reZ;
      Wrz = rez;
    }
  }
  if (__ballot(near)) {
    const double ih = 1.0 / h, ih2 = ih * ih;
    double f1[3] = {0.0, 0.0, 0.0}, f2[3] = {0.0, 0.0, 0.0};
    {
      double tx, ty;
      const int i = phi_index(FTW, 1.0, R * ih, v1 * ih, tx, ty);
      HBG_UNIFORM_ROUNDS(i, near, u, mine, {
        cdouble* c = uniform_tile(F.phi1, u, FTILE);
        if (mine) phi_eval_u(c, tx, ty, f1);
      })
    }
    {
      double tx, ty;
      const int i = phi_index(FTV, 2.0, R * ih, v2 * ih, tx, ty);
      HBG_UNIFORM_ROUNDS(i, near, u, mine, {
        cdouble* c = uniform_tile(F.phi2, u, FTILE);
        if (mine) phi_eval_u(c, tx, ty, f2);
      })
    }
    const double X = K * R, Y = K * v3;
    double wj0 = 0.0, wj1 = 0.0, wyt = 0.0, wdyt = 0.0, V = 0.0, VX = 0.0;
    {
      double t;
      const int i1 = tile1d(X, t);
      const bool act = near && X < (double)XT;
      HBG_UNIFORM_ROUNDS(i1, act, u, mine, {
        cdouble* c = uniform_tile(T.t1d, u, TILE1);
        if (mine) table1d_eval_u(c, t, wj0, wj1, wyt, wdyt);
      })
    }
    {
      double tx, ty;
      const int i2 = tile2d(X, Y, tx, ty);
      const bool act = near && X < (double)XT && Y < (double)YT;
      HBG_UNIFORM_ROUNDS(i2, act, u, mine, {
        cdouble* c = uniform_tile(T.t2d, u, TILE2);
        if (mine) table2d_eval_u(c, tx, ty, V, VX);
      })
    }
    if (near) {
      WaveK w = wave_kernel_tab(X, Y, wj0, wj1, wyt, wdyt, V, VX);
      double FY = -w.F - 2.0 / sqrt(X * X + Y * Y);
      if (R == 0.0 && v3 == 0.0) {
        // coincident points on z = 0 (an interior panel and its own centroid): the limit of
        // Ft + 2 psi, psi = log((Y + R0) / K), and no derivatives (green_fin.rs:816-822)
        w.F = -2.0 * (GAMMA - LOG2 + log(K));
        w.FX = 0.0;
        FY = 0.0;
      }
      Wr = ih * (f1[0] + f2[0]) + K * w.F;
      WrR = ih2 * (f1[1] + f2[1]) + K * K * w.FX;
      const double A1z = -sg * ih2 * f1[2];
      const double Bz = ih2 * f2[2] - K * K * FY;
      WrZ = A1z + Bz;
      Wrz = -A1z + Bz;
    }
  }
  o.Wr = Wr;
  o.WrR = WrR;
  o.WrZ = WrZ;
  o.Wrz = Wrz;
  o.Wi = -f.A0 * E * j0;
  o.WiR = f.A0 * E * k0 * j1;
  o.WiZ = -f.A0 * Ez * j0;
  o.Wiz = -f.A0 * Ef * j0;
  return o;
}

// The propagating mode of the pair: a = A0 E, its derivatives in z (field) and zeta (source),
// the Bessel functions at k0 R. With it the real part splits as Wr = T - a Y0(k0 R) (T the
// evanescent remainder, smooth over R / h), and Wi = -a J0(k0 R) exactly.
struct FinProp {
  double a, az, af, j0, y0, j1, y1;
};

__device__ __forceinline__ FinProp fin_prop(const FinFreq* __restrict__ fp, const FinTables& F, const WaveTables& T,
                                            int k, int n, double R) {
  const FinFreq& f = *fp;
  const double k0 = f.k0;
  const double4 pz = F.pf[k], pq = F.pf[n];
  const double e3 = pz.x * pq.x, e4 = pz.y * pq.z, e5 = pz.z * pq.y, e6 = pz.w * pq.w;
  FinProp p;
  p.a = f.A0 * (e3 + e4 + e5 + e6);
  p.az = f.A0 * k0 * (e3 + e4 - e5 - e6);
  p.af = f.A0 * k0 * (e3 - e4 + e5 - e6);
  bessel_jy01(T, k0 * R, p.j0, p.y0, p.j1, p.y1);
  return p;
}

}  // namespace hbg
