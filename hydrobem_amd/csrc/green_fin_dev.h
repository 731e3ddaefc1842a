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

namespace hbg {

// ---- the frequency limits in finite depth (green_fin.rs FreqMode::Zero / Infinite): Wi = 0 and
// the regular part Wr = G - sum_p s_p / r_p of the Green function that the limit's boundary
// conditions on z = 0 and z = -h fix:
//   omega = 0 (zero): dG/dz = 0 on both, s_p = +1 for all six Rankine terms,
//     G = -(2/h) (log(R/2h) + gamma) + (4/h) sum_n cos k_n(z+h) cos k_n(zeta+h) K0(k_n R), k_n = n pi / h;
//   omega = inf: G = 0 on z = 0, dG/dz = 0 on z = -h, s_p = (+, +, -, -, -, -),
//     G = (4/h) sum_n cos k_n(z+h) cos k_n(zeta+h) K0(k_n R), k_n = (n - 1/2) pi / h.
// R >= h/2: that series, 24 roots, stopped at k_n R >= 18, as green_fin.rs:256-322 evaluates it.
// R < h/2 (where the reference reads generated tables of the smooth remainder): the same G as
// the sum of its images, the source's reflections in z = 0 and z = -h,
//   G = sum_m e_m [1/rho(d1 + 2 m h) + t_m / rho(s + 2 m h)],  d1 = z - zeta, s = z + zeta,
//   rho(x) = sqrt(R^2 + x^2), e_m = 1, t_m = 1 (omega = 0) or e_m = (-1)^m, t_m = -1 (omega = inf),
// the six Rankine terms being the images m = -1, 0, 1 of the first family and m = 0, 1, 2 of the
// second. The terms |m| <= LIM_M are summed; the four runs beyond go in closed form: at omega = 0
// by Euler-Maclaurin about m = M + 1/2 (the integral of 1/rho in asinh, the log divergence of the
// runs cancelling against the regularization that the series' constant -(2/h) gamma fixes), at
// omega = inf by Boole's summation of the alternating runs about m = M + 1; with the third
// derivative terms kept the truncation is below 1e-11 of the kernel. Checked against the
// mpmath values of the reference's tests/data/green_fin_reference.txt (modes 1 and 2).
constexpr int LIM_M = 32;
constexpr double EULER_GAMMA = 0.57721566490153286061;

// 1/rho and its x derivatives up to the fourth, and the R derivatives of f, f', f''' at x
struct RhoD {
  double f0, f1, f2, f3, f4, f0R, f1R, f3R;
};
__device__ __forceinline__ RhoD rho_derivs(double R, double x) {
  const double r2 = R * R + x * x, ir = 1.0 / sqrt(r2), ir2 = 1.0 / r2;
  const double ir3 = ir * ir2, ir5 = ir3 * ir2, ir7 = ir5 * ir2, ir9 = ir7 * ir2;
  const double x2 = x * x;
  RhoD d;
  d.f0 = ir;
  d.f1 = -x * ir3;
  d.f2 = -ir3 + 3.0 * x2 * ir5;
  d.f3 = 9.0 * x * ir5 - 15.0 * x2 * x * ir7;
  d.f4 = 9.0 * ir5 - 90.0 * x2 * ir7 + 105.0 * x2 * x2 * ir9;
  d.f0R = -R * ir3;
  d.f1R = 3.0 * x * R * ir5;
  d.f3R = -45.0 * x * R * ir7 + 105.0 * x2 * x * R * ir9;
  return d;
}

__device__ __noinline__ FinK fin_limit_kernel(bool zero, double h, double R, double z, double zeta,
                                              const double* __restrict__ ktab) {
  const double d1 = z - zeta, s = z + zeta;
  const double sg = d1 > 0.0 ? 1.0 : (d1 < 0.0 ? -1.0 : 0.0);
  FinK o{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (R >= 0.5 * h) {
    const double v[6] = {fabs(d1), s + 2.0 * h, -s, d1 + 2.0 * h, -d1 + 2.0 * h, s + 4.0 * h};
    const double sp[6] = {1.0, 1.0, zero ? 1.0 : -1.0, zero ? 1.0 : -1.0, zero ? 1.0 : -1.0, zero ? 1.0 : -1.0};
    const double dvz[6] = {-sg, 1.0, -1.0, -1.0, 1.0, 1.0};  // d v_p / d zeta
    const double dvf[6] = {sg, 1.0, -1.0, 1.0, -1.0, 1.0};   // d v_p / d z
    double re = 0.0, reR = 0.0, reZ = 0.0, rez = 0.0;
    const double c = 4.0 / h;
    for (int n = 1; n <= FIN_ROOTS; ++n) {
      const double k = ((double)n - (zero ? 0.0 : 0.5)) * PI / h, x = k * R;
      if (x >= SERIES_STOP) break;
      double K0, K1, sz, cz, ss, cs;
      k01(ktab, x, K0, K1);
      sincos(k * (z + h), &sz, &cz);
      sincos(k * (zeta + h), &ss, &cs);
      re += c * cz * cs * K0;
      reR -= c * cz * cs * k * K1;
      reZ -= c * cz * k * ss * K0;
      rez -= c * k * sz * cs * K0;
    }
    if (zero) {
      re += -2.0 / h * (log(R / (2.0 * h)) + EULER_GAMMA);
      reR += -2.0 / (h * R);
    }
    for (int p = 0; p < 6; ++p) {
      const double r2 = R * R + v[p] * v[p], r = sqrt(r2), ir3 = sp[p] / (r2 * r);
      re -= sp[p] / r;
      reR += R * ir3;
      reZ += v[p] * dvz[p] * ir3;
      rez += v[p] * dvf[p] * ir3;
    }
    o.Wr = re;
    o.WrR = reR;
    o.WrZ = reZ;
    o.Wrz = rez;
    return o;
  }
  // the images: family a (offset d1: dx/dzeta = -1, dx/dz = 1), family b (offset s: both 1)
  double re = 0.0, reR = 0.0, reZ = 0.0, rez = 0.0;
  for (int m = -LIM_M; m <= LIM_M; ++m) {
    const double e = zero ? 1.0 : ((m & 1) ? -1.0 : 1.0);
    if (m < -1 || m > 1) {
      const double x = d1 + 2.0 * m * h, r2 = R * R + x * x, ir = 1.0 / sqrt(r2), ir3 = ir / r2;
      re += e * ir;
      reR -= e * R * ir3;
      reZ += e * x * ir3;
      rez -= e * x * ir3;
    }
    if (m < 0 || m > 2) {
      const double t = zero ? e : -e;
      const double x = s + 2.0 * m * h, r2 = R * R + x * x, ir = 1.0 / sqrt(r2), ir3 = ir / r2;
      re += t * ir;
      reR -= t * R * ir3;
      reZ -= t * x * ir3;
      rez -= t * x * ir3;
    }
  }
  // the four runs beyond |m| = LIM_M, each sum_{m > M} w_m / rho(c + 2 m h): offset c, its
  // derivatives in zeta and z, and the weight of the run's first term at omega = inf
  const double cs4[4] = {d1, -d1, s, -s};
  const double dcz[4] = {-1.0, 1.0, 1.0, -1.0};
  const double dcf[4] = {1.0, -1.0, 1.0, -1.0};
  const double h2 = 2.0 * h, h23 = h2 * h2 * h2;
  for (int i = 0; i < 4; ++i) {
    const double c = cs4[i];
    double T, TR, Tc;
    if (zero) {
      // Euler-Maclaurin about M + 1/2: the integral (1/2h)[log(2h/R) - asinh(u/R)] (the
      // log(2h/R) of the regularization), + f'/24 - 7 f'''/5760 (in m)
      const double u = c + (2.0 * LIM_M + 1.0) * h, q = sqrt(u * u + R * R);
      const RhoD d = rho_derivs(R, u);
      T = log(h2 / (u + q)) / h2 + h2 * d.f1 / 24.0 - 7.0 * h23 * d.f3 / 5760.0;
      TR = -(R / q) / (u + q) / h2 + h2 * d.f1R / 24.0 - 7.0 * h23 * d.f3R / 5760.0;
      Tc = -1.0 / (q * h2) + h2 * d.f2 / 24.0 - 7.0 * h23 * d.f4 / 5760.0;
    } else {
      // Boole about m = M + 1: f/2 - f'/4 + f'''/48 (in m), the run's first sign
      const double w = (((LIM_M + 1) & 1) ? -1.0 : 1.0) * (i < 2 ? 1.0 : -1.0);
      const RhoD d = rho_derivs(R, c + 2.0 * (LIM_M + 1) * h);
      T = w * (0.5 * d.f0 - 0.25 * h2 * d.f1 + h23 * d.f3 / 48.0);
      TR = w * (0.5 * d.f0R - 0.25 * h2 * d.f1R + h23 * d.f3R / 48.0);
      Tc = w * (0.5 * d.f1 - 0.25 * h2 * d.f2 + h23 * d.f4 / 48.0);
    }
    re += T;
    reR += TR;
    reZ += Tc * dcz[i];
    rez += Tc * dcf[i];
  }
  if (zero) re -= 2.0 / h * EULER_GAMMA;
  o.Wr = re;
  o.WrR = reR;
  o.WrZ = reZ;
  o.Wrz = rez;
  return o;
}

}  // namespace hbg
