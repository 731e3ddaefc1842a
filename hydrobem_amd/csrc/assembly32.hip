// The FP32 far-field assembly of the precision path (HBG_PRECISION_MIXED32): the influence
// matrices S and A = 2 pi I - D of one frequency in complex64, the wave kernel evaluated in FP32
// from FP32 tables of lower degree, the statics read as FP32 records, and the self and adjacent
// pairs (centroids closer than NEAR_RATIO times the sum of the panels' radii) redone in FP64
// (statics, wave kernel and accumulation) by a second kernel that also keeps, per such entry,
// the difference between its FP64 value and the complex64 value stored, so that the right-hand
// sides and the refinement's residuals see the FP64 entries there.
//
// The entries are those of lo_frequency.rs:399-537 with low_order.rs:738-769 (as the FP64 kernels
// of frequency.hip compute them); the FP32 kernel follows the same forms (green_inf.rs:340-349 the
// identity form of the infinite-depth kernel, green_fin.rs:256-322 and 1007-1095 the finite-depth
// series and the Newman decomposition) on tables truncated from the FP64 ones: each FP64 tile's
// monomial expansion is converted to Chebyshev form, cut to the FP32 degree and converted back
// (the dropped Chebyshev coefficients bound the truncation error: at most 6e-9 of the values for
// the infinite-depth tiles of degree 6), then rounded to FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "green_fin_dev.h"
#include "green_inf_dev.h"
#include "hbg_internal.h"
#include "pair_sort.h"

namespace hbg {

// ------------------------------------------------------------------ FP32 tables
constexpr int D2F = 6, M2F = D2F + 1, TILE2F = 2 * M2F * M2F;  // V, V_X: [f][y power][x power]
constexpr int D1F = 6, M1F = D1F + 1, TILE1F = 4 * M1F;        // J0, J1, Yt, Yt': [f][power]
constexpr int FDF = 7, MFF = FDF + 1, FTILEF = 3 * MFF * MFF;  // Phi: [value, d/dxb, d/dvb][vb power][xb power]
constexpr int KDF = 11, KTILEF = 2 * (KDF + 1);                 // K0, K1: [f][power]
constexpr double NEAR_RATIO = 2.0;  // pairs in FP64: |c_k - c_n| < NEAR_RATIO (radius_k + radius_n)

// The FP32 evaluation's square roots, reciprocals and logarithms by the hardware's instructions
// (v_sqrt_f32, v_rcp_f32, v_log_f32, about 1 ulp) instead of the correctly rounded library
// sequences: the FP32 path's accuracy is that of its tables and of FP32 arithmetic (about 1e-7
// relative), not half an ulp here. The exponential stays the library's (its argument reaches 48,
// where the instruction's argument rounding would cost digits).
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

struct WaveTables32 {
  const float* t2d = nullptr;  // [YT][XT][TILE2F]
  const float* t1d = nullptr;  // [XT][TILE1F]
};

struct FinFreq32 {
  float h, K, k0, A0;
  float kn[FIN_ROOTS], cn[FIN_ROOTS];
};

struct FinTables32 {
  const float* phi1;   // [FTS][FTW][FTILEF]
  const float* phi2;   // [FTS][FTV][FTILEF]
  const float* ktab;   // [KTILES][KTILEF]
  const float4* pf;    // [N] propagating factors
  const float2* pcs;   // [N][FIN_ROOTS] (cos, sin) k_n (z + h)
};

// centroid and area, normal and radius of a panel in FP32
struct PanelF {
  float4 ca;
  float4 nr;
};

FinFreq32 fin_freq32(const FinFreq& f) {
  FinFreq32 g;
  g.h = (float)f.h;
  g.K = (float)f.K;
  g.k0 = (float)f.k0;
  g.A0 = (float)f.A0;
  for (int m = 0; m < FIN_ROOTS; ++m) {
    g.kn[m] = (float)f.kn[m];
    g.cn[m] = (float)f.cn[m];
  }
  return g;
}

// The map of a degree-D monomial expansion on [-1, 1] to the monomial coefficients of its
// Chebyshev truncation to degree d: P [D + 1][d + 1], P[a][a'] = the coefficient of t^a' in the
// truncation of t^a (t^a = 2^(1 - a) sum_j C(a, j) T_(a - 2j), the middle term halved).
std::vector<double> truncation_matrix(int D, int d) {
  std::vector<long double> Tm((size_t)(D + 1) * (D + 1), 0.0L);  // Tm[k][a]: t^a coefficient of T_k
  Tm[0] = 1.0L;
  if (D >= 1) Tm[(size_t)1 * (D + 1) + 1] = 1.0L;
  for (int k = 2; k <= D; ++k)
    for (int a = 0; a <= k; ++a) {
      long double v = -Tm[(size_t)(k - 2) * (D + 1) + a];
      if (a > 0) v += 2.0L * Tm[(size_t)(k - 1) * (D + 1) + a - 1];
      Tm[(size_t)k * (D + 1) + a] = v;
    }
  std::vector<double> P((size_t)(D + 1) * (d + 1), 0.0);
  for (int a = 0; a <= D; ++a) {
    // t^a in Chebyshev form: m[k]
    std::vector<long double> m(D + 1, 0.0L);
    long double binom = 1.0L, scale = std::ldexp(1.0L, 1 - a);  // a = 0: 2 x the halved T_0 term = 1
    for (int j = 0; 2 * j <= a; ++j) {
      const int k = a - 2 * j;
      long double c = binom * scale;
      if (k == 0) c *= 0.5L;
      m[k] += c;
      binom = binom * (a - j) / (j + 1);
    }
    for (int ap = 0; ap <= d; ++ap) {
      long double v = 0.0L;
      for (int k = ap; k <= d; ++k) v += m[k] * Tm[(size_t)k * (D + 1) + ap];
      P[(size_t)a * (d + 1) + ap] = (double)v;
    }
  }
  return P;
}

namespace {

// 2D tiles [nf][M][M] (y power, x power) -> [nf][m][m] through P [M][m] on both axes, in FP32
__global__ void truncate2d_kernel(const double* __restrict__ src, size_t sstride, int nf, int M,
                                  float* __restrict__ dst, size_t dstride, int m, const double* __restrict__ P,
                                  int ntiles) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)nf * m * m;
  if (id >= per * ntiles) return;
  const size_t tile = id / per;
  const int r = (int)(id % per), f = r / (m * m), bp = (r / m) % m, ap = r % m;
  const double* c = src + tile * sstride + (size_t)f * M * M;
  double v = 0.0;
  for (int b = 0; b < M; ++b) {
    double row = 0.0;
    for (int a = 0; a < M; ++a) row += P[(size_t)a * m + ap] * c[(size_t)b * M + a];
    v += P[(size_t)b * m + bp] * row;
  }
  dst[tile * dstride + ((size_t)f * m + bp) * m + ap] = (float)v;
}

// 1D tiles [M][nf] (power, function interleaved) -> [nf][m]
__global__ void truncate1d_kernel(const double* __restrict__ src, size_t sstride, int nf, int M,
                                  float* __restrict__ dst, size_t dstride, int m, const double* __restrict__ P,
                                  int ntiles) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)nf * m;
  if (id >= per * ntiles) return;
  const size_t tile = id / per;
  const int r = (int)(id % per), f = r / m, ap = r % m;
  const double* c = src + tile * sstride;
  double v = 0.0;
  for (int a = 0; a < M; ++a) v += P[(size_t)a * m + ap] * c[(size_t)a * nf + f];
  dst[tile * dstride + (size_t)f * m + ap] = (float)v;
}

__global__ void panels32_kernel(const PanelD* __restrict__ P, int N, PanelF* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const PanelD& p = P[j];
  PanelF q;
  q.ca = make_float4((float)p.c[0], (float)p.c[1], (float)p.c[2], (float)p.area);
  q.nr = make_float4((float)p.n[0], (float)p.n[1], (float)p.n[2], (float)p.radius);
  out[j] = q;
}

__global__ void factors32_kernel(const double4* __restrict__ pf, const double2* __restrict__ pcs, int N,
                                 float4* __restrict__ pf32, float2* __restrict__ pcs32) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)N) {
    const double4 v = pf[i];
    pf32[i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
  }
  if (i < (size_t)N * FIN_ROOTS) {
    const double2 v = pcs[i];
    pcs32[i] = make_float2((float)v.x, (float)v.y);
  }
}

// ------------------------------------------------------------------ FP32 device evaluation
typedef __attribute__((address_space(4))) const float cfloat;

__device__ __forceinline__ cfloat* uniform_tile32(const float* base, int u, int stride) {
  cfloat* c = (cfloat*)(base + (size_t)u * stride);
  asm volatile("" : "+s"(c));
  return c;
}

template <int M>
__device__ __forceinline__ void powers32(float t, float* pw) {
  pw[0] = 1.0f;
  pw[1] = t;
#pragma unroll
  for (int a = 2; a < M; ++a) pw[a] = pw[a - 1] * t;
}

// sum_a c[a] t^a with the powers given, each term one FMA with a scalar coefficient
template <int M>
__device__ __forceinline__ float poly32_u(cfloat* c, const float* pw) {
  float r = c[0];
#pragma unroll
  for (int a = 1; a < M; ++a) r = fmaf(c[a], pw[a], r);
  return r;
}

// sum_b ty^b sum_a c[b M + a] tx^a (the x powers given)
template <int M>
__device__ __forceinline__ float poly2d32_u(cfloat* c, const float* pwx, float ty) {
  float s = 0.0f;
#pragma unroll
  for (int b = M - 1; b >= 0; --b) s = fmaf(s, ty, poly32_u<M>(c + b * M, pwx));
  return s;
}

constexpr float PI_F = 3.14159265358979323846f;
constexpr float GAMMA_F = 0.5772156649015329f;
constexpr float LOG2_F = 0.69314718055994530942f;

__device__ __forceinline__ int tile1d32(float X, float& t) {
  int i = (int)X;
  i = i < 0 ? 0 : (i >= XT ? XT - 1 : i);
  t = 2.0f * (X - (float)i) - 1.0f;
  return i;
}

__device__ __forceinline__ int tile2d32(float X, float Y, float& tx, float& ty) {
  int ix = (int)X, iy = (int)Y;
  ix = ix < 0 ? 0 : (ix >= XT ? XT - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= YT ? YT - 1 : iy);
  tx = 2.0f * (X - (float)ix) - 1.0f;
  ty = 2.0f * (Y - (float)iy) - 1.0f;
  return iy * XT + ix;
}

// (J0, J1, Yt, Yt') at X from the FP32 1D tiles, wavefront-uniform rounds
__device__ __forceinline__ void table1d32_u(const float* t1d, bool act, float X, float& j0, float& j1, float& yt,
                                            float& dyt) {
  float t;
  const int i1 = tile1d32(X, t);
  HBG_UNIFORM_ROUNDS(i1, act, u, mine, {
    cfloat* c = uniform_tile32(t1d, u, TILE1F);
    if (mine) {
      float pw[M1F];
      powers32<M1F>(t, pw);
      j0 = poly32_u<M1F>(c, pw);
      j1 = poly32_u<M1F>(c + M1F, pw);
      yt = poly32_u<M1F>(c + 2 * M1F, pw);
      dyt = poly32_u<M1F>(c + 3 * M1F, pw);
    }
  })
}

// (V, V_X) at (X, Y) from the FP32 2D tiles, wavefront-uniform rounds
__device__ __forceinline__ void table2d32_u(const float* t2d, bool act, float X, float Y, float& V, float& VX) {
  float tx, ty;
  const int i2 = tile2d32(X, Y, tx, ty);
  HBG_UNIFORM_ROUNDS(i2, act, u, mine, {
    cfloat* c = uniform_tile32(t2d, u, TILE2F);
    if (mine) {
      float pw[M2F];
      powers32<M2F>(tx, pw);
      V = poly2d32_u<M2F>(c, pw, ty);
      VX = poly2d32_u<M2F>(c + M2F * M2F, pw, ty);
    }
  })
}

struct WaveK32 {
  float F, FX, I, IX;
};

// The identity form inside the tables (X < XT, Y < YT; the host keeps every pair there).
__device__ __forceinline__ WaveK32 wave_tab32(float X, float Y, float j0, float j1, float yt, float dyt, float V,
                                              float VX) {
  WaveK32 w;
  const float R0 = fsqrt(X * X + Y * Y);
  const float eY = expf(-Y);
  const float lg = flog(Y + R0) + (GAMMA_F - LOG2_F);
  const float iR0 = frcp(R0);
  w.F = -2.0f * eY * j0 * lg - PI_F * eY * yt - 2.0f * R0 * V;
  w.FX = 2.0f * eY * j1 * lg - 2.0f * eY * j0 * X * iR0 * frcp(Y + R0) - PI_F * eY * dyt - 2.0f * (X * iR0) * V -
         2.0f * R0 * VX;
  w.I = -2.0f * PI_F * eY * j0;
  w.IX = 2.0f * PI_F * eY * j1;
  return w;
}

__device__ __forceinline__ WaveK32 wave_kernel32_u(const WaveTables32& T, bool valid, float X, float Y) {
  float j0 = 0.0f, j1 = 0.0f, yt = 0.0f, dyt = 0.0f, V = 0.0f, VX = 0.0f;
  table1d32_u(T.t1d, valid, X, j0, j1, yt, dyt);
  table2d32_u(T.t2d, valid, X, Y, V, VX);
  return wave_tab32(X, Y, j0, j1, yt, dyt, V, VX);
}

// The tile of x in [1/2, 18) of the K0, K1 table
__device__ __forceinline__ int ktile32(float x, float& t) {
  int i;
  if (x < 1.0f) {
    i = 0;
    t = 4.0f * (x - 0.5f) - 1.0f;
  } else {
    i = (int)x;
    if (i > KTILES - 1) i = KTILES - 1;
    t = 2.0f * (x - (float)i) - 1.0f;
  }
  return i;
}

__device__ __forceinline__ int phi_index32(int ny, float ymax, float xb, float vb, float& tx, float& ty) {
  const float ux = xb * (FTS / 0.5f), uy = vb * ((float)ny / ymax);
  int ix = (int)ux, iy = (int)uy;
  ix = ix < 0 ? 0 : (ix > FTS - 1 ? FTS - 1 : ix);
  iy = iy < 0 ? 0 : (iy > ny - 1 ? ny - 1 : iy);
  tx = 2.0f * (ux - (float)ix) - 1.0f;
  ty = 2.0f * (uy - (float)iy) - 1.0f;
  return ix * ny + iy;
}

__device__ __forceinline__ void phi_eval32_u(cfloat* c, float tx, float ty, float* out) {
  float pw[MFF];
  powers32<MFF>(tx, pw);
#pragma unroll
  for (int q = 0; q < 3; ++q) out[q] = poly2d32_u<MFF>(c + q * MFF * MFF, pw, ty);
}

// (Phi, dPhi/dxb, dPhi/dvb) of the FP32 tile family, wavefront-uniform rounds
__device__ __forceinline__ void phi32_u(const float* tab, int ny, float ymax, bool act, float xb, float vb,
                                        float* out) {
  float tx, ty;
  const int i = phi_index32(ny, ymax, xb, vb, tx, ty);
  HBG_UNIFORM_ROUNDS(i, act, u, mine, {
    cfloat* c = uniform_tile32(tab, u, FTILEF);
    if (mine) phi_eval32_u(c, tx, ty, out);
  })
}

struct FinK32 {
  float Wr, WrR, WrZ, Wrz, Wi, WiR, WiZ, Wiz;
};

// The finite-depth kernel of fin_kernel_u (green_fin_dev.h) in FP32 on the FP32 tables.
__device__ __forceinline__ FinK32 fin_kernel32_u(const FinFreq32* __restrict__ fp, const FinTables32& F,
                                                 const WaveTables32& T, bool valid, int k, int n, float R, float z,
                                                 float zeta) {
  const FinFreq32& f = *fp;
  const float h = f.h, K = f.K, k0 = f.k0, A0 = f.A0;
  const float d1 = z - zeta;
  const float sg = d1 > 0.0f ? 1.0f : (d1 < 0.0f ? -1.0f : 0.0f);
  const float v1 = fabsf(d1), v2 = z + zeta + 2.0f * h, v3 = -(z + zeta);
  const float4 pz = F.pf[k], pq = F.pf[n];
  const float e3 = pz.x * pq.x, e4 = pz.y * pq.z, e5 = pz.z * pq.y, e6 = pz.w * pq.w;
  const float E = e3 + e4 + e5 + e6, Ez = k0 * (e3 + e4 - e5 - e6), Ef = k0 * (e3 - e4 + e5 - e6);
  const float x = k0 * R;
  const bool series = valid && R >= 0.5f * h, near = valid && !(R >= 0.5f * h);
  float j0 = 0.0f, j1 = 0.0f, yt = 0.0f, dyt = 0.0f;
  table1d32_u(T.t1d, valid, x, j0, j1, yt, dyt);
  float Wr = 0.0f, WrR = 0.0f, WrZ = 0.0f, Wrz = 0.0f;
  if (__ballot(series)) {
    float y0 = 0.0f, y1 = 0.0f;
    if (series) {
      const float L = flog(0.5f * x) + GAMMA_F;
      y0 = yt + 2.0f / PI_F * L * j0;
      y1 = -dyt - 2.0f / PI_F * (j0 * frcp(x) - L * j1);
    }
    float re = -A0 * E * y0, reR = A0 * E * k0 * y1, reZ = -A0 * Ez * y0, rez = -A0 * Ef * y0;
    const float2* cz = F.pcs + (size_t)k * FIN_ROOTS;
    const float2* cq = F.pcs + (size_t)n * FIN_ROOTS;
    bool on = series;
    for (int m = 0; m < FIN_ROOTS; ++m) {
      const float kk = f.kn[m], xm = kk * R;
      on = on && xm < (float)SERIES_STOP;
      if (!__ballot(on)) break;
      float t, K0 = 0.0f, K1 = 0.0f;
      const int it = ktile32(xm, t);
      HBG_UNIFORM_ROUNDS(it, on, u, mine, {
        cfloat* c = uniform_tile32(F.ktab, u, KTILEF);
        if (mine) {
          float pw[KDF + 1];
          powers32<KDF + 1>(t, pw);
          K0 = poly32_u<KDF + 1>(c, pw);
          K1 = poly32_u<KDF + 1>(c + KDF + 1, pw);
        }
      })
      if (on) {
        const float2 a = cz[m], b = cq[m];
        const float c = f.cn[m];
        re += c * a.x * b.x * K0;
        reR -= c * a.x * b.x * kk * K1;
        reZ -= c * a.x * kk * b.y * K0;
        rez -= c * kk * a.y * b.x * K0;
      }
    }
    if (series) {
      const float v[6] = {v1, v2, v3, d1 + 2.0f * h, -d1 + 2.0f * h, z + zeta + 4.0f * h};
      const float dvz[6] = {-sg, 1.0f, -1.0f, -1.0f, 1.0f, 1.0f}, dvf[6] = {sg, 1.0f, -1.0f, 1.0f, -1.0f, 1.0f};
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        const float ir = rsqrtf(R * R + v[p] * v[p]), ir3 = ir * ir * ir;
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
    const float ih = frcp(h), ih2 = ih * ih;
    float f1[3] = {0.0f, 0.0f, 0.0f}, f2[3] = {0.0f, 0.0f, 0.0f};
    phi32_u(F.phi1, FTW, 1.0f, near, R * ih, v1 * ih, f1);
    phi32_u(F.phi2, FTV, 2.0f, near, R * ih, v2 * ih, f2);
    const float X = K * R, Y = K * v3;
    float wj0 = 0.0f, wj1 = 0.0f, wyt = 0.0f, wdyt = 0.0f, V = 0.0f, VX = 0.0f;
    table1d32_u(T.t1d, near, X, wj0, wj1, wyt, wdyt);
    table2d32_u(T.t2d, near, X, Y, V, VX);
    if (near) {
      const WaveK32 w = wave_tab32(X, Y, wj0, wj1, wyt, wdyt, V, VX);
      const float FY = -w.F - 2.0f * __builtin_amdgcn_rsqf(X * X + Y * Y);
      Wr = ih * (f1[0] + f2[0]) + K * w.F;
      WrR = ih2 * (f1[1] + f2[1]) + K * K * w.FX;
      const float A1z = -sg * ih2 * f1[2];
      const float Bz = ih2 * f2[2] - K * K * FY;
      WrZ = A1z + Bz;
      Wrz = -A1z + Bz;
    }
  }
  FinK32 o;
  o.Wr = Wr;
  o.WrR = WrR;
  o.WrZ = WrZ;
  o.Wrz = Wrz;
  o.Wi = -A0 * E * j0;
  o.WiR = A0 * E * k0 * j1;
  o.WiZ = -A0 * Ez * j0;
  o.Wiz = -A0 * Ef * j0;
  return o;
}

// one entry (row = field point, col = source panel) of S and A in complex64
// (A2, nullable: a second copy of A, ld lds: the one the refinement's residuals read after the LU
// has factored A in place; the solve copies A there after the assembly in one coalesced pass,
// which costs less than a third stream of scattered writes here)
__device__ __forceinline__ void put_entry32(float2* __restrict__ S, float2* __restrict__ A, float2* __restrict__ A2,
                                            int lds, int lda, int row, int col, float4 e, float Ar, float nz, float a,
                                            float c2, float Wr, float WrR, float WrZ, float Wi, float WiR, float WiZ,
                                            float free_term) {
  S[(size_t)col * lds + row] = make_float2(e.x + Ar * Wr - c2 * e.z, Ar * Wi);
  const float dre = e.y + Ar * (-WrR * a + WrZ * nz) - c2 * e.w, dim = Ar * (-WiR * a + WiZ * nz);
  const float2 v = make_float2((row == col ? free_term : 0.0f) - dre, -dim);
  A[(size_t)col * lda + row] = v;
  if (A2) A2[(size_t)col * lds + row] = v;
}

// Infinite depth, FP32: every unordered pair once (64 x 64 tiles above the diagonal), the pairs
// sorted by the V tile of (X, Y) as assemble_kernel does.
__global__ void __launch_bounds__(256) assemble32_kernel(const PanelF* __restrict__ P, int N, float K, float logc,
                                                         WaveTables32 T, const float4* __restrict__ st,
                                                         float2* __restrict__ S, float2* __restrict__ A,
                                                         float2* __restrict__ A2, int lda, float free_term) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  __shared__ PairSort L;
  const unsigned total = sort_tile_pairs(L, bi, bj, N, [&](int k, int n) {
    const float4 ck = P[k].ca, cn = P[n].ca;
    const float dx = ck.x - cn.x, dy = ck.y - cn.y;
    float tx, ty;
    return tile2d32(K * fsqrt(dx * dx + dy * dy), -K * (ck.z + cn.z), tx, ty);
  });
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float c2 = 2.0f * K * logc, K2 = K * K;
  const int chunks = (int)((total + 63) / 64);
  for (int ch = wave; ch < chunks; ch += 4) {
    const unsigned pos = (unsigned)ch * 64 + lane;
    const bool valid = pos < total;
    const int p = valid ? L.order[pos] : 0;
    const int k = bi * ASM_TILE + (p & 63), n = bj * ASM_TILE + (p >> 6);
    const PanelF pk = P[k], pn = P[n];
    const float dx = pk.ca.x - pn.ca.x, dy = pk.ca.y - pn.ca.y;
    const float R = fsqrt(dx * dx + dy * dy);
    const float X = K * R, Y = -K * (pk.ca.z + pn.ca.z);
    const WaveK32 w = wave_kernel32_u(T, valid, X, Y);
    if (valid) {
      const float R0 = fsqrt(X * X + Y * Y);
      const float FY = -w.F - 2.0f * frcp(R0), IY = -w.I;
      const float Wr = K * w.F, WrR = K2 * w.FX, WrZ = -K2 * FY;
      const float Wi = K * w.I, WiR = K2 * w.IX, WiZ = -K2 * IY;
      const float an = R > 0.0f ? (dx * pn.nr.x + dy * pn.nr.y) * frcp(R) : 0.0f;
      put_entry32(S, A, A2, N, lda, k, n, st[(size_t)n * N + k], pn.ca.w, pn.nr.z, an, c2, Wr, WrR, WrZ, Wi, WiR, WiZ,
                  free_term);
      if (k != n) {
        const float ak = R > 0.0f ? -(dx * pk.nr.x + dy * pk.nr.y) * frcp(R) : 0.0f;
        put_entry32(S, A, A2, N, lda, n, k, st[(size_t)k * N + n], pk.ca.w, pk.nr.z, ak, c2, Wr, WrR, WrZ, Wi, WiR, WiZ,
                    free_term);
      }
    }
  }
}

constexpr int FS_NEAR32 = FTS * FTW * FTV;

// Finite depth, FP32: the pairs sorted as in assemble_fin_kernel (Phi tiles near the axis, bins
// of R / h in the series).
__global__ void __launch_bounds__(256) assemble_fin32_kernel(const PanelF* __restrict__ P, int N,
                                                             const FinFreq32* __restrict__ f, FinTables32 F,
                                                             WaveTables32 T, float logc,
                                                             const float4* __restrict__ st, float2* __restrict__ S,
                                                             float2* __restrict__ A, float2* __restrict__ A2, int lda,
                                                             float free_term) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  __shared__ PairSort L;
  const float h = f->h, ih = frcp(h);
  const unsigned total = sort_tile_pairs(L, bi, bj, N, [&](int k, int n) {
    const float4 ck = P[k].ca, cn = P[n].ca;
    const float dx = ck.x - cn.x, dy = ck.y - cn.y;
    const float R = fsqrt(dx * dx + dy * dy);
    if (R >= 0.5f * h) {
      const float b = (R * ih - 0.5f) * 32.0f;
      return FS_NEAR32 + (b < 1023.0f ? (int)b : 1023);
    }
    float tx, ty;
    const int i1 = phi_index32(FTW, 1.0f, R * ih, fabsf(ck.z - cn.z) * ih, tx, ty);
    const int i2 = phi_index32(FTV, 2.0f, R * ih, (ck.z + cn.z + 2.0f * h) * ih, tx, ty);
    return i1 * FTV + i2 % FTV;
  });
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float c2 = 2.0f * f->K * logc;
  const int chunks = (int)((total + 63) / 64);
  for (int ch = wave; ch < chunks; ch += 4) {
    const unsigned pos = (unsigned)ch * 64 + lane;
    const bool valid = pos < total;
    const int p = valid ? L.order[pos] : 0;
    const int k = bi * ASM_TILE + (p & 63), n = bj * ASM_TILE + (p >> 6);
    const PanelF pk = P[k], pn = P[n];
    const float dx = pk.ca.x - pn.ca.x, dy = pk.ca.y - pn.ca.y;
    const float R = fsqrt(dx * dx + dy * dy);
    const FinK32 w = fin_kernel32_u(f, F, T, valid, k, n, R, pk.ca.z, pn.ca.z);
    if (valid) {
      const float an = R > 0.0f ? (dx * pn.nr.x + dy * pn.nr.y) * frcp(R) : 0.0f;
      put_entry32(S, A, A2, N, lda, k, n, st[(size_t)n * N + k], pn.ca.w, pn.nr.z, an, c2, w.Wr, w.WrR, w.WrZ, w.Wi,
                  w.WiR, w.WiZ, free_term);
      if (k != n) {
        const float ak = R > 0.0f ? -(dx * pk.nr.x + dy * pk.nr.y) * frcp(R) : 0.0f;
        put_entry32(S, A, A2, N, lda, n, k, st[(size_t)k * N + n], pk.ca.w, pk.nr.z, ak, c2, w.Wr, w.WrR, w.Wrz, w.Wi,
                    w.WiR, w.Wiz, free_term);
      }
    }
  }
}

// The near entries in FP64, one thread per entry (row k, column n) of the list: S and A of the
// pair from the FP64 statics and the FP64 kernel (frequency.hip's forms), stored into the
// complex64 matrices, and their differences from the stored complex64 values into dS, dA.
// finite: the finite-depth kernel (f, F) else the infinite-depth one at K.
__global__ void __launch_bounds__(256) near64_kernel(const PanelD* __restrict__ P, int N, const int2* __restrict__ ent,
                                                     int nnz, int finite, double K, const FinFreq* __restrict__ f,
                                                     FinTables F, WaveTables T, double logc,
                                                     const StaticD* __restrict__ st, float2* __restrict__ S,
                                                     float2* __restrict__ A, float2* __restrict__ A2, int lda,
                                                     double free_term, double2* __restrict__ dS,
                                                     double2* __restrict__ dA) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int k = ent[e].x, n = ent[e].y;
  const PanelD* pk = P + k;
  const PanelD* pn = P + n;
  const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
  const double R = sqrt(dx * dx + dy * dy);
  double Wr, WrR, WrZ, Wi, WiR, WiZ, KK;
  if (finite) {
    const FinK w = fin_kernel(f, F, T, k, n, R, pk->c[2], pn->c[2]);
    Wr = w.Wr;
    WrR = w.WrR;
    WrZ = w.WrZ;
    Wi = w.Wi;
    WiR = w.WiR;
    WiZ = w.WiZ;
    KK = f->K;
  } else {
    const double X = K * R, Y = -K * (pk->c[2] + pn->c[2]);
    const WaveK w = wave_kernel(T, X, Y);
    const double R0 = sqrt(X * X + Y * Y);
    const double FY = -w.F - 2.0 / R0, IY = -w.I;
    Wr = K * w.F;
    WrR = K * K * w.FX;
    WrZ = -K * K * FY;
    Wi = K * w.I;
    WiR = K * K * w.IX;
    WiZ = -K * K * IY;
    KK = K;
  }
  const double c2 = 2.0 * KK * logc;
  const StaticD s = st[(size_t)n * N + k];
  const double Ar = pn->area, a = R > 0.0 ? (dx * pn->n[0] + dy * pn->n[1]) / R : 0.0, nz = pn->n[2];
  const double2 sv = make_double2(s.rs + Ar * Wr - c2 * s.dl, Ar * Wi);
  const double dre = s.rd + Ar * (-WrR * a + WrZ * nz) - c2 * s.dd, dim = Ar * (-WiR * a + WiZ * nz);
  const double2 av = make_double2((k == n ? free_term : 0.0) - dre, -dim);
  const float2 s32 = make_float2((float)sv.x, (float)sv.y), a32 = make_float2((float)av.x, (float)av.y);
  S[(size_t)n * N + k] = s32;
  A[(size_t)n * lda + k] = a32;
  if (A2) A2[(size_t)n * N + k] = a32;  // nullable: the solve copies A whole after the assembly
  dS[e] = make_double2(sv.x - (double)s32.x, sv.y - (double)s32.y);
  dA[e] = make_double2(av.x - (double)a32.x, av.y - (double)a32.y);
}

// Row sums over the near entries: for row k (one wavefront; lanes over the row's entries, 64 at a
// time, then a fixed-order butterfly), out[c * ldo + k] (+)= sign * sum_e d_e v[c * N + col_e] for
// every column c < R of v. accumulate 0 stores the sum, 1 adds it. blockIdx.y: the system (d, v
// and out advanced by their strides).
__global__ void __launch_bounds__(256) near_rows_kernel(const int* __restrict__ rowptr, const int2* __restrict__ ent,
                                                        const double2* __restrict__ d, size_t dstride,
                                                        const double2* __restrict__ v, size_t vstride, int N, int R,
                                                        double sign, double2* __restrict__ out, size_t ostride,
                                                        int ldo, int accumulate) {
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= N) return;
  d += blockIdx.y * dstride;
  v += blockIdx.y * vstride;
  out += blockIdx.y * ostride;
  const int e0 = rowptr[k], e1 = rowptr[k + 1];
  for (int c = 0; c < R; ++c) {
    double2 acc = make_double2(0.0, 0.0);
    for (int e = e0 + lane; e < e1; e += 64) {
      const double2 a = d[e], x = v[(size_t)c * N + ent[e].y];
      acc.x = fma(a.x, x.x, fma(-a.y, x.y, acc.x));
      acc.y = fma(a.x, x.y, fma(a.y, x.x, acc.y));
    }
    for (int off = 32; off >= 1; off >>= 1) {
      acc.x += __shfl_xor(acc.x, off);
      acc.y += __shfl_xor(acc.y, off);
    }
    if (lane == 0) {
      double2* o = out + (size_t)c * ldo + k;
      const double2 prev = accumulate ? *o : make_double2(0.0, 0.0);
      *o = make_double2(prev.x + sign * acc.x, prev.y + sign * acc.y);
    }
  }
}

// the FP32 evaluators at points, for the tests: infinite depth out [n][4] = Ft, Ft_X, Gi, Gi_X
__global__ void wave_points32_kernel(WaveTables32 T, const double* __restrict__ X, const double* __restrict__ Y,
                                     int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const float x = valid ? (float)X[i] : 1.0f, y = valid ? (float)Y[i] : 1.0f;
  const WaveK32 w = wave_kernel32_u(T, valid, x, y);
  if (valid) {
    out[4 * i] = w.F;
    out[4 * i + 1] = w.FX;
    out[4 * i + 2] = w.I;
    out[4 * i + 3] = w.IX;
  }
}

// finite depth out [n][8] = Wr, WrR, WrZ, Wrz, Wi, WiR, WiZ, Wiz as fin_points_eval_kernel: the
// factors of point i's depths at "panels" i (z) and n + i (zeta)
__global__ void fin_points32_kernel(const FinFreq32* __restrict__ f, FinTables32 F, WaveTables32 T,
                                    const double* __restrict__ R, const double* __restrict__ z,
                                    const double* __restrict__ zeta, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const int ii = valid ? i : 0;
  const FinK32 w = fin_kernel32_u(f, F, T, valid, ii, n + ii, (float)R[ii], (float)z[ii], (float)zeta[ii]);
  if (valid) {
    double* o = out + 8 * (size_t)i;
    o[0] = w.Wr;
    o[1] = w.WrR;
    o[2] = w.WrZ;
    o[3] = w.Wrz;
    o[4] = w.Wi;
    o[5] = w.WiR;
    o[6] = w.WiZ;
    o[7] = w.Wiz;
  }
}

}  // namespace

// ------------------------------------------------------------------ host side
// The FP32 tables of a device from its FP64 ones (once per device).
void build_tables32(const WaveTables& T64, const double* ktab64, float* t2d32, float* t1d32, float* ktab32,
                    hipStream_t stream) {
  const std::vector<double> P2 = truncation_matrix(D2, D2F), P1 = truncation_matrix(D1, D1F),
                            PK = truncation_matrix(KD, KDF);
  double* dP = nullptr;
  const size_t n2 = P2.size(), n1 = P1.size(), nk = PK.size();
  HBG_HIP(hipMalloc(&dP, (n2 + n1 + nk) * sizeof(double)));
  HBG_HIP(hipMemcpy(dP, P2.data(), n2 * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dP + n2, P1.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dP + n2 + n1, PK.data(), nk * sizeof(double), hipMemcpyHostToDevice));
  const size_t o2 = (size_t)XT * YT * 2 * M2F * M2F, o1 = (size_t)XT * 4 * M1F, ok = (size_t)KTILES * 2 * (KDF + 1);
  truncate2d_kernel<<<(unsigned)((o2 + 255) / 256), 256, 0, stream>>>(T64.t2d, TILE2, 2, M2, t2d32, TILE2F, M2F, dP,
                                                                       XT * YT);
  truncate1d_kernel<<<(unsigned)((o1 + 255) / 256), 256, 0, stream>>>(T64.t1d, TILE1, 4, M1, t1d32, TILE1F, M1F,
                                                                       dP + n2, XT);
  truncate1d_kernel<<<(unsigned)((ok + 255) / 256), 256, 0, stream>>>(ktab64, KTILE, 2, KD + 1, ktab32, KTILEF,
                                                                       KDF + 1, dP + n2 + n1, KTILES);
  HBG_HIP(hipGetLastError());
  HBG_HIP(hipStreamSynchronize(stream));
  HBG_HIP(hipFree(dP));
}
size_t tables32_t2d_elems() { return (size_t)XT * YT * TILE2F; }
size_t tables32_t1d_elems() { return (size_t)XT * TILE1F; }
size_t tables32_ktab_elems() { return (size_t)KTILES * KTILEF; }
size_t fin32_phi1_elems() { return (size_t)FTS * FTW * FTILEF; }
size_t fin32_phi2_elems() { return (size_t)FTS * FTV * FTILEF; }
size_t panel32_bytes() { return sizeof(PanelF); }
size_t finfreq32_bytes() { return sizeof(FinFreq32); }
double near_ratio() { return NEAR_RATIO; }

// The truncation of the FP64 Phi tiles of a frequency to the FP32 ones (dP: the device copy of
// truncation_matrix(FD, FDF)), and the FP32 copies of the panels' factors.
std::vector<double> fin32_truncation() { return truncation_matrix(FD, FDF); }
void launch_fin_tables32(const double* phi1, const double* phi2, const double* dP, float* phi1_32, float* phi2_32,
                         const double4* pf, const double2* pcs, int N, float4* pf32, float2* pcs32,
                         hipStream_t stream) {
  const size_t n1 = fin32_phi1_elems(), n2 = fin32_phi2_elems();
  truncate2d_kernel<<<(unsigned)((n1 + 255) / 256), 256, 0, stream>>>(phi1, FTILE, 3, FM, phi1_32, FTILEF, MFF, dP,
                                                                       FTS * FTW);
  truncate2d_kernel<<<(unsigned)((n2 + 255) / 256), 256, 0, stream>>>(phi2, FTILE, 3, FM, phi2_32, FTILEF, MFF, dP,
                                                                       FTS * FTV);
  const size_t np = (size_t)N * FIN_ROOTS;
  factors32_kernel<<<(unsigned)((np + 255) / 256), 256, 0, stream>>>(pf, pcs, N, pf32, pcs32);
  HBG_HIP(hipGetLastError());
}

void launch_panels32(const PanelD* P, int N, void* out, hipStream_t stream) {
  panels32_kernel<<<(N + 255) / 256, 256, 0, stream>>>(P, N, (PanelF*)out);
  HBG_HIP(hipGetLastError());
}

void launch_assemble32(const void* P32, int N, double K, double logc, const float* t2d32, const float* t1d32,
                       const float4* st32, float2* S, float2* A, float2* A2, int lda, double free_term,
                       hipStream_t stream) {
  const unsigned nt = (unsigned)((N + ASM_TILE - 1) / ASM_TILE);
  WaveTables32 T{t2d32, t1d32};
  assemble32_kernel<<<dim3(nt, nt, 1), 256, 0, stream>>>((const PanelF*)P32, N, (float)K, (float)logc, T, st32, S, A,
                                                         A2, lda, (float)free_term);
  HBG_HIP(hipGetLastError());
}

void launch_assemble_fin32(const void* P32, int N, const void* f32, const float* phi1_32, const float* phi2_32,
                           const float* ktab32, const float4* pf32, const float2* pcs32, const float* t2d32,
                           const float* t1d32, double logc, const float4* st32, float2* S, float2* A, float2* A2,
                           int lda, double free_term, hipStream_t stream) {
  const unsigned nt = (unsigned)((N + ASM_TILE - 1) / ASM_TILE);
  WaveTables32 T{t2d32, t1d32};
  FinTables32 F{phi1_32, phi2_32, ktab32, pf32, pcs32};
  assemble_fin32_kernel<<<dim3(nt, nt, 1), 256, 0, stream>>>((const PanelF*)P32, N, (const FinFreq32*)f32, F, T,
                                                             (float)logc, st32, S, A, A2, lda, (float)free_term);
  HBG_HIP(hipGetLastError());
}

void launch_near64(const PanelD* P, int N, const int2* ent, int nnz, bool finite, double K, const FinFreq* f,
                   FinTables F, WaveTables T, double logc, const StaticD* st, float2* S, float2* A, float2* A2,
                   int lda, double free_term, double2* dS, double2* dA, hipStream_t stream) {
  if (nnz <= 0) return;
  near64_kernel<<<(nnz + 255) / 256, 256, 0, stream>>>(P, N, ent, nnz, finite ? 1 : 0, K, f, F, T, logc, st, S, A,
                                                       A2, lda, free_term, dS, dA);
  HBG_HIP(hipGetLastError());
}

void launch_near_rows(const int* rowptr, const int2* ent, const double2* d, size_t dstride, const double2* v,
                      size_t vstride, int N, int R, double sign, double2* out, size_t ostride, int ldo, int accumulate,
                      int nsys, hipStream_t stream) {
  near_rows_kernel<<<dim3((N + 3) / 4, nsys), 256, 0, stream>>>(rowptr, ent, d, dstride, v, vstride, N, R, sign, out,
                                                               ostride, ldo, accumulate);
  HBG_HIP(hipGetLastError());
}

void launch_wave_points32(const float* t2d32, const float* t1d32, const double* X, const double* Y, int n,
                          double* out, hipStream_t stream) {
  WaveTables32 T{t2d32, t1d32};
  wave_points32_kernel<<<(n + 255) / 256, 256, 0, stream>>>(T, X, Y, n, out);
  HBG_HIP(hipGetLastError());
}

void launch_fin_points32(const void* f32, const float* phi1_32, const float* phi2_32, const float* ktab32,
                         const float4* pf32, const float2* pcs32, const float* t2d32, const float* t1d32,
                         const double* R, const double* z, const double* zeta, int n, double* out,
                         hipStream_t stream) {
  WaveTables32 T{t2d32, t1d32};
  FinTables32 F{phi1_32, phi2_32, ktab32, pf32, pcs32};
  fin_points32_kernel<<<(n + 255) / 256, 256, 0, stream>>>((const FinFreq32*)f32, F, T, R, z, zeta, n, out);
  HBG_HIP(hipGetLastError());
}

void* make_finfreq32(const FinFreq& f, void* host_out) {
  *(FinFreq32*)host_out = fin_freq32(f);
  return host_out;
}

}  // namespace hbg
