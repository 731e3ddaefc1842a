// The infinite-depth wave-kernel tables: Chebyshev fits of V, V_X on unit tiles of
// [0, XT) x [0, YT) and of J0, J1, Yt, Yt' on unit intervals of [0, XT), stored as monomial
// coefficients in the tile variable (the layout green_inf_dev.h reads).
//
// The functions are those of the reference's table generator (scripts/green_reference.py
// docstring; green_inf.rs:1-12):
//   Yt(X) = Y0(X) - (2/pi)(log(X/2) + gamma) J0(X)                  (entire)
//   V(X, Y) = e^-Y N(X, Y) / R0,  N = pi H0(X)/2 + int_0^Y (e^t - J0(X)) (X^2 + t^2)^-1/2 dt.
// N and dN/dX are computed by Gauss-Legendre quadrature after the substitution t = X sinh u,
// which removes the X-scale feature of the integrand at t = 0; H0 and H1 by their integral
// representations (A&S 12.1.7). The node values are evaluated on the GPU (one thread per node)
// when a device is present, otherwise on the host; the fits run on the host in long double.
#include <cmath>
#include <cstring>
#include <vector>

#include "hbg_internal.h"

namespace hbg {

namespace {

struct GL16 {
  double x[16], w[16];
};

// Gauss-Legendre nodes of order 16 (Newton iteration on P_16 from cos(pi (i + 3/4)/(n + 1/2))).
GL16 make_gl16() {
  GL16 g;
  const int n = 16;
  for (int i = 0; i < n / 2; ++i) {
    double z = std::cos(PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) <= 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    g.x[i] = -z;
    g.x[n - 1 - i] = z;
    g.w[i] = g.w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  return g;
}

__host__ __device__ inline void struve01_q(double x, const GL16& g, double& h0, double& h1) {
  const int panels = 8 + (int)(x / 3.0);
  const double a = 0.5 * PI / panels;
  double s0 = 0.0, s1 = 0.0;
  for (int q = 0; q < panels; ++q) {
    const double t0 = q * a;
    for (int i = 0; i < 16; ++i) {
      const double t = t0 + 0.5 * a * (g.x[i] + 1.0);
      const double w = 0.5 * a * g.w[i];
      const double sn = sin(x * cos(t));
      const double st = sin(t);
      s0 += w * sn;
      s1 += w * sn * st * st;
    }
  }
  h0 = 2.0 / PI * s0;
  h1 = 2.0 * x / PI * s1;
}

__host__ __device__ inline double one_minus_j0_q(double x) {
  if (x < 1.0) {
    const double q = 0.25 * x * x;
    double term = 1.0, s = 0.0;
    for (int k = 1; k < 30; ++k) {
      term *= -q / ((double)k * k);
      s -= term;
      if (fabs(term) < 1e-18 * fabs(s)) break;
    }
    return s;
  }
  return 1.0 - j0(x);
}

// (V, V_X) at X > 0, Y >= 0
__host__ __device__ inline void v_vx_q(double X, double Y, const GL16& g, double* out) {
  double h0, h1;
  struve01_q(X, g, h0, h1);
  const double j1v = j1(X);
  const double omj0 = one_minus_j0_q(X);
  double N = 0.5 * PI * h0;
  double NX = 1.0 - 0.5 * PI * h1;
  const double R0 = sqrt(X * X + Y * Y);
  if (Y > 0.0) {
    const double U = asinh(Y / X);
    double i1 = 0.0, i2 = 0.0, a = 0.0;
    while (a < U) {
      double w = fmin(1.0, U - a);
      while (w > 1e-3 && X * cosh(a + w) * w > 0.5) w *= 0.5;
      const double b = fmin(U, a + w);
      const double h = 0.5 * (b - a);
      for (int i = 0; i < 16; ++i) {
        const double u = a + h * (g.x[i] + 1.0);
        const double ch = cosh(u);
        const double e = expm1(X * sinh(u));
        i1 += h * g.w[i] * e;
        i2 += h * g.w[i] * e / (X * ch * ch);
      }
      a = b;
    }
    N += i1 + U * omj0;
    NX += j1v * U - (i2 + omj0 / X * (Y / R0));
  }
  const double eY = exp(-Y);
  out[0] = eY * N / R0;
  out[1] = eY * (NX / R0 - N * X / (R0 * R0 * R0));
}

// (J0, J1, Yt, Yt') at X > 0
__host__ __device__ inline void line_q(double X, double* out) {
  const double j0v = j0(X), j1v = j1(X);
  double yt, dyt;
  if (X <= 8.0) {
    const double q = 0.25 * X * X;
    double pk = 1.0, hk = 0.0, s = 0.0, ds = 0.0;
    for (int k = 1; k < 80; ++k) {
      pk *= q / ((double)k * k);
      hk += 1.0 / k;
      const double t = ((k & 1) ? 1.0 : -1.0) * hk * pk;
      s += t;
      ds += t * k;
      if (pk * hk < 1e-19 * fabs(s) && k > 4) break;
    }
    yt = 2.0 / PI * s;
    dyt = 2.0 / PI * ds / q * 0.5 * X;
  } else {
    const double y0v = y0(X), y1v = y1(X);
    const double L = log(0.5 * X) + GAMMA;
    yt = y0v - 2.0 / PI * L * j0v;
    dyt = -y1v - 2.0 / PI * (j0v / X - L * j1v);
  }
  out[0] = j0v;
  out[1] = j1v;
  out[2] = yt;
  out[3] = dyt;
}

__global__ void nodes2d_kernel(const double* __restrict__ xy, int n, GL16 g, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  v_vx_q(xy[2 * i], xy[2 * i + 1], g, out + 2 * i);
}

__global__ void nodes1d_kernel(const double* __restrict__ x, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  line_q(x[i], out + 4 * i);
}

// Chebyshev coefficients of the values at the M nodes cos(pi (i + 1/2)/M) -> monomial
// coefficients of the same polynomial in t (long double throughout).
void cheb_to_monomial(const long double* cheb, int M, long double* mono) {
  // T_k as monomials by T_{k+1} = 2 t T_k - T_{k-1}
  std::vector<long double> Tkm1(M, 0.0L), Tk(M, 0.0L), Tkp1(M, 0.0L);
  for (int a = 0; a < M; ++a) mono[a] = 0.0L;
  Tkm1[0] = 1.0L;  // T_0
  mono[0] += cheb[0];
  if (M == 1) return;
  Tk[1] = 1.0L;  // T_1
  for (int a = 0; a < M; ++a) mono[a] += cheb[1] * Tk[a];
  for (int k = 1; k + 1 < M; ++k) {
    for (int a = 0; a < M; ++a) Tkp1[a] = (a > 0 ? 2.0L * Tk[a - 1] : 0.0L) - Tkm1[a];
    for (int a = 0; a < M; ++a) mono[a] += cheb[k + 1] * Tkp1[a];
    Tkm1 = Tk;
    Tk = Tkp1;
  }
}

// values f[i] at nodes i -> Chebyshev coefficients c_k = (2/M) sum f_i T_k(t_i), c_0 halved
void fit1(const long double* f, int M, long double* c) {
  static thread_local std::vector<long double> cosm[32];
  std::vector<long double>& T = cosm[M];
  if (T.empty()) {
    T.resize((size_t)M * M);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < M; ++k)
      for (int i = 0; i < M; ++i) T[(size_t)k * M + i] = cosl(pi * (long double)k * ((long double)i + 0.5L) / M);
  }
  for (int k = 0; k < M; ++k) {
    long double s = 0.0L;
    for (int i = 0; i < M; ++i) s += f[i] * T[(size_t)k * M + i];
    c[k] = s * 2.0L / (long double)M * (k == 0 ? 0.5L : 1.0L);
  }
}

}  // namespace

void v_vx_node(double X, double Y, double* out2) {
  static const GL16 g = make_gl16();
  v_vx_q(X, Y, g, out2);
}

void line_node(double X, double* out4) { line_q(X, out4); }

void build_wave_tables(bool on_device, std::vector<double>& t2d, std::vector<double>& t1d) {
  const GL16 g = make_gl16();
  std::vector<double> tn2(M2), tn1(M1);
  for (int i = 0; i < M2; ++i) tn2[i] = std::cos(PI * (i + 0.5) / M2);
  for (int i = 0; i < M1; ++i) tn1[i] = std::cos(PI * (i + 0.5) / M1);
  // node coordinates: 2D [iy][ix][jy][jx], 1D [ix][j]
  const int n2 = XT * YT * M2 * M2, n1 = XT * M1;
  std::vector<double> xy(2 * (size_t)n2), x1(n1), v2(2 * (size_t)n2), v1(4 * (size_t)n1);
  for (int iy = 0; iy < YT; ++iy)
    for (int ix = 0; ix < XT; ++ix)
      for (int jy = 0; jy < M2; ++jy)
        for (int jx = 0; jx < M2; ++jx) {
          size_t o = (((size_t)iy * XT + ix) * M2 + jy) * M2 + jx;
          xy[2 * o] = ix + 0.5 * (tn2[jx] + 1.0);
          xy[2 * o + 1] = iy + 0.5 * (tn2[jy] + 1.0);
        }
  for (int ix = 0; ix < XT; ++ix)
    for (int j = 0; j < M1; ++j) x1[(size_t)ix * M1 + j] = ix + 0.5 * (tn1[j] + 1.0);
  if (on_device) {
    double *dxy, *dv2, *dx1, *dv1;
    HBG_HIP(hipMalloc(&dxy, xy.size() * sizeof(double)));
    HBG_HIP(hipMalloc(&dv2, v2.size() * sizeof(double)));
    HBG_HIP(hipMalloc(&dx1, x1.size() * sizeof(double)));
    HBG_HIP(hipMalloc(&dv1, v1.size() * sizeof(double)));
    HBG_HIP(hipMemcpy(dxy, xy.data(), xy.size() * sizeof(double), hipMemcpyHostToDevice));
    HBG_HIP(hipMemcpy(dx1, x1.data(), x1.size() * sizeof(double), hipMemcpyHostToDevice));
    nodes2d_kernel<<<(n2 + 127) / 128, 128>>>(dxy, n2, g, dv2);
    HBG_HIP(hipGetLastError());
    nodes1d_kernel<<<(n1 + 127) / 128, 128>>>(dx1, n1, dv1);
    HBG_HIP(hipGetLastError());
    HBG_HIP(hipMemcpy(v2.data(), dv2, v2.size() * sizeof(double), hipMemcpyDeviceToHost));
    HBG_HIP(hipMemcpy(v1.data(), dv1, v1.size() * sizeof(double), hipMemcpyDeviceToHost));
    HBG_HIP(hipFree(dxy));
    HBG_HIP(hipFree(dv2));
    HBG_HIP(hipFree(dx1));
    HBG_HIP(hipFree(dv1));
  } else {
    for (int i = 0; i < n2; ++i) v_vx_q(xy[2 * (size_t)i], xy[2 * (size_t)i + 1], g, &v2[2 * (size_t)i]);
    for (int i = 0; i < n1; ++i) line_q(x1[i], &v1[4 * (size_t)i]);
  }
  // 2D fits: per tile and function, Chebyshev in x for each node row, then in y; monomials.
  t2d.assign((size_t)XT * YT * TILE2, 0.0);
  for (int tile = 0; tile < XT * YT; ++tile) {
    for (int f = 0; f < 2; ++f) {
      long double vals[M2][M2], cx[M2][M2], cxy[M2][M2], mono[M2][M2], tmp[M2], out[M2];
      for (int jy = 0; jy < M2; ++jy)
        for (int jx = 0; jx < M2; ++jx) vals[jy][jx] = v2[2 * (((size_t)tile * M2 + jy) * M2 + jx) + f];
      for (int jy = 0; jy < M2; ++jy) {
        fit1(vals[jy], M2, tmp);
        cheb_to_monomial(tmp, M2, out);
        for (int a = 0; a < M2; ++a) cx[jy][a] = out[a];  // x monomial coefficient a at node row jy
      }
      for (int a = 0; a < M2; ++a) {
        long double col[M2];
        for (int jy = 0; jy < M2; ++jy) col[jy] = cx[jy][a];
        fit1(col, M2, tmp);
        cheb_to_monomial(tmp, M2, out);
        for (int b = 0; b < M2; ++b) cxy[b][a] = out[b];
      }
      for (int b = 0; b < M2; ++b)
        for (int a = 0; a < M2; ++a) mono[b][a] = cxy[b][a];
      double* dst = &t2d[(size_t)tile * TILE2 + (size_t)f * M2 * M2];
      for (int b = 0; b < M2; ++b)
        for (int a = 0; a < M2; ++a) dst[b * M2 + a] = (double)mono[b][a];
    }
  }
  t1d.assign((size_t)XT * TILE1, 0.0);
  for (int ix = 0; ix < XT; ++ix)
    for (int f = 0; f < 4; ++f) {
      long double vals[M1], tmp[M1], out[M1];
      for (int j = 0; j < M1; ++j) vals[j] = v1[4 * ((size_t)ix * M1 + j) + f];
      fit1(vals, M1, tmp);
      cheb_to_monomial(tmp, M1, out);
      for (int a = 0; a < M1; ++a) t1d[(size_t)ix * TILE1 + 4 * a + f] = (double)out[a];
    }
}

}  // namespace hbg
