// Per-frequency data of the finite-depth kernel, built on the GPU (green_fin.rs:874-952
// make_finite_depth / build_tiles / fin_point, restated without the reference's Q tables):
//   - the dispersion roots and series coefficients (host, dispersion.rs:19-95);
//   - the Phi1 and Phi2 tiles: at the Chebyshev nodes of every tile one thread evaluates Q by
//     Gauss-Legendre quadrature of its principal-value integral along the path
//     0 -> i d -> T + i d -> T -> T + 24 (above the real poles k = x0 and k = H, below the first
//     imaginary pole i x1, x1 > pi/2; the real part of the path integral is the principal value
//     since the integrand is real on the real axis), adds the Ft terms from the infinite-depth
//     tables, and one workgroup per tile fits the value and both derivatives (monomial
//     coefficients from the node values by one matrix, M F M^T);
//   - per panel the propagating-mode factors and cos, sin k_n (z + h).
// The scaled K0, K1 tiles are frequency-independent (host, once per process).
#include <cmath>
#include <vector>

#include "green_fin_dev.h"
#include "hbg_internal.h"

namespace hbg {

namespace {

struct cx {
  double r, i;
};
__device__ __forceinline__ cx operator+(cx a, cx b) { return cx{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cx operator-(cx a, cx b) { return cx{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cx operator*(cx a, cx b) { return cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cx operator*(double s, cx a) { return cx{s * a.r, s * a.i}; }
__device__ __forceinline__ cx cdivx(cx a, cx b) {
  const double d = b.r * b.r + b.i * b.i;
  return cx{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}
__device__ __forceinline__ cx cexpx(cx z) {
  double s, c;
  sincos(z.i, &s, &c);
  const double e = exp(z.r);
  return cx{e * c, e * s};
}
__device__ __forceinline__ void csinhcosh(cx z, cx& sh, cx& ch) {
  double s, c;
  sincos(z.i, &s, &c);
  const double a = sinh(z.r), b = cosh(z.r);
  sh = cx{a * c, b * s};
  ch = cx{b * c, a * s};
}
// J0, J1 of a complex argument by their power series (|z| up to about 25 here)
__device__ __forceinline__ void cj01x(cx z, cx& j0, cx& j1) {
  const cx q = -0.25 * (z * z);
  cx t0{1.0, 0.0}, t1{1.0, 0.0}, s0{1.0, 0.0}, s1{1.0, 0.0};
  for (int m = 1; m < 90; ++m) {
    t0 = (1.0 / ((double)m * m)) * (t0 * q);
    t1 = (1.0 / ((double)m * (m + 1))) * (t1 * q);
    s0 = s0 + t0;
    s1 = s1 + t1;
    if (m > 3 && fabs(t0.r) + fabs(t0.i) < 1e-18 * (fabs(s0.r) + fabs(s0.i)) &&
        fabs(t1.r) + fabs(t1.i) < 1e-18 * (fabs(s1.r) + fabs(s1.i)))
      break;
  }
  j0 = s0;
  j1 = (0.5 * z) * s1;
}

// Q, dQ/dxb, dQ/dvb by the path integral (see the file header)
__device__ void q_path(double H, double x0, double xb, double vb, const double* __restrict__ g16, double* q) {
  const double* gx = g16;
  const double* gw = g16 + 16;
  cx s0{0, 0}, s1{0, 0}, s2{0, 0};
  auto seg = [&](cx a, cx b, int panels) {
    const cx d = (1.0 / panels) * (b - a);
    for (int p = 0; p < panels; ++p) {
      const cx pa = a + (double)p * d;
      const cx hh = 0.5 * d;
      for (int i = 0; i < 16; ++i) {
        const cx k = pa + (gx[i] + 1.0) * hh;
        const cx dk = gw[i] * hh;
        cx sh, ch, shv, chv, j0, j1;
        csinhcosh(k, sh, ch);
        csinhcosh(vb * k, shv, chv);
        cj01x(xb * k, j0, j1);
        const cx kp = k + cx{H, 0.0}, km = k - cx{H, 0.0};
        const cx num = (kp * kp) * cexpx(-3.0 * k);
        const cx den = (k * sh - H * ch) * km;
        const cx wgt = cdivx(num, den) * dk;
        s0 = s0 + wgt * chv * j0;
        s1 = s1 - wgt * chv * k * j1;
        s2 = s2 + wgt * k * shv * j0;
      }
    }
  };
  if (fmin(x0, H) > 28.0) {
    seg(cx{0, 0}, cx{28.0, 0}, 56);
  } else {
    const double d = 0.7, T = fmax(x0, H) + 3.0;
    double y = fmin(0.05, 1e-3 * sqrt(H));
    seg(cx{0, 0}, cx{0, y}, 1);
    while (2.0 * y < d) {
      seg(cx{0, y}, cx{0, 2.0 * y}, 1);
      y *= 2.0;
    }
    seg(cx{0, y}, cx{0, d}, 1);
    seg(cx{0, d}, cx{T, d}, (int)ceil(T) + 1);
    seg(cx{T, d}, cx{T, 0}, 1);
    seg(cx{T, 0}, cx{T + 24.0, 0}, 12);
  }
  q[0] = s0.r;
  q[1] = s1.r;
  q[2] = s2.r;
}

// node values of the Phi tiles: node index -> (tile, jy, jx); out [tile][q][jy][jx]
__global__ void __launch_bounds__(64) fin_nodes_kernel(FinFreq f, WaveTables T, const double* __restrict__ g16,
                                                       double* __restrict__ out1, double* __restrict__ out2) {
  const int n1 = FTS * FTW * FM * FM, n2 = FTS * FTV * FM * FM;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n1 + n2) return;
  const int which = id < n1 ? 1 : 2;
  const int local = which == 1 ? id : id - n1;
  const int ny = which == 1 ? FTW : FTV;
  const double ymax = which == 1 ? 1.0 : 2.0;
  const int tile = local / (FM * FM), node = local % (FM * FM);
  const int jy = node / FM, jx = node % FM;
  const int ix = tile / ny, iy = tile % ny;
  const double tx = cos(PI * (jx + 0.5) / FM), ty = cos(PI * (jy + 0.5) / FM);
  const double xb = (0.5 / FTS) * (ix + 0.5 * (tx + 1.0));
  const double vb = (ymax / ny) * (iy + 0.5 * (ty + 1.0));
  double q[3];
  q_path(f.H, f.x0, xb, vb, g16, q);
  const double H = f.H, X = H * xb, H2 = H * H;
  if (which == 1) {
    const double Ya = H * (2.0 - vb), Yb = H * (2.0 + vb);
    const WaveK a = wave_kernel(T, X, Ya), b = wave_kernel(T, X, Yb);
    const double FYa = -a.F - 2.0 / sqrt(X * X + Ya * Ya), FYb = -b.F - 2.0 / sqrt(X * X + Yb * Yb);
    q[0] += H * (a.F + b.F);
    q[1] += H2 * (a.FX + b.FX);
    q[2] += H2 * (-FYa + FYb);
  } else {
    const double Yb = H * (2.0 + vb);
    const WaveK b = wave_kernel(T, X, Yb);
    const double FYb = -b.F - 2.0 / sqrt(X * X + Yb * Yb);
    q[0] += H * b.F;
    q[1] += H2 * b.FX;
    q[2] += H2 * FYb;
  }
  double* o = (which == 1 ? out1 : out2) + (size_t)tile * FTILE;
  for (int k = 0; k < 3; ++k) o[k * FM * FM + node] = q[k];
}

// monomial coefficients of every tile and value: C = M1 F M1^T, in place over the node values
__global__ void __launch_bounds__(128) fin_fit_kernel(double* __restrict__ tabs, int ntiles,
                                                      const double* __restrict__ M1) {
  __shared__ double F[FM][FM], W[FM][FM], Ms[FM][FM];
  const int item = blockIdx.x;  // tile * 3 + q
  if (item >= ntiles * 3) return;
  double* p = tabs + (size_t)item * FM * FM;
  for (int e = threadIdx.x; e < FM * FM; e += blockDim.x) {
    F[e / FM][e % FM] = p[e];
    Ms[e / FM][e % FM] = M1[e];
  }
  __syncthreads();
  // W = F M1^T: W[jy][a] = sum_jx F[jy][jx] M1[a][jx]
  for (int e = threadIdx.x; e < FM * FM; e += blockDim.x) {
    const int jy = e / FM, a = e % FM;
    double s = 0.0;
    for (int jx = 0; jx < FM; ++jx) s += F[jy][jx] * Ms[a][jx];
    W[jy][a] = s;
  }
  __syncthreads();
  // C[b][a] = sum_jy M1[b][jy] W[jy][a]
  for (int e = threadIdx.x; e < FM * FM; e += blockDim.x) {
    const int b = e / FM, a = e % FM;
    double s = 0.0;
    for (int jy = 0; jy < FM; ++jy) s += Ms[b][jy] * W[jy][a];
    p[e] = s;
  }
}

__global__ void fin_points_kernel(FinFreq f, const PanelD* __restrict__ P, int N, double4* __restrict__ pf,
                                  double2* __restrict__ pcs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const double z = P[j].c[2], h = f.h, k0 = f.k0;
  pf[j] = make_double4(exp(k0 * z), exp(-k0 * (z + h)), exp(-k0 * (h - z)), exp(-k0 * (z + 2.0 * h)));
  for (int m = 0; m < FIN_ROOTS; ++m) {
    double s, c;
    sincos(f.kn[m] * (z + h), &s, &c);
    pcs[(size_t)j * FIN_ROOTS + m] = make_double2(c, s);
  }
}

// x0 tanh x0 = H and xn tan xn = -H (dispersion.rs:19-95): safeguarded Newton on the brackets
// max(H, sqrt H) <= x0 <= (H + sqrt(H^2 + 4H))/2 and xn = n pi - d, 0 < d < pi/2
double root_x0(double H) {
  if (!(H > 0.0)) return 0.0;
  if (H >= 40.0) return H;
  double lo = std::max(H, std::sqrt(H)), hi = 0.5 * (H + std::sqrt(H * H + 4.0 * H));
  double x = std::min(std::max(H < 1.0 ? std::sqrt(H) * (1.0 + H / 6.0) : H, lo), hi), prev = -1.0;
  for (int it = 0; it < 200; ++it) {
    const double t = std::tanh(x), fv = x * t - H;
    (fv < 0.0 ? lo : hi) = x;
    double xn = x - fv / (t + x * (1.0 - t * t));
    if (!(xn >= lo && xn <= hi)) xn = 0.5 * (lo + hi);
    if (xn == x || xn == prev) break;
    prev = x;
    x = xn;
  }
  return x;
}

double root_xn(double H, int n) {
  const double npi = n * PI;
  if (!(H > 0.0)) return npi;
  double lo = 0.0, hi = 0.5 * PI, d = std::atan(H / (npi - 0.25 * PI)), prev = -1.0;
  for (int it = 0; it < 200; ++it) {
    const double s = std::sin(d), c = std::cos(d), g = (npi - d) * s - H * c;
    (g < 0.0 ? lo : hi) = d;
    double dn = d - g / ((npi - d) * c + (H - 1.0) * s);
    if (!(dn > lo && dn < hi)) dn = 0.5 * (lo + hi);
    if (dn == d || dn == prev) break;
    prev = d;
    d = dn;
  }
  return npi - d;
}

}  // namespace

FinFreq make_fin_freq(double h, double K) {
  FinFreq f;
  f.h = h;
  f.K = K;
  f.H = K * h;
  f.x0 = root_x0(f.H);
  f.k0 = f.x0 / h;
  f.A0 = 2.0 * PI * f.k0 / (-std::expm1(-4.0 * f.x0) + 4.0 * f.x0 * std::exp(-2.0 * f.x0));
  for (int n = 1; n <= FIN_ROOTS; ++n) {
    const double k = root_xn(f.H, n) / h;
    f.kn[n - 1] = k;
    f.cn[n - 1] = 4.0 * (k * k + K * K) / (k * k * h + K * K * h - K);
  }
  return f;
}

double dispersion_root(double H, int n) { return n == 0 ? root_x0(H) : root_xn(H, n); }

// node values -> monomial coefficients in one dimension (long double): M1 = A^T Tm
std::vector<double> fit_matrix(int M) {
  std::vector<long double> Tm((size_t)M * M), A((size_t)M * M, 0.0L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < M; ++k)
    for (int j = 0; j < M; ++j)
      Tm[(size_t)k * M + j] = cosl(pi * k * (j + 0.5L) / M) * 2.0L / M * (k == 0 ? 0.5L : 1.0L);
  // A[k][a]: coefficient of t^a in T_k
  A[0] = 1.0L;
  if (M > 1) A[(size_t)1 * M + 1] = 1.0L;
  for (int k = 2; k < M; ++k)
    for (int a = 0; a < M; ++a)
      A[(size_t)k * M + a] = (a > 0 ? 2.0L * A[(size_t)(k - 1) * M + a - 1] : 0.0L) - A[(size_t)(k - 2) * M + a];
  std::vector<double> M1((size_t)M * M);
  for (int a = 0; a < M; ++a)
    for (int j = 0; j < M; ++j) {
      long double s = 0.0L;
      for (int k = 0; k < M; ++k) s += A[(size_t)k * M + a] * Tm[(size_t)k * M + j];
      M1[(size_t)a * M + j] = (double)s;
    }
  return M1;
}

// e^x sqrt(x) K0(x) and K1(x) on the unit tiles of [1/2, 18): the trapezoidal rule of
// K_nu(x) = int_0^inf e^{-x cosh t} cosh(nu t) dt at the Chebyshev nodes, fitted to monomials
std::vector<double> build_k_table() {
  const int M = KD + 1;
  std::vector<double> tab((size_t)KTILES * M * 2);
  const std::vector<double> M1 = fit_matrix(M);
  for (int i = 0; i < KTILES; ++i) {
    const double a = i == 0 ? 0.5 : (double)i, b = i == 0 ? 1.0 : (double)(i + 1);
    std::vector<double> v0(M), v1(M);
    for (int j = 0; j < M; ++j) {
      const double t = std::cos(PI * (j + 0.5) / M), x = a + 0.5 * (b - a) * (t + 1.0);
      long double s0 = 0.5L, s1 = 0.5L;  // e^{-x (cosh t - 1)} at t = 0, halved
      const double step = 0.02;
      for (int q = 1; q < 100000; ++q) {
        const double tt = q * step, c = std::cosh(tt);
        const long double e = expl(-(long double)x * (c - 1.0L));
        s0 += e;
        s1 += e * c;
        if (e * c < 1e-22L) break;
      }
      // e^x sqrt(x) K(x) = sqrt(x) int e^{-x (cosh t - 1)} ... dt
      v0[j] = (double)(s0 * step * sqrtl((long double)x));
      v1[j] = (double)(s1 * step * sqrtl((long double)x));
    }
    for (int p = 0; p < M; ++p) {
      long double c0 = 0.0L, c1 = 0.0L;
      for (int j = 0; j < M; ++j) {
        c0 += (long double)M1[(size_t)p * M + j] * v0[j];
        c1 += (long double)M1[(size_t)p * M + j] * v1[j];
      }
      tab[((size_t)i * M + p) * 2] = (double)c0;
      tab[((size_t)i * M + p) * 2 + 1] = (double)c1;
    }
  }
  return tab;
}

void launch_fin_tables(const FinFreq& f, WaveTables T, const double* g16, double* phi1, double* phi2,
                       const double* M1dev, const PanelD* P, int N, double4* pf, double2* pcs, hipStream_t stream) {
  const int n = (FTS * FTW + FTS * FTV) * FM * FM;
  fin_nodes_kernel<<<(n + 63) / 64, 64, 0, stream>>>(f, T, g16, phi1, phi2);
  HBG_HIP(hipGetLastError());
  fin_fit_kernel<<<FTS * FTW * 3, 128, 0, stream>>>(phi1, FTS * FTW, M1dev);
  fin_fit_kernel<<<FTS * FTV * 3, 128, 0, stream>>>(phi2, FTS * FTV, M1dev);
  HBG_HIP(hipGetLastError());
  fin_points_kernel<<<(N + 255) / 256, 256, 0, stream>>>(f, P, N, pf, pcs);
  HBG_HIP(hipGetLastError());
}

}  // namespace hbg
