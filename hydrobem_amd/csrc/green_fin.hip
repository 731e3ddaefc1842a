// Per-frequency data of the finite-depth kernel, built on the GPU (green_fin.rs:874-952
// make_finite_depth / build_tiles / fin_point, restated without the reference's Q tables):
//   - the dispersion roots and series coefficients (host, dispersion.rs:19-95);
//   - the Phi1 and Phi2 tiles: at the Chebyshev nodes of every tile Q by Gauss-Legendre
//     quadrature of its principal-value integral along the path
//     0 -> i d -> T + i d -> T -> T + 24 (above the real poles k = x0 and k = H, below the first
//     imaginary pole i x1, x1 > pi/2; the real part of the path integral is the principal value
//     since the integrand is real on the real axis), factored over the nodes (fq_* kernels), plus
//     the Ft terms from the infinite-depth tables; one workgroup per tile fits the value and both
//     derivatives (monomial coefficients from the node values by one matrix, M F M^T);
//   - per panel the propagating-mode factors and cos, sin k_n (z + h).
// The K0, K1 tiles are frequency-independent (host, once per process).
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

// Q at the tile nodes, factored. The path and its points do not depend on the node, and along
// the path the integrand of Q (and of its two derivatives) is w(k) times a factor of xb (J0(xb k),
// J1(xb k)) times a factor of vb (cosh(vb k), sinh(vb k)), w(k) = (k + H)^2 e^{-3k} dk /
// [(k sinh k - H cosh k)(k - H)]. The nodes of a family of tiles are all pairs of its xb values
// (FTS tiles x FM nodes) and vb values (FTW or FTV tiles x FM nodes), so per frequency the factors
// are evaluated once per (point, xb) and (point, vb) and each node is three dot products along
// the path:
//   Q = Re sum_i [w chv]_i J0_i,  Q_xb = -Re sum_i [w chv k]_i J1_i,  Q_vb = Re sum_i [w k shv]_i J0_i,
// summed in the order of the points (segment by segment, Gauss point by point).
constexpr int FQ_NPAN = 64, FQ_NP = FQ_NPAN * 16;  // panels of 16 Gauss points along the path, at most
constexpr int FQ_NX = FTS * FM;                     // xb values
constexpr int FQ_NV1 = FTW * FM, FQ_NV = FQ_NV1 + FTV * FM;  // vb values of Phi1, then of Phi2
// vb rows per workgroup of the sums (one: 264 workgroups for 256 CUs; each sum runs in the same
// point order whatever the rows per workgroup), points per pass
constexpr int FQ_VB = 1, FQ_CHUNK = 256;

// Work layout (doubles): point count, points k_i and weights w_i, factors J [FQ_NP][FQ_NX][J0, J1],
// factors V [FQ_NV][FQ_NP][w chv, w chv k, w k shv]
constexpr size_t FQ_OFF_PTS = 2, FQ_OFF_J = FQ_OFF_PTS + 4 * (size_t)FQ_NP,
                 FQ_OFF_V = FQ_OFF_J + 4 * (size_t)FQ_NP * FQ_NX, FQ_WORK = FQ_OFF_V + 6 * (size_t)FQ_NV * FQ_NP;

__device__ __forceinline__ double xb_value(int x) {
  const int ix = x / FM, jx = x % FM;
  return (0.5 / FTS) * (ix + 0.5 * (cos(PI * (jx + 0.5) / FM) + 1.0));
}
__device__ __forceinline__ double vb_value(int v) {
  const bool one = v < FQ_NV1;
  const int l = one ? v : v - FQ_NV1, iy = l / FM, jy = l % FM;
  const double ymax = one ? 1.0 : 2.0;
  const int ny = one ? FTW : FTV;
  return (ymax / ny) * (iy + 0.5 * (cos(PI * (jy + 0.5) / FM) + 1.0));
}

// The path 0 -> i d -> T + i d -> T -> T + 24 (or the real axis to 28 when both poles lie beyond
// 28) in panels, each 16 Gauss points: one workgroup of FQ_NP threads, thread 16 p + g the point g
// of panel p; points past the last panel get weight zero.
__global__ void __launch_bounds__(FQ_NP) fq_points_kernel(double H, double x0, const double* __restrict__ g16,
                                                          double* __restrict__ work) {
  __shared__ cx pa[FQ_NPAN], pd[FQ_NPAN];
  __shared__ int npan;
  if (threadIdx.x == 0) {
    int np = 0;
    auto seg = [&](cx a, cx b, int panels) {
      const cx d = (1.0 / panels) * (b - a);
      for (int p = 0; p < panels && np < FQ_NPAN; ++p) {
        pa[np] = a + (double)p * d;
        pd[np] = d;
        ++np;
      }
    };
    if (fmin(x0, H) > 28.0) {
      seg(cx{0, 0}, cx{28.0, 0}, 56);
    } else {
      const double d = 0.7, T = fmax(x0, H) + 3.0;
      double y = fmin(0.05, 1e-3 * sqrt(H));
      seg(cx{0, 0}, cx{0, y}, 1);
      while (2.0 * y < d && np < FQ_NPAN) {
        seg(cx{0, y}, cx{0, 2.0 * y}, 1);
        y *= 2.0;
      }
      seg(cx{0, y}, cx{0, d}, 1);
      seg(cx{0, d}, cx{T, d}, (int)ceil(T) + 1);
      seg(cx{T, d}, cx{T, 0}, 1);
      seg(cx{T, 0}, cx{T + 24.0, 0}, 12);
    }
    npan = np;
    work[0] = 16.0 * np;
  }
  __syncthreads();
  const int p = threadIdx.x / 16, i = threadIdx.x % 16;
  cx k{1.0, 0.0}, w{0.0, 0.0};
  if (p < npan) {
    const cx hh = 0.5 * pd[p];
    k = pa[p] + (g16[i] + 1.0) * hh;
    const cx dk = g16[16 + i] * hh;
    cx sh, ch;
    csinhcosh(k, sh, ch);
    const cx kp = k + cx{H, 0.0}, km = k - cx{H, 0.0};
    const cx num = (kp * kp) * cexpx(-3.0 * k);
    const cx den = (k * sh - H * ch) * km;
    w = cdivx(num, den) * dk;
  }
  double* o = work + FQ_OFF_PTS + 4 * (size_t)threadIdx.x;
  o[0] = k.r;
  o[1] = k.i;
  o[2] = w.r;
  o[3] = w.i;
}

// J0(xb k_i), J1(xb k_i) for every point and xb value: thread = (point, xb)
__global__ void fq_jfactor_kernel(double* __restrict__ work) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = (int)work[0];
  if (e >= np * FQ_NX) return;
  const int i = e / FQ_NX, x = e % FQ_NX;
  const double* pt = work + FQ_OFF_PTS + 4 * (size_t)i;
  const cx k{pt[0], pt[1]};
  cx j0, j1;
  cj01x(xb_value(x) * k, j0, j1);
  double* o = work + FQ_OFF_J + 4 * (size_t)e;
  o[0] = j0.r;
  o[1] = j0.i;
  o[2] = j1.r;
  o[3] = j1.i;
}

// w chv, w chv k, w k shv for every vb value and point: thread = (vb, point)
__global__ void fq_vfactor_kernel(double* __restrict__ work) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = (int)work[0];
  if (e >= FQ_NV * FQ_NP) return;
  const int v = e / FQ_NP, i = e % FQ_NP;
  if (i >= np) return;
  const double* pt = work + FQ_OFF_PTS + 4 * (size_t)i;
  const cx k{pt[0], pt[1]}, w{pt[2], pt[3]};
  cx shv, chv;
  csinhcosh(vb_value(v) * k, shv, chv);
  const cx v0 = w * chv, v0k = v0 * k, v1 = (w * k) * shv;
  double* o = work + FQ_OFF_V + 6 * (size_t)e;
  o[0] = v0.r;
  o[1] = v0.i;
  o[2] = v0k.r;
  o[3] = v0k.i;
  o[4] = v1.r;
  o[5] = v1.i;
}

// The node values: Q by the dot products, plus the Ft terms of the infinite-depth tables;
// out [tile][q][jy][jx]. Workgroup: FQ_VB vb rows (their factors staged in LDS a chunk of points
// at a time), thread x the xb value.
__global__ void __launch_bounds__(128) fq_nodes_kernel(FinFreq f, WaveTables T, const double* __restrict__ work,
                                                       double* __restrict__ out1, double* __restrict__ out2) {
  __shared__ cx Vs[FQ_VB][FQ_CHUNK][3];
  const int v0 = blockIdx.x * FQ_VB, x = threadIdx.x;
  const int np = (int)work[0];
  cx s[FQ_VB][3];
#pragma unroll
  for (int r = 0; r < FQ_VB; ++r) s[r][0] = s[r][1] = s[r][2] = cx{0.0, 0.0};
  for (int c0 = 0; c0 < np; c0 += FQ_CHUNK) {
    const int nc = min(FQ_CHUNK, np - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < FQ_VB * nc; e += blockDim.x) {
      const int r = e / nc, i = e % nc;
      if (v0 + r < FQ_NV) {
        const double* src = work + FQ_OFF_V + 6 * ((size_t)(v0 + r) * FQ_NP + c0 + i);
        Vs[r][i][0] = cx{src[0], src[1]};
        Vs[r][i][1] = cx{src[2], src[3]};
        Vs[r][i][2] = cx{src[4], src[5]};
      }
    }
    __syncthreads();
    if (x < FQ_NX)
      for (int i = 0; i < nc; ++i) {
        const double* jp = work + FQ_OFF_J + 4 * ((size_t)(c0 + i) * FQ_NX + x);
        const cx j0{jp[0], jp[1]}, j1{jp[2], jp[3]};
#pragma unroll
        for (int r = 0; r < FQ_VB; ++r) {
          s[r][0] = s[r][0] + Vs[r][i][0] * j0;
          s[r][1] = s[r][1] - Vs[r][i][1] * j1;
          s[r][2] = s[r][2] + Vs[r][i][2] * j0;
        }
      }
  }
  if (x >= FQ_NX) return;
  const double H = f.H, H2 = H * H, xb = xb_value(x), X = H * xb;
  const int ix = x / FM, jx = x % FM;
#pragma unroll
  for (int r = 0; r < FQ_VB; ++r) {
    const int v = v0 + r;
    if (v >= FQ_NV) break;
    const bool one = v < FQ_NV1;
    const int l = one ? v : v - FQ_NV1, iy = l / FM, jy = l % FM, ny = one ? FTW : FTV;
    const double vb = vb_value(v);
    double q[3] = {s[r][0].r, s[r][1].r, s[r][2].r};
    if (one) {
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
    double* o = (one ? out1 : out2) + (size_t)(ix * ny + iy) * FTILE;
    const int node = jy * FM + jx;
    for (int k = 0; k < 3; ++k) o[k * FM * FM + node] = q[k];
  }
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

// K0(x) and K1(x) on the unit tiles of [1/2, 18): the trapezoidal rule of
// K_nu(x) = int_0^inf e^{-x cosh t} cosh(nu t) dt at the Chebyshev nodes, fitted to monomials
// (the tiles are narrow enough for the exponential decay: no scaling, no exp per root in the
// kernels)
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
      // K(x) = e^{-x} int e^{-x (cosh t - 1)} ... dt
      const long double ex = expl(-(long double)x);
      v0[j] = (double)(s0 * step * ex);
      v1[j] = (double)(s1 * step * ex);
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

size_t fin_tables_work() { return FQ_WORK; }

// work: fin_tables_work() doubles of device scratch
void launch_fin_tables(const FinFreq& f, WaveTables T, const double* g16, double* phi1, double* phi2,
                       const double* M1dev, const PanelD* P, int N, double4* pf, double2* pcs, double* work,
                       hipStream_t stream) {
  fq_points_kernel<<<1, FQ_NP, 0, stream>>>(f.H, f.x0, g16, work);
  fq_jfactor_kernel<<<(FQ_NP * FQ_NX + 255) / 256, 256, 0, stream>>>(work);
  fq_vfactor_kernel<<<(FQ_NV * FQ_NP + 255) / 256, 256, 0, stream>>>(work);
  fq_nodes_kernel<<<(FQ_NV + FQ_VB - 1) / FQ_VB, 128, 0, stream>>>(f, T, work, phi1, phi2);
  HBG_HIP(hipGetLastError());
  fin_fit_kernel<<<FTS * FTW * 3, 128, 0, stream>>>(phi1, FTS * FTW, M1dev);
  fin_fit_kernel<<<FTS * FTV * 3, 128, 0, stream>>>(phi2, FTS * FTV, M1dev);
  HBG_HIP(hipGetLastError());
  fin_points_kernel<<<(N + 255) / 256, 256, 0, stream>>>(f, P, N, pf, pcs);
  HBG_HIP(hipGetLastError());
}

}  // namespace hbg
