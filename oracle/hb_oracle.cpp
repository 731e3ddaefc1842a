// FP64 CPU oracle of HydroBEM's low-order frequency-domain solve (test infrastructure only).
//
// This file is the checker of the HIP library in hydrobem_amd/: it restates, in plain C++ and
// double precision, the algorithm of the reference's diagnostic double mode (HB_DIAG_DOUBLE):
//   - panel rule            crates/hydrobem/src/geometry.rs:120-199            (make_panel)
//   - panel geometry        crates/hydrobem/src/panel_integrals.rs:145-217     (make_panel_geom)
//   - Gauss-Legendre        crates/hydrobem/src/quadrature.rs:17-59
//   - Rankine closed forms  crates/hydrobem/src/panel_integrals.rs:226-275     (diag: to 1000 radii)
//   - log singularity       crates/hydrobem/src/panel_integrals.rs:380-450, Adaptive :69-140
//   - static rows           crates/hydrobem/src/lo_static.rs:47-119            (infinite depth)
//   - wave kernel           crates/hydrobem/src/green_inf.rs:316-368           (identity form)
//   - pair kernel / apply   crates/hydrobem/src/low_order.rs:627-648, 717-769
//   - assembly              crates/hydrobem/src/lo_frequency.rs:399-537        (G = 1, wet panels)
//   - boundary values       crates/hydrobem/src/lo_frequency.rs:214-273, low_order.rs:1401-1485
//   - RHS product + LU      crates/hydrobem/src/lo_frequency.rs:823-955, linalg.rs:134-252
//   - forces, Haskind       crates/hydrobem/src/lo_frequency.rs:1039-1171
//
// Where the reference reads generated Chebyshev tables (V, V_X, Yt, Yt' of green_inf.rs), this
// oracle evaluates the same functions directly from their definitions (scripts/green_reference.py
// docstring): Yt by its power series or libm's y0/j0, V and V_X by Gauss-Legendre quadrature of
// the defining integral after the substitution t = X sinh u, H0 and H1 by their integral
// representations. It is therefore more accurate than the tables (2e-10) and independent of any
// table layout. Everything is double precision; nothing here is shipped or called by the product.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "hb_oracle.h"

typedef std::complex<double> cd;

namespace {

const double PI = 3.14159265358979323846;
const double GAMMA = 0.5772156649015329;  // constants.rs:4
const double LOG2 = 0.69314718055994530942;

struct V3 {
  double x[3];
  double& operator[](int i) { return x[i]; }
  double operator[](int i) const { return x[i]; }
};
inline V3 v3(double a, double b, double c) { V3 r; r.x[0] = a; r.x[1] = b; r.x[2] = c; return r; }
inline V3 operator+(V3 a, V3 b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 operator-(V3 a, V3 b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 operator*(double s, V3 a) { return v3(s * a[0], s * a[1], s * a[2]); }
inline double dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(V3 a, V3 b) {
  return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// ------------------------------------------------------------------ quadrature.rs:17-59
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  double nf = n;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(PI * (i + 0.75) / (nf + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = nf * (z * p1 - p0) / (z * z - 1.0);
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
    dp = nf * (z * p1 - p0) / (z * z - 1.0);
    x[i] = -z;
    x[n - 1 - i] = z;
    double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

struct GaussCache {
  std::vector<double> x[17], w[17];
  GaussCache() {
    for (int n = 1; n <= 16; ++n) gauss_legendre(n, x[n], w[n]);
  }
};
const GaussCache& gauss() {
  static GaussCache c;
  return c;
}

// ------------------------------------------------------------------ geometry.rs:120-199
const double COINCIDENCE_TOL = 1e-6;
const double CONVEXITY_TOL = 1e-6;

struct Panel {
  V3 raw[4], proj[4], poly[4], normal, centroid;
  int distinct = 0;
  double area = 0, diameter = 0;
};

int make_panel(const double* v12, Panel& out) {
  for (int i = 0; i < 12; ++i)
    if (!std::isfinite(v12[i])) return ORC_ERR_VALUE;
  for (int i = 0; i < 4; ++i) out.raw[i] = v3(v12[3 * i], v12[3 * i + 1], v12[3 * i + 2]);
  double diameter = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) diameter = std::max(diameter, norm(out.raw[i] - out.raw[j]));
  out.diameter = diameter;
  if (diameter == 0.0) return ORC_ERR_VALUE;
  double tol = COINCIDENCE_TOL * diameter;
  int keep[4], m = 0;
  for (int i = 0; i < 4; ++i)
    if (norm(out.raw[i] - out.raw[(i + 1) % 4]) > tol) keep[m++] = i;
  if (m < 3) return ORC_ERR_VALUE;
  for (int a = 0; a < m; ++a)
    for (int b = a + 1; b < m; ++b)
      if (norm(out.raw[keep[a]] - out.raw[keep[b]]) <= tol) return ORC_ERR_VALUE;
  V3 c = 0.25 * (out.raw[0] + out.raw[1] + out.raw[2] + out.raw[3]);
  V3 n = cross(out.raw[2] - out.raw[0], out.raw[3] - out.raw[1]);
  double nn = norm(n);
  if (0.5 * nn <= COINCIDENCE_TOL * diameter * diameter) return ORC_ERR_VALUE;
  V3 nh = (1.0 / nn) * n;
  for (int i = 0; i < 4; ++i) out.proj[i] = out.raw[i] - dot(out.raw[i] - c, nh) * nh;
  out.distinct = m;
  for (int a = 0; a < m; ++a) out.poly[a] = out.proj[keep[a]];
  for (int a = 0; a < m; ++a) {
    V3 e1 = out.poly[(a + 1) % m] - out.poly[a];
    V3 e2 = out.poly[(a + 2) % m] - out.poly[(a + 1) % m];
    double s = dot(cross(e1, e2), nh) / (norm(e1) * norm(e2));
    if (!(s > CONVEXITY_TOL)) return ORC_ERR_VALUE;
  }
  double area = 0.0;
  V3 moment = v3(0, 0, 0);
  for (int a = 1; a < m - 1; ++a) {
    double t = 0.5 * dot(cross(out.poly[a] - out.poly[0], out.poly[a + 1] - out.poly[0]), nh);
    area += t;
    moment = moment + (t / 3.0) * (out.poly[0] + out.poly[a] + out.poly[a + 1]);
  }
  out.area = area;
  out.centroid = (1.0 / area) * moment;
  out.normal = nh;
  return ORC_OK;
}

// ------------------------------------------------------------------ panel_integrals.rs:22-217
struct PanelGeom {
  int nv = 0;
  V3 v[4], c, n;
  double area = 0, radius = 0;
  double elen[4];
  V3 em[4];
  V3 gx[9];
  double gw[9];
  V3 q4[4];  // the four (projected) corners of the bilinear map of the 3 x 3 rule
};

// patch_geometry.rs:14-22
void bilinear(const V3* vt, double u, double v, V3& x, double& jac) {
  double n0 = 0.25 * (1.0 - u) * (1.0 - v), n1 = 0.25 * (1.0 + u) * (1.0 - v);
  double n2 = 0.25 * (1.0 + u) * (1.0 + v), n3 = 0.25 * (1.0 - u) * (1.0 + v);
  x = n0 * vt[0] + n1 * vt[1] + n2 * vt[2] + n3 * vt[3];
  V3 xu = (0.25 * (1.0 - v)) * (vt[1] - vt[0]) + (0.25 * (1.0 + v)) * (vt[2] - vt[3]);
  V3 xv = (0.25 * (1.0 - u)) * (vt[3] - vt[0]) + (0.25 * (1.0 + u)) * (vt[2] - vt[1]);
  jac = norm(cross(xu, xv));
}

PanelGeom make_panel_geom(const V3* q, V3 normal, V3 centroid, double area) {
  PanelGeom g;
  g.n = normal;
  g.c = centroid;
  g.area = area;
  double diam = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) diam = std::max(diam, norm(q[i] - q[j]));
  for (int i = 0; i < 4; ++i)
    if (norm(q[i] - q[(i + 1) % 4]) > COINCIDENCE_TOL * diam) g.v[g.nv++] = q[i];
  for (int i = 0; i < g.nv; ++i) g.radius = std::max(g.radius, norm(g.v[i] - centroid));
  for (int e = 0; e < g.nv; ++e) {
    V3 edge = g.v[(e + 1) % g.nv] - g.v[e];
    g.elen[e] = norm(edge);
    g.em[e] = (1.0 / g.elen[e]) * cross(edge, normal);
  }
  for (int i = 0; i < 4; ++i) g.q4[i] = q[i];
  const GaussCache& G = gauss();
  double sum = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double jac;
      bilinear(q, G.x[3][i], G.x[3][j], g.gx[3 * i + j], jac);
      g.gw[3 * i + j] = G.w[3][i] * G.w[3][j] * jac;
      sum += g.gw[3 * i + j];
    }
  for (int k = 0; k < 9; ++k) g.gw[k] *= area / sum;
  return g;
}

// the bilinear map of the distinct vertices (panel_integrals.rs:52-56)
void panel_map(const PanelGeom& p, V3* m) {
  m[0] = p.v[0];
  m[1] = p.v[1];
  m[2] = p.v[2];
  m[3] = p.nv == 4 ? p.v[3] : p.v[2];
}

// panel_integrals.rs:226-275 with the diagnostic-mode far switch at 1000 radii (:231)
void rankine(const PanelGeom& p, V3 x, double& S, double& D) {
  V3 rc = x - p.c;
  double dist = norm(rc);
  if (dist > 1000.0 * p.radius) {
    S = 0.0;
    D = 0.0;
    for (int k = 0; k < 9; ++k) {
      V3 r = x - p.gx[k];
      double ir = 1.0 / norm(r);
      S += p.gw[k] * ir;
      D += p.gw[k] * dot(p.n, r) * ir * ir * ir;
    }
    return;
  }
  double w = dot(rc, p.n);
  int nv = p.nv;
  V3 av[4];
  double r[4];
  for (int i = 0; i < nv; ++i) {
    av[i] = p.v[i] - x;
    r[i] = norm(av[i]);
  }
  double omega = 0.0;
  if (std::fabs(w) > 1e-12 * p.radius) {
    for (int i = 1; i < nv - 1; ++i) {
      V3 a = av[0], b = av[i], c = av[i + 1];
      double num = dot(a, cross(b, c));
      double den = r[0] * r[i] * r[i + 1] + dot(a, b) * r[i + 1] + dot(a, c) * r[i] + dot(b, c) * r[0];
      omega -= 2.0 * std::atan2(num, den);
    }
  }
  double s = 0.0;
  for (int e = 0; e < nv; ++e) {
    int f = (e + 1) % nv;
    double d = dot(av[e], p.em[e]);
    if (std::fabs(d) <= 1e-14 * p.radius) continue;
    double q = r[e] * r[f] + dot(av[e], av[f]);
    if (!(q > 0.0)) continue;
    double sm = r[e] + r[f] + p.elen[e];
    s += d * std::log(sm * sm / (2.0 * q));
  }
  S = s - w * omega;
  D = omega;
}

// panel_integrals.rs:281-286
inline void psi_terms(V3 xi, V3 p, V3 n, double& psi, double& dpsi) {
  V3 d = xi - p;
  double R = norm(d);
  double den = R - d[2];
  psi = std::log(den);
  dpsi = (dot(d, n) / R - n[2]) / den;
}

// panel_integrals.rs:69-140: adaptive quadrisection of the 4 x 4 Gauss rule
struct AdaptiveLog {
  const V3* map;
  V3 image, n;
  double tol_l, tol_d;
  void cell(double u0, double u1, double v0, double v1, double* s) const {
    const GaussCache& G = gauss();
    s[0] = s[1] = 0.0;
    double hu = 0.5 * (u1 - u0), hv = 0.5 * (v1 - v0);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        V3 xi;
        double jac;
        bilinear(map, u0 + hu * (G.x[4][i] + 1.0), v0 + hv * (G.x[4][j] + 1.0), xi, jac);
        double ww = G.w[4][i] * G.w[4][j] * hu * hv * jac;
        double psi, dpsi;
        psi_terms(xi, image, n, psi, dpsi);
        s[0] += ww * psi;
        s[1] += ww * dpsi;
      }
  }
  void adapt(double u0, double u1, double v0, double v1, const double* prev, int depth, double* acc) const {
    double um = 0.5 * (u0 + u1), vm = 0.5 * (v0 + v1);
    double c[4][2];
    cell(u0, um, v0, vm, c[0]);
    cell(um, u1, v0, vm, c[1]);
    cell(um, u1, vm, v1, c[2]);
    cell(u0, um, vm, v1, c[3]);
    double sum[2];
    for (int k = 0; k < 2; ++k) sum[k] = c[0][k] + c[1][k] + c[2][k] + c[3][k];
    bool done = std::fabs(sum[0] - prev[0]) <= tol_l && std::fabs(sum[1] - prev[1]) <= tol_d;
    if (depth >= 14 || done) {
      acc[0] += sum[0];
      acc[1] += sum[1];
      return;
    }
    const double us[4][2] = {{u0, um}, {um, u1}, {um, u1}, {u0, um}};
    const double vs[4][2] = {{v0, vm}, {v0, vm}, {vm, v1}, {vm, v1}};
    for (int q = 0; q < 4; ++q) adapt(us[q][0], us[q][1], vs[q][0], vs[q][1], c[q], depth + 1, acc);
  }
};

// panel_integrals.rs:380-450 (diagnostic mode: no 3 x 3 far rule, :382)
void log_singularity(const PanelGeom& p, V3 image, double& L, double& Ld) {
  double dist = norm(image - p.c);
  V3 map[4];
  panel_map(p, map);
  double top = p.v[0][2];
  for (int i = 1; i < p.nv; ++i) top = std::max(top, p.v[i][2]);
  double r = dist / p.radius;
  double clearance = (image[2] - top) / p.radius;
  int order = 0;
  if (clearance >= 0.5 || r >= 2.0) {
    order = r >= 6.0 ? 4 : r >= 3.0 ? 5 : r >= 2.0 ? 6 : r >= 1.5 ? 8 : r >= 1.0 ? 10 : 0;
  }
  const GaussCache& G = gauss();
  if (order > 0) {
    L = 0.0;
    Ld = 0.0;
    for (int i = 0; i < order; ++i)
      for (int j = 0; j < order; ++j) {
        V3 xi;
        double jac;
        bilinear(map, G.x[order][i], G.x[order][j], xi, jac);
        double psi, dpsi;
        psi_terms(xi, image, p.n, psi, dpsi);
        double ww = G.w[order][i] * G.w[order][j] * jac;
        L += ww * psi;
        Ld += ww * dpsi;
      }
    return;
  }
  double dmin = std::max(dist - p.radius, 1e-3 * p.radius);
  double scale = std::fabs(std::log(dist + p.radius)) + 1.0;
  AdaptiveLog a;
  a.map = map;
  a.image = image;
  a.n = p.n;
  a.tol_l = 1e-13 * p.area * scale;
  a.tol_d = 1e-13 * p.area / dmin;
  double whole[2], acc[2] = {0.0, 0.0};
  a.cell(-1.0, 1.0, -1.0, 1.0, whole);
  a.adapt(-1.0, 1.0, -1.0, 1.0, whole, 0, acc);
  L = acc[0];
  Ld = acc[1];
}

// lo_static.rs:47-55
bool log_pair(const PanelGeom& p, V3 x, double tol) {
  double dx = x[0] - p.c[0], dy = x[1] - p.c[1];
  double zz = std::fabs(x[2] + p.c[2]);
  double d = std::sqrt(dx * dx + dy * dy + zz * zz) + zz;
  return d * d * tol < 0.6 * p.radius * p.radius;
}

// lo_static.rs:59-119, wetted panels (no surface panels in z = 0); finite depth h adds the bottom
// images of image_points (:29-38): (x, y, -z - 2h) and its image (x, y, z + 2h), (x, y, z - 2h),
// (x, y, -z - 4h)
// inf (optional): the omega = infinity row (:84-95), the integrals of 1/r1 + 1/r2 - (1/r3 + ... +
// 1/r6) in finite depth, of 1/r - 1/r' in infinite depth
void static_entry(const PanelGeom& p, V3 x, bool logint, double tol, double depth, double* e,
                  double* inf = nullptr) {
  double s0, d0, sb, db;
  rankine(p, x, s0, d0);
  rankine(p, v3(x[0], x[1], -x[2]), sb, db);
  e[0] = s0 + sb;
  e[1] = d0 + db;
  if (inf) {
    inf[0] = s0 - sb;
    inf[1] = d0 - db;
  }
  if (std::isfinite(depth)) {
    const double h = depth;
    double s2, d2, s4, d4, s5, d5, s6, d6;
    rankine(p, v3(x[0], x[1], -x[2] - 2.0 * h), s2, d2);
    rankine(p, v3(x[0], x[1], x[2] + 2.0 * h), s4, d4);
    rankine(p, v3(x[0], x[1], x[2] - 2.0 * h), s5, d5);
    rankine(p, v3(x[0], x[1], -x[2] - 4.0 * h), s6, d6);
    e[0] += (s2 + s4) + s5 + s6;
    e[1] += (d2 + d4) + d5 + d6;
    if (inf) {
      inf[0] = s0 + s2 - sb - s4 - s5 - s6;
      inf[1] = d0 + d2 - db - d4 - d5 - d6;
    }
  }
  e[2] = 0.0;
  e[3] = 0.0;
  if (logint && log_pair(p, x, tol)) {
    V3 xb = v3(x[0], x[1], -x[2]);
    double L, Ld, psi, dpsi;
    log_singularity(p, xb, L, Ld);
    psi_terms(p.c, xb, p.n, psi, dpsi);
    e[2] = L - p.area * psi;
    e[3] = Ld - p.area * dpsi;
  }
}

// ------------------------------------------------------------------ special functions
// Struve H0, H1 (A&S 12.1.7): H0 = (2/pi) int_0^{pi/2} sin(x cos t) dt,
// H1 = (2x/pi) int_0^{pi/2} sin(x cos t) sin^2 t dt; composite Gauss-Legendre, exact to rounding.
void struve01(double x, double& h0, double& h1) {
  const GaussCache& G = gauss();
  int panels = 8 + (int)(x / 3.0);
  double a = 0.5 * PI / panels;
  double s0 = 0.0, s1 = 0.0;
  for (int q = 0; q < panels; ++q) {
    double t0 = q * a;
    for (int i = 0; i < 16; ++i) {
      double t = t0 + 0.5 * a * (G.x[16][i] + 1.0);
      double w = 0.5 * a * G.w[16][i];
      double sn = std::sin(x * std::cos(t));
      double st = std::sin(t);
      s0 += w * sn;
      s1 += w * sn * st * st;
    }
  }
  h0 = 2.0 / PI * s0;
  h1 = 2.0 * x / PI * s1;
}

// 1 - J0(x), without cancellation for small x
double one_minus_j0(double x) {
  if (x < 1.0) {
    double q = 0.25 * x * x, term = 1.0, s = 0.0;
    for (int k = 1; k < 30; ++k) {
      term *= -q / ((double)k * k);
      s -= term;
      if (std::fabs(term) < 1e-18 * std::fabs(s)) break;
    }
    return s;
  }
  return 1.0 - ::j0(x);
}

// Yt = Y0 - (2/pi)(log(x/2) + gamma) J0 and its derivative (green_reference.py Yt)
void yt_and_derivative(double x, double& yt, double& dyt) {
  if (x == 0.0) {
    yt = 0.0;
    dyt = 0.0;
    return;
  }
  if (x <= 8.0) {
    // (2/pi) sum_{k>=1} (-1)^{k+1} H_k q^k / (k!)^2, q = x^2/4
    double q = 0.25 * x * x, pk = 1.0, hk = 0.0, s = 0.0, ds = 0.0;
    for (int k = 1; k < 80; ++k) {
      pk *= q / ((double)k * k);  // q^k / (k!)^2
      hk += 1.0 / k;
      double sg = (k % 2) ? 1.0 : -1.0;
      double t = sg * hk * pk;
      s += t;
      ds += t * k;  // d/dq q^k = k q^{k-1}; times dq/dx = x/2 below
      if (pk * hk < 1e-19 * std::fabs(s) && k > 4) break;
    }
    yt = 2.0 / PI * s;
    dyt = 2.0 / PI * ds / q * 0.5 * x;
    return;
  }
  double j0v = ::j0(x), j1v = ::j1(x), y0v = ::y0(x), y1v = ::y1(x);
  double L = std::log(0.5 * x) + GAMMA;
  yt = y0v - 2.0 / PI * L * j0v;
  dyt = -y1v - 2.0 / PI * (j0v / x - L * j1v);
}

// Ein(y) = int_0^y (e^t - 1)/t dt = sum y^k / (k k!) (the series below 40; Ei(y) - log y - gamma above)
double ein(double y) {
  if (y > 40.0) return std::expint(y) - std::log(y) - GAMMA;
  double term = 1.0, s = 0.0;
  for (int k = 1; k < 400; ++k) {
    term *= y / k;
    double t = term / k;
    s += t;
    if (t < 1e-18 * s) break;
  }
  return s;
}

// N(X, Y) = pi H0(X)/2 + int_0^Y (e^t - J0(X)) / sqrt(X^2 + t^2) dt and dN/dX, X > 0
// (green_reference.py V_VX: V = exp(-Y) N / R0), by Gauss-Legendre in u, t = X sinh u.
void n_and_nx(double X, double Y, double& N, double& NX) {
  double h0, h1;
  struve01(X, h0, h1);
  double j1v = ::j1(X);
  double omj0 = one_minus_j0(X);
  N = 0.5 * PI * h0;
  NX = 1.0 - 0.5 * PI * h1;  // (pi/2) H0' with H0' = 2/pi - H1
  if (Y <= 0.0) return;
  double U = std::asinh(Y / X);
  const GaussCache& G = gauss();
  double i1 = 0.0, i2 = 0.0;  // int expm1(X sinh u) du, int expm1(X sinh u) / (X cosh^2 u) du
  double a = 0.0;
  while (a < U) {
    double w = std::min(1.0, U - a);
    while (w > 1e-3 && X * std::cosh(a + w) * w > 0.5) w *= 0.5;
    double b = std::min(U, a + w);
    double h = 0.5 * (b - a);
    for (int i = 0; i < 16; ++i) {
      double u = a + h * (G.x[16][i] + 1.0);
      double sh = std::sinh(u), ch = std::cosh(u);
      double e = std::expm1(X * sh);
      i1 += h * G.w[16][i] * e;
      i2 += h * G.w[16][i] * e / (X * ch * ch);
    }
    a = b;
  }
  double R0 = std::sqrt(X * X + Y * Y);
  N += i1 + U * omj0;
  NX += j1v * U - (i2 + omj0 / X * (Y / R0));
}

}  // namespace

// ------------------------------------------------------------------ wave kernel (green_inf.rs:316-368)
// Far from the free surface (Y > 300, where e^Y of the identity form would overflow) Havelock's form (3b) with P = int_0^Y e^{t-Y} / sqrt(X^2 + t^2) dt
// written as int_0^Y e^{-s} / sqrt(X^2 + (Y - s)^2) ds (no overflow of e^t for large Y):
// Ft = -pi e^-Y (H0 + Y0) - 2 P, Ft_X = -pi e^-Y (2/pi - H1 - Y1) - 2 P_X; on the axis
// Ft(0, Y) = -2 e^-Y Ei(Y) by its asymptotic series.
void wave_far(double X, double Y, double& F, double& FX) {
  const GaussCache& G = gauss();
  const double top = std::min(Y, 60.0);
  double P = 0.0, PX = 0.0;
  const int panels = (int)std::ceil(top / 2.0);
  for (int q = 0; q < panels; ++q) {
    const double a = top * q / panels, b = top * (q + 1) / panels, h = 0.5 * (b - a);
    for (int i = 0; i < 16; ++i) {
      const double sv = a + h * (G.x[16][i] + 1.0), t = Y - sv;
      const double r2 = X * X + t * t, ir = 1.0 / std::sqrt(r2), e = std::exp(-sv) * h * G.w[16][i];
      P += e * ir;
      PX -= e * X * ir * ir * ir;
    }
  }
  if (X == 0.0) {
    // e^-Y Ei(Y) = sum_k k! / Y^(k+1) (asymptotic, Y > 40)
    double term = 1.0 / Y, s = 0.0;
    for (int k = 0; k < 40; ++k) {
      s += term;
      const double next = term * (k + 1) / Y;
      if (next > term || next < 1e-18 * s) break;
      term = next;
    }
    F = -2.0 * s;
    FX = 0.0;
    return;
  }
  double h0, h1;
  struve01(X, h0, h1);
  const double eY = std::exp(-Y);
  F = -PI * eY * (h0 + ::y0(X)) - 2.0 * P;
  FX = -PI * eY * (2.0 / PI - h1 - ::y1(X)) - 2.0 * PX;
}

extern "C" int orc_wave_kernel(double X, double Y, double* out) {
  if (!(X >= 0.0) || !(Y >= 0.0) || (X == 0.0 && Y == 0.0)) return ORC_ERR_ARGUMENT;
  double R0 = std::sqrt(X * X + Y * Y);
  double eY = std::exp(-Y);
  if (Y > 300.0) {
    double F, FX;
    wave_far(X, Y, F, FX);
    const double j0v = ::j0(X), j1v = ::j1(X);
    out[0] = F;
    out[1] = FX;
    out[2] = -F - 2.0 / R0;
    out[3] = -2.0 * PI * eY * j0v;
    out[4] = 2.0 * PI * eY * j1v;
    out[5] = -out[3];
    out[6] = eY * j0v;
    return ORC_OK;
  }
  double j0v = ::j0(X), j1v = ::j1(X);
  double yt, dyt;
  yt_and_derivative(X, yt, dyt);
  double V, VX;
  if (X == 0.0) {
    V = eY * ein(Y) / Y;
    VX = 0.0;
  } else {
    double N, NX;
    n_and_nx(X, Y, N, NX);
    V = eY * N / R0;
    VX = eY * (NX / R0 - N * X / (R0 * R0 * R0));
  }
  double lg = std::log(Y + R0) + GAMMA - LOG2;
  double F = -2.0 * eY * j0v * lg - PI * eY * yt - 2.0 * R0 * V;
  double FX = 2.0 * eY * j1v * lg - 2.0 * eY * j0v * X / (R0 * (Y + R0)) - PI * eY * dyt -
              2.0 * (X / R0) * V - 2.0 * R0 * VX;
  out[0] = F;
  out[1] = FX;
  out[2] = -F - 2.0 / R0;
  out[3] = -2.0 * PI * eY * j0v;
  out[4] = 2.0 * PI * eY * j1v;
  out[5] = -out[3];
  out[6] = eY * j0v;
  return ORC_OK;
}

extern "C" int orc_v_vx(double X, double Y, double* out) {
  if (!(X >= 0.0) || !(Y >= 0.0) || (X == 0.0 && Y == 0.0)) return ORC_ERR_ARGUMENT;
  double R0 = std::sqrt(X * X + Y * Y);
  double eY = std::exp(-Y);
  if (X == 0.0) {
    out[0] = eY * ein(Y) / Y;
    out[1] = 0.0;
    return ORC_OK;
  }
  double N, NX;
  n_and_nx(X, Y, N, NX);
  out[0] = eY * N / R0;
  out[1] = eY * (NX / R0 - N * X / (R0 * R0 * R0));
  return ORC_OK;
}

extern "C" int orc_yt(double X, double* out) {
  yt_and_derivative(X, out[0], out[1]);
  return ORC_OK;
}

extern "C" int orc_struve01(double x, double* out) {
  struve01(x, out[0], out[1]);
  return ORC_OK;
}

extern "C" int orc_gauss_legendre(int n, double* x, double* w) {
  if (n < 1 || n > 64) return ORC_ERR_ARGUMENT;
  std::vector<double> xs, ws;
  gauss_legendre(n, xs, ws);
  for (int i = 0; i < n; ++i) {
    x[i] = xs[i];
    w[i] = ws[i];
  }
  return ORC_OK;
}

extern "C" int orc_make_panel(const double* v12, double* out) {
  Panel p;
  int st = make_panel(v12, p);
  if (st) return st;
  // out: centroid[3], normal[3], area, distinct, proj[12]
  for (int i = 0; i < 3; ++i) {
    out[i] = p.centroid[i];
    out[3 + i] = p.normal[i];
  }
  out[6] = p.area;
  out[7] = p.distinct;
  for (int i = 0; i < 4; ++i)
    for (int d = 0; d < 3; ++d) out[8 + 3 * i + d] = p.proj[i][d];
  return ORC_OK;
}

static int geom_of(const double* v12, PanelGeom& g) {
  Panel p;
  int st = make_panel(v12, p);
  if (st) return st;
  g = make_panel_geom(p.proj, p.normal, p.centroid, p.area);
  return ORC_OK;
}

extern "C" int orc_rankine(const double* v12, const double* x, double* out) {
  PanelGeom g;
  int st = geom_of(v12, g);
  if (st) return st;
  rankine(g, v3(x[0], x[1], x[2]), out[0], out[1]);
  return ORC_OK;
}

extern "C" int orc_log_singularity(const double* v12, const double* image, double* out) {
  PanelGeom g;
  int st = geom_of(v12, g);
  if (st) return st;
  log_singularity(g, v3(image[0], image[1], image[2]), out[0], out[1]);
  return ORC_OK;
}

// ------------------------------------------------------------------ finite depth
// dispersion.rs:19-95: x0 tanh x0 = H (propagating) and xn tan xn = -H (evanescent), safeguarded
// Newton on the brackets of the reference.
extern "C" double orc_dispersion_x0(double H) {
  if (!(H > 0.0)) return 0.0;
  if (H >= 40.0) return H;
  double lo = std::max(H, std::sqrt(H)), hi = 0.5 * (H + std::sqrt(H * H + 4.0 * H));
  double x = H < 1.0 ? std::sqrt(H) * (1.0 + H / 6.0) : H;
  x = std::min(std::max(x, lo), hi);
  double prev = -1.0;
  for (int it = 0; it < 200; ++it) {
    double t = std::tanh(x), f = x * t - H;
    if (f < 0.0) lo = x; else hi = x;
    double fp = t + x * (1.0 - t * t);
    double xn = x - f / fp;
    if (!(xn >= lo && xn <= hi)) xn = 0.5 * (lo + hi);
    if (xn == x || xn == prev) break;
    prev = x;
    x = xn;
  }
  return x;
}

extern "C" double orc_dispersion_xn(double H, int n) {
  double npi = n * PI;
  if (!(H > 0.0)) return npi;
  if (std::isinf(H)) return npi - 0.5 * PI;
  double lo = 0.0, hi = 0.5 * PI, d = std::atan(H / (npi - 0.25 * PI)), prev = -1.0;
  for (int it = 0; it < 200; ++it) {
    double sd = std::sin(d), cd_ = std::cos(d), g = (npi - d) * sd - H * cd_;
    if (g < 0.0) lo = d; else hi = d;
    double gp = (npi - d) * cd_ + (H - 1.0) * sd;
    double dn = d - g / gp;
    if (!(dn > lo && dn < hi)) dn = 0.5 * (lo + hi);
    if (dn == d || dn == prev) break;
    prev = d;
    d = dn;
  }
  return npi - d;
}

namespace {

const int FIN_ROOTS = 24;         // green_fin.rs:113
const double SERIES_STOP = 18.0;  // green_fin.rs:349

// complex J0, J1 by their power series (|z| <= ~15 here)
void cj01(cd z, cd& j0, cd& j1) {
  cd q = -0.25 * z * z, t0 = 1.0, t1 = 1.0, s0 = 1.0, s1 = 1.0;
  for (int m = 1; m < 80; ++m) {
    t0 *= q / ((double)m * m);
    t1 *= q / ((double)m * (m + 1));
    s0 += t0;
    s1 += t1;
    if (std::abs(t0) < 1e-18 * std::abs(s0) && std::abs(t1) < 1e-18 * std::abs(s1) && m > 3) break;
  }
  j0 = s0;
  j1 = 0.5 * z * s1;
}

// The smooth remainder of Newman's decomposition (green_fin.rs:1-20, 466-497) without tables:
// Q(xb, vb) = PV int_0^inf (k+H)^2 cosh(k vb) e^{-3k} J0(k xb) / [(k sinh k - H cosh k)(k - H)] dk
// (k = kappa h dimensionless) with dQ/dxb and dQ/dvb; the principal value as the real part of
// the integral along 0 -> i d -> T + i d -> T -> inf, above the poles k = x0 and k = H and
// below the first imaginary pole i x1 (x1 > pi/2).
void q_direct(double H, double x0, double xb, double vb, double* q) {
  const GaussCache& G = gauss();
  cd s0(0, 0), s1(0, 0), s2(0, 0);
  auto f = [&](cd k, cd dk) {
    cd num = (k + H) * (k + H) * std::exp(-3.0 * k);
    cd den = (k * std::sinh(k) - H * std::cosh(k)) * (k - H);
    cd ch = std::cosh(k * vb), sh = std::sinh(k * vb);
    cd j0, j1;
    cj01(k * xb, j0, j1);
    cd w = num / den * dk;
    s0 += w * ch * j0;
    s1 += -w * ch * k * j1;
    s2 += w * k * sh * j0;
  };
  auto seg = [&](cd a, cd b, int panels) {
    for (int p = 0; p < panels; ++p) {
      cd pa = a + (b - a) * ((double)p / panels), pb = a + (b - a) * ((double)(p + 1) / panels);
      cd h2 = 0.5 * (pb - pa);
      for (int i = 0; i < 16; ++i) f(pa + h2 * (G.x[16][i] + 1.0), h2 * G.w[16][i]);
    }
  };
  if (std::min(x0, H) > 28.0) {
    seg(cd(0, 0), cd(28.0, 0), 56);
  } else {
    const double d = 0.7, T = std::max(x0, H) + 3.0;
    // the integrand behaves like 1/k for sqrt(H) << |k| << 1 (the log H of Q as H -> 0):
    // geometric panels up the imaginary axis from far below sqrt(H)
    double y = std::min(0.05, 1e-3 * std::sqrt(H));
    seg(cd(0, 0), cd(0, y), 1);
    while (2.0 * y < d) {
      seg(cd(0, y), cd(0, 2.0 * y), 1);
      y *= 2.0;
    }
    seg(cd(0, y), cd(0, d), 1);
    seg(cd(0, d), cd(T, d), (int)std::ceil(T) + 1);
    seg(cd(T, d), cd(T, 0), 1);
    // the tail decays at least like e^{-2k}
    seg(cd(T, 0), cd(T + 24.0, 0), 12);
  }
  q[0] = s0.real();
  q[1] = s1.real();
  q[2] = s2.real();
}

// Phi1 (which 1) and Phi2 (which 2) with their xb and vb derivatives (green_fin.rs:466-497):
// Phi1 = H [Ft(X, 2H - V) + Ft(X, 2H + V)] + Q(xb, vb), Phi2 = H Ft(X, 2H + V) + Q(xb, vb)
void phi_direct(double H, double x0, int which, double xb, double vb, double* out) {
  double q[3];
  q_direct(H, x0, xb, vb, q);
  out[0] = q[0];
  out[1] = q[1];
  out[2] = q[2];
  const double X = H * xb, H2 = H * H;
  double w[7];
  if (which == 1) {
    orc_wave_kernel(X, H * (2.0 - vb), w);
    out[0] += H * w[0];
    out[1] += H2 * w[1];
    out[2] -= H2 * w[2];
  }
  orc_wave_kernel(X, H * (2.0 + vb), w);
  out[0] += H * w[0];
  out[1] += H2 * w[1];
  out[2] += H2 * w[2];
}

struct FinDepth {
  double h = 0, K = 0, H = 0, x0 = 0, k0 = 0, A0 = 0;
  double kn[FIN_ROOTS], cn[FIN_ROOTS];
};

// make_finite_depth (green_fin.rs:874-934), ordinary frequencies
FinDepth make_fd(double h, double K) {
  FinDepth fd;
  fd.h = h;
  fd.K = K;
  fd.H = K * h;
  fd.x0 = orc_dispersion_x0(fd.H);
  fd.k0 = fd.x0 / h;
  fd.A0 = 2.0 * PI * fd.k0 / (-std::expm1(-4.0 * fd.x0) + 4.0 * fd.x0 * std::exp(-2.0 * fd.x0));
  for (int n = 1; n <= FIN_ROOTS; ++n) {
    double k = orc_dispersion_xn(fd.H, n) / h;
    fd.kn[n - 1] = k;
    fd.cn[n - 1] = 4.0 * (k * k + K * K) / (k * k * h + K * K * h - K);
  }
  return fd;
}

// K0, K1 by the integral int_0^inf e^{-x cosh t} (1, cosh t) dt, trapezoidal rule (exponentially
// convergent for this integrand)
void bessel_k01(double x, double& k0v, double& k1v) {
  const double step = 0.05;
  double s0 = 0.5 * std::exp(-x), s1 = s0;
  for (int i = 1; i < 4000; ++i) {
    double t = i * step, c = std::cosh(t), e = std::exp(-x * c);
    s0 += e;
    s1 += e * c;
    if (e < 1e-300 || e * c < 1e-19 * s1) break;
  }
  k0v = s0 * step;
  k1v = s1 * step;
}

// finite_kernel (green_fin.rs:1007-1095) for ordinary frequencies: the regular part
// Wr = Re G - sum_p 1/r_p and Wi = Im G with d/dR, d/dzeta (source), d/dz (field point)
void finite_kernel(const FinDepth& fd, double R, double z, double zeta, double* o) {
  const double h = fd.h, K = fd.K, k0 = fd.k0;
  const double d1 = z - zeta;
  const double sg = d1 > 0.0 ? 1.0 : d1 < 0.0 ? -1.0 : 0.0;
  const double v[6] = {std::fabs(d1), z + zeta + 2.0 * h, -(z + zeta), d1 + 2.0 * h, -d1 + 2.0 * h, z + zeta + 4.0 * h};
  double e3 = std::exp(-k0 * v[2]), e4 = std::exp(-k0 * v[3]), e5 = std::exp(-k0 * v[4]), e6 = std::exp(-k0 * v[5]);
  double E = e3 + e4 + e5 + e6, Ez = k0 * (e3 + e4 - e5 - e6), Ef = k0 * (e3 - e4 + e5 - e6);
  const double x = k0 * R;
  double j0v = ::j0(x), j1v = ::j1(x);
  double Wr, WrR, WrZ, Wrz;
  if (R >= 0.5 * h) {
    double y0v = ::y0(x), y1v = ::y1(x);
    double re = -fd.A0 * E * y0v, reR = fd.A0 * E * k0 * y1v, reZ = -fd.A0 * Ez * y0v, rez = -fd.A0 * Ef * y0v;
    for (int n = 0; n < FIN_ROOTS; ++n) {
      double k = fd.kn[n], xn = k * R;
      if (xn >= SERIES_STOP) break;
      double K0, K1;
      bessel_k01(xn, K0, K1);
      double cz = std::cos(k * (z + h)), sz = std::sin(k * (z + h));
      double cs = std::cos(k * (zeta + h)), ss = std::sin(k * (zeta + h));
      double c = fd.cn[n];
      re += c * cz * cs * K0;
      reR -= c * cz * cs * k * K1;
      reZ -= c * cz * k * ss * K0;
      rez -= c * k * sz * cs * K0;
    }
    const double dvz[6] = {-sg, 1.0, -1.0, -1.0, 1.0, 1.0}, dvf[6] = {sg, 1.0, -1.0, 1.0, -1.0, 1.0};
    for (int p = 0; p < 6; ++p) {
      double r2 = R * R + v[p] * v[p], r = std::sqrt(r2), ir3 = 1.0 / (r2 * r);
      re -= 1.0 / r;
      reR += R * ir3;
      reZ += v[p] * dvz[p] * ir3;
      rez += v[p] * dvf[p] * ir3;
    }
    Wr = re;
    WrR = reR;
    WrZ = reZ;
    Wrz = rez;
  } else {
    const double ih = 1.0 / h, ih2 = ih * ih;
    double f1[3], f2[3];
    phi_direct(fd.H, fd.x0, 1, R * ih, v[0] * ih, f1);
    phi_direct(fd.H, fd.x0, 2, R * ih, v[1] * ih, f2);
    Wr = ih * (f1[0] + f2[0]);
    WrR = ih2 * (f1[1] + f2[1]);
    double A1z = -sg * ih2 * f1[2];
    double Bz = ih2 * f2[2];
    double w[7];
    orc_wave_kernel(K * R, K * v[2], w);
    Wr += K * w[0];
    WrR += K * K * w[1];
    Bz -= K * K * w[2];
    WrZ = A1z + Bz;
    Wrz = -A1z + Bz;
  }
  o[0] = Wr;
  o[1] = WrR;
  o[2] = WrZ;
  o[3] = Wrz;
  o[4] = -fd.A0 * E * j0v;
  o[5] = fd.A0 * E * k0 * j1v;
  o[6] = -fd.A0 * Ez * j0v;
  o[7] = -fd.A0 * Ef * j0v;
}

// The regular part at the frequency limits in finite depth (green_fin.rs FreqMode::Zero with
// zero true, FreqMode::Infinite otherwise; Wi = 0): R >= h/2 the eigenfunction series of
// series_form (:256-322) with k_n = n pi / h (zero) or (n - 1/2) pi / h, c_n = 4/h
// (make_finite_depth :874-902), at omega = 0 the term -(2/h)(log(R/2h) + gamma), the Rankine
// signs +1 (zero) or (+, +, -, -, -, -); R < h/2, where the reference interpolates its generated
// tables of the smooth remainder (L0 and the last region's Q), the same Green function as the sum
// of the source's images in z = 0 and z = -h (family a at d1 + 2mh, family b at s + 2mh, signs
// 1, 1 at omega = 0 and (-1)^m, -(-1)^m at omega = inf), |m| <= 64 summed and the four runs
// beyond in closed form (omega = 0: Euler-Maclaurin about m = 64.5, the asinh integral whose log
// divergence the -(2/h) gamma of the series' regularization absorbs; omega = inf: Boole's sum of
// the alternating runs about m = 65), minus the six Rankine terms. Pinned by the mpmath rows of
// modes 1 and 2 in the reference's tests/data/green_fin_reference.txt.
void finite_limit_kernel(bool zero, double h, double R, double z, double zeta, double* o) {
  const double d1 = z - zeta, s = z + zeta;
  const double sg = d1 > 0.0 ? 1.0 : d1 < 0.0 ? -1.0 : 0.0;
  double re = 0.0, reR = 0.0, reZ = 0.0, rez = 0.0;
  if (R >= 0.5 * h) {
    const double v[6] = {std::fabs(d1), s + 2.0 * h, -s, d1 + 2.0 * h, -d1 + 2.0 * h, s + 4.0 * h};
    const double c = 4.0 / h;
    for (int n = 1; n <= FIN_ROOTS; ++n) {
      const double k = (n - (zero ? 0.0 : 0.5)) * PI / h, xn = k * R;
      if (xn >= SERIES_STOP) break;
      double K0, K1;
      bessel_k01(xn, K0, K1);
      const double cz = std::cos(k * (z + h)), sz = std::sin(k * (z + h));
      const double cs = std::cos(k * (zeta + h)), ss = std::sin(k * (zeta + h));
      re += c * cz * cs * K0;
      reR -= c * cz * cs * k * K1;
      reZ -= c * cz * k * ss * K0;
      rez -= c * k * sz * cs * K0;
    }
    if (zero) {
      re += -2.0 / h * (std::log(R / (2.0 * h)) + GAMMA);
      reR += -2.0 / (h * R);
    }
    const double dvz[6] = {-sg, 1.0, -1.0, -1.0, 1.0, 1.0}, dvf[6] = {sg, 1.0, -1.0, 1.0, -1.0, 1.0};
    for (int p = 0; p < 6; ++p) {
      const double sp = (zero || p < 2) ? 1.0 : -1.0;
      const double r2 = R * R + v[p] * v[p], r = std::sqrt(r2), ir3 = sp / (r2 * r);
      re -= sp / r;
      reR += R * ir3;
      reZ += v[p] * dvz[p] * ir3;
      rez += v[p] * dvf[p] * ir3;
    }
  } else {
    const int M = 64;
    for (int m = -M; m <= M; ++m) {
      const double e = zero ? 1.0 : ((m % 2) ? -1.0 : 1.0);
      if (m < -1 || m > 1) {  // family a: dx/dzeta = -1, dx/dz = 1
        const double x = d1 + 2.0 * m * h, r2 = R * R + x * x, r = std::sqrt(r2), ir3 = 1.0 / (r2 * r);
        re += e / r;
        reR -= e * R * ir3;
        reZ += e * x * ir3;
        rez -= e * x * ir3;
      }
      if (m < 0 || m > 2) {  // family b: dx/dzeta = dx/dz = 1
        const double t = zero ? e : -e;
        const double x = s + 2.0 * m * h, r2 = R * R + x * x, r = std::sqrt(r2), ir3 = 1.0 / (r2 * r);
        re += t / r;
        reR -= t * R * ir3;
        reZ -= t * x * ir3;
        rez -= t * x * ir3;
      }
    }
    // the runs m > M of 1/rho(c + 2 m h): c = d1, -d1 (family a), s, -s (family b)
    const double cs4[4] = {d1, -d1, s, -s}, dcz[4] = {-1.0, 1.0, 1.0, -1.0}, dcf[4] = {1.0, -1.0, 1.0, -1.0};
    const double h2 = 2.0 * h, h23 = h2 * h2 * h2;
    for (int i = 0; i < 4; ++i) {
      // f = 1/rho(x) and its derivatives at the run's start x0
      const double c = cs4[i];
      const double x0 = zero ? c + (2.0 * M + 1.0) * h : c + 2.0 * (M + 1) * h;
      const double r2 = R * R + x0 * x0, r = std::sqrt(r2);
      const double ir3 = 1.0 / (r2 * r), ir5 = ir3 / r2, ir7 = ir5 / r2, ir9 = ir7 / r2, x2 = x0 * x0;
      const double f0 = 1.0 / r, f1 = -x0 * ir3, f2 = -ir3 + 3.0 * x2 * ir5, f3 = 9.0 * x0 * ir5 - 15.0 * x2 * x0 * ir7;
      const double f4 = 9.0 * ir5 - 90.0 * x2 * ir7 + 105.0 * x2 * x2 * ir9;
      const double f0R = -R * ir3, f1R = 3.0 * x0 * R * ir5, f3R = -45.0 * x0 * R * ir7 + 105.0 * x2 * x0 * R * ir9;
      double T, TR, Tc;
      if (zero) {
        T = std::log(h2 / (x0 + r)) / h2 + h2 * f1 / 24.0 - 7.0 * h23 * f3 / 5760.0;
        TR = -(R / r) / (x0 + r) / h2 + h2 * f1R / 24.0 - 7.0 * h23 * f3R / 5760.0;
        Tc = -1.0 / (r * h2) + h2 * f2 / 24.0 - 7.0 * h23 * f4 / 5760.0;
      } else {
        const double w = (((M + 1) % 2) ? -1.0 : 1.0) * (i < 2 ? 1.0 : -1.0);
        T = w * (0.5 * f0 - 0.25 * h2 * f1 + h23 * f3 / 48.0);
        TR = w * (0.5 * f0R - 0.25 * h2 * f1R + h23 * f3R / 48.0);
        Tc = w * (0.5 * f1 - 0.25 * h2 * f2 + h23 * f4 / 48.0);
      }
      re += T;
      reR += TR;
      reZ += Tc * dcz[i];
      rez += Tc * dcf[i];
    }
    if (zero) re -= 2.0 / h * GAMMA;
  }
  o[0] = re;
  o[1] = reR;
  o[2] = reZ;
  o[3] = rez;
  for (int i = 4; i < 8; ++i) o[i] = 0.0;
}

}  // namespace

// mode 1: omega = 0, mode 2: omega = infinity (the reference's green_fin_reference.txt modes)
extern "C" int orc_finite_limit_kernel(int mode, double h, double R, double z, double zeta, double* out8) {
  if ((mode != 1 && mode != 2) || !(h > 0.0) || !(R >= 0.0) || z > 0.0 || zeta > 0.0 || z < -h || zeta < -h)
    return ORC_ERR_ARGUMENT;
  finite_limit_kernel(mode == 1, h, R, z, zeta, out8);
  return ORC_OK;
}

extern "C" int orc_finite_kernel(double h, double K, double R, double z, double zeta, double* out8) {
  if (!(h > 0.0) || !(K > 0.0) || !(R >= 0.0) || z > 0.0 || zeta > 0.0 || z < -h || zeta < -h)
    return ORC_ERR_ARGUMENT;
  FinDepth fd = make_fd(h, K);
  finite_kernel(fd, R, z, zeta, out8);
  return ORC_OK;
}

extern "C" int orc_finite_phi(double H, int which, double xb, double vb, double* out3) {
  phi_direct(H, orc_dispersion_x0(H), which, xb, vb, out3);
  return ORC_OK;
}

// ------------------------------------------------------------------ the problem
struct OrcProblem {
  int N = 0, nb = 0;
  std::vector<PanelGeom> pg;
  std::vector<int> body;
  std::vector<double> nj;  // [N][6]
  double rho = 1025.0, grav = 9.81, depth = INFINITY;
  bool logint = true;
  double logtol = 1e-4;
  std::vector<double> st;     // statics [k][j][4]: rs, rd, dl, dd
  std::vector<double> stinf;  // the omega = infinity rows [k][j][2]
  bool have_statics = false;
};

extern "C" void* orc_problem_create(const double* vertices, int n, const int* body, int nbodies,
                                    const double* rotation_centres, double rho, double grav, double depth,
                                    int logint, double logtol, int* status) {
  OrcProblem* p = new OrcProblem();
  p->N = n;
  p->nb = nbodies;
  p->rho = rho;
  p->grav = grav;
  p->depth = depth;
  p->logint = logint != 0;
  p->logtol = logtol;
  p->pg.resize(n);
  p->body.resize(n);
  p->nj.resize(6 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    Panel pn;
    int st = make_panel(vertices + 12 * (size_t)j, pn);
    if (st) {
      *status = st;
      delete p;
      return nullptr;
    }
    p->pg[j] = make_panel_geom(pn.proj, pn.normal, pn.centroid, pn.area);
    int b = body ? body[j] : 0;
    p->body[j] = b;
    // lo_geometry.rs:799-802: n_j = (n, (c - R) x n)
    V3 rc = v3(rotation_centres[3 * b], rotation_centres[3 * b + 1], rotation_centres[3 * b + 2]);
    V3 mom = cross(pn.centroid - rc, pn.normal);
    for (int d = 0; d < 3; ++d) {
      p->nj[6 * j + d] = pn.normal[d];
      p->nj[6 * j + 3 + d] = mom[d];
    }
  }
  *status = ORC_OK;
  return p;
}

extern "C" void orc_problem_free(void* h) { delete (OrcProblem*)h; }

static void ensure_statics(OrcProblem* p) {
  if (p->have_statics) return;
  int N = p->N;
  p->st.assign((size_t)N * N * 4, 0.0);
  p->stinf.assign((size_t)N * N * 2, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < N; ++k) {
    V3 x = p->pg[k].c;
    for (int j = 0; j < N; ++j)
      static_entry(p->pg[j], x, p->logint, p->logtol, p->depth, &p->st[((size_t)k * N + j) * 4],
                   &p->stinf[((size_t)k * N + j) * 2]);
  }
  p->have_statics = true;
}

extern "C" int orc_statics(void* h, double* out) {
  OrcProblem* p = (OrcProblem*)h;
  ensure_statics(p);
  std::memcpy(out, p->st.data(), p->st.size() * sizeof(double));
  return ORC_OK;
}

namespace {

struct Kern {
  double Wr = 0, WrR = 0, WrZ = 0, Wrz = 0, Wi = 0, WiR = 0, WiZ = 0, Wiz = 0;
};

// low_order.rs:627-648 and :717-730 (infinite depth, ordinary frequency)
Kern pair_kernel(double K, V3 x, V3 xi, double& R) {
  double dx = x[0] - xi[0], dy = x[1] - xi[1];
  R = std::sqrt(dx * dx + dy * dy);
  Kern k;
  if (R == 0.0 && x[2] == 0.0 && xi[2] == 0.0) {
    k.Wr = -2.0 * K * (GAMMA - LOG2 + std::log(K));
    k.Wi = -2.0 * PI * K;
    return k;
  }
  double w[7];
  orc_wave_kernel(K * R, -K * (x[2] + xi[2]), w);
  double K2 = K * K;
  k.Wr = K * w[0];
  k.WrR = K2 * w[1];
  k.WrZ = -K2 * w[2];
  k.Wrz = k.WrZ;
  k.Wi = K * w[3];
  k.WiR = K2 * w[4];
  k.WiZ = -K2 * w[5];
  k.Wiz = k.WiZ;
  return k;
}

// low_order.rs:738-769 (wetted panels: the dipole column is D)
// at omega = infinity (at_inf) the statics' omega = infinity row (:752-753)
void apply_kernel(const OrcProblem* p, int row_k, int j, V3 x, double K, double logc, const Kern& k, double R,
                  bool swapped, cd& S, cd& D, bool at_inf = false) {
  const PanelGeom& pj = p->pg[j];
  double a = R > 0.0 ? ((x[0] - pj.c[0]) * pj.n[0] + (x[1] - pj.c[1]) * pj.n[1]) / R : 0.0;
  const size_t o = (size_t)row_k * p->N + j;
  const double* e = &p->st[o * 4];
  const double rs = at_inf ? p->stinf[o * 2] : e[0], rd = at_inf ? p->stinf[o * 2 + 1] : e[1];
  double A = pj.area;
  double rz = swapped ? k.Wrz : k.WrZ;
  double iz = swapped ? k.Wiz : k.WiZ;
  S = cd(rs + A * k.Wr - 2.0 * K * logc * e[2], A * k.Wi);
  D = cd(rd + A * (-k.WrR * a + rz * pj.n[2]) - 2.0 * K * logc * e[3], A * (-k.WiR * a + iz * pj.n[2]));
}

// lo_frequency.rs:399-537 for G = 1: S(k, n), A(k, n) = 2 pi delta_kn - D(k, n), column-major;
// finite depth: the kernel of finite_kernel (green_fin.rs:1007-1095), the (n, k) entry with the
// derivative in its first point (Wrz, lo_frequency.rs:494-497)
// omega = 0 and omega = infinity: the limits (solve_common.rs:44-71: K = 0, no log terms; the
// kernel of finite_limit_kernel in finite depth, none in infinite depth: 1/r +- 1/r' only,
// low_order.rs:635-637)
void assemble(OrcProblem* p, double omega, std::vector<cd>& S, std::vector<cd>& Am) {
  ensure_statics(p);
  int N = p->N;
  const bool limit = omega == 0.0 || std::isinf(omega), at_inf = std::isinf(omega);
  double K = limit ? 0.0 : omega * omega / p->grav;
  double logc = p->logint && !limit ? 1.0 : 0.0;
  const bool finite = std::isfinite(p->depth);
  FinDepth fd;
  if (finite && !limit) fd = make_fd(p->depth, K);
  S.assign((size_t)N * N, cd(0, 0));
  Am.assign((size_t)N * N, cd(0, 0));
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < N; ++k) {
    V3 xk = p->pg[k].c;
    for (int n = k; n < N; ++n) {
      V3 xn = p->pg[n].c;
      double R;
      Kern w;
      if (finite) {
        double dx = xk[0] - xn[0], dy = xk[1] - xn[1];
        R = std::sqrt(dx * dx + dy * dy);
        double o[8];
        if (limit)
          finite_limit_kernel(!at_inf, p->depth, R, xk[2], xn[2], o);
        else
          finite_kernel(fd, R, xk[2], xn[2], o);
        w.Wr = o[0];
        w.WrR = o[1];
        w.WrZ = o[2];
        w.Wrz = o[3];
        w.Wi = o[4];
        w.WiR = o[5];
        w.WiZ = o[6];
        w.Wiz = o[7];
      } else if (limit) {
        double dx = xk[0] - xn[0], dy = xk[1] - xn[1];
        R = std::sqrt(dx * dx + dy * dy);
      } else {
        w = pair_kernel(K, xk, xn, R);
      }
      cd S1, D1, S2, D2;
      apply_kernel(p, k, n, xk, K, logc, w, R, false, S1, D1, at_inf);
      S[k + (size_t)n * N] = S1;
      Am[k + (size_t)n * N] = (k == n ? cd(2.0 * PI, 0.0) : cd(0, 0)) - D1;
      if (n != k) {
        apply_kernel(p, n, k, xn, K, logc, w, R, true, S2, D2, at_inf);
        S[n + (size_t)k * N] = S2;
        Am[n + (size_t)k * N] = -D2;
      }
    }
  }
}

// Complex LU with partial pivoting (linalg.rs:134-252 semantics: LAPACK ipiv, unit lower L,
// zero pivot -> numerical error); the pivot is the entry of largest |re| + |im| (LAPACK izamax).
int lu_solve(int n, cd* a, int nrhs, cd* b) {
  for (size_t i = 0; i < (size_t)n * n; ++i)
    if (!std::isfinite(a[i].real()) || !std::isfinite(a[i].imag())) return ORC_ERR_NUMERICAL;
  std::vector<int> ipiv(n);
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = -1.0;
    for (int i = k; i < n; ++i) {
      cd v = a[i + (size_t)k * n];
      double m = std::fabs(v.real()) + std::fabs(v.imag());
      if (m > best) {
        best = m;
        piv = i;
      }
    }
    ipiv[k] = piv;
    if (piv != k)
      for (int j = 0; j < n; ++j) std::swap(a[k + (size_t)j * n], a[piv + (size_t)j * n]);
    cd d = a[k + (size_t)k * n];
    if (d == cd(0, 0)) return ORC_ERR_NUMERICAL;
    cd inv = 1.0 / d;
    for (int i = k + 1; i < n; ++i) a[i + (size_t)k * n] *= inv;
#pragma omp parallel for schedule(static) if (n - k > 256)
    for (int j = k + 1; j < n; ++j) {
      cd u = a[k + (size_t)j * n];
      if (u == cd(0, 0)) continue;
      cd* cj = a + (size_t)j * n;
      const cd* ck = a + (size_t)k * n;
      for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * u;
    }
  }
  for (int c = 0; c < nrhs; ++c) {
    cd* x = b + (size_t)c * n;
    for (int k = 0; k < n; ++k)
      if (ipiv[k] != k) std::swap(x[k], x[ipiv[k]]);
    for (int k = 0; k < n; ++k) {
      cd v = x[k];
      for (int i = k + 1; i < n; ++i) x[i] -= a[i + (size_t)k * n] * v;
    }
    for (int k = n - 1; k >= 0; --k) {
      x[k] /= a[k + (size_t)k * n];
      cd v = x[k];
      for (int i = 0; i < k; ++i) x[i] -= a[i + (size_t)k * n] * v;
    }
  }
  for (size_t i = 0; i < (size_t)n * nrhs; ++i)
    if (!std::isfinite(b[i].real()) || !std::isfinite(b[i].imag())) return ORC_ERR_NUMERICAL;
  return ORC_OK;
}

}  // namespace

extern "C" int orc_assemble(void* h, double omega, double* S_out, double* A_out) {
  OrcProblem* p = (OrcProblem*)h;
  if (!(omega >= 0.0)) return ORC_ERR_ARGUMENT;
  std::vector<cd> S, A;
  assemble(p, omega, S, A);
  std::memcpy(S_out, S.data(), S.size() * sizeof(cd));
  std::memcpy(A_out, A.data(), A.size() * sizeof(cd));
  return ORC_OK;
}

extern "C" int orc_lu_solve(int n, double* a, int nrhs, double* b) {
  return lu_solve(n, (cd*)a, nrhs, (cd*)b);
}

// One solve over F frequencies and H headings (degrees). Outputs (double / interleaved complex):
// A, B [F][D][K], X, FK, SC [F][H][D], HK [F][K][H], with D = K = 6 nb (rigid modes of each body,
// lo_frequency.rs:1039-1171; solve_driver.rs:362-387 orders the result arrays so).
extern "C" int orc_solve(void* h, const double* omegas, int nf, const double* headings_deg, int nh, double* Aout,
                         double* Bout, double* Xout, double* FKout, double* SCout, double* HKout) {
  OrcProblem* p = (OrcProblem*)h;
  int N = p->N;
  int D = 6 * p->nb, nrad = D, nrhs = nrad + nh;
  std::vector<cd> S, Am;
  for (int fi = 0; fi < nf; ++fi) {
    double w = omegas[fi];
    if (!(w >= 0.0)) return ORC_ERR_ARGUMENT;
    // omega = 0 and infinity: the radiation limits (no incident wave; B, X, FK, SC, HK are zero)
    const bool limit = w == 0.0 || std::isinf(w);
    const int nhw = limit ? 0 : nh;
    double K = limit ? 0.0 : w * w / p->grav, k = K;
    const bool finite = std::isfinite(p->depth);
    if (finite && !limit) k = orc_dispersion_x0(K * p->depth) / p->depth;
    assemble(p, w, S, Am);
    // boundary values (lo_frequency.rs:214-273; low_order.rs:1458-1468 incident_one, frame at the origin)
    std::vector<cd> phiI((size_t)nh * N), dphiI((size_t)nh * N), u((size_t)nrhs * N, cd(0, 0));
    for (int hh = 0; hh < nhw; ++hh) {
      double beta = headings_deg[hh] * PI / 180.0;
      double cb = std::cos(beta), sb = std::sin(beta);
      for (int j = 0; j < N; ++j) {
        V3 x = p->pg[j].c;
        double phase = -k * (x[0] * cb + x[1] * sb);
        // depth factor (low_order.rs:1401-1413): cosh k(z + h) / cosh kh and its z derivative
        double Z = std::exp(k * x[2]), Zz = k * Z;
        if (finite) {
          const double hh = p->depth, b = std::exp(-k * (x[2] + 2.0 * hh)), d = 1.0 + std::exp(-2.0 * k * hh);
          Zz = k * (Z - b) / d;
          Z = (Z + b) / d;
        }
        cd e = cd(0.0, p->grav / w) * cd(std::cos(phase), std::sin(phase));
        cd phi = e * Z;
        cd g0 = phi * cd(0.0, -k * cb), g1 = phi * cd(0.0, -k * sb), g2 = e * Zz;
        V3 n = p->pg[j].n;
        phiI[(size_t)hh * N + j] = phi;
        dphiI[(size_t)hh * N + j] = g0 * n[0] + g1 * n[1] + g2 * n[2];
      }
    }
    for (int c = 0; c < nrad; ++c) {
      int b = c / 6, mode = c % 6;
      for (int j = 0; j < N; ++j)
        if (p->body[j] == b) u[(size_t)c * N + j] = cd(p->nj[6 * j + mode], 0.0);
    }
    for (int hh = 0; hh < nh; ++hh)
      for (int j = 0; j < N; ++j) u[(size_t)(nrad + hh) * N + j] = -dphiI[(size_t)hh * N + j];
    // rhs = S U, U = -u (lo_frequency.rs:855-893 with G = 1, chi = 1)
    std::vector<cd> rhs((size_t)N * nrhs, cd(0, 0));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nrhs; ++c)
      for (int nn = 0; nn < N; ++nn) {
        cd un = -u[(size_t)c * N + nn];
        if (un == cd(0, 0)) continue;
        const cd* col = &S[(size_t)nn * N];
        cd* r = &rhs[(size_t)c * N];
        for (int kk = 0; kk < N; ++kk) r[kk] += col[kk] * un;
      }
    int st = lu_solve(N, Am.data(), nrhs, rhs.data());
    if (st) return st;
    const std::vector<cd>& phi = rhs;  // [nrhs][N]
    cd iwr(0.0, limit ? 0.0 : w * p->rho);
    for (int i = 0; i < D; ++i) {
      int bi = i / 6, mi = i % 6;
      for (int c = 0; c < nrad; ++c) {
        cd z(0, 0);
        for (int j = 0; j < N; ++j)
          if (p->body[j] == bi) z += p->nj[6 * j + mi] * phi[(size_t)c * N + j] * p->pg[j].area;
        z *= -p->rho;
        size_t o = ((size_t)fi * D + i) * nrad + c;
        Aout[o] = z.real();
        Bout[o] = limit ? 0.0 : -w * z.imag();
      }
      for (int hh = 0; hh < nh; ++hh) {
        cd x(0, 0), fk(0, 0);
        for (int j = 0; j < N; ++j)
          if (p->body[j] == bi) {
            double ni = p->nj[6 * j + mi];
            cd pd = phi[(size_t)(nrad + hh) * N + j] + phiI[(size_t)hh * N + j];
            x += ni * pd * p->pg[j].area;
            fk += ni * phiI[(size_t)hh * N + j] * p->pg[j].area;
          }
        size_t o = ((size_t)fi * nh + hh) * D + i;
        cd X = iwr * x, FK = iwr * fk, SC = iwr * (x - fk);
        Xout[2 * o] = X.real();
        Xout[2 * o + 1] = X.imag();
        FKout[2 * o] = FK.real();
        FKout[2 * o + 1] = FK.imag();
        SCout[2 * o] = SC.real();
        SCout[2 * o + 1] = SC.imag();
      }
    }
    for (int c = 0; c < nrad; ++c) {
      int bi = c / 6, mc = c % 6;
      for (int hh = 0; hh < nh; ++hh) {
        cd v(0, 0);
        for (int j = 0; j < N; ++j) {
          cd ni = p->body[j] == bi ? cd(p->nj[6 * j + mc], 0.0) : cd(0, 0);
          v += (ni * phiI[(size_t)hh * N + j] - phi[(size_t)c * N + j] * dphiI[(size_t)hh * N + j]) * p->pg[j].area;
        }
        cd HK = iwr * v;
        size_t o = ((size_t)fi * nrad + c) * nh + hh;
        HKout[2 * o] = HK.real();
        HKout[2 * o + 1] = HK.imag();
      }
    }
  }
  return ORC_OK;
}
