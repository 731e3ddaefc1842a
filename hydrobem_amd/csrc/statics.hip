// Frequency-independent influence data of the low-order method, FP64 on the GPU.
//
// For every field point k (the centroid of panel k) and panel j (lo_static.rs:59-119, infinite
// depth): rs, rd = Rankine source and dipole integrals of 1/r + 1/r' (closed forms of
// panel_integrals.rs:226-275: Newman 1986 edge logarithms and the Van Oosterom-Strackee solid
// angle, kept to 1000 panel radii as the reference's double mode does, :231), and where the log
// singularity of the wave kernel needs its analytic integral (lo_static.rs:47-55, LOG_BOUND 0.6)
// the corrections dl, dd = int psi dS - A psi(c) and the same for d psi / dn
// (panel_integrals.rs:380-450: Gauss rules of order 4..10 by distance, adaptive quadrisection
// of the 4 x 4 rule near the image point).
//
// Layout: StaticD st[j * N + k] for the pair (field point k, panel j), i.e. column-major in
// (k, j) like the matrices the assembly writes. Thread mapping: the 64 lanes of a wavefront take
// 64 consecutive field points k of one panel j, so the panel record is a wavefront-uniform
// (scalar) load and the stores are coalesced. The pairs that need the adaptive rule are queued
// and finished by one wavefront per pair (adaptive_kernel): its 64 lanes evaluate the four
// child cells of a quadrisection step (16 Gauss points each) at once, and the depth-first stack
// of cells lives in LDS.
#include "../../include/hydrobem_amd.h"
#include "hbg_internal.h"

namespace hbg {

namespace {

struct P3 {
  double x, y, z;
};
__device__ __forceinline__ P3 mk(double x, double y, double z) { return P3{x, y, z}; }
__device__ __forceinline__ P3 sub(P3 a, P3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ double dot3(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ P3 cross3(P3 a, P3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double nrm(P3 a) { return sqrt(dot3(a, a)); }
__device__ __forceinline__ P3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }

// the bilinear map of corners q (at (-1,-1), (1,-1), (1,1), (-1,1)): point and Jacobian
__device__ __forceinline__ void bilinear(const double (*q)[3], double u, double v, P3& x, double& jac) {
  const double n0 = 0.25 * (1.0 - u) * (1.0 - v), n1 = 0.25 * (1.0 + u) * (1.0 - v);
  const double n2 = 0.25 * (1.0 + u) * (1.0 + v), n3 = 0.25 * (1.0 - u) * (1.0 + v);
  x = mk(n0 * q[0][0] + n1 * q[1][0] + n2 * q[2][0] + n3 * q[3][0],
         n0 * q[0][1] + n1 * q[1][1] + n2 * q[2][1] + n3 * q[3][1],
         n0 * q[0][2] + n1 * q[1][2] + n2 * q[2][2] + n3 * q[3][2]);
  const double a = 0.25 * (1.0 - v), b = 0.25 * (1.0 + v), c = 0.25 * (1.0 - u), d = 0.25 * (1.0 + u);
  P3 xu = mk(a * (q[1][0] - q[0][0]) + b * (q[2][0] - q[3][0]), a * (q[1][1] - q[0][1]) + b * (q[2][1] - q[3][1]),
             a * (q[1][2] - q[0][2]) + b * (q[2][2] - q[3][2]));
  P3 xv = mk(c * (q[3][0] - q[0][0]) + d * (q[2][0] - q[1][0]), c * (q[3][1] - q[0][1]) + d * (q[2][1] - q[1][1]),
             c * (q[3][2] - q[0][2]) + d * (q[2][2] - q[1][2]));
  jac = nrm(cross3(xu, xv));
}

// Rankine source S and dipole D of panel p at x (panel_integrals.rs:226-275)
__device__ __forceinline__ void rankine_d(const PanelD* __restrict__ p, P3 x, const double* __restrict__ gl,
                                          double& S, double& D) {
  const P3 c = ld3(p->c), n = ld3(p->n);
  const P3 rc = sub(x, c);
  const double radius = p->radius;
  if (nrm(rc) > 1000.0 * radius) {
    // 3 x 3 Gauss rule on the bilinear map of the corners, weights rescaled to the area
    const double* x3 = gl + 3 * 32;
    const double* w3 = x3 + 16;
    double s = 0.0, d = 0.0, sum = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        P3 g;
        double jac;
        bilinear(p->q4, x3[i], x3[j], g, jac);
        const double ww = w3[i] * w3[j] * jac;
        const P3 r = sub(x, g);
        const double ir = 1.0 / nrm(r);
        s += ww * ir;
        d += ww * dot3(n, r) * ir * ir * ir;
        sum += ww;
      }
    S = s * (p->area / sum);
    D = d * (p->area / sum);
    return;
  }
  const int nv = (int)p->nv;
  const double w = dot3(rc, n);
  P3 av[4];
  double r[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    av[i] = sub(ld3(p->v[i]), x);
    r[i] = nrm(av[i]);
  }
  double omega = 0.0;
  if (fabs(w) > 1e-12 * radius) {
    for (int i = 1; i < nv - 1; ++i) {
      const P3 a = av[0], b = av[i], cc = av[i + 1];
      const double num = dot3(a, cross3(b, cc));
      const double den = r[0] * r[i] * r[i + 1] + dot3(a, b) * r[i + 1] + dot3(a, cc) * r[i] + dot3(b, cc) * r[0];
      omega -= 2.0 * atan2(num, den);
    }
  }
  double s = 0.0;
  for (int e = 0; e < nv; ++e) {
    const int f = (e + 1 == nv) ? 0 : e + 1;
    const double d = dot3(av[e], ld3(p->em[e]));
    if (fabs(d) <= 1e-14 * radius) continue;
    const double q = r[e] * r[f] + dot3(av[e], av[f]);
    if (!(q > 0.0)) continue;
    const double sm = r[e] + r[f] + p->elen[e];
    s += d * log(sm * sm / (2.0 * q));
  }
  S = s - w * omega;
  D = omega;
}

// psi = log(|xi - p| + p_z - xi_z) and its derivative along n at xi (panel_integrals.rs:281-286)
__device__ __forceinline__ void psi_terms_d(P3 xi, P3 img, P3 n, double& psi, double& dpsi) {
  const P3 d = sub(xi, img);
  const double R = nrm(d);
  const double den = R - d.z;
  psi = log(den);
  dpsi = (dot3(d, n) / R - n.z) / den;
}

// the Gauss order of the log integral by distance and clearance, 0: adaptive (panel_integrals.rs:392-419)
__device__ __forceinline__ int log_order(const PanelD* __restrict__ p, P3 img) {
  const double dist = nrm(sub(img, ld3(p->c)));
  const double r = dist / p->radius;
  const double clearance = (img.z - p->top) / p->radius;
  if (clearance >= 0.5 || r >= 2.0)
    return r >= 6.0 ? 4 : r >= 3.0 ? 5 : r >= 2.0 ? 6 : r >= 1.5 ? 8 : r >= 1.0 ? 10 : 0;
  return 0;
}

__device__ __forceinline__ bool log_pair_d(const PanelD* __restrict__ p, P3 x, double tol) {
  const double dx = x.x - p->c[0], dy = x.y - p->c[1];
  const double zz = fabs(x.z + p->c[2]);
  const double d = sqrt(dx * dx + dy * dy + zz * zz) + zz;
  return d * d * tol < 0.6 * p->radius * p->radius;
}

// int log|xi - x| dS over the plane polygon of p for x in its plane: the polygon as the signed
// triangles (x, edge) (panel_integrals.rs:621-648), each in polar coordinates about x: with h the
// distance of x from the edge's line (> 0 inside) and t the coordinate along the edge from the
// foot of the perpendicular, r = sqrt(h^2 + t^2) on the edge, and
//   int int log r r dr dtheta = (h / 4) int (log(h^2 + t^2) - 1) dt = (h / 2) [F(t2) - F(t1)],
//   F(t) = (t log(h^2 + t^2) - 2 t + 2 h atan(t / h)) / 2 - t / 2.
__device__ __forceinline__ double log_in_plane(const PanelD* __restrict__ p, P3 x) {
  const int nv = (int)p->nv;
  double sum = 0.0;
  for (int i = 0; i < nv; ++i) {
    const P3 a = ld3(p->v[i]), b = ld3(p->v[(i + 1) % nv]);
    const double h = dot3(sub(a, x), ld3(p->em[i]));
    if (h == 0.0) continue;
    const double il = 1.0 / p->elen[i];
    const P3 e = mk(il * (b.x - a.x), il * (b.y - a.y), il * (b.z - a.z));
    const double t1 = dot3(sub(a, x), e), t2 = dot3(sub(b, x), e);
    auto F = [h](double t) {
      const double q = h * h + t * t;
      return 0.5 * (t * log(q) - 2.0 * t + 2.0 * h * atan(t / h)) - 0.5 * t;
    };
    sum += 0.5 * h * (F(t2) - F(t1));
  }
  return sum;
}

// The record of the pair (field point x, panel p): returns whether its log corrections need the
// adaptive rule (then dl, dd are left 0 for adaptive_kernel). sinf, dinf: the omega = infinity
// row, the integrals of 1/r - 1/r' (infinite depth) or of 1/r1 + 1/r2 - (1/r3 + ... + 1/r6)
// (finite depth, the six image points).
__device__ __forceinline__ bool pair_statics(const PanelD* __restrict__ p, P3 x, int logint, double tol, double depth,
                                             const double* __restrict__ gl, StaticD& e, double& sinf,
                                             double& dinf) {
  double s0, d0, sb, db;
  rankine_d(p, x, gl, s0, d0);
  const P3 xb = mk(x.x, x.y, -x.z);
  rankine_d(p, xb, gl, sb, db);
  e.rs = s0 + sb;
  e.rd = d0 + db;
  sinf = s0 - sb;
  dinf = d0 - db;
  if (depth > 0.0) {
    // finite depth: the bottom images of lo_static.rs:29-38 (x, y, -z - 2h), (x, y, z + 2h),
    // (x, y, z - 2h), (x, y, -z - 4h), summed in the reference's order
    const double h = depth;
    double s2, d2, s4, d4, s5, d5, s6, d6;
    rankine_d(p, mk(x.x, x.y, -x.z - 2.0 * h), gl, s2, d2);
    rankine_d(p, mk(x.x, x.y, x.z + 2.0 * h), gl, s4, d4);
    rankine_d(p, mk(x.x, x.y, x.z - 2.0 * h), gl, s5, d5);
    rankine_d(p, mk(x.x, x.y, -x.z - 4.0 * h), gl, s6, d6);
    e.rs += (s2 + s4) + s5 + s6;
    e.rd += (d2 + d4) + d5 + d6;
    // the omega = infinity row in finite depth (lo_static.rs:84-95): 1/r1 + 1/r2 - (1/r3 + ... + 1/r6)
    sinf = s0 + s2 - sb - s4 - s5 - s6;
    dinf = d0 + d2 - db - d4 - d5 - d6;
  }
  e.dl = 0.0;
  e.dd = 0.0;
  if (logint && log_pair_d(p, x, tol) && p->interior != 0.0 && x.z == 0.0) {
    // both in z = 0 (lo_static.rs:104-113): psi = log |xi - x| integrated exactly in the plane;
    // at the panel's own centroid the centroid value is part of the kernel's limit; the columns
    // of interior panels are K times their sources, so dd is not used
    const double L = log_in_plane(p, x);
    const double Rr = nrm(sub(ld3(p->c), x));
    e.dl = Rr > 0.0 ? L - p->area * log(Rr) : L;
  } else if (logint && log_pair_d(p, x, tol)) {
    const int order = log_order(p, xb);
    if (order > 0) {
      const double* xs = gl + order * 32;
      const double* ws = xs + 16;
      const double(*q)[3] = p->v;  // the map of the distinct vertices (v[3] = v[2] for triangles)
      const P3 n = ld3(p->n);
      double L = 0.0, Ld = 0.0;
      for (int i = 0; i < order; ++i)
        for (int jj = 0; jj < order; ++jj) {
          P3 xi;
          double jac, psi, dpsi;
          bilinear(q, xs[i], xs[jj], xi, jac);
          psi_terms_d(xi, xb, n, psi, dpsi);
          const double ww = ws[i] * ws[jj] * jac;
          L += ww * psi;
          Ld += ww * dpsi;
        }
      double psi, dpsi;
      psi_terms_d(ld3(p->c), xb, n, psi, dpsi);
      e.dl = L - p->area * psi;
      e.dd = Ld - p->area * dpsi;
    } else {
      return true;
    }
  }
  return false;
}

__device__ __forceinline__ void queue_pair(int2* __restrict__ queue, unsigned* __restrict__ qcount, unsigned qcap,
                                           int k, int j) {
  const unsigned slot = atomicAdd(qcount, 1u);
  if (slot < qcap) queue[slot] = make_int2(k, j);
}

}  // namespace

// gl: Gauss-Legendre nodes and weights [order 0..16][x 16 | w 16]. inf (optional):
// the omega = infinity row of the pair (lo_static.rs:84-96, pair_statics), at inf[j * N + k] as
// (source, dipole).
__global__ void __launch_bounds__(256) statics_kernel(const PanelD* __restrict__ P, const PanelD* __restrict__ Pf,
                                                      int N, int Nsrc, int logint,
                                                      double tol, double depth, const double* __restrict__ gl,
                                                      StaticD* __restrict__ st, int2* __restrict__ queue,
                                                      unsigned* __restrict__ qcount, unsigned qcap,
                                                      double2* __restrict__ inf) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + threadIdx.y);
  if (j >= Nsrc || k >= N) return;
  StaticD e;
  double sinf, dinf;
  if (pair_statics(P + j, ld3(Pf[k].c), logint, tol, depth, gl, e, sinf, dinf)) queue_pair(queue, qcount, qcap, k, j);
  if (inf) inf[(size_t)j * N + k] = make_double2(sinf, dinf);
  st[(size_t)j * N + k] = e;
}

// The records of the listed pairs (field point ent[e].x, panel ent[e].y) only, at their places
// st[j * N + k]: the near entries of the FP32 assembly, whose other pairs take FP32 records.
__global__ void __launch_bounds__(256) statics_list_kernel(const PanelD* __restrict__ P, int N,
                                                           const int2* __restrict__ ent, int nnz, int logint,
                                                           double tol, double depth, const double* __restrict__ gl,
                                                           StaticD* __restrict__ st, int2* __restrict__ queue,
                                                           unsigned* __restrict__ qcount, unsigned qcap) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int k = ent[e].x, j = ent[e].y;
  StaticD r;
  double sinf, dinf;
  if (pair_statics(P + j, ld3(P[k].c), logint, tol, depth, gl, r, sinf, dinf)) queue_pair(queue, qcount, qcap, k, j);
  st[(size_t)j * N + k] = r;
}

// ------------------------------------------------------------------ FP32 statics
// The FP32 records (rs, rd, dl, dd) of the precision path with the FP32 far-field assembly, for
// the pairs whose field point lies at least two radii from the panel (every image point then lies
// farther still: |z + zeta| >= |z - zeta| for z, zeta <= 0, and likewise about the sea floor):
// - the Rankine integrals of the field point and each image by Gauss rules on the bilinear map of
//   the corners, the float32 tiers of the reference's production path (panel_integrals.rs:288-306:
//   orders 6, 5, 4, 3 from 2, 3, 4, 8 radii; 1e-6 relative at each tier's inner boundary), the
//   points taken relative to the centroid in FP64 first and the weights rescaled to the area
//   (panel_integrals.rs:185-214), so that no FP32 difference cancels;
// - the log corrections by the Gauss rules of log_order on the map of the distinct vertices.
// A wavefront takes 64 field points of one panel and evaluates every lane with the highest order
// any of its lanes needs (an order above a pair's tier only adds accuracy), so that the rule
// points are wavefront-uniform scalar loads. The pairs nearer than two radii are written 0: they
// are near entries of the assembly, whose FP64 records replace them.
constexpr int TIER_RANKINE = 9 + 16 + 25 + 36;  // orders 3, 4, 5, 6
constexpr int TIER_LOG = 16 + 25 + 36;          // orders 4, 5, 6
constexpr int TIER_STRIDE = TIER_RANKINE + TIER_LOG + 1;  // float4 per panel (padded to 32 bytes)
__host__ __device__ constexpr int tier_rankine_off(int order) {
  return order == 3 ? 0 : order == 4 ? 9 : order == 5 ? 25 : 50;
}
__host__ __device__ constexpr int tier_log_off(int order) {
  return TIER_RANKINE + (order == 4 ? 0 : order == 5 ? 16 : 41);
}

// rules[j][TIER_STRIDE]: (x - c, y - c, z - c, w) of every rule point of panel j; one workgroup
// per panel, one thread per point
__global__ void __launch_bounds__(192) tier_rules_kernel(const PanelD* __restrict__ P, int N,
                                                         const double* __restrict__ gl, float4* __restrict__ rules) {
  const int j = blockIdx.x, t = threadIdx.x;
  if (j >= N || t >= TIER_STRIDE) return;
  const PanelD* p = P + j;
  float4* out = rules + (size_t)j * TIER_STRIDE;
  if (t == TIER_STRIDE - 1) {
    out[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const P3 c = ld3(p->c);
  const bool rank = t < TIER_RANKINE;
  int order, q;
  if (rank) {
    order = t < 9 ? 3 : t < 25 ? 4 : t < 50 ? 5 : 6;
    q = t - tier_rankine_off(order);
  } else {
    const int u = t - TIER_RANKINE;
    order = u < 16 ? 4 : u < 41 ? 5 : 6;
    q = t - tier_log_off(order);
  }
  const double* xs = gl + order * 32;
  const double* ws = xs + 16;
  const double(*map)[3] = rank ? p->q4 : p->v;
  P3 x;
  double jac;
  bilinear(map, xs[q / order], xs[q % order], x, jac);
  double w = ws[q / order] * ws[q % order] * jac;
  if (rank) {
    double sum = 0.0;
    for (int a = 0; a < order; ++a)
      for (int b = 0; b < order; ++b) {
        P3 y;
        double jb;
        bilinear(map, xs[a], xs[b], y, jb);
        sum += ws[a] * ws[b] * jb;
      }
    w *= p->area / sum;
  }
  out[t] = make_float4((float)(x.x - c.x), (float)(x.y - c.y), (float)(x.z - c.z), (float)w);
}

namespace {

// the highest of the lanes' orders (0 where no lane is active)
__device__ __forceinline__ int wave_max_order(int o) {
  int m = 0;
  for (int v = 3; v <= 6; ++v)
    if (__ballot(o == v) != 0) m = v;
  return m;
}

// Rankine source and dipole of the tier rule of ORDER of a panel at d = x - c (unrolled, so that
// the rule's scalar loads go several at a time)
template <int ORDER>
__device__ __forceinline__ void rankine_tier32_o(const float4* __restrict__ R, float dx, float dy, float dz, float nx,
                                                 float ny, float nz, float& S, float& D) {
  const float4* g = R + tier_rankine_off(ORDER);
  float s = 0.0f, d = 0.0f;
#pragma unroll
  for (int q = 0; q < ORDER * ORDER; ++q) {
    const float4 a = g[q];
    const float rx = dx - a.x, ry = dy - a.y, rz = dz - a.z;
    const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float nr = fmaf(nx, rx, fmaf(ny, ry, nz * rz));
    s = fmaf(a.w, ir, s);
    d = fmaf(a.w * nr, ir * ir * ir, d);
  }
  S = s;
  D = d;
}

__device__ __forceinline__ void rankine_tier32(const float4* __restrict__ R, int order, float dx, float dy, float dz,
                                               float nx, float ny, float nz, float& S, float& D) {
  switch (order) {
    case 3: rankine_tier32_o<3>(R, dx, dy, dz, nx, ny, nz, S, D); break;
    case 4: rankine_tier32_o<4>(R, dx, dy, dz, nx, ny, nz, S, D); break;
    case 5: rankine_tier32_o<5>(R, dx, dy, dz, nx, ny, nz, S, D); break;
    default: rankine_tier32_o<6>(R, dx, dy, dz, nx, ny, nz, S, D); break;
  }
}

// The log corrections dl = int psi dS - A psi(c) and dd = int d psi / dn dS - A dpsi(c) by the
// log rule of ORDER of a panel, the image at e = image - c, in difference form: the sums of
// w (psi_i - psi_c) and w (dpsi_i - dpsi_c) (the rule's weights sum to the area), each difference
// computed without cancellation from the rule point a = xi - c:
//   R_i - R_c = (a.a - 2 a.e) / (R_i + R_c),  den_i - den_c = (R_i - R_c) - a_z  (den = R - d_z),
//   psi_i - psi_c = log1p((den_i - den_c) / den_c),
//   dpsi = g / den with g = d.n / R - n_z:  g_i - g_c = a.n / R_i + (e.n) (R_i - R_c) / (R_i R_c),
//   dpsi_i - dpsi_c = ((g_i - g_c) den_c - g_c (den_i - den_c)) / (den_i den_c),
// so that FP32 keeps the corrections to their own precision (the direct FP32 difference of the two
// terms loses it to cancellation where the corrections are small beside A psi(c)).
template <int ORDER>
__device__ __forceinline__ void log_tier32_o(const float4* __restrict__ R, float ex, float ey, float ez, float nx,
                                             float ny, float nz, float& dl, float& dd) {
  const float4* g = R + tier_log_off(ORDER);
  const float Rc = __builtin_amdgcn_sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez)));
  const float denc = Rc + ez;  // at the centroid d = -e, den = R_c + e_z
  const float idenc = __builtin_amdgcn_rcpf(denc), iRc = __builtin_amdgcn_rcpf(Rc);
  const float en = fmaf(nx, ex, fmaf(ny, ey, nz * ez));
  const float gc = -en * iRc - nz;
  float l = 0.0f, ld = 0.0f;
#pragma unroll
  for (int q = 0; q < ORDER * ORDER; ++q) {
    const float4 a = g[q];
    const float rx = a.x - ex, ry = a.y - ey, rz = a.z - ez;  // xi - image, rz <= 0
    const float Ri = __builtin_amdgcn_sqrtf(fmaf(rx, rx, fmaf(ry, ry, rz * rz)));
    const float iRi = __builtin_amdgcn_rcpf(Ri);
    const float den = Ri - rz;
    const float aa = fmaf(a.x, a.x - 2.0f * ex, fmaf(a.y, a.y - 2.0f * ey, a.z * (a.z - 2.0f * ez)));
    const float dR = aa * __builtin_amdgcn_rcpf(Ri + Rc);
    const float delta = dR - a.z;
    l = fmaf(a.w, log1pf(delta * idenc), l);
    const float gg = fmaf(fmaf(nx, a.x, fmaf(ny, a.y, nz * a.z)), iRi, en * dR * iRi * iRc);
    ld = fmaf(a.w, fmaf(gg, denc, -gc * delta) * __builtin_amdgcn_rcpf(den * denc), ld);
  }
  dl = l;
  dd = ld;
}

__device__ __forceinline__ int rankine_tier_order(float r2) {  // r2: (distance / radius)^2
  return r2 >= 64.0f ? 3 : r2 >= 16.0f ? 4 : r2 >= 9.0f ? 5 : 6;
}

}  // namespace

__global__ void __launch_bounds__(256) statics_far32_kernel(const PanelD* __restrict__ P, int N, int logint,
                                                            double tol, double depth,
                                                            const float4* __restrict__ rules,
                                                            float4* __restrict__ st32) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + threadIdx.y);
  if (j >= N) return;  // wavefront-uniform: every lane below takes part in the ballots
  const bool live = k < N;
  const PanelD* p = P + j;
  const P3 x = ld3(P[live ? k : 0].c);
  const P3 c = ld3(p->c);
  const double rad = p->radius;
  const P3 d0 = sub(x, c);
  const bool act = live && dot3(d0, d0) >= 4.0 * rad * rad;
  const float4* R = rules + (size_t)j * TIER_STRIDE;
  const float nx = (float)p->n[0], ny = (float)p->n[1], nz = (float)p->n[2];
  const float dx = (float)d0.x, dy = (float)d0.y;
  const float irad2 = (float)(1.0 / (rad * rad));
  // the image points' z - c_z: the field point, its free-surface image and, in finite depth, the
  // bottom images of lo_static.rs:29-38 in the reference's order
  const int nim = depth > 0.0 ? 6 : 2;
  // sums in the reference's order: (s0 + sb) + (((s2 + s4) + s5) + s6)
  float sa = 0.0f, da = 0.0f, sbt = 0.0f, dbt = 0.0f;
#pragma unroll 1
  for (int m = 0; m < nim; ++m) {
    const double h = depth;
    const double zi = m == 0 ? x.z : m == 1 ? -x.z : m == 2 ? -x.z - 2.0 * h : m == 3 ? x.z + 2.0 * h
                    : m == 4 ? x.z - 2.0 * h : -x.z - 4.0 * h;
    const float dz = (float)(zi - c.z);
    const float r2 = (fmaf(dx, dx, fmaf(dy, dy, dz * dz))) * irad2;
    const int order = wave_max_order(act ? rankine_tier_order(r2) : 0);
    float si = 0.0f, di = 0.0f;
    if (order > 0) rankine_tier32(R, order, dx, dy, dz, nx, ny, nz, si, di);
    if (m < 2) {
      sa += si;
      da += di;
    } else {
      sbt += si;
      dbt += di;
    }
  }
  float4 e;
  e.x = nim == 6 ? sa + sbt : sa;
  e.y = nim == 6 ? da + dbt : da;
  e.z = 0.0f;
  e.w = 0.0f;
  // the log corrections dl = int psi dS - A psi(c), dd likewise for d psi / dn, psi at the image
  // (x, y, -z); beyond two radii log_order takes 6, 5 or 4 points a side (6 from two radii up)
  const P3 xb = mk(x.x, x.y, -x.z);
  const bool lg = act && logint && log_pair_d(p, x, tol);
  int lo = 0;
  if (lg) {
    const double r = nrm(sub(xb, c)) / rad;
    lo = r >= 6.0 ? 4 : r >= 3.0 ? 5 : 6;
  }
  const int order = wave_max_order(lo);
  if (order > 0) {
    const float ex = dx, ey = dy, ez = (float)(xb.z - c.z);  // the image relative to the centroid
    float dl = 0.0f, dd = 0.0f;
    switch (order) {
      case 4: log_tier32_o<4>(R, ex, ey, ez, nx, ny, nz, dl, dd); break;
      case 5: log_tier32_o<5>(R, ex, ey, ez, nx, ny, nz, dl, dd); break;
      default: log_tier32_o<6>(R, ex, ey, ez, nx, ny, nz, dl, dd); break;
    }
    if (lg) {
      e.z = dl;
      e.w = dd;
    }
  }
  if (!act) e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (live) st32[(size_t)j * N + k] = e;
}

// One wavefront per queued pair: the adaptive 4 x 4 Gauss rule of panel_integrals.rs:69-140 /
// :436-449, depth-first in the reference's order of the quarters, to at most 14 levels.
constexpr int ASTACK = 48;  // 3 * 14 + 1 entries suffice for a depth-first quadrisection of depth 14

__global__ void __launch_bounds__(64) adaptive_kernel(const PanelD* __restrict__ P, const PanelD* __restrict__ Pf, int N,
                                                      const double* __restrict__ gl, StaticD* __restrict__ st,
                                                      const int2* __restrict__ queue, unsigned count) {
  __shared__ double stk[ASTACK][7];  // u0, u1, v0, v1, prev L, prev Ld, depth
  __shared__ int top;
  const int lane = threadIdx.x;
  const unsigned item = blockIdx.x;
  if (item >= count) return;
  const int2 kj = queue[item];
  const int k = kj.x, j = kj.y;
  const PanelD* p = P + j;
  const P3 x = ld3(Pf[k].c);
  const P3 xb = mk(x.x, x.y, -x.z);
  const P3 n = ld3(p->n);
  const double dist = nrm(sub(xb, ld3(p->c)));
  const double dmin = fmax(dist - p->radius, 1e-3 * p->radius);
  const double scale = fabs(log(dist + p->radius)) + 1.0;
  const double tol_l = 1e-13 * p->area * scale, tol_d = 1e-13 * p->area / dmin;
  const double* x4 = gl + 4 * 32;
  const double* w4 = x4 + 16;
  const int q = lane >> 4, gi = (lane >> 2) & 3, gj = lane & 3;
  const double(*map)[3] = p->v;

  // the 4 x 4 rule on cell (a0, a1) x (b0, b1), summed over the 16 lanes of each quarter group
  auto cell_point = [&](double a0, double a1, double b0, double b1, double& l, double& ld) {
    const double hu = 0.5 * (a1 - a0), hv = 0.5 * (b1 - b0);
    P3 xi;
    double jac, psi, dpsi;
    bilinear(map, a0 + hu * (x4[gi] + 1.0), b0 + hv * (x4[gj] + 1.0), xi, jac);
    psi_terms_d(xi, xb, n, psi, dpsi);
    const double ww = w4[gi] * w4[gj] * hu * hv * jac;
    l = ww * psi;
    ld = ww * dpsi;
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) {
      l += __shfl_xor(l, off, 16);
      ld += __shfl_xor(ld, off, 16);
    }
  };
  double wl, wld;
  cell_point(-1.0, 1.0, -1.0, 1.0, wl, wld);
  wl = __shfl(wl, 0);
  wld = __shfl(wld, 0);
  if (lane == 0) {
    stk[0][0] = -1.0;
    stk[0][1] = 1.0;
    stk[0][2] = -1.0;
    stk[0][3] = 1.0;
    stk[0][4] = wl;
    stk[0][5] = wld;
    stk[0][6] = 0.0;
    top = 1;
  }
  __syncthreads();
  double accL = 0.0, accD = 0.0;
  while (top > 0) {
    const int t = top - 1;
    const double u0 = stk[t][0], u1 = stk[t][1], v0 = stk[t][2], v1 = stk[t][3];
    const double pl = stk[t][4], pd = stk[t][5];
    const int depth = (int)stk[t][6];
    __syncthreads();
    const double um = 0.5 * (u0 + u1), vm = 0.5 * (v0 + v1);
    // quarter q: (u0,um,v0,vm), (um,u1,v0,vm), (um,u1,vm,v1), (u0,um,vm,v1)
    const double a0 = (q == 1 || q == 2) ? um : u0, a1 = (q == 1 || q == 2) ? u1 : um;
    const double b0 = (q >= 2) ? vm : v0, b1 = (q >= 2) ? v1 : vm;
    double cl, cd;
    cell_point(a0, a1, b0, b1, cl, cd);
    const double c0l = __shfl(cl, 0), c1l = __shfl(cl, 16), c2l = __shfl(cl, 32), c3l = __shfl(cl, 48);
    const double c0d = __shfl(cd, 0), c1d = __shfl(cd, 16), c2d = __shfl(cd, 32), c3d = __shfl(cd, 48);
    const double sl = c0l + c1l + c2l + c3l, sd = c0d + c1d + c2d + c3d;
    const bool done = fabs(sl - pl) <= tol_l && fabs(sd - pd) <= tol_d;
    if (depth >= 14 || done) {
      accL += sl;
      accD += sd;
      if (lane == 0) top = t;
    } else if (lane == 0) {
      // push the quarters in reverse order so that quarter 0 is refined first
      const double qa0[4] = {u0, um, um, u0}, qa1[4] = {um, u1, u1, um};
      const double qb0[4] = {v0, v0, vm, vm}, qb1[4] = {vm, vm, v1, v1};
      const double ql[4] = {c0l, c1l, c2l, c3l}, qd[4] = {c0d, c1d, c2d, c3d};
      int s = t;
      for (int r = 3; r >= 0 && s < ASTACK; --r, ++s) {
        stk[s][0] = qa0[r];
        stk[s][1] = qa1[r];
        stk[s][2] = qb0[r];
        stk[s][3] = qb1[r];
        stk[s][4] = ql[r];
        stk[s][5] = qd[r];
        stk[s][6] = depth + 1;
      }
      top = s;
    }
    __syncthreads();
  }
  if (lane == 0) {
    double psi, dpsi;
    psi_terms_d(ld3(p->c), xb, n, psi, dpsi);
    StaticD& e = st[(size_t)j * N + k];
    e.dl = accL - p->area * psi;
    e.dd = accD - p->area * dpsi;
  }
}

// field points: the centroids of the N records Pf (the panels themselves, or the points of the
// field values); sources: the first Nsrc panels of P (a symmetric body or walls: its panels and
// their images); st[j * N + k] for the pair (field k, source j)
void launch_statics(const PanelD* P, const PanelD* Pf, int N, int Nsrc, int logint, double tol, double depth,
                    const double* gl, StaticD* st, int2* queue, unsigned* qcount, unsigned qcap, double2* inf,
                    hipStream_t stream) {
  dim3 block(64, 4);
  dim3 grid((N + 63) / 64, (Nsrc + 3) / 4);
  statics_kernel<<<grid, block, 0, stream>>>(P, Pf, N, Nsrc, logint, tol, depth, gl, st, queue, qcount, qcap, inf);
  HBG_HIP(hipGetLastError());
}

void launch_adaptive(const PanelD* P, const PanelD* Pf, int N, const double* gl, StaticD* st, const int2* queue,
                     unsigned count, hipStream_t stream) {
  if (count == 0) return;
  adaptive_kernel<<<count, 64, 0, stream>>>(P, Pf, N, gl, st, queue, count);
  HBG_HIP(hipGetLastError());
}

void launch_statics_list(const PanelD* P, int N, const int2* ent, int nnz, int logint, double tol, double depth,
                         const double* gl, StaticD* st, int2* queue, unsigned* qcount, unsigned qcap,
                         hipStream_t stream) {
  if (nnz <= 0) return;
  statics_list_kernel<<<(nnz + 255) / 256, 256, 0, stream>>>(P, N, ent, nnz, logint, tol, depth, gl, st, queue, qcount,
                                                             qcap);
  HBG_HIP(hipGetLastError());
}

size_t tier_rules_elems(int N) { return (size_t)N * TIER_STRIDE; }

// rules: tier_rules_elems(N) float4 of workspace; st32[j * N + k] for every pair
void launch_statics_far32(const PanelD* P, int N, int logint, double tol, double depth, const double* gl,
                          float4* rules, float4* st32, hipStream_t stream) {
  tier_rules_kernel<<<N, 192, 0, stream>>>(P, N, gl, rules);
  HBG_HIP(hipGetLastError());
  statics_far32_kernel<<<dim3((N + 63) / 64, (N + 3) / 4), dim3(64, 4), 0, stream>>>(P, N, logint, tol, depth, rules,
                                                                                     st32);
  HBG_HIP(hipGetLastError());
}

}  // namespace hbg
