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

}  // namespace

// gl: Gauss-Legendre nodes and weights [order 0..16][x 16 | w 16]. inf (optional, infinite depth):
// the omega = infinity row of the pair, the integrals of 1/r - 1/r' (lo_static.rs:89-96), at
// inf[j * N + k] as (source, dipole).
__global__ void __launch_bounds__(256) statics_kernel(const PanelD* __restrict__ P, const PanelD* __restrict__ Pf,
                                                      int N, int Nsrc, int logint,
                                                      double tol, double depth, const double* __restrict__ gl,
                                                      StaticD* __restrict__ st, int2* __restrict__ queue,
                                                      unsigned* __restrict__ qcount, unsigned qcap,
                                                      double2* __restrict__ inf) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + threadIdx.y);
  if (j >= Nsrc || k >= N) return;
  const PanelD* p = P + j;
  const P3 x = ld3(Pf[k].c);
  double s0, d0, sb, db;
  rankine_d(p, x, gl, s0, d0);
  const P3 xb = mk(x.x, x.y, -x.z);
  rankine_d(p, xb, gl, sb, db);
  StaticD e;
  e.rs = s0 + sb;
  e.rd = d0 + db;
  if (inf) inf[(size_t)j * N + k] = make_double2(s0 - sb, d0 - db);
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
  }
  e.dl = 0.0;
  e.dd = 0.0;
  if (logint && log_pair_d(p, x, tol)) {
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
      const unsigned slot = atomicAdd(qcount, 1u);
      if (slot < qcap) queue[slot] = make_int2(k, j);
    }
  }
  st[(size_t)j * N + k] = e;
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
  if (inf && depth > 0.0) throw Error{HBG_ERR_UNSUPPORTED, "the omega = infinity statics in finite depth"};
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

}  // namespace hbg
