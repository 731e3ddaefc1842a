// One frequency of the low-order method on the GPU, FP64: the boundary values, the influence
// matrices, the right-hand sides S U and the forces (lo_frequency.rs:214-273, 399-537, 823-955,
// 1039-1171; low_order.rs:627-648, 717-790, 1401-1485), for wetted panels without symmetry
// (one character, chi = 1) in infinite depth.
#include "green_fin_dev.h"
#include "green_inf_dev.h"
#include "hbg_internal.h"
#include "pair_sort.h"

namespace hbg {

void launch_near_rows(const int* rowptr, const int2* ent, const double2* d, size_t dstride, const double2* v,
                      size_t vstride, int N, int R, double sign, double2* out, size_t ostride, int ldo, int accumulate,
                      int nsys, hipStream_t stream);

// The matrices of an ordinary frequency: for the field point k (centroid of panel k) and panel n,
//   S(k, n) = rs + A_n Wr - 2 K logc dl + i A_n Wi
//   D(k, n) = rd + A_n (-WrR a + WrZ n_z) - 2 K logc dd + i A_n (-WiR a + WiZ n_z)
//   A(k, n) = 2 pi delta_kn - D(k, n)
// with the wave kernel at the centroids (one-point rule, low_order.rs:738-769),
// a = ((x_k - c_n)_h . n_h) / R and WrZ the derivative in the source depth. Column-major outputs
// with leading dimension lda.
//
// The kernel of a pair is symmetric: (n, k) has the same value and R-derivative, and its
// source-depth derivative is the field-depth derivative of (k, n) (the same one in infinite
// depth, where the kernel depends on z + zeta). So each unordered pair is evaluated once: a
// workgroup takes a 64 x 64 tile (k-block i <= n-block j; the diagonal tiles the pairs k <= n),
// the 64 lanes of a wavefront along k and 16 n per wavefront, and writes (k, n) (coalesced) and
// (n, k) (one column per lane; the wavefront's consecutive n fill the same cache lines).
// rint, cint: the kind of the row's and of the column's panel (0 body, 1 interior free surface,
// 2 lid). The row's free term is -4 pi on the interior free surface (-2 free_term; Lee, Newman &
// Zhu 1996 eq. 25) and 4 pi on a lid (lo_frequency.rs:85-97); the column's "dipole" is K S on the
// interior free surface (dG/dzeta = K G on zeta = 0, low_order.rs:771-785) and -(K - kappa) S on a
// lid, kappa = w^2 / (g + i w eps) = K / (1 + i eps w / g) with the column panel's eps (ceps) and
// wg = w / g (low_order.rs:811-819, solve_common.rs:52-54); eps = 0 gives kappa = K, a column 0.
__device__ __forceinline__ double2 lid_column_coef(double K, double ceps, double wg) {
  const double e = ceps * wg, d = 1.0 / (1.0 + e * e);
  return make_double2(K * d - K, -K * e * d);  // kappa - K, kappa = K (1 - i e) / (1 + e^2)
}

__device__ __forceinline__ void put_entry(double2* __restrict__ S, double2* __restrict__ A, float2* __restrict__ A32,
                                          int lda, int row, int col, const StaticD& e, double Ar, double nz, double a,
                                          double c2, double Wr, double WrR, double WrZ, double Wi, double WiR,
                                          double WiZ, double free_term, double rint = 0.0, double cint = 0.0,
                                          double K = 0.0, double ceps = 0.0, double wg = 0.0) {
  const double2 sv = make_double2(e.rs + Ar * Wr - c2 * e.dl, Ar * Wi);
  S[(size_t)col * lda + row] = sv;
  double dre = e.rd + Ar * (-WrR * a + WrZ * nz) - c2 * e.dd, dim = Ar * (-WiR * a + WiZ * nz);
  if (cint == 1.0) {
    dre = K * sv.x;
    dim = K * sv.y;
  } else if (cint != 0.0) {
    const double2 c = lid_column_coef(K, ceps, wg);
    dre = c.x * sv.x - c.y * sv.y;
    dim = c.x * sv.y + c.y * sv.x;
  }
  const double ft = rint == 0.0 ? free_term : (rint == 1.0 ? -2.0 * free_term : 2.0 * free_term);
  const double2 v = make_double2((row == col ? ft : 0.0) - dre, -dim);
  A[(size_t)col * lda + row] = v;
  if (A32) A32[(size_t)col * lda + row] = float2((float)v.x, (float)v.y);
}

// Infinite depth: the pairs sorted by the V tile of (X, Y) = (K R, -K (z + zeta)) (then also
// their 1D tile of X is shared); pairs beyond the tables by their 1D tile, or in one bin.
__global__ void __launch_bounds__(256) assemble_kernel(const PanelD* __restrict__ P, const PanelD* __restrict__ Psrc,
                                                       int N, double K, double logc, WaveTables T,
                                                       const StaticD* __restrict__ st, double2* __restrict__ S,
                                                       double2* __restrict__ A, float2* __restrict__ A32, int lda,
                                                       double free_term, double wg) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  if (blockIdx.z) {  // image block g of a symmetric body: sources g n, no free term
    const size_t g = blockIdx.z;
    Psrc += g * N;
    st += g * N * N;
    S += g * N * N;
    A += g * N * lda;
    free_term = 0.0;
  }
  __shared__ PairSort L;
  static_assert(XT * YT + XT + 1 <= PS_KEYS, "infinite-depth keys");
  const unsigned total = sort_tile_pairs(L, bi, bj, N, [&](int k, int n) {
    const PanelD* pk = P + k;
    const PanelD* pn = Psrc + n;
    const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
    const double X = K * sqrt(dx * dx + dy * dy), Y = -K * (pk->c[2] + pn->c[2]);
    if (X >= (double)XT) return XT * YT + XT;
    double tx, ty;
    if (Y >= (double)YT) return XT * YT + tile1d(X, tx);
    return tile2d(X, Y, tx, ty);
  });
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double c2 = 2.0 * K * logc, K2 = K * K;
  const int chunks = (int)((total + 63) / 64);
  for (int ch = wave; ch < chunks; ch += 4) {
    const unsigned pos = (unsigned)ch * 64 + lane;
    const bool valid = pos < total;
    const int p = valid ? L.order[pos] : 0;
    const int k = bi * ASM_TILE + (p & 63), n = bj * ASM_TILE + (p >> 6);
    const PanelD* pk = P + k;
    const PanelD* pn = Psrc + n;
    const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
    const double R = sqrt(dx * dx + dy * dy);
    const double X = K * R, Y = -K * (pk->c[2] + pn->c[2]);
    const WaveK w = wave_kernel_u(T, valid, X, Y);
    if (valid) {
      const double R0 = sqrt(X * X + Y * Y);
      const double FY = -w.F - 2.0 / R0, IY = -w.I;
      double Wr = K * w.F, WrR = K2 * w.FX, WrZ = -K2 * FY;
      double Wi = K * w.I, WiR = K2 * w.IX, WiZ = -K2 * IY;
      if (R == 0.0 && pk->c[2] == 0.0 && pn->c[2] == 0.0) {
        // coincident points on z = 0 (a panel in z = 0 and its own centroid): the regular part
        // K (Ft + 2 psi), psi = log(r' + |z + zeta|), tends to -2 K (gamma - log 2 + log K)
        // (low_order.rs:634-643); the log part is integrated in the statics, the columns of such
        // panels are K S (the derivatives are not used)
        Wr = -2.0 * K * (GAMMA - LOG2 + log(K));
        Wi = -2.0 * PI * K;
        WrR = WrZ = WiR = WiZ = 0.0;
      }
      const double an = R > 0.0 ? (dx * pn->n[0] + dy * pn->n[1]) / R : 0.0;
      put_entry(S, A, A32, lda, k, n, st[(size_t)n * N + k], pn->area, pn->n[2], an, c2, Wr, WrR, WrZ, Wi, WiR, WiZ,
                free_term, pk->interior, pn->interior, K, pn->lid_eps, wg);
      if (k != n) {
        const double ak = R > 0.0 ? -(dx * pk->n[0] + dy * pk->n[1]) / R : 0.0;
        put_entry(S, A, A32, lda, n, k, st[(size_t)k * N + n], pk->area, pk->n[2], ak, c2, Wr, WrR, WrZ, Wi, WiR,
                  WiZ, free_term, pn->interior, pk->interior, K, pk->lid_eps, wg);
      }
    }
  }
}

// The matrices in finite depth: the same entries with the kernel of green_fin_dev.h, evaluated
// at (field x_k, source c_n); (n, k) takes its field-depth derivative Wrz as its source derivative.
//
// The pairs are sorted by table tile: near the axis (R < h/2) by their Phi tiles (R/h, v1/h,
// v2/h), in the series by R/h in bins of 1/32 (then the K0/K1 tiles of a root span one or two
// unit tiles among a wavefront's lanes).
constexpr int FS_NEAR = FTS * FTW * FTV;  // keys of the near pairs: (ix, iy1, iy2)

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) assemble_fin_kernel(const PanelD* __restrict__ P,
                                                           const PanelD* __restrict__ Psrc, int N,
                                                           const FinFreq* __restrict__ f,
                                                           FinTables F, WaveTables T, double logc,
                                                           const StaticD* __restrict__ st, double2* __restrict__ S,
                                                           double2* __restrict__ A, float2* __restrict__ A32,
                                                           int lda, double free_term, double wg) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  if (blockIdx.z) {  // image block g of a symmetric body: sources g n, no free term
    const size_t g = blockIdx.z;
    Psrc += g * N;
    st += g * N * N;
    S += g * N * N;
    A += g * N * lda;
    free_term = 0.0;
  }
  __shared__ PairSort L;
  static_assert(FS_NEAR + 1024 <= PS_KEYS, "finite-depth keys");
  const double h = f->h, ih = 1.0 / h;
  const unsigned total = sort_tile_pairs(L, bi, bj, N, [&](int k, int n) {
    const PanelD* pk = P + k;
    const PanelD* pn = Psrc + n;
    const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
    const double R = sqrt(dx * dx + dy * dy), z = pk->c[2], zeta = pn->c[2];
    if (R >= 0.5 * h) {
      const double b = (R * ih - 0.5) * 32.0;
      return FS_NEAR + (b < 1023.0 ? (int)b : 1023);
    }
    double tx, ty;
    const int i1 = phi_index(FTW, 1.0, R * ih, fabs(z - zeta) * ih, tx, ty);  // ix FTW + iy1
    const int i2 = phi_index(FTV, 2.0, R * ih, (z + zeta + 2.0 * h) * ih, tx, ty);  // ix FTV + iy2
    return i1 * FTV + i2 % FTV;
  });
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double c2 = 2.0 * f->K * logc;
  const int chunks = (int)((total + 63) / 64);
  for (int ch = wave; ch < chunks; ch += 4) {
    const unsigned pos = (unsigned)ch * 64 + lane;
    const bool valid = pos < total;
    const int p = valid ? L.order[pos] : 0;
    const int k = bi * ASM_TILE + (p & 63), n = bj * ASM_TILE + (p >> 6);
    const PanelD* pk = P + k;
    const PanelD* pn = Psrc + n;
    const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
    const double R = sqrt(dx * dx + dy * dy);
    const FinK w = fin_kernel_u(f, F, T, valid, k, n, R, pk->c[2], pn->c[2]);
    if (valid) {
      const double an = R > 0.0 ? (dx * pn->n[0] + dy * pn->n[1]) / R : 0.0;
      put_entry(S, A, A32, lda, k, n, st[(size_t)n * N + k], pn->area, pn->n[2], an, c2, w.Wr, w.WrR, w.WrZ, w.Wi,
                w.WiR, w.WiZ, free_term, pk->interior, pn->interior, f->K, pn->lid_eps, wg);
      if (k != n) {
        const double ak = R > 0.0 ? -(dx * pk->n[0] + dy * pk->n[1]) / R : 0.0;
        put_entry(S, A, A32, lda, n, k, st[(size_t)k * N + n], pk->area, pk->n[2], ak, c2, w.Wr, w.WrR, w.Wrz, w.Wi,
                  w.WiR, w.Wiz, free_term, pn->interior, pk->interior, f->K, pk->lid_eps, wg);
      }
    }
  }
}

// The finite-depth kernel at n point pairs for the tests: the point depths z[i] at "panel" i and
// zeta[i] at "panel" n + i of the per-point factors; out [n][8] = Wr, WrR, WrZ, Wrz, Wi, WiR, WiZ, Wiz
__global__ void fin_points_eval_kernel(const FinFreq* __restrict__ f, FinTables F, WaveTables T, const double* __restrict__ R,
                                       const double* __restrict__ z, const double* __restrict__ zeta, int n,
                                       double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FinK w = fin_kernel(f, F, T, i, n + i, R[i], z[i], zeta[i]);
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

// The wave kernel at points (X, Y) for the tests: out[7 i ..] = Ft, Ft_X, Ft_Y, Gi, Gi_X, Gi_Y, e^-Y J0
__global__ void wave_points_kernel(WaveTables T, const double* __restrict__ X, const double* __restrict__ Y, long n,
                                   double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = X[i], y = Y[i];
  const WaveK w = wave_kernel(T, x, y);
  const double R0 = sqrt(x * x + y * y);
  double* o = out + 7 * i;
  o[0] = w.F;
  o[1] = w.FX;
  o[2] = -w.F - 2.0 / R0;
  o[3] = w.I;
  o[4] = w.IX;
  o[5] = -w.I;
  o[6] = -w.I / (2.0 * PI);
}

// Incident wave and boundary values (lo_frequency.rs:214-273, low_order.rs:1458-1468, frame at
// the origin): phi_I = (i g / w) e^{k z} e^{-i k (x cos b + y sin b)}; u = n_j on the body of the
// radiation column, -dphi_I/dn for each heading (the scattered formulation).
// phiI, dphiI: [H][N]; u: [R][N], R = nrad + H, nrad = 6 nbodies. walls (bit 0: x = 0, bit 1:
// y = 0): the incident wave includes its reflections in the walls, the sum of the wave at the
// reflected points with the gradient reflected back (low_order.rs:1470-1486).
__global__ void boundary_kernel(const PanelD* __restrict__ P, const int* __restrict__ body,
                                const double* __restrict__ nj, int N, int nrad, int H, double w, double k, double grav,
                                double depth, const double* __restrict__ heads, int walls, double2* __restrict__ phiI,
                                double2* __restrict__ dphiI, double2* __restrict__ u) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const PanelD* p = P + j;
  const double x = p->c[0], y = p->c[1], z = p->c[2];
  for (int c = 0; c < nrad; ++c)
    u[(size_t)c * N + j] = make_double2(body[j] == c / 6 ? nj[6 * (size_t)j + c % 6] : 0.0, 0.0);
  // depth factor (low_order.rs:1401-1413): Z = cosh k(z + h) / cosh kh, e^{k z} in infinite depth
  double Z = exp(k * z), Zz = k * Z;
  if (depth > 0.0) {
    const double b = exp(-k * (z + 2.0 * depth)), dd = 1.0 + exp(-2.0 * k * depth);
    Zz = k * (Z - b) / dd;
    Z = (Z + b) / dd;
  }
  for (int h = 0; h < H; ++h) {
    double sb, cb;
    sincos(heads[h], &sb, &cb);
    const double gw = grav / w;
    double2 phi = make_double2(0.0, 0.0), g0 = phi, g1 = phi, g2 = phi;
    for (int r = 0; r < 4; ++r) {  // the identity, then the wall reflections x, y, xy present
      if ((r & walls) != r) continue;
      const double sx = (r & 1) ? -1.0 : 1.0, sy = (r & 2) ? -1.0 : 1.0;
      const double ph = -k * (sx * x * cb + sy * y * sb);
      double sp, cp;
      sincos(ph, &sp, &cp);
      // e = (i g / w) (cos ph + i sin ph)
      const double2 e = make_double2(-gw * sp, gw * cp);
      const double2 f = make_double2(e.x * Z, e.y * Z);
      phi.x += f.x;
      phi.y += f.y;
      // gradient: phi (-i k cb), phi (-i k sb), e Zz at the reflected point, reflected back
      g0.x += sx * f.y * k * cb;
      g0.y -= sx * f.x * k * cb;
      g1.x += sy * f.y * k * sb;
      g1.y -= sy * f.x * k * sb;
      g2.x += e.x * Zz;
      g2.y += e.y * Zz;
    }
    double2 dn = make_double2(g0.x * p->n[0] + g1.x * p->n[1] + g2.x * p->n[2],
                              g0.y * p->n[0] + g1.y * p->n[1] + g2.y * p->n[2]);
    if (p->interior == 1.0) dn = make_double2(0.0, 0.0);  // the interior free surface carries no velocity
    if (p->interior == 2.0) {
      // a lid: the condition phi_z - kappa phi = 0 of the total potential leaves the boundary value
      // u = (K - kappa) phi_I of the scattered one (lo_frequency.rs:263-268), carried here as
      // dphi_I/dn = -(K - kappa) phi_I (then u = -dphi_I/dn, and the Haskind sum's -phi_c dphi_I/dn
      // is the lid term (K - kappa) phi_c phi_I of lo_frequency.rs:1150-1156)
      const double K = w * w / grav, e = p->lid_eps * w / grav, d = 1.0 / (1.0 + e * e);
      const double cx = K * d - K, cy = -K * e * d;  // kappa - K
      dn = make_double2(cx * phi.x - cy * phi.y, cx * phi.y + cy * phi.x);
    }
    phiI[(size_t)h * N + j] = phi;
    dphiI[(size_t)h * N + j] = dn;
    u[(size_t)(nrad + h) * N + j] = make_double2(-dn.x, -dn.y);
  }
}

// rhs(:, c) = S U(:, c), U = -u (lo_frequency.rs:855-893), in two passes: every workgroup
// multiplies a block of 256 rows of S by one chunk of columns (one row per thread, S read column
// by column, coalesced along the rows, u wavefront-uniform), the partial sums [chunk][c][row] are
// then added in chunk order (deterministic), into the augmented columns of M. Several systems
// (the characters of a symmetric body) go in one launch, one per blockIdx.z. Chunks of 128
// columns up to order 4096 and 256 above keep at least 256 workgroups in flight from order 2048.
constexpr int MAXRHS = 32;
int rhs_chunk_len(int N) { return N <= 4096 ? 128 : 256; }
// RM: right-hand sides held in registers (R <= RM); the columns of S go 8 at a time so that eight
// loads are in flight per lane (the kernel streams S once; its rate is that of the loads in flight),
// and the wavefront-uniform u of 64 columns at a time is staged in LDS and read as broadcasts.
constexpr int RHS_SUB = 64;
__device__ __forceinline__ double2 to_double2(double2 v) { return v; }
__device__ __forceinline__ double2 to_double2(float2 v) { return make_double2((double)v.x, (double)v.y); }

// ST: the element type of S (complex128, or complex64 from the FP32 assembly; the sums in FP64)
template <int RM, class ST = double2>
__global__ void __launch_bounds__(256) rhs_partial_kernel(const ST* __restrict__ S, size_t sstride, int N,
                                                          int lds, const double2* __restrict__ u, size_t ustride,
                                                          int R, int chunk_len, double2* __restrict__ part,
                                                          size_t pstride) {
  __shared__ double2 us[RHS_SUB][RM];
  S += blockIdx.z * sstride;
  u += blockIdx.z * ustride;
  part += blockIdx.z * pstride;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  const bool row = k < N;
  const int kk = row ? k : N - 1;
  double2 acc[RM];
#pragma unroll
  for (int c = 0; c < RM; ++c) acc[c] = make_double2(0.0, 0.0);
  auto column = [&](double2 s, int j) {
#pragma unroll
    for (int c = 0; c < RM; ++c) {
      if (c < R) {
        const double2 v = us[j][c];
        acc[c].x -= s.x * v.x - s.y * v.y;
        acc[c].y -= s.x * v.y + s.y * v.x;
      }
    }
  };
  const int l0 = chunk * chunk_len, l1 = min(N, l0 + chunk_len);
  for (int s0 = l0; s0 < l1; s0 += RHS_SUB) {
    const int ns = min(RHS_SUB, l1 - s0);
    __syncthreads();
    for (int e = threadIdx.x; e < RHS_SUB * R; e += blockDim.x) {
      const int c = e / RHS_SUB, j = e % RHS_SUB;
      if (j < ns) us[j][c] = u[(size_t)c * N + s0 + j];
    }
    __syncthreads();
    int j = 0;
    for (; j + 8 <= ns; j += 8) {
      ST sv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) sv[q] = S[(size_t)(s0 + j + q) * lds + kk];
#pragma unroll
      for (int q = 0; q < 8; ++q) column(to_double2(sv[q]), j + q);
    }
    for (; j < ns; ++j) column(to_double2(S[(size_t)(s0 + j) * lds + kk]), j);
  }
  if (row)
#pragma unroll
    for (int c = 0; c < RM; ++c)
      if (c < R) part[((size_t)chunk * R + c) * N + k] = acc[c];
}

// the sums into column col0 + c of M (leading dimension ldm)
__global__ void rhs_reduce_kernel(const double2* __restrict__ part, size_t pstride, int N, int R, int chunks,
                                  double2* __restrict__ M, size_t mstride, int ldm, int col0) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)R * N) return;
  part += blockIdx.y * pstride;
  M += blockIdx.y * mstride;
  const int c = (int)(e / N), k = (int)(e % N);
  double2 s = make_double2(0.0, 0.0);
  for (int q = 0; q < chunks; ++q) {
    const double2 v = part[((size_t)q * R + c) * N + k];
    s.x += v.x;
    s.y += v.y;
  }
  M[(size_t)(col0 + c) * ldm + k] = s;
}

// A system of order nz stored at order n (leading dimension n) among systems of uneven order:
// the identity on rows and columns [nz, n), zeros elsewhere in them, zero rows [nz, n) of the
// right-hand sides [0, rz) and zero right-hand sides [rz, R) (columns n + c). Its LU is then
// block-diagonal and its solution zero on [nz, n). M32 (optional): the matrix part in complex64.
// One workgroup per column.
__global__ void __launch_bounds__(256) pad_system_kernel(double2* __restrict__ M, float2* __restrict__ M32, int n,
                                                         int nz, int R, int rz) {
  const int c = blockIdx.x;
  if (c >= n + R) return;
  const int lo = (c < nz || (c >= n && c - n < rz)) ? nz : 0;
  for (int r = lo + threadIdx.x; r < n; r += 256) {
    const double v = (c == r) ? 1.0 : 0.0;
    M[(size_t)c * n + r] = make_double2(v, 0.0);
    if (M32 && c < n) M32[(size_t)c * n + r] = float2((float)v, 0.0f);
  }
}

// The forces of one frequency (lo_frequency.rs:1039-1171), each output a sum over the panels by
// one workgroup (fixed-order tree reduction: results independent of timing).
// phi: the solved potentials, column c at M[(N + c) * ldm]. Outputs (double, complex interleaved):
// A, B [D][nrad]; X, FK, SC [H][D]; HK [nrad][H]; D = nrad = 6 nbodies.
__global__ void __launch_bounds__(256) forces_kernel(const PanelD* __restrict__ P, const int* __restrict__ body,
                                                     const double* __restrict__ nj, int N, int nrad, int H,
                                                     double w, double rho, const double2* __restrict__ sol, int lds,
                                                     const double2* __restrict__ phiI,
                                                     const double2* __restrict__ dphiI, double* __restrict__ out) {
  __shared__ double2 red[256];
  __shared__ double2 red2[256];
  const int D = nrad;
  const int nAB = D * nrad, nX = H * D;
  const int item = blockIdx.x;
  const int tid = threadIdx.x;
  double2 s = make_double2(0.0, 0.0), s2 = make_double2(0.0, 0.0);
  int kind, i = 0, c = 0, h = 0;
  if (item < nAB) {
    kind = 0;
    i = item / nrad;
    c = item % nrad;
  } else if (item < nAB + nX) {
    kind = 1;
    h = (item - nAB) / D;
    i = (item - nAB) % D;
  } else {
    kind = 2;
    c = (item - nAB - nX) / H;
    h = (item - nAB - nX) % H;
  }
  for (int j = tid; j < N; j += blockDim.x) {
    const double area = P[j].area;
    if (kind == 0) {
      if (body[j] == i / 6) {
        const double ni = nj[6 * (size_t)j + i % 6] * area;
        const double2 ph = sol[(size_t)c * lds + j];
        s.x += ni * ph.x;
        s.y += ni * ph.y;
      }
    } else if (kind == 1) {
      if (body[j] == i / 6) {
        const double ni = nj[6 * (size_t)j + i % 6] * area;
        const double2 ps = sol[(size_t)(nrad + h) * lds + j];
        const double2 pi = phiI[(size_t)h * N + j];
        s.x += ni * (ps.x + pi.x);
        s.y += ni * (ps.y + pi.y);
        s2.x += ni * pi.x;
        s2.y += ni * pi.y;
      }
    } else {
      const double ncv = body[j] == c / 6 ? nj[6 * (size_t)j + c % 6] : 0.0;
      const double2 pi = phiI[(size_t)h * N + j];
      const double2 ph = sol[(size_t)c * lds + j];
      const double2 dn = dphiI[(size_t)h * N + j];
      // (n_c phi_I - phi_c dphi_I/dn) A
      s.x += (ncv * pi.x - (ph.x * dn.x - ph.y * dn.y)) * area;
      s.y += (ncv * pi.y - (ph.x * dn.y + ph.y * dn.x)) * area;
    }
  }
  red[tid] = s;
  red2[tid] = s2;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) {
      red[tid].x += red[tid + off].x;
      red[tid].y += red[tid + off].y;
      red2[tid].x += red2[tid + off].x;
      red2[tid].y += red2[tid + off].y;
    }
    __syncthreads();
  }
  if (tid != 0) return;
  const double2 z = red[0], z2 = red2[0];
  // i w rho (x + i y) = (-w rho y, w rho x)
  const double wr = w * rho;
  if (kind == 0) {
    // z = -rho sum n_i phi_c A: A = Re z, B = -w Im z
    out[item] = -rho * z.x;                      // added mass
    out[nAB + item] = -w * (-rho * z.y);         // damping
  } else if (kind == 1) {
    const int o = item - nAB;
    double* X = out + 2 * nAB;
    double* FK = X + 2 * nX;
    double* SC = FK + 2 * nX;
    X[2 * o] = -wr * z.y;
    X[2 * o + 1] = wr * z.x;
    FK[2 * o] = -wr * z2.y;
    FK[2 * o + 1] = wr * z2.x;
    SC[2 * o] = -wr * (z.y - z2.y);
    SC[2 * o + 1] = wr * (z.x - z2.x);
  } else {
    const int o = item - nAB - nX;
    double* HK = out + 2 * nAB + 6 * nX;
    HK[2 * o] = -wr * z.y;
    HK[2 * o + 1] = wr * z.x;
  }
}

// A32 (nullable): A in complex64 too. Psrc: the source panels (P: the block of the panels with
// themselves). images > 1 (a symmetric body): blockIdx.z = g assembles the image block g, the
// sources Psrc + g N (P reflected by g), its statics st + g N N, into S + g N N and A + g N lda,
// with no free term for g > 0; its pairs keep the symmetry used here since |c_k - g c_n| =
// |c_n - g c_k| for a reflection. One launch for all blocks fills the chip at sector orders where
// one block alone would not (528 workgroups at order 2048).
void launch_assemble(const PanelD* P, const PanelD* Psrc, int N, double K, double logc, WaveTables T,
                     const StaticD* st, double2* S, double2* A, float2* A32, int lda, double free_term,
                     hipStream_t stream, int images, double wg) {
  const unsigned nt = (unsigned)((N + ASM_TILE - 1) / ASM_TILE);
  if (images > 1 && A32) throw Error{6, "image blocks without a complex64 copy"};
  assemble_kernel<<<dim3(nt, nt, images), 256, 0, stream>>>(P, Psrc, N, K, logc, T, st, S, A, A32, lda, free_term,
                                                            wg);
  HBG_HIP(hipGetLastError());
}

void launch_assemble_fin(const PanelD* P, const PanelD* Psrc, int N, const FinFreq* f, FinTables F, WaveTables T,
                         double logc, const StaticD* st, double2* S, double2* A, float2* A32, int lda,
                         double free_term, hipStream_t stream, int images, double wg) {
  const unsigned nt = (unsigned)((N + ASM_TILE - 1) / ASM_TILE);
  if (images > 1 && A32) throw Error{6, "image blocks without a complex64 copy"};
  assemble_fin_kernel<<<dim3(nt, nt, images), 256, 0, stream>>>(P, Psrc, N, f, F, T, logc, st, S, A, A32, lda,
                                                                free_term, wg);
  HBG_HIP(hipGetLastError());
}

// Field values (lo_frequency.rs:1392-1495, infinite depth): the influence of panel j at point k,
// S_x and the dipole part as the assembly writes it (A_x = -D_x, no free term: the points are off
// the panels), from the statics rows of the points st[j * Np + k] and the same wave kernel. Not
// the pair-symmetric tiles of the assembly (points and panels differ): 64 points per wavefront
// along k, one panel per wavefront, so the panel record and its tile reads are wavefront-uniform.
__global__ void __launch_bounds__(256) field_influence_kernel(const PanelD* __restrict__ Pf,
                                                              const PanelD* __restrict__ P, int Np, int N, double K,
                                                              double logc, WaveTables T,
                                                              const StaticD* __restrict__ st, double2* __restrict__ Sx,
                                                              double2* __restrict__ Ax) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + threadIdx.y);
  if (j >= N) return;
  const bool valid = k < Np;
  const PanelD* pk = Pf + (valid ? k : 0);
  const PanelD* pn = P + j;
  if (pn->interior != 0.0) {  // wavefront-uniform: the interior free surface is not in the sums
    if (valid) Sx[(size_t)j * Np + k] = Ax[(size_t)j * Np + k] = make_double2(0.0, 0.0);
    return;
  }
  const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
  const double R = sqrt(dx * dx + dy * dy);
  const double X = K * R, Y = -K * (pk->c[2] + pn->c[2]);
  const WaveK w = wave_kernel_u(T, valid, X, Y);
  if (!valid) return;
  const double K2 = K * K, c2 = 2.0 * K * logc;
  const double R0 = sqrt(X * X + Y * Y);
  const double FY = -w.F - 2.0 / R0, IY = -w.I;
  const double Wr = K * w.F, WrR = K2 * w.FX, WrZ = -K2 * FY;
  const double Wi = K * w.I, WiR = K2 * w.IX, WiZ = -K2 * IY;
  const double an = R > 0.0 ? (dx * pn->n[0] + dy * pn->n[1]) / R : 0.0;
  put_entry(Sx, Ax, nullptr, Np, k, j, st[(size_t)j * Np + k], pn->area, pn->n[2], an, c2, Wr, WrR, WrZ, Wi, WiR, WiZ,
            0.0);
}

// The same in finite depth: the evaluator of the assembly (fin_kernel_u) with the per-point factors
// of the tables at N + k (the tables hold the panels, then the points).
__global__ void __launch_bounds__(256) field_influence_fin_kernel(const PanelD* __restrict__ Pf,
                                                                  const PanelD* __restrict__ P, int Np, int N,
                                                                  const FinFreq* __restrict__ f, FinTables F,
                                                                  WaveTables T, double logc,
                                                                  const StaticD* __restrict__ st,
                                                                  double2* __restrict__ Sx, double2* __restrict__ Ax) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + threadIdx.y);
  if (j >= N) return;
  const bool valid = k < Np;
  const int kk = valid ? k : 0;
  const PanelD* pk = Pf + kk;
  const PanelD* pn = P + j;
  if (pn->interior != 0.0) {  // wavefront-uniform: the interior free surface is not in the sums
    if (valid) Sx[(size_t)j * Np + k] = Ax[(size_t)j * Np + k] = make_double2(0.0, 0.0);
    return;
  }
  const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
  const double R = sqrt(dx * dx + dy * dy);
  const FinK w = fin_kernel_u(f, F, T, valid, N + kk, j, R, pk->c[2], pn->c[2]);
  if (!valid) return;
  const double an = R > 0.0 ? (dx * pn->n[0] + dy * pn->n[1]) / R : 0.0;
  put_entry(Sx, Ax, nullptr, Np, k, j, st[(size_t)j * Np + k], pn->area, pn->n[2], an, 2.0 * f->K * logc, w.Wr, w.WrR,
            w.WrZ, w.Wi, w.WiR, w.WiZ, 0.0);
}

// The field values of one frequency at every point k (lo_frequency.rs:1461-1491): radiation column
// c: frad = (w^2 / g) (1 / 4 pi) sum_j (phi_j D_x - S_x u_j) with u the body's n_j, over the body
// panels (an interior free surface's panels have S_x = D_x = 0, lo_frequency.rs:1443-1458); heading h
// (the scattered formulation): phi_S = (1 / 4 pi) sum_j (phi_S + phi_I)_j D_x, and the elevations
// eta = -(i w / g) phi of the incident wave at the point, the scattered wave and their sum.
// sol: the nsol x R solution columns (leading dimension nsol); phiI: [H][nsol] on the panels; N
// source panels, panel j taking the values of j mod nsol (walls: an image carries its panel's
// potential and incident wave, nsol = the real panels); walls (bit 0: x = 0, bit 1: y = 0): the
// incident wave at the point with its reflections; out: [Np][nrad + 3 H], per point frad (nrad),
// finc, fsca, ftot (H each).
__global__ void __launch_bounds__(256) field_sum_kernel(const PanelD* __restrict__ Pf, const int* __restrict__ body,
                                                        const double* __restrict__ nj, int Np, int N, int nsol,
                                                        int nrad, int H, double w, double k, double depth,
                                                        double grav, const double* __restrict__ heads, int walls,
                                                        const double2* __restrict__ Sx,
                                                        const double2* __restrict__ Ax,
                                                        const double2* __restrict__ sol,
                                                        const double2* __restrict__ phiI, double2* __restrict__ out) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int c = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + threadIdx.y);
  const int R = nrad + H;
  if (p >= Np || c >= R) return;
  double vr = 0.0, vi = 0.0;
  const double2* phi = sol + (size_t)c * nsol;
  if (c < nrad) {
    for (int j = 0; j < N; ++j) {
      const double2 a = Ax[(size_t)j * Np + p], sv = Sx[(size_t)j * Np + p], f = phi[j % nsol];
      const double u = body[j] == c / 6 ? nj[6 * (size_t)j + c % 6] : 0.0;
      // phi D - S u with D = -A
      vr += -(f.x * a.x - f.y * a.y) - sv.x * u;
      vi += -(f.x * a.y + f.y * a.x) - sv.y * u;
    }
    const double sc = w * w / grav / (4.0 * PI);
    out[(size_t)p * (nrad + 3 * H) + c] = make_double2(sc * vr, sc * vi);
    return;
  }
  const int h = c - nrad;
  const double2* pI = phiI + (size_t)h * nsol;
  for (int j = 0; j < N; ++j) {
    const int js = j % nsol;
    const double2 a = Ax[(size_t)j * Np + p], f = make_double2(phi[js].x + pI[js].x, phi[js].y + pI[js].y);
    vr += -(f.x * a.x - f.y * a.y);
    vi += -(f.x * a.y + f.y * a.x);
  }
  // phi_I at the point: (i g / w) Z(z) e^{-i k (x cos b + y sin b)}, Z = e^{k z} or, in finite
  // depth, cosh k (z + h) / cosh k h with the real root k
  const PanelD* q = Pf + p;
  double sb, cb;
  sincos(heads[h], &sb, &cb);
  double Z = exp(k * q->c[2]);
  if (depth > 0.0) Z = (Z + exp(-k * (q->c[2] + 2.0 * depth))) / (1.0 + exp(-2.0 * k * depth));
  const double gw = grav / w;
  double2 pi_ = make_double2(0.0, 0.0);
  for (int r = 0; r < 4; ++r) {  // the wave and its reflections in the walls present
    if ((r & walls) != r) continue;
    const double sx = (r & 1) ? -1.0 : 1.0, sy = (r & 2) ? -1.0 : 1.0;
    double sp, cp;
    sincos(-k * (sx * q->c[0] * cb + sy * q->c[1] * sb), &sp, &cp);
    pi_.x -= gw * sp * Z;
    pi_.y += gw * cp * Z;
  }
  const double2 ps = make_double2(vr / (4.0 * PI), vi / (4.0 * PI));
  // eta = -(i w / g) phi
  const double e = w / grav;
  const double2 eI = make_double2(e * pi_.y, -e * pi_.x), eS = make_double2(e * ps.y, -e * ps.x);
  double2* o = out + (size_t)p * (nrad + 3 * H) + nrad;
  o[h] = eI;
  o[H + h] = eS;
  o[2 * H + h] = make_double2(eI.x + eS.x, eI.y + eS.y);
}

// f, F (finite depth, depth > 0): the frequency and its tables with the points' factors at N + k;
// kw: the wave number of the incident wave (K in infinite depth, the real root in finite depth)
void launch_field_values(const PanelD* Pf, const PanelD* P, const int* body, const double* nj, int Np, int N, int nsol,
                         int nrad, int H, double w, double K, double kw, double depth, const FinFreq* f, FinTables F,
                         double logc, WaveTables T, double grav, const double* heads, int walls, const StaticD* stf,
                         double2* Sx, double2* Ax, const double2* sol, const double2* phiI, double2* out,
                         hipStream_t stream) {
  const dim3 block(64, 4), grid((Np + 63) / 64, (N + 3) / 4);
  if (depth > 0.0)
    field_influence_fin_kernel<<<grid, block, 0, stream>>>(Pf, P, Np, N, f, F, T, logc, stf, Sx, Ax);
  else
    field_influence_kernel<<<grid, block, 0, stream>>>(Pf, P, Np, N, K, logc, T, stf, Sx, Ax);
  field_sum_kernel<<<dim3((Np + 63) / 64, (nrad + H + 3) / 4), block, 0, stream>>>(
      Pf, body, nj, Np, N, nsol, nrad, H, w, kw, depth, grav, heads, walls, Sx, Ax, sol, phiI, out);
  HBG_HIP(hipGetLastError());
}

// Kochin functions of one frequency (lo_frequency.rs:1279-1372, far_kernel low_order.rs:1417-1433)
// on the angles theta_t = 2 pi t / NK: Green's identity with the far-field kernel Phi =
// Z(z) e^{i k (x cos + y sin)} over the panels (one point each, like the matrices): radiation
// column c: H = sum_j area_j (phi_j dPhi/dn - Phi u_j), u the body's n_j; heading h (the scattered
// formulation): H = sum_j area_j (phi_S + phi_I)_j dPhi/dn; and the theta derivatives Ht of both.
// N panels, panel j taking the potential of j mod nsol (walls: the images carry their panels').
// One thread per (angle, column); the panel loop reads wavefront-uniform panel records.
// out: [2][NK][R] (H, then Ht).
__global__ void __launch_bounds__(256) kochin_kernel(const PanelD* __restrict__ P, const int* __restrict__ body,
                                                     const double* __restrict__ nj, int N, int nsol, int nrad, int H,
                                                     double k,
                                                     double depth, int NK, const double2* __restrict__ sol,
                                                     const double2* __restrict__ phiI, double2* __restrict__ out) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  const int c = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + threadIdx.y);
  const int R = nrad + H;
  if (t >= NK || c >= R) return;
  double sn, cs;
  sincos(2.0 * PI * t / NK, &sn, &cs);
  const double2* phi = sol + (size_t)c * nsol;
  const double2* pI = c >= nrad ? phiI + (size_t)(c - nrad) * nsol : nullptr;
  double hr = 0.0, hi = 0.0, tr = 0.0, ti = 0.0;
  for (int j = 0; j < N; ++j) {
    const PanelD* p = P + j;
    if (p->interior != 0.0) continue;  // the body panels only (lo_frequency.rs:1311-1313)
    const double x = p->c[0], y = p->c[1], z = p->c[2];
    double Z = exp(k * z), Zz = k * Z;
    if (depth > 0.0) {
      const double b = exp(-k * (z + 2.0 * depth)), d = 1.0 + exp(-2.0 * k * depth);
      Zz = k * (Z - b) / d;
      Z = (Z + b) / d;
    }
    double es, ec;
    sincos(k * (x * cs + y * sn), &es, &ec);
    const double dt = -x * sn + y * cs, dnt = -p->n[0] * sn + p->n[1] * cs;
    const double a = p->area, kn = k * (p->n[0] * cs + p->n[1] * sn);
    // g = Z e, gn = e (i k Z (n . dir) + Zz n_z), gt = g i k dt, gnt = gn i k dt + e Z i k dnt
    const double2 g = make_double2(a * Z * ec, a * Z * es);
    const double2 gn = make_double2(a * (Zz * p->n[2] * ec - Z * kn * es), a * (Zz * p->n[2] * es + Z * kn * ec));
    const double2 gt = make_double2(-k * dt * g.y, k * dt * g.x);
    const double2 gnt = make_double2(-k * dt * gn.y - a * Z * k * dnt * es, k * dt * gn.x + a * Z * k * dnt * ec);
    const int js = j % nsol;
    double2 f = phi[js];
    double u = 0.0;
    if (c < nrad)
      u = body[j] == c / 6 ? nj[6 * (size_t)j + c % 6] : 0.0;
    else
      f = make_double2(f.x + pI[js].x, f.y + pI[js].y);
    hr += f.x * gn.x - f.y * gn.y - u * g.x;
    hi += f.x * gn.y + f.y * gn.x - u * g.y;
    tr += f.x * gnt.x - f.y * gnt.y - u * gt.x;
    ti += f.x * gnt.y + f.y * gnt.x - u * gt.y;
  }
  out[(size_t)t * R + c] = make_double2(hr, hi);
  out[((size_t)NK + t) * R + c] = make_double2(tr, ti);
}

void launch_kochin(const PanelD* P, const int* body, const double* nj, int N, int nsol, int nrad, int H, double k,
                   double depth, int NK, const double2* sol, const double2* phiI, double2* out, hipStream_t stream) {
  kochin_kernel<<<dim3((NK + 63) / 64, (nrad + H + 3) / 4), dim3(64, 4), 0, stream>>>(P, body, nj, N, nsol, nrad, H,
                                                                                     k, depth, NK, sol, phiI, out);
  HBG_HIP(hipGetLastError());
}

// The character systems of a symmetric body from its image blocks (lo_frequency.rs:855-943):
// S_c = sum_g chi_c(g) S_g and the matrix part of M_c = sum_g chi_c(g) A_g for every character c,
// with A32 (nullable) the complex64 copy of M_c, for the first nchar characters (all of them for a
// symmetric body; with walls, the first only: the one even under every wall reflection). One pass:
// each entry of the G blocks is read once and each character entry written once. chi: [G][G] row
// c, column g; S_g, A_g at Sg + g N N, Ag + g N N (leading dimension N); M_c at M + c mstride with
// leading dimension ldm (>= N: a system padded to a batch's larger order), S_c at S + c N N.
template <int G>
__global__ void symmetry_combine_kernel(const double2* __restrict__ Sg, const double2* __restrict__ Ag, int N,
                                        const double* __restrict__ chi, int nchar, double2* __restrict__ S,
                                        double2* __restrict__ M, size_t mstride, int ldm, float2* __restrict__ A32) {
  const size_t NN = (size_t)N * N;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NN) return;
  const size_t em = (e / N) * ldm + e % N;
  double2 sg[G], ag[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    sg[g] = Sg[g * NN + e];
    ag[g] = Ag[g * NN + e];
  }
#pragma unroll
  for (int c = 0; c < G; ++c) {
    if (c >= nchar) break;
    double2 sv = make_double2(0.0, 0.0), mv = make_double2(0.0, 0.0);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const double x = chi[c * G + g];
      sv.x += x * sg[g].x;
      sv.y += x * sg[g].y;
      mv.x += x * ag[g].x;
      mv.y += x * ag[g].y;
    }
    S[c * NN + e] = sv;
    M[c * mstride + em] = mv;
    if (A32) A32[c * mstride + em] = float2((float)mv.x, (float)mv.y);
  }
}

void launch_symmetry_combine(const double2* Sg, const double2* Ag, int N, int G, const double* chi, int nchar,
                             double2* S, double2* M, size_t mstride, int ldm, float2* A32, hipStream_t stream) {
  const unsigned blocks = (unsigned)(((size_t)N * N + 255) / 256);
  if (nchar < 1 || nchar > G) throw Error{6, "characters out of range"};
  if (ldm < N) throw Error{6, "leading dimension below the order"};
  if (G == 2)
    symmetry_combine_kernel<2><<<blocks, 256, 0, stream>>>(Sg, Ag, N, chi, nchar, S, M, mstride, ldm, A32);
  else if (G == 4)
    symmetry_combine_kernel<4><<<blocks, 256, 0, stream>>>(Sg, Ag, N, chi, nchar, S, M, mstride, ldm, A32);
  else
    throw Error{6, "symmetry groups have 2 or 4 elements"};
  HBG_HIP(hipGetLastError());
}

// The right-hand-side values of every character c (lo_frequency.rs:873-880): uc[c][col][n] =
// (1 / G) sum_g chi_c(g) u[col][g N + n] for the R columns of u on the G N panels of the whole body
// (the S U product negates them).
__global__ void symmetry_project_kernel(const double2* __restrict__ u, int N, int G, int R,
                                        const double* __restrict__ chi, double2* __restrict__ uc) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * R * G) return;
  const int n = (int)(e % N), col = (int)(e / N % R), c = (int)(e / ((size_t)N * R));
  double2 v = make_double2(0.0, 0.0);
  for (int g = 0; g < G; ++g) {
    const double x = chi[c * G + g];
    const double2 a = u[(size_t)col * G * N + (size_t)g * N + n];
    v.x += x * a.x;
    v.y += x * a.y;
  }
  uc[e] = make_double2(v.x / G, v.y / G);
}

void launch_symmetry_project(const double2* u, int N, int G, int R, const double* chi, double2* uc,
                             hipStream_t stream) {
  symmetry_project_kernel<<<(unsigned)(((size_t)N * R * G + 255) / 256), 256, 0, stream>>>(u, N, G, R, chi, uc);
  HBG_HIP(hipGetLastError());
}

// The potential on the whole body from the character solutions (lo_frequency.rs:931-943):
// phi[col][g N + n] = sum_c chi_c(g) phi_c[col][n], phi_c the solution columns of character c at
// sol + c mstride (leading dimension lds).
__global__ void symmetry_recombine_kernel(const double2* __restrict__ sol, size_t mstride, int lds, int N, int G, int R,
                                          const double* __restrict__ chi, double2* __restrict__ phi) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)G * N * R) return;
  const int m = (int)(e % ((size_t)G * N)), col = (int)(e / ((size_t)G * N));
  const int g = m / N, n = m % N;
  double2 v = make_double2(0.0, 0.0);
  for (int c = 0; c < G; ++c) {
    const double x = chi[c * G + g];
    const double2 a = sol[(size_t)c * mstride + (size_t)col * lds + n];
    v.x += x * a.x;
    v.y += x * a.y;
  }
  phi[e] = v;
}

void launch_symmetry_recombine(const double2* sol, size_t mstride, int lds, int N, int G, int R, const double* chi,
                               double2* phi, hipStream_t stream) {
  symmetry_recombine_kernel<<<(unsigned)(((size_t)G * N * R + 255) / 256), 256, 0, stream>>>(sol, mstride, lds, N, G,
                                                                                            R, chi, phi);
  HBG_HIP(hipGetLastError());
}

void launch_fin_points_eval(const FinFreq* f, FinTables F, WaveTables T, const double* R, const double* z,
                            const double* zeta, int n, double* out, hipStream_t stream) {
  fin_points_eval_kernel<<<(n + 127) / 128, 128, 0, stream>>>(f, F, T, R, z, zeta, n, out);
  HBG_HIP(hipGetLastError());
}

void launch_wave_points(WaveTables T, const double* X, const double* Y, long n, double* out, hipStream_t stream) {
  wave_points_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(T, X, Y, n, out);
  HBG_HIP(hipGetLastError());
}

void launch_boundary(const PanelD* P, const int* body, const double* nj, int N, int nrad, int H, double w, double k,
                     double grav, double depth, const double* heads, double2* phiI, double2* dphiI, double2* u,
                     hipStream_t stream, int walls) {
  boundary_kernel<<<(N + 255) / 256, 256, 0, stream>>>(P, body, nj, N, nrad, H, w, k, grav, depth, heads, walls,
                                                       phiI, dphiI, u);
  HBG_HIP(hipGetLastError());
}

int rhs_chunks(int N) { return (N + rhs_chunk_len(N) - 1) / rhs_chunk_len(N); }

// The matrices of a frequency limit in infinite depth (low_order.rs:635-637, 757-758: the wave
// terms and the log terms vanish, K = 0): S = rs, D = rd (1/r + 1/r') at omega = 0, and with
// phi = 0 on z = 0 the integrals of 1/r - 1/r' (inf) at omega = infinity; A = 2 pi I - D.
// Column-major, leading dimension N (the layout of the statics): entry (k, n) for the N field
// panels k and the Nsrc sources n (a symmetric body or walls: the panels and their images, the
// image g's block at g N N), the 2 pi only on the diagonal of the first block.
__global__ void limit_matrices_kernel(const StaticD* __restrict__ st, const double2* __restrict__ inf, int N,
                                      int Nsrc, double2* __restrict__ S, double2* __restrict__ A) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * Nsrc) return;
  const int k = (int)(e % N), n = (int)(e / N);
  const double s = inf ? inf[e].x : st[e].rs, d = inf ? inf[e].y : st[e].rd;
  S[e] = make_double2(s, 0.0);
  A[e] = make_double2((k == n ? 2.0 * PI : 0.0) - d, 0.0);
}

void launch_limit_matrices(const StaticD* st, const double2* inf, int N, int Nsrc, double2* S, double2* A,
                           hipStream_t stream) {
  limit_matrices_kernel<<<(unsigned)(((size_t)N * Nsrc + 255) / 256), 256, 0, stream>>>(st, inf, N, Nsrc, S, A);
  HBG_HIP(hipGetLastError());
}

// The matrices at a frequency limit in finite depth (lo_frequency.rs with FreqMode::Zero /
// Infinite, apply_kernel low_order.rs:738-769 with K = 0: no log terms): entry (k, n), field point
// the centroid of k, source panel n, S = rs + A_n Wr and A = 2 pi delta_kn - (rd + A_n (-WrR a +
// WrZ n_z)) with the limit's kernel (fin_limit_kernel) and rs, rd the statics (omega = 0) or their
// omega = infinity row (inf). Sources n < Nsrc as in limit_matrices_kernel (the images' records
// follow the panels in P). One thread per entry: this runs once per solve.
__global__ void __launch_bounds__(256) limit_fin_matrices_kernel(const PanelD* __restrict__ P, int N, int Nsrc,
                                                                 double h, int zero, const double* __restrict__ ktab,
                                                                 const StaticD* __restrict__ st,
                                                                 const double2* __restrict__ inf,
                                                                 double2* __restrict__ S, double2* __restrict__ A) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * Nsrc) return;
  const int k = (int)(e % N), n = (int)(e / N);
  const PanelD* pk = P + k;
  const PanelD* pn = P + n;
  const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
  const double R = sqrt(dx * dx + dy * dy);
  const FinK w = fin_limit_kernel(zero != 0, h, R, pk->c[2], pn->c[2], ktab);
  const double a = R > 0.0 ? (dx * pn->n[0] + dy * pn->n[1]) / R : 0.0;
  const double rs = zero ? st[e].rs : inf[e].x, rd = zero ? st[e].rd : inf[e].y;
  S[e] = make_double2(rs + pn->area * w.Wr, 0.0);
  A[e] = make_double2((k == n ? 2.0 * PI : 0.0) - (rd + pn->area * (-w.WrR * a + w.WrZ * pn->n[2])), 0.0);
}

void launch_limit_fin_matrices(const PanelD* P, int N, int Nsrc, double h, bool zero, const double* ktab,
                               const StaticD* st, const double2* inf, double2* S, double2* A, hipStream_t stream) {
  limit_fin_matrices_kernel<<<(unsigned)(((size_t)N * Nsrc + 255) / 256), 256, 0, stream>>>(
      P, N, Nsrc, h, zero ? 1 : 0, ktab, st, inf, S, A);
  HBG_HIP(hipGetLastError());
}

// The frequency limits with an interior free surface (lo_frequency.rs:99-104): its unknowns are
// not used there, their rows and columns the identity (S zero), so the body panels' system is
// solved as without the surface. The nchar systems: S_c at S + c N N, A_c at M + c mstride (both
// leading dimension N).
__global__ void interior_identity_kernel(const PanelD* __restrict__ P, int N, int nchar, double2* __restrict__ S,
                                         double2* __restrict__ M, size_t mstride) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * N) return;
  const int k = (int)(e % N), n = (int)(e / N);
  if (P[k].interior == 0.0 && P[n].interior == 0.0) return;
  for (int c = 0; c < nchar; ++c) {
    S[(size_t)c * N * N + e] = make_double2(0.0, 0.0);
    M[(size_t)c * mstride + e] = make_double2(k == n ? 1.0 : 0.0, 0.0);
  }
}

void launch_interior_identity(const PanelD* P, int N, int nchar, double2* S, double2* M, size_t mstride,
                              hipStream_t stream) {
  interior_identity_kernel<<<(unsigned)(((size_t)N * N + 255) / 256), 256, 0, stream>>>(P, N, nchar, S, M, mstride);
  HBG_HIP(hipGetLastError());
}

// The limit kernels at n point pairs (hbg_finite_kernel at K = 0 and K = infinity): out [n][8]
__global__ void limit_points_kernel(double h, int zero, const double* __restrict__ ktab, int n,
                                    const double* __restrict__ R, const double* __restrict__ z,
                                    const double* __restrict__ zeta, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FinK w = fin_limit_kernel(zero != 0, h, R[i], z[i], zeta[i], ktab);
  double* o = out + 8 * (size_t)i;
  o[0] = w.Wr;
  o[1] = w.WrR;
  o[2] = w.WrZ;
  o[3] = w.Wrz;
  o[4] = o[5] = o[6] = o[7] = 0.0;
}

void launch_limit_points(double h, bool zero, const double* ktab, int n, const double* R, const double* z,
                         const double* zeta, double* out, hipStream_t stream) {
  limit_points_kernel<<<(n + 127) / 128, 128, 0, stream>>>(h, zero ? 1 : 0, ktab, n, R, z, zeta, out);
  HBG_HIP(hipGetLastError());
}

void launch_pad_system(double2* M, float2* M32, int n, int nz, int R, int rz, hipStream_t stream) {
  if (nz >= n && rz >= R) return;
  pad_system_kernel<<<n + R, 256, 0, stream>>>(M, M32, n, nz, R, rz);
  HBG_HIP(hipGetLastError());
}

// part: rhs_chunks(N) * R * N values of scratch
void launch_rhs_batched(const double2* S, size_t sstride, int N, int lds, const double2* u, size_t ustride, int R,
                        double2* M, size_t mstride, int ldm, double2* part, int nsys, hipStream_t stream) {
  const int chunks = rhs_chunks(N), len = rhs_chunk_len(N);
  const size_t pstride = (size_t)chunks * R * N;
  const dim3 g((N + 255) / 256, chunks, nsys);
  if (R <= 8)
    rhs_partial_kernel<8><<<g, 256, 0, stream>>>(S, sstride, N, lds, u, ustride, R, len, part, pstride);
  else if (R <= 16)
    rhs_partial_kernel<16><<<g, 256, 0, stream>>>(S, sstride, N, lds, u, ustride, R, len, part, pstride);
  else
    rhs_partial_kernel<MAXRHS><<<g, 256, 0, stream>>>(S, sstride, N, lds, u, ustride, R, len, part, pstride);
  rhs_reduce_kernel<<<dim3((unsigned)(((size_t)R * N + 255) / 256), nsys), 256, 0, stream>>>(part, pstride, N, R,
                                                                                           chunks, M, mstride, ldm, N);
  HBG_HIP(hipGetLastError());
}

// B (N x R, ld N) = S (-u) with S complex64 (the FP32 assembly's) and, at the near entries
// (rowptr, ent), S + dS; the products summed in FP64
void launch_rhs32(const float2* S, int N, const double2* u, int R, double2* B, double2* part, const int* rowptr,
                  const int2* ent, const double2* dS, hipStream_t stream) {
  const int chunks = rhs_chunks(N), len = rhs_chunk_len(N);
  const size_t pstride = (size_t)chunks * R * N;
  const dim3 g((N + 255) / 256, chunks, 1);
  if (R <= 8)
    rhs_partial_kernel<8, float2><<<g, 256, 0, stream>>>(S, 0, N, N, u, 0, R, len, part, pstride);
  else if (R <= 16)
    rhs_partial_kernel<16, float2><<<g, 256, 0, stream>>>(S, 0, N, N, u, 0, R, len, part, pstride);
  else
    rhs_partial_kernel<MAXRHS, float2><<<g, 256, 0, stream>>>(S, 0, N, N, u, 0, R, len, part, pstride);
  rhs_reduce_kernel<<<dim3((unsigned)(((size_t)R * N + 255) / 256), 1), 256, 0, stream>>>(part, pstride, N, R, chunks,
                                                                                       B, 0, N, 0);
  HBG_HIP(hipGetLastError());
  // the near entries' dS (-u), row by row
  if (rowptr) launch_near_rows(rowptr, ent, dS, 0, u, 0, N, R, -1.0, B, 0, N, 1, 1, stream);
}

void launch_rhs(const double2* S, int N, int lds, const double2* u, int R, double2* M, int ldm, double2* part,
                hipStream_t stream) {
  launch_rhs_batched(S, 0, N, lds, u, 0, R, M, 0, ldm, part, 1, stream);
}

// the solution of [A | B] in the columns [N, N + R) of M (leading dimension ldm)
void launch_forces(const PanelD* P, const int* body, const double* nj, int N, int nrad, int H, double w, double rho,
                   const double2* M, int ldm, const double2* phiI, const double2* dphiI, double* out,
                   hipStream_t stream) {
  const int items = nrad * nrad + H * nrad + nrad * H;
  forces_kernel<<<items, 256, 0, stream>>>(P, body, nj, N, nrad, H, w, rho, M + (size_t)N * ldm, ldm, phiI, dphiI,
                                           out);
  HBG_HIP(hipGetLastError());
}

// the same with the solution given on its own: column c at sol + c lds (the whole body of a
// symmetric solve)
void launch_forces_sol(const PanelD* P, const int* body, const double* nj, int N, int nrad, int H, double w,
                       double rho, const double2* sol, int lds, const double2* phiI, const double2* dphiI, double* out,
                       hipStream_t stream) {
  const int items = nrad * nrad + H * nrad + nrad * H;
  forces_kernel<<<items, 256, 0, stream>>>(P, body, nj, N, nrad, H, w, rho, sol, lds, phiI, dphiI, out);
  HBG_HIP(hipGetLastError());
}

}  // namespace hbg
