// The motions (RAOs) of the rigid bodies, one frequency per workgroup (motion.rs:121-158):
//   (C - w^2 (M + A(w)) + i w (B(w) + B_ext)) xi = X(w)
// on the D = 6 n_bodies modes, complex, solved by Gaussian elimination with partial pivoting on
// the largest |re| + |im| (the reference's lu_solve, linalg.rs:321-328). The system lives in
// LDS; thread i holds row i (D <= 64); the pivot search is a wavefront reduction.
#include <hip/hip_runtime.h>

#include "../../include/hydrobem_amd.h"
#include "hbg_internal.h"

namespace hbg {

namespace {

constexpr int MOT_MAXD = 64, MOT_MAXH = 32;

__global__ void __launch_bounds__(64) motions_kernel(int D, int H, const double* __restrict__ omegas,
                                                     const double* __restrict__ Mm, const double* __restrict__ Cm,
                                                     const double* __restrict__ Bx, const double* __restrict__ Aw,
                                                     const double* __restrict__ Bw, const double2* __restrict__ Xw,
                                                     double2* __restrict__ out, int* __restrict__ status) {
  __shared__ double2 Z[MOT_MAXD][MOT_MAXD + 1];
  __shared__ double2 Y[MOT_MAXD][MOT_MAXH + 1];
  const int f = blockIdx.x, i = threadIdx.x;
  const double w = omegas[f];
  const double* A = Aw + (size_t)f * D * D;
  const double* B = Bw + (size_t)f * D * D;
  const double2* X = Xw + (size_t)f * H * D;
  if (i < D) {
    for (int j = 0; j < D; ++j) {
      const int o = i * D + j;
      Z[i][j] = make_double2(Cm[o] - w * w * (Mm[o] + A[o]), w * (B[o] + (Bx ? Bx[o] : 0.0)));
    }
    for (int h = 0; h < H; ++h) Y[i][h] = X[(size_t)h * D + i];
  }
  __syncthreads();
  int zp = 0;
  for (int j = 0; j < D; ++j) {
    // pivot: the largest |re| + |im| in column j from row j down, the first on ties; a NaN
    // magnitude counts as the largest, so that every lane agrees on a row in [j, D)
    double v = -1.0;
    int r = 0x7fffffff;
    if (i >= j && i < D) {
      v = fabs(Z[i][j].x) + fabs(Z[i][j].y);
      if (v != v) v = INFINITY;
      r = i;
    }
    for (int off = 32; off >= 1; off >>= 1) {
      const double v2 = __shfl_xor(v, off);
      const int r2 = __shfl_xor(r, off);
      if (v2 > v || (v2 == v && r2 < r)) {
        v = v2;
        r = r2;
      }
    }
    if (v == 0.0 || r < j || r >= D) {  // no usable pivot (r outside the rows is never taken)
      if (zp == 0) zp = j + 1;
      continue;
    }
    if (r != j) {  // swap rows j and r: thread c takes column c
      for (int c = i; c < D; c += 64) {
        const double2 t = Z[j][c];
        Z[j][c] = Z[r][c];
        Z[r][c] = t;
      }
      for (int h = i; h < H; h += 64) {
        const double2 t = Y[j][h];
        Y[j][h] = Y[r][h];
        Y[r][h] = t;
      }
    }
    __syncthreads();
    if (i > j && i < D) {
      const double2 p = Z[j][j], a = Z[i][j];
      const double d = p.x * p.x + p.y * p.y;
      const double2 l = make_double2((a.x * p.x + a.y * p.y) / d, (a.y * p.x - a.x * p.y) / d);
      for (int c = j + 1; c < D; ++c) {
        const double2 u = Z[j][c];
        Z[i][c].x -= l.x * u.x - l.y * u.y;
        Z[i][c].y -= l.x * u.y + l.y * u.x;
      }
      for (int h = 0; h < H; ++h) {
        const double2 u = Y[j][h];
        Y[i][h].x -= l.x * u.x - l.y * u.y;
        Y[i][h].y -= l.x * u.y + l.y * u.x;
      }
    }
    __syncthreads();
  }
  if (i == 0) status[f] = zp;
  if (zp) return;
  // back substitution, one right-hand side per thread
  for (int h = i; h < H; h += 64)
    for (int j = D - 1; j >= 0; --j) {
      double2 s = Y[j][h];
      for (int c = j + 1; c < D; ++c) {
        const double2 z = Z[j][c], y = Y[c][h];
        s.x -= z.x * y.x - z.y * y.y;
        s.y -= z.x * y.y + z.y * y.x;
      }
      const double2 p = Z[j][j];
      const double d = p.x * p.x + p.y * p.y;
      Y[j][h] = make_double2((s.x * p.x + s.y * p.y) / d, (s.y * p.x - s.x * p.y) / d);
    }
  __syncthreads();
  if (i < D)
    for (int h = 0; h < H; ++h) out[((size_t)f * H + h) * D + i] = Y[i][h];
}

}  // namespace

int motions_max_dof() { return MOT_MAXD; }
int motions_max_headings() { return MOT_MAXH; }

// All on the device; status[f]: 0, or the zero pivot (1-based) of frequency f.
void launch_motions(int F, int D, int H, const double* omegas, const double* M, const double* C, const double* Bx,
                    const double* A, const double* B, const double2* X, double2* out, int* status,
                    hipStream_t stream) {
  motions_kernel<<<F, 64, 0, stream>>>(D, H, omegas, M, C, Bx, A, B, X, out, status);
  HBG_HIP(hipGetLastError());
}

}  // namespace hbg
