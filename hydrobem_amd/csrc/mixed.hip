// The mixed-precision solve: the FP64-assembled system [A | b] is factored in complex64 on the
// matrix cores (lu.hip) and the solution recovered by iterative refinement in FP64,
//   x0 = (LU)^-1 b (FP32),   r = b - A x (FP64, against the FP64 matrix),   x += (LU)^-1 r,
// until the normwise backward error |r| / (|A| |x| + |b|) (infinity norms) is at most the
// tolerance. No counterpart in the reference (whose single-precision mode factors in complex64
// and stops there, linalg.rs:321-328); the kernels here are the casts, the residual with its
// norms, and the update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "../../include/hydrobem_amd.h"
#include "hbg_internal.h"

namespace hbg {

void launch_near_rows(const int* rowptr, const int2* ent, const double2* d, size_t dstride, const double2* v,
                      size_t vstride, int N, int R, double sign, double2* out, size_t ostride, int ldo, int accumulate,
                      int nsys, hipStream_t stream);

namespace {

constexpr int RES_THREADS = 256;
constexpr int RES_MAXR = 32;
constexpr int RES_JC = 32;  // columns of A per LDS stage

__global__ void cast_kernel(const double2* __restrict__ src, size_t sstride, float2* __restrict__ dst, size_t dstride,
                            size_t count) {
  src += blockIdx.z * sstride;
  dst += blockIdx.z * dstride;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = src[i];
    dst[i] = float2((float)v.x, (float)v.y);
  }
}

// The right-hand sides go to the complex64 LU scaled column by column by a power of two, so that
// their largest |re| + |im| lies in (1/2, 1] (exact both ways): the LU's split-FP16 GEMM holds
// them in FP16 along with the matrix. scale[z][c] = 2^-e, one workgroup per (c, z).
__global__ void __launch_bounds__(256) rhs_scale_kernel(const double2* __restrict__ B, size_t bstride, int n,
                                                        int R, double* __restrict__ scale) {
  __shared__ double red[4];
  const double2* b = B + blockIdx.z * bstride + (size_t)blockIdx.x * n;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(b[i].x) + fabs(b[i].y));
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    int e = 0;
    if (m > 0.0 && isfinite(m)) frexp(m, &e);  // m = f 2^e, f in [1/2, 1)
    scale[(size_t)blockIdx.z * R + blockIdx.x] = ldexp(1.0, -e);
  }
}

// the right-hand sides B (n x R, ld n) into the columns [n, n + R) of M32, times their column scales
__global__ void cast_rhs_kernel(const double2* __restrict__ B, size_t bstride, float2* __restrict__ M32,
                                size_t mstride, int n, int R, const double* __restrict__ scale) {
  const double2* M64 = B + blockIdx.z * bstride;
  M32 += blockIdx.z * mstride + (size_t)n * n;
  const size_t count = (size_t)n * R;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const double s = scale[(size_t)blockIdx.z * R + i / n];
    const double2 v = M64[i];
    M32[i] = float2((float)(v.x * s), (float)(v.y * s));
  }
}

// X (n x R, ld n) = the complex64 columns [n, n + R) of M32, divided by their column scales
__global__ void widen_kernel(const float2* __restrict__ M32, size_t mstride, int n, double2* __restrict__ X,
                             size_t xstride, int R, const double* __restrict__ scale) {
  M32 += blockIdx.z * mstride + (size_t)n * n;
  X += blockIdx.z * xstride;
  const size_t count = (size_t)n * R;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const double s = 1.0 / scale[(size_t)blockIdx.z * R + i / n];
    const float2 v = M32[i];
    X[i] = double2((double)v.x * s, (double)v.y * s);
  }
}

// The residual b - A X of the systems [A | b] (n x (n + R), ld n) in FP64, in two passes:
// residual_part_kernel: the products of column chunk cy (n / RES_SPLIT columns) for one row per
//   thread, X staged through LDS, into part2[z][cy][R + 1][n] (the last row: sum_j |a_ij|);
// residual_sum_kernel: the chunks summed in a fixed order, D = (complex64) r, and per workgroup
//   the maxima over its rows of max_c |r_ic|, sum_j |a_ij|, max_c |x_ic|, max_c |b_ic| (|re| + |im|)
//   in part[z][block][4].
constexpr int RES_SPLIT = 8;

// RM: right-hand sides held in registers (R <= RM); the columns go 8 at a time so that eight
// loads of A are in flight per lane (the kernel streams A once: its rate is that of the loads in
// flight).
// nc (optional): entries (row i, column j) where the matrix is A + dA, dA kept apart in FP64
// (the complex64 matrix of the FP32 assembly and its FP64 near entries), rows in the list's order.
struct NearCorr {
  const int* rowptr = nullptr;   // [n + 1]
  const int2* ent = nullptr;     // (row, column) of each entry
  const double2* dA = nullptr;   // [system][entries]
  size_t stride = 0;             // entries per system in dA
};

__device__ __forceinline__ double2 widen2(double2 v) { return v; }
__device__ __forceinline__ double2 widen2(float2 v) { return double2((double)v.x, (double)v.y); }

template <int RM, class MT>
__global__ void __launch_bounds__(RES_THREADS) residual_part_kernel(const MT* __restrict__ M, size_t mstride,
                                                                    int n, const double2* __restrict__ X,
                                                                    size_t xstride, int R,
                                                                    double2* __restrict__ part2,
                                                                    const int* __restrict__ active) {
  if (active[blockIdx.z] == 0) return;
  __shared__ double2 xs[RES_JC][RM];
  M += blockIdx.z * mstride;
  X += blockIdx.z * xstride;
  const int tid = threadIdx.x;
  const int i = blockIdx.x * RES_THREADS + tid;
  const bool live = i < n;
  const int cw = (n + RES_SPLIT - 1) / RES_SPLIT;
  const int jlo = blockIdx.y * cw, jhi = min(n, jlo + cw);
  double2 acc[RM];
#pragma unroll
  for (int c = 0; c < RM; ++c) acc[c] = double2(0.0, 0.0);
  double rowsum = 0.0;
  auto column = [&](double2 a, int jj) {
    rowsum += fabs(a.x) + fabs(a.y);
#pragma unroll
    for (int c = 0; c < RM; ++c) {
      if (c < R) {
        const double2 x = xs[jj][c];
        acc[c].x = fma(a.x, x.x, fma(-a.y, x.y, acc[c].x));
        acc[c].y = fma(a.x, x.y, fma(a.y, x.x, acc[c].y));
      }
    }
  };
  for (int j0 = jlo; j0 < jhi; j0 += RES_JC) {
    const int jn = min(RES_JC, jhi - j0);
    __syncthreads();
    for (int e = tid; e < RES_JC * RM; e += RES_THREADS) {
      const int jj = e % RES_JC, c = e / RES_JC;
      xs[jj][c] = (jj < jn && c < R) ? X[(size_t)c * n + j0 + jj] : double2(0.0, 0.0);
    }
    __syncthreads();
    if (live) {
      const MT* col = M + (size_t)j0 * n + i;
      int jj = 0;
      for (; jj + 8 <= jn; jj += 8) {
        MT a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = col[(size_t)(jj + u) * n];
#pragma unroll
        for (int u = 0; u < 8; ++u) column(widen2(a[u]), jj + u);
      }
      for (; jj < jn; ++jj) column(widen2(col[(size_t)jj * n]), jj);
    }
  }
  if (!live) return;
  double2* o = part2 + (((size_t)blockIdx.z * (RES_SPLIT + 1) + blockIdx.y) * (R + 1)) * n + i;
#pragma unroll
  for (int c = 0; c < RM; ++c)
    if (c < R) o[(size_t)c * n] = acc[c];
  o[(size_t)R * n] = double2(rowsum, 0.0);
}

// a maximum that keeps a NaN operand as +inf (fmax drops it), so that a non-finite residual,
// solution or right-hand side can never pass the tolerance
__device__ __forceinline__ double max_keep_nan(double acc, double v) { return v != v ? INFINITY : fmax(acc, v); }

// Partial maxima per workgroup: for each right-hand side c, max |r_ic|, max |x_ic|, max |b_ic|
// over its rows (|re| + |im|), then max_i sum_j |a_ij|, into part[z][block][3 R + 1].
__global__ void __launch_bounds__(RES_THREADS) residual_sum_kernel(const double2* __restrict__ B, size_t bstride,
                                                                   int n, const double2* __restrict__ X,
                                                                   size_t xstride, int R,
                                                                   const double2* __restrict__ part2,
                                                                   float2* __restrict__ D, size_t dstride,
                                                                   double* __restrict__ part,
                                                                   const int* __restrict__ active, int chunks) {
  if (active[blockIdx.z] == 0) return;
  __shared__ double red[3 * RES_MAXR + 1][RES_THREADS / 64];
  B += blockIdx.z * bstride;
  X += blockIdx.z * xstride;
  D += blockIdx.z * dstride;
  const int tid = threadIdx.x;
  const int i = blockIdx.x * RES_THREADS + tid;
  const bool live = i < n;
  const double2* p = part2 + ((size_t)blockIdx.z * (RES_SPLIT + 1) * (R + 1)) * n + i;
  double rowsum = 0.0;
  if (live)
    for (int cy = 0; cy < RES_SPLIT; ++cy) rowsum += p[((size_t)cy * (R + 1) + R) * n].x;
  auto reduce = [&](double v, int slot) {
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((tid & 63) == 0) red[slot][tid >> 6] = v;
  };
  for (int c = 0; c < R; ++c) {
    double rm = 0.0, xm = 0.0, bm = 0.0;
    if (live) {
      double2 sacc = double2(0.0, 0.0);
      for (int cy = 0; cy < chunks; ++cy) {  // the column chunks, then the near entries' chunk
        const double2 v = p[((size_t)cy * (R + 1) + c) * n];
        sacc.x += v.x;
        sacc.y += v.y;
      }
      const double2 b = B[(size_t)c * n + i];
      const double2 x = X[(size_t)c * n + i];
      const double rx = b.x - sacc.x, ry = b.y - sacc.y;
      D[(size_t)c * n + i] = float2((float)rx, (float)ry);
      rm = max_keep_nan(0.0, fabs(rx) + fabs(ry));
      xm = max_keep_nan(0.0, fabs(x.x) + fabs(x.y));
      bm = max_keep_nan(0.0, fabs(b.x) + fabs(b.y));
    }
    reduce(rm, 3 * c);
    reduce(xm, 3 * c + 1);
    reduce(bm, 3 * c + 2);
  }
  reduce(max_keep_nan(0.0, rowsum), 3 * R);
  __syncthreads();
  for (int q = tid; q <= 3 * R; q += RES_THREADS) {
    double m = 0.0;
    for (int w = 0; w < RES_THREADS / 64; ++w) m = fmax(m, red[q][w]);
    part[((size_t)blockIdx.z * gridDim.x + blockIdx.x) * (3 * RES_MAXR + 1) + q] = m;
  }
}

// The backward error of every system still refined, column by column (a right-hand side of small
// magnitude is held to the tolerance on its own scale):
//   eta = max_c |r_c| / (|A| |x_c| + |b_c|)  (infinity norms, |re| + |im|),
// 0 for a zero residual, +inf for a non-finite one. One wavefront per system.
__global__ void __launch_bounds__(64) eta_kernel(const double* __restrict__ part, int nblk, int R,
                                                 const int* __restrict__ active, double* __restrict__ eta) {
  const int z = blockIdx.x, tid = threadIdx.x;
  if (active[z] == 0) return;
  const double* p = part + (size_t)z * nblk * (3 * RES_MAXR + 1);
  double rowsum = 0.0;
  for (int k = tid; k < nblk; k += 64) rowsum = fmax(rowsum, p[(size_t)k * (3 * RES_MAXR + 1) + 3 * R]);
  for (int off = 32; off >= 1; off >>= 1) rowsum = fmax(rowsum, __shfl_xor(rowsum, off));
  double e = 0.0;
  for (int c = 0; c < R; ++c) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int k = tid; k < nblk; k += 64)
      for (int q = 0; q < 3; ++q) v[q] = fmax(v[q], p[(size_t)k * (3 * RES_MAXR + 1) + 3 * c + q]);
    for (int q = 0; q < 3; ++q)
      for (int off = 32; off >= 1; off >>= 1) v[q] = fmax(v[q], __shfl_xor(v[q], off));
    const double den = rowsum * v[1] + v[2];
    const double ec = v[0] == 0.0 ? 0.0 : (den > 0.0 ? v[0] / den : INFINITY);
    e = max_keep_nan(e, ec);
  }
  if (tid == 0) eta[z] = e;
}

// X += D (the correction, complex64) for the systems still refined
__global__ void update_kernel(double2* __restrict__ X, size_t xstride, const float2* __restrict__ D, size_t dstride,
                              size_t count, const int* __restrict__ active) {
  if (active[blockIdx.z] == 0) return;
  X += blockIdx.z * xstride;
  D += blockIdx.z * dstride;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const float2 d = D[i];
    double2 x = X[i];
    x.x += (double)d.x;
    x.y += (double)d.y;
    X[i] = x;
  }
}

// B (n x R, ld n, FP64) = X
__global__ void store_kernel(double2* __restrict__ B, size_t bstride, int n, const double2* __restrict__ X,
                             size_t xstride, int R) {
  double2* M = B + blockIdx.z * bstride;
  X += blockIdx.z * xstride;
  const size_t count = (size_t)n * R;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    M[i] = X[i];
}

unsigned grid_for(size_t count) { return (unsigned)std::min<size_t>((count + 255) / 256, 4096); }

}  // namespace

void lu_solve_batched32(float2* M, size_t mstride, int batch, int n, int lda, int nrhs, hipStream_t stream,
                        LuStats* stats, int* status, const int* nsys = nullptr);
void lu_resolve32(const float2* M, size_t mstride, int batch, int n, int lda, float2* X, int ldx, size_t xstride,
                  int nrhs, hipStream_t stream, const int* active);

int residual_blocks(int n) { return (n + RES_THREADS - 1) / RES_THREADS; }
// doubles of the per-workgroup maxima of the residual, per system (the eta of the system last)
size_t residual_max_elems(int n) { return (size_t)residual_blocks(n) * (3 * RES_MAXR + 1) + 1; }
// the chunk partials of the residual, per system
size_t residual_part_elems(int n, int R) { return (size_t)(RES_SPLIT + 1) * (R + 1) * n; }

// Solve the `batch` FP64 systems M64 (n x (n + R), ld n, stride n (n + R)) in mixed precision:
// M32 (same shape, complex64; its first n columns already written by the assembly when
// a32_written) and D (n x R complex64), X (n x R complex128), part (batch x residual_max_elems(n)),
// part2 (batch x residual_part_elems(n, R)) and active (batch ints) are workspaces; the systems
// that meet the tolerance take no further step (their solution is the one checked). On return the
// solution is in the columns
// [n, n + R) of M64 for the systems with status 0; status -1: the complex64 factorization needed
// interchanges or the refinement did not reach the tolerance in max_iter steps (solve it in FP64);
// k > 0: zero pivot k. iters[b], eta[b]: refinement steps and the final backward error.
namespace {

// The refinement of `batch` systems whose matrix for the residual is Amat (n x n, ld n, stride
// astride; FP64, or the complex64 matrix itself, then with the near entries nc), right-hand sides B
// (n x R, ld n, stride bstride; overwritten by the solution of the systems with status 0) and the
// complex64 systems M32 (n x (n + R), stride ms32; its first n columns written when a32_written).
template <class MT>
void mixed_core(const MT* Amat, size_t astride, double2* B, size_t bstride, NearCorr nc, float2* M32, size_t ms32,
                double2* X, float2* D, double* part, double2* part2, double* rscale, int batch, int n, int R,
                bool a32_written, double tol, int max_iter, hipStream_t stream, LuStats* stats, int* status,
                int* iters, double* eta, const int* nsys, int* active) {
  if (R > RES_MAXR) throw Error{HBG_ERR_UNSUPPORTED, "mixed precision: more than 32 right-hand sides"};
  const size_t xs = (size_t)n * R;
  const int nblk = residual_blocks(n);
  if (!a32_written) {  // else the assembly wrote A in complex64 too: the right-hand sides only
    if constexpr (std::is_same<MT, double2>::value)
      cast_kernel<<<dim3(grid_for((size_t)n * n), 1, batch), 256, 0, stream>>>(Amat, astride, M32, ms32,
                                                                                (size_t)n * n);
    else
      throw Error{HBG_ERR_INTERNAL, "mixed precision: a complex64 matrix that was not written"};
  }
  if (R > 0) {
    rhs_scale_kernel<<<dim3(R, 1, batch), 256, 0, stream>>>(B, bstride, n, R, rscale);
    cast_rhs_kernel<<<dim3(grid_for(xs), 1, batch), 256, 0, stream>>>(B, bstride, M32, ms32, n, R, rscale);
  }
  lu_solve_batched32(M32, ms32, batch, n, n, R, stream, stats, status, nsys);
  widen_kernel<<<dim3(grid_for(xs), 1, batch), 256, 0, stream>>>(M32, ms32, n, X, xs, R, rscale);
  // the per-workgroup maxima first, then one eta per system after them
  double* const deta = part + (size_t)batch * nblk * (3 * RES_MAXR + 1);
  std::vector<double> he(batch);
  for (int b = 0; b < batch; ++b) {
    iters[b] = 0;
    eta[b] = INFINITY;
  }
  // act[b]: system b takes the next residual (and, unless it then meets the tolerance, the next
  // correction); the kernels skip the others
  hipEvent_t r0 = nullptr, r1 = nullptr;
  if (stats) {
    HBG_HIP(hipEventCreate(&r0));
    HBG_HIP(hipEventCreate(&r1));
    HBG_HIP(hipEventRecord(r0, stream));
  }
  std::vector<int> act(batch);
  bool open = false;
  for (int b = 0; b < batch; ++b) {
    act[b] = status[b] == 0 ? 1 : 0;
    open = open || act[b];
  }
  HBG_HIP(hipMemcpyAsync(active, act.data(), batch * sizeof(int), hipMemcpyHostToDevice, stream));
  while (open) {
    const dim3 rg(nblk, RES_SPLIT, batch);
    if (R <= 8)
      residual_part_kernel<8, MT><<<rg, RES_THREADS, 0, stream>>>(Amat, astride, n, X, xs, R, part2, active);
    else if (R <= 16)
      residual_part_kernel<16, MT><<<rg, RES_THREADS, 0, stream>>>(Amat, astride, n, X, xs, R, part2, active);
    else
      residual_part_kernel<RES_MAXR, MT><<<rg, RES_THREADS, 0, stream>>>(Amat, astride, n, X, xs, R, part2, active);
    if (nc.rowptr)  // the near entries' dA X into the chunk after the column chunks
      launch_near_rows(nc.rowptr, nc.ent, nc.dA, nc.stride, X, xs, n, R, 1.0, part2 + (size_t)RES_SPLIT * (R + 1) * n,
                       residual_part_elems(n, R), n, 0, batch, stream);
    residual_sum_kernel<<<dim3(nblk, 1, batch), RES_THREADS, 0, stream>>>(B, bstride, n, X, xs, R, part2, D, xs,
                                                                          part, active, RES_SPLIT + (nc.rowptr ? 1 : 0));
    eta_kernel<<<batch, 64, 0, stream>>>(part, nblk, R, active, deta);
    HBG_HIP(hipMemcpyAsync(he.data(), deta, batch * sizeof(double), hipMemcpyDeviceToHost, stream));
    HBG_HIP(hipStreamSynchronize(stream));
    open = false;
    for (int b = 0; b < batch; ++b) {
      if (!act[b]) continue;
      eta[b] = he[b];
      if (stats && iters[b] == 0) stats->first_eta = std::max(stats->first_eta, eta[b] == eta[b] ? eta[b] : INFINITY);
      if (!(eta[b] <= tol) && iters[b] < max_iter)
        open = true;
      else
        act[b] = 0;
    }
    if (!open) break;
    HBG_HIP(hipMemcpyAsync(active, act.data(), batch * sizeof(int), hipMemcpyHostToDevice, stream));
    lu_resolve32(M32, ms32, batch, n, n, D, n, xs, R, stream, active);
    update_kernel<<<dim3(grid_for(xs), 1, batch), 256, 0, stream>>>(X, xs, D, xs, xs, active);
    for (int b = 0; b < batch; ++b)
      if (act[b]) ++iters[b];
  }
  for (int b = 0; b < batch; ++b)
    if (status[b] == 0 && !(eta[b] <= tol)) status[b] = -1;
  store_kernel<<<dim3(grid_for(xs), 1, batch), 256, 0, stream>>>(B, bstride, n, X, xs, R);
  HBG_HIP(hipGetLastError());
  if (stats) {
    HBG_HIP(hipEventRecord(r1, stream));
    HBG_HIP(hipEventSynchronize(r1));
    float t = 0.0f;
    HBG_HIP(hipEventElapsedTime(&t, r0, r1));
    stats->refine_ms += t;
    (void)hipEventDestroy(r0);
    (void)hipEventDestroy(r1);
  }
}

}  // namespace

void mixed_solve_batched(double2* M64, float2* M32, double2* X, float2* D, double* part, double2* part2,
                         double* rscale, int batch, int n, int R, bool a32_written, double tol, int max_iter,
                         hipStream_t stream, LuStats* stats, int* status, int* iters, double* eta, const int* nsys,
                         int* active) {
  const size_t ms = (size_t)n * (n + R);
  mixed_core<double2>(M64, ms, M64 + (size_t)n * n, ms, NearCorr{}, M32, ms, X, D, part, part2, rscale, batch, n, R,
                      a32_written, tol, max_iter, stream, stats, status, iters, eta, nsys, active);
}

// The systems of the FP32 assembly: M32 [A | b] (complex64, stride n (n + R), A written by the
// assembly and factored here in place), A32r the same A unfactored (stride n n), B (n x R FP64
// right-hand sides, stride n R; the solutions on return) and the FP64 near entries of every system
// (rowptr, ent shared; dA with `dstride` entries per system). The residuals are taken in FP64
// against A32r + dA.
void mixed32_solve_batched(float2* M32, const float2* A32r, double2* B, const int* rowptr, const int2* ent,
                           const double2* dA, size_t dstride, double2* X, float2* D, double* part, double2* part2,
                           double* rscale, int batch, int n, int R, double tol, int max_iter, hipStream_t stream,
                           LuStats* stats, int* status, int* iters, double* eta, int* active) {
  const size_t ms = (size_t)n * (n + R);
  NearCorr nc;
  nc.rowptr = rowptr;
  nc.ent = ent;
  nc.dA = dA;
  nc.stride = dstride;
  mixed_core<float2>(A32r, (size_t)n * n, B, (size_t)n * R, rowptr ? nc : NearCorr{}, M32, ms, X, D, part, part2,
                     rscale, batch, n, R, true, tol, max_iter, stream, stats, status, iters, eta, nullptr, active);
}

}  // namespace hbg
