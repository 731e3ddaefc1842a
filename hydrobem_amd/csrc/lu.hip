// Dense complex LU of the system matrices, FP64, on the GPU (linalg.rs:134-252 semantics:
// partial pivoting on the largest |re| + |im| as LAPACK's izamax, LAPACK ipiv, unit lower L,
// "singular system matrix (zero pivot k)").
//
// The matrix is stored column-major with the right-hand sides appended as extra columns
// ([A | B], n x (n + r), leading dimension lda), so that the row interchanges, the block row of U
// and the trailing updates carry the forward substitution L^-1 P B along; a blocked back
// substitution with the stored inverses of U's diagonal blocks then solves U X = L^-1 P B.
//
// Right-looking, in panels of NB columns. The whole matrix is first factored without
// interchanges (diag_kernel on the diagonal block in registers, L21 = A21 U11^-1 by apply_kernel)
// while checking, column by column, that the diagonal entry is the largest |re| + |im| of its
// column at that step of the elimination: then partial pivoting would make no interchange and the
// result IS the partially pivoted factorization. Where the check fails on any panel (its flag),
// the saved matrix is restored and factored again panel by panel, the panels whose check fails
// being restored and pivoted by a tournament (tslu_kernel): every workgroup takes up to 256 rows,
// one row per thread in registers, and eliminates with partial pivoting (a wavefront shuffle
// reduction per column), its NB pivot rows go up a tree of 4-way merges that repeat the
// elimination on the original values of the candidates; the last workgroup to finish a node
// merges it (atomic counters, no waiting). With one leaf this is partial pivoting exactly; with
// more it is tournament pivoting (Grigori, Demmel & Xiang 2011).
// Then U12 = L11^-1 A12 and the trailing update A22 -= L21 U12 (gemm_mfma_kernel, complex FP64
// on the matrix cores), once per block of panels in the pass without interchanges.
#include <atomic>
#include <algorithm>
#include <array>
#include <type_traits>
#include <vector>

#include "gemm_mfma.h"
#include "hbg_internal.h"

namespace hbg {

constexpr int NB = 64;      // panel width
constexpr int LEAF = 256;   // rows per tournament leaf (one per thread)
constexpr int FAN = 4;      // candidate sets per merge
constexpr int MAXRHS_LU = 32;  // right-hand sides carried along

struct TsluWork {
  int* cand;       // [levels][nodes][NB] candidate rows (global indices)
  int* ccount;     // [levels][nodes] number of candidates
  unsigned* ctr;   // [levels][nodes] arrival counters, zeroed before each panel
  double2* top;    // [NB][NB] L11 \ U11 of the root, row-major by pivot step
  int* winners;    // [NB] pivot rows in order
  int* zero_pivot; // first zero pivot (1-based column), 0: none
  int levels;
  int max_nodes;
};

// Strides between the matrices of a batch (blockIdx.z) and their workspaces, in elements.
struct BatchStrides {
  size_t m;     // [A | B]
  int flag;     // per-panel flags and zero pivots
  int ipiv;
  size_t uinv;  // per-panel U11^-1
};

namespace {

__device__ __forceinline__ double mag1(double2 v) { return fabs(v.x) + fabs(v.y); }
__device__ __forceinline__ float mag1(float2 v) { return fabsf(v.x) + fabsf(v.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// (value, row) argmax with the lower row index on a tie
__device__ __forceinline__ void better(double& v, int& r, double v2, int r2) {
  if (v2 > v || (v2 == v && r2 < r && r2 >= 0)) {
    v = v2;
    r = r2;
  }
}

// Partial-pivot elimination of the rows held by the workgroup (row `grow` per thread, -1: none)
// over kb columns: returns the rows chosen at each step in `win` (LDS), `nwin` of them. With
// `store_top` the chosen rows (multipliers left of the diagonal, U on and right of it) are
// written to top[step][*]. All threads of the workgroup must call it. A NaN magnitude counts as
// the largest (every comparison with NaN is false, so otherwise a step whose candidates are all
// NaN would choose no row); the NaN then reaches the solution, which the solve reports.
__device__ void eliminate(double2 (&a)[NB], int grow, int nrows, int kb, int* win, double2* urow, double* redv,
                          int* redr, bool store_top, double2* top, int* zero_pivot, int col0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bool used = grow < 0;
  const int steps = min(kb, nrows);
  for (int j = 0; j < steps; ++j) {
    double2 aj = make_double2(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c == j) aj = a[c];
    double v = used ? -1.0 : mag1(aj);
    if (v != v) v = INFINITY;
    int r = used ? 0x7fffffff : grow;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double v2 = __shfl_xor(v, off);
      const int r2 = __shfl_xor(r, off);
      better(v, r, v2, r2);
    }
    if (lane == 0) {
      redv[wave] = v;
      redr[wave] = r;
    }
    __syncthreads();
    double bv = redv[0];
    int br = redr[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) better(bv, br, redv[w], redr[w]);
    const bool mine = !used && grow == br;
    if (mine) {
#pragma unroll
      for (int c = 0; c < NB; ++c) urow[c] = a[c];
      win[j] = grow;
      if (store_top) {
#pragma unroll
        for (int c = 0; c < NB; ++c) top[j * NB + c] = a[c];
      }
      used = true;
    }
    __syncthreads();
    const double2 piv = urow[j];
    if (bv == 0.0 && tid == 0 && zero_pivot) atomicMin(zero_pivot, col0 + j + 1);
    if (!used) {
      const double2 l = (piv.x == 0.0 && piv.y == 0.0) ? make_double2(0.0, 0.0) : cdiv(aj, piv);
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c > j) {
          const double2 u = urow[c];
          a[c].x -= l.x * u.x - l.y * u.y;
          a[c].y -= l.x * u.y + l.y * u.x;
        } else if (c == j) {
          a[c] = l;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

// The tournament of the panel (k0, kb) over rows [k0, n): grid = number of leaves.
__global__ void __launch_bounds__(LEAF) tslu_kernel(const double2* __restrict__ A, int n, int lda, int k0, int kb,
                                                    TsluWork w, const int* __restrict__ gate) {
  if (*gate == 0) return;
  __shared__ int win[NB];
  __shared__ double2 urow[NB];
  __shared__ double redv[LEAF / 64];
  __shared__ int redr[LEAF / 64];
  __shared__ int rows[LEAF];
  __shared__ int s_nrows;
  __shared__ int s_go;
  const int tid = threadIdx.x;
  const int m = n - k0;
  int nodes = (m + LEAF - 1) / LEAF;  // nodes of the current level
  int node = blockIdx.x;
  int level = 0;
  // leaf rows
  {
    const int r0 = k0 + node * LEAF;
    const int nr = min(LEAF, n - r0);
    if (tid == 0) s_nrows = nr;
    rows[tid] = tid < nr ? r0 + tid : -1;
  }
  __syncthreads();
  while (true) {
    const int grow = rows[tid];
    const int nrows = s_nrows;
    double2 a[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
      a[c] = (grow >= 0 && c < kb) ? A[(size_t)(k0 + c) * lda + grow] : make_double2(0.0, 0.0);
    const bool root = nodes == 1;
    eliminate(a, grow, nrows, kb, win, urow, redv, redr, root, w.top, root ? w.zero_pivot : nullptr, k0);
    const int nw = min(kb, nrows);
    if (root) {
      if (tid < nw) w.winners[tid] = win[tid];
      return;
    }
    // publish this node's candidates and arrive at the parent
    int* cand = w.cand + ((size_t)level * w.max_nodes + node) * NB;
    if (tid < nw) cand[tid] = win[tid];
    if (tid == 0) w.ccount[level * w.max_nodes + node] = nw;
    __threadfence();
    __syncthreads();
    const int parent = node / FAN;
    const int nparent = (nodes + FAN - 1) / FAN;
    const int children = min(FAN, nodes - parent * FAN);
    if (tid == 0) {
      const unsigned prev = atomicAdd(&w.ctr[level * w.max_nodes + parent], 1u);
      s_go = (prev + 1 == (unsigned)children) ? 1 : 0;
    }
    __syncthreads();
    if (!s_go) return;
    __threadfence();
    // the last child merges: gather the children's candidates
    if (tid == 0) {
      int cnt = 0;
      for (int c = 0; c < children; ++c) {
        const int child = parent * FAN + c;
        const int cc = atomicAdd(&w.ccount[level * w.max_nodes + child], 0);
        const int* src = w.cand + ((size_t)level * w.max_nodes + child) * NB;
        for (int i = 0; i < cc; ++i) rows[cnt++] = atomicAdd((int*)&src[i], 0);
      }
      for (int i = cnt; i < LEAF; ++i) rows[i] = -1;
      s_nrows = cnt;
    }
    __syncthreads();
    node = parent;
    nodes = nparent;
    ++level;
  }
}

// The pivot interchanges of the panel in LAPACK's ipiv form from its pivot rows, by one
// wavefront: lane l tracks the position of the panel's original row k0 + l (pos) and the original
// row now at position k0 + l (at); a pivot row outside the block sits where it started until it
// is chosen. Gated by the panel's flag. A pivot row outside [k0, n) (none is chosen so) is taken
// as no interchange and reported as a numerical failure at its column, never used as an index.
__global__ void __launch_bounds__(64) ipiv_kernel(int k0, int kb, int n, const int* __restrict__ winners,
                                                  int* __restrict__ ipiv, int* __restrict__ zero_pivot,
                                                  const int* __restrict__ gate) {
  if (*gate == 0) return;
  const int l = threadIdx.x;
  int pos = k0 + l, at = k0 + l;
  for (int i = 0; i < kb; ++i) {
    int w = winners[i];
    if (w < k0 || w >= n) {
      if (l == 0) atomicMin(zero_pivot, k0 + i + 1);
      w = __shfl(at, i);
    }
    const bool inside = w < k0 + kb;
    const int p = inside ? __shfl(pos, w - k0) : w;  // current position of row w
    const int o = __shfl(at, i);                        // the original block row at position k0 + i
    if (l == i) at = w;
    if (p < k0 + kb && l == p - k0) at = o;
    if (inside && l == w - k0) pos = k0 + i;
    if (l == o - k0) pos = p;
    if (l == 0) ipiv[k0 + i] = p;
  }
}

__global__ void swap_kernel(double2* __restrict__ A, int lda, int ncols, int k0, int kb, const int* __restrict__ ipiv,
                            const double2* __restrict__ top, const int* __restrict__ gate) {
  if (*gate == 0) return;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  double2* c = A + (size_t)col * lda;
  for (int i = 0; i < kb; ++i) {
    const int p = ipiv[k0 + i];
    if (p != k0 + i) {
      const double2 t = c[k0 + i];
      c[k0 + i] = c[p];
      c[p] = t;
    }
  }
  if (col >= k0 && col < k0 + kb) {
    const int cc = col - k0;
    for (int r = 0; r < kb; ++r) c[k0 + r] = top[r * NB + cc];
  }
}

// the panel's columns [k0, k0 + kb), rows [k0, n): saved before the attempt without interchanges
// (dir 0) and restored for the fallback (dir 1, gated)
__global__ void panel_copy_kernel(double2* __restrict__ M, int n, int lda, int k0, int kb, double2* __restrict__ save,
                                  int dir, const int* __restrict__ gate) {
  if (dir == 1 && *gate == 0) return;
  const int m = n - k0;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)m * kb) return;
  const int r = (int)(e % m), c = (int)(e / m);
  double2* a = M + (size_t)(k0 + c) * lda + k0 + r;
  if (dir == 0)
    save[e] = *a;
  else
    *a = save[e];
}

template <typename T, typename C2 = typename Cplx<T>::type>
__device__ __forceinline__ C2 crecip(C2 v) {
  const T d = v.x * v.x + v.y * v.y;
  if (d == 0.0) return C2(0.0, 0.0);
  const T id = T(1) / d;
  return C2(v.x * id, -v.y * id);
}

// The diagonal block (k0, kb): mode 0 factors it without interchanges, checking at every step j
// that no entry below the diagonal exceeds the pivot in |re| + |im| (else the panel's flag is
// set), writes L11 \ U11 back and the identity interchanges; mode 1 (gated) takes the block as the
// tournament wrote it. Both then invert the triangles: Uinv = U11^-1 (kept per panel for the back
// substitution), Linv = L11^-1 (stored column-major: entry (r, c) at c NB + r, so that the
// column slices the MFMA operands take are contiguous), and keep the diagonal of U. Outside kb the
// block is the identity.
//
// The block lives in registers: thread (ti, tj) of the 16 x 16 grid holds the entries
// (ti + 16 a, tj + 16 b), a, b < 4. At step j the owners of row j and column j publish them in
// LDS (two alternating buffers, so one barrier per step) and every thread updates its entries.
template <typename T>
__global__ void __launch_bounds__(256) diag_kernel(typename Cplx<T>::type* __restrict__ M, int lda, int k0, int kb, int mode,
                                                   int* __restrict__ flag, int* __restrict__ zp_fast,
                                                   int* __restrict__ ipiv, typename Cplx<T>::type* __restrict__ Uinv,
                                                   typename Cplx<T>::type* __restrict__ Linv, typename Cplx<T>::type* __restrict__ udiag, BatchStrides bs) {
  typedef typename Cplx<T>::type C2;
  {
    const int z = blockIdx.z;
    M += z * bs.m;
    flag += z * bs.flag;
    zp_fast += z * bs.flag;
    ipiv += z * bs.ipiv;
    Uinv += z * bs.uinv;
    Linv += z * bs.uinv;
    udiag += z * NB;
  }
  if (mode == 1 && *flag == 0) return;
  __shared__ C2 rowb[2][NB], colb[2][NB];
  __shared__ C2 rowu[2][NB], colu[2][NB];
  __shared__ C2 dg[NB];
  __shared__ int s_fail;
  const int tid = threadIdx.x, ti = tid & 15, tj = tid >> 4;
  if (tid == 0) s_fail = 0;
  C2 t[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int r = ti + 16 * a, c = tj + 16 * b;
      t[a][b] = (r < kb && c < kb) ? M[(size_t)(k0 + c) * lda + k0 + r] : C2(r == c ? 1.0 : 0.0, 0.0);
    }
  if (mode == 0) {
    bool fail = false;
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) {
      for (int jj = 0; jj < 16; ++jj) {
        const int j = 16 * jb + jj, buf = jj & 1;
        if (ti == jj) {
#pragma unroll
          for (int b = 0; b < 4; ++b) rowb[buf][tj + 16 * b] = t[jb][b];
        }
        if (tj == jj) {
#pragma unroll
          for (int a = 0; a < 4; ++a) colb[buf][ti + 16 * a] = t[a][jb];
        }
        __syncthreads();
        const C2 piv = rowb[buf][j];
        const T pm = mag1(piv);
        const C2 rp = crecip<T>(piv);
        if (tid == 0 && pm == 0.0 && j < kb) atomicMin(zp_fast, k0 + j + 1);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          if (a < jb) continue;
          if (a == jb && ti <= jj) continue;
          const C2 v = colb[buf][ti + 16 * a];
          if (tj == jj && mag1(v) > pm) fail = true;
          const C2 l = cmul(v, rp);
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            if (b < jb) continue;
            if (b == jb && tj < jj) continue;
            if (b == jb && tj == jj) {
              t[a][b] = l;
            } else {
              const C2 u = rowb[buf][tj + 16 * b];
              t[a][b].x -= l.x * u.x - l.y * u.y;
              t[a][b].y -= l.x * u.y + l.y * u.x;
            }
          }
        }
      }
    }
    if (fail) s_fail = 1;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int r = ti + 16 * a, c = tj + 16 * b;
        if (r < kb && c < kb) M[(size_t)(k0 + c) * lda + k0 + r] = t[a][b];
      }
    for (int i = tid; i < kb; i += 256) ipiv[k0 + i] = k0 + i;
  }
  if (ti == tj) {
#pragma unroll
    for (int a = 0; a < 4; ++a) dg[ti + 16 * a] = t[a][a];
  }
  __syncthreads();
  if (mode == 0 && tid == 0 && s_fail) *flag = 1;
  if (tid < NB) udiag[tid] = dg[tid];
  // Linv: L X = I (unit lower), step j ascending: X[r][*] -= l_rj X[j][*], r > j.
  // Uinv: U X = I, step j descending: X[j][*] *= 1 / u_jj, then X[r][*] -= u_rj X[j][*], r < j.
  // Both in one loop, step s doing j = s for Linv and j = 63 - s for Uinv.
  C2 xl[4][4], xu[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const T v = (a == b && ti == tj) ? 1.0 : 0.0;
      xl[a][b] = C2(v, 0.0);
      xu[a][b] = C2(v, 0.0);
    }
#pragma unroll
  for (int sb = 0; sb < 4; ++sb) {
    const int ub = 3 - sb;
    for (int ss = 0; ss < 16; ++ss) {
      const int buf = ss & 1, uj = 15 - ss;
      const int ju = 16 * ub + uj;
      if (ti == ss) {
#pragma unroll
        for (int b = 0; b < 4; ++b) rowb[buf][tj + 16 * b] = xl[sb][b];
      }
      if (tj == ss) {
#pragma unroll
        for (int a = 0; a < 4; ++a) colb[buf][ti + 16 * a] = t[a][sb];
      }
      if (ti == uj) {
        const C2 rd = crecip<T>(dg[ju]);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          xu[ub][b] = cmul(xu[ub][b], rd);
          rowu[buf][tj + 16 * b] = xu[ub][b];
        }
      }
      if (tj == uj) {
#pragma unroll
        for (int a = 0; a < 4; ++a) colu[buf][ti + 16 * a] = t[a][ub];
      }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        // Linv rows r > jl; row jl of X is zero right of jl
        if (a > sb || (a == sb && ti > ss)) {
          const C2 l = colb[buf][ti + 16 * a];
#pragma unroll
          for (int b = 0; b <= sb; ++b) {
            const C2 x = rowb[buf][tj + 16 * b];
            xl[a][b].x -= l.x * x.x - l.y * x.y;
            xl[a][b].y -= l.x * x.y + l.y * x.x;
          }
        }
        // Uinv rows r < ju; row ju of X is zero left of ju
        if (a < ub || (a == ub && ti < uj)) {
          const C2 u = colu[buf][ti + 16 * a];
#pragma unroll
          for (int b = ub; b < 4; ++b) {
            const C2 x = rowu[buf][tj + 16 * b];
            xu[a][b].x -= u.x * x.x - u.y * x.y;
            xu[a][b].y -= u.x * x.y + u.y * x.x;
          }
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int r = ti + 16 * a, c = tj + 16 * b;
      Uinv[r * NB + c] = xu[a][b];
      Linv[c * NB + r] = xl[a][b];  // L11^-1 column-major (its transpose row-major)
    }
}

// The panel's block products with the triangle inverses, in place:
//   blocks [0, rblocks): L21 = A21 Uinv on AB rows each (rows k0 + kb + AB i .. < lim), checking
//     that |l_ij u_jj| (|re| + |im|) does not exceed |u_jj|: the entry at step j of the
//     elimination below the pivot (else the panel's flag);
//   blocks [rblocks, gridDim.x): U12 = Linv A12 on AB columns each (< ncols).
// gate: run only if the flag is set. The inverse is staged in LDS with all of a thread's loads in
// flight at once. A thread's outputs are 8 consecutive columns of L21 or 8 consecutive rows of
// U12: with split set (complex64, a full panel), they also go into the split-FP16 operands of the
// trailing updates (split8: L21 row r, columns 8 g .. 8 g + 7 at k group g of the row; U12 column
// c, rows 8 g .. 8 g + 7 at k group g of the column), so the panel is not read again.
constexpr int AB = 32;

// The complex64 block products of apply_kernel on the FP32 matrix cores (v_mfma_f32_16x16x4f32,
// the lane layout of gemm_mfma_kernel: operand lane l holds column (row) l % 16 at k = l / 16 of
// the k-step, output item q of lane l is row l % 16, column 4 (l / 16) + q), in place in Xs:
//   right: Xs[c][i] = sum_k Xs[k][i] Uinv[k][c]  (L21 = A21 Uinv; i < AB, c < NB; wave w: columns
//          16 w .., both 16-row blocks),
//   left:  Xs[r][c] = sum_k Linv[r][k] Xs[k][c]  (U12 = Linv A12; r < NB, c < AB; wave w: rows
//          16 w .., both 16-column blocks).
// The triangle's operands come in registers, tf[s] its entry (k = 4 s + l / 16, 16 w + l % 16) of
// Uinv (row-major) or (16 w + l % 16, k) of Linv (column-major): a contiguous 16-lane run either
// way, loaded by apply_fragments. Every thread of the workgroup calls it (it synchronizes).
__device__ __forceinline__ void apply_fragments(const float2* __restrict__ Tm, float2 (&tf)[NB / 4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < NB / 4; ++s) tf[s] = Tm[(4 * s + (lane >> 4)) * NB + 16 * wave + (lane & 15)];
}

__device__ __forceinline__ void apply_mfma(const float2 (&tf)[NB / 4], float2 (&Xs)[NB][AB + 1], bool right) {
  typedef Mfma<float> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
  F::v4 re[2], im[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) re[b] = im[b] = F::v4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < NB / 4; ++s) {
    const int k = 4 * s + kq;
    if (right) {
      const float2 u = tf[s];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const float2 x = Xs[k][16 * b + l16];
        re[b] = F::mma(u.x, x.x, re[b]);
        re[b] = F::mma(-u.y, x.y, re[b]);
        im[b] = F::mma(u.y, x.x, im[b]);
        im[b] = F::mma(u.x, x.y, im[b]);
      }
    } else {
      const float2 l = tf[s];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const float2 x = Xs[k][16 * b + l16];
        re[b] = F::mma(x.x, l.x, re[b]);
        re[b] = F::mma(-x.y, l.y, re[b]);
        im[b] = F::mma(x.y, l.x, im[b]);
        im[b] = F::mma(x.x, l.y, im[b]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 v = float2(re[b][q], im[b][q]);
      if (right)
        Xs[16 * wave + 4 * kq + q][16 * b + l16] = v;
      else
        Xs[16 * wave + l16][16 * b + 4 * kq + q] = v;
    }
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(256) apply_kernel(typename Cplx<T>::type* __restrict__ M, int lda, int lim, int ncols, int k0,
                                                    int kb, int rblocks, const typename Cplx<T>::type* __restrict__ Uinv,
                                                    const typename Cplx<T>::type* __restrict__ Linv,
                                                    const typename Cplx<T>::type* __restrict__ udiag, int* __restrict__ flag,
                                                    const int* __restrict__ gate, BatchStrides bs, SplitOps so, int split) {
  typedef typename Cplx<T>::type C2;
  if constexpr (std::is_same<T, float>::value) {
    so.ls += blockIdx.z * so.lstride;
    so.us += blockIdx.z * so.ustride;
  }
  {
    const int z = blockIdx.z;
    M += z * bs.m;
    Uinv += z * bs.uinv;
    Linv += z * bs.uinv;
    udiag += z * NB;
    if (flag) flag += z * bs.flag;
  }
  if (gate && *gate == 0) return;
  constexpr bool F32 = std::is_same<T, float>::value;
  // complex128: the triangle staged in LDS, Ts[r][k] its entry (r, k); complex64: in registers
  __shared__ C2 Ts[F32 ? 1 : NB][NB + 1];
  __shared__ C2 Xs[NB][AB + 1];  // right: [k][row]; left: [k][col]
  __shared__ int s_fail;
  const int tid = threadIdx.x;
  const bool right = (int)blockIdx.x < rblocks;
  const C2* __restrict__ Tm = right ? Uinv : Linv;
  float2 tf[NB / 4];
  if constexpr (F32) {
    apply_fragments(Tm, tf);
  } else {
    C2 t[NB * NB / 256];
#pragma unroll
    for (int j = 0; j < NB * NB / 256; ++j) t[j] = Tm[tid + 256 * j];
    // Uinv row-major, Linv column-major
#pragma unroll
    for (int j = 0; j < NB * NB / 256; ++j) {
      const int e = tid + 256 * j;
      if (right)
        Ts[e >> 6][e & 63] = t[j];
      else
        Ts[e & 63][e >> 6] = t[j];
    }
  }
  constexpr int VR = NB * AB / 256;  // outputs per thread
  if (right) {
    const int base = k0 + kb + blockIdx.x * AB;
    if (tid == 0) s_fail = 0;
    {
      C2 t[VR];
#pragma unroll
      for (int j = 0; j < VR; ++j) {
        const int e = tid + 256 * j, i = e % AB, k = e / AB;
        t[j] = (base + i < lim && k < kb) ? M[(size_t)(k0 + k) * lda + base + i] : C2(0.0, 0.0);
      }
#pragma unroll
      for (int j = 0; j < VR; ++j) {
        const int e = tid + 256 * j;
        Xs[e / AB][e % AB] = t[j];
      }
    }
    __syncthreads();
    static_assert(VR == 8 && VR * (256 / AB) == NB, "8 consecutive columns per thread");
    const int i = tid % AB, cq = tid / AB;  // row of the block, columns VR cq .. VR cq + 7
    C2 acc[VR];
    if constexpr (std::is_same<T, float>::value) {
      apply_mfma(tf, Xs, true);
#pragma unroll
      for (int v = 0; v < VR; ++v) acc[v] = Xs[VR * cq + v][i];
    } else {
#pragma unroll
      for (int v = 0; v < VR; ++v) acc[v] = C2(0.0, 0.0);
      for (int k = 0; k < NB; ++k) {
        const C2 x = Xs[k][i];
#pragma unroll
        for (int v = 0; v < VR; ++v) {
          const C2 u = Ts[k][VR * cq + v];
          acc[v].x = fma(x.x, u.x, fma(-x.y, u.y, acc[v].x));
          acc[v].y = fma(x.x, u.y, fma(x.y, u.x, acc[v].y));
        }
      }
    }
    bool fail = false;
    if (base + i < lim) {
#pragma unroll
      for (int v = 0; v < VR; ++v) {
        const int c = VR * cq + v;
        if (c >= kb) continue;
        M[(size_t)(k0 + c) * lda + base + i] = acc[v];
        if (flag) {
          const C2 d = udiag[c];
          if (mag1(cmul(acc[v], d)) > mag1(d)) fail = true;
        }
      }
      if constexpr (std::is_same<T, float>::value)
        if (split) {
          h8 part[4];
          split8(acc, part);
          h8* dst = so.ls + (size_t)((k0 - so.K0) / 8 + cq) * 4 * so.pad + base + i;
#pragma unroll
          for (int q = 0; q < 4; ++q) dst[(size_t)q * so.pad] = part[q];
        }
    }
    if (fail) s_fail = 1;
    __syncthreads();
    if (tid == 0 && s_fail && flag) *flag = 1;
  } else {
    const int base = k0 + kb + (blockIdx.x - rblocks) * AB;
    {
      C2 t[VR];
#pragma unroll
      for (int j = 0; j < VR; ++j) {
        const int e = tid + 256 * j, k = e & 63, c = e >> 6;
        t[j] = (base + c < ncols && k < kb) ? M[(size_t)(base + c) * lda + k0 + k] : C2(0.0, 0.0);
      }
#pragma unroll
      for (int j = 0; j < VR; ++j) {
        const int e = tid + 256 * j;
        Xs[e & 63][e >> 6] = t[j];
      }
    }
    __syncthreads();
    const int cl = tid % AB, rg = tid / AB;  // column of the block, rows VR rg .. VR rg + 7
    C2 acc[VR];
    if constexpr (std::is_same<T, float>::value) {
      apply_mfma(tf, Xs, false);
#pragma unroll
      for (int v = 0; v < VR; ++v) acc[v] = Xs[VR * rg + v][cl];
    } else {
#pragma unroll
      for (int v = 0; v < VR; ++v) acc[v] = C2(0.0, 0.0);
      for (int k = 0; k < NB; ++k) {
        const C2 x = Xs[k][cl];
#pragma unroll
        for (int v = 0; v < VR; ++v) {
          const C2 l = Ts[VR * rg + v][k];
          acc[v].x = fma(l.x, x.x, fma(-l.y, x.y, acc[v].x));
          acc[v].y = fma(l.x, x.y, fma(l.y, x.x, acc[v].y));
        }
      }
    }
    const int c = base + cl;
    if (c < ncols) {
#pragma unroll
      for (int v = 0; v < VR; ++v)
        if (VR * rg + v < kb) M[(size_t)c * lda + k0 + VR * rg + v] = acc[v];
      if constexpr (std::is_same<T, float>::value)
        if (split) {
          h8 part[4];
          split8(acc, part);
          h8* dst = so.us + (size_t)((k0 - so.K0) / 8 + rg) * 4 * so.pad + c;
#pragma unroll
          for (int q = 0; q < 4; ++q) dst[(size_t)q * so.pad] = part[q];
        }
    }
  }
}

// After the panel: the zero pivot found without interchanges counts when no fallback ran.
__global__ void panel_finish_kernel(const int* __restrict__ flag, const int* __restrict__ zp_fast,
                                    int* __restrict__ zero_pivot) {
  if (threadIdx.x == 0 && *flag == 0 && *zp_fast != 0x7fffffff) atomicMin(zero_pivot, *zp_fast);
}

// Block substitutions with the stored inverses of the diagonal blocks, on right-hand sides Y
// (column c of system z at Y + z * ystride + c * ldy, rows [0, n)), nrhs <= MAXRHS_LU per pass.

// One panel of a block substitution in one launch. The panel's block of the solution,
// Y_b = Tinv_b W_b (Uinv for U X = Y, Linv for L X = Y), is formed by every workgroup from the
// work copy W of the right-hand sides (its rows [k0, k0 + kb) are not written by this launch) and
// stored into Y by workgroup 0; each workgroup then updates its rows of W,
// W[r0 : r0 + nr] -= M[r0 : r0 + nr, k0 : k0 + kb] Y_b, one row per thread (rows above the block
// for U, below it for L). Grid (max(1, ceil(nr / 256)), 1, batch), 256 threads; nrhs <= MAXRHS_LU.
// active (optional): systems z with active[z] == 0 are left as they are.
template <typename T>
__global__ void __launch_bounds__(256) subst_panel_kernel(const typename Cplx<T>::type* __restrict__ M, int lda,
                                                          size_t mstride, const typename Cplx<T>::type* __restrict__ Tinv,
                                                          size_t tstride, typename Cplx<T>::type* __restrict__ W,
                                                          size_t wstride, int ldw, typename Cplx<T>::type* __restrict__ Y,
                                                          int ldy, size_t ystride, int k0, int kb, int r0, int nr,
                                                          int nrhs, int upper, const int* __restrict__ active) {
  typedef typename Cplx<T>::type C2;
  if (active && active[blockIdx.z] == 0) return;
  __shared__ C2 U[NB][NB + 1];
  __shared__ C2 Xb[NB][MAXRHS_LU + 1];
  M += blockIdx.z * mstride;
  Tinv += blockIdx.z * tstride;
  W += blockIdx.z * wstride;
  Y += blockIdx.z * ystride;
  const int tid = threadIdx.x;
  {
    C2 t[NB * NB / 256];
#pragma unroll
    for (int j = 0; j < NB * NB / 256; ++j) t[j] = Tinv[tid + 256 * j];
    // U[r][k]: Uinv row-major, Linv column-major
#pragma unroll
    for (int j = 0; j < NB * NB / 256; ++j) {
      const int e = tid + 256 * j;
      if (upper)
        U[e >> 6][e & 63] = t[j];
      else
        U[e & 63][e >> 6] = t[j];
    }
  }
  for (int e = tid; e < NB * nrhs; e += 256) {
    const int r = e & (NB - 1), c = e >> 6;
    Xb[r][c] = r < kb ? W[(size_t)c * ldw + k0 + r] : C2(0, 0);
  }
  __syncthreads();
  // Y_b in registers (at most MAXRHS_LU / 4 entries per thread), then over W_b in LDS
  constexpr int PER = NB * MAXRHS_LU / 256;
  C2 yb[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + 256 * q, r = e & (NB - 1), c = e >> 6;
    C2 s = C2(0, 0);
    if (c < nrhs && r < kb) {
      const int lo = upper ? r : 0, hi = upper ? kb : r + 1;
      for (int k = lo; k < hi; ++k) {
        const C2 u = U[r][k], y = Xb[k][c];
        s.x += u.x * y.x - u.y * y.y;
        s.y += u.x * y.y + u.y * y.x;
      }
    }
    yb[q] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + 256 * q, r = e & (NB - 1), c = e >> 6;
    if (c < nrhs) {
      Xb[r][c] = yb[q];
      if (blockIdx.x == 0 && r < kb) Y[(size_t)c * ldy + k0 + r] = yb[q];
    }
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + tid;
  if (i >= nr) return;
  const int row = r0 + i;
  const C2* col = M + (size_t)k0 * lda + row;
  for (int q0 = 0; q0 < nrhs; q0 += 8) {
    C2 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = C2(0, 0);
    auto term = [&](C2 u, int c) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const C2 x = Xb[c][min(q0 + q, MAXRHS_LU - 1)];
        acc[q].x = fma(u.x, x.x, fma(-u.y, x.y, acc[q].x));
        acc[q].y = fma(u.x, x.y, fma(u.y, x.x, acc[q].y));
      }
    };
    int c = 0;
    for (; c + 8 <= kb; c += 8) {
      C2 u[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) u[j] = col[(size_t)(c + j) * lda];
#pragma unroll
      for (int j = 0; j < 8; ++j) term(u[j], c + j);
    }
    for (; c < kb; ++c) term(col[(size_t)c * lda], c);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (q0 + q >= nrhs) break;
      C2* b = W + (size_t)(q0 + q) * ldw + row;
      C2 v = *b;
      v.x -= acc[q].x;
      v.y -= acc[q].y;
      *b = v;
    }
  }
}

// W[z][c][r] = Y[z][c][r] for the nrhs columns (the work copy of the block substitutions)
template <typename C2>
__global__ void copy_rhs_kernel(const C2* __restrict__ Y, int ldy, size_t ystride, C2* __restrict__ W, int ldw,
                                size_t wstride, int n, int nrhs, const int* __restrict__ active) {
  if (active && active[blockIdx.z] == 0) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * nrhs) return;
  const int r = (int)(e % n), c = (int)(e / n);
  W[blockIdx.z * wstride + (size_t)c * ldw + r] = Y[blockIdx.z * ystride + (size_t)c * ldy + r];
}

__global__ void fill_u32(unsigned* p, size_t n, unsigned v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_i32(int* p, size_t n, int v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ host driver
template <typename T>
struct LuWorkspace {
  typedef typename Cplx<T>::type C2;
  int n = 0, batch = 0;
  TsluWork w{};              // the tournament (FP64 only)
  int* ipiv = nullptr;       // [batch][n]
  int* nsys = nullptr;       // [batch]: orders of a batch of uneven systems (GemmBatch)
  int* flags = nullptr;      // [batch][panels]: 1 if the panel needed interchanges
  int* zp_fast = nullptr;    // [batch][panels]: zero pivot found without interchanges
  C2* uinv = nullptr;        // [batch][panels]: U11^-1 (NB x NB), for the back substitution
  C2* linv = nullptr;        // [batch][panels]: L11^-1, for the forward substitution
  C2* udiag = nullptr;       // [batch]
  C2* save = nullptr;        // the panel before the attempt without interchanges (pivoted path)
  C2* backup = nullptr;      // [A | B] of a single solve, restored for the pivoted path
  C2* subw = nullptr;        // [batch][MAXRHS_LU][n]: work copy of the right-hand sides (substitutions)
  size_t backup_elems = 0;
  // the split FP16 operands of a block's panels (complex64 LU, GEMM variants 3, 4 and 5)
  h8* ls = nullptr;
  h8* us = nullptr;
  int split_pad = 0, split_kmax = 0, split_batch = 0;
  unsigned long generation = 0;  // bumped at every reallocation (captured graphs hold the pointers)
  size_t split_stride() const { return (size_t)(split_kmax / 8) * 4 * split_pad; }
  // buffers for systems of order up to n, blocks of up to kmax columns, `batch` systems; zeroed
  // once (rows and columns past the matrix are read as zeros)
  void ensure_split(int n_, int batch_, int kmax) {
    if (ls && split_pad >= n_ + 256 && split_kmax >= kmax && split_batch >= batch_) return;
    (void)hipFree(ls);
    (void)hipFree(us);
    ls = us = nullptr;
    split_pad = std::max(split_pad, n_ + 256);  // the zero rows and columns that 128-wide tiles read past n + R
    split_kmax = std::max(split_kmax, kmax);
    split_batch = std::max(split_batch, batch_);
    const size_t bytes = split_stride() * split_batch * sizeof(h8);
    HBG_HIP(hipMalloc(&ls, bytes));
    HBG_HIP(hipMalloc(&us, bytes));
    HBG_HIP(hipMemset(ls, 0, bytes));
    HBG_HIP(hipMemset(us, 0, bytes));
    ++generation;
  }
  void release() {
    (void)hipFree(ls);
    (void)hipFree(us);
    (void)hipFree(w.cand);
    (void)hipFree(w.ccount);
    (void)hipFree(w.ctr);
    (void)hipFree(w.top);
    (void)hipFree(w.winners);
    (void)hipFree(w.zero_pivot);
    (void)hipFree(ipiv);
    (void)hipFree(nsys);
    (void)hipFree(flags);
    (void)hipFree(zp_fast);
    (void)hipFree(uinv);
    (void)hipFree(linv);
    (void)hipFree(udiag);
    (void)hipFree(save);
    (void)hipFree(backup);
    (void)hipFree(subw);
    const unsigned long g = generation;
    *this = LuWorkspace{};
    generation = g + 1;
  }
  BatchStrides strides() const {
    const int panels = (n + NB - 1) / NB;
    return BatchStrides{0, panels, n, (size_t)panels * NB * NB};
  }
  void ensure(int n_, int batch_) {
    if (n >= n_ && batch >= batch_ && ipiv) return;
    n_ = std::max(n_, n);
    batch_ = std::max(batch_, std::max(batch, 1));
    release();
    const int panels = (n_ + NB - 1) / NB;
    if (std::is_same<T, double>::value) {
      int leaves = (n_ + LEAF - 1) / LEAF;
      int levels = 1, nodes = leaves;
      while (nodes > 1) {
        nodes = (nodes + FAN - 1) / FAN;
        ++levels;
      }
      w.levels = levels;
      w.max_nodes = leaves;
      const size_t slots = (size_t)levels * leaves;
      HBG_HIP(hipMalloc(&w.cand, slots * NB * sizeof(int)));
      HBG_HIP(hipMalloc(&w.ccount, slots * sizeof(int)));
      HBG_HIP(hipMalloc(&w.ctr, slots * sizeof(unsigned)));
      HBG_HIP(hipMalloc(&w.top, (size_t)NB * NB * sizeof(double2)));
      HBG_HIP(hipMalloc(&w.winners, NB * sizeof(int)));
      HBG_HIP(hipMalloc(&w.zero_pivot, sizeof(int)));
      HBG_HIP(hipMalloc(&save, (size_t)n_ * NB * sizeof(C2)));
    }
    HBG_HIP(hipMalloc(&ipiv, (size_t)batch_ * n_ * sizeof(int)));
    HBG_HIP(hipMalloc(&nsys, (size_t)batch_ * sizeof(int)));
    HBG_HIP(hipMalloc(&flags, (size_t)batch_ * panels * sizeof(int)));
    HBG_HIP(hipMalloc(&zp_fast, (size_t)batch_ * panels * sizeof(int)));
    HBG_HIP(hipMalloc(&uinv, (size_t)batch_ * panels * NB * NB * sizeof(C2)));
    HBG_HIP(hipMalloc(&linv, (size_t)batch_ * panels * NB * NB * sizeof(C2)));
    HBG_HIP(hipMalloc(&udiag, (size_t)batch_ * NB * sizeof(C2)));
    HBG_HIP(hipMalloc(&subw, (size_t)batch_ * MAXRHS_LU * n_ * sizeof(C2)));
    n = n_;
    batch = batch_;
  }
};

// tuned settings per arithmetic: [0] complex128, [1] complex64
static int g_gemm_variant_[2] = {1, 1};
static int g_lu_depth_[2] = {4, 4};
static int g_lu_sub_[2] = {0, 0};   // panels per sub-block of a block (0: none)
static int g_lu_sub2_[2] = {0, 0};  // panels per sub-block of a sub-block (0: none)
template <typename T>
constexpr int prec_index() {
  return std::is_same<T, double>::value ? 0 : 1;
}
void lu_set_gemm_variant(int v) {
  g_gemm_variant_[0] = (v == 2 || v == 4 || v == 5) ? 2 : 1;  // complex128 has the FP64 kernels only
  g_gemm_variant_[1] = v;
}
int lu_gemm_variant() { return g_gemm_variant_[0]; }
void lu_set_gemm_variant32(int v) { g_gemm_variant_[1] = v; }
int lu_gemm_variant32() { return g_gemm_variant_[1]; }
void lu_set_depth32(int q) { g_lu_depth_[1] = q; }
int lu_depth32() { return g_lu_depth_[1]; }
void lu_set_depth(int q) { g_lu_depth_[0] = g_lu_depth_[1] = q; }
int lu_depth() { return g_lu_depth_[0]; }
void lu_set_sub2(int q) { g_lu_sub2_[0] = g_lu_sub2_[1] = q; }
void lu_set_sub2_32(int q) { g_lu_sub2_[1] = q; }
int lu_sub2() { return g_lu_sub2_[0]; }
int lu_sub2_32() { return g_lu_sub2_[1]; }
void lu_set_sub(int q) { g_lu_sub_[0] = g_lu_sub_[1] = q; }
void lu_set_sub32(int q) { g_lu_sub_[1] = q; }
int lu_sub() { return g_lu_sub_[0]; }
int lu_sub32() { return g_lu_sub_[1]; }

template <typename T>
static LuWorkspace<T>& workspace() {
  static thread_local LuWorkspace<T> ws;
  return ws;
}

// hipGraphs of the batched LU's launch sequence (no host work inside: fills, not memsets)
static int g_lu_graphs = 1;
void lu_set_graphs(int on) { g_lu_graphs = on; }
int lu_graphs() { return g_lu_graphs; }
// Bumped whenever the library frees device memory that a solve's matrices may live in: a
// captured graph replayed after its matrices' memory was freed and handed out again at the same
// address gave wrong factors (measured: a session freed, a new one with the same buffer addresses,
// its replayed LU 1e5 off the eager one), so a graph is replayed only within the epoch it was
// captured in.
static std::atomic<unsigned long> g_free_epoch{0};
void lu_graphs_note_free() { g_free_epoch.fetch_add(1, std::memory_order_relaxed); }

namespace {
struct GraphKey {
  const void* M;
  size_t mstride;
  int batch, n, lda, nrhs, variant, depth, real_bytes, uneven, sub;
  unsigned long generation, epoch;
  bool operator==(const GraphKey& o) const {
    return M == o.M && mstride == o.mstride && batch == o.batch && n == o.n && lda == o.lda && nrhs == o.nrhs &&
           variant == o.variant && depth == o.depth && real_bytes == o.real_bytes && uneven == o.uneven &&
           sub == o.sub && generation == o.generation && epoch == o.epoch;
  }
};
struct GraphCache {
  std::vector<std::pair<GraphKey, hipGraphExec_t>> items;
  hipGraphExec_t find(const GraphKey& k) {
    for (auto& it : items)
      if (it.first == k) return it.second;
    return nullptr;
  }
  void put(const GraphKey& k, hipGraphExec_t e) {
    if (items.size() >= 16) {  // the oldest goes
      (void)hipGraphExecDestroy(items.front().second);
      items.erase(items.begin());
    }
    items.push_back({k, e});
  }
  void clear() {
    for (auto& it : items) (void)hipGraphExecDestroy(it.second);
    items.clear();
  }
};
GraphCache& graphs() {
  static thread_local GraphCache g;
  return g;
}
}  // namespace

void lu_graphs_clear() { graphs().clear(); }

namespace {
struct EventPool {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  hipEvent_t next() {
    if (used == ev.size()) {
      hipEvent_t e;
      HBG_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
    return ev[used++];
  }
};
EventPool& events() {
  static thread_local EventPool p;
  return p;
}

// A second stream (default priority) for the work of a panel that can run beside its diagonal
// block, with the events that fork it from and join it to the calling stream (also inside a graph
// capture, where they become the graph's edges). One per thread, as the workspaces.
struct AuxStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
AuxStream& aux_stream() {
  static thread_local AuxStream a;
  if (!a.s) {
    HBG_HIP(hipStreamCreateWithFlags(&a.s, hipStreamNonBlocking));
    HBG_HIP(hipEventCreateWithFlags(&a.fork, hipEventDisableTiming));
    HBG_HIP(hipEventCreateWithFlags(&a.join, hipEventDisableTiming));
  }
  return a;
}

struct GemmTimer {
  EventPool* ep;
  LuStats* stats;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::vector<char> big;  // per launch: on the 128 x 128 tiles of gemm_f16s_big_kernel
};

// One launch of the trailing-update GEMM in the given tile variant: 1 64 x 64, 2 128 x 64 (FP32 or
// FP64 matrix cores); complex64 only: 3 64 x 64, 4 128 x 64, 5 128 x 128 (8 waves) on the FP16
// matrix cores with split operands (gemm_f16s_kernel), 6 the 128 x 128 tiles of
// gemm_f16s_big_kernel where both sides exceed 64 (the trailing updates) and 64 x 64 for the
// panels' 64-wide strips, 7 the same with the double-buffered big kernel.
template <typename T>
void gemm_launch(int variant, typename Cplx<T>::type* M, int lda, int r0, int nrows, int c0, int ncolsC, int kf,
                 int kk, const GemmBatch& gb, const SplitOps* so, hipStream_t stream) {
  if (nrows <= 0 || ncolsC <= 0) return;
  const int batch = gb.batch;
  const int tn = (ncolsC + 63) / 64;
  const unsigned g64 = gemm_grid((nrows + 63) / 64, tn, batch), g128 = gemm_grid((nrows + 127) / 128, tn, batch);
  if constexpr (std::is_same<T, float>::value) {
    if (variant >= 3) {
      if (!so || (kf - so->K0) % 8 || kk % H16_KC) throw Error{6, "split-FP16 GEMM without its split operands"};
      const int kg0 = (kf - so->K0) / 8;
      if (variant == 3)
        gemm_f16s_kernel<64, 64><<<g64, 256, 0, stream>>>(M, lda, r0, nrows, c0, ncolsC, *so, kg0, kk, gb);
      else if (variant == 4)
        gemm_f16s_kernel<128, 64><<<g128, 256, 0, stream>>>(M, lda, r0, nrows, c0, ncolsC, *so, kg0, kk, gb);
      else if (variant >= 6) {
        // the panels' 64-wide strips on 128 x 64 or 64 x 128 tiles, the trailing updates on 128 x 128
        const bool db = variant == 7;
        const int tm = nrows <= 64 ? 64 : 128, tn = ncolsC <= 64 ? 64 : 128;
        const unsigned g = gemm_grid((nrows + tm - 1) / tm, (ncolsC + tn - 1) / tn, batch);
        if (tm == 64 && tn == 64)
          gemm_f16s_kernel<64, 64><<<g64, 256, 0, stream>>>(M, lda, r0, nrows, c0, ncolsC, *so, kg0, kk, gb);
        else if (tm == 64)
          (db ? gemm_f16s_big_kernel<64, 128, true> : gemm_f16s_big_kernel<64, 128, false>)<<<g, 256, 0, stream>>>(
              M, lda, r0, nrows, c0, ncolsC, *so, kg0, kk, gb);
        else if (tn == 64)
          (db ? gemm_f16s_big_kernel<128, 64, true> : gemm_f16s_big_kernel<128, 64, false>)<<<g, 256, 0, stream>>>(
              M, lda, r0, nrows, c0, ncolsC, *so, kg0, kk, gb);
        else
          (db ? gemm_f16s_big_kernel<128, 128, true> : gemm_f16s_big_kernel<128, 128, false>)<<<g, 256, 0, stream>>>(
              M, lda, r0, nrows, c0, ncolsC, *so, kg0, kk, gb);
      }
      else
        gemm_f16s_kernel<128, 128, 2, 4><<<gemm_grid((nrows + 127) / 128, (ncolsC + 127) / 128, batch), 512, 0,
                                           stream>>>(M, lda, r0, nrows, c0, ncolsC, *so, kg0, kk, gb);
      return;
    }
  }
  if (variant == 2 || variant == 4)
    gemm_mfma_kernel<T, 128, 64><<<g128, 256, 0, stream>>>(M, lda, r0, nrows, c0, ncolsC, kf, kk, gb);
  else
    gemm_mfma_kernel<T, 64, 64><<<g64, 256, 0, stream>>>(M, lda, r0, nrows, c0, ncolsC, kf, kk, gb);
}

// M[r0 + i][c0 + j] -= sum_k M[r0 + i][kf + k] M[kf + k][c0 + j] (i < nrows, j < ncolsC, k < kk)
// on `batch` matrices, on the matrix cores; timed when stats are kept. Under the split-FP16
// variants the right-hand-side columns ride along in FP16: the caller scales them to |b| <= 1.
template <typename T>
void gemm_sub(typename Cplx<T>::type* M, int lda, int r0, int nrows, int c0, int ncolsC, int kf, int kk,
              const GemmBatch& gb, const SplitOps* so, hipStream_t stream, GemmTimer& gt) {
  const int batch = gb.batch;
  if (nrows <= 0 || ncolsC <= 0 || kk <= 0) return;
  hipEvent_t g0 = nullptr;
  if (gt.stats) {
    g0 = gt.ep->next();
    HBG_HIP(hipEventRecord(g0, stream));
  }
  gemm_launch<T>(g_gemm_variant_[prec_index<T>()], M, lda, r0, nrows, c0, ncolsC, kf, kk, gb, so, stream);
  if (gt.stats) {
    hipEvent_t g1 = gt.ep->next();
    HBG_HIP(hipEventRecord(g1, stream));
    gt.ev.push_back({g0, g1});
    const double fl = 8.0 * batch * (double)nrows * (double)ncolsC * (double)kk;
    gt.stats->gemm_flops += fl;
    gt.stats->gemm_launches += 1;
    const bool big = std::is_same<T, float>::value && g_gemm_variant_[prec_index<T>()] >= 6 && nrows > 64 &&
                     ncolsC > 64;
    gt.big.push_back(big ? 1 : 0);
    if (big) {
      gt.stats->big_flops += fl;
      gt.stats->big_launches += 1;
    }
  }
}

// Every panel of every matrix of the batch without interchanges: the diagonal block,
// L21 = A21 U11^-1 (both checking that partial pivoting would make no interchange; the panel's
// flag is set where it would) and U12 = L11^-1 A12. The panels go in blocks of the tuned depth: before
// a panel, the earlier panels of its block update its columns (rows from the panel down) and its
// rows (columns right of it); after the block one trailing update with K = the block's width, so
// the trailing matrix is read and written once per block rather than once per panel.
// one_stream: the panel rows on the calling stream too (no second stream, no events: the
// allocation-free solve, which creates nothing)
template <typename T>
void factor_optimistic(typename Cplx<T>::type* M, int n, int lda, int ncols, int batch, BatchStrides bs,
                       const int* nsys, hipStream_t stream, LuWorkspace<T>& ws, GemmTimer& gt,
                       bool one_stream = false) {
  const int Q = std::max(1, g_lu_depth_[prec_index<T>()]);
  const GemmBatch gb{bs.m, batch, nsys, n};
  // the split-FP16 GEMM reads its operands from the split buffers, filled panel by panel
  const bool split = std::is_same<T, float>::value && g_gemm_variant_[prec_index<T>()] >= 3;
  SplitOps so{ws.ls, ws.us, ws.split_pad, ws.split_stride(), ws.split_stride(), 0};
  const SplitOps* sop = split ? &so : nullptr;
  // sub-blocks of S panels inside a block: a panel's columns and rows are updated by the earlier
  // panels of its sub-block only; at the end of a sub-block the block's later columns (all rows
  // below) and later rows (the columns past the block) take the sub-block's panels in two wide
  // products (an L-shaped update), so that most of the block's in-block work runs as large GEMMs
  // instead of thin strips whose depth grows panel by panel. The trailing matrix past the block
  // still takes the whole block at once.
  // Levels nest: sub-blocks of S1 panels in a block, of S2 panels in a sub-block (either 0: none);
  // at the end of a sub-block that is not its parent's end, the L-shaped update over the parent's
  // later columns and rows, with the sub-block's panels.
  int sz[3] = {Q, 0, 0}, L = 0;
  for (int v : {g_lu_sub_[prec_index<T>()], g_lu_sub2_[prec_index<T>()]})
    if (v > 0 && v < sz[L]) sz[++L] = v;
  for (int K0 = 0; K0 < n; K0 += Q * NB) {
    const int qend = std::min(n, K0 + Q * NB);
    so.K0 = K0;
    for (int k0 = K0; k0 < qend; k0 += NB) {
      const int p = k0 / NB;
      const int kb = min(NB, n - k0);
      const int m = n - k0;
      // the panel's sub-blocks, level by level: start and end
      int st[3] = {K0, 0, 0}, en[3] = {qend, 0, 0};
      for (int l = 1; l <= L; ++l) {
        st[l] = st[l - 1] + (k0 - st[l - 1]) / (sz[l] * NB) * (sz[l] * NB);
        en[l] = std::min(st[l] + sz[l] * NB, en[l - 1]);
      }
      const int Ks = st[L];  // the panel's finest sub-block
      AuxStream* ax = nullptr;
      if (k0 > Ks && one_stream) {
        gemm_sub<T>(M, lda, k0, m, k0, kb, Ks, k0 - Ks, gb, sop, stream, gt);
        gemm_sub<T>(M, lda, k0, kb, k0 + kb, ncols - k0 - kb, Ks, k0 - Ks, gb, sop, stream, gt);
      } else if (k0 > Ks) {
        gemm_sub<T>(M, lda, k0, m, k0, kb, Ks, k0 - Ks, gb, sop, stream, gt);
        // the panel's rows (the columns right of it) are needed only by apply_kernel: they are
        // updated on the second stream beside the diagonal block (disjoint parts of M)
        ax = &aux_stream();
        HBG_HIP(hipEventRecord(ax->fork, stream));
        HBG_HIP(hipStreamWaitEvent(ax->s, ax->fork, 0));
        gemm_sub<T>(M, lda, k0, kb, k0 + kb, ncols - k0 - kb, Ks, k0 - Ks, gb, sop, ax->s, gt);
        HBG_HIP(hipEventRecord(ax->join, ax->s));
      }
      const size_t po = (size_t)p * NB * NB;
      diag_kernel<T><<<dim3(1, 1, batch), 256, 0, stream>>>(M, lda, k0, kb, 0, ws.flags + p, ws.zp_fast + p, ws.ipiv,
                                                            ws.uinv + po, ws.linv + po, ws.udiag, bs);
      if (ax) HBG_HIP(hipStreamWaitEvent(stream, ax->join, 0));
      const int rb = (m - kb + AB - 1) / AB, cb = (ncols - k0 - kb + AB - 1) / AB;
      if (rb + cb > 0)
        apply_kernel<T><<<dim3(rb + cb, 1, batch), 256, 0, stream>>>(M, lda, n, ncols, k0, kb, rb, ws.uinv + po,
                                                                     ws.linv + po, ws.udiag, ws.flags + p, nullptr, bs,
                                                                     so, split && k0 + NB < n ? 1 : 0);
      // the coarsest sub-block that ends here before its parent does: the L-shaped update of the
      // parent's later columns (all rows below) and later rows (the columns past the parent)
      const int se = k0 + kb;
      for (int l = 1; l <= L; ++l)
        if (se == en[l] && se < en[l - 1]) {
          const int pe = en[l - 1];
          gemm_sub<T>(M, lda, se, n - se, se, pe - se, st[l], se - st[l], gb, sop, stream, gt);
          gemm_sub<T>(M, lda, se, pe - se, pe, ncols - pe, st[l], se - st[l], gb, sop, stream, gt);
          break;
        }
    }
    gemm_sub<T>(M, lda, qend, n - qend, qend, ncols - qend, K0, qend - K0, gb, sop, stream, gt);
  }
}

// One FP64 matrix, panel by panel: the attempt without interchanges, then, gated by the panel's
// flag, the saved panel restored and pivoted by the tournament. The result of a panel without a
// flag is the one partial pivoting gives.
void factor_pivoted(double2* M, int n, int lda, int ncols, hipStream_t stream, LuWorkspace<double>& ws,
                    GemmTimer& gt) {
  const size_t slots = (size_t)ws.w.levels * ws.w.max_nodes;
  const BatchStrides bs = ws.strides();
  for (int k0 = 0, p = 0; k0 < n; k0 += NB, ++p) {
    const int kb = min(NB, n - k0);
    const int m = n - k0;
    int* flag = ws.flags + p;
    const size_t po = (size_t)p * NB * NB;
    double2* uinv = ws.uinv + po;
    double2* linv = ws.linv + po;
    const unsigned pe = (unsigned)(((size_t)m * kb + 255) / 256);
    panel_copy_kernel<<<pe, 256, 0, stream>>>(M, n, lda, k0, kb, ws.save, 0, flag);
    diag_kernel<double><<<1, 256, 0, stream>>>(M, lda, k0, kb, 0, flag, ws.zp_fast + p, ws.ipiv, uinv, linv, ws.udiag,
                                               bs);
    const int rb = (m - kb + AB - 1) / AB, cb = (ncols - k0 - kb + AB - 1) / AB;
    if (rb > 0)
      apply_kernel<double><<<rb, 256, 0, stream>>>(M, lda, n, ncols, k0, kb, rb, uinv, linv, ws.udiag, flag, nullptr,
                                                   bs, SplitOps{}, 0);
    panel_copy_kernel<<<pe, 256, 0, stream>>>(M, n, lda, k0, kb, ws.save, 1, flag);
    fill_u32<<<(unsigned)((slots + 255) / 256), 256, 0, stream>>>(ws.w.ctr, slots, 0u);
    tslu_kernel<<<(m + LEAF - 1) / LEAF, LEAF, 0, stream>>>(M, n, lda, k0, kb, ws.w, flag);
    ipiv_kernel<<<1, 64, 0, stream>>>(k0, kb, n, ws.w.winners, ws.ipiv, ws.w.zero_pivot, flag);
    swap_kernel<<<(ncols + 255) / 256, 256, 0, stream>>>(M, lda, ncols, k0, kb, ws.ipiv, ws.w.top, flag);
    diag_kernel<double><<<1, 256, 0, stream>>>(M, lda, k0, kb, 1, flag, ws.zp_fast + p, ws.ipiv, uinv, linv, ws.udiag,
                                               bs);
    if (rb > 0)
      apply_kernel<double><<<rb, 256, 0, stream>>>(M, lda, n, ncols, k0, kb, rb, uinv, linv, ws.udiag, nullptr, flag,
                                                   bs, SplitOps{}, 0);
    panel_finish_kernel<<<1, 64, 0, stream>>>(flag, ws.zp_fast + p, ws.w.zero_pivot);
    if (cb > 0)
      apply_kernel<double><<<cb, 256, 0, stream>>>(M, lda, n, ncols, k0, kb, 0, uinv, linv, ws.udiag, nullptr,
                                                   nullptr, bs, SplitOps{}, 0);
    gemm_sub<double>(M, lda, k0 + kb, n - k0 - kb, k0 + kb, ncols - k0 - kb, k0, kb, GemmBatch{0, 1, nullptr, n},
                     nullptr, stream, gt);
  }
}

// U X = Y (upper) or L X = Y (lower, unit) on right-hand sides Y (ldy, ystride) of every system of
// the batch, at most MAXRHS_LU columns at a time.
template <typename T>
void substitute(const typename Cplx<T>::type* M, int n, int lda, size_t mstride, typename Cplx<T>::type* Y, int ldy,
                size_t ystride, int nrhs, int batch, bool upper, hipStream_t stream, LuWorkspace<T>& ws,
                const int* active = nullptr) {
  const BatchStrides bs = ws.strides();
  const typename Cplx<T>::type* Tinv = upper ? ws.uinv : ws.linv;
  const int panels = (n + NB - 1) / NB;
  if (n > ws.n || batch > ws.batch) throw Error{6, "substitution beyond the LU workspace"};
  const size_t wstride = (size_t)MAXRHS_LU * ws.n;
  for (int r0 = 0; r0 < nrhs; r0 += MAXRHS_LU) {
    const int nr = min(MAXRHS_LU, nrhs - r0);
    typename Cplx<T>::type* Yc = Y + (size_t)r0 * ldy;
    copy_rhs_kernel<<<dim3((unsigned)(((size_t)n * nr + 255) / 256), 1, batch), 256, 0, stream>>>(
        Yc, ldy, ystride, ws.subw, ws.n, wstride, n, nr, active);
    for (int i = 0; i < panels; ++i) {
      const int p = upper ? panels - 1 - i : i;
      const int k0 = p * NB, kb = min(NB, n - k0);
      const int ur0 = upper ? 0 : k0 + kb, unr = upper ? k0 : n - k0 - kb;
      subst_panel_kernel<T><<<dim3((unsigned)std::max(1, (unr + 255) / 256), 1, batch), 256, 0, stream>>>(
          M, lda, mstride, Tinv + (size_t)p * NB * NB, bs.uinv, ws.subw, wstride, ws.n, Yc, ldy, ystride, k0, kb, ur0,
          unr, nr, upper ? 1 : 0, active);
    }
  }
}

void finish_stats(LuStats* stats, EventPool& ep, hipEvent_t t0, GemmTimer& gt, hipStream_t stream) {
  if (!stats) return;
  hipEvent_t t1 = ep.next();
  HBG_HIP(hipEventRecord(t1, stream));
  HBG_HIP(hipEventSynchronize(t1));
  float ms = 0.0f;
  HBG_HIP(hipEventElapsedTime(&ms, t0, t1));
  stats->total_ms += ms;
  for (size_t i = 0; i < gt.ev.size(); ++i) {
    HBG_HIP(hipEventElapsedTime(&ms, gt.ev[i].first, gt.ev[i].second));
    stats->gemm_ms += ms;
    if (i < gt.big.size() && gt.big[i]) stats->big_ms += ms;
  }
  gt.ev.clear();
  gt.big.clear();
}

template <typename T>
void solve_batched(typename Cplx<T>::type* M, size_t mstride, int batch, int n, int lda, int nrhs, hipStream_t stream,
                   LuStats* stats, int* status, const int* nsys_host = nullptr) {
  EventPool& ep = events();
  ep.used = 0;
  GemmTimer gt{&ep, stats, {}};
  hipEvent_t t0 = nullptr;
  if (stats) {
    t0 = ep.next();
    HBG_HIP(hipEventRecord(t0, stream));
  }
  LuWorkspace<T>& ws = workspace<T>();
  ws.ensure(n, batch);
  if (std::is_same<T, float>::value && g_gemm_variant_[prec_index<T>()] >= 3)
    ws.ensure_split(n, batch, NB * std::max(1, g_lu_depth_[prec_index<T>()]));
  BatchStrides bs = ws.strides();
  bs.m = mstride;
  const int ncols = n + nrhs;
  const int panels = (n + NB - 1) / NB;
  const size_t pf = (size_t)bs.flag * batch;
  // uneven orders: copied before the launches (the captured graph reads them from ws.nsys)
  if (nsys_host) HBG_HIP(hipMemcpyAsync(ws.nsys, nsys_host, batch * sizeof(int), hipMemcpyHostToDevice, stream));
  const int* nsys = nsys_host ? ws.nsys : nullptr;
  auto launches = [&]() {
    fill_i32<<<(unsigned)((pf + 255) / 256), 256, 0, stream>>>(ws.flags, pf, 0);
    fill_i32<<<(unsigned)((pf + 255) / 256), 256, 0, stream>>>(ws.zp_fast, pf, 0x7fffffff);
    factor_optimistic<T>(M, n, lda, ncols, batch, bs, nsys, stream, ws, gt);
    if (nrhs > 0) substitute<T>(M, n, lda, mstride, M + (size_t)n * lda, lda, mstride, nrhs, batch, true, stream, ws);
  };
  if (g_lu_graphs && !stats) {
    // the launch sequence depends only on these; replayed from a captured hipGraph
    const GraphKey key{(const void*)M, mstride, batch, n, lda, nrhs, g_gemm_variant_[prec_index<T>()],
                       g_lu_depth_[prec_index<T>()],
                       std::is_same<T, double>::value ? 8 : 4, nsys ? 1 : 0,
                       g_lu_sub_[prec_index<T>()] + 100 * g_lu_sub2_[prec_index<T>()], ws.generation,
                       g_free_epoch.load(std::memory_order_relaxed)};
    GraphCache& gc = graphs();
    hipGraphExec_t exec = gc.find(key);
    if (!exec) {
      hipGraph_t graph = nullptr;
      HBG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      launches();
      HBG_HIP(hipStreamEndCapture(stream, &graph));
      HBG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      HBG_HIP(hipGraphDestroy(graph));
      gc.put(key, exec);
    }
    HBG_HIP(hipGraphLaunch(exec, stream));
  } else {
    launches();
  }
  HBG_HIP(hipGetLastError());
  std::vector<int> flags(pf), zps(pf);
  HBG_HIP(hipMemcpyAsync(flags.data(), ws.flags, pf * sizeof(int), hipMemcpyDeviceToHost, stream));
  HBG_HIP(hipMemcpyAsync(zps.data(), ws.zp_fast, pf * sizeof(int), hipMemcpyDeviceToHost, stream));
  finish_stats(stats, ep, t0, gt, stream);
  HBG_HIP(hipStreamSynchronize(stream));
  for (int b = 0; b < batch; ++b) {
    bool any = false;
    int zp = 0x7fffffff;
    for (int p = 0; p < panels; ++p) {
      any = any || flags[(size_t)b * bs.flag + p] != 0;
      zp = std::min(zp, zps[(size_t)b * bs.flag + p]);
    }
    status[b] = any ? -1 : (zp == 0x7fffffff ? 0 : zp);
  }
}

}  // namespace

// Factor and solve `batch` systems [A | B] (n x (n + nrhs), leading dimension lda, matrix b at
// M + b * mstride) without interchanges, checking on every panel that partial pivoting would make
// none. status[b]: 0 solved (the partially pivoted result), k > 0 zero pivot k (1-based; the
// column is zero on and below the diagonal, as partial pivoting would find it), -1 a panel needs
// interchanges: the matrix is left overwritten and must be rebuilt and given to lu_solve_pivoted.
//
// nsys: for a batch of systems of uneven order, their orders (host array): system b of order
// nsys[b] <= n stored at order n with the identity on rows and columns [nsys[b], n) and zero
// right-hand-side rows there, so that its LU is block-diagonal; the GEMMs skip the zero updates
// of the identity blocks.
void lu_solve_batched(double2* M, size_t mstride, int batch, int n, int lda, int nrhs, hipStream_t stream,
                      LuStats* stats, int* status, const int* nsys) {
  solve_batched<double>(M, mstride, batch, n, lda, nrhs, stream, stats, status, nsys);
}

// The same in complex64 (FP32 arithmetic, FP32 MFMA); the factors and the inverses of their
// diagonal blocks stay for lu_resolve32 until the next call.
void lu_solve_batched32(float2* M, size_t mstride, int batch, int n, int lda, int nrhs, hipStream_t stream,
                        LuStats* stats, int* status, const int* nsys) {
  solve_batched<float>(M, mstride, batch, n, lda, nrhs, stream, stats, status, nsys);
}

// X = U^-1 L^-1 X with the factors of the last lu_solve_batched32 (systems without interchanges):
// column c of system b at X + b * xstride + c * ldx; systems b with active[b] == 0 (device array,
// optional) are left as they are.
void lu_resolve32(const float2* M, size_t mstride, int batch, int n, int lda, float2* X, int ldx, size_t xstride,
                  int nrhs, hipStream_t stream, const int* active) {
  LuWorkspace<float>& ws = workspace<float>();
  substitute<float>(M, n, lda, mstride, X, ldx, xstride, nrhs, batch, false, stream, ws, active);
  substitute<float>(M, n, lda, mstride, X, ldx, xstride, nrhs, batch, true, stream, ws, active);
  HBG_HIP(hipGetLastError());
}

namespace {

// synthetic systems for the autotuner: |off-diagonal| <= 1 / n, diagonal 2 pi (as the BEM's free
// term), so the pass without interchanges succeeds
template <typename T>
__global__ void synth_kernel(typename Cplx<T>::type* M, size_t count, int n, int lda, unsigned long long seed) {
  typedef typename Cplx<T>::type C2;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
    const size_t within = e % ((size_t)lda * (n + 1));
    const int i = (int)(within % lda), j = (int)(within / lda);
    unsigned long long x = seed ^ (e * 0x9E3779B97F4A7C15ull);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    const double u = (double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    const double v = (double)((x >> 3) & 0xFFFF) / 65536.0 - 0.5;
    const double d = (i == j && j < n) ? 6.283185307179586 : 0.0;
    M[e] = C2((T)(d + 2.0 * u / n), (T)(2.0 * v / n));
  }
}

}  // namespace

// Time the batched LU on synthetic systems (n, nrhs) in three stages: every GEMM tile variant (1,
// 2; complex64 also 3 to 7, the split-FP16 kernels) and block depth (2, 4, 8, 12, 16) at the
// largest batch of 4, 8, 16, 32, 48 that is at most max_batch and fits in mem_bytes; deeper blocks
// with nested sub-blocks (depth, sub, sub of sub) = (16, 8, 0), (24, 8, 0), (24, 12, 0),
// (32, 16, 0), (32, 16, 8), (64, 32, 16) at the fastest variant; the smaller batches at the
// fastest of all. Keeps the fastest variant, depth and sub-blocks and returns the fastest batch
// (by time per system). rows of report: variant, depth, sub-block, its sub-block, batch, ms per
// system (up to max_rows rows).
template <typename T>
int autotune_lu(int n, int nrhs, int max_batch, size_t mem_bytes, hipStream_t stream, double* report, int max_rows,
                int* rows) {
  typedef typename Cplx<T>::type C2;
  const size_t ms = (size_t)n * (n + nrhs);
  int bmax = (int)std::min<size_t>((size_t)std::max(max_batch, 1), std::max<size_t>(1, mem_bytes / (ms * sizeof(C2))));
  std::vector<int> batches;
  for (int b : {4, 8, 16, 32, 48})
    if (b <= bmax) batches.push_back(b);
  if (batches.empty()) batches.push_back(1);
  C2* M = nullptr;
  HBG_HIP(hipMalloc(&M, ms * batches.back() * sizeof(C2)));
  hipEvent_t e0, e1;
  HBG_HIP(hipEventCreate(&e0));
  HBG_HIP(hipEventCreate(&e1));
  int& gv = g_gemm_variant_[prec_index<T>()];
  int& gd = g_lu_depth_[prec_index<T>()];
  int& gs = g_lu_sub_[prec_index<T>()];
  int& gs2 = g_lu_sub2_[prec_index<T>()];
  const int old_v = gv, old_d = gd, old_s = gs, old_s2 = gs2;
  double best = 1e300;
  int bv = old_v, bd = old_d, bs = old_s, bs2 = old_s2, bb = batches.back();
  *rows = 0;
  std::vector<int> status(batches.back());
  try {
    const std::vector<int> variants = std::is_same<T, float>::value ? std::vector<int>{1, 2, 3, 4, 5, 6, 7}
                                                                      : std::vector<int>{1, 2};
    std::vector<std::array<int, 5>> configs;  // variant, depth, sub-block, its sub-block, batch
    for (int v : variants)
      for (int d : {2, 4, 8, 12, 16}) configs.push_back({v, d, 0, 0, batches.back()});
    const std::array<int, 3> nested[] = {{16, 8, 0}, {24, 8, 0}, {24, 12, 0}, {32, 16, 0}, {32, 16, 8}, {64, 32, 16}};
    const size_t stage1 = configs.size(), stage2 = stage1 + sizeof(nested) / sizeof(nested[0]);
    for (size_t c = 0; c < configs.size(); ++c) {
      const int v = configs[c][0], d = configs[c][1], sb = configs[c][2], sb2 = configs[c][3], b = configs[c][4];
      {
          gv = v;
          gd = d;
          gs = sb;
          gs2 = sb2;
          double ms_sys = 0.0;
          for (int rep = 0; rep < 2; ++rep) {  // the first captures the graph
            synth_kernel<T><<<4096, 256, 0, stream>>>(M, ms * b, n, n, 1234567ull + rep);
            HBG_HIP(hipEventRecord(e0, stream));
            solve_batched<T>(M, ms, b, n, n, nrhs, stream, nullptr, status.data());
            HBG_HIP(hipEventRecord(e1, stream));
            HBG_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            HBG_HIP(hipEventElapsedTime(&t, e0, e1));
            ms_sys = t / b;
          }
          if (*rows < max_rows && report) {
            double* r = report + 6 * (*rows);
            r[0] = v;
            r[1] = d;
            r[2] = sb;
            r[3] = sb2;
            r[4] = b;
            r[5] = ms_sys;
            ++*rows;
          }
          if (ms_sys < best) {
            best = ms_sys;
            bv = v;
            bd = d;
            bs = sb;
            bs2 = sb2;
            bb = b;
          }
      }
      // second stage: deeper blocks with nested sub-blocks at the fastest variant; third: the
      // smaller batches at the fastest of all
      if (c + 1 == stage1)
        for (const auto& ds : nested) configs.push_back({bv, ds[0], ds[1], ds[2], batches.back()});
      if (c + 1 == stage2)
        for (size_t i = 0; i + 1 < batches.size(); ++i) configs.push_back({bv, bd, bs, bs2, batches[i]});
    }
  } catch (...) {
    gv = old_v;
    gd = old_d;
    gs = old_s;
    gs2 = old_s2;
    (void)hipFree(M);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  gv = bv;
  gd = bd;
  gs = bs;
  gs2 = bs2;
  graphs().clear();  // the captures hold the tuning buffer
  (void)hipFree(M);
  HBG_HIP(hipEventDestroy(e0));
  HBG_HIP(hipEventDestroy(e1));
  return bb;
}

int lu_autotune(int precision64, int n, int nrhs, int max_batch, size_t mem_bytes, hipStream_t stream, double* report,
                int max_rows, int* rows) {
  return precision64 ? autotune_lu<double>(n, nrhs, max_batch, mem_bytes, stream, report, max_rows, rows)
                     : autotune_lu<float>(n, nrhs, max_batch, mem_bytes, stream, report, max_rows, rows);
}

// One system with partial pivoting where a panel needs it (tournament pivoting over several
// leaves); returns the first zero pivot (1-based) or 0.
int lu_solve_pivoted(double2* M, int n, int lda, int nrhs, hipStream_t stream, LuStats* stats) {
  EventPool& ep = events();
  ep.used = 0;
  GemmTimer gt{&ep, stats, {}};
  hipEvent_t t0 = nullptr;
  if (stats) {
    t0 = ep.next();
    HBG_HIP(hipEventRecord(t0, stream));
  }
  LuWorkspace<double>& ws = workspace<double>();
  ws.ensure(n, 1);
  const int ncols = n + nrhs;
  const int panels = (n + NB - 1) / NB;
  fill_i32<<<1, 64, 0, stream>>>(ws.w.zero_pivot, 1, 0x7fffffff);
  fill_i32<<<(panels + 255) / 256, 256, 0, stream>>>(ws.flags, panels, 0);
  fill_i32<<<(panels + 255) / 256, 256, 0, stream>>>(ws.zp_fast, panels, 0x7fffffff);
  factor_pivoted(M, n, lda, ncols, stream, ws, gt);
  if (nrhs > 0) substitute<double>(M, n, lda, 0, M + (size_t)n * lda, lda, 0, nrhs, 1, true, stream, ws);
  HBG_HIP(hipGetLastError());
  int zp = 0;
  std::vector<int> flags(panels);
  HBG_HIP(hipMemcpyAsync(&zp, ws.w.zero_pivot, sizeof(int), hipMemcpyDeviceToHost, stream));
  HBG_HIP(hipMemcpyAsync(flags.data(), ws.flags, panels * sizeof(int), hipMemcpyDeviceToHost, stream));
  finish_stats(stats, ep, t0, gt, stream);
  HBG_HIP(hipStreamSynchronize(stream));
  if (stats)
    for (int f : flags) stats->fallback_panels += f != 0;
  return zp == 0x7fffffff ? 0 : zp;
}

// One system [A | B] in place: the batched path on a batch of one, a copy of the matrix kept to
// restore it for the pivoted path when a panel needs interchanges. Returns the first zero pivot
// (1-based) or 0.
int lu_solve_device(double2* M, int n, int lda, int nrhs, hipStream_t stream, LuStats* stats) {
  LuWorkspace<double>& ws = workspace<double>();
  ws.ensure(n, 1);
  const size_t whole = (size_t)lda * (n + nrhs);
  if (ws.backup_elems < whole) {
    (void)hipFree(ws.backup);
    ws.backup = nullptr;
    ws.backup_elems = 0;
    HBG_HIP(hipMalloc(&ws.backup, whole * sizeof(double2)));
    ws.backup_elems = whole;
  }
  HBG_HIP(hipMemcpyAsync(ws.backup, M, whole * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  int st = 0;
  lu_solve_batched(M, whole, 1, n, lda, nrhs, stream, stats, &st, nullptr);
  if (st >= 0) return st;
  HBG_HIP(hipMemcpyAsync(M, ws.backup, whole * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  return lu_solve_pivoted(M, n, lda, nrhs, stream, stats);
}

// ---- one FP64 system in a region the caller owns (the allocation-free solve): the workspace of
// a batch of one carved out of the region, the launches eager on one stream.
namespace {
constexpr size_t REGION_ALIGN = 256;
size_t region_take(size_t& off, size_t bytes) {
  const size_t o = (off + REGION_ALIGN - 1) / REGION_ALIGN * REGION_ALIGN;
  off = o + bytes;
  return o;
}
// the workspace's pointers into base (when given) for order n; returns the bytes it spans
size_t carve_workspace(LuWorkspace<double>* ws, char* base, int n) {
  const int panels = (n + NB - 1) / NB;
  int leaves = (n + LEAF - 1) / LEAF, levels = 1, nodes = leaves;
  while (nodes > 1) {
    nodes = (nodes + FAN - 1) / FAN;
    ++levels;
  }
  const size_t slots = (size_t)levels * leaves;
  size_t off = 0;
  const size_t o_cand = region_take(off, slots * NB * sizeof(int));
  const size_t o_ccount = region_take(off, slots * sizeof(int));
  const size_t o_ctr = region_take(off, slots * sizeof(unsigned));
  const size_t o_top = region_take(off, (size_t)NB * NB * sizeof(double2));
  const size_t o_win = region_take(off, NB * sizeof(int));
  const size_t o_zp = region_take(off, sizeof(int));
  const size_t o_save = region_take(off, (size_t)n * NB * sizeof(double2));
  const size_t o_ipiv = region_take(off, (size_t)n * sizeof(int));
  const size_t o_nsys = region_take(off, sizeof(int));
  const size_t o_flags = region_take(off, (size_t)panels * sizeof(int));
  const size_t o_zpf = region_take(off, (size_t)panels * sizeof(int));
  const size_t o_uinv = region_take(off, (size_t)panels * NB * NB * sizeof(double2));
  const size_t o_linv = region_take(off, (size_t)panels * NB * NB * sizeof(double2));
  const size_t o_udiag = region_take(off, (size_t)NB * sizeof(double2));
  const size_t o_subw = region_take(off, (size_t)MAXRHS_LU * n * sizeof(double2));
  if (ws && base) {
    ws->n = n;
    ws->batch = 1;
    ws->w.levels = levels;
    ws->w.max_nodes = leaves;
    ws->w.cand = (int*)(base + o_cand);
    ws->w.ccount = (int*)(base + o_ccount);
    ws->w.ctr = (unsigned*)(base + o_ctr);
    ws->w.top = (double2*)(base + o_top);
    ws->w.winners = (int*)(base + o_win);
    ws->w.zero_pivot = (int*)(base + o_zp);
    ws->save = (double2*)(base + o_save);
    ws->ipiv = (int*)(base + o_ipiv);
    ws->nsys = (int*)(base + o_nsys);
    ws->flags = (int*)(base + o_flags);
    ws->zp_fast = (int*)(base + o_zpf);
    ws->uinv = (double2*)(base + o_uinv);
    ws->linv = (double2*)(base + o_linv);
    ws->udiag = (double2*)(base + o_udiag);
    ws->subw = (double2*)(base + o_subw);
  }
  return off;
}
}  // namespace

size_t lu_region_workspace_bytes(int n) { return carve_workspace(nullptr, nullptr, n); }

// M (n x (n + nrhs), leading dimension lda) solved with the workspace in `region` (at least
// lu_region_workspace_bytes(n) bytes, device memory): pivoted 0 the attempt without interchanges
// (returns 0, the first zero pivot k > 0, or -1: a panel needs interchanges, M overwritten),
// pivoted 1 the partially pivoted factorization (returns 0 or the first zero pivot). No
// allocation, no graph, no second stream; every flag and counter of the workspace is filled by a
// kernel before it is read, so the region's earlier contents do not matter.
int lu_solve_region(double2* M, int n, int lda, int nrhs, void* region, size_t region_bytes, int pivoted,
                    hipStream_t stream) {
  if (lu_region_workspace_bytes(n) > region_bytes) throw Error{6, "LU workspace beyond its region"};
  LuWorkspace<double> ws;  // views into the region: never ensured or released
  carve_workspace(&ws, (char*)region, n);
  const BatchStrides bs = ws.strides();
  EventPool ep;
  GemmTimer gt{&ep, nullptr, {}};
  const int ncols = n + nrhs;
  const int panels = (n + NB - 1) / NB;
  fill_i32<<<(panels + 255) / 256, 256, 0, stream>>>(ws.flags, panels, 0);
  fill_i32<<<(panels + 255) / 256, 256, 0, stream>>>(ws.zp_fast, panels, 0x7fffffff);
  if (pivoted) {
    fill_i32<<<1, 64, 0, stream>>>(ws.w.zero_pivot, 1, 0x7fffffff);
    factor_pivoted(M, n, lda, ncols, stream, ws, gt);
  } else {
    BatchStrides b1 = bs;
    b1.m = (size_t)lda * ncols;
    factor_optimistic<double>(M, n, lda, ncols, 1, b1, nullptr, stream, ws, gt, true);
  }
  if (nrhs > 0) substitute<double>(M, n, lda, 0, M + (size_t)n * lda, lda, 0, nrhs, 1, true, stream, ws);
  HBG_HIP(hipGetLastError());
  std::vector<int> flags(panels), zps(panels);
  int zp = 0x7fffffff;
  HBG_HIP(hipMemcpyAsync(flags.data(), ws.flags, panels * sizeof(int), hipMemcpyDeviceToHost, stream));
  HBG_HIP(hipMemcpyAsync(zps.data(), ws.zp_fast, panels * sizeof(int), hipMemcpyDeviceToHost, stream));
  if (pivoted) HBG_HIP(hipMemcpyAsync(&zp, ws.w.zero_pivot, sizeof(int), hipMemcpyDeviceToHost, stream));
  HBG_HIP(hipStreamSynchronize(stream));
  if (pivoted) return zp == 0x7fffffff ? 0 : zp;
  bool any = false;
  for (int p = 0; p < panels; ++p) {
    any = any || flags[p] != 0;
    zp = std::min(zp, zps[p]);
  }
  return any ? -1 : (zp == 0x7fffffff ? 0 : zp);
}

void lu_release() {
  graphs().clear();
  workspace<double>().release();
  workspace<float>().release();
}

}  // namespace hbg
