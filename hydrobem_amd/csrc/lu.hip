// Dense complex LU of the system matrices, FP64, on the GPU (linalg.rs:134-252 semantics).
//
// The matrix is stored column-major with the right-hand sides appended as extra columns
// ([A | B], n x (n + r), leading dimension lda), so that the row interchanges, the block row of U
// and the trailing updates carry the forward substitution L^-1 P B along; a blocked back
// substitution then solves U X = L^-1 P B.
//
// Right-looking, in panels of NB columns:
//   1. pivot rows of the panel by a tournament (tslu_kernel, one launch): every workgroup
//      takes up to 256 rows of the panel, one row per thread in registers, and eliminates with
//      partial pivoting (the pivot is the entry of largest |re| + |im|, as LAPACK's izamax; on a
//      tie the lower row index), the pivot search a wavefront shuffle reduction plus one LDS
//      exchange across the four wavefronts; its NB pivot rows go up a tree of 4-way merges that
//      repeat the elimination on the original values of the candidates; the last workgroup to
//      finish a node merges it (atomic counters, no waiting). The root's elimination gives the
//      panel's diagonal block L11 \ U11. With one leaf this is partial pivoting exactly; with
//      more it is tournament pivoting (Grigori, Demmel & Xiang 2011), whose pivots can differ
//      from partial pivoting's while keeping its stability in practice;
//   2. the interchanges in LAPACK's ipiv form, applied to every column (swap_kernel);
//   3. L21 = A21 U11^-1 (one row per thread) and U12 = L11^-1 A12 (one column per thread);
//   4. the trailing update A22 -= L21 U12 (gemm_update_kernel).
// A zero pivot leaves the factorization going and records the first such column; the caller
// reports "singular system matrix (zero pivot k)" as linalg.rs:221-233 does.
#include <vector>

#include "hbg_internal.h"

namespace hbg {

constexpr int NB = 64;      // panel width
constexpr int LEAF = 256;   // rows per tournament leaf (one per thread)
constexpr int FAN = 4;      // candidate sets per merge

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

namespace {

__device__ __forceinline__ double mag1(double2 v) { return fabs(v.x) + fabs(v.y); }

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
// written to top[step][*]. All threads of the workgroup must call it.
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
                                                    TsluWork w) {
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

// LAPACK-style interchanges of the panel from its pivot rows, then applied to all columns
// [0, ncols) of the augmented matrix; the root's L11 \ U11 written into the diagonal block.
__global__ void ipiv_kernel(int k0, int kb, const int* __restrict__ winners, int* __restrict__ ipiv) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // positions touched so far: (position, original row now there)
  int pos[2 * NB], who[2 * NB];
  int t = 0;
  auto at = [&](int p) {  // original row now at position p
    for (int i = 0; i < t; ++i)
      if (pos[i] == p) return who[i];
    return p;
  };
  auto where = [&](int r) {  // position of original row r
    for (int i = 0; i < t; ++i)
      if (who[i] == r) return pos[i];
    // untouched rows sit where they started, unless that position was overwritten
    return r;
  };
  auto set = [&](int p, int r) {
    for (int i = 0; i < t; ++i)
      if (pos[i] == p) {
        who[i] = r;
        return;
      }
    pos[t] = p;
    who[t] = r;
    ++t;
  };
  for (int i = 0; i < kb; ++i) {
    const int r = winners[i];
    const int p = where(r);
    ipiv[k0 + i] = p;
    const int o = at(k0 + i);
    set(k0 + i, r);
    set(p, o);
  }
}

__global__ void swap_kernel(double2* __restrict__ A, int lda, int ncols, int k0, int kb, const int* __restrict__ ipiv,
                            const double2* __restrict__ top) {
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

// The inverses of the diagonal block's factors: Uinv = U11^-1 (upper) and Linv = L11^-1 (unit
// lower), one column per thread by substitution in LDS; kb <= NB, unused entries zero.
__global__ void __launch_bounds__(2 * NB) trinv_kernel(const double2* __restrict__ A, int lda, int k0, int kb,
                                                        double2* __restrict__ Uinv, double2* __restrict__ Linv) {
  __shared__ double2 T[NB][NB + 1];
  __shared__ double2 X[NB][NB + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int r = e % NB, c = e / NB;
    T[r][c] = (r < kb && c < kb) ? A[(size_t)(k0 + c) * lda + k0 + r] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int j = tid % NB;
  const bool upper = tid < NB;
  if (j < kb) {
    if (upper) {
      // U X = e_j: X[i][j] for i = j .. 0
      for (int i = kb - 1; i > j; --i) X[i][j] = make_double2(0.0, 0.0);
      const double2 d = T[j][j];
      X[j][j] = cdiv(make_double2(1.0, 0.0), d);
      for (int i = j - 1; i >= 0; --i) {
        double2 s = make_double2(0.0, 0.0);
        for (int l = i + 1; l <= j; ++l) {
          const double2 u = T[i][l], x = X[l][j];
          s.x += u.x * x.x - u.y * x.y;
          s.y += u.x * x.y + u.y * x.x;
        }
        X[i][j] = cdiv(make_double2(-s.x, -s.y), T[i][i]);
      }
      for (int i = 0; i < NB; ++i) Uinv[i * NB + j] = i < kb ? X[i][j] : make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  if (j < kb && !upper) {
    // L X = e_j, L unit lower: X[i][j] for i = j .. kb - 1
    for (int i = 0; i < j; ++i) X[i][j] = make_double2(0.0, 0.0);
    X[j][j] = make_double2(1.0, 0.0);
    for (int i = j + 1; i < kb; ++i) {
      double2 s = make_double2(0.0, 0.0);
      for (int l = j; l < i; ++l) {
        const double2 v = T[i][l], x = X[l][j];
        s.x += v.x * x.x - v.y * x.y;
        s.y += v.x * x.y + v.y * x.x;
      }
      X[i][j] = make_double2(-s.x, -s.y);
    }
    for (int i = 0; i < NB; ++i) Linv[i * NB + j] = i < kb ? X[i][j] : make_double2(0.0, 0.0);
  }
  if (j >= kb) {
    for (int i = 0; i < NB; ++i) (upper ? Uinv : Linv)[i * NB + j] = make_double2(0.0, 0.0);
  }
}

// In-place products with a kb x kb triangle inverse Tm (row-major NB x NB):
//   right: X = X Tm for X = rows [r0, n) of columns [k0, k0 + kb)      (L21 = A21 U11^-1)
//   left:  X = Tm X for X = rows [k0, k0 + kb) of columns [c0, ncols)  (U12 = L11^-1 A12)
// 64 rows (right) or 64 columns (left) per workgroup, the block and Tm staged in LDS.
__global__ void __launch_bounds__(256) tri_apply_kernel(double2* __restrict__ M, int lda, int lim, int k0, int kb,
                                                        const double2* __restrict__ Tm, int right) {
  __shared__ double2 Xs[NB][NB + 1];  // right: [row][k]; left: [k][col]
  __shared__ double2 Ts[NB][NB + 1];
  const int tid = threadIdx.x;
  const int base = (right ? k0 + kb : k0 + kb) + blockIdx.x * NB;  // first row (right) / column (left)
  for (int e = tid; e < NB * NB; e += 256) {
    const int a = e / NB, b = e % NB;
    Ts[a][b] = Tm[a * NB + b];
    if (right) {
      // e = k * NB + i: column k0 + k, row base + i (coalesced along rows)
      const int k = e / NB, i = e % NB;
      Xs[i][k] = (base + i < lim && k < kb) ? M[(size_t)(k0 + k) * lda + base + i] : make_double2(0.0, 0.0);
    } else {
      const int c = e / NB, k = e % NB;
      Xs[k][c] = (base + c < lim && k < kb) ? M[(size_t)(base + c) * lda + k0 + k] : make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  // thread: 4 x 4 outputs, rows ti + 16 u, columns tj + 16 v of the 64 x 64 result
  const int ti = tid & 15, tj = tid >> 4;
  double2 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = make_double2(0.0, 0.0);
  for (int k = 0; k < NB; ++k) {
    double2 a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = right ? Xs[ti + 16 * u][k] : Ts[ti + 16 * u][k];
#pragma unroll
    for (int v = 0; v < 4; ++v) b[v] = right ? Ts[k][tj + 16 * v] : Xs[k][tj + 16 * v];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        acc[u][v].x = fma(a[u].x, b[v].x, fma(-a[u].y, b[v].y, acc[u][v].x));
        acc[u][v].y = fma(a[u].x, b[v].y, fma(a[u].y, b[v].x, acc[u][v].y));
      }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = ti + 16 * u, c = tj + 16 * v;
      if (right) {
        if (base + r < lim && c < kb) M[(size_t)(k0 + c) * lda + base + r] = acc[u][v];
      } else {
        if (base + c < lim && r < kb) M[(size_t)(base + c) * lda + k0 + r] = acc[u][v];
      }
    }
}

// C -= A B on the trailing matrix: rows [r0, n), columns [c0, ncols), K = kb (<= NB); A is
// the block column L21 (column k0..), B the block row U12 (rows k0..). 64 x 64 tiles of C per
// workgroup of 256 threads, 4 x 4 complex outputs per thread, operands staged in LDS.
__global__ void __launch_bounds__(256) gemm_update_kernel(double2* __restrict__ M, int n, int lda, int ncols, int k0,
                                                          int kb) {
  __shared__ double2 As[16][64 + 1];
  __shared__ double2 Bs[16][64 + 1];
  const int r0 = k0 + kb, c0 = k0 + kb;
  const int tr = blockIdx.x * 64, tc = blockIdx.y * 64;
  const int tid = threadIdx.x;
  const int ti = tid & 15, tj = tid >> 4;  // thread's 4 x 4 block: rows ti + 16 u, columns tj + 16 v
  double2 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = make_double2(0.0, 0.0);
  for (int kc = 0; kc < kb; kc += 16) {
    // A tile: 64 rows x 16 k; B tile: 16 k x 64 columns
    for (int e = tid; e < 64 * 16; e += 256) {
      const int i = e & 63, kk = e >> 6;
      const int row = r0 + tr + i, k = kc + kk;
      As[kk][i] = (row < n && k < kb) ? M[(size_t)(k0 + k) * lda + row] : make_double2(0.0, 0.0);
      const int jj = e >> 4, kk2 = e & 15;
      const int col = c0 + tc + jj, k2 = kc + kk2;
      Bs[kk2][jj] = (col < ncols && k2 < kb) ? M[(size_t)col * lda + k0 + k2] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double2 a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = As[kk][ti + 16 * u];
#pragma unroll
      for (int v = 0; v < 4; ++v) b[v] = Bs[kk][tj + 16 * v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          acc[u][v].x = fma(a[u].x, b[v].x, fma(-a[u].y, b[v].y, acc[u][v].x));
          acc[u][v].y = fma(a[u].x, b[v].y, fma(a[u].y, b[v].x, acc[u][v].y));
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int col = c0 + tc + tj + 16 * v;
    if (col >= ncols) continue;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int row = r0 + tr + ti + 16 * u;
      if (row >= n) continue;
      double2* p = M + (size_t)col * lda + row;
      double2 c = *p;
      c.x -= acc[u][v].x;
      c.y -= acc[u][v].y;
      *p = c;
    }
  }
}

// Back substitution U X = Y on the augmented columns [n, n + r): the diagonal block (k0, kb)
// solved by one workgroup per right-hand side, then the rows above updated.
__global__ void __launch_bounds__(NB) backsolve_block_kernel(double2* __restrict__ M, int n, int lda, int k0, int kb) {
  __shared__ double2 y[NB];
  const int rhs = blockIdx.x;
  const int t = threadIdx.x;
  double2* b = M + (size_t)(n + rhs) * lda;
  if (t < kb) y[t] = b[k0 + t];
  __syncthreads();
  for (int r = kb - 1; r >= 0; --r) {
    if (t == 0) y[r] = cdiv(y[r], M[(size_t)(k0 + r) * lda + k0 + r]);
    __syncthreads();
    if (t < r) {
      const double2 u = M[(size_t)(k0 + r) * lda + k0 + t];
      const double2 v = y[r];
      y[t].x -= u.x * v.x - u.y * v.y;
      y[t].y -= u.x * v.y + u.y * v.x;
    }
    __syncthreads();
  }
  if (t < kb) b[k0 + t] = y[t];
}

__global__ void __launch_bounds__(256) backsolve_update_kernel(double2* __restrict__ M, int n, int lda, int k0, int kb,
                                                               int nrhs) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= k0) return;
  for (int r = 0; r < nrhs; ++r) {
    double2* b = M + (size_t)(n + r) * lda;
    double2 s = b[row];
    for (int c = 0; c < kb; ++c) {
      const double2 u = M[(size_t)(k0 + c) * lda + row];
      const double2 x = b[k0 + c];
      s.x -= u.x * x.x - u.y * x.y;
      s.y -= u.x * x.y + u.y * x.x;
    }
    b[row] = s;
  }
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
struct LuWorkspace {
  int n = 0;
  TsluWork w{};
  int* ipiv = nullptr;
  double2* uinv = nullptr;
  double2* linv = nullptr;
  void release() {
    hipFree(w.cand);
    hipFree(w.ccount);
    hipFree(w.ctr);
    hipFree(w.top);
    hipFree(w.winners);
    hipFree(w.zero_pivot);
    hipFree(ipiv);
    hipFree(uinv);
    hipFree(linv);
    *this = LuWorkspace{};
  }
};

static void lu_workspace(LuWorkspace& ws, int n) {
  if (ws.n >= n && ws.ipiv) return;
  ws.release();
  int leaves = (n + LEAF - 1) / LEAF;
  int levels = 1, nodes = leaves;
  while (nodes > 1) {
    nodes = (nodes + FAN - 1) / FAN;
    ++levels;
  }
  ws.w.levels = levels;
  ws.w.max_nodes = leaves;
  const size_t slots = (size_t)levels * leaves;
  HBG_HIP(hipMalloc(&ws.w.cand, slots * NB * sizeof(int)));
  HBG_HIP(hipMalloc(&ws.w.ccount, slots * sizeof(int)));
  HBG_HIP(hipMalloc(&ws.w.ctr, slots * sizeof(unsigned)));
  HBG_HIP(hipMalloc(&ws.w.top, (size_t)NB * NB * sizeof(double2)));
  HBG_HIP(hipMalloc(&ws.w.winners, NB * sizeof(int)));
  HBG_HIP(hipMalloc(&ws.w.zero_pivot, sizeof(int)));
  HBG_HIP(hipMalloc(&ws.ipiv, (size_t)n * sizeof(int)));
  HBG_HIP(hipMalloc(&ws.uinv, (size_t)NB * NB * sizeof(double2)));
  HBG_HIP(hipMalloc(&ws.linv, (size_t)NB * NB * sizeof(double2)));
  ws.n = n;
}

static LuWorkspace& workspace() {
  static thread_local LuWorkspace ws;
  return ws;
}

// Timing of the LU's kernels by HIP events on its stream (when requested): the trailing-update
// GEMMs (their algorithmic flops 8 m n k each) and the whole factorization and solve.
struct LuStats {
  double total_ms = 0.0, gemm_ms = 0.0, gemm_flops = 0.0;
  long gemm_launches = 0;
};

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
}  // namespace

// Factor [A | B] (n x (n + nrhs), leading dimension lda) in place and solve; returns the first
// zero pivot (1-based) or 0. Everything on `stream`; the call waits for the result.
int lu_solve_device(double2* M, int n, int lda, int nrhs, hipStream_t stream, LuStats* stats) {
  EventPool& ep = events();
  ep.used = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> gemm_ev;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (stats) {
    t0 = ep.next();
    HBG_HIP(hipEventRecord(t0, stream));
  }
  LuWorkspace& ws = workspace();
  lu_workspace(ws, n);
  const int ncols = n + nrhs;
  const size_t slots = (size_t)ws.w.levels * ws.w.max_nodes;
  fill_i32<<<1, 64, 0, stream>>>(ws.w.zero_pivot, 1, 0x7fffffff);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = min(NB, n - k0);
    const int m = n - k0;
    fill_u32<<<(unsigned)((slots + 255) / 256), 256, 0, stream>>>(ws.w.ctr, slots, 0u);
    tslu_kernel<<<(m + LEAF - 1) / LEAF, LEAF, 0, stream>>>(M, n, lda, k0, kb, ws.w);
    ipiv_kernel<<<1, 1, 0, stream>>>(k0, kb, ws.w.winners, ws.ipiv);
    swap_kernel<<<(ncols + 255) / 256, 256, 0, stream>>>(M, lda, ncols, k0, kb, ws.ipiv, ws.w.top);
    trinv_kernel<<<1, 2 * NB, 0, stream>>>(M, lda, k0, kb, ws.uinv, ws.linv);
    if (m > kb) tri_apply_kernel<<<(m - kb + NB - 1) / NB, 256, 0, stream>>>(M, lda, n, k0, kb, ws.uinv, 1);
    if (ncols > k0 + kb) {
      tri_apply_kernel<<<(ncols - k0 - kb + NB - 1) / NB, 256, 0, stream>>>(M, lda, ncols, k0, kb, ws.linv, 0);
      if (m > kb) {
        dim3 grid((m - kb + 63) / 64, (ncols - k0 - kb + 63) / 64);
        hipEvent_t g0 = nullptr, g1 = nullptr;
        if (stats) {
          g0 = ep.next();
          HBG_HIP(hipEventRecord(g0, stream));
        }
        gemm_update_kernel<<<grid, 256, 0, stream>>>(M, n, lda, ncols, k0, kb);
        if (stats) {
          g1 = ep.next();
          HBG_HIP(hipEventRecord(g1, stream));
          gemm_ev.push_back({g0, g1});
          stats->gemm_flops += 8.0 * (double)(m - kb) * (double)(ncols - k0 - kb) * (double)kb;
          stats->gemm_launches += 1;
        }
      }
    }
  }
  HBG_HIP(hipGetLastError());
  for (int k0 = ((n - 1) / NB) * NB; k0 >= 0 && nrhs > 0; k0 -= NB) {
    const int kb = min(NB, n - k0);
    backsolve_block_kernel<<<nrhs, NB, 0, stream>>>(M, n, lda, k0, kb);
    if (k0 > 0) backsolve_update_kernel<<<(k0 + 255) / 256, 256, 0, stream>>>(M, n, lda, k0, kb, nrhs);
  }
  HBG_HIP(hipGetLastError());
  if (stats) {
    t1 = ep.next();
    HBG_HIP(hipEventRecord(t1, stream));
  }
  int zp = 0;
  HBG_HIP(hipMemcpyAsync(&zp, ws.w.zero_pivot, sizeof(int), hipMemcpyDeviceToHost, stream));
  HBG_HIP(hipStreamSynchronize(stream));
  if (stats) {
    float ms = 0.0f;
    HBG_HIP(hipEventElapsedTime(&ms, t0, t1));
    stats->total_ms += ms;
    for (auto& g : gemm_ev) {
      HBG_HIP(hipEventElapsedTime(&ms, g.first, g.second));
      stats->gemm_ms += ms;
    }
  }
  return zp == 0x7fffffff ? 0 : zp;
}

void lu_release() { workspace().release(); }

}  // namespace hbg
