// The learned surrogate of the finite-depth wave kernel (no counterpart in the reference; it
// stands beside the series/tile evaluator of green_fin.rs:1007-1095, which stays the fallback and
// the judge of its error).
//
// What is learned. With h the depth, the regular part of the real kernel splits as
//   Wr(R, z, zeta) = T(R, z, zeta) - a(z, zeta) Y0(k0 R),   a = A0 E (the propagating mode),
// and Wi = -a J0(k0 R) exactly. The propagating mode oscillates (k0 R reaches ~400 over the
// domain), so it stays analytic; T, the evanescent remainder (the series over the 24 evanescent
// roots or the Phi integrals, the expensive part), is smooth in log(R / h). The network maps
//   u = (ln(R/h) scaled to [-1, 1] over [1e-2, 20], 2 z/h + 1, 2 zeta/h + 1,
//        ln(nu h) scaled to [-1, 1] over [0.05, 20])
// to y = (h T, h^2 T_R, h^2 T_zeta, h^2 T_z) (zeta the source depth, z the field depth), each divided
// by the scale of nn_scales in surrogate.h and standardized: 4 -> 128 -> 128 -> 128 -> 4, SiLU.
//
// Training (surrogate_train): samples drawn on the device (R/h log-uniform, z/h and zeta/h
// uniform in [-1, 0], nu h log-uniform, one nu h per group of samples), targets from the FP64
// evaluator of green_fin_dev.h, full-batch-free Adam on mean squared error with FP32 master
// weights; every product of the forward and backward passes is an FP32 GEMM on the matrix cores
// (v_mfma_f32_16x16x4f32). Inference (nn_wave) runs one wavefront over 64 samples on BF16 matrix
// cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), BF16 weights staged in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/hydrobem_amd.h"
#include "green_fin_dev.h"
#include "hbg_internal.h"
#include "surrogate.h"

namespace hbg {

size_t fin_tables_work();
void launch_fin_tables(const FinFreq& f, WaveTables T, const double* g16, double* phi1, double* phi2,
                       const double* M1dev, const PanelD* P, int N, double4* pf, double2* pcs, double* work,
                       hipStream_t stream);
FinFreq make_fin_freq(double h, double K);

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ sampling and targets
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double uniform01(unsigned long long seed, unsigned long long i, int d) {
  return (double)(mix64(seed ^ mix64(i * 8 + d)) >> 11) * (1.0 / 9007199254740992.0);
}

// sample i of a group: R/h, z/h, zeta/h (h = 1); the depths as panels i (z) and n + i (zeta)
__global__ void sample_kernel(unsigned long long seed, unsigned long long first, int n, double* __restrict__ R,
                              double* __restrict__ z, double* __restrict__ zeta, PanelD* __restrict__ pts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long id = first + i;
  const double lr = NN_LNR_LO + (NN_LNR_HI - NN_LNR_LO) * uniform01(seed, id, 0);
  R[i] = exp(lr);
  z[i] = -uniform01(seed, id, 1);
  zeta[i] = -uniform01(seed, id, 2);
  pts[i].c[2] = z[i];
  pts[n + i].c[2] = zeta[i];
}

// the network's input and target of each sample (h = 1, K = nu h)
__global__ void target_kernel(const FinFreq* __restrict__ f, FinTables F, WaveTables T, int n,
                              const double* __restrict__ R, const double* __restrict__ z,
                              const double* __restrict__ zeta, float4* __restrict__ X, float4* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = R[i];
  const FinK w = fin_kernel(f, F, T, i, n + i, r, z[i], zeta[i]);
  const FinProp p = fin_prop(f, F, T, i, n + i, r);
  const double k0 = f->k0, nuh = f->K * f->h;
  double s0, s1;
  nn_scales(r, nuh, s0, s1);
  X[i] = nn_input(r, z[i], zeta[i], nuh);
  Y[i] = make_float4((float)((w.Wr + p.a * p.y0) / s0), (float)((w.WrR - p.a * k0 * p.y1) / s1),
                     (float)((w.WrZ + p.az * p.y0) / s1), (float)((w.Wrz + p.af * p.y0) / s1));
}

// ------------------------------------------------------------------ training kernels
// C[m][n] (+)= sum_k A(m, k) B(k, n) (+ bias[n]) with element (r, c) of A at A[r ars + c acs] etc.;
// FP32 on the matrix cores, 64 x 64 tiles per workgroup of 4 waves (2 x 2 blocks of 16 x 16 each),
// operands read straight from global memory (the matrices are small and stay in cache).
__global__ void __launch_bounds__(256) mlp_gemm_kernel(int M, int N, int K, const float* __restrict__ A, int ars,
                                                       int acs, const float* __restrict__ B, int brs, int bcs,
                                                       float* __restrict__ C, int crs, int ccs,
                                                       const float* __restrict__ bias, int accumulate) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = blockIdx.x * 64 + (wave & 1) * 32, c0 = blockIdx.y * 64 + (wave >> 1) * 32;
  v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + (lane >> 4);
    float av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int row = r0 + a * 16 + (lane & 15);
      av[a] = (row < M && k < K) ? A[(size_t)row * ars + (size_t)k * acs] : 0.f;
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = c0 + b * 16 + (lane & 15);
      bv[b] = (col < N && k < K) ? B[(size_t)k * brs + (size_t)col * bcs] : 0.f;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = r0 + a * 16 + 4 * (lane >> 4) + r, col = c0 + b * 16 + (lane & 15);
        if (row < M && col < N) {
          float v = acc[a][b][r] + (bias ? bias[col] : 0.f);
          float* c = C + (size_t)row * crs + (size_t)col * ccs;
          *c = accumulate ? *c + v : v;
        }
      }
}

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// A = silu(Z)
__global__ void silu_kernel(const float* __restrict__ Z, float* __restrict__ A, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i] = Z[i] * sigm(Z[i]);
}

// dZ = dA silu'(Z), in place on dA
__global__ void silu_back_kernel(const float* __restrict__ Z, float* __restrict__ dA, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float s = sigm(Z[i]);
    dA[i] *= s * (1.0f + Z[i] * (1.0f - s));
  }
}

// dY = 2 (Y - T) / (B 4) and the loss sum (FP64 atomics per workgroup)
__global__ void loss_kernel(const float* __restrict__ Yp, const float4* __restrict__ T, int B, float* __restrict__ dY,
                            double* __restrict__ loss) {
  __shared__ double red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < B) {
    const float4 t = T[i];
    const float tt[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int q = 0; q < NN_OUT; ++q) {
      const float d = Yp[(size_t)i * NN_OUT + q] - tt[q];
      dY[(size_t)i * NN_OUT + q] = 2.0f * d / (float)(B * NN_OUT);
      l += (double)d * d;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    atomicAdd(loss, s / (double)(B * NN_OUT));
  }
}

// db[o] = sum_b dZ[b][o]: one workgroup per output, a strided sum and a tree
__global__ void colsum_kernel(const float* __restrict__ dZ, int B, int O, float* __restrict__ db) {
  __shared__ float red[256];
  const int o = blockIdx.x;
  float s = 0.f;
  for (int b = threadIdx.x; b < B; b += 256) s += dZ[(size_t)b * O + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[o] = red[0];
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

// Y = (Y - mu) * inv
__global__ void standardize_kernel(float4* __restrict__ Y, int n, float4 mu, float4 inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 y = Y[i];
  y.x = (y.x - mu.x) * inv.x;
  y.y = (y.y - mu.y) * inv.y;
  y.z = (y.z - mu.z) * inv.z;
  y.w = (y.w - mu.w) * inv.w;
  Y[i] = y;
}

// the batch [first, first + B) of a permuted data set into contiguous buffers
__global__ void gather_kernel(const float4* __restrict__ X, const float4* __restrict__ Y, const int* __restrict__ perm,
                              int first, int B, float* __restrict__ Xb, float4* __restrict__ Yb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const int j = perm[first + i];
  const float4 x = X[j];
  Xb[(size_t)i * 4 + 0] = x.x;
  Xb[(size_t)i * 4 + 1] = x.y;
  Xb[(size_t)i * 4 + 2] = x.z;
  Xb[(size_t)i * 4 + 3] = x.w;
  Yb[i] = Y[j];
}

// ------------------------------------------------------------------ inference
// Packed BF16 weights of nn_wave, in units of v8bf (16 B), lane-major:
//   L1  [ob 8][lane 64]            W1[16 ob + l%16][q] for lanes l < 16, q < 4, else 0
//   L2  [ob 8][s 4][lane 64]       W2[16 ob + l%16][f(s, l / 16, q)]
//   L3  as L2
//   L4  [s 4][lane 64]             W4[l%16][f(s, l / 16, q)] for l%16 < 4, else 0
// with the input order of a K-step f(s, g, q) = 32 s + 4 g + q (q < 4), 32 s + 16 + 4 g + q - 4
// (q >= 4): the features a lane holds after the previous layer. Biases FP32 after them.
constexpr int NN_PK_L1 = 8 * 64, NN_PK_L2 = 8 * 4 * 64, NN_PK_L4 = 4 * 64;
constexpr int NN_PK_V8 = NN_PK_L1 + 2 * NN_PK_L2 + NN_PK_L4;  // v8bf units
constexpr int NN_PK_BIAS = 3 * NN_H + 16;                    // floats

// SiLU with the hardware exp2 and reciprocal (about 1 ulp each)
__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

// The network on the 64 samples of a wavefront (u: this lane's sample); returns this lane's four
// outputs (normalized). All 64 lanes must call it. pk: the packed weights in LDS.
__device__ __forceinline__ float4 nn_wave(const v8bf* __restrict__ pk, const float* __restrict__ bias, float4 u) {
  const int lane = threadIdx.x & 63, g = lane >> 4;
  // inputs of sample 16 sb + lane % 16 to the lanes of group 0 (K positions 0..3)
  v8bf xin[4];
#pragma unroll
  for (int sb = 0; sb < 4; ++sb) {
    const int src = 16 * sb + (lane & 15);
    const float a0 = __shfl(u.x, src), a1 = __shfl(u.y, src), a2 = __shfl(u.z, src), a3 = __shfl(u.w, src);
    v8bf v;
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (__bf16)0.0f;
    if (g == 0) {
      v[0] = (__bf16)a0;
      v[1] = (__bf16)a1;
      v[2] = (__bf16)a2;
      v[3] = (__bf16)a3;
    }
    xin[sb] = v;
  }
  // h[ob][sb]: features 16 ob + 4 g + r (r < 4) of sample 16 sb + lane % 16, as BF16 (4 used of 8)
  __bf16 h[8][4][4];
  // layer 1
#pragma unroll
  for (int ob = 0; ob < 8; ++ob) {
    const v8bf w = pk[ob * 64 + lane];
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xin[sb], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) h[ob][sb][r] = (__bf16)silu(acc[r] + bias[16 * ob + 4 * g + r]);
    }
  }
  // layers 2 and 3
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const v8bf* pw = pk + NN_PK_L1 + layer * NN_PK_L2;
    const float* pb = bias + NN_H * (1 + layer);
    v8bf in[4][4];  // [s][sb]
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int sb = 0; sb < 4; ++sb) {
        v8bf v;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          v[q] = h[2 * s][sb][q];
          v[4 + q] = h[2 * s + 1][sb][q];
        }
        in[s][sb] = v;
      }
#pragma unroll
    for (int ob = 0; ob < 8; ++ob) {
      v4f acc[4];
#pragma unroll
      for (int sb = 0; sb < 4; ++sb) acc[sb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const v8bf w = pw[(ob * 4 + s) * 64 + lane];
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) acc[sb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, in[s][sb], acc[sb], 0, 0, 0);
      }
#pragma unroll
      for (int sb = 0; sb < 4; ++sb)
#pragma unroll
        for (int r = 0; r < 4; ++r) h[ob][sb][r] = (__bf16)silu(acc[sb][r] + pb[16 * ob + 4 * g + r]);
    }
  }
  // layer 4: outputs 4 g + r of sample 16 sb + lane % 16 (group 0 holds the four outputs)
  const v8bf* pw4 = pk + NN_PK_L1 + 2 * NN_PK_L2;
  v4f out[4];
#pragma unroll
  for (int sb = 0; sb < 4; ++sb) out[sb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const v8bf w = pw4[s * 64 + lane];
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
      v8bf v;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[q] = h[2 * s][sb][q];
        v[4 + q] = h[2 * s + 1][sb][q];
      }
      out[sb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, v, out[sb], 0, 0, 0);
    }
  }
  // back to the owner lane 16 sb + j of sample (sb, j): from lane j, register sb
  const float* pb4 = bias + 3 * NN_H;
  float res[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float mine = 0.f;
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
      const float v = __shfl(out[sb][r], lane & 15);
      if (g == sb) mine = v;
    }
    res[r] = mine + pb4[r];
  }
  return make_float4(res[0], res[1], res[2], res[3]);
}

// stage the packed weights and biases of the network into LDS (all threads of the workgroup)
__device__ __forceinline__ void nn_stage(const v8bf* __restrict__ gpk, const float* __restrict__ gbias, v8bf* pk,
                                         float* bias) {
  for (int i = threadIdx.x; i < NN_PK_V8; i += blockDim.x) pk[i] = gpk[i];
  for (int i = threadIdx.x; i < NN_PK_BIAS; i += blockDim.x) bias[i] = gbias[i];
}

// outputs of n samples (normalized inputs X), for held-out errors and the tests
__global__ void __launch_bounds__(256) nn_eval_kernel(const v8bf* __restrict__ gpk, const float* __restrict__ gbias,
                                                      const float4* __restrict__ X, int n, float4* __restrict__ Y) {
  __shared__ v8bf pk[NN_PK_V8];
  __shared__ float bias[NN_PK_BIAS];
  nn_stage(gpk, gbias, pk, bias);
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  const float4 u = i < n ? X[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 y = nn_wave(pk, bias, u);
  if (i < n) Y[i] = y;
}

// The finite-depth S and A of one frequency (as assemble_fin_kernel in frequency.hip) with the
// evanescent remainder from the network where rmin <= R/h <= 20 (rmin >= NN_RMIN, the training
// domain's lower end, 0.1 by default); the other pairs are queued
// (wave-aggregated appends) for fallback_kernel, which evaluates them in FP64. As in frequency.hip
// each unordered pair is taken once (64 x 64 tiles with k-block <= n-block, the diagonal tiles
// the pairs k <= n) and both entries are written: the network's third output is the remainder's
// derivative in the source depth, its fourth the one in the field depth. A wave takes 64 field
// points (k) against NN_PER_WAVE sources (n); the packed weights are staged once per workgroup.
constexpr int NN_PER_WAVE = 16;
__device__ __forceinline__ void write_entry(double2* __restrict__ S, double2* __restrict__ A,
                                            float2* __restrict__ A32, int lda, int row, int col, const StaticD& e,
                                            double Ar, double nz, double a, double c2, double Wr, double WrR,
                                            double WrZ, double Wi, double WiR, double WiZ, double free_term) {
  S[(size_t)col * lda + row] = make_double2(e.rs + Ar * Wr - c2 * e.dl, Ar * Wi);
  const double dre = e.rd + Ar * (-WrR * a + WrZ * nz) - c2 * e.dd, dim = Ar * (-WiR * a + WiZ * nz);
  const double2 v = make_double2((row == col ? free_term : 0.0) - dre, -dim);
  A[(size_t)col * lda + row] = v;
  if (A32) A32[(size_t)col * lda + row] = float2((float)v.x, (float)v.y);
}

// both entries of the pair (k, n), k <= n, from the pair's kernel values: field point k, source
// Psrc + n, and field point n, source Psrc + k (the panels themselves, or their images in one
// reflection g: |c_k - g c_n| = |c_n - g c_k|, and the mirrored entry's normal term is that of
// panel k, g being orthogonal), st, S, A those of the block, free_term on its diagonal
__device__ __forceinline__ void write_pair(const PanelD* __restrict__ P, const PanelD* __restrict__ Psrc, int N,
                                           const StaticD* __restrict__ st, const FinK& w, double c2, int k, int n,
                                           int lda, double2* __restrict__ S, double2* __restrict__ A,
                                           float2* __restrict__ A32, double free_term) {
  const PanelD* pk = P + k;
  const PanelD* pn = Psrc + n;
  const double dx = pk->c[0] - pn->c[0], dy = pk->c[1] - pn->c[1];
  const double R = sqrt(dx * dx + dy * dy);
  const double an = R > 0.0 ? (dx * pn->n[0] + dy * pn->n[1]) / R : 0.0;
  write_entry(S, A, A32, lda, k, n, st[(size_t)n * N + k], pn->area, pn->n[2], an, c2, w.Wr, w.WrR, w.WrZ, w.Wi, w.WiR,
              w.WiZ, free_term);
  if (k != n) {
    const double ak = R > 0.0 ? -(dx * pk->n[0] + dy * pk->n[1]) / R : 0.0;
    write_entry(S, A, A32, lda, n, k, st[(size_t)k * N + n], pk->area, pk->n[2], ak, c2, w.Wr, w.WrR, w.Wrz, w.Wi, w.WiR,
                w.Wiz, free_term);
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
assemble_fin_nn_kernel(const PanelD* __restrict__ P, int N, const FinFreq* __restrict__ f, FinTables F, WaveTables T,
                       double logc, const StaticD* __restrict__ st, double2* __restrict__ S, double2* __restrict__ A,
                       float2* __restrict__ A32, int lda, const v8bf* __restrict__ gpk, const float* __restrict__ gbias,
                       int2* __restrict__ queue, unsigned* __restrict__ qcount, double rmin) {
  const int bi = blockIdx.x, bj = blockIdx.y, g = blockIdx.z;
  if (bi > bj) return;  // workgroup-uniform
  // image block g of a symmetric body or walls: sources g n, no free term (A32 is then null)
  const PanelD* Psrc = P + (size_t)g * N;
  st += (size_t)g * N * N;
  S += (size_t)g * N * N;
  A += (size_t)g * N * lda;
  const double free_term = g ? 0.0 : 2.0 * PI;
  __shared__ v8bf pk[NN_PK_V8];
  __shared__ float bias[NN_PK_BIAS];
  nn_stage(gpk, gbias, pk, bias);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k = bi * 64 + lane;
  const int kk = k < N ? k : N - 1;
  const double h = f->h, K = f->K, k0 = f->k0, nuh = K * h, ih = 1.0 / h;
  const double xk = P[kk].c[0], yk = P[kk].c[1], zk = P[kk].c[2];
  const double c2 = 2.0 * K * logc;
  for (int t = 0; t < NN_PER_WAVE; ++t) {
    const int n = bj * 64 + wave * NN_PER_WAVE + t;
    if (n >= N) break;  // wave-uniform
    const bool live = k < N && !(bi == bj && k > n);
    const PanelD* pn = Psrc + n;
    const double dx = xk - pn->c[0], dy = yk - pn->c[1];
    const double R = sqrt(dx * dx + dy * dy), r = R * ih;
    const bool nn = r >= rmin && r <= 20.0;
    const float4 u = nn_input(nn ? r : 1.0, zk * ih, pn->c[2] * ih, nuh);
    const float4 y = nn_wave(pk, bias, u);  // every lane takes part
    const unsigned long long fb = __ballot(live && !nn);
    if (fb) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(qcount, (unsigned)__popcll(fb));
      base = __shfl(base, 0);
      if (live && !nn) queue[base + (unsigned)__popcll(fb & ((1ull << lane) - 1ull))] = make_int2(k, g * N + n);
    }
    if (!live || !nn) continue;
    const FinProp pp = fin_prop(f, F, T, kk, n, R);
    double s0, s1;
    nn_scales(r, nuh, s0, s1);
    const double Tv = y.x * s0 * ih, sd = s1 * ih * ih;
    FinK w;
    w.Wr = Tv - pp.a * pp.y0;
    w.WrR = y.y * sd + pp.a * k0 * pp.y1;
    w.WrZ = y.z * sd - pp.az * pp.y0;
    w.Wrz = y.w * sd - pp.af * pp.y0;
    w.Wi = -pp.a * pp.j0;
    w.WiR = pp.a * k0 * pp.j1;
    w.WiZ = -pp.az * pp.j0;
    w.Wiz = -pp.af * pp.j0;
    write_pair(P, Psrc, N, st, w, c2, k, n, lda, S, A, A32, free_term);
  }
}

// the queued pairs by the FP64 evaluator (grid-stride over the count the first pass left); a
// queued pair (k, g N + n) is the pair (k, n) of the image block g
__global__ void __launch_bounds__(256) fallback_kernel(const PanelD* __restrict__ P, int N,
                                                       const FinFreq* __restrict__ f, FinTables F, WaveTables T,
                                                       double logc, const StaticD* __restrict__ st,
                                                       double2* __restrict__ S, double2* __restrict__ A,
                                                       float2* __restrict__ A32, int lda,
                                                       const int2* __restrict__ queue,
                                                       const unsigned* __restrict__ qcount) {
  const unsigned count = *qcount;
  const double c2 = 2.0 * f->K * logc;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const int2 q = queue[i];
    const int k = q.x, g = q.y / N, n = q.y % N;
    const PanelD* Psrc = P + (size_t)g * N;
    const double dx = P[k].c[0] - Psrc[n].c[0], dy = P[k].c[1] - Psrc[n].c[1];
    const FinK w = fin_kernel(f, F, T, k, n, sqrt(dx * dx + dy * dy), P[k].c[2], Psrc[n].c[2]);
    write_pair(P, Psrc, N, st + (size_t)g * N * N, w, c2, k, n, lda, S + (size_t)g * N * N, A + (size_t)g * N * lda,
               A32, g ? 0.0 : 2.0 * PI);
  }
}

__global__ void zero_u32_kernel(unsigned* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

// FP64 evanescent remainder (h T, h^2 T_R, h^2 T_zeta, h^2 T_z) of points, for the tests and
// the held-out judge: the evaluator of green_fin_dev.h minus the propagating mode
__global__ void remainder_kernel(const FinFreq* __restrict__ f, FinTables F, WaveTables T, int n,
                                 const double* __restrict__ R, const double* __restrict__ z,
                                 const double* __restrict__ zeta, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = R[i], h = f->h, k0 = f->k0;
  const FinK w = fin_kernel(f, F, T, i, n + i, r, z[i], zeta[i]);
  const FinProp p = fin_prop(f, F, T, i, n + i, r);
  double* o = out + 4 * (size_t)i;
  o[0] = (w.Wr + p.a * p.y0) * h;
  o[1] = (w.WrR - p.a * k0 * p.y1) * h * h;
  o[2] = (w.WrZ + p.az * p.y0) * h * h;
  o[3] = (w.Wrz + p.af * p.y0) * h * h;
}

}  // namespace

// ------------------------------------------------------------------ host side
namespace {

// f(s, g, q): the input feature at position 8 g + q of K-step s
int kperm(int s, int g, int q) { return q < 4 ? 32 * s + 4 * g + q : 32 * s + 16 + 4 * g + (q - 4); }

__bf16 to_bf16(float x) { return (__bf16)x; }

}  // namespace

// Packed inference weights (NN_PK_V8 x 8 BF16, then NN_PK_BIAS floats) from the FP32 parameters
// (NnParams layout: W1 [128][4], b1, W2 [128][128], b2, W3, b3, W4 [4][128], b4).
void nn_pack(const float* prm, std::vector<__bf16>& pk, std::vector<float>& bias) {
  const float* W1 = prm + NN_OFF_W1;
  const float* W2 = prm + NN_OFF_W2;
  const float* W3 = prm + NN_OFF_W3;
  const float* W4 = prm + NN_OFF_W4;
  pk.assign((size_t)NN_PK_V8 * 8, to_bf16(0.f));
  for (int ob = 0; ob < 8; ++ob)
    for (int l = 0; l < 64; ++l)
      for (int q = 0; q < 8; ++q)
        if (l < 16 && q < 4) pk[((size_t)ob * 64 + l) * 8 + q] = to_bf16(W1[(16 * ob + l) * NN_IN + q]);
  for (int layer = 0; layer < 2; ++layer) {
    const float* W = layer == 0 ? W2 : W3;
    const size_t base = NN_PK_L1 + (size_t)layer * NN_PK_L2;
    for (int ob = 0; ob < 8; ++ob)
      for (int s = 0; s < 4; ++s)
        for (int l = 0; l < 64; ++l)
          for (int q = 0; q < 8; ++q)
            pk[((base + (ob * 4 + s) * 64) + l) * 8 + q] =
                to_bf16(W[(16 * ob + (l & 15)) * NN_H + kperm(s, l >> 4, q)]);
  }
  const size_t b4 = NN_PK_L1 + 2 * (size_t)NN_PK_L2;
  for (int s = 0; s < 4; ++s)
    for (int l = 0; l < 64; ++l)
      for (int q = 0; q < 8; ++q)
        if ((l & 15) < NN_OUT) pk[((b4 + s * 64) + l) * 8 + q] = to_bf16(W4[(l & 15) * NN_H + kperm(s, l >> 4, q)]);
  bias.assign(NN_PK_BIAS, 0.f);
  for (int i = 0; i < NN_H; ++i) {
    bias[i] = prm[NN_OFF_B1 + i];
    bias[NN_H + i] = prm[NN_OFF_B2 + i];
    bias[2 * NN_H + i] = prm[NN_OFF_B3 + i];
  }
  for (int i = 0; i < NN_OUT; ++i) bias[3 * NN_H + i] = prm[NN_OFF_B4 + i];
}

NnDevice::~NnDevice() { release(); }

void NnDevice::release() {
  (void)hipFree(pk);
  (void)hipFree(bias);
  pk = nullptr;
  bias = nullptr;
  ready = false;
}

void NnDevice::upload(const float* prm) {
  std::vector<__bf16> hpk;
  std::vector<float> hb;
  nn_pack(prm, hpk, hb);
  if (!pk) HBG_HIP(hipMalloc(&pk, hpk.size() * sizeof(__bf16)));
  if (!bias) HBG_HIP(hipMalloc(&bias, hb.size() * sizeof(float)));
  HBG_HIP(hipMemcpy(pk, hpk.data(), hpk.size() * sizeof(__bf16), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(bias, hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice));
  HBG_HIP(hipDeviceSynchronize());  // the DMA of a pageable copy may outlive hipMemcpy's return
  params.assign(prm, prm + NN_PARAMS);
  ready = true;
}

void nn_eval(const NnDevice& nn, const float4* X, int n, float4* Y, hipStream_t stream) {
  if (n <= 0) return;
  nn_eval_kernel<<<(n + 255) / 256, 256, 0, stream>>>((const v8bf*)nn.pk, nn.bias, X, n, Y);
  HBG_HIP(hipGetLastError());
}

namespace {

template <typename T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t m) {
    if (m <= n) return;
    (void)hipFree(p);
    p = nullptr;
    n = 0;
    HBG_HIP(hipMalloc(&p, std::max<size_t>(m, 1) * sizeof(T)));
    n = m;
  }
  ~Buf() { (void)hipFree(p); }
};

void gemm(int M, int N, int K, const float* A, int ars, int acs, const float* B, int brs, int bcs, float* C, int crs,
          int ccs, const float* bias, int accumulate, hipStream_t st) {
  mlp_gemm_kernel<<<dim3((M + 63) / 64, (N + 63) / 64), 256, 0, st>>>(M, N, K, A, ars, acs, B, brs, bcs, C, crs, ccs,
                                                                      bias, accumulate);
}

unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

// forward of a batch (X [B][4]) with the FP32 parameters P: pre-activations Z1..Z3, activations
// A1..A3 ([B][128]), outputs Yp [B][4]
struct Acts {
  Buf<float> Z[3], A[3], Yp, dA, dB;
};

void forward(const float* P, const float* X, int B, Acts& a, hipStream_t st) {
  for (int l = 0; l < 3; ++l) {
    a.Z[l].ensure((size_t)B * NN_H);
    a.A[l].ensure((size_t)B * NN_H);
  }
  a.Yp.ensure((size_t)B * NN_OUT);
  gemm(B, NN_H, NN_IN, X, NN_IN, 1, P + NN_OFF_W1, 1, NN_IN, a.Z[0].p, NN_H, 1, P + NN_OFF_B1, 0, st);
  silu_kernel<<<blocks((size_t)B * NN_H), 256, 0, st>>>(a.Z[0].p, a.A[0].p, (size_t)B * NN_H);
  gemm(B, NN_H, NN_H, a.A[0].p, NN_H, 1, P + NN_OFF_W2, 1, NN_H, a.Z[1].p, NN_H, 1, P + NN_OFF_B2, 0, st);
  silu_kernel<<<blocks((size_t)B * NN_H), 256, 0, st>>>(a.Z[1].p, a.A[1].p, (size_t)B * NN_H);
  gemm(B, NN_H, NN_H, a.A[1].p, NN_H, 1, P + NN_OFF_W3, 1, NN_H, a.Z[2].p, NN_H, 1, P + NN_OFF_B3, 0, st);
  silu_kernel<<<blocks((size_t)B * NN_H), 256, 0, st>>>(a.Z[2].p, a.A[2].p, (size_t)B * NN_H);
  gemm(B, NN_OUT, NN_H, a.A[2].p, NN_H, 1, P + NN_OFF_W4, 1, NN_H, a.Yp.p, NN_OUT, 1, P + NN_OFF_B4, 0, st);
}

// gradient G (NnParams layout) of the mean squared error of the batch; dY in a.dB ([B][4])
void backward(const float* P, const float* X, int B, Acts& a, float* G, hipStream_t st) {
  a.dA.ensure((size_t)B * NN_H);
  a.dB.ensure((size_t)B * NN_H);
  const float* dY = a.dB.p;
  // layer 4: dW4 [4][128] = dY^T A3, db4, dA3 = dY W4
  gemm(NN_OUT, NN_H, B, dY, 1, NN_OUT, a.A[2].p, NN_H, 1, G + NN_OFF_W4, NN_H, 1, nullptr, 0, st);
  colsum_kernel<<<NN_OUT, 256, 0, st>>>(dY, B, NN_OUT, G + NN_OFF_B4);
  gemm(B, NN_H, NN_OUT, dY, NN_OUT, 1, P + NN_OFF_W4, NN_H, 1, a.dA.p, NN_H, 1, nullptr, 0, st);
  float* cur = a.dA.p;   // dA of the layer being processed
  float* nxt = a.dB.p;   // dY no longer needed: reused for the next dA
  const int offW[3] = {NN_OFF_W1, NN_OFF_W2, NN_OFF_W3}, offB[3] = {NN_OFF_B1, NN_OFF_B2, NN_OFF_B3};
  for (int l = 2; l >= 0; --l) {
    silu_back_kernel<<<blocks((size_t)B * NN_H), 256, 0, st>>>(a.Z[l].p, cur, (size_t)B * NN_H);  // cur = dZ
    const float* in = l > 0 ? a.A[l - 1].p : X;
    const int K = l > 0 ? NN_H : NN_IN;
    gemm(NN_H, K, B, cur, 1, NN_H, in, K, 1, G + offW[l], K, 1, nullptr, 0, st);
    colsum_kernel<<<NN_H, 256, 0, st>>>(cur, B, NN_H, G + offB[l]);
    if (l > 0) {
      gemm(B, NN_H, NN_H, cur, NN_H, 1, P + offW[l], NN_H, 1, nxt, NN_H, 1, nullptr, 0, st);
      std::swap(cur, nxt);
    }
  }
}

void init_params(std::vector<float>& P, unsigned long long seed) {
  P.assign(NN_PARAMS, 0.f);
  unsigned long long s = seed ^ 0x5DEECE66Dull;
  auto rnd = [&]() {  // uniform in [-1, 1)
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  };
  auto fill = [&](int off, int out, int in) {
    const double a = std::sqrt(6.0 / (in + out));
    for (int i = 0; i < out * in; ++i) P[off + i] = (float)(a * rnd());
  };
  fill(NN_OFF_W1, NN_H, NN_IN);
  fill(NN_OFF_W2, NN_H, NN_H);
  fill(NN_OFF_W3, NN_H, NN_H);
  fill(NN_OFF_W4, NN_OUT, NN_H);
}

}  // namespace

// The loss and its gradient for a batch on the device (the backward pass of the training), for
// the gradient check against the CPU restatement.
double nn_gradient(const float* P_host, const float* X_host, const float* Y_host, int B, float* G_host,
                   hipStream_t st) {
  Buf<float> P, X, G;
  Buf<float4> Y;
  Buf<double> loss;
  P.ensure(NN_PARAMS);
  G.ensure(NN_PARAMS);
  X.ensure((size_t)B * NN_IN);
  Y.ensure(B);
  loss.ensure(1);
  HBG_HIP(hipMemcpyAsync(P.p, P_host, NN_PARAMS * sizeof(float), hipMemcpyHostToDevice, st));
  HBG_HIP(hipMemcpyAsync(X.p, X_host, (size_t)B * NN_IN * sizeof(float), hipMemcpyHostToDevice, st));
  HBG_HIP(hipMemcpyAsync(Y.p, Y_host, (size_t)B * sizeof(float4), hipMemcpyHostToDevice, st));
  HBG_HIP(hipMemsetAsync(loss.p, 0, sizeof(double), st));
  Acts a;
  forward(P.p, X.p, B, a, st);
  a.dB.ensure((size_t)B * NN_H);
  loss_kernel<<<blocks(B), 256, 0, st>>>(a.Yp.p, Y.p, B, a.dB.p, loss.p);
  backward(P.p, X.p, B, a, G.p, st);
  HBG_HIP(hipGetLastError());
  double l = 0.0;
  HBG_HIP(hipMemcpyAsync(&l, loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
  HBG_HIP(hipMemcpyAsync(G_host, G.p, NN_PARAMS * sizeof(float), hipMemcpyDeviceToHost, st));
  HBG_HIP(hipStreamSynchronize(st));
  return l;
}

// Training data: n samples in groups of `group` with one nu h each (log-uniform), inputs X and
// targets Y on the device.
void nn_make_data(unsigned long long seed, unsigned long long first, int n, int group, const WaveTables& T,
                  const double* ktab, const double* fitm, const double* g16, float4* X, float4* Y,
                  hipStream_t st) {
  Buf<double> R, z, zeta, phi1, phi2, fwork;
  Buf<PanelD> pts;
  Buf<double4> pf;
  Buf<double2> pcs;
  Buf<FinFreq> ff;
  R.ensure(group);
  z.ensure(group);
  zeta.ensure(group);
  pts.ensure(2 * (size_t)group);
  pf.ensure(2 * (size_t)group);
  pcs.ensure(2 * (size_t)group * FIN_ROOTS);
  phi1.ensure((size_t)FTS * FTW * FTILE);
  phi2.ensure((size_t)FTS * FTV * FTILE);
  fwork.ensure(fin_tables_work());
  ff.ensure(1);
  HBG_HIP(hipMemsetAsync(pts.p, 0, 2 * (size_t)group * sizeof(PanelD), st));
  unsigned long long gs = seed * 0x2545F4914F6CDD1Dull + 0x1234567ull;
  for (int g0 = 0; g0 < n; g0 += group) {
    const int m = std::min(group, n - g0);
    gs = gs * 6364136223846793005ull + 1442695040888963407ull;
    const double u = (double)(gs >> 11) * (1.0 / 9007199254740992.0);
    const double nuh = std::exp(NN_LNNU_LO + (NN_LNNU_HI - NN_LNNU_LO) * u);
    const FinFreq f = make_fin_freq(1.0, nuh);
    sample_kernel<<<blocks(m), 256, 0, st>>>(seed, first + g0, m, R.p, z.p, zeta.p, pts.p);
    launch_fin_tables(f, T, g16, phi1.p, phi2.p, fitm, pts.p, 2 * m, pf.p, pcs.p, fwork.p, st);
    HBG_HIP(hipMemcpyAsync(ff.p, &f, sizeof(FinFreq), hipMemcpyHostToDevice, st));
    FinTables F;
    F.phi1 = phi1.p;
    F.phi2 = phi2.p;
    F.ktab = ktab;
    F.pf = pf.p;
    F.pcs = pcs.p;
    target_kernel<<<(unsigned)((m + 127) / 128), 128, 0, st>>>(ff.p, F, T, m, R.p, z.p, zeta.p, X + g0, Y + g0);
    HBG_HIP(hipGetLastError());
    HBG_HIP(hipStreamSynchronize(st));  // the group's host-side FinFreq is reused next
  }
}

// Train the network: n_train + n_hold samples, `epochs` passes of Adam over minibatches of
// `batch`, learning rate lr0 decayed by cosine to lr0 / 100. The parameters go to P (NnParams
// layout); report: [0] final training loss (normalized MSE), [1] held-out RMS error of the BF16
// network (normalized units), [2..5] held-out max |error| per output (normalized), [6] data
// seconds, [7] training seconds, [8] steps, [9] non-finite targets among the first 2^20 samples,
// [10] non-finite parameters after the training.
void nn_train(unsigned long long seed, int n_train, int n_hold, int epochs, int batch, float lr0, const WaveTables& T,
              const double* ktab, const double* fitm, const double* g16, std::vector<float>& P, double* report,
              NnDevice& nn, hipStream_t st) {
  typedef std::chrono::steady_clock clk;
  const auto t0 = clk::now();
  const int n = n_train + n_hold;
  Buf<float4> X, Y;
  X.ensure(n);
  Y.ensure(n);
  nn_make_data(seed, 0, n, 8192, T, ktab, fitm, g16, X.p, Y.p, st);
  // per-output mean and standard deviation (from up to 2^20 training samples); the network learns
  // the standardized targets and the last layer is scaled back after the training
  double mu[4] = {0, 0, 0, 0}, sd[4] = {0, 0, 0, 0};
  long nonfinite = 0;
  {
    const int m = std::min(n_train, 1 << 20);
    std::vector<float4> hy(m);
    HBG_HIP(hipMemcpyAsync(hy.data(), Y.p, m * sizeof(float4), hipMemcpyDeviceToHost, st));
    HBG_HIP(hipStreamSynchronize(st));
    for (const float4& v : hy) {
      const double a[4] = {v.x, v.y, v.z, v.w};
      if (!(std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]) && std::isfinite(a[3]))) {
        ++nonfinite;
        continue;
      }
      for (int q = 0; q < 4; ++q) {
        mu[q] += a[q];
        sd[q] += a[q] * a[q];
      }
    }
    const double me = (double)std::max<long>(m - nonfinite, 1);
    for (int q = 0; q < 4; ++q) {
      mu[q] /= me;
      sd[q] = std::sqrt(std::max(sd[q] / me - mu[q] * mu[q], 1e-30));
    }
    standardize_kernel<<<blocks(n), 256, 0, st>>>(Y.p, n, make_float4((float)mu[0], (float)mu[1], (float)mu[2], (float)mu[3]),
                                                  make_float4((float)(1.0 / sd[0]), (float)(1.0 / sd[1]),
                                                              (float)(1.0 / sd[2]), (float)(1.0 / sd[3])));
  }
  const auto t1 = clk::now();
  init_params(P, seed);
  Buf<float> dP, dG, dM, dV, Xb;
  Buf<float4> Yb;
  Buf<double> loss;
  Buf<int> perm;
  dP.ensure(NN_PARAMS);
  dG.ensure(NN_PARAMS);
  dM.ensure(NN_PARAMS);
  dV.ensure(NN_PARAMS);
  Xb.ensure((size_t)batch * NN_IN);
  Yb.ensure(batch);
  loss.ensure(1);
  perm.ensure(n_train);
  HBG_HIP(hipMemcpyAsync(dP.p, P.data(), NN_PARAMS * sizeof(float), hipMemcpyHostToDevice, st));
  HBG_HIP(hipMemsetAsync(dM.p, 0, NN_PARAMS * sizeof(float), st));
  HBG_HIP(hipMemsetAsync(dV.p, 0, NN_PARAMS * sizeof(float), st));
  std::vector<int> hperm(n_train);
  for (int i = 0; i < n_train; ++i) hperm[i] = i;
  unsigned long long rs = seed ^ 0xA5A5A5A5ull;
  const int per_epoch = n_train / batch;
  const long total = (long)per_epoch * epochs;
  long step = 0;
  Acts a;
  double last = 0.0;
  const float b1 = 0.9f, b2 = 0.999f;
  for (int e = 0; e < epochs; ++e) {
    for (int i = n_train - 1; i > 0; --i) {  // Fisher-Yates on the host
      rs = rs * 6364136223846793005ull + 1442695040888963407ull;
      std::swap(hperm[i], hperm[(size_t)(rs >> 33) % (size_t)(i + 1)]);
    }
    HBG_HIP(hipStreamSynchronize(st));  // the previous epoch has read perm
    HBG_HIP(hipMemcpyAsync(perm.p, hperm.data(), n_train * sizeof(int), hipMemcpyHostToDevice, st));
    for (int b = 0; b < per_epoch; ++b, ++step) {
      gather_kernel<<<blocks(batch), 256, 0, st>>>(X.p, Y.p, perm.p, b * batch, batch, Xb.p, Yb.p);
      forward(dP.p, Xb.p, batch, a, st);
      a.dB.ensure((size_t)batch * NN_H);
      const bool rec = b == per_epoch - 1;
      if (rec) HBG_HIP(hipMemsetAsync(loss.p, 0, sizeof(double), st));
      loss_kernel<<<blocks(batch), 256, 0, st>>>(a.Yp.p, Yb.p, batch, a.dB.p, loss.p);
      backward(dP.p, Xb.p, batch, a, dG.p, st);
      const double frac = (double)step / (double)std::max<long>(total, 1);
      const float lr = (float)(lr0 * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac))));
      const float c1 = 1.0f / (1.0f - std::pow(b1, (float)(step + 1))),
                  c2 = 1.0f / (1.0f - std::pow(b2, (float)(step + 1)));
      adam_kernel<<<blocks(NN_PARAMS), 256, 0, st>>>(dP.p, dG.p, dM.p, dV.p, NN_PARAMS, lr, b1, b2, c1, c2);
      if (rec) {
        HBG_HIP(hipMemcpyAsync(&last, loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
        HBG_HIP(hipStreamSynchronize(st));
      }
    }
  }
  HBG_HIP(hipGetLastError());
  HBG_HIP(hipMemcpyAsync(P.data(), dP.p, NN_PARAMS * sizeof(float), hipMemcpyDeviceToHost, st));
  HBG_HIP(hipStreamSynchronize(st));
  for (int q = 0; q < NN_OUT; ++q) {  // undo the standardization in the last layer
    for (int i = 0; i < NN_H; ++i) P[NN_OFF_W4 + q * NN_H + i] *= (float)sd[q];
    P[NN_OFF_B4 + q] = (float)(P[NN_OFF_B4 + q] * sd[q] + mu[q]);
  }
  standardize_kernel<<<blocks(n), 256, 0, st>>>(Y.p, n, make_float4(0.f, 0.f, 0.f, 0.f),
                                                make_float4((float)sd[0], (float)sd[1], (float)sd[2], (float)sd[3]));
  standardize_kernel<<<blocks(n), 256, 0, st>>>(Y.p, n, make_float4(-(float)mu[0], -(float)mu[1], -(float)mu[2],
                                                                    -(float)mu[3]),
                                                make_float4(1.f, 1.f, 1.f, 1.f));
  const auto t2 = clk::now();
  nn.upload(P.data());
  // held-out error of the BF16 network
  Buf<float4> Yh;
  Yh.ensure(std::max(n_hold, 1));
  nn_eval(nn, X.p + n_train, n_hold, Yh.p, st);
  std::vector<float4> ht(n_hold), hp(n_hold);
  HBG_HIP(hipMemcpyAsync(ht.data(), Y.p + n_train, n_hold * sizeof(float4), hipMemcpyDeviceToHost, st));
  HBG_HIP(hipMemcpyAsync(hp.data(), Yh.p, n_hold * sizeof(float4), hipMemcpyDeviceToHost, st));
  HBG_HIP(hipStreamSynchronize(st));
  double se = 0.0, mx[4] = {0, 0, 0, 0};
  for (int i = 0; i < n_hold; ++i) {
    const float d[4] = {hp[i].x - ht[i].x, hp[i].y - ht[i].y, hp[i].z - ht[i].z, hp[i].w - ht[i].w};
    for (int q = 0; q < 4; ++q) {
      se += (double)d[q] * d[q];
      mx[q] = std::max(mx[q], (double)std::fabs(d[q]));
    }
  }
  for (int q = 0; q < 4; ++q) mx[q] /= sd[q];  // in units of each output's standard deviation
  se /= (sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2] + sd[3] * sd[3]) / 4.0;
  long bad_params = 0;
  for (float v : P) bad_params += std::isfinite(v) ? 0 : 1;
  report[9] = (double)nonfinite;
  report[10] = (double)bad_params;
  report[0] = last;
  report[1] = n_hold > 0 ? std::sqrt(se / (4.0 * n_hold)) : 0.0;
  for (int q = 0; q < 4; ++q) report[2 + q] = mx[q];
  report[6] = std::chrono::duration<double>(t1 - t0).count();
  report[7] = std::chrono::duration<double>(t2 - t1).count();
  report[8] = (double)step;
}

// queue: images N (N + 1) / 2 pairs of room (every pair may fall back), qcount: one counter.
// images > 1 (a symmetric body or walls): grid z = g assembles the image block g (sources the
// records P + g N, S, A and the statics at g N N, leading dimension N), as assemble_fin_kernel.
void launch_assemble_fin_nn(const PanelD* P, int N, const FinFreq* f, FinTables F, WaveTables T, double logc,
                            const StaticD* st, double2* S, double2* A, float2* A32, int lda, const NnDevice& nn,
                            int2* queue, unsigned* qcount, double rmin, hipStream_t stream, int images) {
  if (images > 1 && (A32 || lda != N)) throw Error{6, "image blocks without a complex64 copy, at leading dimension N"};
  zero_u32_kernel<<<1, 64, 0, stream>>>(qcount);
  const unsigned nt = (unsigned)((N + 63) / 64);
  dim3 grid(nt, nt, images);
  assemble_fin_nn_kernel<<<grid, 256, 0, stream>>>(P, N, f, F, T, logc, st, S, A, A32, lda, (const v8bf*)nn.pk,
                                                   nn.bias, queue, qcount, std::max(rmin, NN_RMIN));
  fallback_kernel<<<2048, 256, 0, stream>>>(P, N, f, F, T, logc, st, S, A, A32, lda, queue, qcount);
  HBG_HIP(hipGetLastError());
}

void launch_remainder(const FinFreq* f, FinTables F, WaveTables T, int n, const double* R, const double* z,
                      const double* zeta, double* out, hipStream_t stream) {
  remainder_kernel<<<(n + 127) / 128, 128, 0, stream>>>(f, F, T, n, R, z, zeta, out);
  HBG_HIP(hipGetLastError());
}

}  // namespace hbg
