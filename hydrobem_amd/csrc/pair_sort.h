// The pair sort of the assembly kernels (FP64 frequency.hip, FP32 assembly32.hip): a workgroup
// takes a 64 x 64 tile of (field point k, source n) pairs and sorts them by table tile in LDS.
#pragma once

#include <hip/hip_runtime.h>

namespace hbg {

constexpr int ASM_TILE = 64;

// The tables make the kernels' cost: a pair reads hundreds of coefficients (the V tile and the 1D
// tile in infinite depth; the two Phi tiles and the V tile near the axis, one K0/K1 tile per root
// in the series of finite depth), and read lane by lane they bound the kernel by the vector
// memory pipeline. So a workgroup first sorts its tile's pairs by table tile (a counting sort in
// LDS), and each wavefront then evaluates 64 consecutive sorted pairs with tile reads that are
// wavefront-uniform: scalar loads, made once per distinct tile of the wavefront.
constexpr int PS_KEYS = 4096;                              // keys of the sort
constexpr int PS_PER_THREAD = ASM_TILE * ASM_TILE / 256;   // pairs per thread of the tile
struct PairSort {
  unsigned cnt[PS_KEYS];
  unsigned short order[ASM_TILE * ASM_TILE];  // pair p = k_local + 64 n_local, in key order
  unsigned wsum[4];
};

// Sorts the pairs (k, n) of the tile (bi, bj) (k <= n on the diagonal tile, both < N) by
// key_of(k, n) in [0, PS_KEYS); returns their number. Order within a key is arbitrary: every
// pair is evaluated on its own.
template <class KeyOf>
__device__ __forceinline__ unsigned sort_tile_pairs(PairSort& L, int bi, int bj, int N, KeyOf key_of) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < PS_KEYS; i += 256) L.cnt[i] = 0;
  __syncthreads();
  unsigned kr[PS_PER_THREAD];  // key << 13 | rank within the key, or ~0 for no pair
#pragma unroll
  for (int j = 0; j < PS_PER_THREAD; ++j) {
    const int p = tid + 256 * j, k = bi * ASM_TILE + (p & 63), n = bj * ASM_TILE + (p >> 6);
    kr[j] = 0xffffffffu;
    if (k < N && n < N && !(bi == bj && k > n)) {
      const int kf = key_of(k, n);
      const unsigned key = kf < 0 ? 0u : (kf >= PS_KEYS ? (unsigned)PS_KEYS - 1 : (unsigned)kf);
      kr[j] = (key << 13) | atomicAdd(&L.cnt[key], 1u);
    }
  }
  __syncthreads();
  // exclusive scan of the counts: thread t owns cnt[16 t .. 16 t + 15]
  constexpr int OWN = PS_KEYS / 256;
  unsigned loc[OWN], sum = 0;
#pragma unroll
  for (int i = 0; i < OWN; ++i) {
    loc[i] = sum;
    sum += L.cnt[OWN * tid + i];
  }
  unsigned incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned o = __shfl_up(incl, off);
    if (lane >= off) incl += o;
  }
  if (lane == 63) L.wsum[wave] = incl;
  __syncthreads();
  unsigned base = incl - sum;
  for (int w = 0; w < wave; ++w) base += L.wsum[w];
  const unsigned total = L.wsum[0] + L.wsum[1] + L.wsum[2] + L.wsum[3];
#pragma unroll
  for (int i = 0; i < OWN; ++i) L.cnt[OWN * tid + i] = base + loc[i];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PS_PER_THREAD; ++j)
    if (kr[j] != 0xffffffffu) L.order[L.cnt[kr[j] >> 13] + (kr[j] & 8191u)] = (unsigned short)(tid + 256 * j);
  __syncthreads();
  return total;
}

}  // namespace hbg
