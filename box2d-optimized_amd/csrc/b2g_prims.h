// b2g_prims.h -- device-wide primitives sized for 64-wide wavefronts:
// exclusive scan of int32 (three kernels), stable LSD radix sort of
// (uint64 key, uint32 value) pairs with 8-bit digits, and small helpers.
//
// Counts may live in device memory (n_dev): kernels read the count at run
// time and loop grid-stride, so the launches need no host round trip.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace b2g {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kScanItems = 4;                       // per thread
constexpr int kScanTile = kBlock * kScanItems;      // 1024 per block tile
constexpr int kSortItems = 4;
constexpr int kSortTile = kBlock * kSortItems;      // 1024 keys per tile
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / kWave); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// inclusive wave scan of ints
__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int l = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_up(v, off, kWave);
    if (l >= off) v += t;
  }
  return v;
}

// exclusive block scan (kBlock threads); returns the block total in *total
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds_waves /* kBlock/kWave */, int* total) {
  const int inc = wave_inclusive_scan(v);
  if (lane_id() == kWave - 1) lds_waves[wave_id()] = inc;
  __syncthreads();
  int wave_off = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    const int s = lds_waves[w];
    if (w < wave_id()) wave_off += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return wave_off + inc - v;
}

// ---- exclusive scan: out[i] = sum(in[0..i)), total -> *total_out -----------
// pass 1: per-tile sums
__global__ void __launch_bounds__(kBlock) k_scan_tiles(const int* __restrict__ in, const int* __restrict__ n_dev,
                                                       int n_host, int* __restrict__ tile_sums) {
  __shared__ int lds[kBlock / kWave];
  const int n = n_dev ? *n_dev : n_host;
  const int tiles = (n + kScanTile - 1) / kScanTile;
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    int s = 0;
    const int base = t * kScanTile + threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
      if (base + k < n) s += in[base + k];
    int total;
    block_exclusive_scan(s, lds, &total);
    if (threadIdx.x == 0) tile_sums[t] = total;
  }
}

// pass 2: one block scans the tile sums in place; writes the grand total
__global__ void __launch_bounds__(kBlock) k_scan_sums(int* __restrict__ tile_sums, const int* __restrict__ n_dev,
                                                      int n_host, int* __restrict__ total_out) {
  __shared__ int lds[kBlock / kWave];
  const int n = n_dev ? *n_dev : n_host;
  const int tiles = (n + kScanTile - 1) / kScanTile;
  int carry = 0;
  for (int base = 0; base < tiles; base += kBlock) {
    const int i = base + (int)threadIdx.x;
    const int v = i < tiles ? tile_sums[i] : 0;
    int total;
    const int ex = block_exclusive_scan(v, lds, &total);
    if (i < tiles) tile_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// pass 3: scan inside each tile and add the tile offset (in and out may alias:
// every thread reads its own items before the block barrier, writes after)
__global__ void __launch_bounds__(kBlock) k_scan_apply(const int* in, const int* __restrict__ n_dev, int n_host,
                                                       const int* __restrict__ tile_sums, int* out) {
  __shared__ int lds[kBlock / kWave];
  const int n = n_dev ? *n_dev : n_host;
  const int tiles = (n + kScanTile - 1) / kScanTile;
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int base = t * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < n ? in[base + k] : 0;
      s += v[k];
    }
    int total;
    int ex = block_exclusive_scan(s, lds, &total) + tile_sums[t];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < n) out[base + k] = ex;
      ex += v[k];
    }
  }
}

// passes 2 and 3 in one launch, for up to kScanFoldTiles tiles: each block
// adds up the raw sums of the tiles before its own (a short reduction), so no
// separate pass scans the tile sums; the block of the last tile writes the
// grand total
constexpr int kScanFoldTiles = 1024;
__global__ void __launch_bounds__(kBlock) k_scan_apply_fold(const int* in, const int* __restrict__ n_dev,
                                                            int n_host, const int* __restrict__ tile_sums, int* out,
                                                            int* __restrict__ total_out) {
  __shared__ int lds[kBlock / kWave];
  const int n = n_dev ? *n_dev : n_host;
  const int tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 0 && blockIdx.x == 0 && threadIdx.x == 0 && total_out) *total_out = 0;
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    int part = 0;
    for (int u = threadIdx.x; u < t; u += blockDim.x) part += tile_sums[u];
    int offset;
    (void)block_exclusive_scan(part, lds, &offset);
    const int base = t * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < n ? in[base + k] : 0;
      s += v[k];
    }
    int total;
    int ex = block_exclusive_scan(s, lds, &total) + offset;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < n) out[base + k] = ex;
      ex += v[k];
    }
    if (t == tiles - 1 && threadIdx.x == 0 && total_out) *total_out = offset + total;
  }
}

// ---- one-block exclusive scan for short arrays: each of the 1024 threads
// owns a contiguous chunk (sum, block scan of the sums, rewrite); one launch
// instead of three. Correct for any n, meant for n up to ~64k. In and out may
// alias: every thread reads its whole chunk before the block barrier.
constexpr int kScan1Threads = 1024;
__global__ void __launch_bounds__(kScan1Threads) k_scan_single(const int* in, const int* __restrict__ n_dev,
                                                               int n_host, int* out, int* __restrict__ total_out) {
  __shared__ int waves[kScan1Threads / kWave];
  const int n = n_dev ? *n_dev : n_host;
  const int chunk = (n + kScan1Threads - 1) / kScan1Threads;
  const int lo = min(n, (int)threadIdx.x * chunk), hi = min(n, lo + chunk);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += in[i];
  const int inc = wave_inclusive_scan(s);
  if (lane_id() == kWave - 1) waves[wave_id()] = inc;
  __syncthreads();
  int before = 0, total = 0;
  for (int w = 0; w < kScan1Threads / kWave; ++w) {
    const int v = waves[w];
    if (w < wave_id()) before += v;
    total += v;
  }
  int run = before + inc - s;
  for (int i = lo; i < hi; ++i) {
    const int v = in[i];
    out[i] = run;
    run += v;
  }
  if (threadIdx.x == 0 && total_out) *total_out = total;
}

// ---- stable LSD radix sort of (key, value), 8-bit digits -------------------
// histogram layout: hist[digit * tiles + tile] so one exclusive scan over the
// whole array yields the scatter base of every (digit, tile).
__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint64_t* __restrict__ keys, int n, int shift,
                                                       int* __restrict__ hist) {
  __shared__ int cnt[kRadix];
  const int tiles = (n + kSortTile - 1) / kSortTile;
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    for (int d = threadIdx.x; d < kRadix; d += kBlock) cnt[d] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
      const int i = t * kSortTile + k * kBlock + (int)threadIdx.x;
      if (i < n) atomicAdd(&cnt[(int)((keys[i] >> shift) & (kRadix - 1))], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kBlock) hist[d * tiles + t] = cnt[d];
    __syncthreads();
  }
}

// Ranks every key of a tile among equal digits in tile order (stable), then
// scatters to offsets[digit * tiles + tile] + rank.
__global__ void __launch_bounds__(kBlock) k_radix_scatter(const uint64_t* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ vals_in, int n, int shift,
                                                          const int* __restrict__ offsets,
                                                          uint64_t* __restrict__ keys_out,
                                                          uint32_t* __restrict__ vals_out) {
  constexpr int kWaves = kBlock / kWave;
  __shared__ int wave_cnt[kWaves][kRadix];
  __shared__ int running[kRadix];
  const int tiles = (n + kSortTile - 1) / kSortTile;
  const int w = wave_id();
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    for (int d = threadIdx.x; d < kRadix; d += kBlock) running[d] = offsets[d * tiles + t];
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < kSortItems; ++k) {
      for (int d = threadIdx.x; d < kWaves * kRadix; d += kBlock) (&wave_cnt[0][0])[d] = 0;
      __syncthreads();
      const int i = t * kSortTile + k * kBlock + (int)threadIdx.x;
      const bool valid = i < n;
      uint64_t key = valid ? keys_in[i] : 0;
      const uint32_t digit = valid ? (uint32_t)((key >> shift) & (kRadix - 1)) : 0u;
      // lanes of this wave with the same digit (8 ballots)
      uint64_t same = __ballot(valid);
#pragma unroll
      for (int b = 0; b < kRadixBits; ++b) {
        const bool bit = (digit >> b) & 1u;
        const uint64_t m = __ballot(bit);
        same &= bit ? m : ~m;
      }
      const int rank_in_wave = __popcll(same & lanemask_lt());
      if (valid && rank_in_wave == 0) wave_cnt[w][digit] = __popcll(same);
      __syncthreads();
      if (valid) {
        int before = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q)
          if (q < w) before += wave_cnt[q][digit];
        const int dst = running[digit] + before + rank_in_wave;
        if (dst >= 0 && dst < n) {
          keys_out[dst] = key;
          vals_out[dst] = vals_in[i];
        }
      }
      __syncthreads();
      for (int d = threadIdx.x; d < kRadix; d += kBlock) {
        int s = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) s += wave_cnt[q][d];
        running[d] += s;
      }
      __syncthreads();
    }
  }
}

// ordered-integer encoding of floats for atomicMin/atomicMax
__device__ __forceinline__ uint32_t float_to_ordered(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

}  // namespace b2g
