// radix_sort.h -- stable LSD radix sort of (cell key, value) pairs for gfx950.
//
// Each pass sorts on one digit of up to 11 bits:
//   k_sort_hist    per-tile digit histogram in LDS
//   k_sort_scan    exclusive scan of the [digit][tile] histogram (one block)
//   k_sort_scatter per-tile stable ranking in LDS and scatter
// A tile is 256 lanes x 16 rounds.  Inside a round each 64-lane wavefront
// ranks its lanes among equal digits with ballots (one ballot per digit
// bit), the four wavefronts of the block are combined through LDS counts,
// and a running per-digit offset in LDS carries the order across rounds, so
// equal keys keep their input order (stability).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace arm {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;
constexpr int SORT_MAX_BITS = 11;
constexpr int SORT_MAX_BINS = 1 << SORT_MAX_BITS;
constexpr int SORT_WAVES = SORT_THREADS / 64;

__global__ __launch_bounds__(SORT_THREADS) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                            int n, int shift, int bins,
                                                            int ntiles,
                                                            uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[SORT_MAX_BINS];
  for (int d = threadIdx.x; d < bins; d += SORT_THREADS) h[d] = 0;
  __syncthreads();
  const int base = blockIdx.x * SORT_TILE;
  const uint32_t mask = (uint32_t)bins - 1u;
#pragma unroll 4
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int idx = base + r * SORT_THREADS + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < bins; d += SORT_THREADS) hist[d * ntiles + blockIdx.x] = h[d];
}

// Exclusive scan of `total` entries in one block of 1024 lanes.
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t* __restrict__ hist, int total) {
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const int chunk = (total + 1023) / 1024;
  const int b = tid * chunk;
  const int e = min(b + chunk, total);
  uint32_t s = 0;
  for (int k = b; k < e; ++k) s += hist[k];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = part[tid] - s;  // exclusive prefix of this chunk
  for (int k = b; k < e; ++k) {
    const uint32_t c = hist[k];
    hist[k] = run;
    run += c;
  }
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n, int shift, int bits,
    int ntiles, const uint32_t* __restrict__ hist) {
  __shared__ uint32_t off[SORT_MAX_BINS];
  __shared__ uint32_t wcnt[SORT_WAVES][SORT_MAX_BINS];
  const int bins = 1 << bits;
  const uint32_t mask = (uint32_t)bins - 1u;
  for (int d = threadIdx.x; d < bins; d += SORT_THREADS) {
    off[d] = hist[d * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) wcnt[w][d] = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int base = blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int idx = base + r * SORT_THREADS + threadIdx.x;
    const bool valid = idx < n;
    const uint32_t key = valid ? kin[idx] : 0u;
    const uint32_t val = valid ? vin[idx] : 0u;
    const uint32_t d = (key >> shift) & mask;
    uint64_t match = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      match &= bit ? bb : ~bb;
    }
    if (!valid) match = 0;
    const uint32_t rank = (uint32_t)__popcll(match & lt);
    const uint32_t cnt = (uint32_t)__popcll(match);
    const bool leader = valid && rank == 0;
    if (leader) wcnt[w][d] = cnt;
    __syncthreads();
    if (valid) {
      uint32_t pos = off[d] + rank;
      for (int w2 = 0; w2 < w; ++w2) pos += wcnt[w2][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    if (leader) atomicAdd(&off[d], cnt);
    __syncthreads();
    if (leader) wcnt[w][d] = 0;
  }
}

}  // namespace arm
