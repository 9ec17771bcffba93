// radix_sort.h -- stable LSD radix sort of (cell key, value) pairs for gfx950.
//
// Each pass sorts on one digit of up to 11 bits with three kernels:
//   k_sort_hist    per-tile digit histogram in LDS          -> hist[d][tile]
//   k_sort_scan    one wavefront per digit: exclusive scan of hist[d][*]
//                  over the tiles, and the digit total tot[d]
//   k_sort_scatter exclusive scan of tot[] in LDS (digit bases), then a
//                  stable ranking of the tile in LDS and the scatter
// A tile is 256 lanes x 16 rounds; each lane loads its 16 elements up front.
// Inside a round each 64-lane wavefront ranks its lanes among equal digits
// with ballots (one per digit bit); the four wavefronts of the block are
// combined through LDS counts and a running per-digit offset carries the
// order across rounds, so equal keys keep their input order (stability).
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
  const int base = blockIdx.x * SORT_TILE;
  const uint32_t mask = (uint32_t)bins - 1u;
  uint32_t k[SORT_ROUNDS];
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int idx = base + r * SORT_THREADS + threadIdx.x;
    k[r] = idx < n ? keys[idx] : 0xffffffffu;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int idx = base + r * SORT_THREADS + threadIdx.x;
    if (idx < n) atomicAdd(&h[(k[r] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < bins; d += SORT_THREADS) hist[d * ntiles + blockIdx.x] = h[d];
}

// Inclusive scan across the 64 lanes of a wavefront.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// One 64-lane block per digit d: hist[d][b] <- sum_{b' < b} hist[d][b'],
// tot[d] <- sum_b hist[d][b].
__global__ __launch_bounds__(64) void k_sort_scan(uint32_t* __restrict__ hist,
                                                  uint32_t* __restrict__ tot, int ntiles) {
  const int d = blockIdx.x;
  const int lane = threadIdx.x;
  uint32_t carry = 0;
  uint32_t* row = hist + (size_t)d * ntiles;
  for (int b0 = 0; b0 < ntiles; b0 += 64) {
    const int b = b0 + lane;
    const uint32_t v = b < ntiles ? row[b] : 0u;
    const uint32_t inc = wave_incl_scan(v);
    if (b < ntiles) row[b] = carry + inc - v;
    carry += __shfl(inc, 63, 64);
  }
  if (lane == 0) tot[d] = carry;
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n, int shift, int bits,
    int ntiles, const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot) {
  __shared__ uint32_t off[SORT_MAX_BINS];
  __shared__ uint32_t wcnt[SORT_WAVES][SORT_MAX_BINS];
  __shared__ uint32_t wsum[SORT_WAVES];
  const int bins = 1 << bits;
  const uint32_t mask = (uint32_t)bins - 1u;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int base = blockIdx.x * SORT_TILE;
  // prefetch the tile
  uint32_t kk[SORT_ROUNDS], vv[SORT_ROUNDS];
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int idx = base + r * SORT_THREADS + threadIdx.x;
    kk[r] = idx < n ? kin[idx] : 0u;
    vv[r] = idx < n ? vin[idx] : 0u;
  }
  // digit bases: exclusive scan of tot[] (each lane owns a contiguous run)
  const int per = (bins + SORT_THREADS - 1) / SORT_THREADS;
  uint32_t run[SORT_MAX_BINS / SORT_THREADS];
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < SORT_MAX_BINS / SORT_THREADS; ++q) {
    const int d = threadIdx.x * per + q;
    run[q] = (q < per && d < bins) ? tot[d] : 0u;
    s += run[q];
  }
  const uint32_t inc = wave_incl_scan(s);
  if (lane == 63) wsum[w] = inc;
  for (int d = threadIdx.x; d < bins; d += SORT_THREADS)
#pragma unroll
    for (int w2 = 0; w2 < SORT_WAVES; ++w2) wcnt[w2][d] = 0;
  __syncthreads();
  uint32_t pre = inc - s;
  for (int w2 = 0; w2 < w; ++w2) pre += wsum[w2];
#pragma unroll
  for (int q = 0; q < SORT_MAX_BINS / SORT_THREADS; ++q) {
    const int d = threadIdx.x * per + q;
    if (q < per && d < bins) {
      off[d] = pre + hist[(size_t)d * ntiles + blockIdx.x];
      pre += run[q];
    }
  }
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int idx = base + r * SORT_THREADS + threadIdx.x;
    const bool valid = idx < n;
    const uint32_t key = kk[r];
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
      vout[pos] = vv[r];
    }
    __syncthreads();
    if (leader) atomicAdd(&off[d], cnt);
    __syncthreads();
    if (leader) wcnt[w][d] = 0;
  }
}

}  // namespace arm
