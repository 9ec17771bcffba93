// cell_sort.h -- counting sort of particles into hash cells for gfx950.
//
// Produces the same (cell, id) order as a stable sort of the cell keys
// listed in id order, and the cell ranges as a by-product:
//   emit_key        (in the kernels that hash positions) key_by_slot[t] and
//                   count[key] += run, where a run is a group of adjacent
//                   lanes of a wavefront with the same key: one atomic per
//                   run; in sorted order most of a cell is one run
//   k_cs_scan1      exclusive scan of the counts -> cell_start[0 .. ncell]
//                   in one launch (each workgroup publishes its tile's sum
//                   and adds up those of the tiles before its own); the
//                   two-launch form k_cs_tile_sum + k_cs_tile_scan serves the
//                   allocation-free entry point and the ghost numbering
//   k_cs_scatter    every particle takes a place in its cell range (count[key]
//                   -= run restores the counters to zero); the order inside a
//                   cell is arbitrary here
//   k_cs_rank       every particle counts the ids of its cell below its own:
//                   that rank is its place in (cell, id) order
// The counters must be zero before the keys are emitted (k_cs_zero) and are
// zero again after the scatter.  Grids of
// up to CS_TILE * CS_MAX_TILES cells are handled.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace arm {

constexpr int CS_THREADS = 256;
constexpr int CS_PER_LANE = 8;
constexpr int CS_TILE = CS_THREADS * CS_PER_LANE;  // cells per scan tile
constexpr int CS_MAX_TILES = 4096;

// Runs of equal keys in adjacent lanes of a wavefront.  All 64 lanes must
// call it.  head: first lane of its run; len: run length (valid on heads);
// hl: the head lane of this lane's run; rank: lane - hl.
struct WaveRun {
  bool head;
  int len, hl, rank;
};

__device__ __forceinline__ WaveRun wave_run(uint32_t c, bool valid) {
  const int lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(c, 1, 64);
  const bool head = valid && (lane == 0 || prev != c);
  const uint64_t hb = __ballot(head || !valid);
  const uint64_t after = lane == 63 ? 0ull : (hb >> (lane + 1));
  const int next = after ? lane + 1 + __builtin_ctzll(after) : 64;
  const uint64_t upto = lane == 63 ? hb : (hb & ((2ull << lane) - 1ull));
  WaveRun r;
  r.head = head;
  r.len = next - lane;
  r.hl = upto ? 63 - __builtin_clzll(upto) : lane;
  r.rank = lane - r.hl;
  return r;
}

// Where the hashing kernels put the keys: indexed by id for the radix sort,
// by slot (plus the per-cell counts) for the counting sort.
struct KeyOut {
  uint32_t* key_by_id;
  uint32_t* slot_by_id;
  uint32_t* key_by_slot;
  int* count;
};

// Stores the key of slot t (particle id); nothing when every pointer of ko
// is null.  All 64 lanes of the wavefront must call it; lanes with
// valid = false only take part in the ballots.
__device__ __forceinline__ void emit_key(const KeyOut& ko, bool valid, int t, int id,
                                         uint32_t key) {
  if (ko.key_by_slot) {
    const WaveRun r = wave_run(key, valid);
    if (valid) ko.key_by_slot[t] = key;
    if (r.head) atomicAdd(&ko.count[key], r.len);
  } else if (ko.key_by_id && valid) {
    ko.key_by_id[id] = key;
    ko.slot_by_id[id] = (uint32_t)t;
  }
}

// The fused steps' halt word (see KParams::halt): the sort of step its does
// nothing once an earlier launch recorded a failure at or before its.
struct Halt {
  const int* word = nullptr;
  int its = 0;
  __device__ __forceinline__ bool on() const {
    if (!word) return false;
    const int h = *word;
    return h != 0 && its >= h;
  }
};

// Clears the per-cell counters (a fill kernel, so that it may be captured).
__global__ __launch_bounds__(CS_THREADS) void k_cs_zero(int* __restrict__ count, int ncell) {
  for (int c = blockIdx.x * CS_THREADS + threadIdx.x; c < ncell; c += gridDim.x * CS_THREADS)
    count[c] = 0;
}

// Exclusive scan of one value per lane across the workgroup (LDS, one
// wavefront-sized partial per wavefront).  Returns the lane's exclusive
// prefix; *total receives the workgroup sum.
template <int THREADS>
__device__ __forceinline__ int block_exclusive_scan(int v, int* __restrict__ wsum, int* total) {
  constexpr int NW = THREADS / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    int s = lane < NW ? wsum[lane] : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(s, d, 64);
      if (lane >= d) s += y;
    }
    if (lane < NW) wsum[lane] = s;  // inclusive prefix of the wavefront sums
  }
  __syncthreads();
  const int before = w > 0 ? wsum[w - 1] : 0;
  *total = wsum[NW - 1];
  return before + x - v;
}

// The CS_PER_LANE consecutive counters of a lane (zero past n): two 16-byte
// loads where the run lies inside the array (p comes from hipMalloc and c0
// is a multiple of CS_PER_LANE, so the loads are aligned).
static_assert(CS_PER_LANE == 8, "load8/store8 move eight counters");
__device__ __forceinline__ void load8(const int* __restrict__ p, int c0, int n, int* v) {
  if (c0 + 8 <= n) {
    const int4 a = *reinterpret_cast<const int4*>(p + c0);
    const int4 b = *reinterpret_cast<const int4*>(p + c0 + 4);
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
    v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (c0 + k < n) ? p[c0 + k] : 0;
  }
}

__device__ __forceinline__ void store8(int* __restrict__ p, int c0, int n, const int* v) {
  if (c0 + 8 <= n) {
    *reinterpret_cast<int4*>(p + c0) = make_int4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<int4*>(p + c0 + 4) = make_int4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (c0 + k < n) p[c0 + k] = v[k];
  }
}

__global__ __launch_bounds__(CS_THREADS) void k_cs_tile_sum(const int* __restrict__ count,
                                                            int ncell, int* __restrict__ tsum,
                                                            Halt hl) {
  if (hl.on()) return;
  __shared__ int wsum[CS_THREADS / 64];
  const int c0 = blockIdx.x * CS_TILE + threadIdx.x * CS_PER_LANE;
  int v[CS_PER_LANE];
  load8(count, c0, ncell, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < CS_PER_LANE; ++k) s += v[k];
  int total;
  block_exclusive_scan<CS_THREADS>(s, wsum, &total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(CS_THREADS) void k_cs_tile_scan(const int* __restrict__ count,
                                                             int ncell, int n,
                                                             const int* __restrict__ tsum,
                                                             int* __restrict__ cstart, Halt hl) {
  if (hl.on()) return;
  __shared__ int wsum[CS_THREADS / 64];
  const int c0 = blockIdx.x * CS_TILE + threadIdx.x * CS_PER_LANE;
  int v[CS_PER_LANE];
  load8(count, c0, ncell, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < CS_PER_LANE; ++k) s += v[k];
  int part = 0;
  for (int j = threadIdx.x; j < (int)blockIdx.x; j += CS_THREADS) part += tsum[j];
  int base;
  block_exclusive_scan<CS_THREADS>(part, wsum, &base);
  __syncthreads();  // wsum is reused
  int total;
  int run = base + block_exclusive_scan<CS_THREADS>(s, wsum, &total);
  int o[CS_PER_LANE];
#pragma unroll
  for (int k = 0; k < CS_PER_LANE; ++k) {
    o[k] = run;
    run += v[k];
  }
  store8(cstart, c0, ncell, o);
  if (blockIdx.x == 0 && threadIdx.x == 0) cstart[ncell] = n;
}

// ---- single-pass scan -----------------------------------------------------
// k_cs_scan1 does the work of k_cs_tile_sum and k_cs_tile_scan in one launch:
// each workgroup publishes its tile's sum in a status word, then adds up the
// words of all the tiles before its own, read by all its lanes at once (the
// grid's workgroups run together, so the words are there within a few
// microseconds).  A word is (epoch << 32) | sum; a word of another epoch (an
// earlier sort) reads as not yet published.  The host gives every sort a new
// epoch (1 .. SCAN_EPOCH_MAX), so the words need no clearing between sorts;
// a launch replayed with the same epoch (a captured graph) would read the
// previous replay's words, so the allocation-free entry point, which a
// caller may capture, keeps the two-launch scan.
constexpr uint32_t SCAN_EPOCH_MAX = 0xffffffffu;  // the host clears the words and restarts at 1
constexpr int SCAN_SPIN_LIMIT = 1 << 14;

__global__ __launch_bounds__(CS_THREADS) void k_cs_scan1(const int* __restrict__ count, int ncell,
                                                         int n, int* __restrict__ cstart,
                                                         uint64_t* __restrict__ tstat,
                                                         uint32_t epoch, Halt hl) {
  if (hl.on()) return;
  __shared__ int wsum[CS_THREADS / 64];
  const int b = blockIdx.x;
  const int c0 = b * CS_TILE + threadIdx.x * CS_PER_LANE;
  int v[CS_PER_LANE];
  load8(count, c0, ncell, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < CS_PER_LANE; ++k) s += v[k];
  int total;
  const int pre = block_exclusive_scan<CS_THREADS>(s, wsum, &total);
  if (threadIdx.x == 0)
    __hip_atomic_store(&tstat[b], ((uint64_t)epoch << 32) | (uint32_t)total, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  // the sums of tiles 0 .. b-1; re-read while one is missing, and after
  // SCAN_SPIN_LIMIT rounds add up their counters instead (slower, the same)
  int part = 0;
  bool ready = false;
  for (int spins = 0; !ready && spins < SCAN_SPIN_LIMIT; ++spins) {
    part = 0;
    bool miss = false;
    for (int q = threadIdx.x; q < b; q += CS_THREADS) {
      const uint64_t w = __hip_atomic_load(&tstat[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      miss = miss || (uint32_t)(w >> 32) != epoch;
      part += (int)(uint32_t)w;
    }
    ready = !__syncthreads_or(miss);
  }
  if (!ready) {
    part = 0;
    for (int c = threadIdx.x; c < min(b * CS_TILE, ncell); c += CS_THREADS) part += count[c];
  }
  __syncthreads();  // wsum is reused
  int excl;
  block_exclusive_scan<CS_THREADS>(part, wsum, &excl);
  int run = excl + pre;
  int o[CS_PER_LANE];
#pragma unroll
  for (int k = 0; k < CS_PER_LANE; ++k) {
    o[k] = run;
    run += v[k];
  }
  store8(cstart, c0, ncell, o);
  if (b == 0 && threadIdx.x == 0) cstart[ncell] = n;
}

// The scattered entries: id and slot of the particle at each position of
// its cell range (two arrays, so that the rank's loop reads 4 bytes per
// entry); entc = its cell.
struct Ent {
  uint32_t* id;
  uint32_t* slot;
};

// Particles in slot order; with id_of null the particle in slot t has id t.
__global__ __launch_bounds__(CS_THREADS) void k_cs_scatter(const uint32_t* __restrict__ key_by_slot,
                                                           const int32_t* __restrict__ id_of,
                                                           int n, const int* __restrict__ cstart,
                                                           int* __restrict__ count, Ent ent,
                                                           uint32_t* __restrict__ entc, Halt hl) {
  if (hl.on()) return;
  const int t = blockIdx.x * CS_THREADS + threadIdx.x;
  const bool valid = t < n;
  const uint32_t c = valid ? key_by_slot[t] : 0xffffffffu;
  const WaveRun r = wave_run(c, valid);
  int old = 0;
  if (r.head) old = atomicSub(&count[c], r.len);
  old = __shfl(old, r.hl, 64);
  if (!valid) return;
  const int pos = cstart[c] + old - 1 - r.rank;
  ent.id[pos] = id_of ? (uint32_t)id_of[t] : (uint32_t)t;
  ent.slot[pos] = (uint32_t)t;
  entc[pos] = c;
}

// Pieces of the cell rows, for the gathers that give each workgroup a run of
// consecutive sorted positions of one row: the sorted range of row y,
// [cstart[y nx], cstart[(y+1) nx]), cut into runs of PIECE positions from its
// start (the last one shorter); pieces[s] = (first position, row), in row
// order, and *npieces their count, at most n / PIECE + ny.
constexpr int PIECE = 128;
struct RowPieces {
  int2* pieces = nullptr;
  int* npieces = nullptr;
  int nx = 0, ny = 0;
};

// Workgroup wg of a grid of ceil(ny / THREADS) lists the pieces of rows
// wg THREADS .. wg THREADS + THREADS-1, one lane per row.  It counts the
// pieces of the rows before its own itself (as k_cs_tile_scan adds up the
// tile sums before its own), so the table needs neither atomics nor a
// cleared counter; the workgroup of the last row writes *npieces.
template <int THREADS>
__device__ __forceinline__ void row_pieces_block(const int* __restrict__ cstart,
                                                 const RowPieces& rp, int wg) {
  __shared__ int wsum[THREADS / 64];
  const int nx = rp.nx;
  const int y0 = wg * THREADS;
  int before = 0;
  for (int y = threadIdx.x; y < y0; y += THREADS)
    before += (cstart[(y + 1) * nx] - cstart[y * nx] + PIECE - 1) / PIECE;
  int base;
  block_exclusive_scan<THREADS>(before, wsum, &base);
  __syncthreads();  // wsum is reused
  const int y = y0 + (int)threadIdx.x;
  const int a = y < rp.ny ? cstart[y * nx] : 0;
  const int b = y < rp.ny ? cstart[(y + 1) * nx] : 0;
  int total;
  int s = base + block_exclusive_scan<THREADS>((b - a + PIECE - 1) / PIECE, wsum, &total);
  for (int p = a; p < b; p += PIECE) rp.pieces[s++] = make_int2(p, y);
  if (threadIdx.x == 0 && y0 + THREADS >= rp.ny) *rp.npieces = base + total;
}

// skip: a cell whose order does not matter (the extra cell of stale ghosts,
// or -1): its entries keep their scatter positions.
__global__ __launch_bounds__(CS_THREADS) void k_cs_rank(Ent ent, const uint32_t* __restrict__ entc,
                                                        int n, const int* __restrict__ cstart,
                                                        uint32_t* __restrict__ skey,
                                                        uint32_t* __restrict__ sval, int skip,
                                                        Halt hl) {
  if (hl.on()) return;
  const int p = blockIdx.x * CS_THREADS + threadIdx.x;
  if (p >= n) return;
  const uint32_t c = entc[p];
  const uint32_t slot = ent.slot[p];
  if ((int)c == skip) {
    skey[p] = c;
    sval[p] = slot;
    return;
  }
  const int a = cstart[c], b = cstart[c + 1];
  const uint32_t id = ent.id[p];
  const uint32_t* __restrict__ ids = ent.id;
  int rank = 0;
  for (int q = a; q < b; ++q) rank += ids[q] < id ? 1 : 0;
  skey[a + rank] = c;
  sval[a + rank] = slot;
}

}  // namespace arm
