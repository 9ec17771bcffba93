// sph3d_tiled.h -- LDS-tiled gathers of the 3D density and force passes.
//
// A workgroup owns a segment of CW consecutive cells of one x-row of cells
// (fixed cy, cz).  Its particles are one contiguous sorted range, and the
// candidates of all of them lie in nine contiguous sorted ranges: the rows
// (cz + dz, cy + dy), dz, dy in -1..1, cells cx0-1 .. cx0+CW.  The workgroup
// copies those nine ranges, and the cell_start entries that delimit their
// cells, into LDS with coalesced loads.  Each lane then handles one
// particle, row by row (z outer, y inner, as the direct path):
//   A  test the row's candidates eight at a time from LDS, with no
//      per-candidate branch, into 32-bit neighbour masks;
//   B  run the pair arithmetic over the set bits in ascending order.
// The density pass records the masks (Mask3); the force pass visits only
// the recorded neighbours.  Neighbour order and arithmetic are those of the
// direct path, so both paths give identical bits.
//
// In the FP32 form a record keeps the cell column as a float, and the cell
// differences in y and z of row k are the constants 1 - k % 3 and 1 - k / 3:
//   dx = (off_i.x - off_j.x) + (cx_i - cx_j) cell
//   dy = (off_i.y - off_j.y) + (1 - k % 3) cell
//   dz = (off_i.z - off_j.z) + (1 - k / 3) cell
// are the operations, with the same roundings, of delta3() on Pos3F.
//
// The kernels are persistent: a list of the segments that hold particles is
// built once per sort, and the workgroups walk it XCD by XCD (the hardware
// hands consecutive workgroups to the eight XCDs in turn), each XCD taking
// one contiguous eighth of the list, so that the y-neighbour rows of a
// segment tend to sit in that XCD's L2 already.
#pragma once

#include "sph3d.h"
#include "sph_tiled.h"

namespace arm {

constexpr int NXCD = 8;

struct Lane3F {
  float cxf, ox, oy, oz;
};

template <typename R, int CAP> struct Dens3Tile;
template <int CAP> struct Dens3Tile<float, CAP> {
  float4 p[CAP + TPAD];  // offx, offy, offz, cell column as float
  float4 v[CAP + TPAD];  // vx, vy, vz, -
  __device__ __forceinline__ void put(int l, const St3<float>& S, const KParams&, int t) {
    const float4 q = S.pf[t];
    const uint32_t c = __float_as_uint(q.w);
    p[l] = make_float4(q.x, q.y, q.z, (float)(c & ((1u << C3_BITS_X) - 1u)));
    v[l] = S.v[t];
  }
  __device__ __forceinline__ Lane3F lane(int l) const {
    const float4 q = p[l];
    return {q.w, q.x, q.y, q.z};
  }
  __device__ __forceinline__ float delta(const KParams& K, const Lane3F& a, int l, float uy,
                                         float uz, float& dx, float& dy, float& dz) const {
    const float4 q = p[l];
    dx = (a.ox - q.x) + (a.cxf - q.w) * K.cellf;
    dy = (a.oy - q.y) + uy;
    dz = (a.oz - q.z) + uz;
    float r2 = dx * dx;
    r2 = r2 + dy * dy;
    return r2 + dz * dz;
  }
  __device__ __forceinline__ float4 vel(int l) const { return v[l]; }
};
template <int CAP> struct Dens3Tile<double, CAP> {
  double4 x[CAP + TPAD], v[CAP + TPAD];
  __device__ __forceinline__ void put(int l, const St3<double>& S, const KParams&, int t) {
    x[l] = S.xyz[t];
    v[l] = S.v[t];
  }
  __device__ __forceinline__ Pos3D lane(int l) const {
    const double4 q = x[l];
    return {q.x, q.y, q.z};
  }
  __device__ __forceinline__ double delta(const KParams&, const Pos3D& a, int l, double, double,
                                          double& dx, double& dy, double& dz) const {
    const double4 q = x[l];
    dx = a.x - q.x;
    dy = a.y - q.y;
    dz = a.z - q.z;
    double r2 = dx * dx;
    r2 = r2 + dy * dy;
    return r2 + dz * dz;
  }
  __device__ __forceinline__ double4 vel(int l) const { return v[l]; }
};

template <typename R, int CAP> struct Force3Tile;
template <int CAP> struct Force3Tile<float, CAP> {
  Dens3Tile<float, CAP> d;
  float4 b[CAP + TPAD];  // Q = 1/rho, A = p/rho, psi, u
  __device__ __forceinline__ void put(int l, const St3<float>& S, const KParams& K, int t) {
    d.put(l, S, K, t);
    float Q, A;
    q_and_a(K, S.rho[t], S.p[t], Q, A);
    b[l] = make_float4(Q, A, S.psi[t], S.uf[t]);
  }
  __device__ __forceinline__ Lane3F lane(int l) const { return d.lane(l); }
  __device__ __forceinline__ float delta(const KParams& K, const Lane3F& a, int l, float uy,
                                         float uz, float& dx, float& dy, float& dz) const {
    return d.delta(K, a, l, uy, uz, dx, dy, dz);
  }
  template <bool PRESS, bool AV, bool AH, bool XS>
  __device__ __forceinline__ void pair(const KParams& K, ForceOwn3<float>& o, int l, float dx,
                                       float dy, float dz, float r2) const {
    const float4 vj = d.v[l];
    const float4 w = b[l];
    force3_pair<float, PRESS, AV, AH, XS>(K, o, dx, dy, dz, r2, vj.x, vj.y, vj.z, w.x, w.y, w.z,
                                          w.w);
  }
};
template <int CAP> struct Force3Tile<double, CAP> {
  Dens3Tile<double, CAP> d;
  double4 b[CAP + TPAD];
  __device__ __forceinline__ void put(int l, const St3<double>& S, const KParams& K, int t) {
    d.put(l, S, K, t);
    double Q, A;
    q_and_a(K, S.rho[t], S.p[t], Q, A);
    b[l] = make_double4(Q, A, S.psi[t], S.u[t]);
  }
  __device__ __forceinline__ Pos3D lane(int l) const { return d.lane(l); }
  __device__ __forceinline__ double delta(const KParams& K, const Pos3D& a, int l, double uy,
                                          double uz, double& dx, double& dy, double& dz) const {
    return d.delta(K, a, l, uy, uz, dx, dy, dz);
  }
  template <bool PRESS, bool AV, bool AH, bool XS>
  __device__ __forceinline__ void pair(const KParams& K, ForceOwn3<double>& o, int l, double dx,
                                       double dy, double dz, double r2) const {
    const double4 vj = d.v[l];
    const double4 w = b[l];
    force3_pair<double, PRESS, AV, AH, XS>(K, o, dx, dy, dz, r2, vj.x, vj.y, vj.z, w.x, w.y,
                                           w.z, w.w);
  }
};

// ---- segment geometry ------------------------------------------------------
// Per row k: first sorted position rs, length rl, LDS offset lo; cs[k][j] =
// cell_start of cell slo + j of row k, j = 0 .. CW + 2.  Kept in LDS so that
// the row loop may index them.
template <int CW> struct Seg3Lds {
  int cs[9][CW + 3];
  int rs[9], rl[9], lo[10];
};

struct Seg3 {
  int row, slo, O0, O1, total;
};

// Segment s of CW cells: s = row * nseg + x-segment, row = cz * ny + cy.
__device__ __forceinline__ void seg3_range(const KParams& K, int s, int nseg, int cw, int& row,
                                           int& cx0, int& cx1) {
  row = s / nseg;
  cx0 = (s - row * nseg) * cw;
  cx1 = min(cx0 + cw, K.nx);
}

// Fills L for segment s (all lanes take part; the caller synchronises).
template <int CW>
__device__ __forceinline__ Seg3 seg3_of(const KParams& K, const int* __restrict__ cstart, int nseg,
                                        int s, Seg3Lds<CW>& L) {
  Seg3 g;
  int cx0, cx1;
  seg3_range(K, s, nseg, CW, g.row, cx0, cx1);
  const int nx = K.nx, ny = K.ny;
  const int cz = g.row / ny, cy = g.row - cz * ny;
  g.O0 = cstart[g.row * nx + cx0];
  g.O1 = cstart[g.row * nx + cx1];
  g.slo = cx0 > 0 ? cx0 - 1 : 0;
  const int shi = cx1 < nx ? cx1 : nx - 1;
  const int ncs = shi + 2 - g.slo;
  int rs[9], rl[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int rz = cz - 1 + k / 3, ry = cy - 1 + k % 3;
    rs[k] = 0;
    rl[k] = 0;
    if (rz >= 0 && rz < K.nz && ry >= 0 && ry < ny) {
      const int base = (rz * ny + ry) * nx + g.slo;
      rs[k] = cstart[base];
      rl[k] = cstart[base + ncs - 1] - rs[k];
    }
  }
  for (int j = threadIdx.x; j < 9 * ncs; j += blockDim.x) {
    const int k = j / ncs, jj = j - k * ncs;
    const int rz = cz - 1 + k / 3, ry = cy - 1 + k % 3;
    if (rz >= 0 && rz < K.nz && ry >= 0 && ry < ny)
      L.cs[k][jj] = cstart[(rz * ny + ry) * nx + g.slo + jj];
  }
  g.total = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (threadIdx.x == 0) {
      L.rs[k] = rs[k];
      L.rl[k] = rl[k];
      L.lo[k] = g.total;
    }
    g.total += rl[k];
  }
  if (threadIdx.x == 0) L.lo[9] = g.total;
  return g;
}

// Copies the nine candidate ranges into the tile, one flat loop over the
// records so that the loads of consecutive iterations are independent.
template <int BLK, int CW, typename Tile, typename State>
__device__ __forceinline__ void stage3(Tile& T, const State& S, const KParams& K,
                                       const Seg3Lds<CW>& L, int total) {
#pragma unroll 4
  for (int q = threadIdx.x; q < total; q += BLK) {
    int k = 0;
#pragma unroll
    for (int r = 1; r < 9; ++r) k += q >= L.lo[r] ? 1 : 0;
    T.put(q, S, K, L.rs[k] + (q - L.lo[k]));
  }
}

// Segments with particles, listed once per sort: list[0 .. *count).  The
// order of the list does not matter (segments write disjoint outputs).
__global__ __launch_bounds__(256) void k3_active_segments(KParams K,
                                                          const int* __restrict__ cstart,
                                                          int nseg, int nblocks, int cw,
                                                          int* __restrict__ list,
                                                          int* __restrict__ count) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  bool act = false;
  if (s < nblocks) {
    int row, cx0, cx1;
    seg3_range(K, s, nseg, cw, row, cx0, cx1);
    act = cstart[row * K.nx + cx0] < cstart[row * K.nx + cx1];
  }
  const uint64_t m = __ballot(act);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(m);
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __builtin_popcountll(m));
  base = __shfl(base, leader, 64);
  if (act) list[base + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = s;
}

// Persistent walk over the listed segments: the blocks of XCD x (block
// index % 8 == x) take the x-th eighth of the list, so each XCD walks one
// contiguous stretch of rows.  Every block leaves the loop once its share
// is done.
struct SegWalk {
  int e, end, step;
};

__device__ __forceinline__ SegWalk seg_walk(int count) {
  const int per_xcd = (count + NXCD - 1) / NXCD;
  const int x = (int)(blockIdx.x % NXCD);
  SegWalk w;
  w.e = x * per_xcd + (int)(blockIdx.x / NXCD);
  w.end = min(count, (x + 1) * per_xcd);
  w.step = (int)gridDim.x / NXCD;
  return w;
}

// Neighbour mask of LDS records l0 .. l0+len-1 (len <= 32): eight
// independent reads and tests per batch; bits past len and the lane's own
// bit are cleared afterwards.
template <typename R, typename Tile, typename LP>
__device__ __forceinline__ uint32_t row_mask3(const KParams& K, const Tile& T, const LP& pi, R uy,
                                              R uz, int l0, int len, int self) {
  const R cut2 = cutoff2<R>(K);
  uint32_t m = 0;
  for (int k0 = 0; k0 < len; k0 += 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      R dx, dy, dz;
      const R r2 = T.delta(K, pi, l0 + k0 + q, uy, uz, dx, dy, dz);
      m |= (r2 < cut2 ? 1u : 0u) << (k0 + q);
    }
  }
  m &= len >= 32 ? 0xffffffffu : ((1u << len) - 1u);
  const int sb = self - l0;
  if (sb >= 0 && sb < 32) m &= ~(1u << sb);
  return m;
}

template <typename R> __device__ __forceinline__ R row_u(const KParams& K, int d) {
  return (R)((float)d * K.cellf);
}
template <> __device__ __forceinline__ double row_u<double>(const KParams&, int) { return 0.0; }

// Calls visit(l, uy, uz) for every neighbour of lane particle i, rows in the
// direct path's order.  With rec, records the masks of particle i.
template <typename R, int CW, typename Tile, typename LP, typename Visit>
__device__ __forceinline__ void for_neighbours3(const KParams& K, const Seg3& g,
                                                const Seg3Lds<CW>& L, const Tile& T,
                                                const LP& pi, int i, uint32_t key, Visit&& visit,
                                                Mask3 rec) {
  const int nx = K.nx;
  const int cx = (int)key - g.row * nx;
  const int a0 = (cx > 0 ? cx - 1 : 0) - g.slo;
  const int a1 = (cx + 1 < nx ? cx + 1 : nx - 1) - g.slo + 1;
  const int self = L.lo[4] + (i - L.rs[4]);
  bool ok = true;
#pragma unroll 1
  for (int k = 0; k < 9; ++k) {
    const int base = L.lo[k] - L.rs[k];
    const R uy = row_u<R>(K, 1 - k % 3), uz = row_u<R>(K, 1 - k / 3);
    const int ta = L.rl[k] ? L.cs[k][a0] : 0;
    const int tb = L.rl[k] ? L.cs[k][a1] : 0;
    if (tb - ta > 32 * M3_WORDS) ok = false;
    uint32_t w[M3_WORDS];
#pragma unroll
    for (int q = 0; q < M3_WORDS; ++q) {
      const int c0 = ta + 32 * q;
      uint32_t m = c0 < tb ? row_mask3<R>(K, T, pi, uy, uz, base + c0, min(32, tb - c0), self)
                           : 0u;
      w[q] = m;
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        visit(base + c0 + b, uy, uz);
      }
    }
    for (int c0 = ta + 32 * M3_WORDS; c0 < tb; c0 += 32) {
      uint32_t m = row_mask3<R>(K, T, pi, uy, uz, base + c0, min(32, tb - c0), self);
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        visit(base + c0 + b, uy, uz);
      }
    }
    if (rec.w)
#pragma unroll
      for (int q = 0; q < M3_WORDS; ++q) rec.put(k * M3_WORDS + q, i, w[q]);
  }
  if (rec.w) rec.put(M3_OK, i, ok ? 1u : 0u);
}

// The same walk driven by the recorded masks.
template <typename R, int CW, typename Visit>
__device__ __forceinline__ void for_masked3(const KParams& K, const Seg3& g, const Seg3Lds<CW>& L,
                                            int i, uint32_t key, Mask3 nbm, Visit&& visit) {
  const int nx = K.nx;
  const int cx = (int)key - g.row * nx;
  const int a0 = (cx > 0 ? cx - 1 : 0) - g.slo;
  uint32_t mw[M3_OK];
#pragma unroll
  for (int q = 0; q < M3_OK; ++q) mw[q] = nbm.get(q, i);  // all loads in flight at once
#pragma unroll 1
  for (int k = 0; k < 9; ++k) {
    if (L.rl[k] == 0) continue;
    const R uy = row_u<R>(K, 1 - k % 3), uz = row_u<R>(K, 1 - k / 3);
    const int l0 = L.lo[k] - L.rs[k] + L.cs[k][a0];
#pragma unroll 1
    for (int q = 0; q < M3_WORDS; ++q) {
      uint32_t m = mw[k * M3_WORDS + q];
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        visit(l0 + 32 * q + b, uy, uz);
      }
    }
  }
}

template <typename R, bool SUM, bool EOS, int BLK, int CW, int CAP>
__global__ __launch_bounds__(BLK) void k3_density_tiled(St3<R> S, KParams K,
                                                        const uint32_t* __restrict__ skey,
                                                        const int* __restrict__ cstart, int nseg,
                                                        const int* __restrict__ list,
                                                        const int* __restrict__ count,
                                                        Mask3 nbm) {
  __shared__ Dens3Tile<R, CAP> T;
  __shared__ Seg3Lds<CW> L;
  const SegWalk w = seg_walk(*count);
  for (int e = w.e; e < w.end; e += w.step) {
    __syncthreads();  // the previous segment is done with L and T
    const Seg3 g = seg3_of<CW>(K, cstart, nseg, list[e], L);
    const bool tiled = g.total <= CAP;
    __syncthreads();
    if (tiled) stage3<BLK, CW>(T, S, K, L, g.total);
    __syncthreads();
    for (int i0 = g.O0; i0 < g.O1; i0 += BLK) {
      const int i = i0 + threadIdx.x;
      if (i >= g.O1) continue;
      if (!tiled) {
        density3_direct<R, SUM, EOS>(S, K, i, skey, cstart, nbm);
        continue;
      }
      const int self = L.lo[4] + (i - L.rs[4]);
      const auto pi = T.lane(self);
      const auto vi = T.vel(self);
      DensOwn<R> o;
      dens_begin<R, SUM>(K, o);
      for_neighbours3<R, CW>(
          K, g, L, T, pi, i, skey[i],
          [&](int l, R uy, R uz) {
            R dx, dy, dz;
            const R r2 = T.delta(K, pi, l, uy, uz, dx, dy, dz);
            const auto vj = T.vel(l);
            dens3_pair<R, SUM>(K, dx, dy, dz, r2, vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, o.sum,
                               o.rs);
          },
          nbm);
      dens_end<R, SUM, EOS>(S, K, i, o);
    }
  }
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS, int BLK, int CW, int CAP>
__global__ __launch_bounds__(BLK) void k3_force_tiled(St3<R> S, KParams K,
                                                      const uint32_t* __restrict__ skey,
                                                      const int* __restrict__ cstart, int nseg,
                                                      const int* __restrict__ list,
                                                      const int* __restrict__ count, Mask3 nbm) {
  __shared__ Force3Tile<R, CAP> T;
  __shared__ Seg3Lds<CW> L;
  const SegWalk w = seg_walk(*count);
  for (int e = w.e; e < w.end; e += w.step) {
    __syncthreads();
    const Seg3 g = seg3_of<CW>(K, cstart, nseg, list[e], L);
    const bool tiled = g.total <= CAP;
    __syncthreads();
    if (tiled) stage3<BLK, CW>(T, S, K, L, g.total);
    __syncthreads();
    for (int i0 = g.O0; i0 < g.O1; i0 += BLK) {
      const int i = i0 + threadIdx.x;
      if (i >= g.O1) continue;
      if (!tiled) {
        force3_direct<R, PRESS, AV, AH, XS>(S, K, i, skey, cstart, nbm);
        continue;
      }
      const int self = L.lo[4] + (i - L.rs[4]);
      const auto pi = T.lane(self);
      ForceOwn3<R> o;
      force3_begin(S, K, i, o);
      auto visit = [&](int l, R uy, R uz) {
        R dx, dy, dz;
        const R r2 = T.delta(K, pi, l, uy, uz, dx, dy, dz);
        T.template pair<PRESS, AV, AH, XS>(K, o, l, dx, dy, dz, r2);
      };
      const uint32_t key = skey[i];
      if (nbm.w && nbm.get(M3_OK, i))
        for_masked3<R, CW>(K, g, L, i, key, nbm, visit);
      else
        for_neighbours3<R, CW>(K, g, L, T, pi, i, key, visit, Mask3{nullptr, 0});
      force3_end<R, PRESS, XS>(S, K, i, o);
    }
  }
}

// Launch shapes: cells per segment, lanes, LDS record capacity.
template <typename R> struct Tile3Cfg;
template <> struct Tile3Cfg<float> {
  static constexpr int BLK = 128, CW = 6, CAP = 1460;
};
template <> struct Tile3Cfg<double> {
  static constexpr int BLK = 128, CW = 6, CAP = 1460;
};

}  // namespace arm
