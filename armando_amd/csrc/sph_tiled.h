// sph_tiled.h -- LDS-tiled gathers of the density and force passes.
//
// A workgroup owns a segment of one cell row: a piece of up to 128
// consecutive sorted particles of the row (or, for the viscous force
// gathers, a run of CW cells).  Its particles are one contiguous sorted
// range, and the candidates of all of them lie in three contiguous sorted
// ranges (the cell row below, the row itself and the row above, over the
// segment's cells and one either side).  The workgroup copies those three
// ranges into LDS with coalesced loads.  Each lane then handles one
// particle, row by row:
//   A  test the row's candidates (r^2 < (2h)^2) eight at a time, reading
//      their positions from LDS with independent loads, into a 32-bit row
//      mask;
//   B  run the pair arithmetic over the set bits in ascending order, so the
//      wavefront does the expensive math only for real neighbours (one walk
//      over the three rows' masks where they are complete).
// The visiting order (rows bottom to top, ascending sorted position) and the
// arithmetic are those of the direct path, so both give identical bits.
// Rows longer than 32 candidates are handled in 32-candidate chunks (their
// masks are not recorded); when the three ranges exceed the LDS capacity the
// workgroup falls back to the direct path.
#pragma once

#include "sph_kernels.h"

namespace arm {

// ---- LDS tiles ------------------------------------------------------------
// Every tile has TPAD spare records after CAP so that the eight-wide
// neighbour test may read up to seven records past a row's end (their bits
// are masked off).  In the FP32 form a record keeps the cell column as a
// float, and the cell-row difference of a candidate row is the constant
// 1 - k of row k of the segment, so
//   dx = (off_i.x - off_j.x) + (cx_i - cx_j) cell
//   dy = (off_i.y - off_j.y) + (1 - k) cell
// are the same operations, with the same roundings, as delta() on PosF.
constexpr int TPAD = 8;

// Byte offset 4 l of a 4-byte LDS record.  Computed out of the compiler's
// sight: otherwise it derives this address from the 16-byte records' l << 4
// with 64-bit integer multiply-adds, two of them per neighbour.
__device__ __forceinline__ uint32_t off4(int l) {
  uint32_t r;
  asm("v_lshlrev_b32 %0, 2, %1" : "=v"(r) : "v"(l));
  return r;
}

template <typename T> __device__ __forceinline__ T lds_at(const T* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Lane particle for tile deltas (FP32 form).
struct LaneF {
  float cxf, ox, oy;
};

// The cell-row term of row k: (float)(cy_i - cy_j) * cell with cy_i - cy_j = 1 - k.
__device__ __forceinline__ float row_uy(const KParams& K, int k) {
  return (float)(1 - k) * K.cellf;
}

__device__ __forceinline__ void tile_delta(const KParams& K, const LaneF& p, float cxj, float oxj,
                                           float oyj, float uy, float& dx, float& dy) {
  const float tx = p.ox - oxj;
  const float ux = (p.cxf - cxj) * K.cellf;
  dx = tx + ux;
  const float ty = p.oy - oyj;
  dy = ty + uy;
}

template <typename R, int CAP> struct DensTile;
template <int CAP> struct DensTile<float, CAP> {
  // the neighbour test reads one 16-byte record per candidate
  float4 a[CAP + TPAD];  // offx, offy, cell column as float, vx
  float vy[CAP + TPAD];
  __device__ __forceinline__ void put(int l, const St<float>& S, const KParams&, int t) {
    const float2 o = S.off[t];
    const float2 v = S.v[t];
    a[l] = make_float4(o.x, o.y, (float)(S.cxy[t] & 0xffffu), v.x);
    vy[l] = v.y;
  }
  __device__ __forceinline__ LaneF lane(int l) const {
    const float4 q = a[l];
    return {q.z, q.x, q.y};
  }
  __device__ __forceinline__ void delta(const KParams& K, const LaneF& p, int l, float uy,
                                        float& dx, float& dy) const {
    const float4 q = a[l];
    tile_delta(K, p, q.z, q.x, q.y, uy, dx, dy);
  }
  __device__ __forceinline__ float2 vel(int l) const {
    return make_float2(a[l].w, lds_at(vy, off4(l)));
  }
};
template <int CAP> struct DensTile<double, CAP> {
  double2 x[CAP + TPAD], v[CAP + TPAD];
  __device__ __forceinline__ void put(int l, const St<double>& S, const KParams&, int t) {
    x[l] = S.xy[t];
    v[l] = S.v[t];
  }
  __device__ __forceinline__ PosD lane(int l) const {
    const double2 q = x[l];
    return {q.x, q.y};
  }
  __device__ __forceinline__ void delta(const KParams&, const PosD& p, int l, double, double& dx,
                                        double& dy) const {
    const double2 q = x[l];
    dx = p.x - q.x;
    dy = p.y - q.y;
  }
  __device__ __forceinline__ double2 vel(int l) const { return v[l]; }
};

// Neighbour W = c T of the physical viscosity, staged only when VI.
template <typename R, int CAP, bool VI> struct TauTile {
  __device__ __forceinline__ void put(int, const St<R>&, int, R) {}
  __device__ __forceinline__ Tau3<R> get(int) const { return Tau3<R>{}; }
};
template <int CAP> struct TauTile<float, CAP, true> {
  float4 w[CAP];
  __device__ __forceinline__ void put(int l, const St<float>& S, int t, float c) {
    const float4 q = reinterpret_cast<const float4*>(S.tau)[t];
    w[l] = make_float4(c * q.x, c * q.y, c * q.z, q.w);
  }
  __device__ __forceinline__ Tau3<float> get(int l) const {
    const float4 q = w[l];
    return {q.x, q.y, q.z};
  }
};
template <int CAP> struct TauTile<double, CAP, true> {
  double2 w01[CAP];
  double w2[CAP];
  __device__ __forceinline__ void put(int l, const St<double>& S, int t, double c) {
    const double* q = S.tau + 4 * (size_t)t;
    w01[l] = make_double2(c * q[0], c * q[1]);
    w2[l] = c * q[2];
  }
  __device__ __forceinline__ Tau3<double> get(int l) const {
    const double2 q = w01[l];
    return {q.x, q.y, w2[l]};
  }
};

template <typename R, int CAP, bool VI = false> struct ForceTile;
template <int CAP, bool VI> struct ForceTile<float, CAP, VI> {
  float4 a[CAP + TPAD];  // offx, offy, vx, vy
  float4 b[CAP + TPAD];  // Q = 1/rho, A = p/rho, psi, u
  float cx[CAP + TPAD];  // cell column as float
  TauTile<float, CAP + TPAD, VI> tw;
  __device__ __forceinline__ void put(int l, const St<float>& S, const KParams& K, int t, int) {
    const float2 o = S.off[t];
    const float2 v = S.v[t];
    a[l] = make_float4(o.x, o.y, v.x, v.y);
    float Q, A;
    q_and_a(K, S.rho[t], S.p[t], Q, A);
    b[l] = make_float4(Q, A, S.psi[t], S.uf[t]);
    cx[l] = (float)(S.cxy[t] & 0xffffu);
    tw.put(l, S, t, visc_c(K, Q));
  }
  __device__ __forceinline__ LaneF lane(int l) const {
    const float4 q = a[l];
    return {lds_at(cx, off4(l)), q.x, q.y};
  }
  __device__ __forceinline__ void delta(const KParams& K, const LaneF& p, int l, float uy,
                                        float& dx, float& dy) const {
    const float4 q = a[l];
    tile_delta(K, p, lds_at(cx, off4(l)), q.x, q.y, uy, dx, dy);
  }
  template <bool PRESS, bool AV, bool AH, bool XS, int PC = -1>
  __device__ __forceinline__ void pair(const KParams& K, ForceOwn<float>& o, int l, float dx,
                                       float dy, float r2) const {
    const float4 q = a[l];
    const float4 w = b[l];
    force_pair<float, PRESS, AV, AH, XS, VI, PC>(K, o, dx, dy, r2, q.z, q.w, w.x, w.y, w.z,
                                                 w.w, tw.get(l));
  }
};
template <int CAP, bool VI> struct ForceTile<double, CAP, VI> {
  double2 x[CAP + TPAD], v[CAP + TPAD], rp[CAP + TPAD], pu[CAP + TPAD];  // rp = (Q, A), pu = (psi, u)
  TauTile<double, CAP + TPAD, VI> tw;
  __device__ __forceinline__ void put(int l, const St<double>& S, const KParams& K, int t, int) {
    x[l] = S.xy[t];
    v[l] = S.v[t];
    double Q, A;
    q_and_a(K, S.rho[t], S.p[t], Q, A);
    rp[l] = make_double2(Q, A);
    pu[l] = make_double2(S.psi[t], S.u[t]);
    tw.put(l, S, t, visc_c(K, Q));
  }
  __device__ __forceinline__ PosD lane(int l) const {
    const double2 q = x[l];
    return {q.x, q.y};
  }
  __device__ __forceinline__ void delta(const KParams&, const PosD& p, int l, double, double& dx,
                                        double& dy) const {
    const double2 q = x[l];
    dx = p.x - q.x;
    dy = p.y - q.y;
  }
  template <bool PRESS, bool AV, bool AH, bool XS, int PC = -1>
  __device__ __forceinline__ void pair(const KParams& K, ForceOwn<double>& o, int l, double dx,
                                       double dy, double r2) const {
    const double2 vv = v[l];
    const double2 a = rp[l];
    const double2 b = pu[l];
    force_pair<double, PRESS, AV, AH, XS, VI>(K, o, dx, dy, r2, vv.x, vv.y, a.x, a.y, b.x, b.y,
                                              tw.get(l));
  }
};

// Positions, velocities and Q = 1/rho for the shear tensor pass.
template <typename R, int CAP> struct TensTile {
  DensTile<R, CAP> d;
  R q[CAP + TPAD];
  __device__ __forceinline__ void put(int l, const St<R>& S, const KParams& K, int t) {
    d.put(l, S, K, t);
    q[l] = (R)1 / dens(K, S.rho[t]);
  }
  __device__ __forceinline__ auto lane(int l) const { return d.lane(l); }
  template <typename LP>
  __device__ __forceinline__ void delta(const KParams& K, const LP& p, int l, R uy, R& dx,
                                        R& dy) const {
    d.delta(K, p, l, uy, dx, dy);
  }
};

// ---- segment geometry ------------------------------------------------------
// A segment: the particles [O0, O1) of cell row cy, which lie in cells
// slo .. shi of that row; the tile holds the rows cy-1 .. cy+1 over those
// cells (rs[k], rl[k]: sorted start and length of row k's range, lo[k]: its
// first tile entry).
struct Segment {
  int cy, slo, O0, O1;
  int rs[3], rl[3], lo[3];
  int total;
};

// How a workgroup finds its segment.
//  CW > 0: a fixed run of CW cells of a row (segment row * nseg + run).
//  CW = 0: a piece of at most PIECE consecutive sorted positions of one row
//          (RowPieces); every wavefront of a piece but the row's last is
//          full.
struct SegArg {
  int nseg;
  const int2* pieces;
  const int* npieces;
};

// Width of the LDS copy of the cell starts (none for pieces, whose cells are
// looked up in global memory: their span is not bounded).
template <int CW> constexpr int csw() { return CW > 0 ? CW + 3 : 1; }

// cell_start of cell c of tile row k (grid row cy-1+k): from the segment's
// LDS copy (CW > 0) or from global memory (CW = 0).
template <int CW> struct CellBounds {
  const int (*cs)[csw<CW>()];
  int slo;
  __device__ __forceinline__ int at(int k, int c) const { return cs[k][c - slo]; }
};
template <> struct CellBounds<0> {
  const int* __restrict__ cstart;
  int base;  // (cy - 1) nx
  int nx;
  __device__ __forceinline__ int at(int k, int c) const { return cstart[base + k * nx + c]; }
};

template <int CW>
__device__ __forceinline__ CellBounds<CW> bounds_of(const KParams& K, const Segment& g,
                                                    const int* __restrict__ cstart,
                                                    const int (*cs)[csw<CW>()]) {
  if constexpr (CW > 0)
    return CellBounds<CW>{cs, g.slo};
  else
    return CellBounds<0>{cstart, (g.cy - 1) * K.nx, K.nx};
}

// The segment of this workgroup; O0 >= O1 when it has no particles.  With
// CW > 0 the workgroup also copies the cell starts of its rows into cs.
template <int CW>
__device__ __forceinline__ Segment segment_of(const KParams& K, const int* __restrict__ cstart,
                                              const uint32_t* __restrict__ skey,
                                              const SegArg& sa, int (*cs)[csw<CW>()]) {
  Segment g;
  const int nx = K.nx;
  int cx0, cx1;  // the particles' cells: cx0 .. cx1-1
  if constexpr (CW > 0) {
    // plain order: most cell runs of a tank are empty, and contiguous
    // shares per XCD would leave the XCDs of the empty rows idle
    g.cy = blockIdx.x / sa.nseg;
    cx0 = (blockIdx.x - g.cy * sa.nseg) * CW;
    cx1 = min(cx0 + CW, nx);
    g.O0 = cstart[g.cy * nx + cx0];
    g.O1 = cstart[g.cy * nx + cx1];
  } else {
    // pieces in even XCD-contiguous shares: neighbouring pieces, which stage
    // the same rows, share an L2 (the grid is a bound on the piece count)
    const int sb = xcd_share(*sa.npieces);
    if (sb < 0) {
      g.O0 = g.O1 = 0;
      return g;
    }
    const int2 pc = sa.pieces[sb];
    g.cy = pc.y;
    g.O0 = pc.x;
    g.O1 = min(pc.x + PIECE, cstart[(pc.y + 1) * nx]);
    cx0 = (int)skey[g.O0] - g.cy * nx;
    cx1 = (int)skey[g.O1 - 1] - g.cy * nx + 1;
  }
  g.slo = cx0 > 0 ? cx0 - 1 : 0;
  const int shi = cx1 < nx ? cx1 : nx - 1;
  g.total = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ry = g.cy - 1 + k;
    if (ry < 0 || ry >= K.ny || g.O0 >= g.O1) {
      g.rs[k] = 0;
      g.rl[k] = 0;
    } else {
      g.rs[k] = cstart[ry * nx + g.slo];
      g.rl[k] = cstart[ry * nx + shi + 1] - g.rs[k];
      if constexpr (CW > 0) {
        const int ncs = shi + 2 - g.slo;
        for (int j = threadIdx.x; j < ncs; j += blockDim.x) cs[k][j] = cstart[ry * nx + g.slo + j];
      }
    }
    g.lo[k] = g.total;
    g.total += g.rl[k];
  }
  return g;
}

// The row pieces on their own (after a sort that does not build them).
__global__ __launch_bounds__(BLOCK) void k_row_pieces(const int* __restrict__ cstart,
                                                      RowPieces rp, Halt hl) {
  if (hl.on()) return;
  row_pieces_block<BLOCK>(cstart, rp, blockIdx.x);
}

// A particle of the first or last cell column of a grid periodic in x: its
// neighbours wrap to the other end, outside the segment's tile, so the
// direct path (which walks the wrapped cells) takes it.  PER = false: the
// kernel was launched for a grid without periodic x.
template <bool PER>
__device__ __forceinline__ bool seam(const KParams& K, const Segment& g, uint32_t key) {
  if (!PER) return false;
  const int cx = (int)key - g.cy * K.nx;
  return cx == 0 || cx == K.nx - 1;
}

// Neighbour mask of candidates l0 .. l0+len-1 (len <= 32) of row k: eight
// independent tile reads and tests per batch, no per-candidate branch; the
// bits past len and the lane's own bit are cleared afterwards.
template <typename R, typename Tile, typename LP>
__device__ __forceinline__ uint32_t row_mask(const KParams& K, const Tile& T, const LP& pi, R uy,
                                             int l0, int len, int self) {
  const R cut2 = cutoff2<R>(K);
  uint32_t m = 0;
  for (int k0 = 0; k0 < len; k0 += 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      R dx, dy;
      T.delta(K, pi, l0 + k0 + q, uy, dx, dy);
      const R r2 = dx * dx + dy * dy;
      m |= (r2 < cut2 ? 1u : 0u) << (k0 + q);
    }
  }
  m &= len >= 32 ? 0xffffffffu : ((1u << len) - 1u);
  const int sb = self - l0;
  if (sb >= 0 && sb < 32) m &= ~(1u << sb);
  return m;
}

// Calls visit(l, uy) for every neighbour of lane particle i, rows bottom to
// top, ascending sorted position within a row (uy: the cell-row term of the
// row).  With rm != nullptr it also returns the neighbour masks of the three
// rows (bit b = candidate ta_k + b) and whether they are complete (every
// row at most 32 candidates).
template <typename R, typename BD, typename Tile, typename LP, typename Visit>
__device__ __forceinline__ void for_neighbours(const KParams& K, const Segment& g, const BD& B,
                                               const Tile& T, const LP& pi, int i, uint32_t key,
                                               Visit&& visit, uint32_t* rm = nullptr,
                                               bool* ok = nullptr) {
  const int nx = K.nx;
  const int cx = (int)key - g.cy * nx;
  const int ca = cx > 0 ? cx - 1 : 0;
  const int cb = (cx + 1 < nx ? cx + 1 : nx - 1) + 1;
  const int self = g.lo[1] + (i - g.rs[1]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (g.rl[k] == 0) continue;
    const R uy = (R)row_uy(K, k);
    const int ta = B.at(k, ca);
    const int tb = B.at(k, cb);
    const int base = g.lo[k] - g.rs[k];
    if (rm && tb - ta > 32) *ok = false;
    for (int c0 = ta; c0 < tb; c0 += 32) {
      const int len = min(32, tb - c0);
      uint32_t m = row_mask<R>(K, T, pi, uy, base + c0, len, self);
      if (rm && c0 == ta) rm[k] = m;
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        visit(base + c0 + b, uy);
      }
    }
  }
}

// One loop over the bits of three row masks in row order (bit b of row k:
// tile entry l_k + b), so that a wavefront runs for its longest lane's
// neighbour count rather than for the sum of the per-row maxima; the order
// of the visits, and so every sum, is that of a row-by-row walk.
template <typename R, typename Visit>
__device__ __forceinline__ void walk_masks(const KParams& K, uint32_t m0, uint32_t m1,
                                           uint32_t m2, int l0, int l1, int l2, Visit&& visit) {
  while (m0 | m1 | m2) {
    const bool s0 = m0 != 0u, s1 = !s0 && m1 != 0u;
    const uint32_t m = s0 ? m0 : (s1 ? m1 : m2);
    const int l = (s0 ? l0 : (s1 ? l1 : l2)) + (int)__builtin_ctz(m);
    const R uy = s0 ? (R)row_uy(K, 0) : (s1 ? (R)row_uy(K, 1) : (R)row_uy(K, 2));
    const uint32_t rest = m & (m - 1u);
    m0 = s0 ? rest : m0;
    m1 = s1 ? rest : m1;
    m2 = (s0 || s1) ? m2 : rest;
    visit(l, uy);
  }
}

// The same order with less work per neighbour: a lane walks one row mask at
// a time and takes the next non-empty row only when the current one is
// exhausted, so the per-neighbour step is a bit scan; the row change runs
// only in the iterations where some lane of the wavefront needs it.  Row k
// has mask m_k, first tile entry l_k and row term u_k; visit(l, u).
template <typename R, typename Visit>
__device__ __forceinline__ void walk_rows(uint32_t m0, uint32_t m1, uint32_t m2, int l0, int l1,
                                          int l2, R u0, R u1, R u2, Visit&& visit) {
  uint32_t m = m0;
  int base = l0;
  R uy = u0;
  while (true) {
    if (m == 0u) {
      const bool t1 = m1 != 0u;
      m = t1 ? m1 : m2;
      base = t1 ? l1 : l2;
      uy = t1 ? u1 : u2;
      m1 = t1 ? m2 : 0u;
      l1 = l2;
      u1 = u2;
      m2 = 0u;
      if (m == 0u) break;
    }
    const int b = (int)__builtin_ctz(m);
    m &= m - 1u;
    visit(base + b, uy);
  }
}

// for_neighbours with the masks of all three rows built first and one walk
// over them (walk_rows): the same visits in the same order, but a wavefront
// runs for its longest lane's neighbour count rather than for the sum over
// the rows of the per-row maxima.  A particle with a row of more than 32
// candidates takes for_neighbours.  The masks go to rm.
template <typename R, typename BD, typename Tile, typename LP, typename Visit>
__device__ __forceinline__ void for_neighbours_rows(const KParams& K, const Segment& g,
                                                    const BD& B, const Tile& T, const LP& pi,
                                                    int i, uint32_t key, Visit&& visit,
                                                    uint32_t* rm, bool* ok) {
  const int nx = K.nx;
  const int cx = (int)key - g.cy * nx;
  const int ca = cx > 0 ? cx - 1 : 0;
  const int cb = (cx + 1 < nx ? cx + 1 : nx - 1) + 1;
  const int self = g.lo[1] + (i - g.rs[1]);
  uint32_t m[3];
  int l[3];
  bool small = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    m[k] = 0u;
    l[k] = 0;
    if (g.rl[k] == 0) continue;
    const int ta = B.at(k, ca);
    const int tb = B.at(k, cb);
    l[k] = g.lo[k] - g.rs[k] + ta;
    small = small && tb - ta <= 32;
    m[k] = row_mask<R>(K, T, pi, (R)row_uy(K, k), l[k], min(32, tb - ta), self);
  }
  if (!small) {
    for_neighbours<R>(K, g, B, T, pi, i, key, visit, rm, ok);
    return;
  }
  walk_rows(m[0], m[1], m[2], l[0], l[1], l[2], (R)row_uy(K, 0), (R)row_uy(K, 1),
            (R)row_uy(K, 2), visit);
  rm[0] = m[0];
  rm[1] = m[1];
  rm[2] = m[2];
}

// FP32 density pass: the neighbour pre-test of candidates l0 .. l0+len-1
// (len <= 32) of one row against K.pre2f, with x = cx cell + off rounded
// once per record (xi, yi: the lane's x, and its y plus the row term); the
// survivors take the exact test in dens_walk_f32.
template <int CAP>
__device__ __forceinline__ uint32_t row_mask_pre(const KParams& K, const DensTile<float, CAP>& T,
                                                 float xi, float yi, int l0, int len, int self) {
  const float pre2 = K.pre2f, cellf = K.cellf;
  uint32_t m = 0;
  for (int k0 = 0; k0 < len; k0 += 8) {
    uint32_t bb = 0u;  // the batch's bits, placed once
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float4 a = T.a[l0 + k0 + q];
      const float dx = xi - __builtin_fmaf(a.z, cellf, a.x);
      const float dy = yi - a.y;
      const float r2 = __builtin_fmaf(dx, dx, dy * dy);
      bb |= (r2 < pre2 ? 1u : 0u) << q;
    }
    m |= bb << k0;
  }
  m &= len >= 32 ? 0xffffffffu : ((1u << len) - 1u);
  const int sb = self - l0;
  if (sb >= 0 && sb < 32) m &= ~(1u << sb);
  return m;
}

// for_neighbours of the FP32 density pass, row by row: the pre-test, then over
// its survivors the exact test (r^2 < (2h)^2 with the direct path's
// arithmetic) and the pair sums, without a branch: a survivor that fails the
// exact test adds nothing (dens_pair with ok = false) and leaves the recorded
// mask.  Same visiting order and sums as for_neighbours; the exact masks in rm.
// Taken by dens_walk_f32 for a particle with a row of more than 32 candidates.
template <bool SUM, int CAP, typename BD>
__device__ __forceinline__ void dens_walk_rows_f32(const KParams& K, const Segment& g, const BD& B,
                                              const DensTile<float, CAP>& T, const LaneF& pi,
                                              float2 vi, int i, uint32_t key, DensOwn<float>& o,
                                              uint32_t* rm, bool* ok) {
  const int nx = K.nx;
  const int cx = (int)key - g.cy * nx;
  const int ca = cx > 0 ? cx - 1 : 0;
  const int cb = (cx + 1 < nx ? cx + 1 : nx - 1) + 1;
  const int self = g.lo[1] + (i - g.rs[1]);
  const float cut2 = K.cut2f;
  const float xi = __builtin_fmaf(pi.cxf, K.cellf, pi.ox);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (g.rl[k] == 0) continue;
    const float uy = row_uy(K, k);
    const float yi = pi.oy + uy;
    const int ta = B.at(k, ca);
    const int tb = B.at(k, cb);
    const int base = g.lo[k] - g.rs[k];
    if (tb - ta > 32) *ok = false;
    for (int c0 = ta; c0 < tb; c0 += 32) {
      const int len = min(32, tb - c0);
      uint32_t m = row_mask_pre(K, T, xi, yi, base + c0, len, self);
      uint32_t w = m;
      while (w) {
        const int b = __builtin_ctz(w);
        w &= w - 1;
        const int l = base + c0 + b;
        float dx, dy;
        T.delta(K, pi, l, uy, dx, dy);
        const float r2 = dx * dx + dy * dy;
        const bool in = r2 < cut2;
        m = in ? m : (m & ~(1u << b));
        const float2 vj = T.vel(l);
        dens_pair<float, SUM>(K, dx, dy, r2, vi.x - vj.x, vi.y - vj.y, o.sum, o.rs, in);
      }
      if (c0 == ta) rm[k] = m;
    }
  }
}

// The FP32 density walk: the pre-test of all three rows first, then one loop
// over the survivors of the three masks in row order (so that a wavefront
// runs for its longest lane's survivor count, not for the sum over the rows
// of the per-row maxima), each through the exact test and the pair sums
// without a branch.  Failed survivors are collected per row (fc, flushed at
// each row change) and taken out of the recorded masks.  Same visiting order
// and sums as for_neighbours; the exact masks in rm.
template <bool SUM, int CAP, typename BD>
__device__ __forceinline__ void dens_walk_f32(const KParams& K, const Segment& g, const BD& B,
                                              const DensTile<float, CAP>& T, const LaneF& pi,
                                              float2 vi, int i, uint32_t key, DensOwn<float>& o,
                                              uint32_t* rm, bool* ok) {
  const int nx = K.nx;
  const int cx = (int)key - g.cy * nx;
  const int ca = cx > 0 ? cx - 1 : 0;
  const int cb = (cx + 1 < nx ? cx + 1 : nx - 1) + 1;
  const int self = g.lo[1] + (i - g.rs[1]);
  const float xi = __builtin_fmaf(pi.cxf, K.cellf, pi.ox);
  uint32_t pm[3];
  int pb[3];
  bool small = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pm[k] = 0u;
    pb[k] = 0;
    if (g.rl[k] == 0) continue;
    const int ta = B.at(k, ca);
    const int tb = B.at(k, cb);
    pb[k] = g.lo[k] - g.rs[k] + ta;
    small = small && tb - ta <= 32;
    pm[k] = row_mask_pre(K, T, xi, pi.oy + row_uy(K, k), pb[k], min(32, tb - ta), self);
  }
  if (!small) {
    dens_walk_rows_f32<SUM>(K, g, B, T, pi, vi, i, key, o, rm, ok);
    return;
  }
  const float cut2 = K.cut2f;
  uint32_t f0 = 0u, f1 = 0u, f2 = 0u, fc = 0u;  // failed survivors: rows 0..2, current row
  int kc = 0;
  uint32_t m = pm[0];
  int base = pb[0];
  float uy = row_uy(K, 0);
  while (true) {
    if (m == 0u) {
      f0 |= kc == 0 ? fc : 0u;
      f1 |= kc == 1 ? fc : 0u;
      f2 |= kc == 2 ? fc : 0u;
      fc = 0u;
      const bool t1 = kc == 0 && pm[1] != 0u;
      const bool t2 = !t1 && kc < 2 && pm[2] != 0u;
      if (!(t1 || t2)) break;
      kc = t1 ? 1 : 2;
      m = t1 ? pm[1] : pm[2];
      base = t1 ? pb[1] : pb[2];
      uy = t1 ? row_uy(K, 1) : row_uy(K, 2);
    }
    const int b = (int)__builtin_ctz(m);
    m &= m - 1u;
    const int l = base + b;
    float dx, dy;
    T.delta(K, pi, l, uy, dx, dy);
    const float r2 = dx * dx + dy * dy;
    const bool in = r2 < cut2;
    fc |= in ? 0u : (1u << b);
    const float2 vj = T.vel(l);
    dens_pair<float, SUM>(K, dx, dy, r2, vi.x - vj.x, vi.y - vj.y, o.sum, o.rs, in);
  }
  rm[0] = pm[0] & ~f0;
  rm[1] = pm[1] & ~f1;
  rm[2] = pm[2] & ~f2;
}

// The same walk driven by masks recorded by the density pass.
// (Rows outside the grid have empty masks: their entries are not looked up.)
template <typename R, typename BD, typename Visit>
__device__ __forceinline__ void for_masked(const KParams& K, const Segment& g, const BD& B,
                                           uint32_t key, uint4 nm, Visit&& visit) {
  const int nx = K.nx;
  const int cx = (int)key - g.cy * nx;
  const int ca = cx > 0 ? cx - 1 : 0;
  walk_masks<R>(K, nm.x, nm.y, nm.z, g.rl[0] ? g.lo[0] - g.rs[0] + B.at(0, ca) : 0,
                g.lo[1] - g.rs[1] + B.at(1, ca), g.rl[2] ? g.lo[2] - g.rs[2] + B.at(2, ca) : 0,
                visit);
}

// The same with walk_rows (a bit scan per neighbour, the row change only
// where a lane needs it).
template <typename R, typename BD, typename Visit>
__device__ __forceinline__ void for_masked_rows(const KParams& K, const Segment& g, const BD& B,
                                                uint32_t key, uint4 nm, Visit&& visit) {
  const int nx = K.nx;
  const int cx = (int)key - g.cy * nx;
  const int ca = cx > 0 ? cx - 1 : 0;
  walk_rows(nm.x, nm.y, nm.z, g.rl[0] ? g.lo[0] - g.rs[0] + B.at(0, ca) : 0,
            g.lo[1] - g.rs[1] + B.at(1, ca), g.rl[2] ? g.lo[2] - g.rs[2] + B.at(2, ca) : 0,
            (R)row_uy(K, 0), (R)row_uy(K, 1), (R)row_uy(K, 2), visit);
}

template <typename R, bool SUM, bool EOS, bool PER, int BLK, int CW, int CAP>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && !PER ? 8 : 1))) void k_density_tiled(St<R> S, KParams K,
                                                       const uint32_t* __restrict__ skey,
                                                       const int* __restrict__ cstart, SegArg sa,
                                                       uint4* __restrict__ nbm) {
  if (halted(K)) return;
  __shared__ DensTile<R, CAP> T;
  __shared__ int cs[3][csw<CW>()];
  const Segment g = segment_of<CW>(K, cstart, skey, sa, cs);
  if (g.O0 >= g.O1) return;
  const auto B = bounds_of<CW>(K, g, cstart, cs);
  const bool tiled = g.total <= CAP;
  if (tiled) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      for (int q = threadIdx.x; q < g.rl[k]; q += BLK) T.put(g.lo[k] + q, S, K, g.rs[k] + q);
  }
  __syncthreads();
  for (int i0 = g.O0; i0 < g.O1; i0 += BLK) {
    const int i = i0 + threadIdx.x;
    if (i >= g.O1) continue;
    if (!tiled || seam<PER>(K, g, skey[i])) {
      density_direct<R, SUM, EOS, PER>(S, K, i, skey, cstart, nbm);
      continue;
    }
    const int self = g.lo[1] + (i - g.rs[1]);
    const auto pi = T.lane(self);
    const auto vi = T.vel(self);
    DensOwn<R> o;
    dens_begin<R, SUM>(K, o);
    uint32_t rm[3] = {0u, 0u, 0u};
    bool ok = true;
    if constexpr (sizeof(R) == 4) {
      dens_walk_f32<SUM>(K, g, B, T, pi, vi, i, skey[i], o, rm, &ok);
    } else {
      for_neighbours_rows<R>(
          K, g, B, T, pi, i, skey[i],
          [&](int l, R uy) {
            R dx, dy;
            T.delta(K, pi, l, uy, dx, dy);
            const R r2 = dx * dx + dy * dy;
            const auto vj = T.vel(l);
            dens_pair<R, SUM>(K, dx, dy, r2, vi.x - vj.x, vi.y - vj.y, o.sum, o.rs);
          },
          rm, &ok);
    }
    nbm[i] = make_uint4(rm[0], rm[1], rm[2], ok ? 1u : 0u);
    dens_end<R, SUM, EOS>(S, K, i, o);
  }
}

template <typename R, bool PER, int BLK, int CW, int CAP>
__global__ __launch_bounds__(BLK) void k_tensor_tiled(St<R> S, KParams K,
                                                      const uint32_t* __restrict__ skey,
                                                      const int* __restrict__ cstart, SegArg sa,
                                                      const uint4* __restrict__ nbm) {
  if (halted(K)) return;
  __shared__ TensTile<R, CAP> T;
  __shared__ int cs[3][csw<CW>()];
  const Segment g = segment_of<CW>(K, cstart, skey, sa, cs);
  if (g.O0 >= g.O1) return;
  const auto B = bounds_of<CW>(K, g, cstart, cs);
  const bool tiled = g.total <= CAP;
  if (tiled) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      for (int q = threadIdx.x; q < g.rl[k]; q += BLK) T.put(g.lo[k] + q, S, K, g.rs[k] + q);
  }
  __syncthreads();
  for (int i0 = g.O0; i0 < g.O1; i0 += BLK) {
    const int i = i0 + threadIdx.x;
    if (i >= g.O1) continue;
    if (!tiled || seam<PER>(K, g, skey[i])) {
      tensor_direct<R, PER>(S, K, i, skey, cstart);
      continue;
    }
    const int self = g.lo[1] + (i - g.rs[1]);
    const auto pi = T.lane(self);
    const auto vi = T.d.vel(self);
    TensOwn<R> o;
    auto visit = [&](int l, R uy) {
      R dx, dy;
      T.delta(K, pi, l, uy, dx, dy);
      const R r2 = dx * dx + dy * dy;
      const auto vj = T.d.vel(l);
      tens_pair(K, o, dx, dy, r2, vj.x - vi.x, vj.y - vi.y, T.q[l]);
    };
    const uint32_t key = skey[i];
    const uint4 nm = nbm ? nbm[i] : make_uint4(0u, 0u, 0u, 0u);
    if (nm.w && sizeof(R) == 4)
      for_masked_rows<R>(K, g, B, key, nm, visit);
    else if (nm.w)
      for_masked<R>(K, g, B, key, nm, visit);
    else
      for_neighbours<R>(K, g, B, T, pi, i, key, visit);
    tens_end(S, K, i, T.q[self], o);
  }
}

// PC: pa_sph and same_e2 fixed at compile time for the FP32 pair body
// (force_pair_f32), -1 to read them from K (k_force_hot fixes them).
template <typename R, bool PRESS, bool AV, bool AH, bool XS, bool VI, bool PER, int BLK, int CW,
          int CAP, int PC = -1>
__global__ __launch_bounds__(BLK) void k_force_tiled(St<R> S, KParams K,
                                                     const uint32_t* __restrict__ skey,
                                                     const int* __restrict__ cstart, SegArg sa,
                                                     const uint4* __restrict__ nbm) {
  if (halted(K)) return;
  __shared__ ForceTile<R, CAP, VI> T;
  __shared__ int cs[3][csw<CW>()];
  const Segment g = segment_of<CW>(K, cstart, skey, sa, cs);
  if (g.O0 >= g.O1) return;
  const auto B = bounds_of<CW>(K, g, cstart, cs);
  const bool tiled = g.total <= CAP;
  if (tiled) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      for (int q = threadIdx.x; q < g.rl[k]; q += BLK)
        T.put(g.lo[k] + q, S, K, g.rs[k] + q, k);
  }
  __syncthreads();
  for (int i0 = g.O0; i0 < g.O1; i0 += BLK) {
    const int i = i0 + threadIdx.x;
    if (i >= g.O1) continue;
    if (!tiled || seam<PER>(K, g, skey[i])) {
      force_direct<R, PRESS, AV, AH, XS, VI, PER>(S, K, i, skey, cstart);
      continue;
    }
    const int self = g.lo[1] + (i - g.rs[1]);
    const auto pi = T.lane(self);
    ForceOwn<R> o;
    force_begin<R, VI>(S, K, i, o);
    auto visit = [&](int l, R uy) {
      R dx, dy;
      T.delta(K, pi, l, uy, dx, dy);
      const R r2 = dx * dx + dy * dy;
      T.template pair<PRESS, AV, AH, XS, PC>(K, o, l, dx, dy, r2);
    };
    const uint32_t key = skey[i];
    const uint4 nm = nbm ? nbm[i] : make_uint4(0u, 0u, 0u, 0u);
    if (nm.w && sizeof(R) == 4) {
      // FP32: the row-refill walk, the row term from the walk
      const int cx = (int)key - g.cy * K.nx;
      const int ca = cx > 0 ? cx - 1 : 0;
      const int l0 = g.rl[0] ? g.lo[0] - g.rs[0] + B.at(0, ca) : 0;
      const int l1 = g.lo[1] - g.rs[1] + B.at(1, ca);
      const int l2 = g.rl[2] ? g.lo[2] - g.rs[2] + B.at(2, ca) : 0;
      const R u0 = (R)row_uy(K, 0), u1 = (R)row_uy(K, 1), u2 = (R)row_uy(K, 2);
      walk_rows(nm.x, nm.y, nm.z, l0, l1, l2, u0, u1, u2, visit);
    } else if (nm.w) {
      // FP64: walk_masks (walk_rows measured 0.6 % slower here)
      for_masked<R>(K, g, B, key, nm, visit);
    } else {
      for_neighbours<R>(K, g, B, T, pi, i, key, visit);
    }
    force_end<R, PRESS, XS, VI>(S, K, i, o);
  }
}

// ---- the hot FP32 force gather ---------------------------------------------
// The dam-break flag set (pressure, artificial viscosity and heat, XSPH, no
// physical viscosity, x not periodic) with pa_sph and same_e2 fixed at
// compile time (PC), on row pieces, with 32-byte records: the record's cell
// column is not staged.  For a neighbour in row k of a lane in cell column
// cx, (cx_i - cx_j) cell is +cell, 0 or -cell as the neighbour's tile entry
// lies before the row's cell cx, in it, or after it (e1_k, e2_k: the entries
// where cells cx and cx+1 start), the exact value the FP32 delta computes.
// The smaller tile lets eight workgroups share a CU.  Particles without
// complete masks take the direct path.
template <int CAP> struct HotForceTile {
  float4 a[CAP];  // offx, offy, vx, vy
  float4 b[CAP];  // Q = 1/rho, A = p/rho, psi, u
  __device__ __forceinline__ void put(int l, const St<float>& S, const KParams& K, int t) {
    const float2 o = S.off[t];
    const float2 v = S.v[t];
    a[l] = make_float4(o.x, o.y, v.x, v.y);
    float Q, A;
    q_and_a(K, S.rho[t], S.p[t], Q, A);
    b[l] = make_float4(Q, A, S.psi[t], S.uf[t]);
  }
};

// A row of the hot walk: first tile entry of the lane's three cells, the
// entries where cells cx and cx+1 start, the row's cell-row term.
struct HotRow {
  int base, e1, e2;
  float uy;
};

// walk_rows two neighbours at a time, with the rows' cell boundaries:
// visit(la, ra, lb, rb, hb) with lb the next neighbour after la in the same
// order when hb (else lb = la, to be ignored), so that the pair arithmetic
// of two neighbours can overlap.
template <typename Visit2>
__device__ __forceinline__ void walk_rows2_hot(uint32_t m0, uint32_t m1, uint32_t m2, HotRow r0,
                                               HotRow r1, HotRow r2, Visit2&& visit) {
  uint32_t m = m0;
  HotRow r = r0;
  auto refill = [&]() {
    const bool t1 = m1 != 0u;
    m = t1 ? m1 : m2;
    r.base = t1 ? r1.base : r2.base;
    r.e1 = t1 ? r1.e1 : r2.e1;
    r.e2 = t1 ? r1.e2 : r2.e2;
    r.uy = t1 ? r1.uy : r2.uy;
    m1 = t1 ? m2 : 0u;
    r1 = r2;
    m2 = 0u;
  };
  while (true) {
    if (m == 0u) {
      refill();
      if (m == 0u) break;
    }
    const int la = r.base + (int)__builtin_ctz(m);
    const HotRow ra = r;
    m &= m - 1u;
    if (m == 0u) refill();
    const bool hb = m != 0u;
    const int lb = hb ? r.base + (int)__builtin_ctz(m) : la;
    HotRow rb;
    rb.e1 = hb ? r.e1 : ra.e1;
    rb.e2 = hb ? r.e2 : ra.e2;
    rb.uy = hb ? r.uy : ra.uy;
    m = hb ? (m & (m - 1u)) : m;
    visit(la, ra, lb, rb, hb);
  }
}

template <int BLK, int CAP, int PC>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) void k_force_hot(
    St<float> S, KParams K, const uint32_t* __restrict__ skey, const int* __restrict__ cstart,
    SegArg sa, const uint4* __restrict__ nbm) {
  if (halted(K)) return;
  __shared__ HotForceTile<CAP> T;
  __shared__ int cs[3][csw<0>()];
  const Segment g = segment_of<0>(K, cstart, skey, sa, cs);
  if (g.O0 >= g.O1) return;
  const auto B = bounds_of<0>(K, g, cstart, cs);
  const bool tiled = g.total <= CAP;
  if (tiled) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      for (int q = threadIdx.x; q < g.rl[k]; q += BLK) T.put(g.lo[k] + q, S, K, g.rs[k] + q);
  }
  __syncthreads();
  for (int i0 = g.O0; i0 < g.O1; i0 += BLK) {
    const int i = i0 + threadIdx.x;
    if (i >= g.O1) continue;
    const uint4 nm = nbm ? nbm[i] : make_uint4(0u, 0u, 0u, 0u);
    if (!tiled || !nm.w) {
      force_direct<float, true, true, true, true, false, false>(S, K, i, skey, cstart);
      continue;
    }
    const int self = g.lo[1] + (i - g.rs[1]);
    const float4 qs = T.a[self];
    ForceOwn<float> o;
    force_begin<float, false>(S, K, i, o);
    const int cx = (int)skey[i] - g.cy * K.nx;
    const int ca = cx > 0 ? cx - 1 : 0;
    HotRow rw[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int d = g.lo[k] - g.rs[k];
      const bool on = k == 1 || g.rl[k] != 0;  // rows outside the grid have empty masks
      rw[k].base = on ? d + B.at(k, ca) : 0;
      rw[k].e1 = on ? d + B.at(k, cx) : 0;
      rw[k].e2 = on ? d + B.at(k, cx + 1) : 0;
      rw[k].uy = row_uy(K, k);
    }
    const float cellf = K.cellf;
    auto terms = [&](int l, const HotRow& r) {
      const float4 q = T.a[l];
      const float4 w = T.b[l];
      const float ux = l < r.e1 ? cellf : (l < r.e2 ? 0.f : -cellf);
      const float tx = qs.x - q.x;
      const float dx = tx + ux;
      const float ty = qs.y - q.y;
      const float dy = ty + r.uy;
      return force_terms_f32<true, true, true, true, PC>(K, o, dx, dy, q.z, q.w, w.x, w.y, w.z,
                                                         w.w);
    };
    // two neighbours per step; each added in walk order
    walk_rows2_hot(nm.x, nm.y, nm.z, rw[0], rw[1], rw[2],
                   [&](int la, const HotRow& ra, int lb, const HotRow& rb, bool hb) {
                     const TermsF ta = terms(la, ra);
                     const TermsF tb = terms(lb, rb);
                     force_acc_f32<true, true, true, true>(o, ta, true);
                     // without a second neighbour G and wf are zeroed (adding
                     // -0 dx to a sum that is never -0 leaves its bits), the
                     // energy rate is skipped
                     const float gb = hb ? tb.G : 0.f, wb = hb ? tb.wf : 0.f;
                     o.fx = __builtin_fmaf(-gb, tb.dx, o.fx);
                     o.fy = __builtin_fmaf(-gb, tb.dy, o.fy);
                     o.du.add_if(tb.d, hb);
                     o.xsx = __builtin_fmaf(-wb, tb.dvx, o.xsx);
                     o.xsy = __builtin_fmaf(-wb, tb.dvy, o.xsy);
                   });
    force_end<float, true, true, false>(S, K, i, o);
  }
}

// Launch shapes: cells per segment (0: row pieces of PIECE particles), lanes,
// LDS record capacity.  Both forms run on row pieces, except the force
// gathers with the physical viscosity (ForceCfg), which run on cell runs.
template <typename R> struct TileCfg;
template <> struct TileCfg<float> {
  static constexpr int BLK = PIECE, CW = 0, CAP = 640;
};
// The hot FP32 force gather (k_force_hot): row pieces; the tile (592 records
// of 32 bytes, 18.9 KB) lets eight workgroups share a CU (a piece's three
// rows hold about 3 x 140 records on the dam-break lattice, 584 at most).
struct HotForceCfg {
  static constexpr int BLK = PIECE, CW = 0, CAP = 592;
};
template <> struct TileCfg<double> {
  static constexpr int BLK = PIECE, CW = 0, CAP = 600;
};
// The force gathers with the physical viscosity stage 16 (FP32) or 24 (FP64)
// more bytes per record: a piece's 600-record tile would leave four (FP32)
// or two (FP64) workgroups per CU, so they run on runs of cells with smaller
// tiles (heated channel: FP32 force 218 -> 164 us, FP64 360 -> 334 us).
template <typename R, bool VI> struct ForceCfg : TileCfg<R> {};
template <> struct ForceCfg<float, true> {
  static constexpr int BLK = 256, CW = 32, CAP = 720;
};
template <> struct ForceCfg<double, true> {
  static constexpr int BLK = 128, CW = 16, CAP = 400;
};

}  // namespace arm
