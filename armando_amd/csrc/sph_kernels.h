// sph_kernels.h -- HIP kernels of the WCSPH step for gfx950.
//
// Every pair pass is a gather: one lane per particle i (sorted order), which
// walks the three cell rows around its cell.  Because keys are ix + iy*nx and
// particles are sorted by (cell, id), the three cells of a row are one
// contiguous range [cell_start[c-1], cell_start[c+2]) of the sorted arrays,
// so each lane reads three contiguous runs.  Each lane writes only its own
// particle: no atomics, bitwise repeatable.
#pragma once

#include "cell_sort.h"
#include "sph_device.h"

namespace arm {

constexpr int BLOCK = 256;

// Zeroes nw 32-bit words.
__global__ __launch_bounds__(BLOCK) void k_zero_words(uint32_t* __restrict__ p, size_t nw) {
  for (size_t k = blockIdx.x * (size_t)BLOCK + threadIdx.x; k < nw;
       k += (size_t)gridDim.x * BLOCK)
    p[k] = 0u;
}

// ------------------------------------------------------------------------
// cell keys (updateHashHost, armando2Dcpu.cu:2080-2101) in FP64
// ------------------------------------------------------------------------
__device__ __forceinline__ int cell_of(const KParams& K, double x, double y, int& ix, int& iy,
                                       bool& ok) {
  const double tx = (x - K.ox) / K.cell;
  const double ty = (y - K.oy) / K.cell;
  ok = (tx >= 0.0) && (ty >= 0.0) && (tx < (double)K.nx) && (ty < (double)K.ny);
  ix = ok ? (int)tx : 0;
  iy = ok ? (int)ty : 0;
  ix = ix < K.nx ? ix : K.nx - 1;
  iy = iy < K.ny ? iy : K.ny - 1;
  ok = ok && ((int)tx < K.nx) && ((int)ty < K.ny);
  return ix + iy * K.nx;
}

// Key of the particle in slot s (emit_key: at its id for the radix sort, at
// its slot with the per-cell counts for the counting sort; a stable sort of
// the keys listed in id order yields (cell, id) order).  FP32 form: also
// refreshes the integer cell and FP32 offset from the FP64 accumulator.
// Ghosts (ids >= nb) in slots below gfirst are those of an earlier step:
// their key is the extra cell ncell, which the sort puts after every real
// cell, so they drop out of the sorted state.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_keys(St<R> S, KParams K, KeyOut ko,
                                                int* __restrict__ err, int gfirst) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = s < K.n;
  uint32_t key = 0;
  int id = 0;
  if (valid && S.id[s] >= K.nb && s < gfirst) {
    id = S.id[s];
    key = (uint32_t)K.ncell;
  } else if (valid) {
    const double2 q = S.xy[s];
    int ix, iy;
    bool ok;
    key = (uint32_t)cell_of(K, q.x, q.y, ix, iy, ok);
    if (!ok) {
      atomicAdd(err, 1);
      raise_halt(K);
    }
    id = S.id[s];
    if constexpr (sizeof(R) == 4) {
      const double bx = K.ox + (double)ix * K.cell;
      const double by = K.oy + (double)iy * K.cell;
      S.cxy[s] = (uint32_t)ix | ((uint32_t)iy << 16);
      S.off[s] = make_float2((float)(q.x - bx), (float)(q.y - by));
    }
  }
  emit_key(ko, valid, s, id, key);
}

// ------------------------------------------------------------------------
// cell_start from sorted keys: cell_start[c] = first sorted position with
// key >= c, for c in [0, ncell].  Sorted position t owns the cells
// (key[t-1], key[t]].  Short gaps are filled by their own lane; long gaps
// (empty regions of the tank) by the whole wavefront, 64 cells per store;
// the cells before the first key and after the last one by every lane of
// the grid, strided.  Must be called by all lanes of the wavefront.
// ------------------------------------------------------------------------
__device__ __forceinline__ void fill_cell_start(const uint32_t* __restrict__ skey, int t,
                                                bool valid, int n, int ncell,
                                                int* __restrict__ cstart) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  const int first = (int)skey[0];
  const int last = (int)skey[n - 1];
  for (int c = g; c <= first; c += stride) cstart[c] = 0;
  for (int c = last + 1 + g; c <= ncell; c += stride) cstart[c] = n;
  int lo = 1, hi = 0;
  if (valid && t > 0) {
    lo = (int)skey[t - 1] + 1;
    hi = (int)skey[t];
  }
  const bool longgap = hi - lo >= 8;
  if (!longgap)
    for (int c = lo; c <= hi; ++c) cstart[c] = t;
  uint64_t m = __ballot(longgap);
  while (m) {
    const int L = __ffsll((unsigned long long)m) - 1;
    const int l0 = __shfl(lo, L, 64), h0 = __shfl(hi, L, 64), tv = __shfl(t, L, 64);
    for (int c = l0 + lane; c <= h0; c += 64) cstart[c] = tv;
    m &= m - 1;
  }
}

// ------------------------------------------------------------------------
// mirror ghosts (flag 512; CERNsource/ghost.f:1-86)
// ------------------------------------------------------------------------
struct Ghosts {
  double2* xprev;  // by slot: position at the last sort (the neighbour search
                   // whose pair list ghost_particles walks)
  int* gsel;       // by slot: slot of the chosen virtual particle, or -1
  int* gflag;      // by real id: 1 if the particle gets a ghost
  int* grank;      // by real id: its ghost's number (exclusive scan of gflag)
  int* gpar;       // by staging slot: the parent's slot (order before the sort)
  int* gwall;      // by staging slot: the virtual particle (index in vxy)
  int* gpid;       // by slot of the sorted order: a ghost's parent id
  int* sid;        // by real id: slot in the sorted order
  uint8_t* vnear;  // by cell: 1 if a virtual particle lies in its 3x3 block
  // the virtual particles (fixed) in (cell, id) order: positions, wall
  // normals, ids and the cell ranges vstart[ncell + 1]
  int* vstart;
  double2* vxy;
  double2* vnrm;
  int* vid;
};

// The fields of ghost slot t (sorted order) from staging slot s: position
// as built, velocity = the parent's (after the predictor when predict) with
// its component along the wall normal n reversed, v - 2 n (v.n)
// (ghost.f:60-79); rho, u, p of the parent.  The EOS pass copies rho, p, u
// again after the equation of state (single_step.f:188-196).
template <typename R>
__device__ __forceinline__ void ghost_slot(const St<R>& A, St<R>& Bs, const KParams& K,
                                           const Ghosts& G, int s, int t, int id,
                                           int predict) {
  using R2 = typename V2<R>::type;
  const int ps = G.gpar[s], ws = G.gwall[s];
  const double2 q = A.xy[s];
  Bs.id[t] = id;
  Bs.xy[t] = q;
  if constexpr (sizeof(R) == 4) {
    int ix, iy;
    bool ok;
    cell_of(K, q.x, q.y, ix, iy, ok);
    const double bx = K.ox + (double)ix * K.cell;
    const double by = K.oy + (double)iy * K.cell;
    Bs.cxy[t] = (uint32_t)ix | ((uint32_t)iy << 16);
    Bs.off[t] = make_float2((float)(q.x - bx), (float)(q.y - by));
  }
  R2 v = A.v[ps];
  R rho = A.rho[ps];
  double u = A.u[ps];
  if (predict) {  // the parent's predictor, as k_reorder_pred applies it
    const double hdt = K.dt / 2.0;
    u = u + hdt * (double)A.du[ps];
    if (u < 0.0) u = 0.0;
    if (K.density_method == 0) rho = rho + (R)hdt * A.drho[ps];
    const R2 a = A.a[ps];
    v.x = v.x + (R)hdt * a.x;
    v.y = v.y + (R)hdt * a.y;
  }
  R2 nrm;
  nrm.x = (R)G.vnrm[ws].x;
  nrm.y = (R)G.vnrm[ws].y;
  R d2 = (R)0;
  d2 = d2 + v.x * nrm.x;
  d2 = d2 + v.y * nrm.y;
  R2 vg;
  vg.x = v.x - (R)2 * nrm.x * d2;
  vg.y = v.y - (R)2 * nrm.y * d2;
  Bs.v[t] = vg;
  Bs.rho[t] = rho;
  Bs.u[t] = u;
  if constexpr (sizeof(R) == 4) Bs.uf[t] = (float)u;
  Bs.p[t] = A.p[ps];
  Bs.flag[t] = 0;
  G.gpid[t] = A.id[ps];
}

// A reorder kernel launched with row pieces (rp.pieces set; cell_start
// written by the sort) has row_blocks(rp) extra workgroups in front, one
// lane per cell row, which list the rows' pieces (they run beside the first
// particles' workgroups).  Returns true in those; *t is the sorted position
// of the lane in the others.
__host__ __device__ inline int row_blocks(const RowPieces& rp) {
  return rp.pieces ? (rp.ny + BLOCK - 1) / BLOCK : 0;
}
__device__ __forceinline__ bool reorder_row_pieces(const int* __restrict__ cstart,
                                                   const RowPieces& rp, int* t) {
  const int rb = row_blocks(rp);
  if ((int)blockIdx.x < rb) {
    row_pieces_block<BLOCK>(cstart, rp, blockIdx.x);
    return true;
  }
  *t = ((int)blockIdx.x - rb) * BLOCK + threadIdx.x;
  return false;
}

// Permute every field from state A to state B (per-routine link_list);
// fill = 0 when the sort already wrote cell_start.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_reorder_all(St<R> A, St<R> Bs, KParams K,
                                                       const uint32_t* __restrict__ skey,
                                                       const uint32_t* __restrict__ sval,
                                                       int* __restrict__ cstart, int fill,
                                                       Ghosts G, RowPieces rp) {
  int t;
  if (reorder_row_pieces(cstart, rp, &t)) return;
  const bool valid = t < K.n;
  if (fill) fill_cell_start(skey, t, valid, K.n, K.ncell, cstart);
  if (!valid) return;
  const int s = (int)sval[t];
  if (G.xprev) {
    const int id = A.id[s];
    if (id >= K.nb) {
      ghost_slot(A, Bs, K, G, s, t, id, 0);
      return;
    }
    G.xprev[t] = A.xy[s];
    if (id < K.np) G.sid[id] = t;
  }
  Bs.id[t] = A.id[s];
  Bs.xy[t] = A.xy[s];
  if constexpr (sizeof(R) == 4) {
    Bs.cxy[t] = A.cxy[s];
    Bs.off[t] = A.off[s];
    Bs.uf[t] = A.uf[s];
  }
  Bs.v[t] = A.v[s];
  Bs.rho[t] = A.rho[s];
  Bs.u[t] = A.u[s];
  Bs.p[t] = A.p[s];
  Bs.psi[t] = A.psi[s];
  Bs.drho[t] = A.drho[s];
  Bs.du[t] = A.du[s];
  Bs.a[t] = A.a[s];
  Bs.av[t] = A.av[s];
  Bs.divv[t] = A.divv[s];
  Bs.rho_min[t] = A.rho_min[s];
  Bs.u_min[t] = A.u_min[s];
  Bs.v_min[t] = A.v_min[s];
  Bs.flag[t] = A.flag[s];
  if (K.sle) Bs.hs[t] = A.hs[s];
}

// Fused reorder + leapfrog predictor (time_integration.f:136-154): the
// predictor does not move particles, so sorting before it is equivalent.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_reorder_pred(St<R> A, St<R> Bs, KParams K,
                                                        const uint32_t* __restrict__ skey,
                                                        const uint32_t* __restrict__ sval,
                                                        int* __restrict__ cstart, int predict,
                                                        int fill, Ghosts G, RowPieces rp) {
  if (halted(K)) return;
  int t;
  if (reorder_row_pieces(cstart, rp, &t)) return;
  using R2 = typename V2<R>::type;
  const bool valid = t < K.n;
  if (fill) fill_cell_start(skey, t, valid, K.n, K.ncell, cstart);
  if (!valid) return;
  const int s = (int)sval[t];
  const int id = A.id[s];
  if (G.xprev) {
    if (id >= K.nb) {
      ghost_slot(A, Bs, K, G, s, t, id, predict);
      return;
    }
    G.xprev[t] = A.xy[s];
    if (id < K.np) G.sid[id] = t;
  }
  Bs.id[t] = id;
  const double2 q = A.xy[s];
  Bs.xy[t] = q;
  if (K.sle) Bs.hs[t] = A.hs[s];
  if constexpr (sizeof(R) == 4) {
    // cell + offset recomputed from the FP64 accumulator (as k_keys does)
    int ix, iy;
    bool ok;
    cell_of(K, q.x, q.y, ix, iy, ok);
    const double bx = K.ox + (double)ix * K.cell;
    const double by = K.oy + (double)iy * K.cell;
    Bs.cxy[t] = (uint32_t)ix | ((uint32_t)iy << 16);
    Bs.off[t] = make_float2((float)(q.x - bx), (float)(q.y - by));
  }
  R2 v = A.v[s];
  R rho = A.rho[s];
  double u = A.u[s];
  uint8_t fl = 0;
  if (predict) {
    const double hdt = K.dt / 2.0;
    const bool rate = is_rate(K, id);
    if (rate) {
      Bs.u_min[t] = u;
      u = u + hdt * (double)A.du[s];
      if (u < 0.0) {
        u = 0.0;
        fl = 1u;
      }
      if (K.density_method == 0) {
        Bs.rho_min[t] = rho;
        rho = rho + (R)hdt * A.drho[s];
      }
    }
    if (id < K.np) {
      const R2 a = A.a[s];
      Bs.v_min[t] = v;
      v.x = v.x + (R)hdt * a.x;
      v.y = v.y + (R)hdt * a.y;
    }
  }
  Bs.v[t] = v;
  Bs.rho[t] = rho;
  Bs.u[t] = u;
  if constexpr (sizeof(R) == 4) Bs.uf[t] = (float)u;
  Bs.flag[t] = fl;
  // the EOS recomputes p of every rate particle; others keep theirs
  if (!is_rate(K, id)) Bs.p[t] = A.p[s];
}

// ------------------------------------------------------------------------
// neighbour walk helpers
// ------------------------------------------------------------------------
// Candidate ranges of the three rows cy-1 .. cy+1 over the x-adjacent cells.
struct Rows {
  int t0[3], t1[3];
};

__device__ __forceinline__ Rows rows_of(const KParams& K, uint32_t key,
                                        const int* __restrict__ cstart) {
  const int cy = (int)key / K.nx;
  const int cx = (int)key - cy * K.nx;
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx + 1 < K.nx ? cx + 1 : K.nx - 1;
  Rows r;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ry = cy - 1 + k;
    if (ry < 0 || ry >= K.ny) {
      r.t0[k] = 0;
      r.t1[k] = 0;
    } else {
      r.t0[k] = cstart[ry * K.nx + x0];
      r.t1[k] = cstart[ry * K.nx + x1 + 1];
    }
  }
  return r;
}

// On a grid periodic in x the first and last columns see the other end:
// the wrapped column's ranges of the same three rows and the shift of its
// image, sh = +nx cells for cx = 0 and -nx for cx = nx-1 (0 elsewhere).
__device__ __forceinline__ Rows wrap_rows(const KParams& K, uint32_t key,
                                          const int* __restrict__ cstart, int& sh) {
  const int cy = (int)key / K.nx;
  const int cx = (int)key - cy * K.nx;
  sh = cx == 0 ? K.nx : cx == K.nx - 1 ? -K.nx : 0;
  const int wc = cx == 0 ? K.nx - 1 : 0;
  Rows r;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ry = cy - 1 + k;
    if (sh == 0 || ry < 0 || ry >= K.ny) {
      r.t0[k] = 0;
      r.t1[k] = 0;
    } else {
      r.t0[k] = cstart[ry * K.nx + wc];
      r.t1[k] = cstart[ry * K.nx + wc + 1];
    }
  }
  return r;
}

// Every candidate of the cell of `key`: f(t, sh, k, t0) for candidate t of
// row k (t0 = the row's first candidate), first the unwrapped rows (sh = 0),
// then, with periodic x, the wrapped column.  The two phases keep the
// unwrapped walk free of the wrap's row bounds; WRAP = false compiles the
// wrap out (callers that never see a periodic grid).
template <bool WRAP = true, typename F>
__device__ __forceinline__ void for_candidates(const KParams& K, uint32_t key,
                                               const int* __restrict__ cstart, F&& f) {
  const Rows rw = rows_of(K, key, cstart);
  for (int k = 0; k < 3; ++k)
    for (int t = rw.t0[k]; t < rw.t1[k]; ++t) f(t, 0, k, rw.t0[k]);
  if (WRAP && K.periodic) {
    int sh;
    const Rows w = wrap_rows(K, key, cstart, sh);
    if (sh)
      for (int k = 0; k < 3; ++k)
        for (int t = w.t0[k]; t < w.t1[k]; ++t) f(t, sh, k, w.t0[k]);
  }
}

// ------------------------------------------------------------------------
// pair arithmetic shared by every gather path (same operations in the same
// order, so the tiled and the direct paths give identical bits)
// ------------------------------------------------------------------------

// density pass: con_density (density.f:166-223) and, for method 1,
// sum_density (density.f:1-62).  dvx = v_i.x - v_j.x.
// ok = false: the pair is no neighbour and adds nothing (the sums keep
// every word, as if it had not been visited).
template <typename R, bool SUM>
__device__ __forceinline__ void dens_pair(const KParams& K, R dx, R dy, R r2, R dvx, R dvy,
                                          Acc<R>& sum, Acc<R>& rs, bool ok = true) {
  if constexpr (sizeof(R) == 4) {
    // m dW.dv = m C1 (q - 2)^3 (dx dvx + dy dvy), m C1 one constant
    const float q = fsqrt(r2) * K.kf.invh;
    const float tq = q - 2.f;
    const float vr = __builtin_fmaf(dvy, dy, dvx * dx);
    sum.add_if(K.kf.mC1 * ((tq * tq) * tq * vr), ok);
    if (SUM) {
      const float s = __builtin_fmaf(-0.5f, q, 1.f);
      const float s2 = s * s;
      rs.add_if(K.kf.mw * ((s2 * s2) * __builtin_fmaf(2.f, q, 1.f)), ok);
    }
    return;
  }
  const R q = fsqrt(r2) * (R)K.invh;
  const R tq = q - (R)2;
  const R F = (R)K.C1 * (tq * tq * tq);
  const R dwx = F * dx, dwy = F * dy;
  const R vcc = fmr(dvy, dwy, dvx * dwx);
  sum.add_if((R)K.mass * vcc, ok);
  if (SUM) {
    const R s = fmr((R)-0.5, q, (R)1);
    const R s2 = s * s;
    rs.add_if((R)K.mass * ((R)K.wnorm * (s2 * s2) * fmr((R)2, q, (R)1)), ok);
  }
}

template <typename R> struct DensOwn {
  Acc<R> sum, rs;
};

template <typename R, bool SUM>
__device__ __forceinline__ void dens_begin(const KParams& K, DensOwn<R>& o) {
  if (SUM) o.rs.add((R)K.mass * (R)K.wnorm);
}

// EOS of one particle from its stored density (Tait, eos.f90:255-283, or
// linear) with the p < pmin clamp, then the art_heat coefficient from the
// raw divergence sum (divv = sum / rho).  Shared by k_eos and the density
// pass with the EOS folded in, so both give the same bits.
template <typename R, typename SS>
__device__ __forceinline__ void eos_point(SS& S, const KParams& K, int i, bool rate, R rho,
                                          R raw_divv) {
  if (rate) {
    R p = K.eos == 1 ? linear_p(K, rho) : tait(K, rho);
    if (p < (R)K.pmin) {
      p = (R)0;
      rho = (sizeof(R) == 4) ? (R)0 : (R)K.rho0;
      S.rho[i] = rho;
      S.flag[i] |= 2u;
    }
    S.p[i] = p;
  }
  const R dv = raw_divv / dens(K, rho);
  S.divv[i] = dv;
  const R ad = dv < (R)0 ? -dv : dv;
  S.psi[i] = (R)K.psi0 + (R)K.g2h2 * (ad - dv);
}

// Writes drho (rate particles), rho (summation) and the raw divergence sum
// of art_heat's first pass (art_heat.f:46-67): with a uniform mass it is the
// negative of the continuity sum; the EOS divides it by rho.  With EOS the
// equation of state follows at once (single_step.f:119-186 comes right after
// the density call).
template <typename R, bool SUM, bool EOS = false, typename SS = St<R>>
__device__ __forceinline__ void dens_end(SS& S, const KParams& K, int i, const DensOwn<R>& o) {
  const bool rate = is_rate(K, S.id[i]);
  R rho = (R)0;
  if (!SUM) {
    S.drho[i] = rate ? o.sum.get() : (R)0;
    if (EOS) rho = S.rho[i];
  } else {
    S.drho[i] = (R)0;
    if (rate) {
      if constexpr (sizeof(R) == 4)
        rho = o.rs.get() - (float)K.rho0;
      else
        rho = o.rs.get();
      S.rho[i] = rho;
    } else if (EOS) {
      rho = S.rho[i];
    }
  }
  if (EOS)
    eos_point(S, K, i, rate, rho, -o.sum.get());
  else
    S.divv[i] = -o.sum.get();
}

// force pass: int_force pressure part (internal_force.f:178-259, 309-430),
// art_visc (art_visc.f:55-101), art_heat second pass (art_heat.f:69-97)
// and av_vel (av_vel.f:37-56), rewritten around per-particle
//   Q = 1/rho and A = p/rho:
//   pa_sph 0: (p_i+p_j)/(rho_i rho_j) = A_i Q_j + A_j Q_i
//   pa_sph 1: p_i/rho_i^2 + p_j/rho_j^2 = A_i Q_i + A_j Q_j
//   1/rho_bar = 2/(rho_i+rho_j) = 2 Q_i Q_j / (Q_i + Q_j)
// and grad_i W = F (x_i - x_j), so (v_i-v_j).grad W = F vr and
// (x_i-x_j).grad W = F r^2.  Per pair:
//   G   = m F (a + Pi_ij)                  (Pi_ij only when vr < 0)
//   a_i -= G (x_i - x_j)
//   du_i += G vr / 2 + m (psi_i+psi_j) F r^2 / (rho_bar (r^2+(h eta)^2)) (u_i-u_j)
//   av_i -= eps m / rho_bar W (v_i - v_j)
// The energy rate is one compensated (FP32) sum.
template <typename R> struct ForceOwn {
  using R2 = typename V2<R>::type;
  R2 v;
  R Q, A, psi, u;
  // FP32 pair body (force_terms_f32): with c4 = f_cm 4 Q, the factors
  // cA = f_cm A, cQ = f_cm Q, cAQ = f_cm A Q, avb = c4 avbh2, ava = c4 avah,
  // c4, c4psi = c4 psi, xw = f_xs 4 Q and xw2 = 2 xw
  R cA, cQ, cAQ, avb, ava, c4, c4psi, xw, xw2;
  R wxx = 0, wyy = 0, wxy = 0, tds = 0;  // physical viscosity (tau of i)
  R fx = 0, fy = 0, xsx = 0, xsy = 0;
  Acc<R> du;
};

// W of a neighbour for the viscous force (unused without viscosity).
template <typename R> struct Tau3 {
  R xx, yy, xy;
};

// Viscous weight of a particle: W = c T with c = eta (pa_sph 0) or
// eta / rho^2 = eta Q^2 (pa_sph 1).
template <typename R> __device__ __forceinline__ R visc_c(const KParams& K, R Q) {
  return K.pa_sph == 0 ? (R)K.eta : (R)K.eta * Q * Q;
}

template <typename R> __device__ __forceinline__ Tau3<R> tau_of(const St<R>& S, int t, R c) {
  const R* w = S.tau + 4 * (size_t)t;
  return {c * w[0], c * w[1], c * w[2]};
}

// Q = 1/rho and A = p/rho of a particle from its stored density.
template <typename R>
__device__ __forceinline__ void q_and_a(const KParams& K, R rho_s, R p, R& Q, R& A) {
  Q = (R)1 / dens(K, rho_s);
  A = p * Q;
}

// The FP32 pair body's factors of the particle and the constants of K.
__device__ __forceinline__ void own_factors_f32(const KParams& K, ForceOwn<float>& o) {
  const float cm = K.kf.cm, q4 = 4.f * o.Q, c4 = cm * q4;
  o.cA = cm * o.A;
  o.cQ = cm * o.Q;
  o.cAQ = o.cA * o.Q;
  o.avb = c4 * K.kf.avbh2;
  o.ava = c4 * K.kf.avah;
  o.c4 = c4;
  o.c4psi = c4 * o.psi;
  o.xw = K.kf.xs * q4;
  o.xw2 = 2.f * o.xw;
}

template <typename R, bool VI = false>
__device__ __forceinline__ void force_begin(const St<R>& S, const KParams& K, int i, ForceOwn<R>& o) {
  o.v = S.v[i];
  q_and_a(K, S.rho[i], S.p[i], o.Q, o.A);
  o.psi = S.psi[i];
  if constexpr (sizeof(R) == 4) {
    o.u = S.uf[i];
    own_factors_f32(K, o);
  } else {
    o.u = (R)S.u[i];
  }
  if (VI) {
    const R* w = S.tau + 4 * (size_t)i;
    const R c = visc_c(K, o.Q);
    o.wxx = c * w[0];
    o.wyy = c * w[1];
    o.wxy = c * w[2];
    o.tds = w[3];
  }
}

// Viscous force of internal_force.f:196-430 (the `visc` branches):
//   pa_sph 0: a_i += m (eta T_i + eta T_j) . grad W / (rho_i rho_j)
//   pa_sph 1: a_i += m (eta T_i / rho_i^2 + eta T_j / rho_j^2) . grad W
// with the per-particle W = eta T (pa_sph 0) or eta T / rho^2 (pa_sph 1).
template <typename R>
__device__ __forceinline__ void visc_pair(const KParams& K, ForceOwn<R>& o, R mF, R dx, R dy,
                                          R Qj, const Tau3<R>& wj) {
  R sxx = o.wxx + wj.xx, syy = o.wyy + wj.yy, sxy = o.wxy + wj.xy;
  if (K.pa_sph == 0) {
    const R qq = o.Q * Qj;
    sxx = qq * sxx;
    syy = qq * syy;
    sxy = qq * sxy;
  }
  o.fx = fmr(mF, fmr(sxy, dy, sxx * dx), o.fx);
  o.fy = fmr(mF, fmr(syy, dy, sxy * dx), o.fy);
}

// The FP32 form's pair terms: the same quantities with fewer operations.
//  - m F = f_cm s^3 with s = 1 - q/2, since (q - 2)^3 = -8 s^3; W = wnorm s^4 (1 + 2q).
//  - imr2 = 2/rho_bar = 4 Q Q_j/(Q + Q_j); the factors 1/2 this doubles go into
//    the constants f_ava, f_avb (artificial viscosity) and f_xs (XSPH).
//  - The energy rate is accumulated doubled (G vr plus twice the heat terms)
//    and halved once in force_end.
//  - The viscosity term uses phi = min(vr, 0) h/(r^2 + (h eta)^2): zero for
//    receding pairs, which then add an exact zero (the reference's vr < 0 test).
//    (f_avb phi + f_ava) phi is evaluated as e (avbh2 e + avah) with
//    e = phi / h, the factors of h folded into the two constants.
//  - The particle's own factors f_cm and 4 Q are applied once per particle
//    (ForceOwn: c4 = f_cm 4 Q and the products with it): per pair remain
//    s^3, Q_j/(Q + Q_j) and their product M, which the viscosity, heat and
//    XSPH terms share (m F imr2 = c4 M; f_xs s^4 (1 + 2q) imr2 =
//    f_xs 4 Q (1 + 2q) s M).
//  - r^2 of the arithmetic is one fused multiply-add (the neighbour test, done
//    before, uses the unfused form).
// PC < 0 reads pa_sph and same_e2 from K; PC >= 0 fixes them at compile time
// (bit 0: pa_sph == 0, bit 1: same_e2) with the same expressions, so both give
// identical bits, and leaves out conduction (its callers run without it).
// The pair terms apart from their accumulation (force_acc_f32): the particle
// gains -G dx, -G dy, the doubled energy rate d and -wf dv.  mF: for the
// physical viscosity, which force_pair_f32 adds before them.
struct TermsF {
  float mF, G, dx, dy, d, wf, dvx, dvy;
};

template <bool PRESS, bool AV, bool AH, bool XS, int PC>
__device__ __forceinline__ TermsF force_terms_f32(const KParams& K, const ForceOwn<float>& o,
                                                  float dx, float dy, float vjx, float vjy,
                                                  float Qj, float Aj, float psij, float uj) {
  const bool pa0 = PC < 0 ? K.pa_sph == 0 : (PC & 1) != 0;
  const bool se = PC < 0 ? K.same_e2 != 0 : (PC & 2) != 0;
  TermsF t;
  t.dx = dx;
  t.dy = dy;
  const float r2 = __builtin_fmaf(dx, dx, dy * dy);
  const float q = fsqrt(r2) * K.kf.invh;
  const float s = __builtin_fmaf(-0.5f, q, 1.0f);
  const float s2 = s * s;
  const float s3 = s2 * s;
  t.mF = K.kf.cm * s3;
  const float dvx = o.v.x - vjx, dvy = o.v.y - vjy;
  t.dvx = dvx;
  t.dvy = dvy;
  const float vr = __builtin_fmaf(dvy, dy, dvx * dx);
  const float M = s3 * (Qj * frcp(o.Q + Qj));
  float G = 0.f;
  if (PRESS)
    G = s3 * (pa0 ? __builtin_fmaf(o.cA, Qj, Aj * o.cQ) : __builtin_fmaf(Aj * Qj, K.kf.cm, o.cAQ));
  float inv_ea = 0.f;
  if (AV) {
    inv_ea = frcp(r2 + K.kf.ave2);
    const float e = __builtin_fminf(vr, 0.f) * inv_ea;
    G = __builtin_fmaf(e * __builtin_fmaf(o.avb, e, o.ava), M, G);
  }
  t.G = 0.f;
  t.d = 0.f;
  if (PRESS || AV) {
    t.G = G;
    float d = G * vr;
    float inv_eh = 0.f;
    if (AH) {
      inv_eh = (AV && se) ? inv_ea : frcp(r2 + K.kf.ahe2);
      d = __builtin_fmaf(__builtin_fmaf(o.c4, psij, o.c4psi) * M * (r2 * inv_eh), o.u - uj, d);
    }
    if (PRESS && PC < 0 && (K.flags & 128)) {  // Cleary-Monaghan conduction
      if (!AH) inv_eh = frcp(r2 + K.kf.ahe2);
      d = __builtin_fmaf(K.kf.cond2 * o.Q * Qj * t.mF * (r2 * inv_eh), o.u - uj, d);
    }
    t.d = d;
  } else if (AH) {
    const float inv_eh = frcp(r2 + K.kf.ahe2);
    t.d = __builtin_fmaf(o.c4, psij, o.c4psi) * M * (r2 * inv_eh) * (o.u - uj);
  }
  t.wf = 0.f;
  if (XS) t.wf = __builtin_fmaf(o.xw2, q, o.xw) * s * M;
  return t;
}

// Adds the terms of one pair when ok (else nothing), without a branch.
template <bool PRESS, bool AV, bool AH, bool XS>
__device__ __forceinline__ void force_acc_f32(ForceOwn<float>& o, const TermsF& t, bool ok) {
  if (PRESS || AV) {
    o.fx = ok ? __builtin_fmaf(-t.G, t.dx, o.fx) : o.fx;
    o.fy = ok ? __builtin_fmaf(-t.G, t.dy, o.fy) : o.fy;
  }
  if (PRESS || AV || AH) o.du.add_if(t.d, ok);
  if (XS) {
    o.xsx = ok ? __builtin_fmaf(-t.wf, t.dvx, o.xsx) : o.xsx;
    o.xsy = ok ? __builtin_fmaf(-t.wf, t.dvy, o.xsy) : o.xsy;
  }
}

template <bool PRESS, bool AV, bool AH, bool XS, bool VI, int PC>
__device__ __forceinline__ void force_pair_f32(const KParams& K, ForceOwn<float>& o, float dx,
                                               float dy, float vjx, float vjy, float Qj,
                                               float Aj, float psij, float uj,
                                               const Tau3<float>& wj) {
  const TermsF t =
      force_terms_f32<PRESS, AV, AH, XS, PC>(K, o, dx, dy, vjx, vjy, Qj, Aj, psij, uj);
  if (VI) visc_pair(K, o, t.mF, dx, dy, Qj, wj);
  force_acc_f32<PRESS, AV, AH, XS>(o, t, true);
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS, bool VI = false, int PC = -1>
__device__ __forceinline__ void force_pair(const KParams& K, ForceOwn<R>& o, R dx, R dy, R r2,
                                           R vjx, R vjy, R Qj, R Aj, R psij, R uj,
                                           const Tau3<R>& wj = Tau3<R>{}) {
  if constexpr (sizeof(R) == 4) {
    force_pair_f32<PRESS, AV, AH, XS, VI, PC>(K, o, dx, dy, vjx, vjy, Qj, Aj, psij, uj, wj);
    return;
  }
  const R m = (R)K.mass;
  const R q = fsqrt(r2) * (R)K.invh;
  const R tq = q - (R)2;
  const R F = (R)K.C1 * (tq * tq * tq);
  const R dvx = o.v.x - vjx, dvy = o.v.y - vjy;
  const R vr = fmr(dvy, dy, dvx * dx);
  const R inv_mrho = (R)2 * o.Q * Qj * frcp(o.Q + Qj);
  R coef = 0;
  if (PRESS) coef = (K.pa_sph == 0) ? fmr(o.A, Qj, Aj * o.Q) : fmr(o.A, o.Q, Aj * Qj);
  R inv_ea = 0;
  if (AV) {
    inv_ea = frcp(r2 + (R)K.av_e2);
    if (vr < (R)0) {
      const R phi = (R)K.h * vr * inv_ea;
      const R pc = fmr((R)K.av_beta, phi, -((R)K.av_alpha * (R)K.c0));
      coef = fmr(pc * phi, inv_mrho, coef);
    }
  }
  if (VI) visc_pair(K, o, m * F, dx, dy, Qj, wj);
  if (PRESS || AV) {
    const R G = m * F * coef;
    o.fx = fmr(-G, dx, o.fx);
    o.fy = fmr(-G, dy, o.fy);
    R dut = (R)0.5 * G * vr;
    R inv_eh = (R)0;
    if (AH) {
      inv_eh = (AV && K.same_e2) ? inv_ea : frcp(r2 + (R)K.ah_e2);
      dut = fmr(m * (o.psi + psij) * F * r2 * inv_mrho * inv_eh, o.u - uj, dut);
    }
    if (PRESS && (K.flags & 128)) {
      // Cleary-Monaghan conduction: m cond Q_i Q_j (u_i - u_j) F r^2/(r^2 + (h eta)^2)
      if (!AH) inv_eh = frcp(r2 + (R)K.ah_e2);
      dut = fmr(m * (R)K.cond * o.Q * Qj * F * r2 * inv_eh, o.u - uj, dut);
    }
    o.du.add(dut);
  } else if (AH) {
    const R inv_eh = frcp(r2 + (R)K.ah_e2);
    o.du.add(m * (o.psi + psij) * F * r2 * inv_mrho * inv_eh * (o.u - uj));
  }
  if (XS) {
    const R s = fmr((R)-0.5, q, (R)1);
    const R s2 = s * s;
    const R w = (R)K.wnorm * (s2 * s2) * fmr((R)2, q, (R)1);
    const R wf = (R)K.eps_m * inv_mrho * w;
    o.xsx = fmr(-wf, dvx, o.xsx);
    o.xsy = fmr(-wf, dvy, o.xsy);
  }
}

template <typename R, bool PRESS, bool XS, bool VI = false>
__device__ __forceinline__ void force_end(St<R>& S, const KParams& K, int i, const ForceOwn<R>& o) {
  using R2 = typename V2<R>::type;
  if (PRESS) {
    const int id = S.id[i];
    const bool grav = (K.flags & 16) && id < K.np;
    R gx, gy;
    if constexpr (sizeof(R) == 4) {
      gx = K.kf.gx;
      gy = K.kf.gy;
    } else {
      gx = (R)K.gx;
      gy = (R)K.gy;
    }
    const R ex = grav ? gx : (R)0, ey = grav ? gy : (R)0;
    R2 acc;
    acc.x = o.fx + ex;
    acc.y = o.fy + ey;
    S.a[i] = acc;
    // du = tds/dt + the pair sums (internal_force.f:433-435) [+ heat source];
    // the FP32 pair body accumulates the sums doubled
    const R sums = sizeof(R) == 4 ? (R)0.5 * o.du.get() : o.du.get();
    R du = VI ? sums + o.tds : sums;
    if ((K.flags & 64) && id < K.np) {
      const double2 q = S.xy[i];
      du = du + (R)(K.hs_q0 * heat_shape(K, q.x, q.y)) * o.Q;
    }
    S.du[i] = du;
  }
  if (XS) {
    R2 w;
    w.x = o.xsx;
    w.y = o.xsy;
    S.av[i] = w;
  }
}

// The deposition heat source after the force pass, in the steps where it
// acts: du += v(x, y, 0) Q with Q = 1/rho as the force pass computes it
// (fluka_heat.f:485-573; exdudt comes last in single_step.f:226).  A kernel
// of its own so that the force gather carries none of its registers.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_deposition(St<R> S, KParams K) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n || S.id[i] >= K.np) return;
  const double2 q = S.xy[i];
  R Q, A;
  q_and_a(K, S.rho[i], S.p[i], Q, A);
  S.du[i] = S.du[i] + (R)deposition_at(K, q.x, q.y, 0.0) * Q;
}

// ------------------------------------------------------------------------
// contact displacement (flag 1024) of FORTRANsource/external_force.f90,
// after the force pass of a step (time_integration.f90:128-142)
// ------------------------------------------------------------------------
// contact_displacement (external_force.f90:191-244): real particle i walks
// its virtual neighbours (unit normal n_j in v) in gather order and pushes
// its displacement d to the contact distance: pni = h/hdr - (x_i - x_j).n_j,
// pnm = d.n_j, d += (pni - pnm) n_j when pni > 0 and pni > pnm.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_contact_dx(St<R> S, KParams K,
                                                      const uint32_t* __restrict__ skey,
                                                      const int* __restrict__ cstart,
                                                      typename V2<R>::type* __restrict__ cd) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  typename V2<R>::type d;
  d.x = (R)0;
  d.y = (R)0;
  if (S.id[i] < K.np) {
    const auto pi = load_pos(S, i);
    const R cut2 = cutoff2<R>(K);
    const R hh = (R)K.contact_h;
    for_candidates<false>(K, skey[i], cstart, [&](int t, int sh, int, int) {
      const int jd = S.id[t];
      if (jd < K.np || jd >= K.nb) return;
      const auto pj = load_pos(S, t);
      R dx, dy;
      delta(pi, pj, K, dx, dy, sh);
      const R r2 = dx * dx + dy * dy;
      if (!(r2 < cut2)) return;
      const auto n = S.v[t];
      R pni = (R)0, pnm = (R)0;
      pni = pni + dx * n.x;
      pnm = pnm + d.x * n.x;
      pni = pni + dy * n.y;
      pnm = pnm + d.y * n.y;
      pni = hh - pni;
      if (pni > (R)0 && pni > pnm) {
        d.x = d.x + (pni - pnm) * n.x;
        d.y = d.y + (pni - pnm) * n.y;
      }
    });
  }
  cd[i] = d;
}

// smooth_vector (external_force.f90:247-339) and the acceleration of
// time_integration.f90:138-142: a_i += 0.5 (W(0) m/rho_i d_i
// + sum over real neighbours W_ij m/rho_i d_j) / dt^2.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_contact_apply(St<R> S, KParams K,
                                                         const uint32_t* __restrict__ skey,
                                                         const int* __restrict__ cstart,
                                                         const typename V2<R>::type* __restrict__ cd) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n || S.id[i] >= K.np) return;
  const auto pi = load_pos(S, i);
  const R cut2 = cutoff2<R>(K);
  const R m = (R)K.mass;
  const R rhoi = dens(K, S.rho[i]);
  const R w0 = (R)K.wnorm;
  const auto di = cd[i];
  R sx = w0 * m / rhoi * di.x;
  R sy = w0 * m / rhoi * di.y;
  for_candidates<false>(K, skey[i], cstart, [&](int t, int sh, int, int) {
    if (t == i || S.id[t] >= K.np) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    if (!(r2 < cut2)) return;
    const R q = fsqrt(r2) * (R)K.invh;
    const R s = (R)1 - (R)0.5 * q;
    const R s2 = s * s;
    const R w = (R)K.wnorm * (s2 * s2) * ((R)1 + (R)2 * q);
    const auto dj = cd[t];
    sx = sx + w * m / rhoi * dj.x;
    sy = sy + w * m / rhoi * dj.y;
  });
  const R dt2 = (R)(K.dt * K.dt);
  auto a = S.a[i];
  a.x = a.x + (R)0.5 * sx / dt2;
  a.y = a.y + (R)0.5 * sy / dt2;
  S.a[i] = a;
}

template <typename R> __device__ __forceinline__ R u_of(const St<R>& S, int t) {
  if constexpr (sizeof(R) == 4) return S.uf[t]; else return (R)S.u[t];
}

// ---- direct path: neighbour data straight from global memory ------------
// Also records the neighbour masks of the three rows (bit b = candidate
// t0 + b) for the force pass; w = 0 marks them incomplete (a row longer than
// 32 candidates).
template <typename R, bool SUM, bool EOS = false, bool WRAP = true>
__device__ __forceinline__ void density_direct(St<R>& S, const KParams& K, int i,
                                               const uint32_t* __restrict__ skey,
                                               const int* __restrict__ cstart,
                                               uint4* __restrict__ nbm) {
  const auto pi = load_pos(S, i);
  const auto vi = S.v[i];
  const R cut2 = cutoff2<R>(K);
  DensOwn<R> o;
  dens_begin<R, SUM>(K, o);
  uint32_t rm[3] = {0u, 0u, 0u};
  bool ok = true;  // masks cover the unwrapped rows of up to 32 candidates
  for_candidates<WRAP>(K, skey[i], cstart, [&](int t, int sh, int k, int t0) {
    if (sh != 0 || t - t0 >= 32) ok = false;
    if (t == i) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    if (!(r2 < cut2)) return;
    if (sh == 0 && t - t0 < 32) rm[k] |= 1u << (t - t0);
    const auto vj = S.v[t];
    dens_pair<R, SUM>(K, dx, dy, r2, vi.x - vj.x, vi.y - vj.y, o.sum, o.rs);
  });
  if (nbm) nbm[i] = make_uint4(rm[0], rm[1], rm[2], ok ? 1u : 0u);
  dens_end<R, SUM, EOS>(S, K, i, o);
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS, bool VI = false, bool WRAP = true>
__device__ __forceinline__ void force_direct(St<R>& S, const KParams& K, int i,
                                             const uint32_t* __restrict__ skey,
                                             const int* __restrict__ cstart) {
  const auto pi = load_pos(S, i);
  ForceOwn<R> o;
  force_begin<R, VI>(S, K, i, o);
  const R cut2 = cutoff2<R>(K);
  for_candidates<WRAP>(K, skey[i], cstart, [&](int t, int sh, int, int) {
    if (t == i) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    if (!(r2 < cut2)) return;
    const auto vj = S.v[t];
    R Qj, Aj;
    q_and_a(K, S.rho[t], S.p[t], Qj, Aj);
    force_pair<R, PRESS, AV, AH, XS, VI>(K, o, dx, dy, r2, vj.x, vj.y, Qj, Aj, S.psi[t],
                                         u_of(S, t),
                                         VI ? tau_of(S, t, visc_c(K, Qj)) : Tau3<R>{});
  });
  force_end<R, PRESS, XS, VI>(S, K, i, o);
}

// ---- per-particle smoothing length (K.sle != 0): the direct gathers with
// the pair constants of mh = (h_i + h_j)/2 (with_h), the neighbour test
// r^2 < (2 mh)^2, the self term and the art_heat coefficient of h_i.  No
// masks are recorded (the force pass tests the candidates again).
template <typename R, bool SUM, bool EOS>
__device__ __forceinline__ void density_direct_h(St<R>& S, const KParams& K, int i,
                                                 const uint32_t* __restrict__ skey,
                                                 const int* __restrict__ cstart,
                                                 uint4* __restrict__ nbm) {
  const auto pi = load_pos(S, i);
  const auto vi = S.v[i];
  const double hi = S.hs[i];
  const KParams Ki = with_h(K, hi);
  DensOwn<R> o;
  dens_begin<R, SUM>(Ki, o);
  for_candidates(K, skey[i], cstart, [&](int t, int sh, int, int) {
    if (t == i) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    const KParams Kp = with_h(K, 0.5 * (hi + S.hs[t]));
    if (!(r2 < cutoff2<R>(Kp))) return;
    const auto vj = S.v[t];
    dens_pair<R, SUM>(Kp, dx, dy, r2, vi.x - vj.x, vi.y - vj.y, o.sum, o.rs);
  });
  if (nbm) nbm[i] = make_uint4(0u, 0u, 0u, 0u);
  dens_end<R, SUM, EOS>(S, Ki, i, o);
}

template <typename R, bool SUM, bool EOS>
__global__ __launch_bounds__(BLOCK) void k_density(St<R> S, KParams K,
                                                   const uint32_t* __restrict__ skey,
                                                   const int* __restrict__ cstart,
                                                   uint4* __restrict__ nbm) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  if (K.sle)
    density_direct_h<R, SUM, EOS>(S, K, i, skey, cstart, nbm);
  else
    density_direct<R, SUM, EOS>(S, K, i, skey, cstart, nbm);
}

// Tait EOS (eos.f90:255-283) or the linear EOS on rate particles, then the
// art_heat coefficient from divv = (raw divergence sum) / rho.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_eos(St<R> S, KParams K) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  if (K.sle)
    eos_point(S, with_h(K, S.hs[i]), i, is_rate(K, S.id[i]), S.rho[i], S.divv[i]);
  else
    eos_point(S, K, i, is_rate(K, S.id[i]), S.rho[i], S.divv[i]);
}

// The new smoothing length of particle i from the force pass
// (hsml.f h_upgrade on real particles): sle 1 h = 2 sqrt(m/rho), sle 2
// h += dt (h/2) dvh with dvh = sum m/rho_j (v_j - v_i).grad W; a step that
// would make h non-positive is undone.  Counts in err[1] the particles whose
// 2h outgrew the cell.
template <typename R>
__device__ __forceinline__ void h_update(St<R>& S, const KParams& K, int i, R dvh,
                                         int* __restrict__ err) {
  const double h = S.hs[i];
  double hn = h;
  if (S.id[i] < K.np) {
    if (K.sle == 1) {
      hn = 2.0 * sqrt(K.mass / (double)dens(K, S.rho[i]));
    } else {
      const double dh = (h / 2.0) * (double)dvh;
      hn = h + K.dt * dh;
      if (hn <= 0.0) hn = hn - K.dt * dh;
    }
  }
  if (2.0 * hn > K.cell) {
    atomicAdd(err + 1, 1);
    raise_halt(K);
  }
  S.hn[i] = hn;
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS, bool VI>
__device__ __forceinline__ void force_direct_h(St<R>& S, const KParams& K, int i,
                                               const uint32_t* __restrict__ skey,
                                               const int* __restrict__ cstart,
                                               int* __restrict__ err) {
  const auto pi = load_pos(S, i);
  const double hi = S.hs[i];
  ForceOwn<R> o;
  force_begin<R, VI>(S, K, i, o);
  R dvh = 0;
  for_candidates(K, skey[i], cstart, [&](int t, int sh, int, int) {
    if (t == i) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    const KParams Kp = with_h(K, 0.5 * (hi + S.hs[t]));
    if (!(r2 < cutoff2<R>(Kp))) return;
    const auto vj = S.v[t];
    R Qj, Aj;
    q_and_a(K, S.rho[t], S.p[t], Qj, Aj);
    if constexpr (sizeof(R) == 4) own_factors_f32(Kp, o);  // the pair's h
    force_pair<R, PRESS, AV, AH, XS, VI>(Kp, o, dx, dy, r2, vj.x, vj.y, Qj, Aj, S.psi[t],
                                         u_of(S, t),
                                         VI ? tau_of(S, t, visc_c(K, Qj)) : Tau3<R>{});
    if (PRESS && K.sle == 2) {
      const R tq = fsqrt(r2) * (R)Kp.invh - (R)2;
      const R F = (R)Kp.C1 * (tq * tq * tq);
      const R vr = (o.v.x - vj.x) * dx + (o.v.y - vj.y) * dy;
      dvh = dvh - (R)K.mass * Qj * F * vr;
    }
  });
  force_end<R, PRESS, XS, VI>(S, K, i, o);
  if (PRESS) h_update(S, K, i, dvh, err);
}

template <typename R, bool PRESS, bool AV, bool AH, bool XS, bool VI>
__global__ __launch_bounds__(BLOCK) void k_force(St<R> S, KParams K,
                                                 const uint32_t* __restrict__ skey,
                                                 const int* __restrict__ cstart,
                                                 int* __restrict__ err) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  if (K.sle)
    force_direct_h<R, PRESS, AV, AH, XS, VI>(S, K, i, skey, cstart, err);
  else
    force_direct<R, PRESS, AV, AH, XS, VI>(S, K, i, skey, cstart);
}

// ------------------------------------------------------------------------
// shear tensor of the physical viscosity (internal_force.f:73-142), with
// dv = v_j - v_i, grad_i W = F (x_i - x_j) and the post-EOS densities:
//   T_xx = sum m/rho_j 2/3 (2 dvx dWx - dvy dWy)
//   T_yy = sum m/rho_j 2/3 (2 dvy dWy - dvx dWx)
//   T_xy = sum m/rho_j (dvx dWy + dvy dWx)
// then tds/dt = eta/(2 rho) (Txx^2 + Tyy^2 + 2 Txy^2) (:144-169).
// ------------------------------------------------------------------------
template <typename R> struct TensOwn {
  R xx = 0, yy = 0, xy = 0;
};

template <typename R>
__device__ __forceinline__ void tens_pair(const KParams& K, TensOwn<R>& o, R dx, R dy, R r2,
                                          R dvx, R dvy, R Qj) {
  const R q = fsqrt(r2) * (R)K.invh;
  const R tq = q - (R)2;
  const R F = (R)K.C1 * (tq * tq * tq);
  const R dwx = F * dx, dwy = F * dy;
  const R mq = (R)K.mass * Qj;
  const R a = dvx * dwx, b = dvy * dwy;
  const R hxx = (R)(2.0 / 3.0) * fmr((R)2, a, -b);
  const R hyy = (R)(2.0 / 3.0) * fmr((R)2, b, -a);
  const R hxy = fmr(dvy, dwx, dvx * dwy);
  o.xx = fmr(mq, hxx, o.xx);
  o.yy = fmr(mq, hyy, o.yy);
  o.xy = fmr(mq, hxy, o.xy);
}

template <typename R>
__device__ __forceinline__ void tens_end(St<R>& S, const KParams& K, int i, R Qi,
                                         const TensOwn<R>& o) {
  const R eta = (R)K.eta;
  const R tt = o.xx * o.xx + o.yy * o.yy + (R)2 * (o.xy * o.xy);
  R* w = S.tau + 4 * (size_t)i;
  w[0] = o.xx;
  w[1] = o.yy;
  w[2] = o.xy;
  w[3] = (R)0.5 * eta * Qi * tt;
}

template <typename R, bool WRAP = true>
__device__ __forceinline__ void tensor_direct(St<R>& S, const KParams& K, int i,
                                              const uint32_t* __restrict__ skey,
                                              const int* __restrict__ cstart) {
  const auto pi = load_pos(S, i);
  const auto vi = S.v[i];
  const R cut2 = cutoff2<R>(K);
  TensOwn<R> o;
  for_candidates<WRAP>(K, skey[i], cstart, [&](int t, int sh, int, int) {
    if (t == i) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    if (!(r2 < cut2)) return;
    const auto vj = S.v[t];
    tens_pair(K, o, dx, dy, r2, vj.x - vi.x, vj.y - vi.y, (R)1 / dens(K, S.rho[t]));
  });
  tens_end(S, K, i, (R)1 / dens(K, S.rho[i]), o);
}

template <typename R>
__device__ __forceinline__ void tensor_direct_h(St<R>& S, const KParams& K, int i,
                                                const uint32_t* __restrict__ skey,
                                                const int* __restrict__ cstart) {
  const auto pi = load_pos(S, i);
  const auto vi = S.v[i];
  const double hi = S.hs[i];
  TensOwn<R> o;
  for_candidates(K, skey[i], cstart, [&](int t, int sh, int, int) {
    if (t == i) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    const KParams Kp = with_h(K, 0.5 * (hi + S.hs[t]));
    if (!(r2 < cutoff2<R>(Kp))) return;
    const auto vj = S.v[t];
    tens_pair(Kp, o, dx, dy, r2, vj.x - vi.x, vj.y - vi.y, (R)1 / dens(K, S.rho[t]));
  });
  tens_end(S, K, i, (R)1 / dens(K, S.rho[i]), o);
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_tensor(St<R> S, KParams K,
                                                  const uint32_t* __restrict__ skey,
                                                  const int* __restrict__ cstart) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  if (K.sle)
    tensor_direct_h(S, K, i, skey, cstart);
  else
    tensor_direct(S, K, i, skey, cstart);
}

// ------------------------------------------------------------------------
// leapfrog (time_integration.f:136-154 predictor, :186-217 corrector)
// ------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_predictor(St<R> S, KParams K) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  const int id = S.id[i];
  const double hdt = K.dt / 2.0;
  S.flag[i] = 0;
  if (is_rate(K, id)) {
    double u = S.u[i];
    S.u_min[i] = u;
    u = u + hdt * (double)S.du[i];
    if (u < 0.0) {
      u = 0.0;
      S.flag[i] = 1u;
    }
    S.u[i] = u;
    if constexpr (sizeof(R) == 4) S.uf[i] = (float)u;
    if (K.density_method == 0) {
      const R rho = S.rho[i];
      S.rho_min[i] = rho;
      S.rho[i] = rho + (R)hdt * S.drho[i];
    }
  }
  if (id < K.np) {
    auto v = S.v[i];
    const auto a = S.a[i];
    S.v_min[i] = v;
    v.x = v.x + (R)hdt * a.x;
    v.y = v.y + (R)hdt * a.y;
    S.v[i] = v;
  }
}

// Corrector; with keys it also hashes the updated position for the next
// step's sort (the reorder then rebuilds the FP32 cell + offset).
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_corrector(St<R> S, KParams K, int its, int keys,
                                                     KeyOut ko, int* __restrict__ err) {
  if (halted(K)) return;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = i < K.n;
  uint32_t key = 0;
  int id = 0;
  if (valid) {
    id = S.id[i];
    const double dt = K.dt, hdt = K.dt / 2.0;
    double2 q = S.xy[i];
    if (is_rate(K, id)) {
      double u;
      if (its != 1) {
        u = S.u_min[i] + dt * (double)S.du[i];
        if (K.density_method == 0) S.rho[i] = S.rho_min[i] + (R)dt * S.drho[i];
      } else {
        u = S.u[i] + hdt * (double)S.du[i];
        if (K.density_method == 0) S.rho[i] = S.rho[i] + (R)hdt * S.drho[i];
      }
      if (u < 0.0) u = 0.0;
      S.u[i] = u;
      if constexpr (sizeof(R) == 4) S.uf[i] = (float)u;
      if (id < K.np) {
        auto v = S.v[i];
        const auto a = S.a[i];
        if (its != 1) {
          const auto vm = S.v_min[i];
          const auto av = S.av[i];
          v.x = vm.x + (R)dt * a.x + av.x;
          v.y = vm.y + (R)dt * a.y + av.y;
        } else {
          v.x = v.x + (R)hdt * a.x;
          v.y = v.y + (R)hdt * a.y;
        }
        S.v[i] = v;
        q.x = q.x + dt * (double)v.x;
        q.y = q.y + dt * (double)v.y;
        if (K.periodic) {  // re-enter on the other side of the period
          const double L = (double)K.nx * K.cell;
          if (q.x >= K.ox + L)
            q.x = q.x - L;
          else if (q.x < K.ox)
            q.x = q.x + L;
        }
        S.xy[i] = q;
      }
    }
    if (keys && id >= K.nb) {
      key = (uint32_t)K.ncell;  // this step's ghost: out of the next sort
    } else if (keys) {
      int ix, iy;
      bool ok;
      key = (uint32_t)cell_of(K, q.x, q.y, ix, iy, ok);
      if (!ok) {
        atomicAdd(err, 1);
        raise_halt(K);
      }
    }
  }
  if (keys) emit_key(ko, valid, i, id, key);
}

// ------------------------------------------------------------------------
// pair census for arm_link_list / arm_get_pairs: neighbours with larger id
// ------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_pair_count(St<R> S, KParams K,
                                                      const uint32_t* __restrict__ skey,
                                                      const int* __restrict__ cstart,
                                                      int* __restrict__ cnt,
                                                      const int* __restrict__ offs,
                                                      int* __restrict__ out_i,
                                                      int* __restrict__ out_j) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  const auto pi = load_pos(S, i);
  const int idi = S.id[i];
  const R cut2 = cutoff2<R>(K);
  int c = 0;
  int o = offs ? offs[i] : 0;
  for_candidates(K, skey[i], cstart, [&](int t, int sh, int, int) {
    const int idj = S.id[t];
    if (idj <= idi) return;
    const auto pj = load_pos(S, t);
    R dx, dy;
    delta(pi, pj, K, dx, dy, sh);
    const R r2 = dx * dx + dy * dy;
    const R cut = K.sle ? cutoff2<R>(with_h(K, 0.5 * (S.hs[i] + S.hs[t]))) : cut2;
    if (!(r2 < cut)) return;
    if (offs) {
      out_i[o + c] = idi;
      out_j[o + c] = idj;
    }
    ++c;
  });
  if (cnt) cnt[i] = c;
}

// ------------------------------------------------------------------------
// FP64 conservation partial sums over real particles, one row per block
// ------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_conserve(St<R> S, KParams K, double* __restrict__ part) {
  __shared__ double sh[5][BLOCK];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double v[5] = {0, 0, 0, 0, 0};
  if (i < K.n && S.id[i] < K.np) {
    const auto vel = S.v[i];
    const double vx = (double)vel.x, vy = (double)vel.y;
    v[0] = K.mass;
    v[1] = K.mass * vx;
    v[2] = K.mass * vy;
    v[3] = 0.5 * K.mass * (vx * vx + vy * vy);
    v[4] = K.mass * S.u[i];
  }
  for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) part[blockIdx.x * 5 + k] = sh[k][0];
}

}  // namespace arm
