// sph_ghost.h -- the mirror ghosts of CERNsource/ghost.f:1-86 on gfx950.
//
// The virtual particles (ids np .. nb-1) carry the unit wall normal in v and
// never move.  Every step, before the sort (time_integration.f:157-161 calls
// ghost_particles after the predictor and before link_list):
//   k_ghost_mark    (once per upload) the cells near a virtual particle
//   k_ghost_select  each real particle takes, among the virtual particles
//                   within 2h of its position at the last sort (the pair
//                   list ghost_particles walks), the nearest one at its
//                   current position; ties go to the smaller id.  The
//                   virtual particles come from a fixed (cell, id) list built
//                   at upload.
//   (scan)          the flags by id give each ghost its number, so ghosts
//                   are numbered in the order of their parents, as ghost.f
//                   appends them
//   k_ghost_make    writes ghost g at staging slot n_old + g: its mirror
//                   position x - 2 n ((x - x_w).n), the parent and wall
//                   slots, and its key for the counting sort
// The ghosts of the previous step carry the extra key ncell and drop out of
// the sort; the reorder builds each new ghost's fields (ghost_slot) and,
// after the EOS, k_ghost_copy gives it the parent's rho, p and u
// (single_step.f:188-196).
#pragma once

#include "sph_kernels.h"

namespace arm {

// The cells whose 3x3 block holds a virtual particle (the walls never move,
// so this is set once per upload): only particles there can have
// candidates.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_ghost_mark(St<R> A, KParams K, Ghosts G, int n) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= n) return;
  const int id = A.id[s];
  if (id < K.np || id >= K.nb) return;
  const double2 q = A.xy[s];
  int ix, iy;
  bool ok;
  cell_of(K, q.x, q.y, ix, iy, ok);
  if (!ok) return;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int cx = ix + dx, cy = iy + dy;
      if (cx >= 0 && cx < K.nx && cy >= 0 && cy < K.ny) G.vnear[cy * K.nx + cx] = 1;
    }
}

// gsel[s]: the chosen virtual particle's index in the fixed (cell, id)
// list, or -1.  The cell of xprev[s] is the key slot s was sorted by
// (skey[s]: the slots keep the order of the last sort until the next one),
// so a particle away from the walls reads four bytes and leaves.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_ghost_select(St<R> A, KParams K, Ghosts G,
                                                        const uint32_t* __restrict__ skey,
                                                        int n_old) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= n_old) return;
  const int id = A.id[s];
  if (id >= K.np) return;
  const uint32_t key = skey[s];
  const double cut = (2.0 * K.h) * (2.0 * K.h);
  int best = -1, bid = 0;
  double bd = 0.0;
  if (G.vnear[key]) {
    const double2 xp = G.xprev[s];
    const double2 q = A.xy[s];
    const Rows rw = rows_of(K, key, G.vstart);
    for (int k = 0; k < 3; ++k)
      for (int t = rw.t0[k]; t < rw.t1[k]; ++t) {
        const int idt = G.vid[t];
        const double2 w = G.vxy[t];
        const double dx = xp.x - w.x, dy = xp.y - w.y;
        double r2 = dx * dx;
        r2 = r2 + dy * dy;
        if (!(r2 < cut)) continue;
        const double ex = w.x - q.x, ey = w.y - q.y;
        double dd = ex * ex;
        dd = dd + ey * ey;
        if (best < 0 || dd < bd || (dd == bd && idt < bid)) {
          best = t;
          bid = idt;
          bd = dd;
        }
      }
  }
  G.gsel[s] = best;
  G.gflag[id] = best >= 0 ? 1 : 0;
}

// Ghost g of the parent in slot s goes to staging slot n_old + g (written
// only below cap; *ng receives the count either way).  err[0] counts ghosts
// outside the hash grid (their key is the extra cell).  keys = 0: no key
// emission (the per-routine link_list hashes every slot itself).
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_ghost_make(St<R> A, KParams K, Ghosts G, int n_old,
                                                      int cap, int keys, KeyOut ko,
                                                      int* __restrict__ err,
                                                      int* __restrict__ ng) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s == 0) *ng = G.grank[K.np - 1] + G.gflag[K.np - 1];
  int id = 0, slot = 0;
  uint32_t key = 0xffffffffu;
  bool valid = false;
  if (s < n_old) {
    id = A.id[s];
    if (id < K.np && G.gsel[s] >= 0) {
      const int g = G.grank[id];
      slot = n_old + g;
      valid = slot < cap;
      if (valid) {
        const int t = G.gsel[s];
        const double2 q = A.xy[s];
        const double2 w = G.vxy[t];
        const double nx = G.vnrm[t].x, ny = G.vnrm[t].y;
        double d1 = 0.0;
        d1 = d1 + (q.x - w.x) * nx;
        d1 = d1 + (q.y - w.y) * ny;
        double2 gq;
        gq.x = q.x - 2.0 * nx * d1;
        gq.y = q.y - 2.0 * ny * d1;
        A.xy[slot] = gq;
        A.id[slot] = K.nb + g;
        G.gpar[slot] = s;
        G.gwall[slot] = t;
        int ix, iy;
        bool ok;
        key = (uint32_t)cell_of(K, gq.x, gq.y, ix, iy, ok);
        if (!ok) {
          atomicAdd(err, 1);
          key = (uint32_t)K.ncell;
        }
        id = K.nb + g;
      }
    }
  }
  if (keys) emit_key(ko, valid, slot, id, key);
}

// After the EOS: every ghost takes rho, p and u of its parent; its own
// velocity divergence (a sum over its neighbours, divided by its density)
// and art_heat coefficient follow the copied density.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_ghost_copy(St<R> S, KParams K, Ghosts G) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= K.n) return;
  if (S.id[t] < K.nb) return;
  const int ps = G.sid[G.gpid[t]];
  const R r_old = S.rho[t], r = S.rho[ps];
  S.rho[t] = r;
  S.p[t] = S.p[ps];
  S.u[t] = S.u[ps];
  if constexpr (sizeof(R) == 4) S.uf[t] = S.uf[ps];
  if (r != r_old) {
    const R dv = S.divv[t] * dens(K, r_old) / dens(K, r);
    S.divv[t] = dv;
    const R ad = dv < (R)0 ? -dv : dv;
    S.psi[t] = (R)K.psi0 + (R)K.g2h2 * (ad - dv);
  }
}

}  // namespace arm
