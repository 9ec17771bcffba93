// sph_kernels.h -- HIP kernels of the WCSPH step for gfx950.
//
// Every pair pass is a gather: one lane per particle i (sorted order), which
// walks the three cell rows around its cell.  Because keys are ix + iy*nx and
// particles are sorted by (cell, id), the three cells of a row are one
// contiguous range [cell_start[c-1], cell_start[c+2]) of the sorted arrays,
// so each lane reads three contiguous runs.  Each lane writes only its own
// particle: no atomics, bitwise repeatable.
#pragma once

#include "sph_device.h"

namespace arm {

constexpr int BLOCK = 256;

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

// Key of the particle in slot s, written at its id (the sort input is listed
// in id order so that a stable sort yields (cell, id) order).  FP32 form:
// also refreshes the integer cell and FP32 offset from the FP64 accumulator.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_keys(St<R> S, KParams K, uint32_t* __restrict__ key_by_id,
                                                uint32_t* __restrict__ slot_by_id,
                                                int* __restrict__ err) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= K.n) return;
  const double2 q = S.xy[s];
  int ix, iy;
  bool ok;
  const int key = cell_of(K, q.x, q.y, ix, iy, ok);
  if (!ok) atomicAdd(err, 1);
  const int id = S.id[s];
  key_by_id[id] = (uint32_t)key;
  slot_by_id[id] = (uint32_t)s;
  if constexpr (sizeof(R) == 4) {
    const double bx = K.ox + (double)ix * K.cell;
    const double by = K.oy + (double)iy * K.cell;
    S.cxy[s] = (uint32_t)ix | ((uint32_t)iy << 16);
    S.off[s] = make_float2((float)(q.x - bx), (float)(q.y - by));
  }
}

// ------------------------------------------------------------------------
// cell_start from sorted keys: cell_start[c] = first sorted position with
// key >= c, for c in [0, ncell]; each lane fills the cells between its
// predecessor's key and its own.
// ------------------------------------------------------------------------
__device__ __forceinline__ void fill_cell_start(const uint32_t* __restrict__ skey, int t, int n,
                                                int ncell, int* __restrict__ cstart) {
  const int k = (int)skey[t];
  const int kp = (t == 0) ? -1 : (int)skey[t - 1];
  for (int c = kp + 1; c <= k; ++c) cstart[c] = t;
  if (t == n - 1)
    for (int c = k + 1; c <= ncell; ++c) cstart[c] = n;
}

// Permute every field from state A to state B (per-routine link_list).
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_reorder_all(St<R> A, St<R> Bs, KParams K,
                                                       const uint32_t* __restrict__ skey,
                                                       const uint32_t* __restrict__ sval,
                                                       int* __restrict__ cstart) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= K.n) return;
  const int s = (int)sval[t];
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
  fill_cell_start(skey, t, K.n, K.ncell, cstart);
}

// Fused reorder + leapfrog predictor (time_integration.f:136-154): the
// predictor does not move particles, so sorting before it is equivalent.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_reorder_pred(St<R> A, St<R> Bs, KParams K,
                                                        const uint32_t* __restrict__ skey,
                                                        const uint32_t* __restrict__ sval,
                                                        int* __restrict__ cstart, int predict) {
  using R2 = typename V2<R>::type;
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= K.n) return;
  const int s = (int)sval[t];
  const int id = A.id[s];
  Bs.id[t] = id;
  Bs.xy[t] = A.xy[s];
  if constexpr (sizeof(R) == 4) {
    Bs.cxy[t] = A.cxy[s];
    Bs.off[t] = A.off[s];
  }
  R2 v = A.v[s];
  R rho = A.rho[s];
  double u = A.u[s];
  if (predict) {
    const double hdt = K.dt / 2.0;
    const bool rate = is_rate(K, id);
    if (rate) {
      Bs.u_min[t] = u;
      u = u + hdt * (double)A.du[s];
      if (u < 0.0) u = 0.0;
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
  Bs.p[t] = A.p[s];
  fill_cell_start(skey, t, K.n, K.ncell, cstart);
}

// ------------------------------------------------------------------------
// neighbour walk helpers
// ------------------------------------------------------------------------
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

// ------------------------------------------------------------------------
// density pass: con_density (density.f:166-223) and, for method 1,
// sum_density (density.f:1-62), gathered.  Also the raw velocity
// divergence sum of art_heat's first pass (art_heat.f:46-67): with a
// uniform mass it is the negative of the continuity sum.
// ------------------------------------------------------------------------
template <typename R, bool SUMMATION>
__global__ __launch_bounds__(BLOCK) void k_density(St<R> S, KParams K,
                                                   const uint32_t* __restrict__ skey,
                                                   const int* __restrict__ cstart) {
  using R2 = typename V2<R>::type;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  const auto pi = load_pos(S, i);
  const R2 vi = S.v[i];
  const Rows rw = rows_of(K, skey[i], cstart);
  const R cut2 = cutoff2<R>(K);
  const R invh = (R)K.invh, C1 = (R)K.C1, m = (R)K.mass, wn = (R)K.wnorm;
  Acc<R> sum;
  Acc<R> rs;
  if (SUMMATION) rs.add(m * wn);
  for (int k = 0; k < 3; ++k) {
    for (int t = rw.t0[k]; t < rw.t1[k]; ++t) {
      if (t == i) continue;
      const auto pj = load_pos(S, t);
      R dx, dy;
      delta(pi, pj, K, dx, dy);
      const R r2 = dx * dx + dy * dy;
      if (!(r2 < cut2)) continue;
      const R q = fsqrt(r2) * invh;
      const R tq = q - (R)2;
      const R F = C1 * (tq * tq * tq);
      const R2 vj = S.v[t];
      const R dwx = F * dx, dwy = F * dy;
      R vcc = (vi.x - vj.x) * dwx;
      vcc = vcc + (vi.y - vj.y) * dwy;
      sum.add(m * vcc);
      if (SUMMATION) {
        const R s = (R)1 - (R)0.5 * q;
        const R s2 = s * s;
        rs.add(m * (wn * (s2 * s2) * ((R)1 + (R)2 * q)));
      }
    }
  }
  const int id = S.id[i];
  const bool rate = is_rate(K, id);
  if (!SUMMATION) S.drho[i] = rate ? sum.get() : (R)0;
  else {
    S.drho[i] = (R)0;
    if (rate) {
      if constexpr (sizeof(R) == 4)
        S.rho[i] = rs.get() - (float)K.rho0;
      else
        S.rho[i] = rs.get();
    }
  }
  S.divv[i] = -sum.get();
}

// Tait EOS (eos.f90:255-283) on rate particles, then the art_heat
// coefficient from divv = (raw divergence sum) / rho.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_eos(St<R> S, KParams K) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  R rho = S.rho[i];
  if (is_rate(K, S.id[i])) {
    R p = tait(K, rho);
    if (p < (R)K.pmin) {
      p = (R)0;
      rho = (sizeof(R) == 4) ? (R)0 : (R)K.rho0;
      S.rho[i] = rho;
    }
    S.p[i] = p;
  }
  const R dv = S.divv[i] / dens(K, rho);
  S.divv[i] = dv;
  const R ad = dv < (R)0 ? -dv : dv;
  S.psi[i] = (R)K.psi0 + (R)K.g2h2 * (ad - dv);
}

// ------------------------------------------------------------------------
// force pass: int_force pressure part (internal_force.f:178-259, 309-430),
// art_visc (art_visc.f:55-101), art_heat second pass (art_heat.f:69-97)
// and av_vel (av_vel.f:37-56), fused over one neighbour walk.
//   PRESS: pressure force and work, plus gravity on real particles
//   AV, AH, XS: artificial viscosity, artificial heat, XSPH
// Outputs are written (not accumulated): a and du when PRESS, av when XS.
// ------------------------------------------------------------------------
template <typename R, bool PRESS, bool AV, bool AH, bool XS>
__global__ __launch_bounds__(BLOCK) void k_force(St<R> S, KParams K,
                                                 const uint32_t* __restrict__ skey,
                                                 const int* __restrict__ cstart) {
  using R2 = typename V2<R>::type;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  const auto pi = load_pos(S, i);
  const R2 vi = S.v[i];
  const R rhoi = dens(K, S.rho[i]);
  const R p_i = S.p[i];
  const R psii = S.psi[i];
  R ui;
  if constexpr (sizeof(R) == 4) ui = S.uf[i]; else ui = (R)S.u[i];
  const Rows rw = rows_of(K, skey[i], cstart);
  const R cut2 = cutoff2<R>(K);
  const R invh = (R)K.invh, C1 = (R)K.C1, m = (R)K.mass, wn = (R)K.wnorm, h = (R)K.h;
  const R av_e2 = (R)K.av_e2, ah_e2 = (R)K.ah_e2, c0 = (R)K.c0;
  const R alpha = (R)K.av_alpha, beta = (R)K.av_beta, eps = (R)K.xsph_eps;
  const R pr_i = p_i / (rhoi * rhoi);
  R fx = 0, fy = 0, avfx = 0, avfy = 0, xsx = 0, xsy = 0;
  Acc<R> due, avdu, ahdu;
  for (int k = 0; k < 3; ++k) {
    for (int t = rw.t0[k]; t < rw.t1[k]; ++t) {
      if (t == i) continue;
      const auto pj = load_pos(S, t);
      R dx, dy;
      delta(pi, pj, K, dx, dy);
      const R r2 = dx * dx + dy * dy;
      if (!(r2 < cut2)) continue;
      const R q = fsqrt(r2) * invh;
      const R tq = q - (R)2;
      const R F = C1 * (tq * tq * tq);
      const R dwx = F * dx, dwy = F * dy;
      const R2 vj = S.v[t];
      const R dvx = vi.x - vj.x, dvy = vi.y - vj.y;
      R vdw = dvx * dwx;
      vdw = vdw + dvy * dwy;
      const R rhoj = dens(K, S.rho[t]);
      const R mrho = (R)0.5 * (rhoi + rhoj);
      if (PRESS) {
        const R pj_ = S.p[t];
        R a;
        if (K.pa_sph == 0)
          a = (p_i + pj_) / (rhoi * rhoj);
        else
          a = pr_i + pj_ / (rhoj * rhoj);
        fx = fx - m * a * dwx;
        fy = fy - m * a * dwy;
        due.add(m * a * vdw);
      }
      if (AV) {
        R vr = dvx * dx;
        vr = vr + dvy * dy;
        if (vr < (R)0) {
          const R phi = (h * vr) / (r2 + av_e2);
          const R piv = (beta * phi - alpha * c0) * phi / mrho;
          avfx = avfx - m * piv * dwx;
          avfy = avfy - m * piv * dwy;
          avdu.add((R)0.5 * m * piv * vdw);
        }
      }
      if (AH) {
        R rdw = dx * dwx;
        rdw = rdw + dy * dwy;
        const R pij = (R)0.5 * (psii + S.psi[t]);
        const R ps = ((R)2 * pij) / (mrho * (r2 + ah_e2)) * rdw;
        R uj;
        if constexpr (sizeof(R) == 4) uj = S.uf[t]; else uj = (R)S.u[t];
        ahdu.add(m * ps * (ui - uj));
      }
      if (XS) {
        const R s = (R)1 - (R)0.5 * q;
        const R s2 = s * s;
        const R w = wn * (s2 * s2) * ((R)1 + (R)2 * q);
        xsx = xsx - eps * m / mrho * dvx * w;
        xsy = xsy - eps * m / mrho * dvy * w;
      }
    }
  }
  if (PRESS) {
    const int id = S.id[i];
    const bool grav = (K.flags & 16) && id < K.np;
    const R ex = grav ? (R)K.gx : (R)0, ey = grav ? (R)K.gy : (R)0;
    R2 acc;
    acc.x = fx + ex + avfx;
    acc.y = fy + ey + avfy;
    S.a[i] = acc;
    S.du[i] = (R)0.5 * due.get() + avdu.get() + ahdu.get();
  }
  if (XS) {
    R2 w;
    w.x = xsx;
    w.y = xsy;
    S.av[i] = w;
  }
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
  if (is_rate(K, id)) {
    double u = S.u[i];
    S.u_min[i] = u;
    u = u + hdt * (double)S.du[i];
    if (u < 0.0) u = 0.0;
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

// Corrector; with keys != nullptr it also hashes the updated position for
// the next step's sort (and, FP32 form, refreshes cell + offset).
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_corrector(St<R> S, KParams K, int its,
                                                     uint32_t* __restrict__ key_by_id,
                                                     uint32_t* __restrict__ slot_by_id,
                                                     int* __restrict__ err) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= K.n) return;
  const int id = S.id[i];
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
      S.xy[i] = q;
    }
  }
  if (key_by_id) {
    int ix, iy;
    bool ok;
    const int key = cell_of(K, q.x, q.y, ix, iy, ok);
    if (!ok) atomicAdd(err, 1);
    key_by_id[id] = (uint32_t)key;
    slot_by_id[id] = (uint32_t)i;
    if constexpr (sizeof(R) == 4) {
      const double bx = K.ox + (double)ix * K.cell;
      const double by = K.oy + (double)iy * K.cell;
      S.cxy[i] = (uint32_t)ix | ((uint32_t)iy << 16);
      S.off[i] = make_float2((float)(q.x - bx), (float)(q.y - by));
    }
  }
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
  const Rows rw = rows_of(K, skey[i], cstart);
  const R cut2 = cutoff2<R>(K);
  int c = 0;
  int o = offs ? offs[i] : 0;
  for (int k = 0; k < 3; ++k)
    for (int t = rw.t0[k]; t < rw.t1[k]; ++t) {
      const int idj = S.id[t];
      if (idj <= idi) continue;
      const auto pj = load_pos(S, t);
      R dx, dy;
      delta(pi, pj, K, dx, dy);
      const R r2 = dx * dx + dy * dy;
      if (!(r2 < cut2)) continue;
      if (offs) {
        out_i[o + c] = idi;
        out_j[o + c] = idj;
      }
      ++c;
    }
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
