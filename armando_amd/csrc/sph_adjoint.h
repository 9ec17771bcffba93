// sph_adjoint.h -- discrete adjoint of the FP64 step (reverse mode).
//
// Differentiates the step exactly as the forward kernels compute it:
//   density   drho_i = sum_j m F vr,  S_i = -sum_j m F vr (divv_i = S_i/rho_i)
//   EOS       p = Tait(rho), clamp p < pmin -> p = 0, rho = rho0
//   psi_i = g1 h c0 + g2 h^2 (|divv_i| - divv_i),  Q = 1/rho,  A = p/rho
//   force     G = m F (A_i Q_j + A_j Q_i + Pi),  a_i = -sum G r + g,
//             du_i = sum G vr/2 + H (u_i - u_j),  av_i = -sum eps m M W dv
//   leapfrog  time_integration.f:136-217 (predictor, corrector, v += av)
// with r = x_i - x_j, vr = (v_i - v_j).r, grad W = F r, M = 2QiQj/(Qi+Qj).
// The neighbour list and the branches (u < 0, p < pmin, vr < 0, |divv|) of
// every step are those of the forward step.
//
// Adjoint variables of the state live in id order (they survive the
// per-step reordering); per-step scratch lives in the step's sorted order.
// Pair terms are gathered per particle: a particle accumulates only the
// derivatives of its own inputs, each pair being visited from both ends, so
// no atomics are needed.  The objective is J = sum_steps dt sum_obj f,
// f = p^2 or u^2; the heat source (flag 64) adds dJ/d hs_q0.
#pragma once

#include "sph_kernels.h"

namespace arm {

// one recorded step: the sorted state after the force pass (before the
// corrector), the keys and the cell ranges of that step
struct Tape {
  int32_t* id;
  double2* xy;
  double2* v;   // predicted (half-step) velocity
  double* rho;  // density after the EOS clamp
  double* u;    // predicted energy
  double* psi;
  double* divv;
  double* p;
  double* drho;
  double* du;
  double2* a;
  double2* av;
  double* rho_min;
  double* u_min;
  double2* v_min;
  uint8_t* flag;
  uint32_t* skey;
  int* cstart;
};

// adjoint variables in id order
struct AdjState {
  double *x, *y, *vx, *vy, *rho, *u;  // of the state after the step (in) ...
  double *bx, *by, *bvx, *bvy, *brho, *bu;  // ... and before it (out)
  double *cax, *cay, *cdrho, *cdu;  // of the rates carried into the next predictor
};

// per-step scratch in sorted order
struct AdjScratch {
  double *lax, *lay, *ldrho, *ldu, *lavx, *lavy, *lp;
  double *lQ, *lA, *lpsi, *lS, *lrhoh, *lu, *lvx, *lvy, *lx, *ly;
  double* lus;  // objective seed of the energy (ARM_OBJ_ENERGY)
};

__global__ __launch_bounds__(BLOCK) void k_tape(St<double> S, KParams K,
                                                const uint32_t* __restrict__ skey,
                                                const int* __restrict__ cstart, Tape T) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const int stride = gridDim.x * BLOCK;
  for (int c = t; c <= K.ncell; c += stride) T.cstart[c] = cstart[c];
  if (t >= K.n) return;
  T.id[t] = S.id[t];
  T.xy[t] = S.xy[t];
  T.v[t] = S.v[t];
  T.rho[t] = S.rho[t];
  T.u[t] = S.u[t];
  T.psi[t] = S.psi[t];
  T.divv[t] = S.divv[t];
  T.p[t] = S.p[t];
  T.drho[t] = S.drho[t];
  T.du[t] = S.du[t];
  T.a[t] = S.a[t];
  T.av[t] = S.av[t];
  T.rho_min[t] = S.rho_min[t];
  T.u_min[t] = S.u_min[t];
  T.v_min[t] = S.v_min[t];
  T.flag[t] = S.flag[t];
  T.skey[t] = skey[t];
}

// objective: jpart[block] += dt sum_{obj} f over the block's particles,
// f = p^2 (kind 0) or u^2 (kind 1)
__global__ __launch_bounds__(BLOCK) void k_objective(St<double> S, KParams K, int kind,
                                                     const uint8_t* __restrict__ objmask,
                                                     double* __restrict__ jpart) {
  __shared__ double sh[BLOCK];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double v = 0.0;
  if (i < K.n && objmask[S.id[i]]) {
    const double f = kind == 1 ? S.u[i] : S.p[i];
    v = K.dt * f * f;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) jpart[blockIdx.x] += sh[0];
}

// 1. corrector adjoint (time_integration.f:186-217): state adjoints after
//    the step -> rate adjoints (sorted) and the direct part of the state
//    adjoints before the step (id order); the objective, scaled by the
//    incoming adjoint seed = dL/dJ, seeds lp.
__global__ __launch_bounds__(BLOCK) void k_adj_corrector(Tape T, KParams K, int its,
                                                         AdjState L, AdjScratch W,
                                                         const uint8_t* __restrict__ objmask,
                                                         int kind, double seed) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const int id = T.id[k];
  const double dt = K.dt, h2 = K.dt / 2.0;
  double lax = L.cax[id], lay = L.cay[id], ldrho = L.cdrho[id], ldu = L.cdu[id];
  double lavx = 0, lavy = 0;
  double bx = L.x[id], by = L.y[id], bvx = L.vx[id], bvy = L.vy[id];
  double brho = 0, bu = 0;
  if (!is_rate(K, id)) {
    brho = L.rho[id];
    bu = L.u[id];
  } else {
    const double c = (its != 1) ? dt : h2;
    brho = L.rho[id];
    ldrho += c * L.rho[id];
    const double ub = (its != 1) ? T.u_min[k] : T.u[k];  // its == 1: u is the state u
    const double uu = ub + c * T.du[k];
    if (!(uu < 0.0)) {
      bu = L.u[id];
      ldu += c * L.u[id];
    }
    if (id < K.np) {
      const double lvtx = L.vx[id] + dt * L.x[id];
      const double lvty = L.vy[id] + dt * L.y[id];
      bvx = lvtx;
      bvy = lvty;
      lax += c * lvtx;
      lay += c * lvty;
      if (its != 1) {
        lavx = lvtx;
        lavy = lvty;
      }
    }
  }
  L.bx[id] = bx;
  L.by[id] = by;
  L.bvx[id] = bvx;
  L.bvy[id] = bvy;
  L.brho[id] = brho;
  L.bu[id] = bu;
  W.lax[k] = lax;
  W.lay[k] = lay;
  W.ldrho[k] = is_rate(K, id) ? ldrho : 0.0;
  W.ldu[k] = ldu;
  W.lavx[k] = lavx;
  W.lavy[k] = lavy;
  W.lp[k] = (objmask[id] && kind == 0) ? seed * 2.0 * dt * T.p[k] : 0.0;
  W.lus[k] = (objmask[id] && kind == 1) ? seed * 2.0 * dt * T.u[k] : 0.0;
}

__device__ __forceinline__ void wendland_q(const KParams& K, double r2, double& q, double& F,
                                           double& Wv) {
  q = sqrt(r2) * K.invh;
  const double t = q - 2.0;
  F = K.C1 * (t * t * t);
  const double s = 1.0 - 0.5 * q;
  Wv = K.wnorm * (s * s) * (s * s) * (1.0 + 2.0 * q);
}

__device__ __forceinline__ Rows tape_rows(const KParams& K, const Tape& T, int k) {
  return rows_of(K, T.skey[k], T.cstart);
}

// 2. force-pass adjoint, gathered: particle k (as i) against every
//    neighbour l (as j); accumulates k's own input adjoints.
template <bool AV, bool AH, bool XS>
__global__ __launch_bounds__(BLOCK) void k_adj_force(Tape T, KParams K, AdjScratch W,
                                                     double* __restrict__ alpha_part,
                                                     double* __restrict__ q0_part) {
  __shared__ double sh[BLOCK];
  double lq0 = 0.0;
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  double lal = 0.0;
  if (k < K.n) {
    const double m = K.mass, h = K.h;
    const double2 xk = T.xy[k], vk = T.v[k];
    const double Qk = 1.0 / T.rho[k], Ak = T.p[k] * Qk;
    const double lak_x = W.lax[k], lak_y = W.lay[k], lduk = W.ldu[k];
    const double lavk_x = W.lavx[k], lavk_y = W.lavy[k];
    const double psik = T.psi[k], uk = T.u[k];
    double lx = 0, ly = 0, lvx = 0, lvy = 0, lQ = 0, lA = 0, lpsi = 0, lu = 0;
    const Rows rw = tape_rows(K, T, k);
    for (int r = 0; r < 3; ++r)
      for (int l = rw.t0[r]; l < rw.t1[r]; ++l) {
        if (l == k) continue;
        const double2 xl = T.xy[l];
        const double rx = xk.x - xl.x, ry = xk.y - xl.y;
        const double r2 = rx * rx + ry * ry;
        if (!(r2 < K.cut2)) continue;
        double q, F, Wv;
        wendland_q(K, r2, q, F, Wv);
        const double2 vl = T.v[l];
        const double dvx = vk.x - vl.x, dvy = vk.y - vl.y;
        const double vr = dvx * rx + dvy * ry;
        const double Ql = 1.0 / T.rho[l], Al = T.p[l] * Ql;
        const double sQ = Qk + Ql;
        const double M = 2.0 * Qk * Ql / sQ;
        double Kc = (K.pa_sph == 0) ? (Ak * Ql + Al * Qk) : (Ak * Qk + Al * Ql);
        const bool av_on = AV && vr < 0.0;
        const double einv = 1.0 / (r2 + K.av_e2);
        const double phi = h * vr * einv;
        if (av_on) Kc += (K.av_beta * phi - K.av_alpha * K.c0) * phi * M;
        const double G = m * F * Kc;
        const double Dax = lak_x - W.lax[l], Day = lak_y - W.lay[l];
        const double sig = lduk + W.ldu[l];
        const double lG = -(Dax * rx + Day * ry) + 0.5 * sig * vr;
        double lrx = -G * Dax, lry = -G * Day;
        double lvr = 0.5 * sig * G;
        double ldvx = 0, ldvy = 0;
        double lF = lG * m * Kc;
        const double lK = lG * m * F;
        double lM = 0, lr2 = 0, lW = 0;
        if (K.pa_sph == 0) {
          lA += lK * Ql;
          lQ += lK * Al;
        } else {
          lA += lK * Qk;
          lQ += lK * Ak;
        }
        if (av_on) {
          const double dPi = (2.0 * K.av_beta * phi - K.av_alpha * K.c0) * M;
          lM += lK * (K.av_beta * phi - K.av_alpha * K.c0) * phi;
          lal += lK * (-K.c0 * phi * M);
          const double lphi = lK * dPi;
          lvr += lphi * h * einv;
          lr2 += lphi * (-phi * einv);
        }
        if (AH) {
          const double eh = 1.0 / (r2 + K.ah_e2);
          const double du_ = uk - T.u[l];
          const double dl = lduk - W.ldu[l];
          const double ps = psik + T.psi[l];
          const double H = m * ps * F * r2 * M * eh;
          const double lH = dl * du_;
          lu += dl * H;
          lpsi += lH * m * F * r2 * M * eh;
          lF += lH * m * ps * r2 * M * eh;
          lr2 += lH * m * ps * F * M * K.ah_e2 * eh * eh;
          lM += lH * m * ps * F * r2 * eh;
        }
        if (XS) {
          const double Dvx = lavk_x - W.lavx[l], Dvy = lavk_y - W.lavy[l];
          const double w = K.xsph_eps * m * M * Wv;
          const double lw = -(Dvx * dvx + Dvy * dvy);
          ldvx += -w * Dvx;
          ldvy += -w * Dvy;
          lM += lw * K.xsph_eps * m * Wv;
          lW += lw * K.xsph_eps * m * M;
        }
        lQ += lM * 2.0 * Ql * Ql / (sQ * sQ);
        const double s = 1.0 - 0.5 * q;
        const double lq = lF * 3.0 * K.C1 * (q - 2.0) * (q - 2.0) + lW * K.wnorm * s * s * s * (-5.0 * q);
        lr2 += lq / (2.0 * h * h * q);
        ldvx += lvr * rx;
        ldvy += lvr * ry;
        lrx += lvr * dvx + 2.0 * lr2 * rx;
        lry += lvr * dvy + 2.0 * lr2 * ry;
        lx += lrx;
        ly += lry;
        lvx += ldvx;
        lvy += ldvy;
      }
    // heat source du_k += q0 E_k Q_k, E = exp(-hs_k |x_k - c|^2)
    if ((K.flags & 64) && T.id[k] < K.np) {
      const double E = heat_shape(K, xk.x, xk.y);
      lq0 = lduk * E * Qk;
      lQ += lduk * K.hs_q0 * E;
      const double lE = lduk * K.hs_q0 * Qk * E * (-2.0 * K.hs_k);
      lx += lE * (xk.x - K.hs_x);
      ly += lE * (xk.y - K.hs_y);
    }
    W.lx[k] = lx;
    W.ly[k] = ly;
    W.lvx[k] = lvx;
    W.lvy[k] = lvy;
    W.lQ[k] = lQ;
    W.lA[k] = lA;
    W.lpsi[k] = lpsi;
    W.lu[k] = lu + W.lus[k];
  }
  // each pair is visited from both ends: half of its alpha derivative each
  sh[threadIdx.x] = 0.5 * lal;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) alpha_part[blockIdx.x] += sh[0];
  __syncthreads();
  sh[threadIdx.x] = lq0;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) q0_part[blockIdx.x] += sh[0];
}

// 3. EOS and art_heat coefficient adjoint (per particle)
__global__ __launch_bounds__(BLOCK) void k_adj_eos(Tape T, KParams K, AdjScratch W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const double rho = T.rho[k];
  const double Q = 1.0 / rho;
  double lp = W.lp[k] + W.lA[k] * Q;
  double lrhoe = -Q * Q * (W.lQ[k] + W.lA[k] * T.p[k]);
  const double dv = T.divv[k];  // S / rho
  const double dpsi = (dv < 0.0) ? -2.0 * K.g2h2 : 0.0;
  const double ldv = W.lpsi[k] * dpsi;
  W.lS[k] = ldv / rho;
  lrhoe += -ldv * dv / rho;
  double lrhoh = 0.0;
  if (is_rate(K, T.id[k]) && !(T.flag[k] & 2u)) {
    const double t = rho / K.rho0;
    const double t2 = t * t;
    lrhoh = lrhoe + lp * 7.0 * K.B / K.rho0 * (t2 * t2 * t2);
  }
  W.lrhoh[k] = lrhoh;
}

// 4. density adjoint, gathered; adds to the position / velocity adjoints
__global__ __launch_bounds__(BLOCK) void k_adj_density(Tape T, KParams K, AdjScratch W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const double m = K.mass, h = K.h;
  const double2 xk = T.xy[k], vk = T.v[k];
  const double ltk = W.ldrho[k] - W.lS[k];
  double lx = 0, ly = 0, lvx = 0, lvy = 0;
  const Rows rw = tape_rows(K, T, k);
  for (int r = 0; r < 3; ++r)
    for (int l = rw.t0[r]; l < rw.t1[r]; ++l) {
      if (l == k) continue;
      const double2 xl = T.xy[l];
      const double rx = xk.x - xl.x, ry = xk.y - xl.y;
      const double r2 = rx * rx + ry * ry;
      if (!(r2 < K.cut2)) continue;
      const double lt = ltk + W.ldrho[l] - W.lS[l];
      double q, F, Wv;
      wendland_q(K, r2, q, F, Wv);
      const double2 vl = T.v[l];
      const double dvx = vk.x - vl.x, dvy = vk.y - vl.y;
      const double vr = dvx * rx + dvy * ry;
      const double lvr = lt * m * F;
      const double lF = lt * m * vr;
      const double lr2 = lF * 3.0 * K.C1 * (q - 2.0) * (q - 2.0) / (2.0 * h * h * q);
      lx += lvr * dvx + 2.0 * lr2 * rx;
      ly += lvr * dvy + 2.0 * lr2 * ry;
      lvx += lvr * rx;
      lvy += lvr * ry;
    }
  W.lx[k] += lx;
  W.ly[k] += ly;
  W.lvx[k] += lvx;
  W.lvy[k] += lvy;
}

// 5. predictor adjoint (time_integration.f:136-154): folds the rate-input
//    adjoints into the state adjoints before the step and into the carried
//    rate adjoints of the previous step
__global__ __launch_bounds__(BLOCK) void k_adj_predictor(Tape T, KParams K, int its, AdjState L,
                                                         AdjScratch W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const int id = T.id[k];
  const double h2 = K.dt / 2.0;
  L.bx[id] += W.lx[k];
  L.by[id] += W.ly[k];
  L.bvx[id] += W.lvx[k];
  L.bvy[id] += W.lvy[k];
  L.brho[id] += W.lrhoh[k];
  double cax = 0, cay = 0, cdrho = 0, cdu = 0;
  if (its != 1 && is_rate(K, id)) {
    cdrho = h2 * W.lrhoh[k];
    if (!(T.flag[k] & 1u)) {
      L.bu[id] += W.lu[k];
      cdu = h2 * W.lu[k];
    }
    if (id < K.np) {
      cax = h2 * W.lvx[k];
      cay = h2 * W.lvy[k];
    }
  } else {
    L.bu[id] += W.lu[k];
  }
  L.cax[id] = cax;
  L.cay[id] = cay;
  L.cdrho[id] = cdrho;
  L.cdu[id] = cdu;
}

}  // namespace arm
