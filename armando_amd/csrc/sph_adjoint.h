// sph_adjoint.h -- discrete adjoint of the step (reverse mode), for the FP64
// plain form (R = double) and the FP32 chosen form (R = float: pair
// arithmetic in FP32 on the cell + offset positions, as the forward; the
// adjoint of the state is accumulated in FP64 like the forward's position
// and energy accumulators).
//
// Differentiates the step exactly as the forward kernels compute it:
//   density   drho_i = sum_j m F vr,  S_i = -sum_j m F vr (divv_i = S_i/rho_i)
//   EOS       p = Tait(rho) or c0^2 rho + p0, clamp p < pmin -> p = 0, rho = rho0
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
// f = p^2 or u^2; the heat source (flag 64) adds dJ/d hs_q0 and the
// physical viscosity (flag 32: shear tensor, viscous force, tds/dt) dJ/d eta.
#pragma once

#include "sph_kernels.h"

namespace arm {

// one recorded step: the sorted state after the force pass (before the
// corrector), the keys and the cell ranges of that step
template <typename R> struct Tape {
  using R2 = typename V2<R>::type;
  int32_t* id;
  double2* xy;
  uint32_t* cxy;  // FP32 form: cell and in-cell offset of the pair arithmetic
  float2* off;
  R2* v;      // predicted (half-step) velocity
  R* rho;     // density after the EOS clamp (FP32 form: rho - rho0)
  double* u;  // predicted energy
  R* psi;
  R* divv;
  R* p;
  R* drho;
  R* du;
  R2* a;
  R2* av;
  R* rho_min;
  double* u_min;
  R2* v_min;
  uint8_t* flag;
  uint32_t* skey;
  int* cstart;
  R* tau;  // physical viscosity: (Txx, Tyy, Txy, tds) per particle, else null
  uint4* nbm;  // the density pass's neighbour masks (three rows + complete flag)
};

__device__ __forceinline__ PosD tape_pos(const Tape<double>& T, int t) {
  const double2 q = T.xy[t];
  return {q.x, q.y};
}
__device__ __forceinline__ PosF tape_pos(const Tape<float>& T, int t) {
  const uint32_t c = T.cxy[t];
  const float2 o = T.off[t];
  return {int(c & 0xffffu), int(c >> 16), o.x, o.y};
}

// adjoint variables in id order
struct AdjState {
  double *x, *y, *vx, *vy, *rho, *u;  // of the state after the step (in) ...
  double *bx, *by, *bvx, *bvy, *brho, *bu;  // ... and before it (out)
  double *cax, *cay, *cdrho, *cdu;  // of the rates carried into the next predictor
};

// per-step scratch in sorted order
template <typename R> struct AdjScratch {
  R *lax, *lay, *ldrho, *ldu, *lavx, *lavy, *lp;
  R *lQ, *lA, *lpsi, *lS, *lrhoh, *lu, *lvx, *lvy, *lx, *ly;
  R* lus;  // objective seed of the energy (ARM_OBJ_ENERGY)
  R *lTxx, *lTyy, *lTxy;  // physical viscosity: adjoint of the shear tensor
};

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_tape(St<R> S, KParams K,
                                                const uint32_t* __restrict__ skey,
                                                const int* __restrict__ cstart,
                                                const uint4* __restrict__ nbm, Tape<R> T) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const int stride = gridDim.x * BLOCK;
  for (int c = t; c <= K.ncell; c += stride) T.cstart[c] = cstart[c];
  if (t >= K.n) return;
  T.id[t] = S.id[t];
  T.xy[t] = S.xy[t];
  if constexpr (sizeof(R) == 4) {
    T.cxy[t] = S.cxy[t];
    T.off[t] = S.off[t];
  }
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
  T.nbm[t] = nbm ? nbm[t] : make_uint4(0u, 0u, 0u, 0u);
  if (T.tau)
    for (int c = 0; c < 4; ++c) T.tau[4 * (size_t)t + c] = S.tau[4 * (size_t)t + c];
}

// objective: jpart[block] += dt sum_{obj} f over the block's particles,
// f = p^2 (kind 0) or u^2 (kind 1)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_objective(St<R> S, KParams K, int kind,
                                                     const uint8_t* __restrict__ objmask,
                                                     double* __restrict__ jpart) {
  __shared__ double sh[BLOCK];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double v = 0.0;
  if (i < K.n && objmask[S.id[i]]) {
    const double f = kind == 1 ? S.u[i] : (double)S.p[i];
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
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_adj_corrector(Tape<R> T, KParams K, int its,
                                                         AdjState L, AdjScratch<R> W,
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
    const double uu = ub + c * (double)T.du[k];
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
  W.lax[k] = (R)lax;
  W.lay[k] = (R)lay;
  W.ldrho[k] = is_rate(K, id) ? (R)ldrho : (R)0;
  W.ldu[k] = (R)ldu;
  W.lavx[k] = (R)lavx;
  W.lavy[k] = (R)lavy;
  W.lp[k] = (objmask[id] && kind == 0) ? (R)(seed * 2.0 * dt * (double)T.p[k]) : (R)0;
  W.lus[k] = (objmask[id] && kind == 1) ? (R)(seed * 2.0 * dt * T.u[k]) : (R)0;
}

template <typename R>
__device__ __forceinline__ void wendland_q(const KParams& K, R r2, R& q, R& F, R& Wv) {
  q = fsqrt(r2) * (R)K.invh;
  const R t = q - (R)2;
  F = (R)K.C1 * (t * t * t);
  const R s = (R)1 - (R)0.5 * q;
  Wv = (R)K.wnorm * (s * s) * (s * s) * ((R)1 + (R)2 * q);
}

template <typename R>
__device__ __forceinline__ Rows tape_rows(const KParams& K, const Tape<R>& T, int k) {
  return rows_of(K, T.skey[k], T.cstart);
}

// The neighbours l of particle k in the forward gathers' order (rows bottom
// to top, ascending position), visit(l, rx, ry, r2): from the forward's
// masks when they are complete (bit b of row r = candidate t0_r + b, walked
// in one loop over the three rows), else by testing every candidate again.
template <typename R, typename LP, typename Visit>
__device__ __forceinline__ void for_tape_neighbours(const KParams& K, const Tape<R>& T, int k,
                                                    const LP& pk, Visit&& visit) {
  const Rows rw = tape_rows(K, T, k);
  const uint4 nm = T.nbm[k];
  if (nm.w) {
    uint32_t m0 = nm.x, m1 = nm.y, m2 = nm.z;
    while (m0 | m1 | m2) {
      const bool s0 = m0 != 0u, s1 = !s0 && m1 != 0u;
      const uint32_t m = s0 ? m0 : (s1 ? m1 : m2);
      const int l = (s0 ? rw.t0[0] : (s1 ? rw.t0[1] : rw.t0[2])) + (int)__builtin_ctz(m);
      const uint32_t rest = m & (m - 1u);
      m0 = s0 ? rest : m0;
      m1 = s1 ? rest : m1;
      m2 = (s0 || s1) ? m2 : rest;
      R rx, ry;
      delta(pk, tape_pos(T, l), K, rx, ry);
      visit(l, rx, ry, rx * rx + ry * ry);
    }
    return;
  }
  const R cut2 = cutoff2<R>(K);
  for (int r = 0; r < 3; ++r)
    for (int l = rw.t0[r]; l < rw.t1[r]; ++l) {
      if (l == k) continue;
      R rx, ry;
      delta(pk, tape_pos(T, l), K, rx, ry);
      const R r2 = rx * rx + ry * ry;
      if (!(r2 < cut2)) continue;
      visit(l, rx, ry, r2);
    }
}

// 2. force-pass adjoint, gathered: particle k (as i) against every
//    neighbour l (as j); accumulates k's own input adjoints.
template <typename R, bool AV, bool AH, bool XS, bool VI>
__global__ __launch_bounds__(BLOCK) void k_adj_force(Tape<R> T, KParams K, AdjScratch<R> W,
                                                     double* __restrict__ alpha_part,
                                                     double* __restrict__ q0_part,
                                                     double* __restrict__ eta_part) {
  __shared__ double sh[BLOCK];
  double lq0 = 0.0, lal = 0.0, leta = 0.0;
  // one L2 per share of the tape; the partials stay indexed by that share
  const int kb = xcd_block();
  const int k = kb * BLOCK + threadIdx.x;
  if (k < K.n) {
    using R2 = typename V2<R>::type;
    const R m = (R)K.mass, h = (R)K.h;
    const auto pk = tape_pos(T, k);
    const R2 vk = T.v[k];
    const R Qk = (R)1 / dens(K, T.rho[k]), Ak = T.p[k] * Qk;
    const R lak_x = W.lax[k], lak_y = W.lay[k], lduk = W.ldu[k];
    const R lavk_x = W.lavx[k], lavk_y = W.lavy[k];
    const R psik = T.psi[k], uk = (R)T.u[k];
    const R c0 = (R)K.c0, al = (R)K.av_alpha, be = (R)K.av_beta;
    R lx = 0, ly = 0, lvx = 0, lvy = 0, lQ = 0, lA = 0, lpsi = 0, lu = 0;
    // physical viscosity: W = c T, c = eta (pa_sph 0) or eta Q^2 (pa_sph 1)
    R Tk[3] = {0, 0, 0}, ck = 0, lWk[3] = {0, 0, 0};
    if (VI) {
      const R* tk = T.tau + 4 * (size_t)k;
      Tk[0] = tk[0];
      Tk[1] = tk[1];
      Tk[2] = tk[2];
      ck = visc_c(K, Qk);
    }
    for_tape_neighbours(K, T, k, pk, [&](int l, R rx, R ry, R r2) {
      R q, F, Wv;
      wendland_q(K, r2, q, F, Wv);
      const R2 vl = T.v[l];
      const R dvx = vk.x - vl.x, dvy = vk.y - vl.y;
      const R vr = dvx * rx + dvy * ry;
      const R Ql = (R)1 / dens(K, T.rho[l]), Al = T.p[l] * Ql;
      const R sQ = Qk + Ql;
      const R M = (R)2 * Qk * Ql / sQ;
      R Kc = (K.pa_sph == 0) ? (Ak * Ql + Al * Qk) : (Ak * Qk + Al * Ql);
      const bool av_on = AV && vr < (R)0;
      const R einv = (R)1 / (r2 + (R)K.av_e2);
      const R phi = h * vr * einv;
      if (av_on) Kc += (be * phi - al * c0) * phi * M;
      const R G = m * F * Kc;
      const R Dax = lak_x - W.lax[l], Day = lak_y - W.lay[l];
      const R sig = lduk + W.ldu[l];
      const R lG = -(Dax * rx + Day * ry) + (R)0.5 * sig * vr;
      R lrx = -G * Dax, lry = -G * Day;
      R lvr = (R)0.5 * sig * G;
      R ldvx = 0, ldvy = 0;
      R lF = lG * m * Kc;
      const R lK = lG * m * F;
      R lM = 0, lr2 = 0, lW = 0;
      if (K.pa_sph == 0) {
        lA += lK * Ql;
        lQ += lK * Al;
      } else {
        lA += lK * Qk;
        lQ += lK * Ak;
      }
      if (av_on) {
        const R dPi = ((R)2 * be * phi - al * c0) * M;
        lM += lK * (be * phi - al * c0) * phi;
        lal += (double)(lK * (-c0 * phi * M));
        const R lphi = lK * dPi;
        lvr += lphi * h * einv;
        lr2 += lphi * (-phi * einv);
      }
      if (AH) {
        const R eh = (R)1 / (r2 + (R)K.ah_e2);
        const R du_ = uk - (R)T.u[l];
        const R dl = lduk - W.ldu[l];
        const R ps = psik + T.psi[l];
        const R H = m * ps * F * r2 * M * eh;
        const R lH = dl * du_;
        lu += dl * H;
        lpsi += lH * m * F * r2 * M * eh;
        lF += lH * m * ps * r2 * M * eh;
        lr2 += lH * m * ps * F * M * (R)K.ah_e2 * eh * eh;
        lM += lH * m * ps * F * r2 * eh;
      }
      if (XS) {
        const R Dvx = lavk_x - W.lavx[l], Dvy = lavk_y - W.lavy[l];
        const R w = (R)K.xsph_eps * m * M * Wv;
        const R lw = -(Dvx * dvx + Dvy * dvy);
        ldvx += -w * Dvx;
        ldvy += -w * Dvy;
        lM += lw * (R)K.xsph_eps * m * Wv;
        lW += lw * (R)K.xsph_eps * m * M;
      }
      if (VI) {
        // a_k += m F S.r, S = (Q_k Q_l if pa_sph 0) (W_k + W_l); a_l gets -that
        const R* tl = T.tau + 4 * (size_t)l;
        const R cl = visc_c(K, Ql);
        R Sr[3] = {ck * Tk[0] + cl * tl[0], ck * Tk[1] + cl * tl[1], ck * Tk[2] + cl * tl[2]};
        const R qq = K.pa_sph == 0 ? Qk * Ql : (R)1;
        const R Sx = qq * Sr[0], Sy = qq * Sr[1], Sxy = qq * Sr[2];
        const R mF = m * F;
        lF += m * (Dax * (Sx * rx + Sxy * ry) + Day * (Sxy * rx + Sy * ry));
        lrx += mF * (Dax * Sx + Day * Sxy);
        lry += mF * (Dax * Sxy + Day * Sy);
        const R lS[3] = {mF * Dax * rx, mF * Day * ry, mF * (Dax * ry + Day * rx)};
        if (K.pa_sph == 0) lQ += Ql * (lS[0] * Sr[0] + lS[1] * Sr[1] + lS[2] * Sr[2]);
        for (int c = 0; c < 3; ++c) lWk[c] += qq * lS[c];
      }
      lQ += lM * (R)2 * Ql * Ql / (sQ * sQ);
      const R s = (R)1 - (R)0.5 * q;
      const R lq = lF * (R)3 * (R)K.C1 * (q - (R)2) * (q - (R)2) +
                   lW * (R)K.wnorm * s * s * s * ((R)-5 * q);
      lr2 += lq / ((R)2 * h * h * q);
      ldvx += lvr * rx;
      ldvy += lvr * ry;
      lrx += lvr * dvx + (R)2 * lr2 * rx;
      lry += lvr * dvy + (R)2 * lr2 * ry;
      lx += lrx;
      ly += lry;
      lvx += ldvx;
      lvy += ldvy;
    });
    if (VI) {
      // W_k = c_k T_k and tds_k = eta/2 Q_k (Txx^2 + Tyy^2 + 2 Txy^2), added to du_k
      const R eta = (R)K.eta;
      const R tw = Tk[0] * lWk[0] + Tk[1] * lWk[1] + Tk[2] * lWk[2];
      const R tt = Tk[0] * Tk[0] + Tk[1] * Tk[1] + (R)2 * Tk[2] * Tk[2];
      const R g = lduk * eta * Qk;
      W.lTxx[k] = ck * lWk[0] + g * Tk[0];
      W.lTyy[k] = ck * lWk[1] + g * Tk[1];
      W.lTxy[k] = ck * lWk[2] + g * (R)2 * Tk[2];
      if (K.pa_sph == 1) lQ += (R)2 * eta * Qk * tw;
      lQ += lduk * (R)0.5 * eta * tt;
      leta = (double)(K.pa_sph == 0 ? tw : Qk * Qk * tw) + (double)(lduk * (R)0.5 * Qk * tt);
    }
    // heat source du_k += q0 E_k Q_k, E = exp(-hs_k |x_k - c|^2)
    if ((K.flags & 64) && T.id[k] < K.np) {
      const double2 xk = T.xy[k];
      const double E = heat_shape(K, xk.x, xk.y);
      lq0 = (double)lduk * E * (double)Qk;
      lQ += lduk * (R)(K.hs_q0 * E);
      const double lE = (double)lduk * K.hs_q0 * (double)Qk * E * (-2.0 * K.hs_k);
      lx += (R)(lE * (xk.x - K.hs_x));
      ly += (R)(lE * (xk.y - K.hs_y));
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
  if (threadIdx.x == 0) alpha_part[kb] += sh[0];
  __syncthreads();
  sh[threadIdx.x] = lq0;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) q0_part[kb] += sh[0];
  if (VI) {
    __syncthreads();
    sh[threadIdx.x] = leta;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) eta_part[kb] += sh[0];
  }
}

// 2b. shear-tensor adjoint (physical viscosity), gathered.  The pair term
//     H(dv, dW), dv = v_l - v_k, dW = F r, r = x_k - x_l, enters
//     T_k with m Q_l and T_l with m Q_k (H is the same from both ends), so
//     particle k collects dv, r and Q_k derivatives of lambda_H =
//     m Q_l lT_k + m Q_k lT_l.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_adj_tensor(Tape<R> T, KParams K, AdjScratch<R> W) {
  const int k = xcd_block() * BLOCK + threadIdx.x;  // one L2 per share of the tape
  if (k >= K.n) return;
  using R2 = typename V2<R>::type;
  const R m = (R)K.mass, h = (R)K.h;
  const auto pk = tape_pos(T, k);
  const R2 vk = T.v[k];
  const R Qk = (R)1 / dens(K, T.rho[k]);
  const R Lk[3] = {W.lTxx[k], W.lTyy[k], W.lTxy[k]};
  const R two3 = (R)(2.0 / 3.0);
  R lx = 0, ly = 0, lvx = 0, lvy = 0, lQ = 0;
  for_tape_neighbours(K, T, k, pk, [&](int l, R rx, R ry, R r2) {
    R q, F, Wv;
    wendland_q(K, r2, q, F, Wv);
    const R2 vl = T.v[l];
    const R dvx = vl.x - vk.x, dvy = vl.y - vk.y;
    const R dWx = F * rx, dWy = F * ry;
    const R a = dvx * dWx, b = dvy * dWy;
    const R H[3] = {two3 * ((R)2 * a - b), two3 * ((R)2 * b - a), dvx * dWy + dvy * dWx};
    const R Ql = (R)1 / dens(K, T.rho[l]);
    const R Ll[3] = {W.lTxx[l], W.lTyy[l], W.lTxy[l]};
    R lH[3];
    for (int c = 0; c < 3; ++c) lH[c] = m * Ql * Lk[c] + m * Qk * Ll[c];
    lQ += m * (H[0] * Ll[0] + H[1] * Ll[1] + H[2] * Ll[2]);
    const R la = two3 * ((R)2 * lH[0] - lH[1]);
    const R lb = two3 * ((R)2 * lH[1] - lH[0]);
    const R ldvx = la * dWx + lH[2] * dWy, ldvy = lb * dWy + lH[2] * dWx;
    const R ldWx = la * dvx + lH[2] * dvy, ldWy = lb * dvy + lH[2] * dvx;
    lvx -= ldvx;
    lvy -= ldvy;
    const R lF = ldWx * rx + ldWy * ry;
    const R lr2 = lF * (R)3 * (R)K.C1 * (q - (R)2) * (q - (R)2) / ((R)2 * h * h * q);
    lx += F * ldWx + (R)2 * lr2 * rx;
    ly += F * ldWy + (R)2 * lr2 * ry;
  });
  W.lx[k] += lx;
  W.ly[k] += ly;
  W.lvx[k] += lvx;
  W.lvy[k] += lvy;
  W.lQ[k] += lQ;
}

// 3. EOS and art_heat coefficient adjoint (per particle)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_adj_eos(Tape<R> T, KParams K, AdjScratch<R> W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const R rho = dens(K, T.rho[k]);
  const R Q = (R)1 / rho;
  R lp = W.lp[k] + W.lA[k] * Q;
  R lrhoe = -Q * Q * (W.lQ[k] + W.lA[k] * T.p[k]);
  const R dv = T.divv[k];  // S / rho
  const R dpsi = (dv < (R)0) ? (R)(-2.0 * K.g2h2) : (R)0;
  const R ldv = W.lpsi[k] * dpsi;
  W.lS[k] = ldv / rho;
  lrhoe += -ldv * dv / rho;
  R lrhoh = 0;
  if (is_rate(K, T.id[k]) && !(T.flag[k] & 2u)) {
    if (K.eos == 1) {
      // linear: dp/drho = c0^2
      lrhoh = lrhoe + lp * (R)K.lin_c2;
    } else {
      // Tait: dp/drho = 7 B / rho0 (rho / rho0)^6
      const R t = rho / (R)K.rho0;
      const R t2 = t * t;
      lrhoh = lrhoe + lp * (R)(7.0 * K.B / K.rho0) * (t2 * t2 * t2);
    }
  }
  W.lrhoh[k] = lrhoh;
}

// 4. density adjoint, gathered; adds to the position / velocity adjoints
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_adj_density(Tape<R> T, KParams K, AdjScratch<R> W) {
  const int k = xcd_block() * BLOCK + threadIdx.x;  // one L2 per share of the tape
  if (k >= K.n) return;
  using R2 = typename V2<R>::type;
  const R m = (R)K.mass, h = (R)K.h;
  const auto pk = tape_pos(T, k);
  const R2 vk = T.v[k];
  const R ltk = W.ldrho[k] - W.lS[k];
  R lx = 0, ly = 0, lvx = 0, lvy = 0;
  for_tape_neighbours(K, T, k, pk, [&](int l, R rx, R ry, R r2) {
    const R lt = ltk + W.ldrho[l] - W.lS[l];
    R q, F, Wv;
    wendland_q(K, r2, q, F, Wv);
    const R2 vl = T.v[l];
    const R dvx = vk.x - vl.x, dvy = vk.y - vl.y;
    const R vr = dvx * rx + dvy * ry;
    const R lvr = lt * m * F;
    const R lF = lt * m * vr;
    const R lr2 = lF * (R)3 * (R)K.C1 * (q - (R)2) * (q - (R)2) / ((R)2 * h * h * q);
    lx += lvr * dvx + (R)2 * lr2 * rx;
    ly += lvr * dvy + (R)2 * lr2 * ry;
    lvx += lvr * rx;
    lvy += lvr * ry;
  });
  W.lx[k] += lx;
  W.ly[k] += ly;
  W.lvx[k] += lvx;
  W.lvy[k] += lvy;
}

// 5. predictor adjoint (time_integration.f:136-154): folds the rate-input
//    adjoints into the state adjoints before the step and into the carried
//    rate adjoints of the previous step
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_adj_predictor(Tape<R> T, KParams K, int its,
                                                         AdjState L, AdjScratch<R> W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const int id = T.id[k];
  const double h2 = K.dt / 2.0;
  const double wlx = W.lx[k], wly = W.ly[k], wlvx = W.lvx[k], wlvy = W.lvy[k];
  const double wlrho = W.lrhoh[k], wlu = W.lu[k];
  L.bx[id] += wlx;
  L.by[id] += wly;
  L.bvx[id] += wlvx;
  L.bvy[id] += wlvy;
  L.brho[id] += wlrho;
  double cax = 0, cay = 0, cdrho = 0, cdu = 0;
  if (its != 1 && is_rate(K, id)) {
    cdrho = h2 * wlrho;
    if (!(T.flag[k] & 1u)) {
      L.bu[id] += wlu;
      cdu = h2 * wlu;
    }
    if (id < K.np) {
      cax = h2 * wlvx;
      cay = h2 * wlvy;
    }
  } else {
    L.bu[id] += wlu;
  }
  L.cax[id] = cax;
  L.cay[id] = cay;
  L.cdrho[id] = cdrho;
  L.cdu[id] = cdu;
}

}  // namespace arm
