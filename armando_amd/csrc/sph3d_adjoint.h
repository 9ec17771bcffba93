// sph3d_adjoint.h -- discrete adjoint of the 3D step, for the FP64 plain
// form and the FP32 chosen form (FP32 pair arithmetic on the cell + offset
// positions, FP64 state adjoints), with the derivation of the 2D adjoint in
// sph_adjoint.h carried to three components:
//   density   drho_i = sum_j m F vr,  S_i = -sum_j m F vr (divv_i = S_i/rho_i)
//   EOS       p = Tait(rho) or c0^2 rho + p0 with the pmin clamp; psi, Q, A
//   force     G = m F (A_i Q_j + A_j Q_i + Pi),  a_i = -sum G r + g,
//             du_i = sum G vr/2 + H (u_i - u_j),  av_i = -sum eps m M W dv
//   leapfrog  time_integration.f:136-217
// r = x_i - x_j (3 components), vr = (v_i - v_j).r, grad W = F r with the
// 3D Wendland constants of KParams.  Neighbour lists and branches are those
// of the forward step.  Pair terms are gathered per particle (each pair is
// visited from both ends), so no atomics are needed.
#pragma once

#include "sph3d.h"

namespace arm {

// one recorded 3D step: the sorted state after the force pass
template <typename R> struct Tape3 {
  using R4 = typename V4<R>::type;
  int32_t* id;
  double4* xyz;
  float4* pf;  // FP32 form: cell column + offsets of the pair arithmetic
  R4* v;
  R* rho;
  double* u;
  R* psi;
  R* divv;
  R* p;
  R* drho;
  R* du;
  R4* a;
  R4* av;
  R* rho_min;
  double* u_min;
  R4* v_min;
  uint8_t* flag;
  uint32_t* skey;
  int* cstart;
  uint32_t* nbm;  // the forward's neighbour masks (Mask3 layout, stride n)
};

__device__ __forceinline__ Pos3D tape_pos3(const Tape3<double>& T, int t) {
  const double4 q = T.xyz[t];
  return {q.x, q.y, q.z};
}
__device__ __forceinline__ Pos3F tape_pos3(const Tape3<float>& T, int t) {
  const float4 q = T.pf[t];
  return {q.w, q.x, q.y, q.z};
}

// state adjoints in id order: after the step (in), before it (out), and of
// the rates carried into the next predictor
struct AdjState3 {
  double *x, *y, *z, *vx, *vy, *vz, *rho, *u;
  double *bx, *by, *bz, *bvx, *bvy, *bvz, *brho, *bu;
  double *cax, *cay, *caz, *cdrho, *cdu;
};

// per-step scratch in sorted order
template <typename R> struct AdjScratch3 {
  R *lax, *lay, *laz, *ldrho, *ldu, *lavx, *lavy, *lavz, *lp;
  R *lQ, *lA, *lpsi, *lS, *lrhoh, *lu, *lvx, *lvy, *lvz, *lx, *ly, *lz;
  R* lus;
};

template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_tape(St3<R> S, KParams K,
                                                 const uint32_t* __restrict__ skey,
                                                 const int* __restrict__ cstart, Mask3 nbm,
                                                 Tape3<R> T) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const int stride = gridDim.x * BLOCK;
  for (int c = t; c <= K.ncell; c += stride) T.cstart[c] = cstart[c];
  if (t >= K.n) return;
  T.id[t] = S.id[t];
  T.xyz[t] = S.xyz[t];
  if constexpr (sizeof(R) == 4) T.pf[t] = S.pf[t];
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
  for (int q = 0; q < M3_TOTAL; ++q)
    T.nbm[(size_t)q * K.n + t] = nbm.w ? nbm.get(q, t) : 0u;
}

// Calls visit(l, uy, uz) for every neighbour l of particle k of the taped
// step, in the forward's order: from the recorded masks when they are
// complete, else by testing the candidates again.
template <typename R, typename Visit>
__device__ __forceinline__ void for_tape_neighbours3(const Tape3<R>& T, const KParams& K, int k,
                                                     Visit&& visit) {
  const Rows9 rw = rows3_of(K, T.skey[k], T.cstart);
  const size_t st = (size_t)K.n;
  if (T.nbm[M3_OK * st + k]) {
    // per row, one loop over the bits of its three words (in word order)
    for (int row = 0; row < 9; ++row) {
      const R uy = row_u3<R>(K, 1 - row % 3), uz = row_u3<R>(K, 1 - row / 3);
      uint32_t m0 = T.nbm[(size_t)(row * M3_WORDS + 0) * st + k];
      uint32_t m1 = T.nbm[(size_t)(row * M3_WORDS + 1) * st + k];
      uint32_t m2 = T.nbm[(size_t)(row * M3_WORDS + 2) * st + k];
      const int t0 = rw.t0[row];
      while (m0 | m1 | m2) {
        const bool s0 = m0 != 0u, s1 = !s0 && m1 != 0u;
        const uint32_t m = s0 ? m0 : (s1 ? m1 : m2);
        const int l = t0 + (s0 ? 0 : (s1 ? 32 : 64)) + (int)__builtin_ctz(m);
        const uint32_t rest = m & (m - 1u);
        m0 = s0 ? rest : m0;
        m1 = s1 ? rest : m1;
        m2 = (s0 || s1) ? m2 : rest;
        visit(l, uy, uz);
      }
    }
    return;
  }
  const auto pk = tape_pos3(T, k);
  const R cut2 = cutoff2<R>(K);
  for (int row = 0; row < 9; ++row) {
    const R uy = row_u3<R>(K, 1 - row % 3), uz = row_u3<R>(K, 1 - row / 3);
    for (int l = rw.t0[row]; l < rw.t1[row]; ++l) {
      if (l == k) continue;
      R d0, d1, d2;
      const R r2 = delta3(pk, tape_pos3(T, l), K, uy, uz, d0, d1, d2);
      if (!(r2 < cut2)) continue;
      visit(l, uy, uz);
    }
  }
}

// objective: jpart[block] += dt sum_{obj} f, f = p^2 (kind 0) or u^2 (kind 1)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_objective(St3<R> S, KParams K, int kind,
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

// 1. corrector adjoint (time_integration.f:186-217)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_adj_corrector(Tape3<R> T, KParams K, int its,
                                                          AdjState3 L, AdjScratch3<R> W,
                                                          const uint8_t* __restrict__ objmask,
                                                          int kind, double seed) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const int id = T.id[k];
  const double dt = K.dt, h2 = K.dt / 2.0;
  double lax = L.cax[id], lay = L.cay[id], laz = L.caz[id];
  double ldrho = L.cdrho[id], ldu = L.cdu[id];
  double lavx = 0, lavy = 0, lavz = 0;
  double bvx = L.vx[id], bvy = L.vy[id], bvz = L.vz[id];
  double brho = L.rho[id], bu = L.u[id];
  if (is_rate(K, id)) {
    const double c = (its != 1) ? dt : h2;
    ldrho += c * L.rho[id];
    const double ub = (its != 1) ? T.u_min[k] : T.u[k];
    const double uu = ub + c * (double)T.du[k];
    bu = 0;
    if (!(uu < 0.0)) {
      bu = L.u[id];
      ldu += c * L.u[id];
    }
    if (id < K.np) {
      const double lvtx = L.vx[id] + dt * L.x[id];
      const double lvty = L.vy[id] + dt * L.y[id];
      const double lvtz = L.vz[id] + dt * L.z[id];
      bvx = lvtx;
      bvy = lvty;
      bvz = lvtz;
      lax += c * lvtx;
      lay += c * lvty;
      laz += c * lvtz;
      if (its != 1) {
        lavx = lvtx;
        lavy = lvty;
        lavz = lvtz;
      }
    }
  }
  L.bx[id] = L.x[id];
  L.by[id] = L.y[id];
  L.bz[id] = L.z[id];
  L.bvx[id] = bvx;
  L.bvy[id] = bvy;
  L.bvz[id] = bvz;
  L.brho[id] = brho;
  L.bu[id] = bu;
  W.lax[k] = (R)lax;
  W.lay[k] = (R)lay;
  W.laz[k] = (R)laz;
  W.ldrho[k] = is_rate(K, id) ? (R)ldrho : (R)0;
  W.ldu[k] = (R)ldu;
  W.lavx[k] = (R)lavx;
  W.lavy[k] = (R)lavy;
  W.lavz[k] = (R)lavz;
  W.lp[k] = (objmask[id] && kind == 0) ? (R)(seed * 2.0 * dt * (double)T.p[k]) : (R)0;
  W.lus[k] = (objmask[id] && kind == 1) ? (R)(seed * 2.0 * dt * T.u[k]) : (R)0;
}

template <typename R>
__device__ __forceinline__ void wendland3_q(const KParams& K, R r2, R& q, R& F, R& Wv) {
  q = fsqrt(r2) * (R)K.invh;
  const R t = q - (R)2;
  F = (R)K.C1 * (t * t * t);
  const R s = (R)1 - (R)0.5 * q;
  Wv = (R)K.wnorm * (s * s) * (s * s) * ((R)1 + (R)2 * q);
}

// 2. force-pass adjoint, gathered: particle k (as i) against every
//    neighbour l (as j); accumulates k's own input adjoints.
template <typename R, bool AV, bool AH, bool XS>
__global__ __launch_bounds__(BLOCK) void k3_adj_force(Tape3<R> T, KParams K, AdjScratch3<R> W,
                                                      double* __restrict__ alpha_part) {
  __shared__ double sh[BLOCK];
  double lal = 0.0;
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k < K.n) {
    const R m = (R)K.mass, h = (R)K.h;
    const auto pk = tape_pos3(T, k);
    const auto vk = T.v[k];
    const R Qk = (R)1 / dens(K, T.rho[k]), Ak = T.p[k] * Qk;
    const R lak[3] = {W.lax[k], W.lay[k], W.laz[k]};
    const R lavk[3] = {W.lavx[k], W.lavy[k], W.lavz[k]};
    const R lduk = W.ldu[k];
    const R psik = T.psi[k], uk = (R)T.u[k];
    const R c0 = (R)K.c0, al = (R)K.av_alpha, be = (R)K.av_beta;
    R lr_acc[3] = {0, 0, 0}, lv_acc[3] = {0, 0, 0};
    R lQ = 0, lA = 0, lpsi = 0, lu = 0;
    for_tape_neighbours3<R>(T, K, k, [&](int l, R uy, R uz) {
        R r[3];
        const R r2 = delta3(pk, tape_pos3(T, l), K, uy, uz, r[0], r[1], r[2]);
        R q, F, Wv;
        wendland3_q(K, r2, q, F, Wv);
        const auto vl = T.v[l];
        const R dv[3] = {vk.x - vl.x, vk.y - vl.y, vk.z - vl.z};
        const R vr = dv[0] * r[0] + dv[1] * r[1] + dv[2] * r[2];
        const R Ql = (R)1 / dens(K, T.rho[l]), Al = T.p[l] * Ql;
        const R sQ = Qk + Ql;
        const R M = (R)2 * Qk * Ql / sQ;
        R Kc = (K.pa_sph == 0) ? (Ak * Ql + Al * Qk) : (Ak * Qk + Al * Ql);
        const bool av_on = AV && vr < (R)0;
        const R einv = (R)1 / (r2 + (R)K.av_e2);
        const R phi = h * vr * einv;
        if (av_on) Kc += (be * phi - al * c0) * phi * M;
        const R G = m * F * Kc;
        const R Da[3] = {lak[0] - W.lax[l], lak[1] - W.lay[l], lak[2] - W.laz[l]};
        const R sig = lduk + W.ldu[l];
        const R lG = -(Da[0] * r[0] + Da[1] * r[1] + Da[2] * r[2]) + (R)0.5 * sig * vr;
        R lr[3] = {-G * Da[0], -G * Da[1], -G * Da[2]};
        R lvr = (R)0.5 * sig * G;
        R ldv[3] = {0, 0, 0};
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
          const R Dv[3] = {lavk[0] - W.lavx[l], lavk[1] - W.lavy[l], lavk[2] - W.lavz[l]};
          const R w = (R)K.xsph_eps * m * M * Wv;
          const R lw = -(Dv[0] * dv[0] + Dv[1] * dv[1] + Dv[2] * dv[2]);
          for (int c = 0; c < 3; ++c) ldv[c] += -w * Dv[c];
          lM += lw * (R)K.xsph_eps * m * Wv;
          lW += lw * (R)K.xsph_eps * m * M;
        }
        lQ += lM * (R)2 * Ql * Ql / (sQ * sQ);
        const R s = (R)1 - (R)0.5 * q;
        const R lq = lF * (R)3 * (R)K.C1 * (q - (R)2) * (q - (R)2) +
                     lW * (R)K.wnorm * s * s * s * ((R)-5 * q);
        lr2 += lq / ((R)2 * h * h * q);
        for (int c = 0; c < 3; ++c) {
          ldv[c] += lvr * r[c];
          lr[c] += lvr * dv[c] + (R)2 * lr2 * r[c];
          lr_acc[c] += lr[c];
          lv_acc[c] += ldv[c];
        }
    });
    W.lx[k] = lr_acc[0];
    W.ly[k] = lr_acc[1];
    W.lz[k] = lr_acc[2];
    W.lvx[k] = lv_acc[0];
    W.lvy[k] = lv_acc[1];
    W.lvz[k] = lv_acc[2];
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
}

// 3. EOS and art_heat coefficient adjoint (per particle)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_adj_eos(Tape3<R> T, KParams K, AdjScratch3<R> W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const R rho = dens(K, T.rho[k]);
  const R Q = (R)1 / rho;
  const R lp = W.lp[k] + W.lA[k] * Q;
  R lrhoe = -Q * Q * (W.lQ[k] + W.lA[k] * T.p[k]);
  const R dv = T.divv[k];
  const R dpsi = (dv < (R)0) ? (R)(-2.0 * K.g2h2) : (R)0;
  const R ldv = W.lpsi[k] * dpsi;
  W.lS[k] = ldv / rho;
  lrhoe += -ldv * dv / rho;
  R lrhoh = 0;
  if (is_rate(K, T.id[k]) && !(T.flag[k] & 2u)) {
    if (K.eos == 1) {
      lrhoh = lrhoe + lp * (R)K.lin_c2;
    } else {
      const R t = rho / (R)K.rho0;
      const R t2 = t * t;
      lrhoh = lrhoe + lp * (R)(7.0 * K.B / K.rho0) * (t2 * t2 * t2);
    }
  }
  W.lrhoh[k] = lrhoh;
}

// 4. density adjoint, gathered; adds to the position / velocity adjoints
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_adj_density(Tape3<R> T, KParams K,
                                                        AdjScratch3<R> W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const R m = (R)K.mass, h = (R)K.h;
  const auto pk = tape_pos3(T, k);
  const auto vk = T.v[k];
  const R ltk = W.ldrho[k] - W.lS[k];
  R lr[3] = {0, 0, 0}, lv[3] = {0, 0, 0};
  for_tape_neighbours3<R>(T, K, k, [&](int l, R uy, R uz) {
      R r[3];
      const R r2 = delta3(pk, tape_pos3(T, l), K, uy, uz, r[0], r[1], r[2]);
      const R lt = ltk + W.ldrho[l] - W.lS[l];
      R q, F, Wv;
      wendland3_q(K, r2, q, F, Wv);
      const auto vl = T.v[l];
      const R dv[3] = {vk.x - vl.x, vk.y - vl.y, vk.z - vl.z};
      const R vr = dv[0] * r[0] + dv[1] * r[1] + dv[2] * r[2];
      const R lvr = lt * m * F;
      const R lF = lt * m * vr;
      const R lr2 = lF * (R)3 * (R)K.C1 * (q - (R)2) * (q - (R)2) / ((R)2 * h * h * q);
      for (int c = 0; c < 3; ++c) {
        lr[c] += lvr * dv[c] + (R)2 * lr2 * r[c];
        lv[c] += lvr * r[c];
      }
  });
  W.lx[k] += lr[0];
  W.ly[k] += lr[1];
  W.lz[k] += lr[2];
  W.lvx[k] += lv[0];
  W.lvy[k] += lv[1];
  W.lvz[k] += lv[2];
}

// 5. predictor adjoint (time_integration.f:136-154)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k3_adj_predictor(Tape3<R> T, KParams K, int its,
                                                          AdjState3 L, AdjScratch3<R> W) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= K.n) return;
  const int id = T.id[k];
  const double h2 = K.dt / 2.0;
  const double wlvx = W.lvx[k], wlvy = W.lvy[k], wlvz = W.lvz[k];
  const double wlrho = W.lrhoh[k], wlu = W.lu[k];
  L.bx[id] += (double)W.lx[k];
  L.by[id] += (double)W.ly[k];
  L.bz[id] += (double)W.lz[k];
  L.bvx[id] += wlvx;
  L.bvy[id] += wlvy;
  L.bvz[id] += wlvz;
  L.brho[id] += wlrho;
  double cax = 0, cay = 0, caz = 0, cdrho = 0, cdu = 0;
  if (its != 1 && is_rate(K, id)) {
    cdrho = h2 * wlrho;
    if (!(T.flag[k] & 1u)) {
      L.bu[id] += wlu;
      cdu = h2 * wlu;
    }
    if (id < K.np) {
      cax = h2 * wlvx;
      cay = h2 * wlvy;
      caz = h2 * wlvz;
    }
  } else {
    L.bu[id] += wlu;
  }
  L.cax[id] = cax;
  L.cay[id] = cay;
  L.caz[id] = caz;
  L.cdrho[id] = cdrho;
  L.cdu[id] = cdu;
}

}  // namespace arm
