/*
 * armando_adjoint.c -- FP64 discrete adjoint of the oracle's time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP adjoint kernels.
 *
 * The forward step differentiated here is the one of armando_oracle.c
 * (time_integration.f:131-235 with the rates of single_step.f), written with
 * the per-particle Q = 1/rho, A = p/rho of the HIP force pass:
 *   density      drho_i = sum_j m F_ij vr_ij            (density.f:201-221)
 *   EOS          p = Tait(rho), clamp p < pmin -> p = 0, rho = rho0
 *   art_heat     divv_i = S_i / rho_i, S_i = -sum_j m F_ij vr_ij,
 *                psi_i = g1 h c0 + g2 h^2 (|divv_i| - divv_i)
 *   force        G_ij = m F_ij (A_i Q_j + A_j Q_i + Pi_ij)        (pa_sph 0)
 *                a_i  = -sum_j G_ij r_ij + g
 *                du_i = sum_j G_ij vr_ij / 2 + H_ij (u_i - u_j)
 *                av_i = -sum_j eps m M_ij W_ij (v_i - v_j)
 *   with r_ij = x_i - x_j, vr_ij = (v_i - v_j).r_ij, grad_i W = F_ij r_ij,
 *   M_ij = 1/rho_bar = 2 Q_i Q_j/(Q_i+Q_j), phi = h vr/(r^2 + (h eta)^2),
 *   Pi = (beta phi - alpha c0) phi M when vr < 0,
 *   H = m (psi_i + psi_j) F r^2 M / (r^2 + (h eta_h)^2).
 * Objective: J = sum over steps of dt * sum_{i in obj} p_i^2, with p the
 * pressure of that step's single_step.
 * The neighbour list of every step is that of the forward step (piecewise
 * constant); the kinks (u < 0, p < pmin, vr < 0, |divv|) take the branch the
 * forward took.  Everything is stored; no checkpoints (small inputs only).
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "armando_oracle.h"

typedef struct {
  /* inputs of the rates */
  double *x, *y, *vx, *vy, *rho, *u;
  /* intermediates */
  double *drho, *S, *rhoe, *p, *psi, *Q, *A;
  int *clamp; /* EOS clamp taken */
  /* outputs */
  double *ax, *ay, *du, *avx, *avy;
} rates_t;

static double* dz(int n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static void rates_alloc(rates_t* R, int n) {
  double** f[] = {&R->drho, &R->S, &R->rhoe, &R->p, &R->psi, &R->Q, &R->A,
                  &R->ax,   &R->ay, &R->du,  &R->avx, &R->avy};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k) *f[k] = dz(n);
  R->clamp = (int*)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
}

static void rates_free(rates_t* R) {
  double* f[] = {R->drho, R->S, R->rhoe, R->p, R->psi, R->Q, R->A,
                 R->ax,   R->ay, R->du,  R->avx, R->avy};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k) free(f[k]);
  free(R->clamp);
}

static inline int is_rate(const orc_params* P, int np, int i) {
  return (P->flags & ORC_DBC) || i < np;
}

/* Wendland pieces: q, F (grad W = F r) and W */
static inline void wend(const orc_params* P, double r2, double* q, double* F, double* W) {
  const double h = P->h;
  const double a = 7.0 / (4.0 * M_PI * h * h);
  const double C1 = 0.625 * a / (h * h);
  *q = sqrt(r2) / h;
  const double t = *q - 2.0;
  *F = C1 * t * t * t;
  const double s = 1.0 - 0.5 * (*q);
  *W = a * s * s * s * s * (1.0 + 2.0 * (*q));
}

static void rates_forward(const orc_params* P, int n, int np, int niac, const int* pi,
                          const int* pj, rates_t* R) {
  const double m = P->mass, h = P->h;
  const double e2a = (h * P->av_eta) * (h * P->av_eta);
  const double e2h = (h * P->ah_eta) * (h * P->ah_eta);
  const double B = P->rho0 * P->c0 * P->c0 / 7.0;
  for (int i = 0; i < n; ++i) {
    R->drho[i] = 0;
    R->S[i] = 0;
    R->ax[i] = R->ay[i] = R->du[i] = R->avx[i] = R->avy[i] = 0;
  }
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    const double rx = R->x[i] - R->x[j], ry = R->y[i] - R->y[j];
    const double r2 = rx * rx + ry * ry;
    double q, F, W;
    wend(P, r2, &q, &F, &W);
    const double vr = (R->vx[i] - R->vx[j]) * rx + (R->vy[i] - R->vy[j]) * ry;
    const double t = m * F * vr;
    R->drho[i] += t;
    R->drho[j] += t;
    R->S[i] -= t;
    R->S[j] -= t;
  }
  for (int i = 0; i < n; ++i) {
    if (!is_rate(P, np, i)) R->drho[i] = 0;
    double rho = R->rho[i];
    double p = 0;
    int cl = 0;
    if (is_rate(P, np, i)) {
      const double tt = rho / P->rho0;
      p = B * (pow(tt, 7.0) - 1.0);
      if (p < P->pmin) {
        p = 0;
        rho = P->rho0;
        cl = 1;
      }
    }
    R->rhoe[i] = rho;
    R->p[i] = p;
    R->clamp[i] = cl;
    const double dv = R->S[i] / rho;
    R->psi[i] = P->ah_g1 * h * P->c0 + P->ah_g2 * h * h * (fabs(dv) - dv);
    R->Q[i] = 1.0 / rho;
    R->A[i] = p / rho;
  }
  const int f_av = (P->flags & ORC_ART_VISC) != 0, f_ah = (P->flags & ORC_ART_HEAT) != 0;
  const int f_xs = (P->flags & ORC_AV_VEL) != 0;
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    const double rx = R->x[i] - R->x[j], ry = R->y[i] - R->y[j];
    const double r2 = rx * rx + ry * ry;
    double q, F, W;
    wend(P, r2, &q, &F, &W);
    const double dvx = R->vx[i] - R->vx[j], dvy = R->vy[i] - R->vy[j];
    const double vr = dvx * rx + dvy * ry;
    const double Qi = R->Q[i], Qj = R->Q[j];
    const double M = 2.0 * Qi * Qj / (Qi + Qj);
    double K = (P->pa_sph == 0) ? (R->A[i] * Qj + R->A[j] * Qi) : (R->A[i] * Qi + R->A[j] * Qj);
    if (f_av && vr < 0) {
      const double phi = h * vr / (r2 + e2a);
      K += (P->av_beta * phi - P->av_alpha * P->c0) * phi * M;
    }
    const double G = m * F * K;
    R->ax[i] -= G * rx;
    R->ay[i] -= G * ry;
    R->ax[j] += G * rx;
    R->ay[j] += G * ry;
    R->du[i] += 0.5 * G * vr;
    R->du[j] += 0.5 * G * vr;
    if (f_ah) {
      const double H = m * (R->psi[i] + R->psi[j]) * F * r2 * M / (r2 + e2h);
      R->du[i] += H * (R->u[i] - R->u[j]);
      R->du[j] += H * (R->u[j] - R->u[i]);
    }
    if (f_xs) {
      const double w = P->xsph_eps * m * M * W;
      R->avx[i] -= w * dvx;
      R->avy[i] -= w * dvy;
      R->avx[j] += w * dvx;
      R->avy[j] += w * dvy;
    }
  }
  if (P->flags & ORC_GRAVITY)
    for (int i = 0; i < np; ++i) {
      R->ax[i] += P->gx;
      R->ay[i] += P->gy;
    }
}

/* Adjoint of rates_forward.  In: lax, lay, ldu, lavx, lavy, ldrho, lp.
   Out (accumulated): lx, ly, lvx, lvy, lrho, lu, and *lalpha. */
static void rates_adjoint(const orc_params* P, int n, int np, int niac, const int* pi,
                          const int* pj, const rates_t* R, const double* lax, const double* lay,
                          const double* ldu, const double* lavx, const double* lavy,
                          const double* ldrho, double* lp, double* lx, double* ly, double* lvx,
                          double* lvy, double* lrho, double* lu, double* lalpha) {
  const double m = P->mass, h = P->h;
  const double e2a = (h * P->av_eta) * (h * P->av_eta);
  const double e2h = (h * P->ah_eta) * (h * P->ah_eta);
  const double a_n = 7.0 / (4.0 * M_PI * h * h);
  const double C1 = 0.625 * a_n / (h * h);
  const double B = P->rho0 * P->c0 * P->c0 / 7.0;
  const int f_av = (P->flags & ORC_ART_VISC) != 0, f_ah = (P->flags & ORC_ART_HEAT) != 0;
  const int f_xs = (P->flags & ORC_AV_VEL) != 0;
  double* lQ = dz(n);
  double* lA = dz(n);
  double* lpsi = dz(n);
  /* 1. force pairs */
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    const double rx = R->x[i] - R->x[j], ry = R->y[i] - R->y[j];
    const double r2 = rx * rx + ry * ry;
    double q, F, W;
    wend(P, r2, &q, &F, &W);
    const double dvx = R->vx[i] - R->vx[j], dvy = R->vy[i] - R->vy[j];
    const double vr = dvx * rx + dvy * ry;
    const double Qi = R->Q[i], Qj = R->Q[j], Ai = R->A[i], Aj = R->A[j];
    const double sQ = Qi + Qj;
    const double M = 2.0 * Qi * Qj / sQ;
    double K = (P->pa_sph == 0) ? (Ai * Qj + Aj * Qi) : (Ai * Qi + Aj * Qj);
    const int av_on = f_av && vr < 0;
    const double einv = 1.0 / (r2 + e2a);
    const double phi = h * vr * einv;
    if (av_on) K += (P->av_beta * phi - P->av_alpha * P->c0) * phi * M;
    const double G = m * F * K;
    const double Dax = lax[i] - lax[j], Day = lay[i] - lay[j];
    const double sig = ldu[i] + ldu[j];
    /* dL/dG, direct dL/dr, dL/dvr */
    const double lG = -(Dax * rx + Day * ry) + 0.5 * sig * vr;
    double lrx = -G * Dax, lry = -G * Day;
    double lvr = 0.5 * sig * G;
    double ldvx = 0, ldvy = 0;
    double lF = lG * m * K;
    const double lK = lG * m * F;
    double lM = 0, lr2 = 0, lW = 0;
    if (P->pa_sph == 0) {
      lA[i] += lK * Qj;
      lA[j] += lK * Qi;
      lQ[j] += lK * Ai;
      lQ[i] += lK * Aj;
    } else {
      lA[i] += lK * Qi;
      lA[j] += lK * Qj;
      lQ[i] += lK * Ai;
      lQ[j] += lK * Aj;
    }
    if (av_on) {
      const double dPi_dphi = (2.0 * P->av_beta * phi - P->av_alpha * P->c0) * M;
      lM += lK * (P->av_beta * phi - P->av_alpha * P->c0) * phi;
      *lalpha += lK * (-P->c0 * phi * M);
      const double lphi = lK * dPi_dphi;
      lvr += lphi * h * einv;
      lr2 += lphi * (-phi * einv);
    }
    if (f_ah) {
      const double eh = 1.0 / (r2 + e2h);
      const double du_ = R->u[i] - R->u[j];
      const double lH = (ldu[i] - ldu[j]) * du_;
      const double ps = R->psi[i] + R->psi[j];
      const double H = m * ps * F * r2 * M * eh;
      lu[i] += (ldu[i] - ldu[j]) * H;
      lu[j] -= (ldu[i] - ldu[j]) * H;
      lpsi[i] += lH * m * F * r2 * M * eh;
      lpsi[j] += lH * m * F * r2 * M * eh;
      lF += lH * m * ps * r2 * M * eh;
      lr2 += lH * m * ps * F * M * e2h * eh * eh;
      lM += lH * m * ps * F * r2 * eh;
    }
    if (f_xs) {
      const double Dvx = lavx[i] - lavx[j], Dvy = lavy[i] - lavy[j];
      const double w = P->xsph_eps * m * M * W;
      const double lw = -(Dvx * dvx + Dvy * dvy);
      ldvx += -w * Dvx;
      ldvy += -w * Dvy;
      lM += lw * P->xsph_eps * m * W;
      lW += lw * P->xsph_eps * m * M;
    }
    /* M = 2 Qi Qj / (Qi + Qj) */
    lQ[i] += lM * 2.0 * Qj * Qj / (sQ * sQ);
    lQ[j] += lM * 2.0 * Qi * Qi / (sQ * sQ);
    /* F(q), W(q), q(r2) */
    const double s = 1.0 - 0.5 * q;
    const double lq = lF * 3.0 * C1 * (q - 2.0) * (q - 2.0) + lW * a_n * s * s * s * (-5.0 * q);
    lr2 += lq / (2.0 * h * h * q);
    /* vr = dv . r, r2 = r . r */
    ldvx += lvr * rx;
    ldvy += lvr * ry;
    lrx += lvr * dvx + 2.0 * lr2 * rx;
    lry += lvr * dvy + 2.0 * lr2 * ry;
    lx[i] += lrx;
    ly[i] += lry;
    lx[j] -= lrx;
    ly[j] -= lry;
    lvx[i] += ldvx;
    lvy[i] += ldvy;
    lvx[j] -= ldvx;
    lvy[j] -= ldvy;
  }
  /* 2. per particle: Q, A, psi, EOS */
  double* lS = dz(n);
  for (int i = 0; i < n; ++i) {
    const double rho = R->rhoe[i];
    lp[i] += lA[i] * R->Q[i];
    double lrhoe = -R->Q[i] * R->Q[i] * (lQ[i] + lA[i] * R->p[i]);
    const double dv = R->S[i] / rho;
    const double dpsi = (dv < 0) ? -2.0 * P->ah_g2 * h * h : 0.0;
    const double ldv = lpsi[i] * dpsi;
    lS[i] = ldv / rho;
    lrhoe += -ldv * R->S[i] / (rho * rho);
    if (is_rate(P, np, i) && !R->clamp[i]) {
      const double tt = R->rho[i] / P->rho0;
      lrho[i] += lrhoe + lp[i] * 7.0 * B / P->rho0 * pow(tt, 6.0);
    }
  }
  /* 3. density pairs: drho_i, drho_j += m F vr; S_i, S_j -= m F vr */
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    const double lt = (is_rate(P, np, i) ? ldrho[i] : 0.0) + (is_rate(P, np, j) ? ldrho[j] : 0.0) -
                      (lS[i] + lS[j]);
    if (lt == 0.0) continue;
    const double rx = R->x[i] - R->x[j], ry = R->y[i] - R->y[j];
    const double r2 = rx * rx + ry * ry;
    double q, F, W;
    wend(P, r2, &q, &F, &W);
    const double dvx = R->vx[i] - R->vx[j], dvy = R->vy[i] - R->vy[j];
    const double vr = dvx * rx + dvy * ry;
    const double lvr = lt * m * F;
    const double lF = lt * m * vr;
    const double lr2 = lF * 3.0 * C1 * (q - 2.0) * (q - 2.0) / (2.0 * h * h * q);
    const double lrx = lvr * dvx + 2.0 * lr2 * rx, lry = lvr * dvy + 2.0 * lr2 * ry;
    lx[i] += lrx;
    ly[i] += lry;
    lx[j] -= lrx;
    ly[j] -= lry;
    lvx[i] += lvr * rx;
    lvy[i] += lvr * ry;
    lvx[j] -= lvr * rx;
    lvy[j] -= lvr * ry;
  }
  free(lS);
  free(lQ);
  free(lA);
  free(lpsi);
}

typedef struct {
  double *x, *y, *vx, *vy, *rho, *u;     /* state at the start of the step */
  double *pax, *pay, *pdrho, *pdu;       /* rates of the previous step */
} snap_t;

static void snap_alloc(snap_t* s, int n) {
  double** f[] = {&s->x, &s->y, &s->vx, &s->vy, &s->rho, &s->u, &s->pax, &s->pay, &s->pdrho, &s->pdu};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k) *f[k] = dz(n);
}

static void snap_free(snap_t* s) {
  double* f[] = {s->x, s->y, s->vx, s->vy, s->rho, s->u, s->pax, s->pay, s->pdrho, s->pdu};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k) free(f[k]);
}

/* half-step inputs of the rates from a snapshot (time_integration.f:136-154) */
static void predict(const orc_params* P, int n, int np, int its, const snap_t* sn, rates_t* R,
                    int* uclamp) {
  const double h2 = P->dt / 2.0;
  for (int i = 0; i < n; ++i) {
    R->x[i] = sn->x[i];
    R->y[i] = sn->y[i];
    R->vx[i] = sn->vx[i];
    R->vy[i] = sn->vy[i];
    R->rho[i] = sn->rho[i];
    R->u[i] = sn->u[i];
    uclamp[i] = 0;
    if (its != 1 && is_rate(P, np, i)) {
      double u = sn->u[i] + h2 * sn->pdu[i];
      if (u < 0) {
        u = 0;
        uclamp[i] = 1;
      }
      R->u[i] = u;
      R->rho[i] = sn->rho[i] + h2 * sn->pdrho[i];
      if (i < np) {
        R->vx[i] = sn->vx[i] + h2 * sn->pax[i];
        R->vy[i] = sn->vy[i] + h2 * sn->pay[i];
      }
    }
  }
}

int orc_adjoint(const orc_params* P, const orc_state* s0, int nsteps, const int* objmask,
                double* J_out, double* dJ_dalpha, double* dJ_dvx0, double* dJ_dvy0) {
  const int n = s0->n, np = s0->np;
  if (P->density_method != 0) return -5; /* continuity density only */
  if ((P->flags & (ORC_VISC | ORC_HEAT_SOURCE | ORC_CONDUCTION | ORC_PERIODIC_X)) || P->sle || P->eos != ORC_EOS_TAIT ||
      P->dep)
    return -5; /* Tait, no viscosity, no heat source */
  const double dt = P->dt, h2 = dt / 2.0;
  snap_t* sn = (snap_t*)calloc((size_t)nsteps + 1, sizeof(snap_t));
  for (int k = 0; k <= nsteps; ++k) snap_alloc(&sn[k], n);
  memcpy(sn[0].x, s0->x, sizeof(double) * n);
  memcpy(sn[0].y, s0->y, sizeof(double) * n);
  memcpy(sn[0].vx, s0->vx, sizeof(double) * n);
  memcpy(sn[0].vy, s0->vy, sizeof(double) * n);
  memcpy(sn[0].rho, s0->rho, sizeof(double) * n);
  memcpy(sn[0].u, s0->u, sizeof(double) * n);
  rates_t R;
  R.x = dz(n); R.y = dz(n); R.vx = dz(n); R.vy = dz(n); R.rho = dz(n); R.u = dz(n);
  rates_alloc(&R, n);
  int* uclamp = (int*)calloc((size_t)n, sizeof(int));
  int* cclamp = (int*)calloc((size_t)n, sizeof(int));
  const int cap = n * 32 + 64;
  int* pi = (int*)malloc(sizeof(int) * cap);
  int* pj = (int*)malloc(sizeof(int) * cap);
  int niac = 0, rc = 0;
  double J = 0;
  /* forward */
  for (int k = 0; k < nsteps && rc == 0; ++k) {
    const int its = k + 1;
    predict(P, n, np, its, &sn[k], &R, uclamp);
    rc = orc_link_list(P, n, R.x, R.y, cap, &niac, pi, pj);
    if (rc) break;
    rates_forward(P, n, np, niac, pi, pj, &R);
    for (int i = 0; i < n; ++i)
      if (objmask[i]) J += dt * R.p[i] * R.p[i];
    snap_t* o = &sn[k + 1];
    for (int i = 0; i < n; ++i) {
      o->x[i] = sn[k].x[i];
      o->y[i] = sn[k].y[i];
      o->vx[i] = sn[k].vx[i];
      o->vy[i] = sn[k].vy[i];
      o->rho[i] = sn[k].rho[i];
      o->u[i] = sn[k].u[i];
      o->pax[i] = R.ax[i];
      o->pay[i] = R.ay[i];
      o->pdrho[i] = R.drho[i];
      o->pdu[i] = R.du[i];
      if (!is_rate(P, np, i)) continue;
      const double c = (its != 1) ? dt : h2;
      double u = sn[k].u[i] + c * R.du[i];
      if (u < 0) u = 0;
      o->u[i] = u;
      o->rho[i] = sn[k].rho[i] + c * R.drho[i];
      if (i < np) {
        if (its != 1) {
          o->vx[i] = sn[k].vx[i] + dt * R.ax[i] + R.avx[i];
          o->vy[i] = sn[k].vy[i] + dt * R.ay[i] + R.avy[i];
        } else {
          o->vx[i] = sn[k].vx[i] + h2 * R.ax[i];
          o->vy[i] = sn[k].vy[i] + h2 * R.ay[i];
        }
        o->x[i] = sn[k].x[i] + dt * o->vx[i];
        o->y[i] = sn[k].y[i] + dt * o->vy[i];
      }
    }
  }
  /* backward */
  double *lx = dz(n), *ly = dz(n), *lvx = dz(n), *lvy = dz(n), *lrho = dz(n), *lu = dz(n);
  double *lcax = dz(n), *lcay = dz(n), *lcdrho = dz(n), *lcdu = dz(n);
  double *lax = dz(n), *lay = dz(n), *ldu = dz(n), *lavx = dz(n), *lavy = dz(n), *ldrho = dz(n);
  double *lp = dz(n), *nlx = dz(n), *nly = dz(n), *nlvx = dz(n), *nlvy = dz(n), *nlrho = dz(n),
         *nlu = dz(n);
  double lalpha = 0;
  for (int k = nsteps - 1; k >= 0 && rc == 0; --k) {
    const int its = k + 1;
    predict(P, n, np, its, &sn[k], &R, uclamp);
    rc = orc_link_list(P, n, R.x, R.y, cap, &niac, pi, pj);
    if (rc) break;
    rates_forward(P, n, np, niac, pi, pj, &R);
    /* corrector adjoint: (lx, lv, lrho, lu) of the step output and
       (lcax ...) of the carried rates -> rate adjoints and the state part */
    for (int i = 0; i < n; ++i) {
      nlx[i] = lx[i];
      nly[i] = ly[i];
      nlvx[i] = lvx[i];
      nlvy[i] = lvy[i];
      nlrho[i] = 0;
      nlu[i] = 0;
      lax[i] = lcax[i];
      lay[i] = lcay[i];
      ldrho[i] = lcdrho[i];
      ldu[i] = lcdu[i];
      lavx[i] = lavy[i] = 0;
      lp[i] = objmask[i] ? 2.0 * dt * R.p[i] : 0.0;
      if (!is_rate(P, np, i)) {
        nlrho[i] = lrho[i];
        nlu[i] = lu[i];
        continue;
      }
      const double c = (its != 1) ? dt : h2;
      nlrho[i] = lrho[i];
      ldrho[i] += c * lrho[i];
      const double uu = sn[k].u[i] + c * R.du[i];
      if (!(uu < 0)) {
        nlu[i] = lu[i];
        ldu[i] += c * lu[i];
      }
      if (i < np) {
        const double lvtx = lvx[i] + dt * lx[i];
        const double lvty = lvy[i] + dt * ly[i];
        nlvx[i] = lvtx;
        nlvy[i] = lvty;
        const double cv = (its != 1) ? dt : h2;
        lax[i] += cv * lvtx;
        lay[i] += cv * lvty;
        if (its != 1) {
          lavx[i] = lvtx;
          lavy[i] = lvty;
        }
      }
    }
    /* rates adjoint -> inputs of the rates (x, v_half, rho_half, u_half) */
    double *rlx = dz(n), *rly = dz(n), *rlvx = dz(n), *rlvy = dz(n), *rlrho = dz(n), *rlu = dz(n);
    rates_adjoint(P, n, np, niac, pi, pj, &R, lax, lay, ldu, lavx, lavy, ldrho, lp, rlx, rly, rlvx,
                  rlvy, rlrho, rlu, &lalpha);
    /* predictor adjoint */
    for (int i = 0; i < n; ++i) {
      nlx[i] += rlx[i];
      nly[i] += rly[i];
      lcax[i] = lcay[i] = lcdrho[i] = lcdu[i] = 0;
      nlvx[i] += rlvx[i];
      nlvy[i] += rlvy[i];
      nlrho[i] += rlrho[i];
      if (its != 1 && is_rate(P, np, i)) {
        lcdrho[i] = h2 * rlrho[i];
        if (!uclamp[i]) {
          nlu[i] += rlu[i];
          lcdu[i] = h2 * rlu[i];
        }
        if (i < np) {
          lcax[i] = h2 * rlvx[i];
          lcay[i] = h2 * rlvy[i];
        }
      } else {
        nlu[i] += rlu[i];
      }
    }
    free(rlx); free(rly); free(rlvx); free(rlvy); free(rlrho); free(rlu);
    memcpy(lx, nlx, sizeof(double) * n);
    memcpy(ly, nly, sizeof(double) * n);
    memcpy(lvx, nlvx, sizeof(double) * n);
    memcpy(lvy, nlvy, sizeof(double) * n);
    memcpy(lrho, nlrho, sizeof(double) * n);
    memcpy(lu, nlu, sizeof(double) * n);
  }
  if (J_out) *J_out = J;
  if (dJ_dalpha) *dJ_dalpha = lalpha;
  if (dJ_dvx0) memcpy(dJ_dvx0, lvx, sizeof(double) * n);
  if (dJ_dvy0) memcpy(dJ_dvy0, lvy, sizeof(double) * n);
  double* fr[] = {lx, ly, lvx, lvy, lrho, lu, lcax, lcay, lcdrho, lcdu, lax, lay, ldu,
                  lavx, lavy, ldrho, lp, nlx, nly, nlvx, nlvy, nlrho, nlu};
  for (size_t k = 0; k < sizeof(fr) / sizeof(fr[0]); ++k) free(fr[k]);
  for (int k = 0; k <= nsteps; ++k) snap_free(&sn[k]);
  free(sn);
  free(R.x); free(R.y); free(R.vx); free(R.vy); free(R.rho); free(R.u);
  rates_free(&R);
  free(uclamp);
  free(cclamp);
  free(pi);
  free(pj);
  return rc;
}
