/*
 * armando_oracle3d.c -- CPU restatement of the WCSPH step in 3D.
 *
 * TEST INFRASTRUCTURE ONLY (see armando_oracle.h).  FP64 gathers over a
 * fixed hash grid of cells of size cell >= 2h, key = ix + iy nx + iz nx ny;
 * particles in (cell, id) order; the 27 neighbour cells of a particle are
 * nine rows (dz, dy in -1..1, z outer) of three consecutive x-cells, each a
 * contiguous range of the sorted order.  The pair terms are those of the 2D
 * restatement with three components:
 *   kernel        kernel.f:31-56 (cubic, dim = 3) and the Wendland C2 3D
 *                 with the normalised a = 21/(16 pi h^3)
 *   sum_density   density.f:1-62        con_density  density.f:166-223
 *   int_force     internal_force.f:178-437 (pressure part, pa_sph 0 and 1)
 *   art_visc      art_visc.f:55-101     art_heat     art_heat.f:42-97
 *   av_vel        av_vel.f:37-56        leapfrog     time_integration.f:131-235
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "armando_oracle.h"

static const double PI_F = 3.1415927410125732;
static const double TWO_THIRDS_F = 0.6666666865348816;

/* W, and F with grad_i W = F (x_i - x_j). */
static double kern3(const orc_params* P, double r, double* F) {
  const double h = P->h, q = r / h;
  if (P->kernel == ORC_K_WENDLAND3D) {
    if (q < 2.0) {
      const double a = 21.0 / (16.0 * M_PI * h * h * h);
      const double s = 1.0 - 0.5 * q, s2 = s * s, t = q - 2.0;
      *F = (0.625 * a / (h * h)) * (t * t * t);
      return a * (s2 * s2) * (1.0 + 2.0 * q);
    }
    *F = 0.0;
    return 0.0;
  }
  /* kernel.f:31-56 with the dim = 3 factor */
  const double factor = 3.0 / (2.0 * PI_F * (h * h * h));
  if (q >= 0.0 && q <= 1.0) {
    *F = factor * (-2.0 + 1.5 * q) / (h * h);
    return factor * (TWO_THIRDS_F - q * q + (q * q * q) / 2.0);
  } else if (q > 1.0 && q <= 2.0) {
    const double g = -factor * 1.0 / 6.0 * 3.0 * ((2.0 - q) * (2.0 - q)) / h;
    *F = g / r;
    return factor * 1.0 / 6.0 * ((2.0 - q) * (2.0 - q) * (2.0 - q));
  }
  *F = 0.0;
  return 0.0;
}

static long long key3(const orc_params* P, double x, double y, double z) {
  const double tx = (x - P->ox) / P->cell, ty = (y - P->oy) / P->cell;
  const double tz = (z - P->oz) / P->cell;
  if (!(tx >= 0.0) || !(ty >= 0.0) || !(tz >= 0.0)) return -1;
  const int ix = (int)tx, iy = (int)ty, iz = (int)tz;
  if (ix >= P->nx || iy >= P->ny || iz >= P->nz) return -1;
  return ix + (long long)P->nx * (iy + (long long)P->ny * iz);
}

int orc3_keys(const orc_params* P, int n, const double* x, const double* y, const double* z,
              int32_t* key) {
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    const long long k = key3(P, x[i], y[i], z[i]);
    key[i] = (int32_t)k;
    if (k < 0) bad++;
  }
  return bad;
}

typedef struct {
  int32_t *key, *sorted, *start;
  int ncell;
} cells3_t;

static int build3(const orc_params* P, int n, const double* x, const double* y,
                  const double* z, cells3_t* C) {
  C->ncell = P->nx * P->ny * P->nz;
  C->key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  C->sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  C->start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)C->ncell + 1));
  const int bad = orc3_keys(P, n, x, y, z, C->key);
  orc_sort_cells(n, C->key, C->ncell, C->sorted, C->start);
  return bad;
}

static void free3(cells3_t* C) {
  free(C->key);
  free(C->sorted);
  free(C->start);
}

/* sorted-position ranges of the nine rows around the cell of key k */
static int rows3(const orc_params* P, const cells3_t* C, int k, int t0[9], int t1[9]) {
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const int cx = k % nx, cy = (k / nx) % ny, cz = k / (nx * ny);
  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < nx ? cx + 1 : nx - 1;
  int nr = 0;
  for (int rz = cz - 1; rz <= cz + 1; ++rz)
    for (int ry = cy - 1; ry <= cy + 1; ++ry) {
      if (rz < 0 || rz >= nz || ry < 0 || ry >= ny) continue;
      const int base = (rz * ny + ry) * nx;
      t0[nr] = C->start[base + x0];
      t1[nr] = C->start[base + x1 + 1];
      nr++;
    }
  return nr;
}

/* ------------------------------------------------------------------ */
/* per-routine outputs in gather form                                   */
/* ------------------------------------------------------------------ */

int orc3_routines(const orc_params* P, int n, const double* x, const double* y,
                  const double* z, const double* vx, const double* vy, const double* vz,
                  const double* rho, const double* p, const double* u, const double* c,
                  double* rho_sum, double* drho, double* if_a, double* if_du, double* av_a,
                  double* av_du, double* ah_du, double* divv, double* xs) {
  cells3_t C;
  int bad = build3(P, n, x, y, z, &C);
  if (bad) {
    free3(&C);
    return bad;
  }
  const double m = P->mass, h = P->h, cut2 = (2.0 * h) * (2.0 * h);
  double F0;
  const double w0 = kern3(P, 0.0, &F0);
  const double av_e2 = (h * P->av_eta) * (h * P->av_eta);
  const double ah_e2 = (h * P->ah_eta) * (h * P->ah_eta);
  /* pass 1: densities and the art_heat divergence */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int t0[9], t1[9];
    const int nr = rows3(P, &C, C.key[i], t0, t1);
    double rs = m * w0, dr = 0.0, dv = 0.0;
    for (int r = 0; r < nr; ++r)
      for (int t = t0[r]; t < t1[r]; ++t) {
        const int j = C.sorted[t];
        if (j == i) continue;
        const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < cut2)) continue;
        double F;
        const double w = kern3(P, sqrt(r2), &F);
        rs += m * w;
        const double vdw = (vx[i] - vx[j]) * F * dx + (vy[i] - vy[j]) * F * dy +
                           (vz[i] - vz[j]) * F * dz;
        dr += m * vdw;
        dv += m / rho[i] * ((vx[j] - vx[i]) * F * dx + (vy[j] - vy[i]) * F * dy +
                            (vz[j] - vz[i]) * F * dz);
      }
    rho_sum[i] = rs;
    drho[i] = dr;
    divv[i] = dv;
  }
  /* pass 2: forces */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int t0[9], t1[9];
    const int nr = rows3(P, &C, C.key[i], t0, t1);
    double fa[3] = {0, 0, 0}, fdu = 0, va[3] = {0, 0, 0}, vdu = 0, hdu = 0, xv[3] = {0, 0, 0};
    for (int r = 0; r < nr; ++r)
      for (int t = t0[r]; t < t1[r]; ++t) {
        const int j = C.sorted[t];
        if (j == i) continue;
        const double d[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(r2 < cut2)) continue;
        double F;
        const double w = kern3(P, sqrt(r2), &F);
        const double dvv[3] = {vx[i] - vx[j], vy[i] - vy[j], vz[i] - vz[j]};
        double vdw = 0.0, vr = 0.0;
        for (int k = 0; k < 3; ++k) {
          vdw += dvv[k] * F * d[k];
          vr += dvv[k] * d[k];
        }
        /* int_force */
        double a;
        if (P->pa_sph == 0)
          a = (p[i] + p[j]) / (rho[i] * rho[j]);
        else
          a = p[i] / (rho[i] * rho[i]) + p[j] / (rho[j] * rho[j]);
        for (int k = 0; k < 3; ++k) fa[k] -= m * a * F * d[k];
        fdu += m * a * vdw;
        /* art_visc */
        const double mrho = 0.5 * (rho[i] + rho[j]);
        if (vr < 0.0) {
          const double phi = (h * vr) / (r2 + av_e2);
          const double mc = 0.5 * (c[i] + c[j]);
          const double piv = (P->av_beta * phi - P->av_alpha * mc) * phi / mrho;
          for (int k = 0; k < 3; ++k) va[k] -= m * piv * F * d[k];
          vdu += 0.5 * m * piv * vdw;
        }
        /* art_heat (second pass) */
        const double psi_j_dv = divv[j];
        const double pi_ = P->ah_g1 * h * c[i] + P->ah_g2 * h * h * (fabs(divv[i]) - divv[i]);
        const double pj_ = P->ah_g1 * h * c[j] + P->ah_g2 * h * h * (fabs(psi_j_dv) - psi_j_dv);
        const double ps = (2.0 * (0.5 * (pi_ + pj_))) / (mrho * (r2 + ah_e2)) * (F * r2);
        hdu += m * ps * (u[i] - u[j]);
        /* av_vel */
        for (int k = 0; k < 3; ++k) xv[k] -= P->xsph_eps * m / mrho * dvv[k] * w;
      }
    for (int k = 0; k < 3; ++k) {
      if_a[3 * i + k] = fa[k];
      av_a[3 * i + k] = va[k];
      xs[3 * i + k] = xv[k];
    }
    if_du[i] = 0.5 * fdu;
    av_du[i] = vdu;
    ah_du[i] = hdu;
  }
  free3(&C);
  return 0;
}

/* ------------------------------------------------------------------ */
/* state, single step, leapfrog                                         */
/* ------------------------------------------------------------------ */

#define ORC3_FIELDS(s)                                                                    \
  &s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->rho, &s->u, &s->p, &s->c, &s->drho, \
      &s->du, &s->ax, &s->ay, &s->az, &s->avx, &s->avy, &s->avz, &s->divv, &s->rho_min,   \
      &s->u_min, &s->vx_min, &s->vy_min, &s->vz_min

orc3_state* orc3_state_new(int n, int np) {
  orc3_state* s = (orc3_state*)calloc(1, sizeof(orc3_state));
  s->n = n;
  s->np = np;
  double** f[] = {ORC3_FIELDS(s)};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k)
    *f[k] = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  return s;
}

void orc3_state_free(orc3_state* s) {
  if (!s) return;
  double** f[] = {ORC3_FIELDS(s)};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k) free(*f[k]);
  free(s);
}

static int nrate3(const orc_params* P, const orc3_state* s) {
  return (P->flags & ORC_DBC) ? s->n : s->np;
}

int orc3_single_step_gather(const orc_params* P, orc3_state* s) {
  if ((P->flags & (ORC_VISC | ORC_HEAT_SOURCE | ORC_CONDUCTION | ORC_PERIODIC_X)) || P->sle)
    return -2; /* 2D only */
  const int n = s->n, np = s->np, nr = nrate3(P, s);
  const double m = P->mass, h = P->h, cut2 = (2.0 * h) * (2.0 * h);
  cells3_t C;
  int bad = build3(P, n, s->x, s->y, s->z, &C);
  if (bad) {
    free3(&C);
    return bad;
  }
  const double *x = s->x, *y = s->y, *z = s->z, *vx = s->vx, *vy = s->vy, *vz = s->vz;
  double F0;
  const double w0 = kern3(P, 0.0, &F0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int t0[9], t1[9];
    const int nrow = rows3(P, &C, C.key[i], t0, t1);
    double dr = 0.0, rs = m * w0;
    for (int r = 0; r < nrow; ++r)
      for (int t = t0[r]; t < t1[r]; ++t) {
        const int j = C.sorted[t];
        if (j == i) continue;
        const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
        double r2 = dx * dx;
        r2 = r2 + dy * dy;
        r2 = r2 + dz * dz;
        if (!(r2 < cut2)) continue;
        double F;
        const double w = kern3(P, sqrt(r2), &F);
        double vcc = (vx[i] - vx[j]) * (F * dx);
        vcc = vcc + (vy[i] - vy[j]) * (F * dy);
        vcc = vcc + (vz[i] - vz[j]) * (F * dz);
        dr = dr + m * vcc;
        rs = rs + m * w;
      }
    if (i < nr) {
      if (P->density_method == 0)
        s->drho[i] = dr;
      else
        s->rho[i] = rs;
    } else {
      s->drho[i] = 0.0;
    }
    s->divv[i] = -dr;
  }
  orc_eos(P, nr, s->rho, s->p, s->c);
  double* psi = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const double dv = s->divv[i] / s->rho[i];
    s->divv[i] = dv;
    psi[i] = P->ah_g1 * h * s->c[i] + P->ah_g2 * h * h * (fabs(dv) - dv);
  }
  const int grav = (P->flags & ORC_GRAVITY) != 0;
  const int f_av = (P->flags & ORC_ART_VISC) != 0;
  const int f_ah = (P->flags & ORC_ART_HEAT) != 0;
  const int f_xs = (P->flags & ORC_AV_VEL) != 0;
  const double *rho = s->rho, *p = s->p, *c = s->c, *u = s->u;
  const double av_e2 = (h * P->av_eta) * (h * P->av_eta);
  const double ah_e2 = (h * P->ah_eta) * (h * P->ah_eta);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int t0[9], t1[9];
    const int nrow = rows3(P, &C, C.key[i], t0, t1);
    double f[3] = {0, 0, 0}, due = 0, avf[3] = {0, 0, 0}, avdu = 0, ahdu = 0;
    double xs[3] = {0, 0, 0};
    const double pr_i = p[i] / (rho[i] * rho[i]);
    for (int r = 0; r < nrow; ++r)
      for (int t = t0[r]; t < t1[r]; ++t) {
        const int j = C.sorted[t];
        if (j == i) continue;
        const double d[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
        double r2 = d[0] * d[0];
        r2 = r2 + d[1] * d[1];
        r2 = r2 + d[2] * d[2];
        if (!(r2 < cut2)) continue;
        double F;
        const double w = kern3(P, sqrt(r2), &F);
        const double dv[3] = {vx[i] - vx[j], vy[i] - vy[j], vz[i] - vz[j]};
        double vdw = dv[0] * (F * d[0]);
        vdw = vdw + dv[1] * (F * d[1]);
        vdw = vdw + dv[2] * (F * d[2]);
        double a;
        if (P->pa_sph == 0)
          a = (p[i] + p[j]) / (rho[i] * rho[j]);
        else
          a = pr_i + p[j] / (rho[j] * rho[j]);
        for (int k = 0; k < 3; ++k) f[k] = f[k] - m * a * (F * d[k]);
        due = due + m * a * vdw;
        const double mrho = 0.5 * (rho[i] + rho[j]);
        if (f_av) {
          double vr = dv[0] * d[0];
          vr = vr + dv[1] * d[1];
          vr = vr + dv[2] * d[2];
          if (vr < 0.0) {
            const double phi = (h * vr) / (r2 + av_e2);
            const double mc = 0.5 * (c[i] + c[j]);
            const double piv = (P->av_beta * phi - P->av_alpha * mc) * phi / mrho;
            for (int k = 0; k < 3; ++k) avf[k] = avf[k] - m * piv * (F * d[k]);
            avdu = avdu + 0.5 * m * piv * vdw;
          }
        }
        if (f_ah) {
          const double pij = 0.5 * (psi[i] + psi[j]);
          const double ps = (2.0 * pij) / (mrho * (r2 + ah_e2)) * (F * r2);
          ahdu = ahdu + m * ps * (u[i] - u[j]);
        }
        if (f_xs)
          for (int k = 0; k < 3; ++k) xs[k] = xs[k] - P->xsph_eps * m / mrho * dv[k] * w;
      }
    const int g = grav && i < np;
    s->ax[i] = f[0] + (g ? P->gx : 0.0) + avf[0];
    s->ay[i] = f[1] + (g ? P->gy : 0.0) + avf[1];
    s->az[i] = f[2] + (g ? P->gz : 0.0) + avf[2];
    s->du[i] = 0.5 * due + avdu + ahdu;
    if (P->dep && P->dep_on && i < np) {
      int in;
      const double v = orc_deposition_at(P, s->x[i], s->y[i], s->z[i], &in);
      if (in) s->du[i] = s->du[i] + v / s->rho[i];
    }
    s->avx[i] = xs[0];
    s->avy[i] = xs[1];
    s->avz[i] = xs[2];
  }
  free(psi);
  free3(&C);
  return 0;
}

void orc3_predictor(const orc_params* P, orc3_state* s) {
  const double dt = P->dt;
  const int nr = nrate3(P, s);
  for (int i = 0; i < nr; ++i) {
    s->u_min[i] = s->u[i];
    s->u[i] = s->u[i] + (dt / 2.0) * s->du[i];
    if (s->u[i] < 0.0) s->u[i] = 0.0;
    if (P->density_method == 0) {
      s->rho_min[i] = s->rho[i];
      s->rho[i] = s->rho[i] + (dt / 2.0) * s->drho[i];
    }
    if (i < s->np) {
      s->vx_min[i] = s->vx[i];
      s->vx[i] = s->vx[i] + (dt / 2.0) * s->ax[i];
      s->vy_min[i] = s->vy[i];
      s->vy[i] = s->vy[i] + (dt / 2.0) * s->ay[i];
      s->vz_min[i] = s->vz[i];
      s->vz[i] = s->vz[i] + (dt / 2.0) * s->az[i];
    }
  }
}

void orc3_corrector(const orc_params* P, orc3_state* s, int its) {
  const double dt = P->dt;
  const int nr = nrate3(P, s);
  for (int i = 0; i < nr; ++i) {
    const int fl = i < s->np;
    if (its != 1) {
      s->u[i] = s->u_min[i] + dt * s->du[i];
      if (s->u[i] < 0.0) s->u[i] = 0.0;
      if (P->density_method == 0) s->rho[i] = s->rho_min[i] + dt * s->drho[i];
      if (fl) {
        s->vx[i] = s->vx_min[i] + dt * s->ax[i] + s->avx[i];
        s->x[i] = s->x[i] + dt * s->vx[i];
        s->vy[i] = s->vy_min[i] + dt * s->ay[i] + s->avy[i];
        s->y[i] = s->y[i] + dt * s->vy[i];
        s->vz[i] = s->vz_min[i] + dt * s->az[i] + s->avz[i];
        s->z[i] = s->z[i] + dt * s->vz[i];
      }
    } else {
      s->u[i] = s->u[i] + (dt / 2.0) * s->du[i];
      if (s->u[i] < 0.0) s->u[i] = 0.0;
      if (P->density_method == 0) s->rho[i] = s->rho[i] + (dt / 2.0) * s->drho[i];
      if (fl) {
        s->vx[i] = s->vx[i] + (dt / 2.0) * s->ax[i];
        s->x[i] = s->x[i] + dt * s->vx[i];
        s->vy[i] = s->vy[i] + (dt / 2.0) * s->ay[i];
        s->y[i] = s->y[i] + dt * s->vy[i];
        s->vz[i] = s->vz[i] + (dt / 2.0) * s->az[i];
        s->z[i] = s->z[i] + dt * s->vz[i];
      }
    }
  }
}

int orc3_run(const orc_params* P, orc3_state* s, int its0, int nsteps) {
  orc_params Q = *P;
  for (int its = its0; its < its0 + nsteps; ++its) {
    Q.dep_on = its * P->dt <= P->dep_tend; /* single_step.f:219 */
    P = &Q;
    if (its != 1) orc3_predictor(P, s);
    const int rc = orc3_single_step_gather(P, s);
    if (rc == -2) return -2;
    if (rc) return -3;
    orc3_corrector(P, s, its);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* pair sets                                                            */
/* ------------------------------------------------------------------ */

static int cmp_i(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

typedef int (*pred3_t)(const void* ctx, int i, int j);

static int pairs3(const orc_params* P, int n, const int32_t* key, pred3_t pred, const void* ctx,
                  int cap, int* niac, int* pair_i, int* pair_j) {
  cells3_t C;
  C.ncell = P->nx * P->ny * P->nz;
  C.key = (int32_t*)key;
  C.sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  C.start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)C.ncell + 1));
  orc_sort_cells(n, key, C.ncell, C.sorted, C.start);
  int* buf = (int*)malloc(sizeof(int) * 16384);
  int cnt = 0, rc = 0;
  for (int i = 0; i < n && rc == 0; ++i) {
    if (key[i] < 0) continue;
    int t0[9], t1[9], nb = 0;
    const int nr = rows3(P, &C, key[i], t0, t1);
    for (int r = 0; r < nr; ++r)
      for (int t = t0[r]; t < t1[r]; ++t) {
        const int j = C.sorted[t];
        if (j > i && pred(ctx, i, j) && nb < 16384) buf[nb++] = j;
      }
    qsort(buf, (size_t)nb, sizeof(int), cmp_i);
    for (int k = 0; k < nb; ++k) {
      if (cnt >= cap) {
        rc = -1;
        break;
      }
      pair_i[cnt] = i;
      pair_j[cnt] = buf[k];
      cnt++;
    }
  }
  *niac = cnt;
  free(buf);
  free(C.sorted);
  free(C.start);
  return rc;
}

typedef struct {
  const double *x, *y, *z;
  double cut2;
} f64_3_ctx;

static int pred_f64_3(const void* v, int i, int j) {
  const f64_3_ctx* c = (const f64_3_ctx*)v;
  const double dx = c->x[i] - c->x[j], dy = c->y[i] - c->y[j], dz = c->z[i] - c->z[j];
  double r2 = dx * dx;
  r2 = r2 + dy * dy;
  r2 = r2 + dz * dz;
  return r2 < c->cut2;
}

int orc3_pairs_f64(const orc_params* P, int n, const double* x, const double* y,
                   const double* z, int cap, int* niac, int* pair_i, int* pair_j) {
  int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  orc3_keys(P, n, x, y, z, key);
  f64_3_ctx ctx = {x, y, z, (2.0 * P->h) * (2.0 * P->h)};
  const int rc = pairs3(P, n, key, pred_f64_3, &ctx, cap, niac, pair_i, pair_j);
  free(key);
  return rc;
}

void orc3_split_f32(const orc_params* P, int n, const double* x, const double* y,
                    const double* z, int32_t* cx, int32_t* cy, int32_t* cz, float* ox,
                    float* oy, float* oz) {
  for (int i = 0; i < n; ++i) {
    const int ix = (int)((x[i] - P->ox) / P->cell), iy = (int)((y[i] - P->oy) / P->cell);
    const int iz = (int)((z[i] - P->oz) / P->cell);
    cx[i] = ix;
    cy[i] = iy;
    cz[i] = iz;
    ox[i] = (float)(x[i] - (P->ox + (double)ix * P->cell));
    oy[i] = (float)(y[i] - (P->oy + (double)iy * P->cell));
    oz[i] = (float)(z[i] - (P->oz + (double)iz * P->cell));
  }
}

typedef struct {
  const int32_t *cx, *cy, *cz;
  const float *ox, *oy, *oz;
  float cellf, cut2f;
} f32_3_ctx;

static int pred_f32_3(const void* v, int i, int j) {
  const f32_3_ctx* c = (const f32_3_ctx*)v;
  volatile float tx = c->ox[i] - c->ox[j];
  volatile float ux = (float)(c->cx[i] - c->cx[j]) * c->cellf;
  volatile float dx = tx + ux;
  volatile float ty = c->oy[i] - c->oy[j];
  volatile float uy = (float)(c->cy[i] - c->cy[j]) * c->cellf;
  volatile float dy = ty + uy;
  volatile float tz = c->oz[i] - c->oz[j];
  volatile float uz = (float)(c->cz[i] - c->cz[j]) * c->cellf;
  volatile float dz = tz + uz;
  volatile float a = dx * dx;
  volatile float b = dy * dy;
  volatile float cc = dz * dz;
  volatile float s = a + b;
  volatile float r2 = s + cc;
  return r2 < c->cut2f;
}

int orc3_pairs_f32(const orc_params* P, int n, const int32_t* cx, const int32_t* cy,
                   const int32_t* cz, const float* ox, const float* oy, const float* oz,
                   int cap, int* niac, int* pair_i, int* pair_j) {
  int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    const int ok = cx[i] >= 0 && cx[i] < P->nx && cy[i] >= 0 && cy[i] < P->ny && cz[i] >= 0 &&
                   cz[i] < P->nz;
    key[i] = ok ? cx[i] + P->nx * (cy[i] + P->ny * cz[i]) : -1;
  }
  f32_3_ctx ctx = {cx, cy, cz, ox, oy, oz, (float)P->cell,
                   (float)((2.0 * P->h) * (2.0 * P->h))};
  const int rc = pairs3(P, n, key, pred_f32_3, &ctx, cap, niac, pair_i, pair_j);
  free(key);
  return rc;
}
