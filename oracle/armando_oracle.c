/*
 * armando_oracle.c -- CPU restatement of Armando's 2D WCSPH time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP library.  Built with
 * -ffp-contract=off so that every product and sum rounds on its own, as
 * the reference's scalar code does.  Citations are to the reference tree.
 */
#define _DEFAULT_SOURCE
#include "armando_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* options.inc:21 assigns a default-real literal to pi, so the Fortran
   value is the REAL(4) rounding of pi widened to REAL(8). */
static const double PI_F = 3.1415927410125732;
/* kernel.f:41 writes 2./3. as a REAL(4) constant. */
static const double TWO_THIRDS_F = 0.6666666865348816;

/* ------------------------------------------------------------------ */
/* kernel                                                               */
/* ------------------------------------------------------------------ */

/* Faithful per-component form of kernel.f (skf = 0) and the Wendland C2 of
   armando2Dcpu.cu:116-172.  dwdx = grad_i W for dx = x_i - x_j. */
/* W and grad W of smoothing length h (kernel.f:31-56 takes h as mh). */
static double kern_h(const orc_params* P, double h, double r, double dx, double dy, double* dwx,
                     double* dwy) {
  const double q = r / h;
  if (P->kernel == ORC_K_WENDLAND2D) {
    /* W = a (1 - q/2)^4 (1 + 2q), dW/dr = (a/h)(5/8) q (q-2)^3, a = 7/(4 pi h^2).
       grad W = dW/dr * dx / r = (5a/(8h^2)) (q-2)^3 dx, regular at r = 0. */
    if (q < 2.0) {
      const double a = 7.0 / (4.0 * M_PI * h * h);
      const double s = 1.0 - 0.5 * q;
      const double s2 = s * s;
      const double t = q - 2.0;
      const double F = (0.625 * a / (h * h)) * (t * t * t);
      *dwx = F * dx;
      *dwy = F * dy;
      return a * (s2 * s2) * (1.0 + 2.0 * q);
    }
    *dwx = 0.0;
    *dwy = 0.0;
    return 0.0;
  }
  /* kernel.f:31-56, cubic spline */
  double factor;
  if (P->kernel == ORC_K_CUBIC2D)
    factor = 15.0 / (7.0 * PI_F * (h * h));
  else
    factor = 3.0 / (2.0 * PI_F * (h * h * h));
  if (q >= 0.0 && q <= 1.0) {
    const double g = factor * (-2.0 + 1.5 * q) / (h * h);
    *dwx = g * dx;
    *dwy = g * dy;
    return factor * (TWO_THIRDS_F - q * q + (q * q * q) / 2.0);
  } else if (q > 1.0 && q <= 2.0) {
    const double g = -factor * 1.0 / 6.0 * 3.0 * ((2.0 - q) * (2.0 - q)) / h;
    *dwx = g * (dx / r);
    *dwy = g * (dy / r);
    return factor * 1.0 / 6.0 * ((2.0 - q) * (2.0 - q) * (2.0 - q));
  }
  *dwx = 0.0;
  *dwy = 0.0;
  return 0.0;
}

static double kern(const orc_params* P, double r, double dx, double dy, double* dwx,
                   double* dwy) {
  return kern_h(P, P->h, r, dx, dy, dwx, dwy);
}

double orc_kernel(const orc_params* P, double r, double* F) {
  double dwx, dwy;
  /* probe with dx = r so that dwdx = F * r */
  double w = kern(P, r, r, 0.0, &dwx, &dwy);
  *F = (r > 0.0) ? dwx / r : 0.0;
  if (r == 0.0) {
    double wx, wy;
    kern(P, 0.0, 1.0, 0.0, &wx, &wy);
    *F = wx;
  }
  return w;
}

/* ------------------------------------------------------------------ */
/* neighbour search                                                     */
/* ------------------------------------------------------------------ */

int orc_link_list(const orc_params* P, int nt, const double* x, const double* y, int cap,
                  int* niac, int* pair_i, int* pair_j) {
  /* link_list.f:29-221 for dim = 2 and a uniform h (skf = 0: scale_k = 2) */
  const int scale_k = 2;
  double gmin[2] = {1.0e30, 1.0e30}, gmax[2] = {-1.0e30, -1.0e30};
  double mh = 0.0;
  for (int i = 0; i < nt; ++i) {
    if (x[i] < gmin[0]) gmin[0] = x[i];
    if (x[i] > gmax[0]) gmax[0] = x[i];
    if (y[i] < gmin[1]) gmin[1] = y[i];
    if (y[i] > gmax[1]) gmax[1] = y[i];
    if (P->h > mh) mh = P->h;
  }
  int gn[2];
  for (int d = 0; d < 2; ++d) {
    gmin[d] = gmin[d] - mh;
    gmax[d] = gmax[d] + mh;
    gn[d] = (int)((gmax[d] - gmin[d]) / (scale_k * mh)) + 1;
  }
  const long gnt = (long)gn[0] * gn[1];
  int* gindex = (int*)calloc((size_t)gnt + 1, sizeof(int));
  int* gpart = (int*)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int));
  int* gcell = (int*)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int));
  if (!gindex || !gpart || !gcell) {
    free(gindex); free(gpart); free(gcell);
    return -2;
  }
  for (int i = 0; i < nt; ++i) {
    int ix = (int)((x[i] - gmin[0]) / (scale_k * mh));
    int iy = (int)((y[i] - gmin[1]) / (scale_k * mh));
    gcell[i] = gn[0] * iy + ix;
    gindex[gcell[i]]++;
  }
  /* running start of each cell; particles are placed in increasing i */
  int counter = 0;
  for (long g = 0; g < gnt; ++g) {
    int c = gindex[g];
    gindex[g] = counter;
    counter += c;
  }
  gindex[gnt] = counter;
  {
    int* fill = (int*)malloc((size_t)gnt * sizeof(int) + sizeof(int));
    memcpy(fill, gindex, (size_t)gnt * sizeof(int));
    for (int i = 0; i < nt; ++i) gpart[fill[gcell[i]]++] = i;
    free(fill);
  }
  const double cut2 = (scale_k * mh) * (scale_k * mh);
  int n = 0;
  int rc = 0;
  for (int i = 0; i < nt - 1 && rc == 0; ++i) {
    int ijk[2], lo[2], hi[2];
    ijk[0] = (int)((x[i] - gmin[0]) / (scale_k * mh));
    ijk[1] = (int)((y[i] - gmin[1]) / (scale_k * mh));
    for (int d = 0; d < 2; ++d) {
      lo[d] = ijk[d] - 1;
      hi[d] = ijk[d] + 1;
      if (lo[d] < 0) lo[d] = 0;
      if (hi[d] >= gn[d]) hi[d] = gn[d] - 1;
    }
    for (int jg = lo[1]; jg <= hi[1] && rc == 0; ++jg) {
      for (int ig = lo[0]; ig <= hi[0] && rc == 0; ++ig) {
        const long g = (long)gn[0] * jg + ig;
        for (int k = gindex[g]; k < gindex[g + 1]; ++k) {
          const int j = gpart[k];
          if (j > i) {
            const double dx = x[i] - x[j];
            const double dy = y[i] - y[j];
            double dr2 = dx * dx;
            dr2 = dr2 + dy * dy;
            if (dr2 < cut2) {
              if (n < cap) {
                pair_i[n] = i;
                pair_j[n] = j;
                n++;
              } else {
                rc = -1;
                break;
              }
            }
          }
        }
      }
    }
  }
  *niac = n;
  free(gindex);
  free(gpart);
  free(gcell);
  return rc;
}

int orc_direct_find(const orc_params* P, int nt, const double* x, const double* y, int cap,
                    int* niac, int* pair_i, int* pair_j) {
  /* direct_find.f:36-60: all i < j with r^2 < (scale_k * mh)^2 */
  const double cut2 = (2.0 * P->h) * (2.0 * P->h);
  int n = 0;
  for (int i = 0; i < nt - 1; ++i)
    for (int j = i + 1; j < nt; ++j) {
      const double dx = x[i] - x[j];
      const double dy = y[i] - y[j];
      double dr2 = dx * dx;
      dr2 = dr2 + dy * dy;
      if (dr2 < cut2) {
        if (n >= cap) {
          *niac = n;
          return -1;
        }
        pair_i[n] = i;
        pair_j[n] = j;
        n++;
      }
    }
  *niac = n;
  return 0;
}

static int key_of(const orc_params* P, double x, double y) {
  /* armando2Dcpu.cu:2080-2101 (updateHashHost) in FP64 */
  const double tx = (x - P->ox) / P->cell;
  const double ty = (y - P->oy) / P->cell;
  if (!(tx >= 0.0) || !(ty >= 0.0)) return -1;
  const int ix = (int)tx;
  const int iy = (int)ty;
  if (ix >= P->nx || iy >= P->ny) return -1;
  return ix + iy * P->nx;
}

int orc_keys(const orc_params* P, int n, const double* x, const double* y, int32_t* key) {
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    key[i] = key_of(P, x[i], y[i]);
    if (key[i] < 0) bad++;
  }
  return bad;
}

void orc_sort_cells(int n, const int32_t* key, int ncell, int32_t* sorted_id,
                    int32_t* cell_start) {
  /* counting sort; input is listed in id order so the result is ordered
     by (cell, id), the order a stable radix sort of (key, id) produces */
  memset(cell_start, 0, sizeof(int32_t) * (size_t)(ncell + 1));
  for (int i = 0; i < n; ++i)
    if (key[i] >= 0) cell_start[key[i] + 1]++;
  for (int c = 0; c < ncell; ++c) cell_start[c + 1] += cell_start[c];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncell > 0 ? ncell : 1));
  memcpy(fill, cell_start, sizeof(int32_t) * (size_t)ncell);
  for (int i = 0; i < n; ++i)
    if (key[i] >= 0) sorted_id[fill[key[i]]++] = i;
  free(fill);
}

/* ------------------------------------------------------------------ */
/* faithful pair routines                                               */
/* ------------------------------------------------------------------ */

static inline void pair_geom(const orc_params* P, int i, int j, const double* x,
                             const double* y, double* dx, double* dy, double* dr2,
                             double* w, double* dwx, double* dwy) {
  *dx = x[i] - x[j];
  *dy = y[i] - y[j];
  double r2 = (*dx) * (*dx);
  r2 = r2 + (*dy) * (*dy);
  *dr2 = r2;
  *w = kern(P, sqrt(r2), *dx, *dy, dwx, dwy);
}

void orc_sum_density(const orc_params* P, int nrate, int niac, const int* pi, const int* pj,
                     const double* x, const double* y, double* rho) {
  /* density.f:1-62 */
  double dwx, dwy;
  const double selfdens = kern(P, 0.0, 0.0, 0.0, &dwx, &dwy);
  for (int i = 0; i < nrate; ++i) rho[i] = selfdens * P->mass;
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    double dx, dy, dr2, w;
    pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
    if (i < nrate) rho[i] = rho[i] + P->mass * w;
    if (j < nrate) rho[j] = rho[j] + P->mass * w;
  }
}

void orc_con_density(const orc_params* P, int nrate, int niac, const int* pi, const int* pj,
                     const double* x, const double* y, const double* vx, const double* vy,
                     double* drho) {
  /* density.f:166-223 */
  for (int i = 0; i < nrate; ++i) drho[i] = 0.0;
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    double dx, dy, dr2, w, dwx, dwy;
    pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
    double vcc = 0.0;
    vcc = vcc + (vx[i] - vx[j]) * dwx;
    vcc = vcc + (vy[i] - vy[j]) * dwy;
    if (i < nrate) drho[i] = drho[i] + P->mass * vcc;
    if (j < nrate) drho[j] = drho[j] + P->mass * vcc;
  }
}

void orc_int_force(const orc_params* P, int nt, int niac, const int* pi, const int* pj,
                   const double* x, const double* y, const double* vx, const double* vy,
                   const double* rho, const double* p, double* ax, double* ay, double* du) {
  /* internal_force.f:57-71 (zeroing), :73-142 (shear tensor, visc), :144-169
     (tdsdt), :178-259 (pa_sph 0), :309-430 (pa_sph 1), :433-435 (du) */
  const double m = P->mass;
  const int visc = (P->flags & ORC_VISC) != 0;
  const double eta = P->visc_eta;
  double *txx = NULL, *tyy = NULL, *txy = NULL;
  for (int i = 0; i < nt; ++i) {
    ax[i] = 0.0;
    ay[i] = 0.0;
    du[i] = 0.0;
  }
  if (visc) {
    txx = (double*)calloc((size_t)(nt > 0 ? nt : 1), sizeof(double));
    tyy = (double*)calloc((size_t)(nt > 0 ? nt : 1), sizeof(double));
    txy = (double*)calloc((size_t)(nt > 0 ? nt : 1), sizeof(double));
    for (int k = 0; k < niac; ++k) {
      const int i = pi[k], j = pj[k];
      double dx, dy, dr2, w, dwx, dwy;
      pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
      const double dvx = vx[j] - vx[i], dvy = vy[j] - vy[i];
      double hxx = 2.0 * dvx * dwx - dvy * dwy;
      double hyy = 2.0 * dvy * dwy - dvx * dwx;
      const double hxy = dvx * dwy + dvy * dwx;
      hxx = 2.0 / 3.0 * hxx;
      hyy = 2.0 / 3.0 * hyy;
      txx[i] = txx[i] + m * hxx / rho[j];
      tyy[i] = tyy[i] + m * hyy / rho[j];
      txy[i] = txy[i] + m * hxy / rho[j];
      txx[j] = txx[j] + m * hxx / rho[i];
      tyy[j] = tyy[j] + m * hyy / rho[i];
      txy[j] = txy[j] + m * hxy / rho[i];
    }
  }
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    double dx, dy, dr2, w, dwx, dwy;
    pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
    double he = 0.0, hv;
    if (P->pa_sph == 0) {
      hv = -(p[i] + p[j]) * dwx;
      he = he + (vx[j] - vx[i]) * hv;
      if (visc)
        hv = hv + (eta * txx[i] + eta * txx[j]) * dwx + (eta * txy[i] + eta * txy[j]) * dwy;
      hv = hv / (rho[i] * rho[j]);
      ax[i] = ax[i] + m * hv;
      ax[j] = ax[j] - m * hv;
      hv = -(p[i] + p[j]) * dwy;
      he = he + (vy[j] - vy[i]) * hv;
      if (visc)
        hv = hv + (eta * txy[i] + eta * txy[j]) * dwx + (eta * tyy[i] + eta * tyy[j]) * dwy;
      hv = hv / (rho[i] * rho[j]);
      ay[i] = ay[i] + m * hv;
      ay[j] = ay[j] - m * hv;
      he = he / (rho[i] * rho[j]);
    } else {
      const double ri2 = rho[i] * rho[i], rj2 = rho[j] * rho[j];
      const double a = p[i] / ri2 + p[j] / rj2;
      hv = -a * dwx;
      he = he + (vx[j] - vx[i]) * hv;
      if (visc)
        hv = hv + (eta * txx[i] / ri2 + eta * txx[j] / rj2) * dwx +
             (eta * txy[i] / ri2 + eta * txy[j] / rj2) * dwy;
      ax[i] = ax[i] + m * hv;
      ax[j] = ax[j] - m * hv;
      hv = -a * dwy;
      he = he + (vy[j] - vy[i]) * hv;
      if (visc)
        hv = hv + (eta * txy[i] / ri2 + eta * txy[j] / rj2) * dwx +
             (eta * tyy[i] / ri2 + eta * tyy[j] / rj2) * dwy;
      ay[i] = ay[i] + m * hv;
      ay[j] = ay[j] - m * hv;
    }
    du[i] = du[i] + m * he;
    du[j] = du[j] + m * he;
  }
  for (int i = 0; i < nt; ++i) {
    double tds = 0.0;
    if (visc) {
      tds = txx[i] * txx[i] + tyy[i] * tyy[i] + 2.0 * txy[i] * txy[i];
      tds = 0.5 * eta / rho[i] * tds;
    }
    du[i] = tds + 0.5 * du[i];
  }
  free(txx);
  free(tyy);
  free(txy);
}

void orc_art_visc(const orc_params* P, int niac, const int* pi, const int* pj,
                  const double* x, const double* y, const double* vx, const double* vy,
                  const double* rho, const double* c, double* ax, double* ay, double* du) {
  /* art_visc.f:55-101 with alpha, beta, eta from the parameters */
  const double m = P->mass;
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    double dx, dy, dr2, w, dwx, dwy;
    pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
    const double mh = P->h;
    double vr = 0.0, vdwdx = 0.0, dv;
    dv = vx[i] - vx[j];
    vr = vr + dv * dx;
    vdwdx = vdwdx + dv * dwx;
    dv = vy[i] - vy[j];
    vr = vr + dv * dy;
    vdwdx = vdwdx + dv * dwy;
    if (vr < 0.0) {
      const double phi = (mh * vr) / (dr2 + (mh * P->av_eta) * (mh * P->av_eta));
      const double mc = 0.5 * (c[i] + c[j]);
      const double mrho = 0.5 * (rho[i] + rho[j]);
      const double piv = (P->av_beta * phi - P->av_alpha * mc) * phi / mrho;
      ax[i] = ax[i] - m * piv * dwx;
      ax[j] = ax[j] + m * piv * dwx;
      ay[i] = ay[i] - m * piv * dwy;
      ay[j] = ay[j] + m * piv * dwy;
      du[i] = du[i] + 0.5 * m * piv * vdwdx;
      du[j] = du[j] + 0.5 * m * piv * vdwdx;
    }
  }
}

void orc_art_heat(const orc_params* P, int nt, int niac, const int* pi, const int* pj,
                  const double* x, const double* y, const double* vx, const double* vy,
                  const double* rho, const double* u, const double* c, double* du,
                  double* divv_out) {
  /* art_heat.f:42-97 with g1, g2, eta from the parameters */
  const double m = P->mass;
  double* divv = (double*)calloc((size_t)(nt > 0 ? nt : 1), sizeof(double));
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    double dx, dy, dr2, w, dwx, dwy;
    pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
    double vdwdx = 0.0;
    vdwdx = vdwdx + (vx[j] - vx[i]) * dwx;
    vdwdx = vdwdx + (vy[j] - vy[i]) * dwy;
    divv[i] = divv[i] + m / rho[i] * vdwdx;
    divv[j] = divv[j] + m / rho[j] * vdwdx;
  }
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    double dx, dy, dr2, w, dwx, dwy;
    pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
    const double mh = P->h, h = P->h;
    const double mrho = 0.5 * (rho[i] + rho[j]);
    double rdwdx = 0.0;
    rdwdx = rdwdx + dx * dwx;
    rdwdx = rdwdx + dy * dwy;
    const double psi_i = P->ah_g1 * h * c[i] + P->ah_g2 * h * h * (fabs(divv[i]) - divv[i]);
    const double psi_j = P->ah_g1 * h * c[j] + P->ah_g2 * h * h * (fabs(divv[j]) - divv[j]);
    const double psi_ij = 0.5 * (psi_i + psi_j);
    const double psi =
        (2.0 * psi_ij) / (mrho * (dr2 + (mh * P->ah_eta) * (mh * P->ah_eta))) * rdwdx;
    du[i] = du[i] + m * psi * (u[i] - u[j]);
    du[j] = du[j] + m * psi * (u[j] - u[i]);
  }
  if (divv_out) memcpy(divv_out, divv, sizeof(double) * (size_t)nt);
  free(divv);
}

void orc_av_vel(const orc_params* P, int niac, const int* pi, const int* pj,
                const double* x, const double* y, const double* vx, const double* vy,
                const double* rho, double* avx, double* avy) {
  /* av_vel.f:37-56 with eps from the parameters */
  const double m = P->mass, eps = P->xsph_eps;
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    double dx, dy, dr2, w, dwx, dwy;
    pair_geom(P, i, j, x, y, &dx, &dy, &dr2, &w, &dwx, &dwy);
    const double mrho = (rho[i] + rho[j]) / 2.0;
    double dv = vx[i] - vx[j];
    avx[i] = avx[i] - eps * m / mrho * dv * w;
    avx[j] = avx[j] + eps * m / mrho * dv * w;
    dv = vy[i] - vy[j];
    avy[i] = avy[i] - eps * m / mrho * dv * w;
    avy[j] = avy[j] + eps * m / mrho * dv * w;
  }
}

void orc_contact_displacement(const orc_params* P, int np, int nv, const double* x,
                              const double* y, const double* vx, const double* vy, int niac,
                              const int* pi, const int* pj, double* dx, double* dy) {
  /* external_force.f90:218-242 */
  const double hh = P->h / P->contact_hdr;
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    if (j >= np && j < np + nv && i < np) {
      double pni = 0.0, pnm = 0.0;
      pni = pni + (x[i] - x[j]) * vx[j];
      pnm = pnm + dx[i] * vx[j];
      pni = pni + (y[i] - y[j]) * vy[j];
      pnm = pnm + dy[i] * vy[j];
      pni = hh - pni;
      if (pni > 0.0 && pni > pnm) {
        dx[i] = dx[i] + (pni - pnm) * vx[j];
        dy[i] = dy[i] + (pni - pnm) * vy[j];
      }
    }
  }
}

void orc_smooth_vector(const orc_params* P, int np, const double* x, const double* y,
                       const double* rho, int niac, const int* pi, const int* pj, double* vx,
                       double* vy) {
  /* external_force.f90:276-337 */
  const double m = P->mass;
  double* ox = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  double* oy = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  double wx, wy;
  const double w0 = kern(P, 0.0, 0.0, 0.0, &wx, &wy);
  for (int i = 0; i < np; ++i) {
    ox[i] = w0 * m / rho[i] * vx[i];
    oy[i] = w0 * m / rho[i] * vy[i];
  }
  for (int k = 0; k < niac; ++k) {
    const int i = pi[k], j = pj[k];
    if (i < np && j < np) {
      const double ddx = x[i] - x[j], ddy = y[i] - y[j];
      const double r = sqrt(ddx * ddx + ddy * ddy);
      const double w = kern(P, r, ddx, ddy, &wx, &wy);
      ox[j] = ox[j] + w * m / rho[j] * vx[i];
      oy[j] = oy[j] + w * m / rho[j] * vy[i];
      ox[i] = ox[i] + w * m / rho[i] * vx[j];
      oy[i] = oy[i] + w * m / rho[i] * vy[j];
    }
  }
  memcpy(vx, ox, sizeof(double) * (size_t)np);
  memcpy(vy, oy, sizeof(double) * (size_t)np);
  free(ox);
  free(oy);
}

static inline double tait_p(const orc_params* P, double rho) {
  /* eos.f90:274: p = rho0 c0^2 / 7 ((rho/rho0)^7 - 1) */
  const double t = rho / P->rho0;
  const double t2 = t * t;
  const double t4 = t2 * t2;
  return P->rho0 * P->c0 * P->c0 / 7.0 * (t4 * t2 * t - 1.0);
}

/* Gaussian heat source rate per unit mass at a particle. */
static inline double heat_source(const orc_params* P, double x, double y, double rho) {
  const double dx = x - P->hs_x, dy = y - P->hs_y;
  const double k = 1.0 / (2.0 * P->hs_sigma * P->hs_sigma);
  return P->hs_q0 * exp(-k * (dx * dx + dy * dy)) / rho;
}

/* One axis of fluka_heat.f:256-286: the two bins around the position and
   the local coordinate t in [-1, 1) between their centres (bin ir: lower or
   upper half), clamped to the map. */
static inline void dep_axis(double s, int n, int* lo, int* hi, double* t) {
  const int ir = (int)s; /* Fortran int(): truncation */
  double xr = s - ir;
  if (xr >= 0.5) {
    *lo = ir;
    *hi = ir + 1;
    xr = 2.0 * xr - 2.0;
  } else {
    *lo = ir - 1;
    *hi = ir;
    xr = 2.0 * xr;
  }
  if (*lo < 0) *lo = 0;
  if (*hi >= n) *hi = n - 1;
  *t = xr;
}

double orc_deposition_at(const orc_params* P, double x, double y, double z, int* inside) {
  /* fluka_heat.f:237-337; with z = 0 the 2D branch :488-573 ((0 - oz)/dz is
     its -oz/dz, and the z bounds hold for a map that spans z = 0) */
  const int* n = P->dep_n;
  const double* o = P->dep_o;
  const double* d = P->dep_d;
  if (P->dep_type == 1) {
    /* R-phi-z binning (fluka_heat.f:341-347 in 3D, :593-595 in 2D): the
       map's axes are r, phi in degrees and z.  The reference's z coordinate
       divides by the phi bin width (:359, :608) and its 3D branch picks the
       bin pairs again after clamping them (:398-420); both are taken here
       as the Cartesian branch does them: z by the z bin width, the clamped
       pairs kept. */
    const double r = sqrt(x * x + y * y);
    const double phi = 180.0 / PI_F * atan2(y, x); /* options.inc's pi, a REAL(4) literal */
    x = r;
    y = phi;
  }
  *inside = x >= o[0] && y >= o[1] && z >= o[2] && x <= o[0] + n[0] * d[0] &&
            y <= o[1] + n[1] * d[1] && z <= o[2] + n[2] * d[2];
  if (!*inside) return 0.0;
  int i0, i1, j0, j1, k0, k1;
  double a, b, c;
  dep_axis((x - o[0]) / d[0], n[0], &i0, &i1, &a);
  dep_axis((y - o[1]) / d[1], n[1], &j0, &j1, &b);
  dep_axis((z - o[2]) / d[2], n[2], &k0, &k1, &c);
  const size_t sx = 1, sy = (size_t)n[0], sz = (size_t)n[0] * (size_t)n[1];
#define DEP(i, j, k) P->dep[(size_t)(i) * sx + (size_t)(j) * sy + (size_t)(k) * sz]
  const double f1 = 0.125 * (1.0 - a) * (1.0 - b) * (1.0 - c);
  const double f2 = 0.125 * (1.0 + a) * (1.0 - b) * (1.0 - c);
  const double f3 = 0.125 * (1.0 + a) * (1.0 + b) * (1.0 - c);
  const double f4 = 0.125 * (1.0 - a) * (1.0 + b) * (1.0 - c);
  const double f5 = 0.125 * (1.0 - a) * (1.0 - b) * (1.0 + c);
  const double f6 = 0.125 * (1.0 + a) * (1.0 - b) * (1.0 + c);
  const double f7 = 0.125 * (1.0 + a) * (1.0 + b) * (1.0 + c);
  const double f8 = 0.125 * (1.0 - a) * (1.0 + b) * (1.0 + c);
  const double v = f1 * DEP(i0, j0, k0) + f2 * DEP(i1, j0, k0) + f3 * DEP(i1, j1, k0) +
                   f4 * DEP(i0, j1, k0) + f5 * DEP(i0, j0, k1) + f6 * DEP(i1, j0, k1) +
                   f7 * DEP(i1, j1, k1) + f8 * DEP(i0, j1, k1);
#undef DEP
  return v;
}

void orc_fluka_heat(const orc_params* P, int np, const double* x, const double* y,
                    const double* rho, double* dedt) {
  /* fluka_heat.f:480-483: dedt reset, then the binned value / rho inside */
  for (int i = 0; i < np; ++i) {
    int in;
    const double v = orc_deposition_at(P, x[i], y[i], 0.0, &in);
    dedt[i] = in ? v / rho[i] : 0.0;
  }
}

void orc3_fluka_heat(const orc_params* P, int np, const double* x, const double* y,
                     const double* z, const double* rho, double* dedt) {
  /* fluka_heat.f:233-337, with exdudt zeroed at the start of the step
     (single_step.f:92-96) */
  for (int i = 0; i < np; ++i) {
    int in;
    const double v = orc_deposition_at(P, x[i], y[i], z[i], &in);
    dedt[i] = in ? v / rho[i] : 0.0;
  }
}

static inline int dep_active(const orc_params* P) { return P->dep != NULL && P->dep_on; }

static inline double deposition(const orc_params* P, double x, double y, double rho) {
  int in;
  const double v = orc_deposition_at(P, x, y, 0.0, &in);
  return in ? v / rho : 0.0;
}

void orc_eos_tait(const orc_params* P, int n, double* rho, double* p, double* c) {
  /* eos.f90:255-283 */
  for (int i = 0; i < n; ++i) {
    p[i] = tait_p(P, rho[i]);
    c[i] = P->c0;
    if (p[i] < P->pmin) {
      p[i] = 0.0;
      rho[i] = P->rho0;
    }
  }
}

void orc_eos(const orc_params* P, int n, double* rho, double* p, double* c) {
  if (P->eos != ORC_EOS_LINEAR) {
    orc_eos_tait(P, n, rho, p, c);
    return;
  }
  /* eos.f eos_poly with only a1 = rho0 c0^2 and an offset:
     p = c0^2 rho + p0; the same clamp as the Tait EOS */
  const double c2 = P->c0 * P->c0;
  for (int i = 0; i < n; ++i) {
    p[i] = c2 * rho[i] + P->eos_p0;
    c[i] = P->c0;
    if (p[i] < P->pmin) {
      p[i] = 0.0;
      rho[i] = P->rho0;
    }
  }
}

/* ------------------------------------------------------------------ */
/* state, single step, integration                                      */
/* ------------------------------------------------------------------ */

static orc_state* state_alloc(int cap, int n, int np) {
  orc_state* s = (orc_state*)calloc(1, sizeof(orc_state));
  s->n = n;
  s->np = np;
  s->nb = n;
  s->cap = cap;
  double** f[] = {&s->x,   &s->y,   &s->vx,  &s->vy,   &s->rho,     &s->u,
                  &s->p,   &s->c,   &s->drho, &s->du,  &s->ax,      &s->ay,
                  &s->avx, &s->avy, &s->divv, &s->rho_min, &s->u_min, &s->vx_min,
                  &s->vy_min, &s->h};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k)
    *f[k] = (double*)calloc((size_t)(cap > 0 ? cap : 1), sizeof(double));
  return s;
}

orc_state* orc_state_new(int n, int np) { return state_alloc(n, n, np); }

orc_state* orc_state_new_ghosts(int np, int nv) {
  orc_state* s = state_alloc(np + nv + np, np + nv, np);
  s->xp = (double*)calloc((size_t)(np > 0 ? np : 1), sizeof(double));
  s->yp = (double*)calloc((size_t)(np > 0 ? np : 1), sizeof(double));
  s->gpar = (int*)calloc((size_t)(np > 0 ? np : 1), sizeof(int));
  return s;
}

void orc_state_free(orc_state* s) {
  if (!s) return;
  double* f[] = {s->x,   s->y,   s->vx,  s->vy,   s->rho,     s->u,      s->p,
                 s->c,   s->drho, s->du, s->ax,   s->ay,      s->avx,    s->avy,
                 s->divv, s->rho_min, s->u_min, s->vx_min, s->vy_min, s->h};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); ++k) free(f[k]);
  free(s->xp);
  free(s->yp);
  free(s->gpar);
  free(s);
}

int orc_ghost_particles(const orc_params* P, orc_state* s) {
  /* ghost.f:1-86 (the loops over the pair list become a loop over the
     virtual particles with the pair test of the last search) */
  const int np = s->np, nb = s->nb;
  const double* xp = s->have_prev ? s->xp : s->x;
  const double* yp = s->have_prev ? s->yp : s->y;
  const double cut = (2.0 * P->h) * (2.0 * P->h);
  double *x = s->x, *y = s->y, *vx = s->vx, *vy = s->vy;
  int ng = 0;
  for (int i = 0; i < np; ++i) {
    int best = -1;
    double bd = 0.0;
    for (int j = np; j < nb; ++j) {
      const double dx = xp[i] - x[j], dy = yp[i] - y[j];
      double r2 = dx * dx;
      r2 = r2 + dy * dy;
      if (!(r2 < cut)) continue;
      /* ghost.f:42-49: squared distance at the current positions */
      const double ex = x[j] - x[i], ey = y[j] - y[i];
      double dd = ex * ex;
      dd = dd + ey * ey;
      if (best < 0 || dd < bd) {
        best = j;
        bd = dd;
      }
    }
    if (best < 0) continue;
    const int g = nb + ng, j = best;
    if (g >= s->cap) return -1;
    /* ghost.f:60-79 */
    double d1 = 0.0, d2 = 0.0;
    d1 = d1 + (x[i] - x[j]) * vx[j];
    d1 = d1 + (y[i] - y[j]) * vy[j];
    d2 = d2 + vx[i] * vx[j];
    d2 = d2 + vy[i] * vy[j];
    x[g] = x[i] - 2.0 * vx[j] * d1;
    y[g] = y[i] - 2.0 * vy[j] * d1;
    vx[g] = vx[i] - 2.0 * vx[j] * d2;
    vy[g] = vy[i] - 2.0 * vy[j] * d2;
    s->rho[g] = s->rho[i];
    s->p[g] = s->p[i];
    s->u[g] = s->u[i];
    s->c[g] = s->c[i];
    s->h[g] = s->h[i];
    s->gpar[ng] = i;
    ++ng;
  }
  s->n = nb + ng;
  for (int i = 0; i < np; ++i) {
    s->xp[i] = x[i];
    s->yp[i] = y[i];
  }
  s->have_prev = 1;
  return ng;
}

/* single_step.f:188-196: after the EOS the ghosts take rho, p and u of
   their parents. */
static void ghost_copy(orc_state* s) {
  if (!s->gpar) return;
  for (int k = s->nb; k < s->n; ++k) {
    const int i = s->gpar[k - s->nb];
    s->rho[k] = s->rho[i];
    s->p[k] = s->p[i];
    s->u[k] = s->u[i];
  }
}

static int nrate_of(const orc_params* P, const orc_state* s) {
  return (P->flags & ORC_DBC) ? s->nb : s->np;
}

int orc_single_step(const orc_params* P, orc_state* s, int niac, const int* pi,
                    const int* pj) {
  /* single_step.f:92-234 (a pair list carries no periodic shift, and the
     reference has no conduction: both live in the gather twin only) */
  if ((P->flags & (ORC_CONDUCTION | ORC_PERIODIC_X)) || P->sle) return -2;
  const int n = s->n, np = s->np, nr = nrate_of(P, s);
  double* avdv_x = (double*)calloc((size_t)n, sizeof(double));
  double* avdv_y = (double*)calloc((size_t)n, sizeof(double));
  double* avdu = (double*)calloc((size_t)n, sizeof(double));
  double* ahdu = (double*)calloc((size_t)n, sizeof(double));
  for (int i = 0; i < n; ++i) {
    s->avx[i] = 0.0;
    s->avy[i] = 0.0;
    s->divv[i] = 0.0;
  }
  if (P->density_method == 0) {
    for (int i = 0; i < n; ++i) s->drho[i] = 0.0;
    orc_con_density(P, nr, niac, pi, pj, s->x, s->y, s->vx, s->vy, s->drho);
  } else {
    orc_sum_density(P, nr, niac, pi, pj, s->x, s->y, s->rho);
  }
  orc_eos(P, nr, s->rho, s->p, s->c);
  ghost_copy(s);
  orc_int_force(P, n, niac, pi, pj, s->x, s->y, s->vx, s->vy, s->rho, s->p, s->ax, s->ay,
                s->du);
  if (P->flags & ORC_ART_VISC)
    orc_art_visc(P, niac, pi, pj, s->x, s->y, s->vx, s->vy, s->rho, s->c, avdv_x, avdv_y,
                 avdu);
  if (P->flags & ORC_ART_HEAT)
    orc_art_heat(P, n, niac, pi, pj, s->x, s->y, s->vx, s->vy, s->rho, s->u, s->c, ahdu,
                 s->divv);
  if (P->flags & ORC_AV_VEL)
    orc_av_vel(P, niac, pi, pj, s->x, s->y, s->vx, s->vy, s->rho, s->avx, s->avy);
  const int grav = (P->flags & ORC_GRAVITY) != 0;
  for (int i = 0; i < n; ++i) {
    const double ex = (grav && i < np) ? P->gx : 0.0;
    const double ey = (grav && i < np) ? P->gy : 0.0;
    s->ax[i] = s->ax[i] + ex + avdv_x[i];
    s->ay[i] = s->ay[i] + ey + avdv_y[i];
    s->du[i] = s->du[i] + avdu[i] + ahdu[i];
    if ((P->flags & ORC_HEAT_SOURCE) && i < np)
      s->du[i] = s->du[i] + heat_source(P, s->x[i], s->y[i], s->rho[i]);
    if (dep_active(P) && i < np) s->du[i] = s->du[i] + deposition(P, s->x[i], s->y[i], s->rho[i]);
  }
  if (P->flags & ORC_CONTACT) {
    /* time_integration.f90:128-142: d from the virtual neighbours, smoothed,
       a += 0.5 d / dt^2 */
    double* cdx = (double*)calloc((size_t)(np > 0 ? np : 1), sizeof(double));
    double* cdy = (double*)calloc((size_t)(np > 0 ? np : 1), sizeof(double));
    orc_contact_displacement(P, np, s->nb - np, s->x, s->y, s->vx, s->vy, niac, pi, pj, cdx,
                             cdy);
    orc_smooth_vector(P, np, s->x, s->y, s->rho, niac, pi, pj, cdx, cdy);
    for (int i = 0; i < np; ++i) {
      s->ax[i] = s->ax[i] + 0.5 * cdx[i] / (P->dt * P->dt);
      s->ay[i] = s->ay[i] + 0.5 * cdy[i] / (P->dt * P->dt);
    }
    free(cdx);
    free(cdy);
  }
  free(avdv_x);
  free(avdv_y);
  free(avdu);
  free(ahdu);
  return 0;
}

/* Cells of the fixed grid, built once per gather call. */
typedef struct {
  int32_t* key;
  int32_t* sorted;
  int32_t* start;
  int ncell;
} cells_t;

static int build_cells(const orc_params* P, const orc_state* s, cells_t* C) {
  C->ncell = P->nx * P->ny;
  C->key = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->n);
  C->sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->n);
  C->start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)C->ncell + 1));
  int bad = orc_keys(P, s->n, s->x, s->y, C->key);
  if (bad) return bad;
  orc_sort_cells(s->n, C->key, C->ncell, C->sorted, C->start);
  return 0;
}

static void free_cells(cells_t* C) {
  free(C->key);
  free(C->sorted);
  free(C->start);
}

/* The cells cx-1, cx, cx+1 of cell row ry, in that order, as ranges of the
   sorted order with their x shift in periods: one range, or two when the
   grid is periodic in x and cx is the first or last column (the wrapped
   cell nx-1 comes first for cx = 0, the wrapped cell 0 last for
   cx = nx-1). */
typedef struct {
  int t0, t1, sh;
} span_t;

static int row_spans(const orc_params* P, const int32_t* start, int cx, int ry, span_t sp[2]) {
  const int nx = P->nx, base = ry * nx;
  if ((P->flags & ORC_PERIODIC_X) && cx == 0) {
    sp[0].t0 = start[base + nx - 1];
    sp[0].t1 = start[base + nx];
    sp[0].sh = 1;
    sp[1].t0 = start[base];
    sp[1].t1 = start[base + 2];
    sp[1].sh = 0;
    return 2;
  }
  if ((P->flags & ORC_PERIODIC_X) && cx == nx - 1) {
    sp[0].t0 = start[base + nx - 2];
    sp[0].t1 = start[base + nx];
    sp[0].sh = 0;
    sp[1].t0 = start[base];
    sp[1].t1 = start[base + 1];
    sp[1].sh = -1;
    return 2;
  }
  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < nx ? cx + 1 : nx - 1;
  sp[0].t0 = start[base + x0];
  sp[0].t1 = start[base + x1 + 1];
  sp[0].sh = 0;
  return 1;
}

/* x_i - x_j across sh periods: (x_i - x_j) + sh (nx cell). */
static double dx_of(const orc_params* P, double xi, double xj, int sh) {
  const double d = xi - xj;
  return sh ? d + (double)sh * ((double)P->nx * P->cell) : d;
}

/* Smoothing length of pair (i, j): the uniform h, or with P->sle the mean
   mh = (h_i + h_j)/2 of every reference pair routine (density.f:54,
   internal_force.f:187, art_visc.f:64, art_heat.f:78, av_vel.f:46). */
static inline double pair_h(const orc_params* P, const orc_state* s, int i, int j) {
  return P->sle ? 0.5 * (s->h[i] + s->h[j]) : P->h;
}

static inline double own_h(const orc_params* P, const orc_state* s, int i) {
  return P->sle ? s->h[i] : P->h;
}

int orc_single_step_gather(const orc_params* P, orc_state* s) {
  const int n = s->n, np = s->np, nr = nrate_of(P, s);
  const double m = P->mass;
  cells_t C;
  int bad = build_cells(P, s, &C);
  if (bad) {
    free_cells(&C);
    return bad;
  }
  const int nx = P->nx, ny = P->ny;
  const double* x = s->x;
  const double* y = s->y;
  const double* vx = s->vx;
  const double* vy = s->vy;
  /* pass A: continuity (or summation) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const int key = C.key[i];
    const int cx = key % nx, cy = key / nx;
    double w0dx, w0dy;
    const double w0 = kern_h(P, own_h(P, s, i), 0.0, 0.0, 0.0, &w0dx, &w0dy);
    double dr = 0.0, rs = m * w0;
    for (int ry = cy - 1; ry <= cy + 1; ++ry) {
      if (ry < 0 || ry >= ny) continue;
      span_t sp[2];
      const int ns = row_spans(P, C.start, cx, ry, sp);
      for (int q = 0; q < ns; ++q)
      for (int t = sp[q].t0; t < sp[q].t1; ++t) {
        const int j = C.sorted[t];
        if (j == i) continue;
        const double dx = dx_of(P, x[i], x[j], sp[q].sh), dy = y[i] - y[j];
        double r2 = dx * dx;
        r2 = r2 + dy * dy;
        const double hij = pair_h(P, s, i, j);
        if (!(r2 < (2.0 * hij) * (2.0 * hij))) continue;
        double dwx, dwy;
        const double w = kern_h(P, hij, sqrt(r2), dx, dy, &dwx, &dwy);
        double vcc = (vx[i] - vx[j]) * dwx;
        vcc = vcc + (vy[i] - vy[j]) * dwy;
        dr = dr + m * vcc;
        rs = rs + m * w;
      }
    }
    if (i < nr) {
      if (P->density_method == 0)
        s->drho[i] = dr;
      else
        s->rho[i] = rs;
    } else {
      s->drho[i] = 0.0;
    }
    s->divv[i] = -dr; /* divided by rho after the EOS */
  }
  orc_eos(P, nr, s->rho, s->p, s->c);
  ghost_copy(s);
  double* psi = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const double dv = s->divv[i] / s->rho[i];
    s->divv[i] = dv;
    const double hi = own_h(P, s, i);
    psi[i] = P->ah_g1 * hi * s->c[i] + P->ah_g2 * hi * hi * (fabs(dv) - dv);
  }
  /* shear tensor (internal_force.f:73-142) with the post-EOS densities */
  const int visc = (P->flags & ORC_VISC) != 0;
  const double eta = P->visc_eta;
  double* tt = visc ? (double*)calloc((size_t)n * 3, sizeof(double)) : NULL;
  if (visc) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      const int key = C.key[i];
      const int cx = key % nx, cy = key / nx;
      double txx = 0.0, tyy = 0.0, txy = 0.0;
      for (int ry = cy - 1; ry <= cy + 1; ++ry) {
        if (ry < 0 || ry >= ny) continue;
        span_t sp[2];
        const int ns = row_spans(P, C.start, cx, ry, sp);
        for (int q = 0; q < ns; ++q)
        for (int t = sp[q].t0; t < sp[q].t1; ++t) {
          const int j = C.sorted[t];
          if (j == i) continue;
          const double dx = dx_of(P, x[i], x[j], sp[q].sh), dy = y[i] - y[j];
          double r2 = dx * dx;
          r2 = r2 + dy * dy;
          const double hij = pair_h(P, s, i, j);
          if (!(r2 < (2.0 * hij) * (2.0 * hij))) continue;
          double dwx, dwy;
          kern_h(P, hij, sqrt(r2), dx, dy, &dwx, &dwy);
          const double dvx = vx[j] - vx[i], dvy = vy[j] - vy[i];
          const double hxx = 2.0 / 3.0 * (2.0 * dvx * dwx - dvy * dwy);
          const double hyy = 2.0 / 3.0 * (2.0 * dvy * dwy - dvx * dwx);
          const double hxy = dvx * dwy + dvy * dwx;
          txx = txx + m * hxx / s->rho[j];
          tyy = tyy + m * hyy / s->rho[j];
          txy = txy + m * hxy / s->rho[j];
        }
      }
      tt[3 * i] = txx;
      tt[3 * i + 1] = tyy;
      tt[3 * i + 2] = txy;
    }
  }
  const int grav = (P->flags & ORC_GRAVITY) != 0;
  const int f_av = (P->flags & ORC_ART_VISC) != 0;
  const int f_ah = (P->flags & ORC_ART_HEAT) != 0;
  const int f_xs = (P->flags & ORC_AV_VEL) != 0;
  const int f_cd = (P->flags & ORC_CONDUCTION) != 0;
  const double* rho = s->rho;
  const double* p = s->p;
  const double* c = s->c;
  const double* u = s->u;
  /* h update (hsml.f): dh/dt = (h/dim) sum m/rho_j (v_j - v_i).grad W, or
     h = 2 (m/rho)^(1/dim); dim 2 for this step (3 reproduces the reference
     as shipped, options.inc dim = 3) */
  const double hdim = P->dim == 3 ? 3.0 : 2.0;
  double* hnew = P->sle ? (double*)malloc(sizeof(double) * (size_t)n) : NULL;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const int key = C.key[i];
    const int cx = key % nx, cy = key / nx;
    double fx = 0.0, fy = 0.0, due = 0.0, avfx = 0.0, avfy = 0.0, avdu = 0.0, ahdu = 0.0;
    double cdu = 0.0, xsx = 0.0, xsy = 0.0, dvh = 0.0;
    const double pr_i = p[i] / (rho[i] * rho[i]);
    for (int ry = cy - 1; ry <= cy + 1; ++ry) {
      if (ry < 0 || ry >= ny) continue;
      span_t sp[2];
      const int ns = row_spans(P, C.start, cx, ry, sp);
      for (int q = 0; q < ns; ++q)
      for (int t = sp[q].t0; t < sp[q].t1; ++t) {
        const int j = C.sorted[t];
        if (j == i) continue;
        const double dx = dx_of(P, x[i], x[j], sp[q].sh), dy = y[i] - y[j];
        double r2 = dx * dx;
        r2 = r2 + dy * dy;
        const double h = pair_h(P, s, i, j);
        if (!(r2 < (2.0 * h) * (2.0 * h))) continue;
        const double av_e2 = (h * P->av_eta) * (h * P->av_eta);
        const double ah_e2 = (h * P->ah_eta) * (h * P->ah_eta);
        double dwx, dwy;
        const double w = kern_h(P, h, sqrt(r2), dx, dy, &dwx, &dwy);
        const double dvx = vx[i] - vx[j], dvy = vy[i] - vy[j];
        if (P->sle == 2) {
          double vdh = (vx[j] - vx[i]) * dwx;
          vdh = vdh + (vy[j] - vy[i]) * dwy;
          dvh = dvh + m / rho[j] * vdh;
        }
        double vdw = dvx * dwx;
        vdw = vdw + dvy * dwy;
        double a;
        if (P->pa_sph == 0)
          a = (p[i] + p[j]) / (rho[i] * rho[j]);
        else
          a = pr_i + p[j] / (rho[j] * rho[j]);
        fx = fx - m * a * dwx;
        fy = fy - m * a * dwy;
        due = due + m * a * vdw;
        if (visc) {
          const double* ti = tt + 3 * i;
          const double* tj = tt + 3 * j;
          double sxx, syy, sxy;
          if (P->pa_sph == 0) {
            const double rr = rho[i] * rho[j];
            sxx = eta * (ti[0] + tj[0]) / rr;
            syy = eta * (ti[1] + tj[1]) / rr;
            sxy = eta * (ti[2] + tj[2]) / rr;
          } else {
            const double ri2 = rho[i] * rho[i], rj2 = rho[j] * rho[j];
            sxx = eta * ti[0] / ri2 + eta * tj[0] / rj2;
            syy = eta * ti[1] / ri2 + eta * tj[1] / rj2;
            sxy = eta * ti[2] / ri2 + eta * tj[2] / rj2;
          }
          fx = fx + m * (sxx * dwx + sxy * dwy);
          fy = fy + m * (sxy * dwx + syy * dwy);
        }
        const double mrho = 0.5 * (rho[i] + rho[j]);
        if (f_av) {
          double vr = dvx * dx;
          vr = vr + dvy * dy;
          if (vr < 0.0) {
            const double phi = (h * vr) / (r2 + av_e2);
            const double mc = 0.5 * (c[i] + c[j]);
            const double piv = (P->av_beta * phi - P->av_alpha * mc) * phi / mrho;
            avfx = avfx - m * piv * dwx;
            avfy = avfy - m * piv * dwy;
            avdu = avdu + 0.5 * m * piv * vdw;
          }
        }
        if (f_ah) {
          double rdw = dx * dwx;
          rdw = rdw + dy * dwy;
          const double pij = 0.5 * (psi[i] + psi[j]);
          const double ps = (2.0 * pij) / (mrho * (r2 + ah_e2)) * rdw;
          ahdu = ahdu + m * ps * (u[i] - u[j]);
        }
        if (f_cd) {
          /* Cleary & Monaghan (1999), uniform k: 4 k k/(k + k) = 2 k */
          double rdw = dx * dwx;
          rdw = rdw + dy * dwy;
          const double kk = 4.0 * P->cond_k * P->cond_k / (P->cond_k + P->cond_k);
          const double dT = u[i] / P->cp - u[j] / P->cp;
          cdu = cdu + m * kk * dT / (rho[i] * rho[j]) * rdw / (r2 + ah_e2);
        }
        if (f_xs) {
          xsx = xsx - P->xsph_eps * m / mrho * dvx * w;
          xsy = xsy - P->xsph_eps * m / mrho * dvy * w;
        }
      }
    }
    const double ex = (grav && i < np) ? P->gx : 0.0;
    const double ey = (grav && i < np) ? P->gy : 0.0;
    s->ax[i] = fx + ex + avfx;
    s->ay[i] = fy + ey + avfy;
    double tds = 0.0;
    if (visc) {
      const double* ti = tt + 3 * i;
      tds = 0.5 * eta / rho[i] * (ti[0] * ti[0] + ti[1] * ti[1] + 2.0 * ti[2] * ti[2]);
    }
    s->du[i] = (tds + 0.5 * due) + avdu + ahdu + cdu;
    if ((P->flags & ORC_HEAT_SOURCE) && i < np) s->du[i] += heat_source(P, x[i], y[i], rho[i]);
    if (dep_active(P) && i < np) s->du[i] += deposition(P, x[i], y[i], rho[i]);
    s->avx[i] = xsx;
    s->avy[i] = xsy;
    if (hnew) {
      double hn = s->h[i];
      if (i < np && P->sle == 1) {
        hn = 2.0 * pow(m / rho[i], (double)(1.0f / (float)hdim));
      } else if (i < np && P->sle == 2) {
        const double dh = (s->h[i] / hdim) * dvh;
        hn = s->h[i] + P->dt * dh;
        if (hn <= 0.0) hn = hn - P->dt * dh;
      }
      hnew[i] = hn;
    }
  }
  if (hnew) {
    memcpy(s->h, hnew, sizeof(double) * (size_t)n);
    free(hnew);
  }
  if (P->flags & ORC_CONTACT) {
    /* contact_displacement and smooth_vector (external_force.f90:191-339)
       in gather order: d_i from the virtual neighbours, then
       a_i += 0.5 (W(0) m/rho_i d_i + sum_real W_ij m/rho_i d_j) / dt^2 */
    double* cdx = (double*)calloc((size_t)n, sizeof(double));
    double* cdy = (double*)calloc((size_t)n, sizeof(double));
    const double hh = P->h / P->contact_hdr, cut2 = (2.0 * P->h) * (2.0 * P->h);
    const int nbv = s->nb;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < np; ++i) {
      const int key = C.key[i];
      const int cx = key % nx, cy = key / nx;
      double ddx = 0.0, ddy = 0.0;
      for (int ry = cy - 1; ry <= cy + 1; ++ry) {
        if (ry < 0 || ry >= ny) continue;
        span_t sp[2];
        const int ns = row_spans(P, C.start, cx, ry, sp);
        for (int q = 0; q < ns; ++q)
        for (int t = sp[q].t0; t < sp[q].t1; ++t) {
          const int j = C.sorted[t];
          if (j < np || j >= nbv) continue;
          const double rx = dx_of(P, x[i], x[j], sp[q].sh), ry_ = y[i] - y[j];
          double r2 = rx * rx;
          r2 = r2 + ry_ * ry_;
          if (!(r2 < cut2)) continue;
          double pni = 0.0, pnm = 0.0;
          pni = pni + rx * vx[j];
          pnm = pnm + ddx * vx[j];
          pni = pni + ry_ * vy[j];
          pnm = pnm + ddy * vy[j];
          pni = hh - pni;
          if (pni > 0.0 && pni > pnm) {
            ddx = ddx + (pni - pnm) * vx[j];
            ddy = ddy + (pni - pnm) * vy[j];
          }
        }
      }
      cdx[i] = ddx;
      cdy[i] = ddy;
    }
    double w0x, w0y;
    const double w0 = kern(P, 0.0, 0.0, 0.0, &w0x, &w0y);
    const double dt2 = P->dt * P->dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < np; ++i) {
      const int key = C.key[i];
      const int cx = key % nx, cy = key / nx;
      double sx = w0 * m / s->rho[i] * cdx[i], sy = w0 * m / s->rho[i] * cdy[i];
      for (int ry = cy - 1; ry <= cy + 1; ++ry) {
        if (ry < 0 || ry >= ny) continue;
        span_t sp[2];
        const int ns = row_spans(P, C.start, cx, ry, sp);
        for (int q = 0; q < ns; ++q)
        for (int t = sp[q].t0; t < sp[q].t1; ++t) {
          const int j = C.sorted[t];
          if (j == i || j >= np) continue;
          const double rx = dx_of(P, x[i], x[j], sp[q].sh), ry_ = y[i] - y[j];
          double r2 = rx * rx;
          r2 = r2 + ry_ * ry_;
          if (!(r2 < cut2)) continue;
          double wx, wy;
          const double w = kern(P, sqrt(r2), rx, ry_, &wx, &wy);
          sx = sx + w * m / s->rho[i] * cdx[j];
          sy = sy + w * m / s->rho[i] * cdy[j];
        }
      }
      s->ax[i] = s->ax[i] + 0.5 * sx / dt2;
      s->ay[i] = s->ay[i] + 0.5 * sy / dt2;
    }
    free(cdx);
    free(cdy);
  }
  free(tt);
  free(psi);
  free_cells(&C);
  return 0;
}

void orc_predictor(const orc_params* P, orc_state* s) {
  /* time_integration.f:136-154; walls (DBC) integrate rho and u only */
  const double dt = P->dt;
  const int nr = nrate_of(P, s);
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
    }
  }
}

/* Periodic in x: a position that left [ox, ox + nx cell) re-enters on the
   other side. */
static double wrap_x(const orc_params* P, double x) {
  if (!(P->flags & ORC_PERIODIC_X)) return x;
  const double L = (double)P->nx * P->cell;
  if (x >= P->ox + L) return x - L;
  if (x < P->ox) return x + L;
  return x;
}

void orc_corrector(const orc_params* P, orc_state* s, int its) {
  /* time_integration.f:186-217 */
  const double dt = P->dt;
  const int nr = nrate_of(P, s);
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
        s->x[i] = wrap_x(P, s->x[i]);
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
        s->x[i] = wrap_x(P, s->x[i]);
      }
    }
  }
}

int orc_run(const orc_params* P, orc_state* s, int its0, int nsteps, int gather) {
  int cap = 0, niac = 0;
  int *pi = NULL, *pj = NULL;
  if (!gather) {
    cap = s->n * 64 + 64;
    pi = (int*)malloc(sizeof(int) * (size_t)cap);
    pj = (int*)malloc(sizeof(int) * (size_t)cap);
  }
  int rc = 0;
  orc_params Q = *P;
  for (int its = its0; its < its0 + nsteps && rc == 0; ++its) {
    Q.dep_on = its * P->dt <= P->dep_tend; /* single_step.f:219 */
    P = &Q;
    if (its != 1) orc_predictor(P, s);
    if ((P->flags & ORC_GHOSTS) && s->gpar && orc_ghost_particles(P, s) < 0) {
      rc = -4;
      break;
    }
    if (gather) {
      rc = orc_single_step_gather(P, s);
      if (rc) rc = -3;
    } else {
      rc = orc_link_list(P, s->n, s->x, s->y, cap, &niac, pi, pj);
      if (rc == 0) rc = orc_single_step(P, s, niac, pi, pj);
    }
    if (rc == 0) orc_corrector(P, s, its);
  }
  free(pi);
  free(pj);
  return rc;
}

void orc_conservation(const orc_params* P, const orc_state* s, double out[5]) {
  double mass = 0, mx = 0, my = 0, ke = 0, ie = 0;
  for (int i = 0; i < s->np; ++i) {
    mass += P->mass;
    mx += P->mass * s->vx[i];
    my += P->mass * s->vy[i];
    ke += 0.5 * P->mass * (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i]);
    ie += P->mass * s->u[i];
  }
  out[0] = mass;
  out[1] = mx;
  out[2] = my;
  out[3] = ke;
  out[4] = ie;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ------------------------------------------------------------------ */
/* FP32 cell + offset criterion                                         */
/* ------------------------------------------------------------------ */

void orc_split_f32(const orc_params* P, int n, const double* x, const double* y,
                   int32_t* cx, int32_t* cy, float* offx, float* offy) {
  for (int i = 0; i < n; ++i) {
    const int ix = (int)((x[i] - P->ox) / P->cell);
    const int iy = (int)((y[i] - P->oy) / P->cell);
    const double bx = P->ox + (double)ix * P->cell;
    const double by = P->oy + (double)iy * P->cell;
    cx[i] = ix;
    cy[i] = iy;
    offx[i] = (float)(x[i] - bx);
    offy[i] = (float)(y[i] - by);
  }
}

static int cmp_int(const void* a, const void* b) {
  const int ia = *(const int*)a, ib = *(const int*)b;
  return (ia > ib) - (ia < ib);
}

/* generic lexicographic half list over the fixed grid using a predicate */
typedef int (*pair_pred)(const void* ctx, int i, int j, int sh);

static int half_pairs(const orc_params* P, int n, const int32_t* cx, const int32_t* cy,
                      pair_pred pred, const void* ctx, int cap, int* niac, int* pair_i,
                      int* pair_j) {
  const int nx = P->nx, ny = P->ny, ncell = nx * ny;
  int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ncell + 1));
  for (int i = 0; i < n; ++i) {
    key[i] = (cx[i] >= 0 && cx[i] < nx && cy[i] >= 0 && cy[i] < ny) ? cx[i] + cy[i] * nx : -1;
  }
  orc_sort_cells(n, key, ncell, sorted, start);
  int* buf = (int*)malloc(sizeof(int) * 4096);
  int cnt = 0, rc = 0;
  for (int i = 0; i < n && rc == 0; ++i) {
    if (key[i] < 0) continue;
    int nb = 0;
    for (int ry = cy[i] - 1; ry <= cy[i] + 1; ++ry) {
      if (ry < 0 || ry >= ny) continue;
      span_t sp[2];
      const int ns = row_spans(P, start, cx[i], ry, sp);
      for (int q = 0; q < ns; ++q)
        for (int t = sp[q].t0; t < sp[q].t1; ++t) {
          const int j = sorted[t];
          if (j > i && pred(ctx, i, j, sp[q].sh) && nb < 4096) buf[nb++] = j;
        }
    }
    qsort(buf, (size_t)nb, sizeof(int), cmp_int);
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
  free(key);
  free(sorted);
  free(start);
  return rc;
}

typedef struct {
  const int32_t *cx, *cy;
  const float *ox, *oy;
  float cellf, cut2f;
  int nx;
} f32_ctx;

static int pred_f32(const void* vctx, int i, int j, int sh) {
  const f32_ctx* c = (const f32_ctx*)vctx;
  const float tx = c->ox[i] - c->ox[j];
  const float ux = (float)(c->cx[i] - c->cx[j] + sh * c->nx) * c->cellf;
  const float dx = tx + ux;
  const float ty = c->oy[i] - c->oy[j];
  const float uy = (float)(c->cy[i] - c->cy[j]) * c->cellf;
  const float dy = ty + uy;
  const float a = dx * dx;
  const float b = dy * dy;
  const float r2 = a + b;
  return r2 < c->cut2f;
}

int orc_pairs_f32(const orc_params* P, int n, const int32_t* cx, const int32_t* cy,
                  const float* offx, const float* offy, int cap, int* niac, int* pair_i,
                  int* pair_j) {
  f32_ctx c;
  c.cx = cx;
  c.cy = cy;
  c.ox = offx;
  c.oy = offy;
  c.cellf = (float)P->cell;
  c.cut2f = (float)((2.0 * P->h) * (2.0 * P->h));
  c.nx = P->nx;
  return half_pairs(P, n, cx, cy, pred_f32, &c, cap, niac, pair_i, pair_j);
}

typedef struct {
  const double *x, *y;
  double cut2;
  const orc_params* P;
} f64_ctx;

static int pred_f64(const void* vctx, int i, int j, int sh) {
  const f64_ctx* c = (const f64_ctx*)vctx;
  const double dx = dx_of(c->P, c->x[i], c->x[j], sh), dy = c->y[i] - c->y[j];
  double r2 = dx * dx;
  r2 = r2 + dy * dy;
  return r2 < c->cut2;
}

int orc_pairs_f64(const orc_params* P, int n, const double* x, const double* y, int cap,
                  int* niac, int* pair_i, int* pair_j) {
  int32_t* cx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* cy = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    const double tx = (x[i] - P->ox) / P->cell, ty = (y[i] - P->oy) / P->cell;
    cx[i] = (tx >= 0.0) ? (int)tx : -1;
    cy[i] = (ty >= 0.0) ? (int)ty : -1;
  }
  f64_ctx c;
  c.x = x;
  c.y = y;
  c.cut2 = (2.0 * P->h) * (2.0 * P->h);
  c.P = P;
  int rc = half_pairs(P, n, cx, cy, pred_f64, &c, cap, niac, pair_i, pair_j);
  free(cx);
  free(cy);
  return rc;
}

/* ------------------------------------------------------------------ */
/* FP32 C/CUDA path restatements                                        */
/* ------------------------------------------------------------------ */

void orc_clist_f32(int n, const float* x, const float* y, float ox, float oy, float size,
                   int nx, int ny, float smooth, int maxn, int32_t* sorted_idx,
                   int32_t* list) {
  /* updateHashHost (armando2Dcpu.cu:2080-2101), stable (key, index) sort
     (CUDPP radix / qsort, :2177,:2197), sortArraysHost (:2049-2064),
     updateSetsHost (:2103-2127), updateListHost (:2129-2160) */
  /* The dump in ArmandoList/out.txt was written by the host code of an i686
     build (lib/libcudpp_i686), where float expressions are evaluated in x87
     extended precision (FLT_EVAL_METHOD 2); the cell index is therefore
     computed in long double before truncation. */
  const int ncell = nx * ny;
  int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const long double tx = ((long double)x[i] - (long double)ox) / (long double)size;
    const long double ty = ((long double)y[i] - (long double)oy) / (long double)size;
    key[i] = (int)tx + (int)ty * nx;
  }
  int32_t* start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ncell + 1));
  orc_sort_cells(n, key, ncell, sorted_idx, start);
  float* sx = (float*)malloc(sizeof(float) * (size_t)n);
  float* sy = (float*)malloc(sizeof(float) * (size_t)n);
  for (int t = 0; t < n; ++t) {
    sx[t] = x[sorted_idx[t]];
    sy[t] = y[sorted_idx[t]];
  }
  for (int t = 0; t < n; ++t) {
    for (int k = 0; k < maxn; ++k) list[(size_t)t * maxn + k] = t;
    const int ic = key[sorted_idx[t]];
    int np = 0;
    for (int j = -1; j <= 1; ++j)
      for (int i = -1; i <= 1; ++i) {
        const int jc = ic + i + j * nx;
        if (jc < 0 || jc >= ncell) continue;
        for (int jp = start[jc]; jp < start[jc + 1]; ++jp) {
          const float dx = sx[t] - sx[jp];
          const float dy = sy[t] - sy[jp];
          const float dr = sqrtf(dx * dx + dy * dy);
          if (((double)dr < 2.0 * (double)smooth) && (np < maxn)) {
            list[(size_t)t * maxn + np] = jp;
            np++;
          }
        }
      }
  }
  free(sx);
  free(sy);
  free(key);
  free(start);
}

static float deriv_wendland_f32(float r, float h) {
  /* kernelDerivWendland, armando2Dcpu.cu:147-172 (float with double
     sub-expressions as written) */
  const float q = r / h;
  const float alpha = (float)(7.0 / (4.0 * M_PI * (double)h * (double)h));
  float dwdr = 0.0f;
  if (q < 2) {
    dwdr = (float)(5.0 / 8.0 * (double)q * pow(((double)q - 2.0), 3));
    dwdr *= alpha / h;
  }
  return dwdr;
}

void orc_kat_interaction_f32(int n, int maxn, const int32_t* list, const float* x,
                             const float* y, const float* vx, const float* vy,
                             const float* rho, const float* p, float smooth, float mass,
                             float sound, float* drho, float* ax, float* ay) {
  /* kerInteraction, armando2Dlist.cu:611-686 */
  for (int ip = 0; ip < n; ++ip) {
    float iDensityDot = 0.0f, iVelDotX = 0.0f, iVelDotY = 0.0f;
    for (int il = 0; il < maxn; ++il) {
      const int jp = list[(size_t)ip * maxn + il];
      const float dx = x[ip] - x[jp];
      const float dy = y[ip] - y[jp];
      const float dz = 0.0f;
      float dr = sqrtf(dx * dx + dy * dy + dz * dz);
      if ((double)dr < 0.1 * (double)smooth) dr = (float)(100.0 * (double)smooth);
      const float dwdr = deriv_wendland_f32(dr, smooth);
      float dvr = 0.0f;
      dvr += (x[ip] - x[jp]) * (vx[ip] - vx[jp]);
      dvr += (y[ip] - y[jp]) * (vy[ip] - vy[jp]);
      iDensityDot += mass * dvr * dwdr / dr;
      float f = -(p[ip] + p[jp]) / (rho[ip] * rho[jp]);
      iVelDotX += mass * f * dwdr * (x[ip] - x[jp]) / dr;
      iVelDotY += mass * f * dwdr * (y[ip] - y[jp]) / dr;
      f = rho[ip] - rho[jp];
      f = (float)((double)f * (2.0 * (double)sound / (double)(rho[ip] + rho[jp])));
      iDensityDot += mass * f * dwdr;
      f = (dvr < 0) ? dvr : 0.0f;
      f = (float)((double)f * ((double)smooth /
                               ((double)(dr * dr) + 0.01 * (double)smooth * (double)smooth)));
      f = (float)((double)f * (2. * (double)sound / (double)(rho[ip] + rho[jp])));
      f = (float)((double)f * 0.03);
      iVelDotX += mass * f * dwdr * (x[ip] - x[jp]) / dr;
      iVelDotY += mass * f * dwdr * (y[ip] - y[jp]) / dr;
    }
    drho[ip] = iDensityDot;
    ax[ip] = iVelDotX;
    ay[ip] = iVelDotY;
  }
}
