/*
 * armando_oracle.h -- CPU restatement of Armando's 2D WCSPH time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * library (armando_amd/csrc).  Only tests, the smoke check and the CPU
 * baseline leg of the benchmark load it; the product path never does.
 *
 * Two families of routines:
 *
 *  1. "Faithful" FP64 routines that restate the CERN Fortran hot path
 *     routine by routine, with the same half pair list (i < j), the same
 *     pair order and the same scatter-to-both-ends accumulation:
 *       link_list      CERNsource/link_list.f:1-223
 *       kernel         CERNsource/kernel.f:1-107 (cubic) and the Wendland C2
 *                      of ArmandoList/armando2Dcpu.cu:116-172
 *       sum_density    CERNsource/density.f:1-62
 *       con_density    CERNsource/density.f:166-223
 *       int_force      CERNsource/internal_force.f:1-437 (pressure part)
 *       art_visc       CERNsource/art_visc.f:1-103
 *       art_heat       CERNsource/art_heat.f:1-99
 *       av_vel         CERNsource/av_vel.f:1-58
 *       eos (Tait)     FORTRANsource/eos.f90:91-102,255-283
 *       time_integration / single_step
 *                      CERNsource/time_integration.f:131-235,
 *                      CERNsource/single_step.f:92-234
 *     Walls follow the dynamic boundary of ArmandoList/armando2Dcpu.cu:
 *     832-907 (density and energy integrated, velocity held) and gravity
 *     follows updateForcesHost (:937-958).
 *
 *  2. A "gather" FP64 twin of single_step that visits, for every particle,
 *     its full neighbour set in (cell, id) order (OpenMP parallel).  Same
 *     mathematics, different summation order; used for large runs.
 *
 *  3. FP32 restatements of the C/CUDA benchmark path used only to pin
 *     conventions against the reference's fixtures: the cell hash, stable
 *     sort, start/stop sets and MAXN list (armando2Dcpu.cu:2080-2160) and the
 *     continuity + pressure + shock-correction interaction of the known
 *     answer test (ArmandoList/armando2Dlist.cu:611-686, 2451-2548).
 *
 *  4. The neighbour criterion of the FP32 "cell + offset" form of the HIP
 *     library, restated in float arithmetic so that neighbour sets can be
 *     compared bit for bit.
 *
 * All indices are 0-based.  Particles 0..np-1 are fluid, np..n-1 are wall
 * (virtual) particles, as in time_integration.f:104 (np real, then nv).
 */
#ifndef ARMANDO_ORACLE_H
#define ARMANDO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_K_WENDLAND2D = 0, /* alpha = 7/(4 pi h^2), support 2h */
  ORC_K_CUBIC2D = 1,    /* kernel.f skf=0, dim=2 factor, REAL(4) constants */
  ORC_K_CUBIC3D = 2,    /* kernel.f skf=0, dim=3 factor (for the as-shipped dim=3 build) */
  ORC_K_WENDLAND3D = 3  /* alpha = 21/(16 pi h^3), support 2h */
};

enum {
  ORC_ART_VISC = 1,  /* art_visc.f */
  ORC_ART_HEAT = 2,  /* art_heat.f */
  ORC_AV_VEL = 4,    /* av_vel.f (XSPH) */
  ORC_DBC = 8,       /* walls integrate density and energy (armando2Dcpu.cu:832-907) */
  ORC_GRAVITY = 16,  /* body force on fluid (armando2Dcpu.cu:937-958) */
  ORC_VISC = 32,     /* physical viscosity: shear tensor, viscous force and
                        tds/dt of internal_force.f:73-171, uniform eta */
  ORC_HEAT_SOURCE = 64, /* du += hs_q0 exp(-|x - c|^2/(2 hs_sigma^2)) / rho on real
                          particles (the exdudt slot of single_step.f:217-222) */
  ORC_CONDUCTION = 128, /* Cleary-Monaghan heat conduction, uniform conductivity k:
                           du_i += sum m 4 k k/(k + k) (T_i - T_j)/(rho_i rho_j)
                                   (x_ij . grad_i W)/(r^2 + (h ah_eta)^2), T = u/cp */
  ORC_PERIODIC_X = 256, /* periodic in x with period nx cell from ox: the cells
                           cx - 1 and cx + 1 wrap, x_i - x_j gains -+ nx cell */
  ORC_GHOSTS = 512,     /* mirror ghosts of ghost.f: the virtual particles (np .. nb-1)
                           carry the wall normal in v; every step each real particle
                           with a virtual neighbour gets a ghost mirrored across the
                           plane of the nearest one (orc_ghost_particles) */
  ORC_CONTACT = 1024    /* contact displacement of FORTRANsource/external_force.f90:
                           the virtual particles carry the wall normal in v; after the
                           forces every real particle gains
                           a += 0.5 S(d)/dt^2, d from orc_contact_displacement and
                           S = orc_smooth_vector (time_integration.f90:128-142) */
};

enum {
  ORC_EOS_TAIT = 0,  /* p = rho0 c0^2/7 ((rho/rho0)^7 - 1), eos.f90:274 */
  ORC_EOS_LINEAR = 1 /* p = c0^2 rho + eos_p0 (the a1 term of eos_poly, eos.f) */
};

typedef struct {
  int kernel;            /* ORC_K_* */
  int flags;             /* ORC_* bit set */
  int pa_sph;            /* 0: -(pi+pj)/(ri rj), 1: -(pi/ri^2 + pj/rj^2) */
  int density_method;    /* 0: continuity, 1: summation */
  double h;              /* uniform smoothing length */
  double mass;           /* uniform particle mass */
  double rho0, c0, pmin; /* Tait: p = rho0 c0^2/7 ((rho/rho0)^7 - 1), c = c0 */
  double av_alpha, av_beta, av_eta; /* Monaghan artificial viscosity */
  double ah_g1, ah_g2, ah_eta;      /* Noh artificial heat */
  double xsph_eps;                  /* XSPH epsilon */
  double gx, gy;                    /* body force on fluid particles */
  double dt;
  /* fixed hash grid (C path convention): key = ix + iy*nx,
     ix = (int)((x - ox)/cell) */
  double ox, oy, cell;
  int nx, ny;
  double visc_eta; /* dynamic viscosity (ORC_VISC) */
  int eos;         /* ORC_EOS_* ; c = c0 for both, clamp p < pmin -> 0, rho0 */
  double eos_p0;   /* ORC_EOS_LINEAR pressure offset */
  double hs_q0, hs_x, hs_y, hs_sigma; /* ORC_HEAT_SOURCE */
  int dim;         /* 2 (default when 0) or 3: the orc3_* routines */
  double oz, gz;   /* 3D: grid origin in z, body force in z */
  int nz;          /* 3D: grid cells in z (key = ix + iy nx + iz nx ny) */
  double cond_k, cp; /* ORC_CONDUCTION: conductivity (W/m/K), heat capacity (J/kg/K) */
  int sle;           /* smoothing-length evolution (hsml.f): 0 fixed h (s->h unused),
                        1 h = 2 (m/rho)^(1/dim), 2 dh/dt = (h/dim) div v; with 1 or 2
                        every pair uses mh = (h_i + h_j)/2 and real particles update
                        s->h after the forces (gather twin only) */
  /* binned energy deposition (fluka_heat.f:201-704, Cartesian binning,
     tfluka = 0; the 2D routines read the slice z = 0): with dep != NULL and
     dep_on, real particles
     gain du += (trilinear map value) / rho after every other term (the
     exdudt slot of single_step.f:217-226).  dep holds dep_n[0] dep_n[1]
     dep_n[2] bin values (W/m^3), x fastest; dep_o and dep_d are the origin
     and bin size.  orc_run sets dep_on for step its to its dt <= dep_tend
     (single_step.f:219); a direct single-step call uses dep_on as given. */
  const double* dep;
  int dep_n[3];
  double dep_o[3], dep_d[3];
  double dep_tend;
  int dep_on;
  double contact_hdr; /* ORC_CONTACT: the contact distance is h / contact_hdr
                         (hdr of FORTRANsource/options.inc, 1.2) */
  int dep_type; /* the binning of the deposition map: 0 Cartesian (tfluka = 0), 1 R-phi-z
                   (tfluka = 1 or 7; fluka_heat.f:341-481 in 3D, :593-704 in 2D): the axes
                   are r = sqrt(x^2 + y^2), phi = atan2(y, x) in degrees and z, with the
                   bin pairs clamped to the map as in the Cartesian branch */
} orc_params;

/* Particle state, SoA, FP64. */
typedef struct {
  int n, np;
  double *x, *y, *vx, *vy, *rho, *u, *p, *c;
  /* rates of the last single_step */
  double *drho, *du, *ax, *ay, *avx, *avy, *divv;
  /* leapfrog save arrays */
  double *rho_min, *u_min, *vx_min, *vy_min;
  /* per-particle smoothing length (P->sle != 0) */
  double* h;
  /* ORC_GHOSTS (orc_state_new_ghosts): particles nb .. n-1 are the ghosts of
     the current step, ghost nb + g mirrors real particle gpar[g]; xp, yp hold
     the positions of the last neighbour search (have_prev = 0: none yet);
     cap = allocated particles.  Without ghosts nb = cap = n. */
  int nb, cap, have_prev;
  double *xp, *yp;
  int* gpar;
} orc_state;

/* fluka_heat.f:485-573 (dim = 2, tfluka = 0): dedt[i] of real particles
   0 .. np-1, 0 outside the binning; the z slice of the 2D branch is z = 0. */
void orc_fluka_heat(const orc_params* P, int np, const double* x, const double* y,
                    const double* rho, double* dedt);
/* The interpolated map value (before the division by rho) at (x, y, z)
   (fluka_heat.f:235-337, the dim = 3 branch; the 2D branch is the same at
   z = 0), and whether the point lies inside the binning. */
double orc_deposition_at(const orc_params* P, double x, double y, double z, int* inside);

/* ---- kernel ---------------------------------------------------------- */
/* w and the factor F with grad_i W = F * (x_i - x_j) (kernel.f returns
   dwdx = F*dx).  Returns w. */
double orc_kernel(const orc_params* P, double r, double* F);

/* ---- neighbour search ------------------------------------------------ */
/* link_list.f restated for dim=2 (data-dependent grid, scale_k=2, half
   list j>i, same visiting order).  Returns 0, or -1 if cap is exceeded. */
int orc_link_list(const orc_params* P, int nt, const double* x, const double* y,
                  int cap, int* niac, int* pair_i, int* pair_j);
/* direct_find.f analogue: O(n^2) brute force, pairs (i<j) in lexicographic
   order.  Returns 0 or -1 on cap overflow. */
int orc_direct_find(const orc_params* P, int nt, const double* x, const double* y,
                    int cap, int* niac, int* pair_i, int* pair_j);

/* Fixed-grid cell keys (FP64 form).  Returns number of particles outside
   the grid (their key is set to -1). */
int orc_keys(const orc_params* P, int n, const double* x, const double* y, int32_t* key);
/* Stable sort of particles by key, ties by id (input listed in id order):
   sorted_id[t] = id at sorted position t; cell_start has ncell+1 entries
   with cell_start[c] = first sorted position whose key >= c. */
void orc_sort_cells(int n, const int32_t* key, int ncell, int32_t* sorted_id,
                    int32_t* cell_start);

/* ---- faithful pair routines (half list, scatter) --------------------- */
void orc_sum_density(const orc_params* P, int nrate, int niac, const int* pi, const int* pj,
                     const double* x, const double* y, double* rho);
void orc_con_density(const orc_params* P, int nrate, int niac, const int* pi, const int* pj,
                     const double* x, const double* y, const double* vx, const double* vy,
                     double* drho);
/* int_force; zeroes ax, ay, du for [0, nt) first.  With ORC_VISC: the
   shear tensor sums (internal_force.f:73-142), tds/dt = eta/(2 rho) T:T
   (:144-169) and the viscous force (eta_i T_i + eta_j T_j).grad W; finishes
   with du = tdsdt + 0.5*du (:433-435). */
void orc_int_force(const orc_params* P, int nt, int niac, const int* pi, const int* pj,
                   const double* x, const double* y, const double* vx, const double* vy,
                   const double* rho, const double* p, double* ax, double* ay, double* du);
/* art_visc accumulates into ax, ay, du. */
void orc_art_visc(const orc_params* P, int niac, const int* pi, const int* pj,
                  const double* x, const double* y, const double* vx, const double* vy,
                  const double* rho, const double* c, double* ax, double* ay, double* du);
/* art_heat accumulates into du; divv (length nt, may be NULL) receives pass 1. */
void orc_art_heat(const orc_params* P, int nt, int niac, const int* pi, const int* pj,
                  const double* x, const double* y, const double* vx, const double* vy,
                  const double* rho, const double* u, const double* c, double* du,
                  double* divv);
/* av_vel accumulates into avx, avy. */
void orc_av_vel(const orc_params* P, int niac, const int* pi, const int* pj,
                const double* x, const double* y, const double* vx, const double* vy,
                const double* rho, double* avx, double* avy);
/* contact_displacement (FORTRANsource/external_force.f90:191-244): over the
   pairs (i real, j virtual) in list order, pni = h/hdr - (x_i - x_j).n_j,
   pnm = d_i.n_j and, when pni > 0 and pni > pnm, d_i += (pni - pnm) n_j,
   with n_j = (vx_j, vy_j); d accumulates into (dx, dy). */
void orc_contact_displacement(const orc_params* P, int np, int nv, const double* x,
                              const double* y, const double* vx, const double* vy, int niac,
                              const int* pi, const int* pj, double* dx, double* dy);
/* smooth_vector (FORTRANsource/external_force.f90:247-339) on the real
   particles: vec_i <- W(0) m/rho_i vec_i + sum over real pairs W_ij m/rho_i
   vec_j (each pair at both ends). */
void orc_smooth_vector(const orc_params* P, int np, const double* x, const double* y,
                       const double* rho, int niac, const int* pi, const int* pj, double* vx,
                       double* vy);
/* Tait EOS on [0, n): p, c, and the pmin clamp (p = 0, rho = rho0). */
void orc_eos_tait(const orc_params* P, int n, double* rho, double* p, double* c);
/* The selected EOS (P->eos) with the same clamp. */
void orc_eos(const orc_params* P, int n, double* rho, double* p, double* c);

/* ---- state and step ---------------------------------------------------- */
orc_state* orc_state_new(int n, int np);
/* A state of np real and nv virtual particles with room for np ghosts. */
orc_state* orc_state_new_ghosts(int np, int nv);
void orc_state_free(orc_state* s);
/* ghost.f:1-86 with a candidate list restated: a real particle i's
   candidates are the virtual particles j with |xp_i - x_j|^2 < (2h)^2 (the
   pairs of the last neighbour search, time_integration.f:157-161 calls it
   before the new link_list); the nearest at the current positions wins
   (ties: the smaller index).  Ghosts are appended in the order of their
   parents: x_g = x_i - 2 n (x_i - x_j).n, v_g = v_i - 2 n v_i.n with n = v_j,
   and rho, p, u, c, h of the parent.  Then xp, yp = x, y (the search of this
   step).  Returns the ghost count, or -1 if the capacity is exceeded. */
int orc_ghost_particles(const orc_params* P, orc_state* s);
/* single_step.f restated with the faithful half-list routines. */
int orc_single_step(const orc_params* P, orc_state* s, int niac, const int* pi, const int* pj);
/* Same rates, gather form over the fixed grid (OpenMP).  Returns the
   number of particles outside the grid (0 on success). */
int orc_single_step_gather(const orc_params* P, orc_state* s);
/* time_integration.f predictor (its != 1) and corrector. */
void orc_predictor(const orc_params* P, orc_state* s);
void orc_corrector(const orc_params* P, orc_state* s, int its);
/* nsteps of time_integration starting at step number its0 (1-based).
   gather=0: faithful serial half-list path; gather=1: gather path.  With
   ORC_GHOSTS the ghosts are rebuilt after the predictor of every step.
   Returns 0, or a negative error. */
int orc_run(const orc_params* P, orc_state* s, int its0, int nsteps, int gather);
/* FP64 conservation sums: mass, momentum x, momentum y, kinetic energy,
   internal energy (fluid particles). */
void orc_conservation(const orc_params* P, const orc_state* s, double out[5]);
/* Number of threads OpenMP would use (1 without OpenMP). */
int orc_num_threads(void);
void orc_set_threads(int nthreads);

/* ---- discrete adjoint (armando_adjoint.c) ------------------------------ */
/* nsteps of time_integration from s0 (its = 1 .. nsteps; s0 is not
   modified), objective J = sum_steps dt sum_{objmask[i]} p_i^2, and its
   gradients with respect to av_alpha and the initial velocities (the
   neighbour list of every step is held fixed).  Returns 0, -5 for
   summation density, or a link_list error. */
int orc_adjoint(const orc_params* P, const orc_state* s0, int nsteps, const int* objmask,
                double* J, double* dJ_dalpha, double* dJ_dvx0, double* dJ_dvy0);

/* ---- FP32 cell + offset form ----------------------------------------- */
/* Split FP64 positions into integer cell coordinates and FP32 in-cell
   offsets: ix = (int)((x-ox)/cell), off = (float)(x - (ox + ix*cell)). */
void orc_split_f32(const orc_params* P, int n, const double* x, const double* y,
                   int32_t* cx, int32_t* cy, float* offx, float* offy);
/* Half pair list (i<j, ids) under the FP32 criterion
   dx = (offx_i - offx_j) + (float)(cx_i - cx_j) * cellf, same for y,
   r2 = dx*dx + dy*dy (no fused multiply-add), pair iff r2 < r2cut_f with
   cellf = (float)cell and r2cut_f = (float)((2h)*(2h)).  Pairs are emitted
   in lexicographic (i, j) order.  Returns 0 or -1 on cap overflow. */
int orc_pairs_f32(const orc_params* P, int n, const int32_t* cx, const int32_t* cy,
                  const float* offx, const float* offy, int cap, int* niac,
                  int* pair_i, int* pair_j);
/* Half pair list (i<j) under the FP64 criterion r2 = dx*dx + dy*dy <
   (2h)*(2h), lexicographic order, computed with the fixed grid. */
int orc_pairs_f64(const orc_params* P, int n, const double* x, const double* y,
                  int cap, int* niac, int* pair_i, int* pair_j);

/* ---- FP32 C/CUDA path restatements (fixtures) ------------------------ */
/* updateHashHost + stable sort + updateSetsHost + updateListHost
   (armando2Dcpu.cu:2080-2160) in float arithmetic.  Inputs are in the
   original order; outputs are in sorted order: sorted_idx[t] = original
   index, list[t*maxn + k] (padded with t). */
void orc_clist_f32(int n, const float* x, const float* y, float ox, float oy, float size,
                   int nx, int ny, float smooth, int maxn, int32_t* sorted_idx,
                   int32_t* list);
/* kerInteraction (armando2Dlist.cu:611-686) in float over a list built by
   orc_clist_f32; arrays in sorted order.  Writes (not accumulates). */
void orc_kat_interaction_f32(int n, int maxn, const int32_t* list, const float* x,
                             const float* y, const float* vx, const float* vy,
                             const float* rho, const float* p, float smooth, float mass,
                             float sound, float* drho, float* ax, float* ay);

/* ---- 3D (dim = 3), oracle/armando_oracle3d.c --------------------------- */
/* Particle state in 3D, SoA, FP64. */
typedef struct {
  int n, np;
  double *x, *y, *z, *vx, *vy, *vz, *rho, *u, *p, *c;
  double *drho, *du, *ax, *ay, *az, *avx, *avy, *avz, *divv;
  double *rho_min, *u_min, *vx_min, *vy_min, *vz_min;
} orc3_state;

orc3_state* orc3_state_new(int n, int np);
void orc3_state_free(orc3_state* s);
/* 3D cell keys ix + iy nx + iz nx ny; returns particles outside the grid. */
int orc3_keys(const orc_params* P, int n, const double* x, const double* y, const double* z,
              int32_t* key);
/* Unique pairs i<j with r^2 < (2h)^2 (FP64), sorted by (i, j). */
int orc3_pairs_f64(const orc_params* P, int n, const double* x, const double* y,
                   const double* z, int cap, int* niac, int* pair_i, int* pair_j);
/* FP32 form: integer cell and FP32 in-cell offset, and the pairs of the
   criterion dx = (off_i - off_j) + (float)(c_i - c_j) * (float)cell. */
void orc3_split_f32(const orc_params* P, int n, const double* x, const double* y,
                    const double* z, int32_t* cx, int32_t* cy, int32_t* cz, float* ox,
                    float* oy, float* oz);
int orc3_pairs_f32(const orc_params* P, int n, const int32_t* cx, const int32_t* cy,
                   const int32_t* cz, const float* ox, const float* oy, const float* oz,
                   int cap, int* niac, int* pair_i, int* pair_j);
/* The per-routine outputs of one step in gather form (every particle counted
   as rate particle): sum_density rho_sum, con_density drho, int_force a/du
   (P->pa_sph), art_visc a/du, art_heat du (+ divv), av_vel av.  Arrays of n. */
int orc3_routines(const orc_params* P, int n, const double* x, const double* y,
                  const double* z, const double* vx, const double* vy, const double* vz,
                  const double* rho, const double* p, const double* u, const double* c,
                  double* rho_sum, double* drho, double* if_a, double* if_du, double* av_a,
                  double* av_du, double* ah_du, double* divv, double* xs);
/* single_step (gather twin, (cell, id) order) and time_integration in 3D. */
int orc3_single_step_gather(const orc_params* P, orc3_state* s);
void orc3_predictor(const orc_params* P, orc3_state* s);
void orc3_corrector(const orc_params* P, orc3_state* s, int its);
int orc3_run(const orc_params* P, orc3_state* s, int its0, int nsteps);
/* fluka_heat.f:235-337 (dim = 3, tfluka = 0): dedt of particles 0 .. np-1,
   0 outside the binning. */
void orc3_fluka_heat(const orc_params* P, int np, const double* x, const double* y,
                     const double* z, const double* rho, double* dedt);

#ifdef __cplusplus
}
#endif
#endif
