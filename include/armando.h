/*
 * armando.h -- C ABI of the MI355X (gfx950) WCSPH time step.
 *
 * One extern "C" entry point per routine of the reference's step
 *   link_list -> density -> eos -> internal_force + art_visc + art_heat
 *   -> av_vel -> time_integration
 * with the same argument meaning, operating on particle state that a
 * context keeps resident in HBM.  Plain pointers and sizes only.
 *
 * Conventions (reference file:line in brackets):
 *  - Particles 0..np-1 are real (fluid), np..np+nv-1 are virtual (wall)
 *    particles appended after them [CERNsource/time_integration.f:104].
 *  - Host arrays use the Fortran layout x(dim, n): x and y of particle i
 *    are adjacent [CERNsource/time_integration.f:62].  Indices are 0-based.
 *  - Scalars that the reference hard-codes as PARAMETERs (alpha, beta, eta
 *    of art_visc.f:38-44; g1, g2, eta of art_heat.f:38-40; eps of
 *    av_vel.f:35) are fields of arm_params.
 *  - Every call returns ARM_OK (0) or a negative ARM_ERR_* code; the
 *    reference's `stop` on overflow [CERNsource/link_list.f:63-67,169-173]
 *    becomes ARM_ERR_CAPACITY / ARM_ERR_DOMAIN, never a silent truncation.
 *  - A context is used from one host thread at a time; all work is ordered
 *    on the context's HIP stream and every call returns after it finished.
 */
#ifndef ARMANDO_H
#define ARMANDO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARM_OK 0
#define ARM_ERR_ARG (-1)         /* bad argument */
#define ARM_ERR_HIP (-2)         /* HIP runtime error */
#define ARM_ERR_CAPACITY (-3)    /* a caller-sized output is too small */
#define ARM_ERR_DOMAIN (-4)      /* a particle left the hash grid */
#define ARM_ERR_UNSUPPORTED (-5) /* option outside what this build implements */
#define ARM_ERR_STATE (-6)       /* call order violated (e.g. no link_list yet) */

/* precision forms */
#define ARM_FP64 0 /* plain form: every field and pair term in FP64 */
#define ARM_FP32 1 /* chosen form: FP32 pair arithmetic, cell + FP32 offset
                      positions, Kahan-compensated rate sums, FP64 position
                      and energy accumulators */

/* flags (single_step.f logicals) */
#define ARM_ART_VISC 1 /* visc_artificial -> art_visc.f */
#define ARM_ART_HEAT 2 /* heat_artificial -> art_heat.f */
#define ARM_AV_VEL 4   /* average_velocity -> av_vel.f (XSPH) */
#define ARM_DBC 8      /* walls integrate density and energy, velocity held
                          [ArmandoList/armando2Dcpu.cu:832-907] */
#define ARM_GRAVITY 16 /* body force (gx, gy) on real particles
                          [ArmandoList/armando2Dcpu.cu:937-958] */
#define ARM_VISC 32    /* visc -> physical viscosity in int_force: shear
                          tensor, viscous force, tds/dt
                          [CERNsource/internal_force.f:73-171, 196-430] */
#define ARM_HEAT_SOURCE 64 /* volumetric Gaussian heat source on real particles:
                              du += hs_q0 exp(-|x - (hs_x, hs_y)|^2 / (2 hs_sigma^2)) / rho
                              (the external heating slot exdudt of
                              CERNsource/single_step.f:217-222) */

#define ARM_CONDUCTION 128 /* Cleary-Monaghan heat conduction with a uniform conductivity:
                              du_i += sum m 4 k k/(k + k) (T_i - T_j)/(rho_i rho_j)
                                      (x_ij . grad_i W)/(r^2 + (h ah_eta)^2), T = u/cp
                              (no counterpart in the reference) */
#define ARM_PERIODIC_X 256 /* periodic in x with period nx*cell from ox: the neighbour
                              cells wrap and positions re-enter on the other side
                              (nx >= 3; the reference's wrap is commented out,
                              ArmandoList/armando2Dlist.cu:856-870) */
#define ARM_CONTACT 1024   /* contact displacement of the other tree's boundary
                              [FORTRANsource/external_force.f90:191-339, applied in
                              time_integration.f90:128-142]: the virtual particles
                              carry the unit wall normal in v; after the forces each
                              real particle gains a += 0.5 S(d)/dt^2, d the push to
                              the contact distance h/contact_hdr from its virtual
                              neighbours, S the kernel smoothing over its real
                              neighbours (2D, without ARM_DBC, ARM_GHOSTS,
                              ARM_PERIODIC_X or sle; no adjoint) */
#define ARM_GHOSTS 512     /* mirror ghosts [CERNsource/ghost.f:1-86]: the virtual
                              particles carry the unit wall normal in v; every step
                              each real particle with a virtual neighbour gets a ghost
                              mirrored across the nearest one's plane (2D, without
                              ARM_DBC, ARM_PERIODIC_X or sle; the capacity must hold
                              np + nv plus twice the ghost count) */

/* equations of state */
#define ARM_EOS_TAIT 0   /* p = rho0 c0^2/7 ((rho/rho0)^7 - 1), c = c0
                            [FORTRANsource/eos.f90:255-283] */
#define ARM_EOS_LINEAR 1 /* p = c0^2 rho + eos_p0, c = c0 (eos_poly of
                            CERNsource/eos.f with only the a1 term) */

typedef struct arm_params {
  int precision;      /* ARM_FP64 or ARM_FP32 */
  int flags;          /* ARM_* bit set */
  int pa_sph;         /* 0: -(p_i+p_j)/(rho_i rho_j), 1: -(p_i/rho_i^2 + p_j/rho_j^2)
                         [CERNsource/internal_force.f:193-335] */
  int density_method; /* 0: con_density, 1: sum_density [single_step.f:108-117] */
  double h;           /* smoothing length (uniform; sle = 0) */
  double mass;        /* particle mass (uniform) */
  double rho0, c0, pmin; /* Tait [FORTRANsource/eos.f90:255-283] */
  double av_alpha, av_beta, av_eta;
  double ah_g1, ah_g2, ah_eta;
  double xsph_eps;
  double gx, gy;
  double dt;
  /* hash grid: cell key = ix + iy*nx, ix = (int)((x - ox)/cell)
     [ArmandoList/armando2Dcpu.cu:2080-2101]; cell >= 2h required */
  double ox, oy, cell;
  int nx, ny;
  double visc_eta; /* dynamic viscosity for ARM_VISC (eta of single_step.f:126-184) */
  int eos;         /* ARM_EOS_*; both clamp p < pmin to p = 0, rho = rho0 */
  double eos_p0;   /* ARM_EOS_LINEAR pressure offset */
  double hs_q0, hs_x, hs_y, hs_sigma; /* ARM_HEAT_SOURCE amplitude (W/m^3), centre, width */
  /* dimension: 2 (or 0) or 3.  With 3 the host arrays are x(3, n), v(3, n),
     dvdt(3, n), av(3, n); the hash key is ix + iy*nx + iz*nx*ny with
     iz = (int)((z - oz)/cell) < nz and the 27 neighbour cells of
     CERNsource/link_list.f:49-99 (dim = 3) / Code/armando.cu:3142-3188;
     the kernel is Wendland C2 3D, alpha = 21/(16 pi h^3); gz is the body
     force along z.  ARM_VISC, ARM_HEAT_SOURCE, ARM_CONDUCTION and
     ARM_PERIODIC_X are 2D only. */
  int dim;
  double oz, gz;
  int nz;
  double cond_k, cp; /* ARM_CONDUCTION: conductivity (W/m/K), heat capacity (J/kg/K) */
  int sle; /* smoothing-length evolution [CERNsource/hsml.f:1-76, called at the end of
              single_step.f:225]: 0 fixed h; 1 h = 2 sqrt(m/rho); 2 dh/dt = (h/dim) div v.
              With 1 or 2 (2D only) h is per particle (initially prm.h, or
              arm_upload_h), every pair uses mh = (h_i + h_j)/2 and the direct
              gathers run; 2h must stay within the cell (ARM_ERR_DOMAIN otherwise). */
  double contact_hdr; /* ARM_CONTACT: the contact distance is h / contact_hdr (hdr = 1.2 of
                         FORTRANsource/options.inc) */
} arm_params;

typedef struct arm_ctx arm_ctx;

/* Context with room for `capacity` particles on the current HIP device. */
int arm_create(const arm_params* prm, int capacity, arm_ctx** out);
void arm_destroy(arm_ctx* ctx);
/* Text of the last error on this context ("" if none). */
const char* arm_last_error(const arm_ctx* ctx);
/* Number of particles outside the grid seen by the last failing call. */
int arm_domain_violations(const arm_ctx* ctx);

/* Upload the state; x, v are x(dim, np+nv) (Fortran layout); rho, u have
   np+nv entries.  Replaces [CERNsource/sph.f:104-108] allocation + input. */
int arm_upload(arm_ctx* ctx, int np, int nv, const double* x, const double* v,
               const double* rho, const double* u);
/* Download in the original particle order.  Any pointer may be NULL. */
int arm_download(arm_ctx* ctx, double* x, double* v, double* rho, double* p, double* u);
/* Rates of the last single_step, original order; any pointer may be NULL.
   dvdt and av are (dim, n); divv is the velocity divergence of art_heat. */
int arm_download_rates(arm_ctx* ctx, double* drho, double* dvdt, double* dudt, double* av,
                       double* divv);

/* Per-particle smoothing lengths h(n) in the original order (sle != 0), the
   reference's h array [CERNsource/time_integration.f:62]: set after
   arm_upload, read back after steps. */
int arm_upload_h(arm_ctx* ctx, const double* h);
int arm_download_h(arm_ctx* ctx, double* h);

/* link_list(nt, x, h, niac, pair_i, pair_j, skf) [CERNsource/link_list.f:1]:
   hashes, sorts by (cell, particle) and builds cell ranges on the device.
   niac (may be NULL) receives the number of unique pairs i<j within 2h. */
int arm_link_list(arm_ctx* ctx, int* niac);
/* Materialise the unique pairs (i<j, original ids), sorted by (i, j), for
   parity checks; ARM_ERR_CAPACITY if more than cap. */
int arm_get_pairs(arm_ctx* ctx, int cap, int* niac, int* pair_i, int* pair_j);
/* Sorted cell keys, sorted particle ids and cell_start[nx*ny + 1]
   (cell c holds sorted positions cell_start[c] .. cell_start[c+1]-1).
   key_sorted and id_sorted receive the live particle count: np + nv, plus
   with ARM_GHOSTS the ghosts of the current state (arm_ghost_count), whose
   ids are np + nv + g. */
int arm_get_cells(arm_ctx* ctx, int32_t* key_sorted, int32_t* id_sorted, int32_t* cell_start);

/* sum_density / con_density [CERNsource/density.f:1,166] (method 1 / 0). */
int arm_density(arm_ctx* ctx, int method);
/* EOS on the current density: Tait [FORTRANsource/eos.f90:255-283] or the
   linear EOS (prm.eos), with the p < pmin clamp; also the art_heat
   coefficient from the density pass's velocity divergence. */
int arm_eos(arm_ctx* ctx);
/* int_force [internal_force.f:1] plus, by flags, art_visc [art_visc.f:1]
   and art_heat [art_heat.f:1], fused in one gather; adds gravity when
   ARM_GRAVITY is set in the context flags, and the external heating of
   single_step.f:217-222 (ARM_HEAT_SOURCE, arm_set_deposition) to du.  With ARM_VISC in flags a shear
   tensor pass runs first and the viscous force and tds/dt are added
   [internal_force.f:73-171, 196-430]; with ARM_CONTACT (a context created with it) the
   contact acceleration [FORTRANsource/time_integration.f90:128-142] follows. */
int arm_internal_force(arm_ctx* ctx, int flags);
/* av_vel [CERNsource/av_vel.f:1] (XSPH velocity correction). */
int arm_av_vel(arm_ctx* ctx);
/* single_step [CERNsource/single_step.f:1]: density, eos, forces, av_vel. */
int arm_single_step(arm_ctx* ctx);
/* The two halves of the leapfrog around single_step
   [CERNsource/time_integration.f:136-154 and :186-217]. */
int arm_predictor(arm_ctx* ctx);
int arm_corrector(arm_ctx* ctx, int its);
/* time_integration [CERNsource/time_integration.f:131-235] for steps
   its0 .. its0+nsteps-1 (refresh_step = 1), fused on the device.  With
   ARM_GHOSTS every step makes its ghosts after the predictor, as
   time_integration.f:157-161 does.  Like link_list.f:63-67's stop, the run
   ends at the first step whose corrector puts a particle outside the hash
   grid (or whose force pass grows 2h beyond the cell, sle != 0): the later
   steps do not run, the call returns ARM_ERR_DOMAIN, arm_last_error names
   the step, and the state (arm_download) is the one that step left, with
   the offending particles outside the grid.  Stepping on needs a state
   inside the grid again (arm_upload). */
int arm_time_integration(arm_ctx* ctx, int its0, int nsteps);

/* ghost_particles(np, nv, ng, x, v, h, mass, rho, p, niac, pair_i, pair_j,
   index_ghost) [CERNsource/ghost.f:1]: for each real particle, the virtual
   particles within 2h at the last arm_link_list are its candidates and the
   nearest one now (ties: smaller id) defines the mirror; the ghost gets
   x - 2 n ((x - x_w).n), v - 2 n (v.n) and the parent's rho, p, u.  Ghosts
   are numbered in parent order (ids np+nv+g) and replace the previous set
   at the next arm_link_list (which must follow; an arm_link_list without a
   new set drops the ghosts); after the EOS they take the parent's rho, p,
   u again [single_step.f:188-196].  ng (may be NULL) receives their count. */
int arm_ghost_particles(arm_ctx* ctx, int* ng);
/* Number of ghosts of the current state. */
int arm_ghost_count(arm_ctx* ctx, int* ng);
/* The ghosts in number order (any pointer may be NULL): x, v as (2, ng),
   rho, p, u, and the parent's id. */
int arm_download_ghosts(arm_ctx* ctx, double* x, double* v, double* rho, double* p, double* u,
                        int32_t* parent);

/* Discrete adjoint of time_integration (no counterpart in the reference).
   From the current state as step 1, runs nsteps forward, storing the state
   every ckpt_every steps, with the objective
       J = sum over steps of dt * sum_{i in obj_ids} p_i^2
   (p of that step's single_step), then runs backward one checkpoint
   segment at a time: the segment is recomputed forward onto a tape and the
   hand-written adjoint kernels of the corrector, force, EOS, density and
   predictor run in reverse.  The neighbour list of each step is that of the
   forward step.  Outputs (any may be NULL): J, dJ/d av_alpha, dJ/dv0 as
   (dim, n) in the original order, and the forward / backward device times
   in ms (dJ/d hs_q0: arm_rollout_param_grads).  The context is left at the
   state after nsteps.  Either form (the FP32 form's reverse pair arithmetic
   is FP32, its state adjoint FP64); continuity density; 2D and 3D (3D:
   without the physical viscosity and the heat source, which are 2D only);
   no conduction or periodic boundary. */
int arm_adjoint_rollout(arm_ctx* ctx, int nsteps, int ckpt_every, const int32_t* obj_ids,
                        int nobj, double* J, double* dJ_dalpha, double* dJ_dv0, double* ms_fwd,
                        double* ms_bwd);
/* The two halves of arm_adjoint_rollout, for an autograd wrapper: the
   forward keeps its checkpoints in the context; the backward scales the
   objective adjoint by seed (= dL/dJ) and may be called again. */
int arm_rollout_forward(arm_ctx* ctx, int nsteps, int ckpt_every, const int32_t* obj_ids,
                        int nobj, double* J, double* ms);
int arm_rollout_backward(arm_ctx* ctx, double seed, double* dJ_dalpha, double* dJ_dv0,
                         double* ms);
/* Replace the physical scalars (mass, rho0, c0, pmin, viscosity, heat, XSPH,
   gravity, dt, heat source) of a context; grid, h, precision and flags must
   not change.  Drops the checkpoints of a previous rollout. */
int arm_update_params(arm_ctx* ctx, const arm_params* prm);
/* Energy deposition heat source [CERNsource/fluka_heat.f:201-704, called at
   single_step.f:219-222], Cartesian binning (tfluka = 0): in every step
   its with its*dt <= t_end each real particle gains
   du += v(x) / rho, v the trilinear interpolation between bin centres
   of the map at the particle (2D: at (x, y, z = 0)), clamped to the edge
   bins, 0 outside the binning.  data holds nbin[0]*nbin[1]*nbin[2] values in W/m^3 with x
   fastest (the datafluka array after read_binning_*: FLUKA values times
   np ef/(dt sf^3)); origin and cell are the binning's corner and bin size
   (m); in 2D z = 0 must lie within the z range.  data = NULL switches the source
   off.  The step number is the its of arm_time_integration, or for the
   routine-by-routine calls the one after the last arm_corrector (1 after
   arm_upload).  The rollouts return ARM_ERR_UNSUPPORTED while a map is
   set. */
int arm_set_deposition(arm_ctx* ctx, const double* origin, const double* cell, const int* nbin,
                       const double* data, double t_end);
/* The same with the binning type `tfluka` of fluka_heat.f (its `type`
   argument [CERNsource/fluka_heat.f:201-233, tfluka set by read_binning_*
   :47-75]): 0 Cartesian, as arm_set_deposition; 1 or 7 R-phi-z
   [fluka_heat.f:341-481 (3D), :593-704 (2D)]: the axes of origin, cell,
   nbin and data are r = sqrt(x^2 + y^2) (m), phi = atan2(y, x) in degrees
   and z (m), r fastest, with the bins, weights and clamping of the
   Cartesian branch on those coordinates.  (The reference's R-phi-z branch
   scales z by the phi bin width, :359 and :608, and in 3D picks the bin
   pairs again after clamping, :398-420, which can index past the map;
   neither is followed.)  Other types: ARM_ERR_UNSUPPORTED. */
int arm_set_deposition_binning(arm_ctx* ctx, int type, const double* origin, const double* cell,
                               const int* nbin, const double* data, double t_end);
/* Objective of the next rollout: J = sum over steps of dt * sum_{obj} f_i
   with f = p^2 (ARM_OBJ_PRESSURE, the default) or u^2 (ARM_OBJ_ENERGY), taken
   after the EOS of each step.  No counterpart in the reference. */
#define ARM_OBJ_PRESSURE 0
#define ARM_OBJ_ENERGY 1
int arm_set_objective(arm_ctx* ctx, int kind);
/* Parameter gradients of the last arm_rollout_backward (scaled by its seed):
   out[ARM_GRAD_ALPHA] = dJ/d av_alpha, out[ARM_GRAD_Q0] = dJ/d hs_q0,
   out[ARM_GRAD_VISC] = dJ/d visc_eta (with ARM_VISC).  n is the length of
   out.  No counterpart in the reference. */
#define ARM_GRAD_ALPHA 0
#define ARM_GRAD_Q0 1
#define ARM_GRAD_VISC 2
#define ARM_NGRAD 3
int arm_rollout_param_grads(arm_ctx* ctx, double* out, int n);

/* FP64 sums over real particles: mass, momentum x, momentum y, kinetic
   energy, internal energy. */
int arm_conservation(arm_ctx* ctx, double out[5]);
/* The same in either dimension with the z momentum: mass, momentum x, y, z,
   kinetic energy, internal energy (momentum z = 0 in 2D). */
int arm_conservation3(arm_ctx* ctx, double out[6]);

/* Per-pass device time accumulated since the last reset, in ms, measured
   with HIP events on the context stream; counts[] = launches per pass.
   Pass indices: see ARM_PASS_*.  n is the array length. */
#define ARM_PASS_KEYS 0      /* cell keys of the current positions */
#define ARM_PASS_SORT 1      /* sort by (cell, id): counting sort over the cells
                                (radix sort for grids above 8.4M cells) */
#define ARM_PASS_REORDER 2   /* permute (+ predictor) + cell ranges */
#define ARM_PASS_DENSITY 3   /* density gather (with the EOS inside single_step and
                                time_integration) */
#define ARM_PASS_EOS 4       /* EOS + art_heat coefficient (arm_eos alone) */
#define ARM_PASS_FORCE 5     /* [shear tensor +] internal_force + art_visc + art_heat
                                + av_vel gather */
#define ARM_PASS_INTEGRATE 6 /* predictor / corrector (+ keys of the next step) */
#define ARM_PASS_COUNT 7
int arm_pass_times(arm_ctx* ctx, int n, double* ms, int* counts);
int arm_reset_pass_times(arm_ctx* ctx);
/* Enable (1) or disable (0) per-pass event timing. */
int arm_set_timing(arm_ctx* ctx, int on);
/* No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  The cell keys, the (cell, id) counting sort and the
   cell ranges of arm_link_list + arm_get_cells, without a context, on device
   buffers the caller owns:
     x           device, x(2, n) FP64 (Fortran layout), particle t = id t;
     key_sorted  device, n entries: sorted cell keys;
     id_sorted   device, n entries: particle ids in (cell, id) order;
     cell_start  device, nx*ny + 1 entries, 16-byte aligned: cell c holds
                 sorted positions cell_start[c] .. cell_start[c+1]-1;
     outside     device, 1 entry, or NULL: receives the number of particles
                 outside the grid (arm_link_list's ARM_ERR_DOMAIN; such a
                 particle is put in cell 0 and the result is not valid);
     stream      hipStream_t to enqueue on (NULL: the null stream).
   Of prm only dim (2 or 0), ox, oy, cell, nx, ny are read.  The call
   enqueues the work and returns without waiting; it calls no allocator.
   The sort's workspace is a static device array of about 20 MB, one per
   GPU, kept until the process ends and shared by every call on that GPU:
   calls that overlap on one GPU race.  The first call on a device looks
   up that device's copy (make it outside a graph capture).  Limits:
   n <= ARM_STATIC_MAX_PARTICLES and nx*ny <= ARM_STATIC_MAX_CELLS
   (ARM_ERR_CAPACITY beyond them, before any work); 3D: ARM_ERR_UNSUPPORTED. */
#define ARM_STATIC_MAX_PARTICLES (1 << 20)
#define ARM_STATIC_MAX_CELLS (1 << 20)
int arm_link_list_static(const arm_params* prm, int n, const double* x, int32_t* key_sorted,
                         int32_t* id_sorted, int32_t* cell_start, int32_t* outside,
                         void* stream);

/* Select the gather implementation: 1 (default) LDS-tiled segments of cell
   rows (pieces of 128 sorted particles of a row, or runs of cells), 0 direct
   per-particle gathers from global memory.  Both visit neighbours
   in the same order with the same arithmetic and give identical results.
   In 3D the direct gathers are the only ones (the switch has no effect). */
int arm_set_gather_path(arm_ctx* ctx, int tiled);

#ifdef __cplusplus
}
#endif
#endif
