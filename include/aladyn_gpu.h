/*
 * aladyn_gpu.h -- C-ABI of the MI355X (gfx950) implementation of ALaDyn's
 * laser-plasma PIC step (lpf2_evolve).
 *
 * The reference has no plugin API: the hot path is plain Fortran module calls
 * inside lpf2_evolve (src/dynamics/pic_evolve_in_time.f90:43-107).  Each entry
 * point below replaces one of those calls (or the whole step) and is bound
 * from Fortran with bind(C) interfaces (aladyn_amd/fortran/aladyn_gpu.f90).
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero ALADYN_E* code
 *    otherwise; aladyn_last_error() returns the message.  The Fortran shim
 *    copies a nonzero code into `ier` so the reference's exit path
 *    (ALaDyn.F90:99-102) fires.
 *  - Host arrays are borrowed only during a call and use the reference's own
 *    Fortran layouts (column-major):
 *      part(ld, 2*curr_ndim+1)    x,y[,z],px,py[,pz],W       (struct_def.f90:26-28)
 *      ebf(n1p, n2p, n3p, nfield) Ex,Ey,[Ez,Bx,By,]Bz        (fstruct_data.f90:27)
 *      jc(n1p, n2p, n3p, curr_ndim)                          (fstruct_data.f90:28)
 *    with n1p = nx+5, n2p = ny+5, n3p = nz+5 (1 in 2-D) (array_alloc.f90:56-80)
 *    and the interior at Fortran indices ix1 = 3 .. ix2 = nx+2 (C index 2..).
 *    W is the packed FP64 slot of precision_def.F90:33-40 (int16 charge,
 *    int16 part_ind, FP32 weight).
 *  - Device buffers belong to the context.  All work is queued on the
 *    context's HIP stream; get_* and the reductions synchronise it.
 *  - One host thread per context.  Contexts are independent.
 */
#ifndef ALADYN_GPU_H
#define ALADYN_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALADYN_OK 0
#define ALADYN_EINVAL 1   /* bad argument or unsupported configuration */
#define ALADYN_EHIP 2     /* HIP runtime error */
#define ALADYN_ENOMEM 3   /* device allocation failed */
#define ALADYN_ESTATE 4   /* call out of order (e.g. step before put) */
#define ALADYN_ERANGE 5   /* particle outside the box at put_species */

#define ALADYN_PREC_F64 64 /* tile-sorted 64-byte records, FP64 grids and push */
#define ALADYN_PREC_F32 32 /* tile-sorted 32-byte records (int32 node + FP32
                              offset), FP32 grids, FP64 current accumulation */

#define ALADYN_MAX_SPECIES 4 /* spec(4) in pstruct_data.f90:39 */

/* Run constants of one single-rank sub-domain, named as in common_param.f90
 * and grid_param.f90.  The derived quantities the reference computes in
 * set_grid / set_initial_param / set_field_param (indices, inverses, djc,
 * se_coeff, box limits) are computed by aladyn_create from these. */
typedef struct aladyn_config {
  int32_t ndim;          /* 2 or 3 (2-D is P-polarised: nfield 3, curr_ndim 2) */
  int32_t nx, ny, nz;    /* interior cells (nz ignored in 2-D) */
  int32_t ibx, iby, ibz; /* boundary types; 0 (open) is supported */
  int32_t der_ord;       /* 2, 3 or 4 */
  int32_t precision;     /* ALADYN_PREC_F64 or ALADYN_PREC_F32 */
  int32_t nsp;           /* species run per step (nsp_run), 1..4 */
  double dx, dy, dz;     /* cell sizes (dz ignored in 2-D) */
  double dt;             /* dt_loc */
  double cfl;            /* only used by der_ord 3 */
  double ompe;           /* 4*pi*rc0*1e-3 = 0.035411 (set_init_param.f90:254-255) */
  double lorentz_fact[ALADYN_MAX_SPECIES]; /* 1/m per species */
  /* tile path tuning (0 = default) */
  int32_t tile[3];       /* tile extent in nodes per axis */
  int32_t block;         /* threads per tile workgroup: 0 = one per node-cell;
                            twice the node-cells is compiled for 8x8x4 and
                            16x16 tiles (FP32 only in 3-D) */
  int32_t sort_every;    /* steps between full re-layouts of the tile store */
  /* reserved[0] != 0 keeps each particle's original index, so that
   * get_species returns particles in put/inject order (costs 4 B/particle of
   * HBM traffic per step); the rest must be 0 */
  int32_t reserved[5];
} aladyn_config;

typedef struct aladyn_ctx aladyn_ctx;

/* Creates a context on the current HIP device.  Replaces the allocation of
 * ebf/jc/ebfp in v_alloc (array_alloc.f90:56-125) for the device copy. */
int aladyn_create(const aladyn_config *cfg, aladyn_ctx **out);
int aladyn_destroy(aladyn_ctx *ctx);
const char *aladyn_last_error(const aladyn_ctx *ctx);

/* Copies species ic (0-based) in: the counterpart of filling spec(ic)%part
 * (part_distribute, init_part_distrib.f90:2475-2553).  Positions must lie in
 * the particle box (loc_*grid%gmin, gmax] (ALADYN_ERANGE otherwise). */
int aladyn_put_species(aladyn_ctx *ctx, int32_t ic, const double *part,
                       int64_t ld, int64_t np);
/* Number of particles of species ic (loc_npart, mpi_var.F90:35). */
int aladyn_species_count(aladyn_ctx *ctx, int32_t ic, int64_t *np);
/* Copies species ic back into part(ld, 2*curr_ndim+1); *np gets the count
 * (which shrinks when particles leave the box, cell_part_dist,
 * mpi_part_interface.f90:525-753).  ld must be >= the count. */
int aladyn_get_species(aladyn_ctx *ctx, int32_t ic, double *part, int64_t ld,
                       int64_t *np);
/* Field and current transfer in the ebf / jc layouts above. */
int aladyn_put_fields(aladyn_ctx *ctx, const double *ebf);
int aladyn_get_fields(aladyn_ctx *ctx, double *ebf);
int aladyn_get_current(aladyn_ctx *ctx, double *jc);

/* One lpf2_evolve(iter) (pic_evolve_in_time.f90:37-109) with part = .true.:
 * pfields_prepare, jc = 0, per species {set_lpf_acc, field_charge_multiply,
 * [init_lpf_momenta], lpf_momenta_and_positions, cell_part_dist,
 * curr_accumulate}, curr_mpi_collect, advance_lpf_fields.  initial_time != 0
 * applies the first-step momentum back-shift (ALaDyn.F90:67). */
int aladyn_lpf2_evolve(aladyn_ctx *ctx, int32_t iter, int32_t initial_time);
/* Changes the time step of the following steps: Lp_cycle's trim of the last
 * step, dt_loc = tmax - tnow when tnow + dt_loc >= tmax (ALaDyn.F90:101).
 * 0 < dt < every cell size (ALADYN_EINVAL otherwise). */
int aladyn_set_dt(aladyn_ctx *ctx, double dt);
/* No counterpart in the reference (its steps are Fortran calls).  For
 * callers whose steps are short enough that launching them costs as much
 * as running them (many small runs stepped in turn): with enable != 0,
 * aladyn_lpf2_evolve replays the step's device work as a HIP graph,
 * captured once per state of the particle stores and of the box (a
 * re-layout, a window move or a put makes the next step capture again).
 * The first step (initial_time) and steps with diagnostics on run as usual.
 * Results are those of the ordinary path; aladyn_kernel_timing does not
 * time graph steps. */
int aladyn_set_graphs(aladyn_ctx *ctx, int32_t enable);

/* Many independent runs stepped together (density or a0 scans).  No
 * counterpart in the reference, which runs one simulation per process.  The
 * contexts (each put with one species and its fields, all of one precision,
 * dimension, tile shape and threads per tile, on one device) keep their own
 * grids and stores; aladyn_batch_lpf2_evolve advances every one of them by
 * one lpf2_evolve with each kernel of the step launched once for all runs:
 * the particle kernel over the work list of every run's tiles, the grid
 * kernels with one row of workgroups per run.  With use_graphs != 0 the
 * step's launches are captured once and replayed as a HIP graph.  Steps are
 * queued on the batch's own stream: call aladyn_batch_synchronize before
 * using a context alone again.  Destroying the batch leaves the contexts. */
typedef struct aladyn_batch aladyn_batch;
int aladyn_batch_create(aladyn_ctx *const *ctxs, int32_t n, int32_t use_graphs,
                        aladyn_batch **out);
int aladyn_batch_lpf2_evolve(aladyn_batch *batch, int32_t iter, int32_t initial_time);
int aladyn_batch_synchronize(aladyn_batch *batch);
int aladyn_batch_destroy(aladyn_batch *batch);
const char *aladyn_batch_last_error(const aladyn_batch *batch);

/* The stages of the step, for drivers that interleave their own work:
 * pfields_prepare(ebf, nfield, 2, 2) (curr_and_fields_util.f90:156-174) */
int aladyn_pfields_prepare(aladyn_ctx *ctx);
/* jc(:,:,:,:) = 0 (pic_evolve_in_time.f90:73) */
int aladyn_curr_clean(aladyn_ctx *ctx);
/* Species loop body for species ic (pic_evolve_in_time.f90:77-89): gather
 * (set_lpf_acc), field_charge_multiply, init_lpf_momenta if initial_time,
 * lpf_momenta_and_positions, cell_part_dist(.false.), curr_accumulate
 * (Esirkepov, iform 0).  Fused into one particle kernel. */
int aladyn_particle_stage(aladyn_ctx *ctx, int32_t ic, int32_t initial_time);
/* curr_mpi_collect(jc) on one rank (curr_and_fields_util.f90:104-152). */
int aladyn_curr_mpi_collect(aladyn_ctx *ctx);
/* advance_lpf_fields(ebf, jc, 0) (curr_and_fields_util.f90:176-260). */
int aladyn_advance_lpf_fields(aladyn_ctx *ctx);

/* lp_window_xshift(witr, init_iter) (window.f90:276-347) with w_speed, on
 * the device copy.  init_iter == 0 resets the saved window state (xlapse,
 * wi1) and does nothing else (window.f90:284-288).  Otherwise
 * xlapse += w_speed*dt*witr and the box moves by
 * *nshx = nint(xlapse/dx) - wi1 cells: the field interior shifts left by nshx
 * with zeros entering on the right (fields_left_xshift,
 * mpi_field_interface.f90:299-359), the box limits move by dx*nshx, and
 * particles left behind the new left face are removed
 * (cell_part_dist(.true.), mpi_part_interface.f90:525-596).  The caller then
 * adds the plasma entering on the right with aladyn_inject_species, as
 * particles_inject does from the caller's own target tables
 * (window.f90:113-270).  ALADYN_ERANGE if nshx reaches the box length (the
 * reference's ier = 2). */
int aladyn_lp_window_xshift(aladyn_ctx *ctx, int32_t witr, int32_t init_iter,
                            double w_speed, int32_t *nshx);
/* Appends np particles part(ld, 2*curr_ndim+1) to species ic, in the current
 * box frame (add_particles, window.f90:39-111); positions must lie in the
 * box.  They are moved from the next step on. */
int aladyn_inject_species(aladyn_ctx *ctx, int32_t ic, const double *part,
                          int64_t ld, int64_t np);
/* Registers the target table of species ic, from which each
 * aladyn_lp_window_xshift call injects the plasma entering on the right on
 * the device (particles_inject and add_particles, window.f90:39-139, cold
 * target: t0_pl = 0): xpt[0..nxp) the ascending x of the plasma columns
 * (xpt(:,ic)), ypt[0..nyp) and zpt[0..nzp) the transverse positions of the
 * box (loc_ypt, loc_zpt; zpt and nzp ignored in 2-D), wghpt[nxp] the column
 * weights and wghyz[nyp*nzp] the transverse weights (loc_wghyz, j fastest),
 * nptx the columns already in the box (nptx(ic)) and [targ_in, targ_end] the
 * target's x extent.  After a move, while targ_in <= xmax < targ_end,
 * columns nptx+1 .. i2 are appended with the reference's selection (1-based:
 * the scan stops at the first xpt > xmax, i2 is the column before it, or
 * that column itself when it is the last; nothing is added unless i2 > i1;
 * nptx becomes i2), in the reference's order (column fastest, then y, then
 * z), with p = 0, weight real(wghpt*wghyz, sp) and the species' charge and
 * part_ind.  Columns outside (xmin, xmax] are consumed but not added (the
 * reference would remove them before their first deposit).  nxp = 0 clears
 * the table.  The species must have been put first (ALADYN_ESTATE at the
 * move otherwise); ypt/zpt outside the box give ALADYN_ERANGE. */
int aladyn_set_injection(aladyn_ctx *ctx, int32_t ic, const double *xpt, int64_t nxp,
                         const double *ypt, int32_t nyp, const double *zpt, int32_t nzp,
                         const double *wghpt, const double *wghyz, int64_t nptx,
                         double targ_in, double targ_end);
/* The columns of species ic's target table consumed so far (nptx(ic)). */
int aladyn_injection_state(aladyn_ctx *ctx, int32_t ic, int64_t *nptx);
/* Places the box at x in (xmin, xmin + nx*dx], as restart does with the
 * dumped x(1) (pic_dump.f90:520-531) after the window has moved; the saved
 * window state stays zero, as the reference's does on restart.  Only before
 * any put_species (ALADYN_ESTATE otherwise): stored particles are relative
 * to the box. */
int aladyn_set_box_origin(aladyn_ctx *ctx, double xmin);
/* The current particle box: out = {xmin, xmax, ymin, ymax, zmin, zmax}
 * (loc_*grid%gmin/gmax; x moves with the window). */
int aladyn_box(aladyn_ctx *ctx, double out[6]);

/* The envelope (ponderomotive) model, env_lpf2_evolve
 * (env_evolve_in_time.f90:196-361), for species 0 on one rank: one laser
 * colour, no ionization, no fluid.  put_envelope copies in
 * env(n1p, n2p, n3p, 4) = A_R^n, A_I^n, A_R^{n-1}, A_I^{n-1} (the envelope
 * array of v_alloc, array_alloc.f90:108-111) and the laser frequency oml
 * (set_init_param.f90:291, 2 pi/lam0); get_envelope copies it out. */
int aladyn_put_envelope(aladyn_ctx *ctx, double oml, const double *env);
int aladyn_get_envelope(aladyn_ctx *ctx, double *env);
/* One env_lpf2_evolve(iter): |A^n|^2/2 and its gradient (env_amp_prepare);
 * per particle of species 0, E, B, |A|^2/2 and its gradient at x^n
 * (set_env_acc), the envelope pusher's momenta (lpf_env_momenta) and the
 * density source (set_env_density); the envelope advance A^n -> A^{n+1}
 * (env_den_collect, advance_lpf_envelope with env_maxw_solve); |A|^2/2 at
 * t^{n+1/2} (env_fields_average); per particle, its gradient at x^n
 * (set_env_grad_interp), the positions (lpf_env_positions), removal at the
 * faces (cell_part_dist), the Esirkepov current; curr_mpi_collect and
 * advance_lpf_fields.  Needs put_fields and put_envelope (ALADYN_ESTATE). */
int aladyn_env_lpf2_evolve(aladyn_ctx *ctx, int32_t iter);

/* energy_momenta for species ic (diag_part_and_fields.f90:172-219), FP64
 * device reduction: ek[0..6] = ek(1:7), *ekmax = max(gamma-1). */
int aladyn_energy_momenta(aladyn_ctx *ctx, int32_t ic, double ek[7],
                          double *ekmax);
/* Field energy per component as in Envar (diag_part_and_fields.f90:499-670,
 * unstretched): out[c] = dgvol * sum_interior ebf(c)^2, FP64 reduction. */
int aladyn_field_energy(aladyn_ctx *ctx, double out[6]);
/* set_grid_charge for species ic (grid_part_util.f90:240-335), accumulated in
 * FP64 on the device and copied into rho(n1p, n2p, n3p) (added to it). */
int aladyn_grid_charge(aladyn_ctx *ctx, int32_t ic, double *rho);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same as
 * aladyn_grid_charge, but its device staging of rho is a fixed array of
 * 2^23 FP64 values per GPU held by the library (64 MiB, set up at
 * aladyn_create), so the call allocates nothing (aladyn_grid_charge
 * allocates its diagnostic buffers on first use).  Grids of more than 2^23
 * nodes (n1p*n2p*n3p) give ALADYN_EINVAL.  Calls that overlap on one GPU
 * share that array and race. */
int aladyn_grid_charge_noalloc(aladyn_ctx *ctx, int32_t ic, double *rho);
/* The bunch fields at the start of a run, added to the context's fields
 * (init_beam_part_distrib.f90:538-602, uniform grid, one rank; replaces its
 * set_grid_charge + fft_3d_psolv / fft_2d_psolv + initial_beam_fields +
 * ebf += ebf_bunch): the charge density of species ic (the bunch), its
 * potential by the sine-transform Poisson solve (psolve.f90:117-247, ft_mod
 * = 2, sym = 1), then (Ex, Ey, Ez, By, Bz) of initial_beam_fields
 * (init_grid_fields.f90:32-131), in FP64.  gam2 = gamma^2 and bet0 = beta of
 * the bunch.  Allocates and frees its work arrays, waits for the stream. */
int aladyn_beam_fields(aladyn_ctx *ctx, int32_t ic, double gam2, double bet0);
/* Discrete continuity residual of the last particle_stage on the device in
 * FP64: r = (rho^{n+1} - rho^n) + (1/curr_ndim) sum_d D_d J_d(raw) on the
 * interior, with rho from the particle positions before and after the stage
 * and J the raw (pre-djc) current of that stage.  out[0] = max|r|,
 * out[1] = max|rho^{n+1} - rho^n|.  Requires aladyn_set_diagnostics(ctx, 1)
 * before the stage (it keeps rho^n and the raw current).  The raw current is
 * kept before curr_clean or curr_mpi_collect change jc, so the residual of
 * the last stage of an lpf2_evolve step can be read after the step; the next
 * particle_stage, set_diagnostics or grid_charge call replaces it. */
int aladyn_continuity_residual(aladyn_ctx *ctx, double out[2]);
int aladyn_set_diagnostics(aladyn_ctx *ctx, int32_t enable);

/* Waits for all queued work of the context. */
int aladyn_synchronize(aladyn_ctx *ctx);
/* Kernel/launch bookkeeping for measurement: average duration (ms) of the
 * particle kernel over the steps since the last reset, measured with HIP
 * events on the context stream. */
int aladyn_kernel_timing(aladyn_ctx *ctx, int32_t reset, double *particle_ms,
                         double *field_ms, int64_t *launches);
/* The same for the envelope step's two particle kernels: the gather + push +
 * density deposit (push_ms) and the gradient gather + position update +
 * Esirkepov deposit + re-binning (move_ms). */
int aladyn_env_kernel_timing(aladyn_ctx *ctx, int32_t reset, double *push_ms, double *move_ms,
                             int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
