/*
 * snap_hip.h -- C ABI of the MI355X (gfx950) mini-KBA solve library.
 *
 * Drop-in boundary for the GPU task variants of Legion-SNAP's steady-state
 * solve.  Each snap_hip_run_* entry point replaces one `extern` host function
 * that the reference's GPU task bodies call (cited per declaration), with the
 * same argument meaning.  Legion types become plain structs:
 *
 *   Realm::AffineAccessor<T,N>  ->  snap_accessorN_t { base, stride[N] }:
 *       the element of point p is at  (char*)base + sum_i p[i]*stride[i]
 *       (byte strides, absolute points, exactly Realm's affine form);
 *   std::vector<Accessor>       ->  pointer + count (one accessor per group);
 *   Rect<3>                     ->  snap_rect3_t (inclusive lo/hi);
 *   DeferredValue<bool>         ->  int32_t* in device memory (0 or 1);
 *   the Realm task stream       ->  snap_stream_t (a hipStream_t; NULL is
 *                                   the null stream).
 *
 * Record types: MomentQuad is a record of 4 doubles and MomentTriple of 3
 * (snap_types.h:30-50).  For cmom = nmom^2 > 4 the records widen: qtot holds
 * cmom doubles, fluxm / q2grpm hold cmom-1, slgg / s_xs hold nmom (one per
 * Legendre order).  Elements inside a record are contiguous doubles.
 *
 * Errors: every entry point returns SNAP_OK (0) or a negative code; nothing
 * calls exit() (the reference asserts or exits, gpu_sweep.cu:84-88,
 * gpu_outer.cu:218-220).  snap_hip_last_error() gives the message.
 * Launches are asynchronous on the given stream.
 */
#ifndef SNAP_HIP_H
#define SNAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNAP_OK 0
#define SNAP_ERR_INVALID (-1)      /* bad argument */
#define SNAP_ERR_UNSUPPORTED (-2)  /* valid in the reference, not built here */
#define SNAP_ERR_HIP (-3)          /* a HIP runtime call failed */
#define SNAP_ERR_NO_CONTEXT (-4)   /* initialize_gpu_context not called on this device */
#define SNAP_ERR_NO_MEMORY (-5)    /* device allocation failed */

#define SNAP_PRECISION_FP64 0        /* FP64 sweep with fused multiply-adds */
#define SNAP_PRECISION_MIXED 1       /* FP32 sweep, FP64 moments/source/convergence */
#define SNAP_PRECISION_FP64_EXACT 2  /* FP64 sweep in the reference CPU operation order,
                                        every multiply and add rounded separately */

typedef void *snap_stream_t;

typedef struct { void *base; int64_t stride[1]; } snap_accessor1_t;
typedef struct { void *base; int64_t stride[2]; } snap_accessor2_t;
typedef struct { void *base; int64_t stride[3]; } snap_accessor3_t;
typedef struct { int64_t lo[3]; int64_t hi[3]; } snap_rect3_t;

/* ---------------------------------------------------------------------- */
/* Library                                                                 */
/* ---------------------------------------------------------------------- */

/* Message of the last error returned on the calling thread. */
const char *snap_hip_last_error(void);
/* Library version string. */
const char *snap_hip_version(void);
/* The angular quadrature of Snap::compute_derived_globals (snap.cc:1171-1342,
 * 3-D): mu, eta, xi, w [num_angles] and ec [8][cmom][num_angles], cmom =
 * num_moments^2, host arrays.  What a caller passes to
 * snap_hip_initialize_gpu_context. */
int snap_hip_quadrature(int num_angles, int num_moments, double *mu, double *eta, double *xi,
                        double *w, double *ec);

/* ---------------------------------------------------------------------- */
/* Drop-in task entry points                                               */
/* ---------------------------------------------------------------------- */

/* Replaces initialize_gpu_context (decl init.cc:215-220, def
 * gpu_sweep.cu:71-273), called by InitGPUSweep::gpu_implementation
 * (init.cc:224-236).  Uploads the quadrature to the current device and sizes
 * the per-stream sweep workspace for one chunk.  ec is [num_octants][cmom]
 * [num_angles] with cmom = num_moments^2 (octant stride cmom*num_angles).
 * Calling it again on the same device replaces the context. */
int snap_hip_initialize_gpu_context(const double *ec, const double *mu,
                                    const double *eta, const double *xi,
                                    const double *w, int num_angles,
                                    int num_moments, int num_octants,
                                    int nx_per_chunk, int ny_per_chunk,
                                    int nz_per_chunk);

/* Arguments of run_gpu_sweep (decl sweep.cc:1524-1546, def
 * gpu_sweep.cu:2046-2295), in the same order. */
typedef struct {
  int64_t origin[3];              /* chunk lower bound (dom.bounds.lo) */
  snap_accessor3_t qtot;          /* RO  cmom doubles per cell */
  snap_accessor3_t flux;          /* RW  1 double per cell; the sweep ADDS its contribution */
  snap_accessor3_t fluxm;         /* RW  cmom-1 doubles per cell (base NULL when cmom == 1) */
  snap_accessor3_t dinv;          /* RO  num_angles doubles per cell */
  snap_accessor3_t time_flux_in;  /* RO  num_angles doubles per cell; read when vdelt != 0 */
  snap_accessor3_t time_flux_out; /* WO  num_angles doubles per cell; written when vdelt != 0 */
  snap_accessor3_t t_xs;          /* RO  1 double per cell */
  snap_accessor2_t ghostx;        /* RW  num_angles doubles at point (y,z); base NULL = vacuum */
  snap_accessor2_t ghosty;        /* RW  num_angles doubles at point (x,z); base NULL = vacuum */
  snap_accessor2_t ghostz;        /* RW  num_angles doubles at point (x,y); base NULL = vacuum */
  snap_accessor3_t qim;           /* RO  MMS angular source, num_angles doubles per cell (mms_source = 1) */
  int x_range, y_range, z_range;
  int corner;                     /* octant: bit0 +x, bit1 +y, bit2 +z */
  int stride_x_positive, stride_y_positive, stride_z_positive;
  int mms_source;                 /* 1: add qim to the angular source (sweep.cc:347-351) */
  int num_moments;                /* nmom; cmom = nmom^2 */
  double hi, hj, hk;              /* 2/dx, 2/dy, 2/dz */
  double vdelt;                   /* 2/(dt v_g) of this group; 0.0 = steady state */
  int num_angles;
  int fixup;                      /* negative-flux fixup on/off */
} snap_sweep_args_t;

/* Replaces run_gpu_sweep: one (corner, group) diamond-difference sweep over
 * one chunk, reading the upstream ghost faces and writing the downstream
 * ones, adding the octant's scalar flux and flux moments into flux / fluxm
 * with device atomics (the reference's relaxed-coherence
 * SIMULTANEOUS read-write, sweep.cc:54-55, gpu_sweep.cu:1705). */
int snap_hip_run_gpu_sweep(const snap_sweep_args_t *args, snap_stream_t stream);

/* One MiniKBA task's group range in one launch series (MiniKBAArgs
 * {corner, group_start, group_stop}, sweep.h:29-37; the task body's loop that
 * calls run_gpu_sweep once per group, sweep.cc:1581-1617).  groups[i] holds
 * run_gpu_sweep's arguments for group group_start + i; all share the chunk,
 * corner and controls, and each field's accessors must have equal strides
 * and equally spaced bases over the groups (the group fields of one Legion
 * SOA instance are) -- else SNAP_ERR_UNSUPPORTED, and the caller sweeps group
 * by group.  Steady sweeps only (vdelt = 0, no MMS source).  precision:
 * SNAP_PRECISION_FP64_EXACT (results bit-equal to one run_gpu_sweep per
 * group) or SNAP_PRECISION_FP64 (fused multiply-adds; up to 64 angles on the
 * one-wavefront tile kernel, which forms dinv from t_xs and does not read the
 * dinv accessor).  All groups' sweeps of the range share every launch, as
 * the solver's do. */
int snap_hip_run_gpu_sweep_groups(const snap_sweep_args_t *groups, int num_groups, int precision,
                                  snap_stream_t stream);

/* snap_hip_run_gpu_sweep without any device allocation: the same kernels and
 * results, with the plan tables and tile-face workspace in a fixed scratch
 * array of 64 MiB that each device holds until the process ends (chunks whose
 * faces and tables exceed it get SNAP_ERR_UNSUPPORTED).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  initialize_gpu_context must have
 * run on the device.  Every call on one device shares its scratch array:
 * calls that overlap on one device (other streams, other threads) race;
 * calls on one stream are ordered.  The array needs no clearing: each call
 * writes its tables and every face slot before reading it. */
int snap_hip_run_gpu_sweep_noalloc(const snap_sweep_args_t *args, snap_stream_t stream);

/* Replaces run_flux0_outer_source (decl outer.cc:220-225, def
 * gpu_outer.cu:104-223): qo0[g] = qi0[g] + sum_{g'!=g} slgg[g][mat][g'].l0 * flux0[g'].
 * slgg[g] is a 2-D accessor at point (material, g') of nmom doubles. */
int snap_hip_run_flux0_outer_source(snap_rect3_t subgrid_bounds,
                                    const snap_accessor3_t *fa_qi0,
                                    const snap_accessor3_t *fa_flux0,
                                    const snap_accessor2_t *fa_slgg,
                                    const snap_accessor3_t *fa_qo0,
                                    snap_accessor3_t fa_mat, int num_groups,
                                    snap_stream_t stream);

/* Replaces run_fluxm_outer_source (decl outer.cc:226-231, def
 * gpu_outer.cu:320-441): qom[g][m] = sum_{g'!=g} slgg[g][mat][g'].L(m) * fluxm[g'][m]
 * for every flux moment m = 1..cmom-1 (L(m) its Legendre order from lma). */
int snap_hip_run_fluxm_outer_source(snap_rect3_t subgrid_bounds,
                                    const snap_accessor3_t *fa_fluxm,
                                    const snap_accessor2_t *fa_slgg,
                                    const snap_accessor3_t *fa_qom,
                                    snap_accessor3_t fa_mat, int num_groups,
                                    int num_moments, const int lma[4],
                                    snap_stream_t stream);

/* Replaces run_inner_source_single_moment (decl inner.cc:163-167, def
 * gpu_inner.cu:44-60): qtot.l0 = q2grp0 + flux0 * s_xs.l0 (one group). */
int snap_hip_run_inner_source_single_moment(snap_rect3_t subgrid_bounds,
                                            snap_accessor3_t fa_sxs,
                                            snap_accessor3_t fa_flux0,
                                            snap_accessor3_t fa_q2grp0,
                                            snap_accessor3_t fa_qtot,
                                            snap_stream_t stream);

/* Replaces run_inner_source_multi_moment (decl inner.cc:168-175, def
 * gpu_inner.cu:97-117): adds qtot[m] = q2grpm[m-1] + fluxm[m-1]*s_xs.L(m). */
int snap_hip_run_inner_source_multi_moment(snap_rect3_t subgrid_bounds,
                                           snap_accessor3_t fa_sxs,
                                           snap_accessor3_t fa_flux0,
                                           snap_accessor3_t fa_q2grp0,
                                           snap_accessor3_t fa_fluxm,
                                           snap_accessor3_t fa_q2grpm,
                                           snap_accessor3_t fa_qtot,
                                           int num_moments, const int lma[4],
                                           snap_stream_t stream);

/* Replaces run_inner_convergence (decl inner.cc:327-331, def
 * gpu_inner.cu:206-236): *result = 1 when every |flux0/flux0pi - 1| <= epsi
 * (|flux0| when |flux0pi| < 1e-12), else 0.  result is device memory. */
int snap_hip_run_inner_convergence(snap_rect3_t subgrid_bounds, int32_t *result,
                                   const snap_accessor3_t *fa_flux0,
                                   const snap_accessor3_t *fa_flux0pi,
                                   int num_groups, double epsi,
                                   snap_stream_t stream);

/* Replaces run_outer_convergence (decl outer.cc:406-410, def
 * gpu_outer.cu:533-563); the caller passes epsi = 100*convergence_eps as
 * TestOuterConvergence::gpu_implementation does (outer.cc:439). */
int snap_hip_run_outer_convergence(snap_rect3_t subgrid_bounds, int32_t *result,
                                   const snap_accessor3_t *fa_flux0,
                                   const snap_accessor3_t *fa_flux0po,
                                   int num_groups, double epsi,
                                   snap_stream_t stream);

/* Replaces run_expand_cross_section (decl expxs.cc:118-122, def
 * gpu_expxs.cu:46-266): xs[g][p] = sig[g][mat[p]]. */
int snap_hip_run_expand_cross_section(const snap_accessor1_t *sig_ptrs,
                                      snap_accessor3_t fa_mat,
                                      const snap_accessor3_t *fa_xs,
                                      int num_groups, snap_rect3_t subgrid_bounds,
                                      snap_stream_t stream);

/* Replaces run_expand_scattering_cross_section (decl expxs.cc:253-258, def
 * gpu_expxs.cu:289-531): xs[g][p] = slgg[g][mat[p]][group_start+g], a record
 * of record_doubles doubles (4 for MomentQuad). */
int snap_hip_run_expand_scattering_cross_section(const snap_accessor2_t *fa_slgg,
                                                 snap_accessor3_t fa_mat,
                                                 const snap_accessor3_t *fa_xs,
                                                 int num_groups,
                                                 snap_rect3_t subgrid_bounds,
                                                 int group_start, int record_doubles,
                                                 snap_stream_t stream);

/* Replaces run_geometry_param (decl expxs.cc:395-399, def
 * gpu_sweep.cu:313-606): dinv[g][p][a] = 1/(xs + vdelt[g] + hi*mu[a] +
 * hj*eta[a] + hk*xi[a]) with the angles of the current context.  vdelts is a
 * host array of num_groups values. */
int snap_hip_run_geometry_param(const snap_accessor3_t *fa_xs,
                                const snap_accessor3_t *fa_dinv,
                                const double *vdelts, int num_groups, double hi,
                                double hj, double hk, snap_rect3_t subgrid_bounds,
                                int num_angles, snap_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Legion-free solver: the transport_solve loop (snap.cc:225-568)          */
/* ---------------------------------------------------------------------- */

typedef struct {
  int nx, ny, nz;
  double lx, ly, lz;
  int ng, nang, nmom, nmat;
  const int32_t *mat;       /* host [nz][ny][nx], 0-based material index */
  const double *sigt;       /* host [nmat][ng] */
  const double *slgg;       /* host [ng_to][nmat][ng_from][nmom] */
  const double *qi;         /* host [ng][nz][ny][nx] */
  const double *sigt_cell;  /* host [ng][nz][ny][nx] or NULL: per-cell total xs */
  double epsi;
  int iitm, oitm;
  int fixup;
  /* time dependence (Snap::initialize_velocity snap.cc:716-738, the time-step
   * loop snap.cc:400-541): nsteps steps of dt = tf / nsteps with
   * vdelt_g = 2 / (dt v_g), v_g = ng - g; timedep = 0 is one steady step */
  int timedep;
  double tf;
  int nsteps;               /* >= 1 (0 counts as 1) */
  /* manufactured-solution source (src_opt 3, mms.cc): the solver builds the
   * fixed and angular sources itself and qi may be NULL (nmom = 1 and
   * material cross sections only) */
  int mms;
} snap_problem_t;

#define SNAP_DINV_DEFAULT 0     /* the one-wavefront tile sweep forms dinv from t_xs;
                                   the multi-wave tile sweeps stream it (MIXED reads the
                                   FP64-formed dinv rounded once to FP32) */
#define SNAP_DINV_STREAM 1      /* sweep reads the dinv array (reference data flow) */
#define SNAP_DINV_RECOMPUTE 2   /* sweep forms dinv from t_xs per cell and angle */

typedef struct {
  int precision;            /* SNAP_PRECISION_FP64, _FP64_EXACT or _MIXED */
  int tile[3];              /* sweep tile edges (x,y,z); 0 = choose */
  int waves_per_tile;       /* waves per tile workgroup; 0 = choose */
  int first_sweep_fp64;     /* MIXED: run the first inner's sweep in FP64 */
  int dinv_mode;            /* SNAP_DINV_DEFAULT, _STREAM or _RECOMPUTE */
  int angles_per_lane;      /* sweep angles per GPU lane, 1 or 2; 0 = choose */
  int refine;               /* MIXED: 1 = defect correction (below); 0 = FP32
                               sweeps of the whole source */
  int outer_ordered;        /* 1 = outer source in the CPU's summation order even
                               when not FP64_EXACT (default: FP64 matrix cores) */
  int wide_sweep;           /* 1 = one workgroup per cell and sweep (the only form past
                               128 angles per octant; up to 2048); 0 = tiled when it fits */
  int sweep_kernel;         /* SNAP_SWEEP_AUTO, _TILED (multi-wave workgroups per tile)
                               or _WAVE_TILE (one wavefront per 4x4xTZ tile, steady
                               problems of up to 64 angles; tile[2] sets TZ, default 8).
                               AUTO takes the wave tile for steady problems of 33..64
                               angles unless a multi-wave option (tile x/y, waves, angles
                               per lane) is set */
  double fp64_polish;       /* MIXED: run an inner's sweep in FP64 when the previous inner
                               of the same outer changed the scalar flux by at most
                               fp64_polish * epsi (max relative change), or by at most
                               1e-5 (FP32 sweeps' own rounding level) whatever epsi
                               asks; 0 = never */
  int groups_per_wave;      /* one-wavefront tile sweeps: 0 = choose (two groups per
                               wavefront, lanes 0-31 and 32-63, for fused sweeps of up
                               to 32 angles with an even group count and x and y
                               extents multiples of 4; bit-equal results), 1 = one */
} snap_solve_options_t;

enum { SNAP_SWEEP_AUTO = 0, SNAP_SWEEP_TILED = 1, SNAP_SWEEP_WAVE_TILE = 2 };

/* MIXED with refine = 1 (iterative refinement of each outer iteration): the
 * first inner sweep of an outer iteration runs in FP64 with the fixup and is
 * kept as the base (source q0, fluxes phi0).  Each later inner computes the
 * FP64 defect q - q0 of its source, sweeps it in FP32 without the fixup (the
 * sweep is linear in its source when no flux is fixed) and adds the result
 * to phi0 in FP64. */

typedef struct {
  int inner_total;          /* inner iterations run */
  int outer_total;          /* outer iterations run */
  int converged;            /* outer convergence reached before oitm */
  int sweeps_fp64;          /* inner sweeps done in FP64 */
  int sweeps_fp32;          /* inner sweeps done in FP32 (MIXED) */
  double sweep_ms;          /* GPU time in sweep kernels, all inners */
  double solve_ms;          /* GPU time of the whole loop */
  double units_per_sweep;   /* cells * 8*nang * ng */
  double grind_ns;          /* sweep_ms*1e6 / (units_per_sweep * inner_total) */
  int time_steps;           /* time steps run; converged = every step converged */
  int groups_per_wave;      /* groups per wavefront of the inner sweeps (2: see
                               snap_solve_options_t.groups_per_wave) */
} snap_solve_stats_t;

typedef struct snap_solver snap_solver_t;

/* Upload the problem, expand cross sections and geometry (snap.cc:419-450). */
int snap_hip_solver_create(const snap_problem_t *problem,
                           const snap_solve_options_t *options,
                           snap_solver_t **solver);
/* Run the outer/inner loop from a zero flux to convergence or oitm, once per
 * time step (the flux carries over between steps; the time-edge fluxes start
 * at zero and alternate between two arrays, snap.cc:288-298, 497-498). */
int snap_hip_solver_run(snap_solver_t *solver, snap_solve_stats_t *stats);
/* MMSCompare (mms.cc:497-541, snap.cc:542-559) of the current flux against
 * the manufactured solution: result = {max, min, mean} of |phi/ref - 1|
 * (|phi| where ref < 1e-12) over all cells and groups.  The reference
 * divides its sum by nx*ichunk * ny*npey * nz*npez cells (snap.cc:546);
 * here the mean is over the ng*nx*ny*nz values.  Needs an mms problem. */
int snap_hip_solver_mms_compare(snap_solver_t *solver, double result[3]);
/* The manufactured scalar flux [ng][cells] the compare uses. */
int snap_hip_solver_get_mms_reference(snap_solver_t *solver, double *ref_flux);

/* One convergence test of the last run, in order: the events the reference's
 * ConvergenceMonad binds (convergence.cc:153-236).  kind 0 is an inner test
 * (after every inner iteration), kind 1 an outer test (every outer iteration
 * after the first, snap.cc:519-525).  t_us is the host time at which the
 * test's result was known, in microseconds from the start of the run. */
typedef struct {
  int kind;
  int time_step, outer, inner;   /* 0-based loop indices */
  int converged;
  double t_us;
} snap_iteration_record_t;
/* Copy up to `capacity` records of the last run; *count gets the total. */
int snap_hip_solver_get_log(snap_solver_t *solver, snap_iteration_record_t *records,
                            int capacity, int *count);

/* The reference mapper's tunables (Snap::SnapMapper::select_tunable_value,
 * mapper.cc:101-172) for a deck: out[0] outer runahead (2), out[1] inner
 * runahead (iitm), out[2] energy groups per sweep task, for `processors`
 * GPUs (or CPUs), out[3] SMs per sweep for a device with `sms` SMs.  The
 * solver issues all groups of an inner iteration as one job and breaks at
 * the first converged test, which the predicated runahead reproduces. */
int snap_hip_mapper_tunables(int num_groups, int max_inner_iters, int processors, int sms,
                             int32_t out[4]);
/* Copy the scalar flux [ng][cells] and flux moments [ng][cells][cmom-1]
 * (may be NULL) to host memory. */
int snap_hip_solver_get_flux(snap_solver_t *solver, double *flux0, double *fluxm);
/* Time `count` inner sweeps (all octants and groups) in the given precision
 * on the source left by the last run, after one untimed sweep.  Writes the
 * mean GPU time of one inner sweep, including the sum of its octant partials
 * into the flux moments (which the reference does inside its sweep kernel),
 * the mean duration of one launch of the sweep kernel and the launches per
 * inner sweep (one per tile diagonal and stream: the one-wavefront tile sweep
 * runs half of the corners on a second stream).  MIXED needs a solver
 * created with precision MIXED. */
int snap_hip_solver_time_sweeps(snap_solver_t *solver, int count, int precision,
                                double *ms_per_sweep, double *ms_per_launch,
                                int *launches_per_sweep);
/* Gradient of a detector response R = sum_g sum_c resp[g][c] * flux0[g][c]
 * at the solver's current (converged) flux with respect to the total cross
 * section t_xs[g][c], for nmom = 1.  No counterpart in the reference (the
 * adjoint-sweep pattern; callers that differentiate a solve).
 * precision: SNAP_PRECISION_FP64, or SNAP_PRECISION_MIXED for FP32 reverse
 * sweeps.  Runs one FP64 forward sweep of the current flux's source (its
 * tile faces are the checkpoint), then reverse sweeps of the adjoint source
 * iteration until max|d zeta| <= tol * max|zeta| (with FP32 reverse sweeps
 * also once the change, below 1e-5 of max|zeta|, stops shrinking: their
 * rounding floor); each reverse sweep
 * recomputes its tile's forward fluxes and fixup decisions in FP64 from the
 * checkpoint (an FP32 recompute flips decisions near zero), and runs the
 * reverse recurrence in FP64 or FP32.  resp and grad are
 * host arrays [ng][nz][ny][nx]; response (optional) receives R. */
int snap_hip_solver_gradient(snap_solver_t *solver, const double *resp, int precision,
                             double tol, int maxit, double *grad, double *response,
                             int *iterations);
/* The sweep plan the solver chose: plan[0..2] tile edges (x,y,z), plan[3]
 * waves per tile workgroup, plan[4] angles per lane, plan[5] cells per
 * wavefront (0 for the one-cell-per-workgroup sweep, whose tiles are 1x1x1;
 * 16 with plan[3] = 1 for the one-wavefront tile sweep, whose wavefront
 * holds a 4 x 4 plane of its tile). */
int snap_hip_solver_plan(snap_solver_t *solver, int32_t plan[6]);
void snap_hip_solver_destroy(snap_solver_t *solver);

/* ------------------------------------------------------------------------
 * Batches of independent problems (no counterpart in the reference; for
 * callers with many small fixed-source problems of uneven size).  The
 * problems must share ng, nang, nmom, epsi, iitm, oitm and fixup; meshes,
 * materials and sources may differ.  Each problem follows its own iteration
 * control exactly as snap_hip_solver_run does; the batch stops when every
 * problem has converged or reached oitm.
 * ------------------------------------------------------------------------ */
typedef struct snap_batch snap_batch_t;

typedef struct {
  int precision;            /* SNAP_PRECISION_FP64 or SNAP_PRECISION_MIXED */
  int tile[3];              /* sweep tile edges; 0 = autotune on the batch */
  int waves_per_tile;       /* 0 = choose */
  int angles_per_lane;      /* 0 = choose (autotuned with the tile) */
  int use_graph;            /* 1 = capture one outer iteration in a hipGraph and replay it */
  double fp64_polish;       /* MIXED: a problem sweeps its next inner iteration in FP64
                               once its previous inner iteration of the same outer
                               changed its flux by at most fp64_polish * epsi or
                               1e-5, whichever is larger (max |flux0/flux0p - 1|);
                               0 = FP32 sweeps throughout */
  int sweep_kernel;         /* SNAP_SWEEP_AUTO (the autotuner times one-wavefront 4x4xTZ
                               tiles, TZ 4, 8 or 16, up to 64 angles, and else the
                               multi-wave tiles), _TILED (only the multi-wave tiles)
                               or _WAVE_TILE (up to 64 angles; tile[2] sets TZ, else
                               timed) */
  int groups_per_wave;      /* one-wavefront tiles: 0 = choose (two groups per wavefront
                               where snap_solve_options_t's form applies, every
                               problem's x and y extents multiples of 4), 1 = one */
} snap_batch_options_t;

typedef struct {
  int inner_total;          /* inner iterations of this problem */
  int outer_total;          /* outer iterations of this problem */
  int converged;
  int sweeps_fp64;          /* inner sweeps run in FP64 (all of them at FP64) */
  double time_ms;           /* device clock: from the start of the run to the end of
                               the outer iteration in which it converged (to the end
                               of the run when it did not) */
} snap_problem_stats_t;

typedef struct {
  int outer_iterations;     /* outer iterations run by the batch */
  int converged;            /* problems converged */
  double solve_ms;          /* wall time of the run */
  double units;             /* sum over problems of cells * 8*nang * ng * inner iterations */
  double grind_ns;          /* solve_ms / units, in ns */
  int tile[3];              /* tile shape used */
  int waves;
  int angles_per_lane;
  double autotune_ms;       /* time the tile autotuner took at creation (0 when the
                               tile came from the process's cache of earlier batches
                               of the same shapes) */
  int plan_cached;          /* 1 = tile taken from that cache */
  int groups_per_wave;      /* groups per wavefront of the sweeps */
} snap_batch_stats_t;

int snap_hip_batch_create(const snap_problem_t *problems, int count,
                          const snap_batch_options_t *options, snap_batch_t **batch);
int snap_hip_batch_run(snap_batch_t *batch, snap_batch_stats_t *stats,
                       snap_problem_stats_t *per_problem /* [count] or NULL */);
/* flux0 [ng][cells of the problem], fluxm [ng][cells][cmom-1] (either NULL) */
int snap_hip_batch_get_flux(snap_batch_t *batch, int problem, double *flux0, double *fluxm);
void snap_hip_batch_destroy(snap_batch_t *batch);

#ifdef __cplusplus
}
#endif
#endif
