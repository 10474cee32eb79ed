/*
 * snap_oracle.h -- CPU oracle for the Legion-SNAP mini-KBA steady-state solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a Legion-free FP64 restatement of the
 * reference's CPU task bodies and of its transport_solve control flow, which
 * the HIP library's results are tested against: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path never calls it.
 *
 * Data layout (dense, x fastest, as the reference's SOA layout
 * snap.cc:1456-1475 with one field per energy group):
 *   cell index c = (z*ny + y)*nx + x
 *   per-group scalar fields      f[g][c]
 *   per-group angle records      f[g][c][nang]           (dinv)
 *   per-group moment records     f[g][c][cmom]           (qtot)
 *   per-group flux moments       f[g][c][cmom-1]         (fluxm, q2grpm)
 *   in-group scattering          s_xs[g][c][nmom]        (by Legendre order)
 *   material tables              sigt[mat][g]
 *   group-to-group scattering    slgg[g_to][mat][g_from][nmom]  (snap.cc:596-708)
 *   expansion coefficients       ec[oct][cmom][nang]     (octant stride cmom*nang)
 */
#ifndef SNAP_ORACLE_H
#define SNAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;
  double lx, ly, lz;
  int ng, nang, nmom, nmat;
  const int32_t *mat;       /* [cells] 0-based material index */
  const double *sigt;       /* [nmat][ng] */
  const double *slgg;       /* [ng][nmat][ng][nmom] */
  const double *qi;         /* [ng][cells] fixed isotropic source */
  const double *sigt_cell;  /* optional [ng][cells]: replaces sigt[mat[c]][g] */
  double epsi;
  int iitm, oitm;
  int fixup;
  int timedep;              /* time dependent: vdelt_g = 2/(dt (ng-g)) (snap.cc:716-738) */
  double tf;                /* total simulated time; dt = tf / nsteps (snap.cc:1174) */
  int nsteps;               /* time steps (0 counts as 1) */
  int mms;                  /* manufactured-solution source (src_opt 3): qi is ignored */
  int sweep_order;          /* 0: generic variant's operation order (sweep.cc:242-616);
                               1: AVX variant's (sweep.cc:1037-1520, nang % 4 == 0) */
} oracle_problem_t;

typedef struct {
  int inner_total;          /* inner iterations executed (sweeps of all octants and groups) */
  int outer_total;          /* outer iterations executed */
  int converged;            /* 1 if the outer test passed before oitm */
  int fixup_cells;          /* (cell,octant,group) sweeps whose fixup loop changed pc, summed over the solve */
  double sweep_seconds;     /* wall time spent in sweeps */
  double solve_seconds;     /* wall time of the whole outer/inner loop */
  int time_steps;           /* time steps run; converged = every step converged */
} oracle_stats_t;

/* compute_derived_globals, snap.cc:1171-1342 (3-D branch). */
int snap_oracle_quadrature(int nang, int nmom, double *mu, double *eta,
                           double *xi, double *w, double *ec);

/* ExpandCrossSection, expxs.cc:85-115 */
void snap_oracle_expand_xs(int ncells, int ng, const int32_t *mat,
                           const double *sig, double *xs);
/* ExpandScatteringCrossSection, expxs.cc:220-250 */
void snap_oracle_expand_scat_xs(int ncells, int ng, int nmat, int nmom,
                                const int32_t *mat, const double *slgg,
                                double *s_xs);
/* CalculateGeometryParam, expxs.cc:358-392 */
void snap_oracle_geometry_param(int ncells, int ng, int nang, const double *t_xs,
                                const double *vdelt, double hi, double hj,
                                double hk, const double *mu, const double *eta,
                                const double *xi, double *dinv);
/* CalcOuterSource, outer.cc:107-217 (all cmom-1 moments). */
void snap_oracle_outer_source(int ncells, int ng, int nmat, int nmom,
                              const int32_t *mat, const double *qi,
                              const double *slgg, const double *flux0,
                              const double *fluxm, double *q2grp0,
                              double *q2grpm);
/* CalcInnerSource, inner.cc:106-160 */
void snap_oracle_inner_source(int ncells, int ng, int nmom, const double *s_xs,
                              const double *flux0, const double *q2grp0,
                              const double *fluxm, const double *q2grpm,
                              double *qtot);
/* TestInnerConvergence inner.cc:283-324 (scale 1) / TestOuterConvergence
 * outer.cc:354-403 (scale 100).  Returns 1 when converged. */
int snap_oracle_converged(int n, const double *flux0, const double *flux0p,
                          double epsi);

/* One (octant, group) mini-KBA sweep, sweep.cc:242-616, vacuum boundaries
 * unless face inflows are given.  Writes the octant's contribution to the
 * scalar flux (part_flux[c]) and to the flux moments (part_fluxm[c][cmom-1]).
 * ghost_{x,y,z}_in: optional [ny*nz|nx*nz|nx*ny][nang] inflow faces indexed in
 * the reference's ghost-point order (ghostx (y,z), ghosty (x,z), ghostz (x,y));
 * ghost_*_out: optional outflow faces in the same order.
 * Returns the number of cells whose fixup loop ran at least once. */
int snap_oracle_sweep(int nx, int ny, int nz, int nang, int cmom, int corner,
                      double hi, double hj, double hk, const double *mu,
                      const double *eta, const double *xi, const double *w,
                      const double *ec, const double *qtot_g,
                      const double *dinv_g, const double *t_xs_g, int fixup,
                      const double *ghost_x_in, const double *ghost_y_in,
                      const double *ghost_z_in, double *ghost_x_out,
                      double *ghost_y_out, double *ghost_z_out,
                      double *part_flux, double *part_fluxm);
/* The same sweep in the AVX variant's operation order (sweep.cc:1037-1520;
 * the differences are listed in snap_oracle.c).  Returns -1 unless
 * nang % 4 == 0. */
int snap_oracle_sweep_avx(int nx, int ny, int nz, int nang, int cmom, int corner,
                          double hi, double hj, double hk, const double *mu,
                          const double *eta, const double *xi, const double *w,
                          const double *ec, const double *qtot_g,
                          const double *dinv_g, const double *t_xs_g, int fixup,
                          const double *ghost_x_in, const double *ghost_y_in,
                          const double *ghost_z_in, double *ghost_x_out,
                          double *ghost_y_out, double *ghost_z_out,
                          double *part_flux, double *part_fluxm);

/* The sweep above from vacuum inflow, with the time-dependent terms (vdelt
 * != 0: inflow time_flux_in and outflow time_flux_out, [cells][nang] of this
 * corner and group) and the MMS angular source qim_g ([cells][nang] or
 * NULL); ghost_*_out as above (optional). */
int snap_oracle_sweep_td(int nx, int ny, int nz, int nang, int cmom, int corner, double hi,
                         double hj, double hk, const double *mu, const double *eta,
                         const double *xi, const double *w, const double *ec,
                         const double *qtot_g, const double *dinv_g, const double *t_xs_g,
                         int fixup, double vdelt, const double *time_flux_in,
                         double *time_flux_out, const double *qim_g, double *ghost_x_out,
                         double *ghost_y_out, double *ghost_z_out, double *part_flux,
                         double *part_fluxm);

/* MMSInitFlux, MMSInitSource and MMSInitTimeDependent (mms.cc:81-409) for
 * nmom = 1: ref_flux [ng][cells] (scaled to tf - dt/2 when time dependent),
 * the unscaled angular source qim [8][ng][cells][nang] and the fixed source
 * qi [ng][cells] (phi/v_g when time dependent, else 0).  Returns -1 for
 * nmom != 1. */
int snap_oracle_mms_setup(const oracle_problem_t *p, double *ref_flux, double *qim, double *qi);
/* MMSCompare (mms.cc:497-541): result = {max, min, mean} of |phi/ref - 1|. */
void snap_oracle_mms_compare(int n, const double *flux, const double *ref, double result[3]);

/* transport_solve loop, snap.cc:386-541: nsteps time steps of the outer /
 * inner iteration (one step when steady).  flux0 [ng][cells],
 * fluxm [ng][cells][cmom-1] (may be NULL when nmom == 1).  nthreads <= 0
 * means all OpenMP threads. */
int snap_oracle_solve(const oracle_problem_t *p, double *flux0, double *fluxm,
                      oracle_stats_t *stats, int nthreads);

/* Time `ninner` inner iterations' sweeps (all octants and groups) of problem
 * p starting from a zero flux: the CPU baseline sample. Returns seconds. */
/* Reverse-mode derivative of one sweep (nmom = 1) with the forward fixup
 * decisions held fixed: adds sum_a dR/dq to u_g and sum_a dR/dt_xs to grad_g
 * for the direct response w_a * zeta_g[c] of pc. */
void snap_oracle_adjoint_sweep(int nx, int ny, int nz, int nang, int corner, double hi,
                               double hj, double hk, const double *mu, const double *eta,
                               const double *xi, const double *w, const double *qtot_g,
                               const double *t_xs_g, int fixup, const double *zeta_g,
                               double *u_g, double *grad_g);
/* dR/dt_xs[g][c] of R = sum resp[g][c] * flux0[g][c] at the converged flux0
 * (nmom = 1), by adjoint source iteration to tol.  Returns iterations or -1. */
int snap_oracle_adjoint(const oracle_problem_t *p, const double *flux0, const double *resp,
                        double tol, int maxit, double *grad, double *zeta_out, int nthreads);

double snap_oracle_time_sweeps(const oracle_problem_t *p, int ninner,
                               int nthreads);

#ifdef __cplusplus
}
#endif
#endif
