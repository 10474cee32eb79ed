/*
 * snap_oracle.c -- CPU oracle for the Legion-SNAP mini-KBA steady-state solve.
 *
 * TEST INFRASTRUCTURE ONLY (see snap_oracle.h).  Every function restates one
 * reference task body; the citation in front of each names the file:lines it
 * follows.  Compiled with -ffp-contract=off so that every multiply and add
 * rounds separately, as the reference's SSE/AVX intrinsics do.
 *
 * Deliberate, documented departures from the reference (each is a reference
 * defect listed with its citation):
 *  - ec is stored with octant stride cmom*nang (snap.cc:1320-1333 writes cmom
 *    moments into an nmom*nang stride);
 *  - the source expansion and moment reduction run over l < cmom
 *    (sweep.cc:361 loops forever; sweep.cc:1135 and outer.cc:205 use nmom);
 *  - all flux moments are re-zeroed each sweep (snap.cc:494 zeroes flux0 only);
 *  - octant contributions are added in octant order 0..7, one deterministic
 *    instance of the reference's unordered SumReduction (snap.cc:1832-1880);
 *  - in the AVX order (below) the flux moments sum all four lanes,
 *    (l0 + l1) + (l2 + l3): the AVX variant applies _mm256_hadd_pd twice and
 *    keeps 2 (l0 + l1), dropping lanes 2 and 3 (sweep.cc:1485-1492).
 *
 * Two operation orders of the sweep body (the `avx` argument of sweep_impl):
 *  - generic: MiniKBATask::cpu_implementation, sweep.cc:242-616;
 *  - AVX: MiniKBATask::avx_implementation, sweep.cc:1037-1520, the variant
 *    the reference registers on x86 with AVX (sweep.cc:165-179); it needs
 *    nang % 4 == 0 (sweep.cc:1064).  Angles a = 4j + k sit in lane k of
 *    vector j.  It differs from the generic order in (i) the fixup update's
 *    products, psi_in * (mu * (hi * (1 + hv))) and a denominator
 *    t_xs + ((mu hi hv_x + eta hj hv_y) + xi hk hv_z) (sweep.cc:1349-1393),
 *    (ii) NaN tests (!(f >= 0) counts as negative, sweep.cc:1229-1232; a
 *    denominator is kept only where pc > 0, sweep.cc:1371-1375), and (iii)
 *    the angle sums: each lane adds its angles in order from 0, then
 *    total = (l0 + l1) + (l2 + l3) (sweep.cc:1465-1475).
 */
#include "snap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* compute_derived_globals, snap.cc:1171-1342, num_dims == 3.
 * mu accumulates by dm (1210-1212), eta counts down (1214-1216),
 * xi = sqrt(1 - (mu^2 + eta^2)) (1219-1222), w = 0.125/nang (1233-1234),
 * ec[oct][mom][a] = (is*mu)^(2l-1) * (ks*xi*js*eta)^m (1312-1336). */
int snap_oracle_quadrature(int nang, int nmom, double *mu, double *eta,
                           double *xi, double *w, double *ec) {
  if (nang < 1 || nmom < 1 || nmom > 4) return -1;
  const int cmom = nmom * nmom;
  const double dm = 1.0 / (double)nang;
  mu[0] = 0.5 * dm;
  for (int i = 1; i < nang; i++) mu[i] = mu[i - 1] + dm;
  eta[0] = 1.0 - 0.5 * dm;
  for (int i = 1; i < nang; i++) eta[i] = eta[i - 1] - dm;
  for (int i = 0; i < nang; i++) {
    const double t = mu[i] * mu[i] + eta[i] * eta[i];
    xi[i] = sqrt(1.0 - t);
  }
  for (int i = 0; i < nang; i++) w[i] = 0.125 / (double)nang;
  memset(ec, 0, sizeof(double) * (size_t)8 * cmom * nang);
  for (int kd = 0; kd < 2; kd++) {
    const int ks = kd ? 1 : -1;
    for (int jd = 0; jd < 2; jd++) {
      const int js = jd ? 1 : -1;
      for (int id = 0; id < 2; id++) {
        const int is = id ? 1 : -1;
        const int oct = 4 * kd + 2 * jd + id;
        double *e = ec + (size_t)oct * cmom * nang;
        for (int a = 0; a < nang; a++) e[a] = 1.0;
        int moment = 1;
        for (int l = 1; l < nmom; l++) {
          for (int m = 0; m < 2 * (l + 1) - 1; m++) {
            for (int a = 0; a < nang; a++)
              e[moment * nang + a] = pow(is * mu[a], 2 * (l + 1) - 3) *
                                     pow(ks * xi[a] * js * eta[a], m);
            moment++;
          }
        }
      }
    }
  }
  return 0;
}

/* ExpandCrossSection::cpu_implementation, expxs.cc:85-115. */
void snap_oracle_expand_xs(int ncells, int ng, const int32_t *mat,
                           const double *sig, double *xs) {
  for (int g = 0; g < ng; g++)
    for (int c = 0; c < ncells; c++) xs[(size_t)g * ncells + c] = sig[(size_t)mat[c] * ng + g];
}

/* ExpandScatteringCrossSection::cpu_implementation, expxs.cc:220-250:
 * s_xs[g][c] = slgg[g][mat[c]][g] (the in-group element, every order l). */
void snap_oracle_expand_scat_xs(int ncells, int ng, int nmat, int nmom,
                                const int32_t *mat, const double *slgg,
                                double *s_xs) {
  for (int g = 0; g < ng; g++)
    for (int c = 0; c < ncells; c++) {
      const double *src = slgg + (((size_t)g * nmat + mat[c]) * ng + g) * nmom;
      double *dst = s_xs + ((size_t)g * ncells + c) * nmom;
      for (int l = 0; l < nmom; l++) dst[l] = src[l];
    }
}

/* CalculateGeometryParam::cpu_implementation, expxs.cc:358-392:
 * dinv = 1/(xs + vdelt + hi*mu + hj*eta + hk*xi), added left to right. */
void snap_oracle_geometry_param(int ncells, int ng, int nang, const double *t_xs,
                                const double *vdelt, double hi, double hj,
                                double hk, const double *mu, const double *eta,
                                const double *xi, double *dinv) {
  for (int g = 0; g < ng; g++) {
    const double vd = vdelt ? vdelt[g] : 0.0;
    for (int c = 0; c < ncells; c++) {
      const double xs = t_xs[(size_t)g * ncells + c];
      double *d = dinv + ((size_t)g * ncells + c) * nang;
      for (int a = 0; a < nang; a++)
        d[a] = 1.0 / (xs + vd + hi * mu[a] + hj * eta[a] + hk * xi[a]);
    }
  }
}

/* Legendre order of flux moment index m >= 1 (lma = {1,3,5,7}, snap.cc:1307). */
static int moment_order(int m) {
  int l = 1, start = 1;
  while (m >= start + (2 * l + 1)) {
    start += 2 * l + 1;
    l++;
  }
  return l;
}

/* CalcOuterSource::cpu_implementation, outer.cc:107-217.
 * q2grp0[g1][c] = qi[g1][c] + sum_{g2 != g1, ascending} slgg[g1][mat][g2][0]*flux0[g2][c]
 * q2grpm[g1][c][m-1] = sum_{g2 != g1} slgg[g1][mat][g2][L(m)]*fluxm[g2][c][m-1],
 * for all m = 1..cmom-1 (the reference stops at nmom-1, outer.cc:205). */
void snap_oracle_outer_source(int ncells, int ng, int nmat, int nmom,
                              const int32_t *mat, const double *qi,
                              const double *slgg, const double *flux0,
                              const double *fluxm, double *q2grp0,
                              double *q2grpm) {
  const int cmom = nmom * nmom;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < ncells; c++) {
    const int mt = mat[c];
    for (int g1 = 0; g1 < ng; g1++) {
      double qo0 = qi[(size_t)g1 * ncells + c];
      for (int g2 = 0; g2 < ng; g2++) {
        if (g1 == g2) continue;
        const double cs = slgg[(((size_t)g1 * nmat + mt) * ng + g2) * nmom];
        qo0 += cs * flux0[(size_t)g2 * ncells + c];
      }
      q2grp0[(size_t)g1 * ncells + c] = qo0;
    }
    if (cmom > 1) {
      for (int g1 = 0; g1 < ng; g1++) {
        double *qom = q2grpm + ((size_t)g1 * ncells + c) * (cmom - 1);
        for (int m = 1; m < cmom; m++) {
          const int l = moment_order(m);
          double acc = 0.0;
          for (int g2 = 0; g2 < ng; g2++) {
            if (g1 == g2) continue;
            const double cs = slgg[(((size_t)g1 * nmat + mt) * ng + g2) * nmom + l];
            acc += cs * fluxm[((size_t)g2 * ncells + c) * (cmom - 1) + (m - 1)];
          }
          qom[m - 1] = acc;
        }
      }
    }
  }
}

/* CalcInnerSource::cpu_implementation, inner.cc:106-160:
 * qtot[0] = q2grp0 + flux0*s_xs[0]; qtot[m] = q2grpm[m-1] + fluxm[m-1]*s_xs[L(m)]. */
void snap_oracle_inner_source(int ncells, int ng, int nmom, const double *s_xs,
                              const double *flux0, const double *q2grp0,
                              const double *fluxm, const double *q2grpm,
                              double *qtot) {
  const int cmom = nmom * nmom;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < ncells; c++) {
    for (int g = 0; g < ng; g++) {
      const size_t gc = (size_t)g * ncells + c;
      const double *sx = s_xs + gc * nmom;
      double *q = qtot + gc * cmom;
      q[0] = q2grp0[gc] + flux0[gc] * sx[0];
      for (int m = 1; m < cmom; m++)
        q[m] = q2grpm[gc * (cmom - 1) + m - 1] +
               fluxm[gc * (cmom - 1) + m - 1] * sx[moment_order(m)];
    }
  }
}

/* TestInnerConvergence inner.cc:283-324 and TestOuterConvergence
 * outer.cc:354-403: df = |flux0/flux0p - 1|, or |flux0| when |flux0p| < 1e-12;
 * not converged as soon as df > epsi (NaN compares false: converged). */
int snap_oracle_converged(int n, const double *flux0, const double *flux0p,
                          double epsi) {
  const double tolr = 1.0e-12;
  for (int i = 0; i < n; i++) {
    double fp = flux0p[i];
    double df = 1.0;
    if (fabs(fp) < tolr) {
      fp = 1.0;
      df = 0.0;
    }
    df = fabs((flux0[i] / fp) - df);
    if (df > epsi) return 0;
  }
  return 1;
}

/* MiniKBATask::cpu_implementation, sweep.cc:242-616 (generic variant).
 * vdelt != 0 adds the time terms (sweep.cc:414-419, 461-479, 512-519,
 * 529-533): time_flux_in/out are [cells][nang] of this (corner, group);
 * qim (NULL unless MMS) is the angular source [cells][nang] added to psi
 * (sweep.cc:347-351). */
static int sweep_impl(int nx, int ny, int nz, int nang, int cmom, int corner, double hi,
                      double hj, double hk, const double *mu, const double *eta,
                      const double *xi, const double *w, const double *ec,
                      const double *qtot_g, const double *dinv_g, const double *t_xs_g,
                      int fixup, const double *ghost_x_in, const double *ghost_y_in,
                      const double *ghost_z_in, double *ghost_x_out, double *ghost_y_out,
                      double *ghost_z_out, double *part_flux, double *part_fluxm,
                      double vdelt, const double *time_flux_in, double *time_flux_out,
                      const double *qim_g, int avx) {
  if (avx && nang % 4 != 0) return -1;
  const int sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
  const size_t ab = (size_t)nang;
  double *buf = (double *)malloc(sizeof(double) * ab * 16);
  double *psi = buf, *pc = buf + ab, *psii = buf + 2 * ab, *psij = buf + 3 * ab,
         *psik = buf + 4 * ab, *hv_x = buf + 5 * ab, *hv_y = buf + 6 * ab,
         *hv_z = buf + 7 * ab, *fx_hv_x = buf + 8 * ab, *fx_hv_y = buf + 9 * ab,
         *fx_hv_z = buf + 10 * ab, *tmp = buf + 11 * ab, *hv_t = buf + 12 * ab,
         *fx_hv_t = buf + 13 * ab, *tin = buf + 14 * ab;
  double *yflux_pencil = (double *)malloc(sizeof(double) * ab * nx);
  double *zflux_plane = (double *)malloc(sizeof(double) * ab * nx * ny);
  const double *ecb = ec + (size_t)corner * cmom * nang;
  const double tolr = 1.0e-12;
  int fixed_cells = 0;
  (void)tmp;

  for (int z = 0; z < nz; z++) {
    for (int y = 0; y < ny; y++) {
      for (int x = 0; x < nx; x++) {
        const int px = sxp ? x : nx - 1 - x;
        const int py = syp ? y : ny - 1 - y;
        const int pz = szp ? z : nz - 1 - z;
        const size_t c = ((size_t)pz * ny + py) * nx + px;
        /* angular source (sweep.cc:331-345) */
        const double *quad = qtot_g + c * cmom;
        for (int a = 0; a < nang; a++) psi[a] = quad[0];
        for (int l = 1; l < cmom; l++)
          for (int a = 0; a < nang; a++) psi[a] += ecb[l * nang + a] * quad[l];
        /* MMS angular source (sweep.cc:347-351) */
        if (qim_g)
          for (int a = 0; a < nang; a++) psi[a] += qim_g[c * ab + a];
        for (int a = 0; a < nang; a++) pc[a] = psi[a];
        /* x inflow (sweep.cc:356-364) */
        if (x == 0) {
          if (ghost_x_in)
            memcpy(psii, ghost_x_in + ((size_t)pz * ny + py) * ab, sizeof(double) * ab);
          else
            memset(psii, 0, sizeof(double) * ab);
        }
        for (int a = 0; a < nang; a++) pc[a] += psii[a] * mu[a] * hi;
        /* y inflow (sweep.cc:365-376) */
        if (y == 0) {
          if (ghost_y_in)
            memcpy(psij, ghost_y_in + ((size_t)pz * nx + px) * ab, sizeof(double) * ab);
          else
            memset(psij, 0, sizeof(double) * ab);
        } else {
          memcpy(psij, yflux_pencil + (size_t)x * ab, sizeof(double) * ab);
        }
        for (int a = 0; a < nang; a++) pc[a] += psij[a] * eta[a] * hj;
        /* z inflow (sweep.cc:377-388) */
        if (z == 0) {
          if (ghost_z_in)
            memcpy(psik, ghost_z_in + ((size_t)py * nx + px) * ab, sizeof(double) * ab);
          else
            memset(psik, 0, sizeof(double) * ab);
        } else {
          memcpy(psik, zflux_plane + ((size_t)y * nx + x) * ab, sizeof(double) * ab);
        }
        for (int a = 0; a < nang; a++) pc[a] += psik[a] * xi[a] * hk;
        /* time-edge inflow (sweep.cc:414-419) */
        if (vdelt != 0.0) {
          memcpy(tin, time_flux_in + c * ab, sizeof(double) * ab);
          for (int a = 0; a < nang; a++) pc[a] += vdelt * tin[a];
        }
        /* multiply by dinv (sweep.cc:420-423) */
        const double *dinv = dinv_g + c * ab;
        for (int a = 0; a < nang; a++) pc[a] *= dinv[a];

        if (fixup) {
          /* negative flux fixup, sweep.cc:424-520 */
          unsigned old_negative_fluxes = 0;
          int ran = 0;
          for (int a = 0; a < nang; a++) hv_x[a] = 1.0;
          for (int a = 0; a < nang; a++) hv_y[a] = 1.0;
          for (int a = 0; a < nang; a++) hv_z[a] = 1.0;
          for (int a = 0; a < nang; a++) hv_t[a] = 1.0;
          const double t_xs = t_xs_g[c];
          while (1) {
            unsigned negative_fluxes = 0;
            for (int a = 0; a < nang; a++) {
              fx_hv_x[a] = 2.0 * pc[a] - psii[a];
              if (avx ? !(fx_hv_x[a] >= 0.0) : fx_hv_x[a] < 0.0) {
                hv_x[a] = 0.0;
                negative_fluxes++;
              }
            }
            for (int a = 0; a < nang; a++) {
              fx_hv_y[a] = 2.0 * pc[a] - psij[a];
              if (avx ? !(fx_hv_y[a] >= 0.0) : fx_hv_y[a] < 0.0) {
                hv_y[a] = 0.0;
                negative_fluxes++;
              }
            }
            for (int a = 0; a < nang; a++) {
              fx_hv_z[a] = 2.0 * pc[a] - psik[a];
              if (avx ? !(fx_hv_z[a] >= 0.0) : fx_hv_z[a] < 0.0) {
                hv_z[a] = 0.0;
                negative_fluxes++;
              }
            }
            if (vdelt != 0.0) {
              for (int a = 0; a < nang; a++) {
                fx_hv_t[a] = 2.0 * pc[a] - tin[a];
                if (avx ? !(fx_hv_t[a] >= 0.0) : fx_hv_t[a] < 0.0) {
                  hv_t[a] = 0.0;
                  negative_fluxes++;
                }
              }
            }
            if (negative_fluxes == old_negative_fluxes) break;
            old_negative_fluxes = negative_fluxes;
            ran = 1;
            if (avx) {
              /* sweep.cc:1302-1393 */
              for (int a = 0; a < nang; a++) {
                double sum = psii[a] * (mu[a] * (hi * (1.0 + hv_x[a])));
                sum = sum + psij[a] * (eta[a] * (hj * (1.0 + hv_y[a])));
                sum = sum + psik[a] * (xi[a] * (hk * (1.0 + hv_z[a])));
                double den3 = (mu[a] * hi) * hv_x[a] + (eta[a] * hj) * hv_y[a];
                den3 = den3 + (xi[a] * hk) * hv_z[a];
                if (vdelt != 0.0) {
                  sum = sum + tin[a] * (vdelt * (1.0 + hv_t[a]));
                  den3 = den3 + vdelt * hv_t[a];
                }
                pc[a] = psi[a] + 0.5 * sum;
                double den = t_xs + den3;
                if (!(pc[a] > 0.0)) den = 0.0;
                pc[a] = (den >= tolr) ? pc[a] / den : 0.0;
              }
            } else if (vdelt != 0.0) {
              for (int a = 0; a < nang; a++) {
                pc[a] = psi[a] + 0.5 * (psii[a] * mu[a] * hi * (1.0 + hv_x[a]) +
                                        psij[a] * eta[a] * hj * (1.0 + hv_y[a]) +
                                        psik[a] * xi[a] * hk * (1.0 + hv_z[a]) +
                                        tin[a] * vdelt * (1.0 + hv_t[a]));
                double den = (pc[a] <= 0.0) ? 0.0
                                            : (t_xs + mu[a] * hi * hv_x[a] +
                                               eta[a] * hj * hv_y[a] +
                                               xi[a] * hk * hv_z[a] + vdelt * hv_t[a]);
                if (den < tolr)
                  pc[a] = 0.0;
                else
                  pc[a] /= den;
              }
            } else {
              for (int a = 0; a < nang; a++) {
                pc[a] = psi[a] + 0.5 * (psii[a] * mu[a] * hi * (1.0 + hv_x[a]) +
                                        psij[a] * eta[a] * hj * (1.0 + hv_y[a]) +
                                        psik[a] * xi[a] * hk * (1.0 + hv_z[a]));
                double den = (pc[a] <= 0.0) ? 0.0
                                            : (t_xs + mu[a] * hi * hv_x[a] +
                                               eta[a] * hj * hv_y[a] +
                                               xi[a] * hk * hv_z[a]);
                if (den < tolr)
                  pc[a] = 0.0;
                else
                  pc[a] /= den;
              }
            }
          }
          fixed_cells += ran;
          for (int a = 0; a < nang; a++) psii[a] = fx_hv_x[a] * hv_x[a];
          for (int a = 0; a < nang; a++) psij[a] = fx_hv_y[a] * hv_y[a];
          for (int a = 0; a < nang; a++) psik[a] = fx_hv_z[a] * hv_z[a];
          if (vdelt != 0.0)
            for (int a = 0; a < nang; a++) time_flux_out[c * ab + a] = fx_hv_t[a] * hv_t[a];
        } else {
          /* sweep.cc:521-534 */
          for (int a = 0; a < nang; a++) psii[a] = 2.0 * pc[a] - psii[a];
          for (int a = 0; a < nang; a++) psij[a] = 2.0 * pc[a] - psij[a];
          for (int a = 0; a < nang; a++) psik[a] = 2.0 * pc[a] - psik[a];
          if (vdelt != 0.0)
            for (int a = 0; a < nang; a++) time_flux_out[c * ab + a] = 2.0 * pc[a] - tin[a];
        }
        /* outflow faces (sweep.cc:536-566) */
        if (x == nx - 1 && ghost_x_out)
          memcpy(ghost_x_out + ((size_t)pz * ny + py) * ab, psii, sizeof(double) * ab);
        if (y == ny - 1) {
          if (ghost_y_out)
            memcpy(ghost_y_out + ((size_t)pz * nx + px) * ab, psij, sizeof(double) * ab);
        } else {
          memcpy(yflux_pencil + (size_t)x * ab, psij, sizeof(double) * ab);
        }
        if (z == nz - 1) {
          if (ghost_z_out)
            memcpy(ghost_z_out + ((size_t)py * nx + px) * ab, psik, sizeof(double) * ab);
        } else {
          memcpy(zflux_plane + ((size_t)y * nx + x) * ab, psik, sizeof(double) * ab);
        }
        /* moment reductions (sweep.cc:568-591; AVX lanes, sweep.cc:1465-1493) */
        if (avx) {
          /* lane k of vector j holds angle 4j + k */
          double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
          for (int a = 0; a < nang; a += 4) {
            psi[a] = pc[a] * w[a];
            psi[a + 1] = pc[a + 1] * w[a + 1];
            psi[a + 2] = pc[a + 2] * w[a + 2];
            psi[a + 3] = pc[a + 3] * w[a + 3];
            l0 = l0 + psi[a];
            l1 = l1 + psi[a + 1];
            l2 = l2 + psi[a + 2];
            l3 = l3 + psi[a + 3];
          }
          part_flux[c] = (l0 + l1) + (l2 + l3);
          for (int l = 1; l < cmom; l++) {
            const double *e = ecb + (size_t)l * nang;
            l0 = l1 = l2 = l3 = 0.0;
            for (int a = 0; a < nang; a += 4) {
              l0 = l0 + psi[a] * e[a];
              l1 = l1 + psi[a + 1] * e[a + 1];
              l2 = l2 + psi[a + 2] * e[a + 2];
              l3 = l3 + psi[a + 3] * e[a + 3];
            }
            part_fluxm[c * (cmom - 1) + (l - 1)] = (l0 + l1) + (l2 + l3);
          }
          continue;
        }
        double total = 0.0;
        for (int a = 0; a < nang; a++) {
          psi[a] = w[a] * pc[a];
          total += psi[a];
        }
        part_flux[c] = total;
        if (cmom > 1) {
          for (int l = 1; l < cmom; l++) {
            const double *e = ecb + (size_t)l * nang;
            total = 0.0;
            for (int a = 0; a < nang; a++) total += e[a] * psi[a];
            part_fluxm[c * (cmom - 1) + (l - 1)] = total;
          }
        }
      }
    }
  }
  free(zflux_plane);
  free(yflux_pencil);
  free(buf);
  return fixed_cells;
}

int snap_oracle_sweep(int nx, int ny, int nz, int nang, int cmom, int corner,
                      double hi, double hj, double hk, const double *mu,
                      const double *eta, const double *xi, const double *w,
                      const double *ec, const double *qtot_g,
                      const double *dinv_g, const double *t_xs_g, int fixup,
                      const double *ghost_x_in, const double *ghost_y_in,
                      const double *ghost_z_in, double *ghost_x_out,
                      double *ghost_y_out, double *ghost_z_out,
                      double *part_flux, double *part_fluxm) {
  return sweep_impl(nx, ny, nz, nang, cmom, corner, hi, hj, hk, mu, eta, xi, w, ec, qtot_g,
                    dinv_g, t_xs_g, fixup, ghost_x_in, ghost_y_in, ghost_z_in, ghost_x_out,
                    ghost_y_out, ghost_z_out, part_flux, part_fluxm, 0.0, NULL, NULL, NULL, 0);
}

int snap_oracle_sweep_avx(int nx, int ny, int nz, int nang, int cmom, int corner,
                          double hi, double hj, double hk, const double *mu,
                          const double *eta, const double *xi, const double *w,
                          const double *ec, const double *qtot_g,
                          const double *dinv_g, const double *t_xs_g, int fixup,
                          const double *ghost_x_in, const double *ghost_y_in,
                          const double *ghost_z_in, double *ghost_x_out,
                          double *ghost_y_out, double *ghost_z_out,
                          double *part_flux, double *part_fluxm) {
  return sweep_impl(nx, ny, nz, nang, cmom, corner, hi, hj, hk, mu, eta, xi, w, ec, qtot_g,
                    dinv_g, t_xs_g, fixup, ghost_x_in, ghost_y_in, ghost_z_in, ghost_x_out,
                    ghost_y_out, ghost_z_out, part_flux, part_fluxm, 0.0, NULL, NULL, NULL, 1);
}

int snap_oracle_sweep_td(int nx, int ny, int nz, int nang, int cmom, int corner, double hi,
                         double hj, double hk, const double *mu, const double *eta,
                         const double *xi, const double *w, const double *ec,
                         const double *qtot_g, const double *dinv_g, const double *t_xs_g,
                         int fixup, double vdelt, const double *time_flux_in,
                         double *time_flux_out, const double *qim_g, double *ghost_x_out,
                         double *ghost_y_out, double *ghost_z_out, double *part_flux,
                         double *part_fluxm) {
  return sweep_impl(nx, ny, nz, nang, cmom, corner, hi, hj, hk, mu, eta, xi, w, ec, qtot_g,
                    dinv_g, t_xs_g, fixup, NULL, NULL, NULL, ghost_x_out, ghost_y_out,
                    ghost_z_out, part_flux, part_fluxm, vdelt, time_flux_in, time_flux_out,
                    qim_g, 0);
}

/* ------------------------------------------------------------------------ */
/* Solve loop (snap.cc:225-568, steady state, one time step)                 */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nx, ny, nz, ncells, ng, nang, nmom, cmom;
  double hi, hj, hk;
  double *mu, *eta, *xi, *w, *ec;
  double *t_xs, *s_xs, *dinv;
  double *part_flux, *part_fluxm; /* [8][ng][cells] and [8][ng][cells][cmom-1] */
  double dt;                      /* tf / nsteps (snap.cc:1174) */
  double *vdelt;                  /* [ng] */
  /* time-dependent sweeps: the step's inflow and outflow time-edge fluxes,
   * [8][ng][cells][nang] each (NULL when vdelt is zero) */
  const double *tflux_in;
  double *tflux_out;
  const double *qim;              /* MMS angular source [8][ng][cells][nang] or NULL */
  int avx;                        /* the AVX variant's operation order */
} oracle_state_t;

static int problem_steps(const oracle_problem_t *p) { return p->nsteps > 0 ? p->nsteps : 1; }

static int state_init(oracle_state_t *s, const oracle_problem_t *p) {
  memset(s, 0, sizeof(*s));
  if (p->sweep_order != 0 && (p->sweep_order != 1 || p->nang % 4 != 0)) return -1;
  s->avx = p->sweep_order;
  s->nx = p->nx;
  s->ny = p->ny;
  s->nz = p->nz;
  s->ncells = p->nx * p->ny * p->nz;
  s->ng = p->ng;
  s->nang = p->nang;
  s->nmom = p->nmom;
  s->cmom = p->nmom * p->nmom;
  /* snap.cc:1178-1180 */
  s->hi = 2.0 / (p->lx / (double)p->nx);
  s->hj = 2.0 / (p->ly / (double)p->ny);
  s->hk = 2.0 / (p->lz / (double)p->nz);
  const size_t na = (size_t)p->nang;
  s->mu = (double *)malloc(sizeof(double) * na);
  s->eta = (double *)malloc(sizeof(double) * na);
  s->xi = (double *)malloc(sizeof(double) * na);
  s->w = (double *)malloc(sizeof(double) * na);
  s->ec = (double *)malloc(sizeof(double) * 8 * s->cmom * na);
  if (snap_oracle_quadrature(p->nang, p->nmom, s->mu, s->eta, s->xi, s->w, s->ec)) return -1;
  const size_t gc = (size_t)s->ng * s->ncells;
  s->t_xs = (double *)malloc(sizeof(double) * gc);
  s->s_xs = (double *)malloc(sizeof(double) * gc * s->nmom);
  s->dinv = (double *)malloc(sizeof(double) * gc * na);
  s->part_flux = (double *)malloc(sizeof(double) * 8 * gc);
  s->part_fluxm = s->cmom > 1 ? (double *)malloc(sizeof(double) * 8 * gc * (s->cmom - 1)) : NULL;
  s->vdelt = (double *)calloc((size_t)s->ng, sizeof(double));
  if (!s->t_xs || !s->s_xs || !s->dinv || !s->part_flux || !s->vdelt) return -1;
  /* Snap::initialize_velocity (snap.cc:716-738): v_g = ng - g and
   * vdelt_g = 2 / (dt v_g) when time dependent, else 0 */
  s->dt = p->tf / (double)problem_steps(p);
  if (p->timedep)
    for (int g = 0; g < s->ng; g++) s->vdelt[g] = 2.0 / (s->dt * (double)(s->ng - g));
  /* ExpandCrossSection / ExpandScatteringCrossSection / CalculateGeometryParam
   * (snap.cc:419-450); a per-cell total cross section replaces the gather. */
  if (p->sigt_cell)
    memcpy(s->t_xs, p->sigt_cell, sizeof(double) * gc);
  else
    snap_oracle_expand_xs(s->ncells, s->ng, p->mat, p->sigt, s->t_xs);
  snap_oracle_expand_scat_xs(s->ncells, s->ng, p->nmat, s->nmom, p->mat, p->slgg, s->s_xs);
  snap_oracle_geometry_param(s->ncells, s->ng, p->nang, s->t_xs, s->vdelt, s->hi, s->hj, s->hk,
                             s->mu, s->eta, s->xi, s->dinv);
  return 0;
}

static void state_free(oracle_state_t *s) {
  free(s->mu);
  free(s->eta);
  free(s->xi);
  free(s->w);
  free(s->ec);
  free(s->t_xs);
  free(s->s_xs);
  free(s->dinv);
  free(s->part_flux);
  free(s->part_fluxm);
  free(s->vdelt);
}

/* perform_sweeps, snap.cc:849-887: 8 corners x all groups, independent tasks.
 * Each task writes its own partial; partials are summed in octant order. */
static int perform_sweeps(oracle_state_t *s, int fixup, const double *qtot,
                          double *flux0, double *fluxm) {
  const int ng = s->ng, nc = s->ncells, cm = s->cmom;
  const size_t gc = (size_t)ng * nc;
  int fixed = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : fixed)
  for (int task = 0; task < 8 * ng; task++) {
    const int corner = task / ng, g = task % ng;
    const size_t ta = ((size_t)corner * gc + (size_t)g * nc) * s->nang;
    const double vd = s->vdelt[g];
    fixed += sweep_impl(
        s->nx, s->ny, s->nz, s->nang, cm, corner, s->hi, s->hj, s->hk, s->mu, s->eta, s->xi,
        s->w, s->ec, qtot + (size_t)g * nc * cm, s->dinv + (size_t)g * nc * s->nang,
        s->t_xs + (size_t)g * nc, fixup, NULL, NULL, NULL, NULL, NULL, NULL,
        s->part_flux + (size_t)corner * gc + (size_t)g * nc,
        cm > 1 ? s->part_fluxm + ((size_t)corner * gc + (size_t)g * nc) * (cm - 1) : NULL, vd,
        vd != 0.0 ? s->tflux_in + ta : NULL, vd != 0.0 ? s->tflux_out + ta : NULL,
        s->qim ? s->qim + ta : NULL, s->avx);
  }
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < gc; i++) {
    double acc = 0.0;
    for (int o = 0; o < 8; o++) acc += s->part_flux[(size_t)o * gc + i];
    flux0[i] = acc;
  }
  if (cm > 1) {
    const size_t gm = gc * (cm - 1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < gm; i++) {
      double acc = 0.0;
      for (int o = 0; o < 8; o++) acc += s->part_fluxm[(size_t)o * gm + i];
      fluxm[i] = acc;
    }
  }
  return fixed;
}

/* ------------------------------------------------------------------------ */
/* Method of manufactured solutions (mms.cc)                                */
/* ------------------------------------------------------------------------ */

#define MMS_PI 3.14159265358979 /* the reference's PI (snap.h:37) */

/* mms_trigint (mms.cc:34-57): cell averages over the edges cb[0..lc] of
 * sin(d x) (COS form: (cos(d cb_i) - cos(d cb_i+1)) / (d del)) and of
 * d cos(d x) (SIN form: (sin(d cb_i+1) - sin(d cb_i)) / del). */
static void mms_trigint(int cos_form, int lc, double d, double del, const double *cb, double *fn) {
  const double denom = d * del;
  for (int i = 0; i < lc; i++) {
    if (cos_form) {
      fn[i] = cos(d * cb[i]) - cos(d * cb[i + 1]);
      fn[i] /= denom;
    } else {
      fn[i] = sin(d * cb[i + 1]) - sin(d * cb[i]);
      fn[i] /= del;
    }
  }
}

/* Cell edges of one axis from 0 in steps of del, summed as mms.cc:108-117
 * does (the reference steps the y and z edges by dx; for the cubic cells of
 * every reference deck that is the same as dy and dz). */
static void mms_edges(int n, double del, double *b) {
  b[0] = 0.0;
  for (int i = 1; i <= n; i++) b[i] = b[i - 1] + del;
}

int snap_oracle_mms_setup(const oracle_problem_t *p, double *ref_flux, double *qim, double *qi) {
  if (p->nmom != 1) return -1; /* the reference's MMS moments are defined for nmom = 1 only */
  const int nx = p->nx, ny = p->ny, nz = p->nz, ng = p->ng, nang = p->nang;
  const size_t nc = (size_t)nx * ny * nz, gc = (size_t)ng * nc;
  double *mu = (double *)malloc(sizeof(double) * nang), *eta = (double *)malloc(sizeof(double) * nang);
  double *xi = (double *)malloc(sizeof(double) * nang), *w = (double *)malloc(sizeof(double) * nang);
  double *ec = (double *)malloc(sizeof(double) * 8 * nang);
  double *ib = (double *)malloc(sizeof(double) * (nx + 1)), *jb = (double *)malloc(sizeof(double) * (ny + 1));
  double *kb = (double *)malloc(sizeof(double) * (nz + 1));
  double *cx = (double *)malloc(sizeof(double) * nx), *sx = (double *)malloc(sizeof(double) * nx);
  double *cy = (double *)malloc(sizeof(double) * ny), *sy = (double *)malloc(sizeof(double) * ny);
  double *cz = (double *)malloc(sizeof(double) * nz), *sz = (double *)malloc(sizeof(double) * nz);
  snap_oracle_quadrature(nang, 1, mu, eta, xi, w, ec);
  const double a = MMS_PI / p->lx, b = MMS_PI / p->ly, cc = MMS_PI / p->lz;
  const double dx = p->lx / (double)nx, dy = p->ly / (double)ny, dz = p->lz / (double)nz;
  mms_edges(nx, dx, ib);
  mms_edges(ny, dy, jb);
  mms_edges(nz, dz, kb);
  mms_trigint(1, nx, a, dx, ib, cx);
  mms_trigint(0, nx, a, dx, ib, sx);
  mms_trigint(1, ny, b, dy, jb, cy);
  mms_trigint(0, ny, b, dy, jb, sy);
  mms_trigint(1, nz, cc, dz, kb, cz);
  mms_trigint(0, nz, cc, dz, kb, sz);
  /* MMSInitFlux (mms.cc:81-148): phi_g = g * tx * ty * tz, g 1-based */
  for (int g = 0; g < ng; g++)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          ref_flux[(size_t)g * nc + ((size_t)k * ny + j) * nx + i] =
              (double)(g + 1) * cx[i] * cy[j] * cz[k];
  /* MMSInitSource (mms.cc:206-344), per corner: the streaming, collision
   * and scattering terms of the manufactured flux */
  for (int corner = 0; corner < 8; corner++) {
    const double is = (corner & 1) ? 1.0 : -1.0, js = (corner & 2) ? 1.0 : -1.0,
                 ks = (corner & 4) ? 1.0 : -1.0;
    for (int g = 0; g < ng; g++)
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++) {
            const size_t c = ((size_t)k * ny + j) * nx + i;
            const int m = p->mat[c];
            const double sigt = p->sigt[(size_t)m * ng + g];
            const double flux_update = sigt * ref_flux[(size_t)g * nc + c];
            double *q = qim + (((size_t)corner * ng + g) * nc + c) * nang;
            for (int an = 0; an < nang; an++) {
              double v = 0.0;
              v += (double)(g + 1) * is * mu[an] * sx[i] * cy[j] * cz[k];
              v += flux_update;
              v += (double)(g + 1) * js * eta[an] * cx[i] * sy[j] * cz[k];
              v += (double)(g + 1) * ks * xi[an] * cx[i] * cy[j] * sz[k];
              for (int gp = 0; gp < ng; gp++)
                v -= p->slgg[(((size_t)g * p->nmat + m) * ng + gp) * p->nmom] *
                     ref_flux[(size_t)gp * nc + c];
              q[an] = v;
            }
          }
  }
  /* MMSInitTimeDependent (mms.cc:382-409): qi = phi / v_g, phi scaled to
   * the last step's mid-time tf - dt/2 */
  if (p->timedep) {
    const double dt = p->tf / (double)problem_steps(p);
    const double t_scale = p->tf - 0.5 * dt;
    for (int g = 0; g < ng; g++) {
      const double vg = (double)(ng - g);
      for (size_t c = 0; c < nc; c++) {
        const double f = ref_flux[(size_t)g * nc + c];
        qi[(size_t)g * nc + c] = f / vg;
        ref_flux[(size_t)g * nc + c] = f * t_scale;
      }
    }
  } else {
    memset(qi, 0, sizeof(double) * gc);
  }
  free(mu); free(eta); free(xi); free(w); free(ec);
  free(ib); free(jb); free(kb);
  free(cx); free(sx); free(cy); free(sy); free(cz); free(sz);
  return 0;
}

/* MMSCompare (mms.cc:497-541): df = |phi/ref - 1| (|phi| where ref < 1e-12);
 * result = {max, min, sum / n}. */
void snap_oracle_mms_compare(int n, const double *flux, const double *ref, double result[3]) {
  const double tolr = 1.0e-12;
  double mx = -INFINITY, mn = INFINITY, sum = 0.0;
  for (int i = 0; i < n; i++) {
    double r = ref[i], df = 1.0;
    if (r < tolr) {
      r = 1.0;
      df = 0.0;
    }
    df = fabs(flux[i] / r - df);
    if (df > mx) mx = df;
    if (df < mn) mn = df;
    sum += df;
  }
  result[0] = mx;
  result[1] = mn;
  result[2] = n > 0 ? sum / (double)n : 0.0;
}

/* Snap::transport_solve (snap.cc:386-541): time steps of the outer/inner
 * loop.  The flux is not reset between steps; the time-edge fluxes
 * alternate between two arrays (time_flux_even / _odd, snap.cc:288-298,
 * 497-498) that start at zero. */
int snap_oracle_solve(const oracle_problem_t *p, double *flux0, double *fluxm,
                      oracle_stats_t *stats, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  oracle_state_t s;
  oracle_stats_t st;
  memset(&st, 0, sizeof(st));
  if (state_init(&s, p)) {
    state_free(&s);
    return -1;
  }
  const int ng = s.ng, nc = s.ncells, cm = s.cmom, nsteps = problem_steps(p);
  const size_t gc = (size_t)ng * nc, ta = 8 * gc * (size_t)s.nang;
  double *flux0po = (double *)calloc(gc, sizeof(double));
  double *flux0pi = (double *)calloc(gc, sizeof(double));
  double *q2grp0 = (double *)calloc(gc, sizeof(double));
  double *qtot = (double *)calloc(gc * cm, sizeof(double));
  double *q2grpm = cm > 1 ? (double *)calloc(gc * (cm - 1), sizeof(double)) : NULL;
  double *fluxm_local = NULL;
  double *tflux[2] = {NULL, NULL}, *qim = NULL, *qi_mms = NULL, *ref = NULL;
  const double *qi = p->qi;
  if (cm > 1 && !fluxm) fluxm = fluxm_local = (double *)calloc(gc * (cm - 1), sizeof(double));
  if (p->timedep) {
    tflux[0] = (double *)calloc(ta, sizeof(double));
    tflux[1] = (double *)calloc(ta, sizeof(double));
  }
  int rc = 0;
  if (p->mms) {
    qim = (double *)malloc(sizeof(double) * ta);
    qi_mms = (double *)malloc(sizeof(double) * gc);
    ref = (double *)malloc(sizeof(double) * gc);
    rc = snap_oracle_mms_setup(p, ref, qim, qi_mms);
    qi = qi_mms;
    s.qim = qim;
  }
  memset(flux0, 0, sizeof(double) * gc);
  if (cm > 1) memset(fluxm, 0, sizeof(double) * gc * (cm - 1));

  const double t0 = now_seconds();
  int all_converged = 1;
  for (int cy = 0; cy < nsteps && rc == 0; cy++) {
    /* even steps read time_flux_even and write _odd; odd steps the reverse */
    if (p->timedep) {
      s.tflux_in = tflux[cy & 1];
      s.tflux_out = tflux[(cy & 1) ^ 1];
    }
    /* MMSScale (snap.cc:452-467): the manufactured source at the step's
     * mid-time, scaled in place step after step */
    if (p->mms) {
      const double f = (cy == 0) ? s.dt * 0.5 : (double)(2 * cy + 1) / (double)(2 * cy - 1);
      for (size_t i = 0; i < ta; i++) qim[i] *= f;
    }
    int inner_converged = 0, converged = 0;
    for (int otno = 0; otno < p->oitm; otno++) {
      /* CalcOuterSource, then save_fluxes flux0 -> flux0po (snap.cc:477-481) */
      snap_oracle_outer_source(nc, ng, p->nmat, s.nmom, p->mat, qi, p->slgg, flux0, fluxm,
                               q2grp0, q2grpm);
      memcpy(flux0po, flux0, sizeof(double) * gc);
      st.outer_total++;
      inner_converged = 0;
      for (int inno = 0; inno < p->iitm; inno++) {
        /* snap.cc:490-503 */
        snap_oracle_inner_source(nc, ng, s.nmom, s.s_xs, flux0, q2grp0, fluxm, q2grpm, qtot);
        memcpy(flux0pi, flux0, sizeof(double) * gc);
        const double ts = now_seconds();
        st.fixup_cells += perform_sweeps(&s, p->fixup, qtot, flux0, fluxm);
        st.sweep_seconds += now_seconds() - ts;
        st.inner_total++;
        inner_converged = snap_oracle_converged((int)gc, flux0, flux0pi, p->epsi);
        if (inner_converged) break;
      }
      /* snap.cc:519-538: the outer test is skipped on the first outer and
       * fails whenever the last inner did not converge (outer.cc:362-365). */
      if (otno == 0) continue;
      const int outer_converged =
          inner_converged && snap_oracle_converged((int)gc, flux0, flux0po, 100.0 * p->epsi);
      if (outer_converged) {
        converged = 1;
        break;
      }
    }
    all_converged &= converged;
    st.time_steps++;
  }
  st.converged = all_converged && rc == 0;
  st.solve_seconds = now_seconds() - t0;
  if (stats) *stats = st;
  free(flux0po);
  free(flux0pi);
  free(q2grp0);
  free(qtot);
  free(q2grpm);
  free(fluxm_local);
  free(tflux[0]);
  free(tflux[1]);
  free(qim);
  free(qi_mms);
  free(ref);
  state_free(&s);
  return rc;
}

double snap_oracle_time_sweeps(const oracle_problem_t *p, int ninner, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  if (p->timedep || p->mms) return -1.0; /* steady-state sweeps only */
  oracle_state_t s;
  if (state_init(&s, p)) {
    state_free(&s);
    return -1.0;
  }
  const int ng = s.ng, nc = s.ncells, cm = s.cmom;
  const size_t gc = (size_t)ng * nc;
  double *flux0 = (double *)calloc(gc, sizeof(double));
  double *fluxm = cm > 1 ? (double *)calloc(gc * (cm - 1), sizeof(double)) : NULL;
  double *qtot = (double *)calloc(gc * cm, sizeof(double));
  /* the first inner's source: q2grp0 = qi, qtot[0] = qi (flux is zero) */
  for (int g = 0; g < ng; g++)
    for (int c = 0; c < nc; c++) qtot[((size_t)g * nc + c) * cm] = p->qi[(size_t)g * nc + c];
  double total = 0.0;
  for (int it = 0; it < ninner; it++) {
    const double ts = now_seconds();
    perform_sweeps(&s, p->fixup, qtot, flux0, fluxm);
    total += now_seconds() - ts;
  }
  free(flux0);
  free(fluxm);
  free(qtot);
  state_free(&s);
  return total;
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the steady-state solve (no counterpart in the reference; this  */
/* is the FP64 reference gradient the GPU adjoint is checked against)       */
/* ------------------------------------------------------------------------ */

/* Per cell-angle state of a forward sweep that the reverse pass needs. */
enum { AF_OUTX = 1, AF_OUTY = 2, AF_OUTZ = 4, AF_PCX = 8, AF_PCY = 16, AF_PCZ = 32,
       AF_FIXED = 64, AF_ZERO = 128 };

/* Reverse-mode derivative of one (corner, group) sweep with isotropic
 * source qtot_g (nmom = 1).  The forward cell equations are
 *   pc = N / D,  N = q + sum_d h_d k_d psi_in,d,  D = t_xs + sum_d h_d v_d
 *   psi_out,d = (2 pc - psi_in,d) o_d
 * with h_d = (mu, eta, xi) * (hi, hj, hk); plain diamond difference has
 * k = v = o = 1 (D = 1/dinv); a fixed cell has k_d = (1 + v_d)/2 with the
 * pc-pass flags v_d, outflow masks o_d and pc = 0 when zeroed.  The fixup's
 * decisions are those of the forward state (held fixed).  The direct
 * response of pc is w_a zeta_g[c].  Adds sum_a dR/dq to u_g[c] and
 * sum_a dR/dt_xs to grad_g[c]. */
void snap_oracle_adjoint_sweep(int nx, int ny, int nz, int nang, int corner, double hi,
                               double hj, double hk, const double *mu, const double *eta,
                               const double *xi, const double *w, const double *qtot_g,
                               const double *t_xs_g, int fixup, const double *zeta_g,
                               double *u_g, double *grad_g) {
  const int sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
  const size_t ab = (size_t)nang, nc = (size_t)nx * ny * nz;
  double *pcs = (double *)malloc(sizeof(double) * nc * ab);
  double *dens = (double *)malloc(sizeof(double) * nc * ab);
  unsigned char *fl = (unsigned char *)malloc(nc * ab);
  double *fx = (double *)calloc(ab, sizeof(double));
  double *fyp = (double *)calloc(ab * nx, sizeof(double));
  double *fzp = (double *)calloc(ab * nx * ny, sizeof(double));
  double *pc = (double *)malloc(sizeof(double) * ab * 4);
  double *in = pc + ab; /* 3 x nang inflows */
  const double tolr = 1.0e-12;
  /* ---- forward pass in sweep order, as snap_oracle_sweep */
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const int px = sxp ? x : nx - 1 - x, py = syp ? y : ny - 1 - y, pz = szp ? z : nz - 1 - z;
        const size_t c = ((size_t)pz * ny + py) * nx + px;
        const size_t t = ((size_t)z * ny + y) * nx + x; /* sweep-order index */
        const double q = qtot_g[c], sig = t_xs_g[c];
        double *pi = in, *pj = in + ab, *pk = in + 2 * ab;
        for (int a = 0; a < nang; a++) {
          pi[a] = (x == 0) ? 0.0 : fx[a];
          pj[a] = (y == 0) ? 0.0 : fyp[(size_t)x * ab + a];
          pk[a] = (z == 0) ? 0.0 : fzp[((size_t)y * nx + x) * ab + a];
        }
        double hvx[256], hvy[256], hvz[256], pvx[256], pvy[256], pvz[256];
        int fixed = 0;
        for (int a = 0; a < nang; a++) {
          const double d = 1.0 / (((sig + hi * mu[a]) + hj * eta[a]) + hk * xi[a]);
          pc[a] = (((q + pi[a] * mu[a] * hi) + pj[a] * eta[a] * hj) + pk[a] * xi[a] * hk) * d;
          dens[t * ab + a] = 1.0 / d;
          hvx[a] = hvy[a] = hvz[a] = 1.0;
          pvx[a] = pvy[a] = pvz[a] = 1.0;
          fl[t * ab + a] = AF_OUTX | AF_OUTY | AF_OUTZ;
        }
        if (fixup) {
          unsigned old_neg = 0;
          for (;;) {
            unsigned neg = 0;
            for (int a = 0; a < nang; a++) {
              if (2.0 * pc[a] - pi[a] < 0.0) { hvx[a] = 0.0; neg++; }
              if (2.0 * pc[a] - pj[a] < 0.0) { hvy[a] = 0.0; neg++; }
              if (2.0 * pc[a] - pk[a] < 0.0) { hvz[a] = 0.0; neg++; }
            }
            if (neg == old_neg) break;
            old_neg = neg;
            fixed = 1;
            for (int a = 0; a < nang; a++) {
              pvx[a] = hvx[a];
              pvy[a] = hvy[a];
              pvz[a] = hvz[a];
              const double p = q + 0.5 * (pi[a] * mu[a] * hi * (1.0 + hvx[a]) +
                                          pj[a] * eta[a] * hj * (1.0 + hvy[a]) +
                                          pk[a] * xi[a] * hk * (1.0 + hvz[a]));
              const double den = (p <= 0.0) ? 0.0
                                             : (sig + mu[a] * hi * hvx[a] + eta[a] * hj * hvy[a] +
                                                xi[a] * hk * hvz[a]);
              dens[t * ab + a] = sig + mu[a] * hi * hvx[a] + eta[a] * hj * hvy[a] + xi[a] * hk * hvz[a];
              if (den < tolr) {
                pc[a] = 0.0;
                fl[t * ab + a] |= AF_ZERO;
              } else {
                pc[a] = p / den;
                fl[t * ab + a] &= (unsigned char)~AF_ZERO;
              }
            }
          }
        }
        for (int a = 0; a < nang; a++) {
          unsigned char f = fl[t * ab + a] & AF_ZERO;
          if (hvx[a] != 0.0) f |= AF_OUTX;
          if (hvy[a] != 0.0) f |= AF_OUTY;
          if (hvz[a] != 0.0) f |= AF_OUTZ;
          if (fixed) {
            f |= AF_FIXED;
            if (pvx[a] != 0.0) f |= AF_PCX;
            if (pvy[a] != 0.0) f |= AF_PCY;
            if (pvz[a] != 0.0) f |= AF_PCZ;
          }
          fl[t * ab + a] = f;
          pcs[t * ab + a] = pc[a];
          fx[a] = (2.0 * pc[a] - pi[a]) * hvx[a];
          fyp[(size_t)x * ab + a] = (2.0 * pc[a] - pj[a]) * hvy[a];
          fzp[((size_t)y * nx + x) * ab + a] = (2.0 * pc[a] - pk[a]) * hvz[a];
        }
      }
  /* ---- reverse pass: adjoint faces flow against the sweep direction */
  memset(fx, 0, sizeof(double) * ab);
  memset(fyp, 0, sizeof(double) * ab * nx);
  memset(fzp, 0, sizeof(double) * ab * nx * ny);
  for (int z = nz - 1; z >= 0; z--)
    for (int y = ny - 1; y >= 0; y--)
      for (int x = nx - 1; x >= 0; x--) {
        const int px = sxp ? x : nx - 1 - x, py = syp ? y : ny - 1 - y, pz = szp ? z : nz - 1 - z;
        const size_t c = ((size_t)pz * ny + py) * nx + px;
        const size_t t = ((size_t)z * ny + y) * nx + x;
        double su = 0.0, sg = 0.0;
        for (int a = 0; a < nang; a++) {
          const unsigned f = fl[t * ab + a];
          /* adjoints of this cell's outflows: the downstream cells' inflow
           * adjoints (zero past the far boundary) */
          const double bx = (x == nx - 1) ? 0.0 : fx[a];
          const double by = (y == ny - 1) ? 0.0 : fyp[(size_t)x * ab + a];
          const double bz = (z == nz - 1) ? 0.0 : fzp[((size_t)y * nx + x) * ab + a];
          const double ox = (f & AF_OUTX) ? 1.0 : 0.0, oy = (f & AF_OUTY) ? 1.0 : 0.0,
                       oz = (f & AF_OUTZ) ? 1.0 : 0.0;
          const double bpc = w[a] * zeta_g[c] + 2.0 * (ox * bx + oy * by + oz * bz);
          double bn = 0.0, bd = 0.0;
          if (!(f & AF_ZERO)) {
            const double den = dens[t * ab + a];
            bn = bpc / den;
            bd = -bpc * pcs[t * ab + a] / den;
          }
          double kx = 1.0, ky = 1.0, kz = 1.0;
          if (f & AF_FIXED) {
            kx = (f & AF_PCX) ? 1.0 : 0.5;
            ky = (f & AF_PCY) ? 1.0 : 0.5;
            kz = (f & AF_PCZ) ? 1.0 : 0.5;
          }
          fx[a] = -ox * bx + bn * mu[a] * hi * kx;
          fyp[(size_t)x * ab + a] = -oy * by + bn * eta[a] * hj * ky;
          fzp[((size_t)y * nx + x) * ab + a] = -oz * bz + bn * xi[a] * hk * kz;
          su += bn;
          sg += bd;
        }
        u_g[c] += su;
        grad_g[c] += sg;
      }
  free(pcs);
  free(dens);
  free(fl);
  free(fx);
  free(fyp);
  free(fzp);
  free(pc);
}

/* Gradient of R = sum_g sum_c resp[g][c] * flux0[g][c] at the converged
 * solution flux0 of problem p (nmom = 1) with respect to t_xs[g][c].
 * Adjoint source iteration zeta = resp + S^T u(zeta), u = sum over the 8*ng
 * adjoint sweeps, until max|dzeta| <= tol * max|zeta|.  Returns the number
 * of adjoint iterations, or -1. */
int snap_oracle_adjoint(const oracle_problem_t *p, const double *flux0, const double *resp,
                        double tol, int maxit, double *grad, double *zeta_out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  if (p->nmom != 1 || p->nang > 256 || p->timedep || p->mms) return -1;
  oracle_state_t s;
  if (state_init(&s, p)) {
    state_free(&s);
    return -1;
  }
  const int ng = s.ng, nc = s.ncells;
  const size_t gc = (size_t)ng * nc;
  double *q2 = (double *)calloc(gc, sizeof(double));
  double *qtot = (double *)calloc(gc, sizeof(double));
  double *zeta = (double *)malloc(sizeof(double) * gc);
  double *zn = (double *)malloc(sizeof(double) * gc);
  double *u = (double *)malloc(sizeof(double) * 8 * gc);
  double *gp = (double *)malloc(sizeof(double) * 8 * gc);
  /* the forward source of the converged flux: q = qi + S flux0 */
  snap_oracle_outer_source(nc, ng, p->nmat, 1, p->mat, p->qi, p->slgg, flux0, NULL, q2, NULL);
  snap_oracle_inner_source(nc, ng, 1, s.s_xs, flux0, q2, NULL, NULL, qtot);
  memcpy(zeta, resp, sizeof(double) * gc);
  int it = 0, done = 0;
  for (; it < maxit && !done; it++) {
    memset(u, 0, sizeof(double) * 8 * gc);
    memset(gp, 0, sizeof(double) * 8 * gc);
#pragma omp parallel for schedule(dynamic, 1)
    for (int task = 0; task < 8 * ng; task++) {
      const int corner = task / ng, g = task % ng;
      snap_oracle_adjoint_sweep(s.nx, s.ny, s.nz, s.nang, corner, s.hi, s.hj, s.hk, s.mu, s.eta,
                                s.xi, s.w, qtot + (size_t)g * nc, s.t_xs + (size_t)g * nc,
                                p->fixup, zeta + (size_t)g * nc,
                                u + (size_t)corner * gc + (size_t)g * nc,
                                gp + (size_t)corner * gc + (size_t)g * nc);
    }
    /* octant partials in order; zeta_new[g2] = resp[g2] + sum_g S[g][mat][g2] u[g] */
    double dmax = 0.0, zmax = 0.0;
    for (int g2 = 0; g2 < ng; g2++)
      for (int c = 0; c < nc; c++) {
        const int mt = p->mat[c];
        double acc = resp[(size_t)g2 * nc + c];
        for (int g = 0; g < ng; g++) {
          double ug = 0.0;
          for (int o = 0; o < 8; o++) ug += u[(size_t)o * gc + (size_t)g * nc + c];
          acc += p->slgg[(((size_t)g * p->nmat + mt) * ng + g2)] * ug;
        }
        zn[(size_t)g2 * nc + c] = acc;
        const double d = fabs(acc - zeta[(size_t)g2 * nc + c]);
        if (d > dmax) dmax = d;
        if (fabs(acc) > zmax) zmax = fabs(acc);
      }
    for (size_t i = 0; i < gc; i++) {
      double acc = 0.0;
      for (int o = 0; o < 8; o++) acc += gp[(size_t)o * gc + i];
      grad[i] = acc;
    }
    done = dmax <= tol * zmax;
    double *tmp = zeta;
    zeta = zn;
    zn = tmp;
  }
  /* grad holds the sweep with the last zeta before the update; one more
   * sweep evaluates it at the converged zeta */
  memset(gp, 0, sizeof(double) * 8 * gc);
  memset(u, 0, sizeof(double) * 8 * gc);
#pragma omp parallel for schedule(dynamic, 1)
  for (int task = 0; task < 8 * ng; task++) {
    const int corner = task / ng, g = task % ng;
    snap_oracle_adjoint_sweep(s.nx, s.ny, s.nz, s.nang, corner, s.hi, s.hj, s.hk, s.mu, s.eta,
                              s.xi, s.w, qtot + (size_t)g * nc, s.t_xs + (size_t)g * nc, p->fixup,
                              zeta + (size_t)g * nc, u + (size_t)corner * gc + (size_t)g * nc,
                              gp + (size_t)corner * gc + (size_t)g * nc);
  }
  for (size_t i = 0; i < gc; i++) {
    double acc = 0.0;
    for (int o = 0; o < 8; o++) acc += gp[(size_t)o * gc + i];
    grad[i] = acc;
  }
  if (zeta_out) memcpy(zeta_out, zeta, sizeof(double) * gc);
  free(q2);
  free(qtot);
  free(zeta);
  free(zn);
  free(u);
  free(gp);
  state_free(&s);
  return done ? it : -1;
}
