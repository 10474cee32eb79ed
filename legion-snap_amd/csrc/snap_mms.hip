// snap_mms.hip -- the method of manufactured solutions (reference: mms.cc,
// driven from snap.cc:372-385, 452-467, 542-559) for the solver.
//
// The manufactured scalar flux of group g (1-based) is g * X(i) Y(j) Z(k),
// where X(i) is the average of sin(pi x / lx) over cell i (MMSInitFlux,
// mms.cc:81-148).  The angular source that makes it the exact solution of
// the transport equation is, per corner and angle, the streaming terms
// (+-mu) g dX Y Z + ..., the collision term sigt * phi and minus the
// scattering into g (MMSInitSource, mms.cc:206-344).  Time dependence makes
// the flux linear in time: qi = phi / v_g and the reference flux is taken at
// tf - dt/2 (MMSInitTimeDependent, mms.cc:382-409), and the angular source is
// rescaled at every step to its mid-time (MMSScale, snap.cc:452-467).
//
// The per-axis averages are small host tables; the angular source
// [8][ng][cells][nang] is formed on the device, one thread per element, in
// the reference's operation order (the library is built without contraction).
#include <cmath>
#include <cstring>

#include "snap_internal.h"

namespace snap {

namespace {
constexpr double kMmsPi = 3.14159265358979;  // the reference's PI (snap.h:37)

// mms_trigint (mms.cc:34-57): the cell average of sin(d x) (cos form) or of
// d cos(d x) (sin form) between the edges cb[i] and cb[i+1]
void trigint(bool cos_form, int lc, double d, double del, const double *cb, double *fn) {
  const double denom = d * del;
  for (int i = 0; i < lc; i++) {
    if (cos_form) {
      fn[i] = std::cos(d * cb[i]) - std::cos(d * cb[i + 1]);
      fn[i] /= denom;
    } else {
      fn[i] = std::sin(d * cb[i + 1]) - std::sin(d * cb[i]);
      fn[i] /= del;
    }
  }
}

// cell edges from 0 in steps of del, accumulated as mms.cc:108-117 does
void edges(int n, double del, std::vector<double> &b) {
  b.assign(n + 1, 0.0);
  for (int i = 1; i <= n; i++) b[i] = b[i - 1] + del;
}
}  // namespace

__global__ void k_mms_source(const double *__restrict__ tab, Extent3 e, int ng, int nang,
                             int nmat, int nmom, const double *__restrict__ mu,
                             const double *__restrict__ eta, const double *__restrict__ xi,
                             const int32_t *__restrict__ mat, const double *__restrict__ sigt,
                             const double *__restrict__ slgg, const double *__restrict__ ref,
                             double *__restrict__ qim) {
  const int64_t cells = e.cells();
  const int64_t n = 8 * (int64_t)ng * cells * nang;
  const double *cx = tab, *sx = cx + e.nx, *cy = sx + e.nx, *sy = cy + e.ny, *cz = sy + e.ny,
               *sz = cz + e.nz;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(t % nang);
    int64_t r = t / nang;
    const int64_t c = r % cells;
    r /= cells;
    const int g = (int)(r % ng);
    const int corner = (int)(r / ng);
    const int i = (int)(c % e.nx), j = (int)((c / e.nx) % e.ny), k = (int)(c / ((int64_t)e.nx * e.ny));
    const double is = (corner & 1) ? 1.0 : -1.0, js = (corner & 2) ? 1.0 : -1.0,
                 ks = (corner & 4) ? 1.0 : -1.0;
    const int m = mat[c];
    const double gg = (double)(g + 1);
    const double flux_update = sigt[(size_t)m * ng + g] * ref[(size_t)g * cells + c];
    double v = 0.0;
    v += gg * is * mu[a] * sx[i] * cy[j] * cz[k];
    v += flux_update;
    v += gg * js * eta[a] * cx[i] * sy[j] * cz[k];
    v += gg * ks * xi[a] * cx[i] * cy[j] * sz[k];
    const double *sl = slgg + (((size_t)g * nmat + m) * ng) * nmom;
    for (int gp = 0; gp < ng; gp++) v -= sl[(size_t)gp * nmom] * ref[(size_t)gp * cells + c];
    qim[t] = v;
  }
}

__global__ void k_scale(double *__restrict__ x, int64_t n, double f) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    x[t] *= f;
}

int mms_prepare(const snap_problem_t &p, MmsState *m) {
  if (p.nmom != 1)
    return fail(SNAP_ERR_UNSUPPORTED, "the manufactured solution is defined for nmom = 1");
  if (p.sigt_cell)
    return fail(SNAP_ERR_UNSUPPORTED, "the manufactured solution uses material cross sections");
  const int nx = p.nx, ny = p.ny, nz = p.nz, ng = p.ng;
  const size_t nc = (size_t)nx * ny * nz;
  const double dx = p.lx / (double)nx, dy = p.ly / (double)ny, dz = p.lz / (double)nz;
  std::vector<double> ib, jb, kb;
  edges(nx, dx, ib);
  edges(ny, dy, jb);
  edges(nz, dz, kb);
  std::vector<double> tab(2 * (size_t)(nx + ny + nz));
  double *cx = tab.data(), *sx = cx + nx, *cy = sx + nx, *sy = cy + ny, *cz = sy + ny, *sz = cz + nz;
  trigint(true, nx, kMmsPi / p.lx, dx, ib.data(), cx);
  trigint(false, nx, kMmsPi / p.lx, dx, ib.data(), sx);
  trigint(true, ny, kMmsPi / p.ly, dy, jb.data(), cy);
  trigint(false, ny, kMmsPi / p.ly, dy, jb.data(), sy);
  trigint(true, nz, kMmsPi / p.lz, dz, kb.data(), cz);
  trigint(false, nz, kMmsPi / p.lz, dz, kb.data(), sz);
  m->ref.assign((size_t)ng * nc, 0.0);
  for (int g = 0; g < ng; g++)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          m->ref[(size_t)g * nc + ((size_t)k * ny + j) * nx + i] = (double)(g + 1) * cx[i] * cy[j] * cz[k];
  SNAP_HIP_TRY(hipMalloc((void **)&m->tab, sizeof(double) * tab.size()));
  SNAP_HIP_TRY(hipMalloc((void **)&m->ref_dev, sizeof(double) * m->ref.size()));
  SNAP_HIP_TRY(hipMemcpy(m->tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  SNAP_HIP_TRY(hipMemcpy(m->ref_dev, m->ref.data(), sizeof(double) * m->ref.size(),
                         hipMemcpyHostToDevice));
  m->qi.assign((size_t)ng * nc, 0.0);
  if (p.timedep) {
    const int steps = p.nsteps > 0 ? p.nsteps : 1;
    const double dt = p.tf / (double)steps;
    const double t_scale = p.tf - 0.5 * dt;
    for (int g = 0; g < ng; g++) {
      const double vg = (double)(ng - g);
      for (size_t c = 0; c < nc; c++) {
        const double f = m->ref[(size_t)g * nc + c];
        m->qi[(size_t)g * nc + c] = f / vg;
        m->ref[(size_t)g * nc + c] = f * t_scale;
      }
    }
  }
  return SNAP_OK;
}

void mms_free(MmsState *m) {
  if (m->tab) (void)hipFree(m->tab);
  if (m->ref_dev) (void)hipFree(m->ref_dev);
  m->tab = m->ref_dev = nullptr;
}

int mms_source(const MmsState &m, const QuadDev &q, Extent3 e, int ng, int nmat, int nmom,
               const int32_t *mat, const double *sigt, const double *slgg, double *qim,
               hipStream_t st) {
  const int64_t n = 8 * (int64_t)ng * e.cells() * q.nang;
  hipLaunchKernelGGL(k_mms_source, dim3(grid_for(n, 256)), dim3(256), 0, st, m.tab, e, ng,
                     q.nang, nmat, nmom, q.mu, q.eta, q.xi, mat, sigt, slgg, m.ref_dev, qim);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

int launch_scale(double *x, int64_t n, double f, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n, 256)), dim3(256), 0, st, x, n, f);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

void mms_compare(const double *flux, const double *ref, size_t n, double result[3]) {
  const double tolr = 1.0e-12;  // mms.cc:510
  double mx = -INFINITY, mn = INFINITY, sum = 0.0;
  for (size_t i = 0; i < n; i++) {
    double r = ref[i], df = 1.0;
    if (r < tolr) {
      r = 1.0;
      df = 0.0;
    }
    df = std::fabs(flux[i] / r - df);
    if (df > mx) mx = df;
    if (df < mn) mn = df;
    sum += df;
  }
  result[0] = mx;
  result[1] = mn;
  result[2] = n ? sum / (double)n : 0.0;
}

}  // namespace snap
