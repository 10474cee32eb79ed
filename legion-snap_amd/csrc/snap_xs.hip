// snap_xs.hip -- cross-section expansion and the geometry parameter dinv.
//
// Element-wise gathers over cells x groups (x angles for dinv), HBM bound.
// One thread per output element with the fastest index (angle, then x) on
// consecutive lanes, so dense-layout stores are fully coalesced.
#include "snap_internal.h"

namespace snap {

// ExpandCrossSection (expxs.cc:85-115; GPU gpu_expxs.cu:21-44):
// xs[g][c] = sig[g][mat[c]].
__global__ void k_expand_xs(Acc2 sig, Acc3 mat, Acc3 xs, Extent3 e, int ng) {
  const int64_t cells = e.cells();
  const int64_t n = cells * ng;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t % cells;
    const int g = (int)(t / cells);
    const int x = (int)(c % e.nx), y = (int)((c / e.nx) % e.ny), z = (int)(c / ((int64_t)e.nx * e.ny));
    const int m = *(const int32_t *)mat.at(x, y, z, 0);
    *(double *)xs.at(x, y, z, g) = *(const double *)sig.at(m, 0, g);
  }
}

int launch_expand_xs(Acc2 sig, Acc3 mat, Acc3 xs, Extent3 e, int ng, hipStream_t st) {
  const int64_t n = e.cells() * ng;
  hipLaunchKernelGGL(k_expand_xs, dim3(grid_for(n, 256)), dim3(256), 0, st, sig, mat, xs, e, ng);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// ExpandScatteringCrossSection (expxs.cc:220-250; GPU gpu_expxs.cu:268-286):
// xs[g][c] = slgg[g][mat[c]][group_start + g], a record of `rec` doubles.
__global__ void k_expand_scat(Acc2 slgg, int group_start, Acc3 mat, Acc3 xs, Extent3 e, int ng,
                              int rec) {
  const int64_t cells = e.cells();
  const int64_t n = cells * ng * rec;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % rec);
    const int64_t r = t / rec;
    const int64_t c = r % cells;
    const int g = (int)(r / cells);
    const int x = (int)(c % e.nx), y = (int)((c / e.nx) % e.ny), z = (int)(c / ((int64_t)e.nx * e.ny));
    const int m = *(const int32_t *)mat.at(x, y, z, 0);
    ((double *)xs.at(x, y, z, g))[l] = ((const double *)slgg.at(m, group_start + g, g))[l];
  }
}

int launch_expand_scat(Acc2 slgg, int group_start, Acc3 mat, Acc3 xs, Extent3 e, int ng,
                       int rec, hipStream_t st) {
  const int64_t n = e.cells() * ng * rec;
  hipLaunchKernelGGL(k_expand_scat, dim3(grid_for(n, 256)), dim3(256), 0, st, slgg, group_start,
                     mat, xs, e, ng, rec);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// CalculateGeometryParam (expxs.cc:358-392; GPU gpu_sweep.cu:276-311):
// dinv[g][c][a] = 1/(xs + vdelt + hi*mu + hj*eta + hk*xi), summed left to
// right as the CPU variant does, in FP64; the FP32 form rounds the FP64 value.
template <typename OutT>
__global__ void k_geometry(Acc3 xs, Acc3 dinv, double vd, Extent3 e, int g0, int ng, int nang,
                           double hi, double hj, double hk, const double *__restrict__ mu,
                           const double *__restrict__ eta, const double *__restrict__ xi) {
  const int64_t cells = e.cells();
  const int64_t n = cells * ng * nang;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(t % nang);
    const int64_t r = t / nang;
    const int64_t c = r % cells;
    const int g = g0 + (int)(r / cells);
    const int x = (int)(c % e.nx), y = (int)((c / e.nx) % e.ny), z = (int)(c / ((int64_t)e.nx * e.ny));
    const double xsv = *(const double *)xs.at(x, y, z, g);
    const double v = 1.0 / (xsv + vd + hi * mu[a] + hj * eta[a] + hk * xi[a]);
    ((OutT *)dinv.at(x, y, z, g))[a] = (OutT)v;
  }
}

int launch_geometry(Acc3 xs, Acc3 dinv, const double *vdelt_host, Extent3 e, int ng, int nang,
                    double hi, double hj, double hk, const QuadDev &q, bool fp32_out,
                    hipStream_t st) {
  if (q.nang != nang)
    return fail(SNAP_ERR_INVALID, "geometry: context has %d angles, call has %d", q.nang, nang);
  bool any_vdelt = false;
  if (vdelt_host)
    for (int g = 0; g < ng; g++) any_vdelt |= (vdelt_host[g] != 0.0);
  // one launch for all groups when vdelt is zero (steady state), else one per group
  const int chunks = any_vdelt ? ng : 1;
  const int per = any_vdelt ? 1 : ng;
  for (int ch = 0; ch < chunks; ch++) {
    const double vd = any_vdelt ? vdelt_host[ch] : 0.0;
    const int64_t n = e.cells() * per * nang;
    if (fp32_out)
      hipLaunchKernelGGL(k_geometry<float>, dim3(grid_for(n, 256)), dim3(256), 0, st, xs, dinv, vd,
                         e, ch * per, per, nang, hi, hj, hk, q.mu, q.eta, q.xi);
    else
      hipLaunchKernelGGL(k_geometry<double>, dim3(grid_for(n, 256)), dim3(256), 0, st, xs, dinv,
                         vd, e, ch * per, per, nang, hi, hj, hk, q.mu, q.eta, q.xi);
    SNAP_HIP_TRY(hipGetLastError());
  }
  return SNAP_OK;
}

}  // namespace snap
