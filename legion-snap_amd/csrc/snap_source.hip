// snap_source.hip -- outer (group-to-group) and inner sources, convergence
// tests and the octant reduction of the sweep's partial fluxes.
//
// Every sum runs in the order the reference's CPU task bodies use (ascending
// source group, skipping the in-group term), so the FP64 results equal the
// CPU arithmetic operation for operation.
#include "snap_internal.h"

namespace snap {

struct OrderTable {
  int8_t order[16];  // Legendre order of flux moment m (m = 1..cmom-1)
};

static OrderTable make_order_table(const int *lma, int nmom) {
  OrderTable t{};
  int m = 1;
  for (int l = 1; l < nmom; l++)
    for (int j = 0; j < lma[l] && m < 16; j++) t.order[m++] = (int8_t)l;
  return t;
}

__device__ inline void cell_xyz(int64_t c, const Extent3 &e, int &x, int &y, int &z) {
  x = (int)(c % e.nx);
  y = (int)((c / e.nx) % e.ny);
  z = (int)(c / ((int64_t)e.nx * e.ny));
}

// CalcOuterSource, scalar part (outer.cc:146-171; GPU gpu_outer.cu:22-74):
// qo0[g1] = qi[g1] + sum_{g2 != g1} slgg[g1][mat][g2].l0 * flux0[g2].
__global__ void k_outer_source(Acc3 qi, Acc3 flux0, Acc2 slgg, Acc3 qo0, Acc3 mat, Extent3 e,
                               int ng) {
  const int64_t cells = e.cells();
  const int64_t n = cells * ng;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t % cells;
    const int g1 = (int)(t / cells);
    int x, y, z;
    cell_xyz(c, e, x, y, z);
    const int m = *(const int32_t *)mat.at(x, y, z, 0);
    double q = *(const double *)qi.at(x, y, z, g1);
    for (int g2 = 0; g2 < ng; g2++) {
      if (g2 == g1) continue;
      const double cs = *(const double *)slgg.at(m, g2, g1);
      q += cs * *(const double *)flux0.at(x, y, z, g2);
    }
    *(double *)qo0.at(x, y, z, g1) = q;
  }
}

int launch_outer_source(Acc3 qi, Acc3 flux0, Acc2 slgg, Acc3 qo0, Acc3 mat, Extent3 e, int ng,
                        hipStream_t st) {
  const int64_t n = e.cells() * ng;
  hipLaunchKernelGGL(k_outer_source, dim3(grid_for(n, 256)), dim3(256), 0, st, qi, flux0, slgg,
                     qo0, mat, e, ng);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// CalcOuterSource, moment part (outer.cc:175-214; GPU gpu_outer.cu:225-293),
// over all cmom-1 moments: qom[g1][m-1] = sum_{g2 != g1} slgg[g1][mat][g2].L(m) * fluxm[g2][m-1].
__global__ void k_outer_source_m(Acc3 fluxm, Acc2 slgg, Acc3 qom, Acc3 mat, Extent3 e, int ng,
                                 int cmom, OrderTable ord) {
  const int64_t cells = e.cells();
  const int nm = cmom - 1;
  const int64_t n = cells * ng * nm;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int mi = (int)(t % nm);
    const int64_t r = t / nm;
    const int64_t c = r % cells;
    const int g1 = (int)(r / cells);
    int x, y, z;
    cell_xyz(c, e, x, y, z);
    const int m = *(const int32_t *)mat.at(x, y, z, 0);
    const int l = ord.order[mi + 1];
    double acc = 0.0;
    for (int g2 = 0; g2 < ng; g2++) {
      if (g2 == g1) continue;
      const double cs = ((const double *)slgg.at(m, g2, g1))[l];
      acc += cs * ((const double *)fluxm.at(x, y, z, g2))[mi];
    }
    ((double *)qom.at(x, y, z, g1))[mi] = acc;
  }
}

int launch_outer_source_m(Acc3 fluxm, Acc2 slgg, Acc3 qom, Acc3 mat, Extent3 e, int ng, int cmom,
                          const int *lma, int nmom, hipStream_t st) {
  if (cmom <= 1) return SNAP_OK;
  if (cmom > 16) return fail(SNAP_ERR_INVALID, "cmom %d > 16", cmom);
  const OrderTable ord = make_order_table(lma, nmom);
  const int64_t n = e.cells() * ng * (cmom - 1);
  hipLaunchKernelGGL(k_outer_source_m, dim3(grid_for(n, 256)), dim3(256), 0, st, fluxm, slgg,
                     qom, mat, e, ng, cmom, ord);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// CalcInnerSource (inner.cc:106-160; GPU gpu_inner.cu:22-95):
// qtot[0] = q2grp0 + flux0*s_xs[0]; qtot[m] = q2grpm[m-1] + fluxm[m-1]*s_xs[L(m)].
__global__ void k_inner_source(Acc3 sxs, Acc3 flux0, Acc3 q2grp0, Acc3 fluxm, Acc3 q2grpm,
                               Acc3 qtot, Extent3 e, int ng, int cmom, OrderTable ord) {
  const int64_t cells = e.cells();
  const int64_t n = cells * ng * cmom;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(t % cmom);
    const int64_t r = t / cmom;
    const int64_t c = r % cells;
    const int g = (int)(r / cells);
    int x, y, z;
    cell_xyz(c, e, x, y, z);
    const double *sx = (const double *)sxs.at(x, y, z, g);
    double v;
    if (m == 0) {
      v = *(const double *)q2grp0.at(x, y, z, g) + *(const double *)flux0.at(x, y, z, g) * sx[0];
    } else {
      v = ((const double *)q2grpm.at(x, y, z, g))[m - 1] +
          ((const double *)fluxm.at(x, y, z, g))[m - 1] * sx[ord.order[m]];
    }
    ((double *)qtot.at(x, y, z, g))[m] = v;
  }
}

int launch_inner_source(Acc3 sxs, Acc3 flux0, Acc3 q2grp0, Acc3 fluxm, Acc3 q2grpm, Acc3 qtot,
                        Extent3 e, int ng, int cmom, const int *lma, int nmom, hipStream_t st) {
  if (cmom > 16) return fail(SNAP_ERR_INVALID, "cmom %d > 16", cmom);
  const OrderTable ord = make_order_table(lma, nmom);
  const int64_t n = e.cells() * ng * cmom;
  hipLaunchKernelGGL(k_inner_source, dim3(grid_for(n, 256)), dim3(256), 0, st, sxs, flux0,
                     q2grp0, fluxm, q2grpm, qtot, e, ng, cmom, ord);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// The same sum over the solver's dense arrays (sxs [cg][nmom], flux0 and
// q2grp0 [cg], fluxm and q2grpm [cg][cmom-1], qtot [cg][cmom] with
// cg = g * cells + cell): one thread per qtot element in memory order, the
// flat index split by the compile-time moment count instead of six 64-bit
// divisions per element.
template <int CMOM, typename Index>
__global__ void k_inner_source_dense(const double *__restrict__ sxs, const double *__restrict__ flux0,
                                     const double *__restrict__ q2grp0,
                                     const double *__restrict__ fluxm,
                                     const double *__restrict__ q2grpm, double *__restrict__ qtot,
                                     Index n, int nmom, OrderTable ord, const int32_t *skip) {
  constexpr int NM = CMOM - 1;
  if (skip && *skip) return;
  for (Index t = (Index)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (Index)gridDim.x * blockDim.x) {
    const Index cg = t / CMOM;
    const int m = (int)(t - cg * CMOM);
    const double *sx = sxs + cg * nmom;
    double v;
    if (m == 0)
      v = q2grp0[cg] + flux0[cg] * sx[0];
    else
      v = q2grpm[cg * NM + (m - 1)] + fluxm[cg * NM + (m - 1)] * sx[ord.order[m]];
    qtot[t] = v;
  }
}

template <int CMOM>
static void enqueue_inner_dense(const double *sxs, const double *flux0, const double *q2grp0,
                                const double *fluxm, const double *q2grpm, double *qtot, int64_t n,
                                int nmom, const OrderTable &ord, hipStream_t st,
                                const int32_t *skip) {
  const dim3 grid(grid_for(n, 256)), block(256);
  if (n <= (int64_t)UINT32_MAX - 65536 * 8 * 256)
    hipLaunchKernelGGL((k_inner_source_dense<CMOM, uint32_t>), grid, block, 0, st, sxs, flux0,
                       q2grp0, fluxm, q2grpm, qtot, (uint32_t)n, nmom, ord, skip);
  else
    hipLaunchKernelGGL((k_inner_source_dense<CMOM, int64_t>), grid, block, 0, st, sxs, flux0,
                       q2grp0, fluxm, q2grpm, qtot, n, nmom, ord, skip);
}

int launch_inner_source_dense(const double *sxs, const double *flux0, const double *q2grp0,
                              const double *fluxm, const double *q2grpm, double *qtot, int64_t cg,
                              int nmom, hipStream_t st, const int32_t *skip) {
  const int cmom = nmom * nmom;
  if (nmom < 1 || nmom > 4) return fail(SNAP_ERR_INVALID, "nmom %d outside 1..4", nmom);
  if (cmom > 1 && (!fluxm || !q2grpm)) return fail(SNAP_ERR_INVALID, "moment arrays missing");
  const int lma[4] = {1, 3, 5, 7};
  const OrderTable ord = make_order_table(lma, nmom);
  const int64_t n = cg * cmom;
  if (n <= 0) return SNAP_OK;
  switch (cmom) {
    case 1: enqueue_inner_dense<1>(sxs, flux0, q2grp0, fluxm, q2grpm, qtot, n, nmom, ord, st, skip); break;
    case 4: enqueue_inner_dense<4>(sxs, flux0, q2grp0, fluxm, q2grpm, qtot, n, nmom, ord, st, skip); break;
    case 9: enqueue_inner_dense<9>(sxs, flux0, q2grp0, fluxm, q2grpm, qtot, n, nmom, ord, st, skip); break;
    default: enqueue_inner_dense<16>(sxs, flux0, q2grp0, fluxm, q2grpm, qtot, n, nmom, ord, st, skip); break;
  }
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// TestInnerConvergence / TestOuterConvergence (inner.cc:283-324,
// outer.cc:354-403; GPU gpu_inner.cu:119-204, gpu_outer.cu:443-531):
// a cell fails when |flux0/flux0p - 1| > epsi (|flux0| when |flux0p| < 1e-12).
// *result must hold 1 on entry; any failing block clears it.
__global__ void k_convergence(Acc3 flux0, Acc3 flux0p, Extent3 e, int ng, double epsi,
                              int32_t *result, const int32_t *skip) {
  if (skip && *skip) return;
  const int64_t cells = e.cells();
  const int64_t n = cells * ng;
  int bad = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t % cells;
    const int g = (int)(t / cells);
    int x, y, z;
    cell_xyz(c, e, x, y, z);
    double fp = *(const double *)flux0p.at(x, y, z, g);
    double df = 1.0;
    if (fabs(fp) < 1.0e-12) {
      fp = 1.0;
      df = 0.0;
    }
    df = fabs((*(const double *)flux0.at(x, y, z, g) / fp) - df);
    bad |= (df > epsi) ? 1 : 0;
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicAnd(result, 0);
}

__global__ void k_fill_int(int32_t *p, int32_t v, const int32_t *skip) {
  if (skip && *skip) return;
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

int launch_fill_int(int32_t *p, int32_t v, hipStream_t st, const int32_t *skip) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, p, v, skip);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

int launch_convergence(Acc3 flux0, Acc3 flux0p, Extent3 e, int ng, double epsi, int32_t *result,
                       hipStream_t st, const int32_t *skip) {
  SNAP_TRY(launch_fill_int(result, 1, st, skip));
  const int64_t n = e.cells() * ng;
  hipLaunchKernelGGL(k_convergence, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, flux0,
                     flux0p, e, ng, epsi, result, skip);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// Largest relative change max |flux0/flux0p - 1| of the convergence test's
// quantity (|flux0| where |flux0p| < 1e-12), as the bit pattern of the
// non-negative double (bit patterns of non-negative doubles order like the
// values).  *out must hold 0 on entry.
__global__ void k_max_change(const double *__restrict__ f, const double *__restrict__ fp,
                             int64_t n, unsigned long long *out) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double p = fp[i], d = 1.0;
    if (fabs(p) < 1.0e-12) {
      p = 1.0;
      d = 0.0;
    }
    const double v = fabs(f[i] / p - d);
    m = (v > m || v != v) ? v : m;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double v = __shfl_xor(m, o, 64);
    m = (v > m || v != v) ? v : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(fabs(m)));
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

int launch_max_change(const double *f, const double *fp, int64_t n, unsigned long long *out,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, out, 0ull);
  hipLaunchKernelGGL(k_max_change, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, f, fp, n, out);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// Sum of the eight octant partials in octant order 0..7.
__global__ void k_octant_reduce(const double *__restrict__ part, double *__restrict__ out,
                                const double *__restrict__ base, int64_t n, const int32_t *skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
#pragma unroll
    for (int o = 0; o < 8; o++) acc += part[(int64_t)o * n + i];
    out[i] = base ? base[i] + acc : acc;
  }
}

int launch_octant_reduce(const double *part, double *out, int64_t n, hipStream_t st,
                         const double *base, const int32_t *skip) {
  if (n <= 0) return SNAP_OK;
  hipLaunchKernelGGL(k_octant_reduce, dim3(grid_for(n, 256)), dim3(256), 0, st, part, out, base, n,
                     skip);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

__global__ void k_copy_doubles(double *__restrict__ dst, const double *__restrict__ src, int64_t n,
                               const int32_t *skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int launch_copy(double *dst, const double *src, int64_t n, hipStream_t st, const int32_t *skip) {
  if (n <= 0) return SNAP_OK;
  hipLaunchKernelGGL(k_copy_doubles, dim3(grid_for(n, 256)), dim3(256), 0, st, dst, src, n, skip);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

__global__ void k_subtract(double *__restrict__ y, const double *__restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] - x[i];
}

int launch_subtract(double *y, const double *x, int64_t n, hipStream_t st) {
  if (n <= 0) return SNAP_OK;
  hipLaunchKernelGGL(k_subtract, dim3(grid_for(n, 256)), dim3(256), 0, st, y, x, n);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

}  // namespace snap
