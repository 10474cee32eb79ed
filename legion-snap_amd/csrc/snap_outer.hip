// snap_outer.hip -- the outer (group-to-group) scattering source of the
// solver's dense layout (CalcOuterSource, outer.cc:107-217; GPU
// gpu_outer.cu:22-441):
//
//   q2grp0[g1][c]     = qi[g1][c] + sum_{g2 != g1} slgg[g1][mat c][g2][0]    * flux0[g2][c]
//   q2grpm[g1][c][m-1] =            sum_{g2 != g1} slgg[g1][mat c][g2][L(m)] * fluxm[g2][c][m-1]
//
// Per cell this is a product of the material's ng x ng scattering matrix
// (zero diagonal) with the cell's group fluxes, one matrix per Legendre
// order.  Two kernels:
//
//  * k_outer_ordered: one thread per (cell, moment) with the ng group fluxes
//    in registers, sums in ascending g2 with separate roundings -- the CPU
//    arithmetic bit for bit (the exact solve);
//  * k_outer_mfma: the same contraction on the FP64 matrix cores.  A
//    workgroup stages the moments of 16 consecutive cells in LDS; each
//    wavefront multiplies one 16-row block of a material's matrix (A, 16 x 4
//    per step) into the 16 cells' fluxes (B, 4 x 16) with
//    v_mfma_f64_16x16x4f64, ng/4 steps per output block.  Operand lanes:
//    A[i][k] and B[k][j] in lane i + 16k / j + 16k; D[i][j] in lane
//    j + 16*(i % 4), register i / 4.  The outputs go through LDS, so each
//    group's 16 cells x cmom moments reach HBM as contiguous runs.
//
// Layouts (solver): flux0, qi, q2grp0 [g][cell]; fluxm, q2grpm
// [g][cell][cmom-1]; slgg [g1][mat][g2][nmom]; mat [cell].
#include "snap_internal.h"

namespace snap {

namespace {

struct Orders {
  int8_t l[16];  // Legendre order of moment m (0 for m = 0)
};

Orders orders_for(int nmom) {
  Orders o{};
  int m = 0;
  for (int l = 0; l < nmom; l++)
    for (int j = 0; j < 2 * l + 1 && m < 16; j++) o.l[m++] = (int8_t)l;
  return o;
}

// ---------------------------------------------------------------------------
// order-preserving kernel
// ---------------------------------------------------------------------------
template <int NGM>
__global__ void __launch_bounds__(256)
    k_outer_ordered(const double *__restrict__ qi, const double *__restrict__ flux0,
                    const double *__restrict__ fluxm, const double *__restrict__ slgg,
                    const int32_t *__restrict__ mat, double *__restrict__ q0,
                    double *__restrict__ qm, int64_t cells, int ng, int nmat, int nmom, int cmom,
                    Orders ord) {
  // thread t: m = t % cmom (0 = scalar part), cell = t / cmom
  const int64_t n = cells * cmom;
  const int nm = cmom - 1;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / cmom;
    const int m = (int)(t - c * cmom);
    const int mt = mat[c];
    const int l = ord.l[m];
    double ph[NGM];
#pragma unroll
    for (int g2 = 0; g2 < NGM; g2++)
      ph[g2] = (g2 < ng) ? (m == 0 ? flux0[g2 * cells + c] : fluxm[(g2 * cells + c) * nm + m - 1])
                         : 0.0;
    for (int g1 = 0; g1 < ng; g1++) {
      const double *srow = slgg + ((size_t)(g1 * nmat + mt) * ng) * nmom + l;
      double acc = (m == 0) ? qi[g1 * cells + c] : 0.0;
#pragma unroll
      for (int g2 = 0; g2 < NGM; g2++)
        if (g2 < ng && g2 != g1) acc += srow[g2 * nmom] * ph[g2];
      if (m == 0)
        q0[g1 * cells + c] = acc;
      else
        qm[(g1 * cells + c) * nm + m - 1] = acc;
    }
  }
}

// group count above every register-resident variant: fluxes read per term
__global__ void __launch_bounds__(256)
    k_outer_ordered_any(const double *__restrict__ qi, const double *__restrict__ flux0,
                        const double *__restrict__ fluxm, const double *__restrict__ slgg,
                        const int32_t *__restrict__ mat, double *__restrict__ q0,
                        double *__restrict__ qm, int64_t cells, int ng, int nmat, int nmom,
                        int cmom, Orders ord) {
  const int64_t n = cells * cmom;
  const int nm = cmom - 1;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / cmom;
    const int m = (int)(t - c * cmom);
    const int mt = mat[c];
    const int l = ord.l[m];
    for (int g1 = 0; g1 < ng; g1++) {
      const double *srow = slgg + ((size_t)(g1 * nmat + mt) * ng) * nmom + l;
      double acc = (m == 0) ? qi[g1 * cells + c] : 0.0;
      for (int g2 = 0; g2 < ng; g2++) {
        if (g2 == g1) continue;
        const double f = (m == 0) ? flux0[g2 * cells + c] : fluxm[(g2 * cells + c) * nm + m - 1];
        acc += srow[g2 * nmom] * f;
      }
      if (m == 0)
        q0[g1 * cells + c] = acc;
      else
        qm[(g1 * cells + c) * nm + m - 1] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// matrix-core kernel
// ---------------------------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kChunk = 16;   // cells per workgroup (the MFMA's N)
constexpr int kOuterWaves = 16;
constexpr int kMaxKSteps = 16;  // K steps of 4 groups: the matrix-core kernel takes ng <= 64
constexpr int kMomBlock = 16;   // moments per workgroup (all of P3)

// A workgroup takes 16 cells (the MFMA's N) and every moment (m = 0 is the
// scalar flux) of every group.  LDS: fluxes ph[g2][m][cell]; with STAGE the
// outputs of sgt 16-group blocks ob[row][cell][m]; then mat[16].  STAGE
// writes each group's 16 cells x (cmom-1) moments as one contiguous run
// after the blocks' products (a lane's own result is one moment of one
// cell, 8 bytes at a (cmom-1)*8-byte stride); without it (when no block's
// outputs fit in LDS) every lane stores its results directly.
template <bool STAGE>
__global__ void __launch_bounds__(64 * kOuterWaves)
    k_outer_mfma(const double *__restrict__ qi, const double *__restrict__ flux0,
                 const double *__restrict__ fluxm, const double *__restrict__ slgg,
                 const int32_t *__restrict__ mat, double *__restrict__ q0,
                 double *__restrict__ qm, int64_t cells, int ng, int nmat, int nmom, int cmom,
                 Orders ord, int sgt) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int mb = cmom;  // <= kMomBlock
  double *ph = (double *)smem_raw;
  double *ob = ph + (size_t)ng * mb * kChunk;
  int *smat = (int *)(ob + (STAGE ? (size_t)sgt * 16 * kChunk * mb : 0));
  const int nm = cmom - 1;
  const int64_t c0 = (int64_t)blockIdx.x * kChunk;
  const int nc = (int)min((int64_t)kChunk, cells - c0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

  // ---- stage the chunk's fluxes; each thread issues kStage independent
  // loads before it writes any of them to LDS.  Element e = (g2, cell j,
  // moment m), moments fastest: a cell's moments are adjacent in memory.
  constexpr int kStage = 8;
  const int per_g = kChunk * mb;
  const int total = ng * per_g;
  for (int e0 = threadIdx.x; e0 < total; e0 += blockDim.x * kStage) {
    double v[kStage];
    int dst[kStage];
#pragma unroll
    for (int u = 0; u < kStage; u++) {
      const int e = e0 + u * blockDim.x;
      dst[u] = -1;
      v[u] = 0.0;
      if (e < total) {
        const int g2 = e / per_g, r = e - g2 * per_g, j = r / mb, m = r - j * mb;
        dst[u] = (g2 * mb + m) * kChunk + j;
        if (j < nc)
          v[u] = (m == 0) ? flux0[g2 * cells + c0 + j] : fluxm[(g2 * cells + c0 + j) * nm + m - 1];
      }
    }
#pragma unroll
    for (int u = 0; u < kStage; u++)
      if (dst[u] >= 0) ph[dst[u]] = v[u];
  }
  for (int j = threadIdx.x; j < kChunk; j += blockDim.x) smat[j] = (j < nc) ? mat[c0 + j] : -1;
  __syncthreads();

  // ---- tasks (16-row block gt of g1, moment m), sgt blocks per pass (all of
  // them without STAGE); wavefronts take turns
  const int ngt = (ng + 15) / 16, nks = (ng + 3) / 4;
  const int i = lane & 15, kq = lane >> 4;  // A row / B column, and K quarter
  const int my_mat = smat[i];
  const int per_pass = STAGE ? sgt : ngt;
  for (int gs = 0; gs < ngt; gs += per_pass) {
    const int nt = min(per_pass, ngt - gs);
    for (int task = wave; task < nt * mb; task += kOuterWaves) {
      const int gt = gs + task / mb, m = task - (task / mb) * mb;
      const int l = ord.l[m];
      const int g1a = gt * 16 + i;
      for (int mt = 0; mt < nmat; mt++) {
        // materials present among the chunk's cells (lanes 0..15 vote)
        if (__builtin_amdgcn_ballot_w64(lane < 16 && my_mat == mt) == 0) continue;
        // the block's A fragments first (all loads in flight together), then
        // the chain of matrix products
        double av[kMaxKSteps];
#pragma unroll
        for (int ks = 0; ks < kMaxKSteps; ks++) {
          const int g2 = ks * 4 + kq;
          av[ks] = (ks < nks && g1a < ng && g2 < ng && g1a != g2)
                       ? slgg[((size_t)(g1a * nmat + mt) * ng + g2) * nmom + l]
                       : 0.0;
        }
        f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < kMaxKSteps; ks++) {
          if (ks >= nks) break;
          const int g2 = ks * 4 + kq;
          const double b = (g2 < ng) ? ph[(g2 * mb + m) * kChunk + i] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], b, acc, 0, 0, 0);
        }
        // D[row][col] sits in lane col + 16*(row % 4), register row / 4
        // (col = cell, row = g1 within the block)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = 4 * r + kq, g1 = gt * 16 + row;
          const int j = i;
          if (g1 < ng && j < nc && smat[j] == mt) {
            if constexpr (STAGE) {
              ob[(((gt - gs) * 16 + row) * kChunk + j) * mb + m] = acc[r];
            } else {
              const int64_t c = c0 + j;
              if (m == 0)
                q0[g1 * cells + c] = qi[g1 * cells + c] + acc[r];
              else
                qm[(g1 * cells + c) * nm + m - 1] = acc[r];
            }
          }
        }
      }
    }
    if constexpr (STAGE) {
      __syncthreads();
      const int rows = min(nt * 16, ng - gs * 16);  // groups gs*16 .. gs*16+rows-1
      for (int e = threadIdx.x; e < rows * nc; e += blockDim.x) {
        const int row = e / nc, j = e - row * nc;
        const int64_t o = (int64_t)(gs * 16 + row) * cells + c0 + j;
        q0[o] = qi[o] + ob[(row * kChunk + j) * mb];
      }
      // a group's moments of the chunk's cells are one run of nc*nm doubles
      const int run = nc * nm;
      for (int e = threadIdx.x; e < rows * run; e += blockDim.x) {
        const int row = e / run, r = e - row * run, j = r / nm, k = r - j * nm;
        qm[((int64_t)(gs * 16 + row) * cells + c0) * nm + r] = ob[(row * kChunk + j) * mb + k + 1];
      }
      __syncthreads();  // the pass's outputs are read before the next pass's
    }
  }
}

}  // namespace

// LDS of k_outer_mfma staging sgt 16-group blocks of outputs per pass (0: none)
size_t outer_mfma_lds(int ng, int cmom, int sgt) {
  return (size_t)ng * cmom * kChunk * sizeof(double) +
         (size_t)sgt * 16 * kChunk * cmom * sizeof(double) + kChunk * sizeof(int);
}

int launch_outer_dense(const double *qi, const double *flux0, const double *fluxm,
                       const double *slgg, const int32_t *mat, double *q0, double *qm,
                       int64_t cells, int ng, int nmat, int nmom, bool ordered, hipStream_t st) {
  const int cmom = nmom * nmom;
  if (nmom < 1 || nmom > 4) return fail(SNAP_ERR_INVALID, "nmom %d outside 1..4", nmom);
  if (cmom > 1 && (!fluxm || !qm)) return fail(SNAP_ERR_INVALID, "moment arrays missing");
  const Orders ord = orders_for(nmom);
  constexpr size_t kLdsMax = 160 * 1024;
  if (!ordered && ng <= 4 * kMaxKSteps && cmom <= kMomBlock && outer_mfma_lds(ng, cmom, 0) <= kLdsMax) {
    // as many 16-group output blocks per pass as LDS holds (all when it can)
    const int ngt = (ng + 15) / 16;
    int sgt = ngt;
    while (sgt > 0 && outer_mfma_lds(ng, cmom, sgt) > kLdsMax) sgt--;
    const size_t lds = outer_mfma_lds(ng, cmom, sgt);
    const void *k = sgt ? (const void *)k_outer_mfma<true> : (const void *)k_outer_mfma<false>;
    SNAP_TRY(ensure_dynamic_lds(k, lds));
    const int64_t chunks = (cells + kChunk - 1) / kChunk;
    if (chunks > INT32_MAX) return fail(SNAP_ERR_UNSUPPORTED, "too many cells");
    if (sgt)
      hipLaunchKernelGGL(k_outer_mfma<true>, dim3((unsigned)chunks), dim3(64 * kOuterWaves), lds, st,
                         qi, flux0, fluxm, slgg, mat, q0, qm, cells, ng, nmat, nmom, cmom, ord, sgt);
    else
      hipLaunchKernelGGL(k_outer_mfma<false>, dim3((unsigned)chunks), dim3(64 * kOuterWaves), lds, st,
                         qi, flux0, fluxm, slgg, mat, q0, qm, cells, ng, nmat, nmom, cmom, ord, 0);
  } else {
    const int64_t n = cells * cmom;
    const dim3 grid(grid_for(n, 256)), block(256);
    if (ng <= 8)
      hipLaunchKernelGGL(k_outer_ordered<8>, grid, block, 0, st, qi, flux0, fluxm, slgg, mat, q0,
                         qm, cells, ng, nmat, nmom, cmom, ord);
    else if (ng <= 16)
      hipLaunchKernelGGL(k_outer_ordered<16>, grid, block, 0, st, qi, flux0, fluxm, slgg, mat, q0,
                         qm, cells, ng, nmat, nmom, cmom, ord);
    else if (ng <= 32)
      hipLaunchKernelGGL(k_outer_ordered<32>, grid, block, 0, st, qi, flux0, fluxm, slgg, mat, q0,
                         qm, cells, ng, nmat, nmom, cmom, ord);
    else
      hipLaunchKernelGGL(k_outer_ordered_any, grid, block, 0, st, qi, flux0, fluxm, slgg, mat, q0,
                         qm, cells, ng, nmat, nmom, cmom, ord);
  }
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

}  // namespace snap
