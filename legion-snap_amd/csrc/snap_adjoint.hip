// snap_adjoint.hip -- reverse (adjoint) sweep of the steady-state solve.
//
// No counterpart in the reference.  Gradient of a detector response
// R = sum_g sum_c r[g][c] * flux0[g][c] at a converged solution with respect
// to the total cross section t_xs[g][c] (for a per-cell absorption field,
// d/dsiga_c = sum_g d/dt_xs[g][c]).  nmom = 1 (isotropic scattering).
//
// With the forward fixup's decisions held fixed, one sweep is linear in its
// source and inflows: per cell and direction
//   pc = N / D,  N = q + sum_d h_d k_d psi_in,d,  D = t_xs + sum_d h_d v_d,
//   psi_out,d = (2 pc - psi_in,d) o_d
// (plain diamond difference: k = v = o = 1; a fixed cell: pc-pass flags v_d,
// k_d = (1 + v_d)/2, outflow masks o_d, pc = 0 when zeroed).  The reverse
// cell recurrence, against the sweep direction, is
//   bpc = w_a zeta_c + 2 sum_d o_d b_out,d
//   bN = bpc / D,  bD = -bpc pc / D   (0 when zeroed)
//   b_in,d = -o_d b_out,d + bN h_d k_d,
// with bN summed over angles giving dR/dq_c and bD giving dR/dt_xs_c.  The
// adjoint source iteration zeta = r + S^T u(zeta) (u = sum of bN) converges
// like the forward one.
//
// Precision: the forward recompute (pc and the fixup's decisions) is FP64
// from an FP64 checkpoint; the reverse recurrence runs in AReal (FP64, or
// FP32 for the FP32 reverse sweeps).  The decisions fix which linear map is
// differentiated, and an FP32 forward flips some of them (a face near zero),
// which moves the gradient along that face's characteristic by far more
// than FP32 rounding; with FP64 decisions the FP32 reverse sweep differs
// from the FP64 one by rounding only.
//
// Checkpointing: the forward sweep leaves every tile's outflow faces in its
// face buffer (the tile-edge planes).  The reverse kernel takes one tile per
// workgroup, recomputes the tile's forward pc and fixup decisions from those
// faces into LDS, then walks the tile's cell diagonals backwards with the
// adjoint faces in LDS and hands its low-side adjoint faces to the upstream
// tiles through a second face buffer.  Tile diagonals launch in reverse
// order.  No angular flux is stored beyond one tile.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "snap_internal.h"

namespace snap {

namespace {

constexpr int kAdjSteps = 4;     // cell groups per wave and diagonal
constexpr unsigned kAfOutX = 1u, kAfOutY = 2u, kAfOutZ = 4u, kAfPcX = 8u, kAfPcY = 16u,
                   kAfPcZ = 32u, kAfFixed = 64u, kAfZero = 128u;

struct AdjArgs {
  Extent3 ext;
  int TX, TY, TZ, nTx, nTy, nTz, ntiles;
  int nang, ncellplanes, ncells_tile, sb, ng_job, fixup;
  double hi, hj, hk;
  const double *mu, *eta, *xi, *w;
  const double *q;       // [ng][cells] forward source
  const double *txs;     // [ng][cells]
  const double *zeta;    // [ng][cells]
  double *part_u;        // [8][ng][cells] sum over angles of dR/dq
  double *part_grad;     // [8][ng][cells] sum over angles of dR/dt_xs
  const void *faces;     // forward tile faces (the checkpoint)
  void *adj_faces;       // adjoint tile faces
  const int *plane_tiles;
  const int *cell_off;
  const int16_t *cells;
};

struct AdjLds {
  size_t pct, flt, tcr, coff, bat, cix, total;
};

__host__ __device__ inline size_t a16(size_t b) { return (b + 15) & ~size_t(15); }

// LDS: the tile's faces (forward FP64, then adjoint AReal of arb bytes in
// the same space), forward pc in AReal, decisions, tables, sums.
__host__ __device__ inline AdjLds adj_lds(int TX, int TY, int TZ, int nang, int W, int SB,
                                          int ncells_tile, int ncp, int arb) {
  AdjLds L;
  size_t o = a16((size_t)(TY * TZ + TX * TZ + TX * TY) * nang * sizeof(double));
  L.pct = o;
  o = a16(o + (size_t)ncells_tile * nang * arb);
  L.flt = o;
  o = a16(o + (size_t)ncells_tile * nang);
  L.tcr = o;
  o = a16(o + (size_t)ncells_tile * 16);
  L.coff = o;
  o = a16(o + (size_t)(ncp + 1) * 4);
  L.bat = o;
  o = a16(o + (size_t)W * SB * 2 * nang * 8);
  L.cix = o;
  o = a16(o + (size_t)W * SB * 4);
  L.total = o;
  return L;
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of the ballots inside the lane's cell segment, summed
// (the cell-global count of negative outflows).
template <int CPW>
__device__ inline unsigned seg_count3(unsigned long long bx, unsigned long long by,
                                      unsigned long long bz, unsigned long long m0, int nang,
                                      int cl) {
  if constexpr (CPW == 1) {
    return (unsigned)(__popcll(bx) + __popcll(by) + __popcll(bz));
  } else {
    unsigned res = 0;
#pragma unroll
    for (int cs = 0; cs < CPW; cs++) {
      const unsigned long long m = m0 << (cs * nang);
      const unsigned n = (unsigned)(__popcll(bx & m) + __popcll(by & m) + __popcll(bz & m));
      res = (cl == cs) ? n : res;
    }
    return res;
  }
}

template <typename AReal, int CPW>
__global__ void __launch_bounds__(1024) k_adjoint_tiles(AdjArgs A, int plane_off) {
  using Real = double;  // the forward recompute
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int TX = A.TX, TY = A.TY, TZ = A.TZ, nang = A.nang, SB = A.sb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const int ncp = A.ncellplanes;
  const AdjLds lay = adj_lds(TX, TY, TZ, nang, W, SB, A.ncells_tile, ncp, (int)sizeof(AReal));
  const int fxn = TY * TZ * nang, fyn = TX * TZ * nang, fzn = TX * TY * nang;
  Real *fx = (Real *)smem_raw;
  Real *fy = fx + fxn;
  Real *fz = fy + fyn;
  AReal *pct = (AReal *)(smem_raw + lay.pct);                 // [tile cell][angle] forward pc
  unsigned char *flt = smem_raw + lay.flt;                    // [tile cell][angle] decisions
  int4 *tcr = (int4 *)(smem_raw + lay.tcr);                   // cell, face offsets
  int *toff = (int *)(smem_raw + lay.coff);
  double *bat = (double *)(smem_raw + lay.bat) + (size_t)wave * SB * 2 * nang;  // [SB][2][nang]
  int *cix = (int *)(smem_raw + lay.cix) + wave * SB;

  const int s = blockIdx.y;
  const int corner = s / A.ng_job, g = s % A.ng_job;
  const int tp = A.plane_tiles[plane_off + blockIdx.x];
  const int tx = tp & 1023, ty = (tp >> 10) & 1023, tz = (tp >> 20) & 1023;
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int nx = A.ext.nx, ny = A.ext.ny, nz = A.ext.nz;
  const int ex = min(TX, nx - x0), ey = min(TY, ny - y0), ez = min(TZ, nz - z0);
  const bool sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
  const int tile_id = (tz * A.nTy + ty) * A.nTx + tx;
  const int per_tile = fxn + fyn + fzn;
  const Real *faces_s = (const Real *)A.faces + (size_t)s * A.ntiles * per_tile;
  AReal *adj_s = (AReal *)A.adj_faces + (size_t)s * A.ntiles * per_tile;
  const int ncells = nx * ny * nz;
  const double *qg = A.q + (size_t)g * ncells;
  const double *tg = A.txs + (size_t)g * ncells;
  const double *zg = A.zeta + (size_t)g * ncells;
  double *pug = A.part_u + ((size_t)corner * A.ng_job + g) * ncells;
  double *pgg = A.part_grad + ((size_t)corner * A.ng_job + g) * ncells;

  const int cl = (CPW == 1) ? 0 : lane / nang;
  const int a = lane - cl * nang;
  const bool lane_ok = cl < CPW && a < nang;
  const int aa = lane_ok ? a : 0;
  const unsigned long long seg0 = (nang >= 64) ? ~0ull : ((1ull << nang) - 1ull);
  const int per_step = W * CPW;

  // ---- tables and the forward inflow faces of the tile (checkpoint)
  for (int idx = threadIdx.x; idx < A.ncells_tile; idx += blockDim.x) {
    const int pk = A.cells[idx];
    const int i = pk & 31, j = (pk >> 5) & 31, k = (pk >> 10) & 31;
    int cell = -1;
    if (i < ex && j < ey && k < ez) {
      const int X = x0 + i, Y = y0 + j, Z = z0 + k;
      const int px = sxp ? X : nx - 1 - X, py = syp ? Y : ny - 1 - Y, pz = szp ? Z : nz - 1 - Z;
      cell = (pz * ny + py) * nx + px;
    }
    tcr[idx] = make_int4(cell, (k * TY + j) * nang, (k * TX + i) * nang, (j * TX + i) * nang);
  }
  for (int idx = threadIdx.x; idx <= ncp; idx += blockDim.x) toff[idx] = A.cell_off[idx];
  {
    const Real *ux = faces_s + (size_t)(tile_id - 1) * per_tile;
    const Real *uy = faces_s + (size_t)(tile_id - A.nTx) * per_tile + fxn;
    const Real *uz = faces_s + (size_t)(tile_id - A.nTx * A.nTy) * per_tile + fxn + fyn;
    for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) fx[idx] = tx > 0 ? ux[idx] : Real(0);
    for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) fy[idx] = ty > 0 ? uy[idx] : Real(0);
    for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) fz[idx] = tz > 0 ? uz[idx] : Real(0);
  }
  const double mu_d = A.mu[aa], eta_d = A.eta[aa], xi_d = A.xi[aa];
  const Real cx = mu_d * A.hi, cy = eta_d * A.hj, cz = xi_d * A.hk;
  const double w_d = A.w[aa];
  const Real tolr = (Real)1.0e-12;
  __syncthreads();

  // ---- forward recompute of the tile: pc and the fixup's decisions
  for (int d = 0; d < ncp; ++d) {
    const int c0 = toff[d], c1 = toff[d + 1];
    for (int st = 0; st < kAdjSteps; ++st) {
      const int my = c0 + st * per_step + wave * CPW + cl;
      if (c0 + st * per_step >= c1) break;
      const int4 rec = (lane_ok && my < c1) ? tcr[my] : make_int4(-1, 0, 0, 0);
      const int cell = rec.x;
      if (cell >= 0) {
        const Real q = qg[cell], txs = tg[cell];
        Real *sx = fx + rec.y + a;
        Real *sy = fy + rec.z + a;
        Real *sz = fz + rec.w + a;
        const Real pi = *sx, pj = *sy, pk = *sz;
        const Real dv = Real(1) / (((txs + cx) + cy) + cz);
        Real pc = q;
        pc = __builtin_fma(pi, cx, pc);
        pc = __builtin_fma(pj, cy, pc);
        pc = __builtin_fma(pk, cz, pc);
        pc *= dv;
        bool hx = true, hy = true, hz = true;
        unsigned f = 0;
        if (A.fixup) {
          auto count = [&]() -> unsigned {
            const Real p2 = Real(2) * pc;
            const bool nxn = p2 - pi < Real(0), nyn = p2 - pj < Real(0), nzn = p2 - pk < Real(0);
            hx = hx && !nxn;
            hy = hy && !nyn;
            hz = hz && !nzn;
            return seg_count3<CPW>(__builtin_amdgcn_ballot_w64(nxn), __builtin_amdgcn_ballot_w64(nyn),
                                   __builtin_amdgcn_ballot_w64(nzn), seg0, nang, cl);
          };
          unsigned old_neg = count();
          if (old_neg != 0) {
            for (int it = 1; it < 3 * nang + 2; ++it) {
              const Real vx = hx ? Real(1) : Real(0), vy = hy ? Real(1) : Real(0),
                         vz = hz ? Real(1) : Real(0);
              const Real t = pi * cx * (Real(1) + vx) + pj * cy * (Real(1) + vy) +
                             pk * cz * (Real(1) + vz);
              const Real p = q + Real(0.5) * t;
              const Real den = (p <= Real(0)) ? Real(0) : (txs + cx * vx + cy * vy + cz * vz);
              const bool zero = den < tolr;
              pc = zero ? Real(0) : p / den;
              f = kAfFixed | (hx ? kAfPcX : 0u) | (hy ? kAfPcY : 0u) | (hz ? kAfPcZ : 0u) |
                  (zero ? kAfZero : 0u);
              const unsigned neg = count();
              if (neg == old_neg) break;
              old_neg = neg;
            }
          }
        }
        f |= (hx ? kAfOutX : 0u) | (hy ? kAfOutY : 0u) | (hz ? kAfOutZ : 0u);
        const Real p2 = Real(2) * pc;
        *sx = (p2 - pi) * (hx ? Real(1) : Real(0));
        *sy = (p2 - pj) * (hy ? Real(1) : Real(0));
        *sz = (p2 - pk) * (hz ? Real(1) : Real(0));
        pct[(size_t)my * nang + a] = (AReal)pc;
        flt[(size_t)my * nang + a] = (unsigned char)f;
      }
    }
    __syncthreads();
  }

  // ---- adjoint inflow: the downstream tiles' low-side adjoint faces (zero
  // past the chunk's far side, where outflows leave), in the face space the
  // forward recompute has finished with
  AReal *bfx = (AReal *)smem_raw;
  AReal *bfy = bfx + fxn;
  AReal *bfz = bfy + fyn;
  {
    const AReal *dx = adj_s + (size_t)(tile_id + 1) * per_tile;
    const AReal *dy = adj_s + (size_t)(tile_id + A.nTx) * per_tile + fxn;
    const AReal *dz = adj_s + (size_t)(tile_id + A.nTx * A.nTy) * per_tile + fxn + fyn;
    const bool hx1 = tx + 1 < A.nTx, hy1 = ty + 1 < A.nTy, hz1 = tz + 1 < A.nTz;
    for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) bfx[idx] = hx1 ? dx[idx] : AReal(0);
    for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) bfy[idx] = hy1 ? dy[idx] : AReal(0);
    for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) bfz[idx] = hz1 ? dz[idx] : AReal(0);
  }
  __syncthreads();
  const AReal acx = (AReal)cx, acy = (AReal)cy, acz = (AReal)cz;

  // ---- reverse pass over the tile's cell diagonals
  int nbuf = 0;
  auto flush = [&]() {
    // per-cell sums over angles in ascending angle order
    wave_sync();
    for (int o = lane; o < nbuf * 2; o += 64) {
      const int kk = o >> 1, which = o & 1;
      const int cell = cix[kk];
      if (cell < 0) continue;
      const double *v = bat + ((size_t)kk * 2 + which) * nang;
      double sum = 0.0;
      for (int b = 0; b < nang; b++) sum += v[b];
      if (which == 0)
        pug[cell] = sum;
      else
        pgg[cell] = sum;
    }
    wave_sync();
    nbuf = 0;
  };
  for (int d = ncp - 1; d >= 0; --d) {
    const int c0 = toff[d], c1 = toff[d + 1];
    for (int st = 0; st < kAdjSteps; ++st) {
      if (c0 + st * per_step >= c1) break;
      const int my = c0 + st * per_step + wave * CPW + cl;
      const int4 rec = (lane_ok && my < c1) ? tcr[my] : make_int4(-1, 0, 0, 0);
      const int cell = rec.x;
      double bn_out = 0.0, bd_out = 0.0;
      if (cell >= 0) {
        using R = AReal;
        const unsigned f = flt[(size_t)my * nang + a];
        const R pc = pct[(size_t)my * nang + a];
        const R txs = (R)tg[cell];
        R *sx = bfx + rec.y + a;
        R *sy = bfy + rec.z + a;
        R *sz = bfz + rec.w + a;
        const R bx = *sx, by = *sy, bz = *sz;
        const R ox = (f & kAfOutX) ? R(1) : R(0), oy = (f & kAfOutY) ? R(1) : R(0),
                oz = (f & kAfOutZ) ? R(1) : R(0);
        const R bpc = (R)(w_d * zg[cell]) + R(2) * (ox * bx + oy * by + oz * bz);
        R kx = R(1), ky = R(1), kz = R(1), den;
        if (f & kAfFixed) {
          const R vx = (f & kAfPcX) ? R(1) : R(0), vy = (f & kAfPcY) ? R(1) : R(0),
                  vz = (f & kAfPcZ) ? R(1) : R(0);
          kx = R(0.5) * (R(1) + vx);
          ky = R(0.5) * (R(1) + vy);
          kz = R(0.5) * (R(1) + vz);
          den = txs + acx * vx + acy * vy + acz * vz;
        } else {
          den = ((txs + acx) + acy) + acz;
        }
        R bn = R(0), bd = R(0);
        if (!(f & kAfZero)) {
          bn = bpc / den;
          bd = -bn * pc;
        }
        *sx = -ox * bx + bn * acx * kx;
        *sy = -oy * by + bn * acy * ky;
        *sz = -oz * bz + bn * acz * kz;
        bn_out = (double)bn;
        bd_out = (double)bd;
      }
      if (lane_ok) {
        bat[((size_t)(nbuf + cl) * 2) * nang + a] = bn_out;
        bat[((size_t)(nbuf + cl) * 2 + 1) * nang + a] = bd_out;
        if (a == 0) cix[nbuf + cl] = cell;
      }
      nbuf += CPW;
      if (nbuf + CPW > SB) flush();
    }
    __syncthreads();
  }
  if (nbuf > 0) flush();

  // ---- low-side adjoint faces to the upstream tiles
  AReal *out = adj_s + (size_t)tile_id * per_tile;
  for (int idx = threadIdx.x; idx < per_tile; idx += blockDim.x) out[idx] = bfx[idx];
}

typedef void (*AdjKernel)(AdjArgs, int);

template <typename Real>
AdjKernel adj_kernel_for(int cpw) {
  switch (cpw) {
    case 1: return k_adjoint_tiles<Real, 1>;
    case 2: return k_adjoint_tiles<Real, 2>;
    case 3: return k_adjoint_tiles<Real, 3>;
    case 4: return k_adjoint_tiles<Real, 4>;
    case 5: return k_adjoint_tiles<Real, 5>;
    case 6: return k_adjoint_tiles<Real, 6>;
    case 8: return k_adjoint_tiles<Real, 8>;
    default: return nullptr;
  }
}

// zeta_new[g2][c] = r[g2][c] + sum_g S[g][mat c][g2] u[g][c] (g ascending),
// and the max-norms of the change and of zeta_new (non-negative doubles
// compare as their bit patterns)
__global__ void k_adjoint_update(const double *__restrict__ r, const double *__restrict__ u,
                                 const double *__restrict__ slgg, const int32_t *__restrict__ mat,
                                 const double *__restrict__ zeta, double *__restrict__ zn,
                                 unsigned long long *norms, int64_t cells, int ng, int nmat,
                                 int nmom) {
  const int64_t n = cells * ng;
  unsigned long long dmax = 0, zmax = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int g2 = (int)(t / cells);
    const int64_t c = t - (int64_t)g2 * cells;
    const int mt = mat[c];
    double acc = r[t];
    for (int g = 0; g < ng; g++)
      acc += slgg[((size_t)(g * nmat + mt) * ng + g2) * nmom] * u[(int64_t)g * cells + c];
    zn[t] = acc;
    const double dz = fabs(acc - zeta[t]), az = fabs(acc);
    dmax = max(dmax, (unsigned long long)__double_as_longlong(dz));
    zmax = max(zmax, (unsigned long long)__double_as_longlong(az));
  }
  // wave maxima, then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) {
    dmax = max(dmax, (unsigned long long)__shfl_xor((long long)dmax, off));
    zmax = max(zmax, (unsigned long long)__shfl_xor((long long)zmax, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&norms[0], dmax);
    atomicMax(&norms[1], zmax);
  }
}

__global__ void k_clear_u64(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ void k_to_float(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

}  // namespace

int adjoint_plan_fits(const SweepPlan &p, int nang, int real_bytes, int *cpw, int *waves,
                      int *sb) {
  if (nang < 1 || nang > 64)
    return fail(SNAP_ERR_UNSUPPORTED, "the adjoint sweep handles 1..64 angles per octant");
  int c = 0;
  for (int v : {1, 2, 3, 4, 5, 6, 8})
    if (v * nang <= 64) c = v;
  const int w = std::max(4, (p.maxdiag + kAdjSteps * c - 1) / (kAdjSteps * c));
  if (w > 16) return fail(SNAP_ERR_INVALID, "adjoint tile needs more than 16 waves");
  const int b = c;  // the per-cell sums are flushed after every step
  const size_t lds = adj_lds(p.TX, p.TY, p.TZ, nang, w, b, p.TX * p.TY * p.TZ, p.ncellplanes,
                             real_bytes).total;
  if (lds > 160 * 1024)
    return fail(SNAP_ERR_INVALID, "adjoint tile %dx%dx%d needs %zu B of LDS", p.TX, p.TY, p.TZ, lds);
  *cpw = c;
  *waves = w;
  *sb = b;
  return SNAP_OK;
}

int adjoint_enqueue(const SweepPlan &p, const QuadDev &quad, int ng, int fixup, double hi,
                    double hj, double hk, const double *q, const double *txs, const double *zeta,
                    double *part_u, double *part_grad, const void *faces, void *adj_faces,
                    bool fp32, hipStream_t st) {
  int cpw, waves, sb;
  SNAP_TRY(adjoint_plan_fits(p, p.nang, fp32 ? 4 : 8, &cpw, &waves, &sb));
  AdjArgs A;
  memset(&A, 0, sizeof(A));
  A.ext = p.ext;
  A.TX = p.TX;
  A.TY = p.TY;
  A.TZ = p.TZ;
  A.nTx = p.nTx;
  A.nTy = p.nTy;
  A.nTz = p.nTz;
  A.ntiles = p.ntiles;
  A.nang = p.nang;
  A.ncellplanes = p.ncellplanes;
  A.ncells_tile = p.TX * p.TY * p.TZ;
  A.sb = sb;
  A.ng_job = ng;
  A.fixup = fixup;
  A.hi = hi;
  A.hj = hj;
  A.hk = hk;
  A.mu = quad.mu;
  A.eta = quad.eta;
  A.xi = quad.xi;
  A.w = quad.w;
  A.q = q;
  A.txs = txs;
  A.zeta = zeta;
  A.part_u = part_u;
  A.part_grad = part_grad;
  A.faces = faces;
  A.adj_faces = adj_faces;
  A.plane_tiles = p.plane_tiles;
  A.cell_off = p.cell_off;
  A.cells = p.cells;
  const AdjKernel k = fp32 ? adj_kernel_for<float>(cpw) : adj_kernel_for<double>(cpw);
  if (!k) return fail(SNAP_ERR_INVALID, "no adjoint kernel for %d cells per wave", cpw);
  const size_t lds = adj_lds(p.TX, p.TY, p.TZ, p.nang, waves, sb, A.ncells_tile, p.ncellplanes,
                             fp32 ? 4 : 8).total;
  SNAP_TRY(ensure_dynamic_lds((const void *)k, lds));
  for (int D = p.nplanes - 1; D >= 0; D--) {
    const int count = p.plane_off[D + 1] - p.plane_off[D];
    hipLaunchKernelGGL(k, dim3(count, 8 * ng), dim3(64 * waves), lds, st, A, p.plane_off[D]);
  }
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

int launch_adjoint_update(const double *r, const double *u, const double *slgg,
                          const int32_t *mat, const double *zeta, double *zn,
                          unsigned long long *norms, int64_t cells, int ng, int nmat, int nmom,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(64), 0, st, norms, 2);
  const int64_t n = cells * ng;
  hipLaunchKernelGGL(k_adjoint_update, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, r, u, slgg,
                     mat, zeta, zn, norms, cells, ng, nmat, nmom);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

int launch_to_float(const double *a, float *b, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_float, dim3(grid_for(n, 256)), dim3(256), 0, st, a, b, n);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

}  // namespace snap
