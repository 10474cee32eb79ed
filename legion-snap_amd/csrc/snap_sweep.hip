// snap_sweep.hip -- the diamond-difference mini-KBA sweep with negative-flux
// fixup (reference: MiniKBATask, sweep.cc:242-616 CPU, gpu_sweep.cu:1395-1742
// GPU), written for gfx950.
//
// Decomposition.  One sweep = one (octant, group) pass over a chunk of
// nx*ny*nz cells.  The chunk is cut into tiles of TX*TY*TZ cells; tile
// (tx,ty,tz) in sweep-order coordinates depends on (tx-1,ty,tz), (tx,ty-1,tz)
// and (tx,ty,tz-1), so all tiles on one tile diagonal tx+ty+tz = D are
// independent.  Each launch processes one tile diagonal for every sweep of the
// job (grid.x = tiles on the diagonal, grid.y = sweeps), so the launch order
// is the dependency order and no workgroup ever waits on another.
//
// Inside a workgroup the tile is walked by cell diagonals i+j+k = d.  The
// three inflow faces of the tile live in LDS and are updated in place (the
// cell at (i,j,k) reads face_x[j,k] written by (i-1,j,k) on the previous
// diagonal and overwrites it with its own outflow), with one workgroup barrier
// per cell diagonal.  Tile faces go to HBM only at tile entry and exit.
//
// Lanes carry angles: a 64-lane wavefront holds CPW = 64/nang cells of the
// current diagonal, lane = cell_slot*nang + angle.  The fixup's negative-flux
// count is cell-global over angles (sweep.cc:437-505), so each pass counts with
// three wave ballots masked to the cell's lane segment.  Flux moments are
// reduced per cell through LDS in ascending angle order, as the CPU variant
// sums them (sweep.cc:568-591).
//
// The FP64 instantiation keeps the reference's operation order everywhere and
// the library is built with -ffp-contract=off, so its results are the CPU
// arithmetic bit for bit.  The FP32 instantiation (mixed precision) runs the
// per-angle arithmetic and faces in float and the moment sums in double.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "snap_internal.h"

namespace snap {

// ---------------------------------------------------------------------------
// LDS layout of one tile workgroup (byte offsets), shared by host and device
// ---------------------------------------------------------------------------
struct SweepLds {
  size_t ecs, ect, pwb, cix, qsl, cells, coff, total;  // byte offsets from the start of LDS
};

__host__ __device__ inline size_t a16(size_t b) { return (b + 15) & ~size_t(15); }

constexpr int kMaxSteps = 4;

// faces [x|y|z] (Real) | ec of the corner [cmom][nang] and transposed
// [nang][cmom] | per-wave w*pc batches [W][SB][nang] | their cell indices
// [W][SB] | per-wave source moments [W][kMaxSteps][CPW][cmom] | the tile's
// cell list [ncells] and diagonal offsets [ncp+1]
__host__ __device__ inline SweepLds sweep_lds(int TX, int TY, int TZ, int nang, int cmom, int W,
                                              int SB, int ncells_tile, int ncp, int cpw, int rb) {
  SweepLds L;
  size_t o = a16((size_t)(TY * TZ + TX * TZ + TX * TY) * nang * rb);
  L.ecs = o;
  o += (size_t)cmom * nang * 8;
  L.ect = o;
  o += (size_t)cmom * nang * 8;
  L.pwb = o;
  o += (size_t)W * SB * nang * 8;
  L.qsl = o;
  o += (size_t)W * kMaxSteps * cpw * cmom * 8;
  L.cix = o;
  o = a16(o + (size_t)W * SB * 4);
  L.cells = o;
  o = a16(o + (size_t)ncells_tile * 2);
  L.coff = o;
  o = a16(o + (size_t)(ncp + 1) * 4);
  L.total = o;
  return L;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
size_t SweepPlan::lds_bytes(int real_bytes, bool recompute_dinv) const {
  (void)recompute_dinv;
  return sweep_lds(TX, TY, TZ, nang, cmom, waves, sb, TX * TY * TZ, ncellplanes, cpw, real_bytes).total;
}

size_t SweepPlan::face_doubles_per_sweep() const {
  return (size_t)ntiles * (TY * TZ + TX * TZ + TX * TY) * nang;
}

static const int kCpwChoices[] = {1, 2, 3, 4, 5, 6, 8};

static int pick_cpw(int nang) {
  int best = 1;
  for (int c : kCpwChoices)
    if (c * nang <= 64) best = c;
  return best;
}

void plan_default_tile(Extent3 e, int nang, int cmom, int *TX, int *TY, int *TZ, int *waves) {
  // 8x8x4 tiles: 18 cell diagonals of at most 32 cells, two workgroups per
  // CU with FP64 faces and double-buffered staging at 36 angles.  Clamp to
  // the chunk.
  *TX = std::min(8, e.nx);
  *TY = std::min(8, e.ny);
  *TZ = std::min(4, e.nz);
  *waves = 8;
  (void)nang;
  (void)cmom;
}

void plan_free(SweepPlan *p) {
  if (p->plane_tiles) (void)hipFree(p->plane_tiles);
  if (p->cell_off) (void)hipFree(p->cell_off);
  if (p->cells) (void)hipFree(p->cells);
  *p = SweepPlan();
}

int plan_create(SweepPlan *p, Extent3 e, int nang, int cmom, int TX, int TY, int TZ, int waves) {
  plan_free(p);
  if (e.nx < 1 || e.ny < 1 || e.nz < 1) return fail(SNAP_ERR_INVALID, "empty chunk");
  if (nang < 1 || nang > 64)
    return fail(SNAP_ERR_UNSUPPORTED, "the tiled sweep handles 1..64 angles per octant, got %d",
                nang);
  if (cmom < 1 || cmom > 16) return fail(SNAP_ERR_INVALID, "cmom must be 1..16, got %d", cmom);
  if (TX < 1 || TY < 1 || TZ < 1 || TX > 32 || TY > 32 || TZ > 32)
    return fail(SNAP_ERR_INVALID, "tile edges must be 1..32");
  if (waves < 1 || waves > 16) return fail(SNAP_ERR_INVALID, "waves per tile must be 1..16");
  p->ext = e;
  p->TX = TX;
  p->TY = TY;
  p->TZ = TZ;
  p->nTx = (e.nx + TX - 1) / TX;
  p->nTy = (e.ny + TY - 1) / TY;
  p->nTz = (e.nz + TZ - 1) / TZ;
  if (p->nTx > 1023 || p->nTy > 1023 || p->nTz > 1023)
    return fail(SNAP_ERR_INVALID, "too many tiles along one axis");
  p->ntiles = p->nTx * p->nTy * p->nTz;
  p->nang = nang;
  p->cmom = cmom;
  p->waves = waves;
  p->cpw = pick_cpw(nang);
  // cells per moment batch: enough (cell, moment) sums to fill a wavefront
  {
    int sbm = std::min(64 / cmom, 8);
    if (sbm < 1) sbm = 1;
    p->sb = std::max(p->cpw, (sbm / p->cpw) * p->cpw);
  }
  // tile diagonals in sweep order
  p->nplanes = p->nTx + p->nTy + p->nTz - 2;
  std::vector<int> tiles;
  p->plane_off.assign(p->nplanes + 1, 0);
  for (int D = 0; D < p->nplanes; D++) {
    p->plane_off[D] = (int)tiles.size();
    for (int tz = 0; tz < p->nTz; tz++)
      for (int ty = 0; ty < p->nTy; ty++) {
        const int tx = D - ty - tz;
        if (tx >= 0 && tx < p->nTx) tiles.push_back(tx | (ty << 10) | (tz << 20));
      }
  }
  p->plane_off[p->nplanes] = (int)tiles.size();

  // cell diagonals inside a full tile
  p->ncellplanes = TX + TY + TZ - 2;
  std::vector<int16_t> cells;
  std::vector<int> coff(p->ncellplanes + 1, 0);
  p->maxdiag = 0;
  for (int d = 0; d < p->ncellplanes; d++) {
    coff[d] = (int)cells.size();
    for (int k = 0; k < TZ; k++)
      for (int j = 0; j < TY; j++) {
        const int i = d - j - k;
        if (i >= 0 && i < TX) cells.push_back((int16_t)(i | (j << 5) | (k << 10)));
      }
    p->maxdiag = std::max(p->maxdiag, (int)cells.size() - coff[d]);
  }
  coff[p->ncellplanes] = (int)cells.size();
  // every wavefront owns at most kMaxSteps cell groups of a diagonal
  const int need_waves = (p->maxdiag + kMaxSteps * p->cpw - 1) / (kMaxSteps * p->cpw);
  if (p->waves < need_waves) p->waves = need_waves;
  if (p->waves > 16)
    return fail(SNAP_ERR_INVALID, "tile %dx%dx%d needs %d waves per workgroup (> 16)", TX, TY, TZ,
                p->waves);
  if (p->lds_bytes(8, false) > 160 * 1024)
    return fail(SNAP_ERR_INVALID, "tile %dx%dx%d with %d angles needs %zu B of LDS (> 160 KiB)",
                TX, TY, TZ, nang, p->lds_bytes(8, false));

  SNAP_HIP_TRY(hipMalloc((void **)&p->plane_tiles, sizeof(int) * tiles.size()));
  SNAP_HIP_TRY(hipMalloc((void **)&p->cell_off, sizeof(int) * coff.size()));
  SNAP_HIP_TRY(hipMalloc((void **)&p->cells, sizeof(int16_t) * cells.size()));
  SNAP_HIP_TRY(hipMemcpy(p->plane_tiles, tiles.data(), sizeof(int) * tiles.size(),
                         hipMemcpyHostToDevice));
  SNAP_HIP_TRY(hipMemcpy(p->cell_off, coff.data(), sizeof(int) * coff.size(),
                         hipMemcpyHostToDevice));
  SNAP_HIP_TRY(hipMemcpy(p->cells, cells.data(), sizeof(int16_t) * cells.size(),
                         hipMemcpyHostToDevice));
  return SNAP_OK;
}

// ---------------------------------------------------------------------------
// kernel
// ---------------------------------------------------------------------------
// Arguments shared by both variants; the dense variant (the solver's own
// [g][cell][record] layout) addresses everything from these base pointers.
struct SweepArgs {
  Extent3 ext;
  int TX, TY, TZ, nTx, nTy, nTz, ntiles;
  int nang, cmom, ncellplanes, ncells_tile, sb;
  int ng_job, corner0, fixup, recompute_dinv;
  double hi, hj, hk;
  const double *mu, *eta, *xi, *w, *ec;
  const double *q;        // dense: [ng_job][cells][cmom]
  const void *dinv;       // dense: [ng_job][cells][nang] of Real
  const void *txs;        // dense: [ng_job][cells] of Real
  double *part_flux;      // [8][ng_job][cells]
  double *part_fluxm;     // [8][ng_job][cells][cmom-1]
  void *faces;
  const int *plane_tiles;
  const int *cell_off;
  const int16_t *cells;
};

// Arguments of the general variant (the drop-in entry point): Realm affine
// accessors, ghost faces, and flux accumulation by device atomics.
struct SweepGeneral {
  Acc3 qtot, dinv, txs, flux, fluxm;
  Acc2 gx, gy, gz;
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a*b + c: one rounding (fused) in the fast variants, two in the exact one
template <bool EXACT>
__device__ inline double madd(double a, double b, double c) {
  if constexpr (EXACT)
    return a * b + c;
  else
    return __builtin_fma(a, b, c);
}
__device__ inline float maddf(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Tile sweep.  Real: arithmetic and face type.  CPW: cells per wavefront.
// EXACT: the reference's operation order with separate roundings (bit-equal
// to the CPU oracle); otherwise fused multiply-adds.  GEN: general accessors,
// ghost faces and atomic accumulation (the drop-in entry point).
template <typename Real, int CPW, bool EXACT, bool GEN>
__global__ void __launch_bounds__(1024) k_sweep_tiles(SweepArgs A, SweepGeneral G, int plane_off) {
  constexpr bool kFP64 = std::is_same<Real, double>::value;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int TX = A.TX, TY = A.TY, TZ = A.TZ, nang = A.nang, cmom = A.cmom, SB = A.sb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const bool recompute = A.recompute_dinv != 0;
  const int ncp = A.ncellplanes;
  const SweepLds lay = sweep_lds(TX, TY, TZ, nang, cmom, W, SB, A.ncells_tile, ncp, CPW, (int)sizeof(Real));
  const int fxn = TY * TZ * nang, fyn = TX * TZ * nang, fzn = TX * TY * nang;
  Real *fx = (Real *)smem_raw;
  Real *fy = fx + fxn;
  Real *fz = fy + fyn;
  double *ecs = (double *)(smem_raw + lay.ecs);   // [cmom][nang]
  double *ect = (double *)(smem_raw + lay.ect);   // [nang][cmom]
  double *pwb = (double *)(smem_raw + lay.pwb) + wave * SB * nang;  // this wave's w*pc batch
  int *cix = (int *)(smem_raw + lay.cix) + wave * SB;               // batch cell indices
  double *qsl = (double *)(smem_raw + lay.qsl) + wave * kMaxSteps * CPW * cmom;  // source moments
  int16_t *tcells = (int16_t *)(smem_raw + lay.cells);              // cell list of the tile
  int *toff = (int *)(smem_raw + lay.coff);                         // its diagonal offsets

  const int s = blockIdx.y;
  const int corner = A.corner0 + s / A.ng_job;
  const int g = s % A.ng_job;
  const int tp = A.plane_tiles[plane_off + blockIdx.x];
  const int tx = tp & 1023, ty = (tp >> 10) & 1023, tz = (tp >> 20) & 1023;
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int nx = A.ext.nx, ny = A.ext.ny, nz = A.ext.nz;
  const int ex = min(TX, nx - x0), ey = min(TY, ny - y0), ez = min(TZ, nz - z0);
  const bool sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
  const int tile_id = (tz * A.nTy + ty) * A.nTx + tx;
  const int per_tile = fxn + fyn + fzn;
  Real *faces_s = (Real *)A.faces + (size_t)s * A.ntiles * per_tile;
  const int ncells = nx * ny * nz;
  // dense per-sweep bases
  const double *qg = A.q + (size_t)g * ncells * cmom;
  const Real *dg = (const Real *)A.dinv + (size_t)g * ncells * nang;
  const Real *tg = (const Real *)A.txs + (size_t)g * ncells;
  double *pfg = A.part_flux + ((size_t)corner * A.ng_job + g) * ncells;
  double *pmg = A.part_fluxm + ((size_t)corner * A.ng_job + g) * ncells * (cmom - 1);

  // ---- per-lane constants: lane = cell slot * nang + angle
  const int cl = (CPW == 1) ? 0 : lane / nang;
  const int a = lane - cl * nang;
  const bool lane_ok = (cl < CPW) && (a < nang);
  const unsigned long long seg =
      lane_ok ? (((nang >= 64) ? ~0ull : ((1ull << nang) - 1ull)) << (cl * nang)) : 0ull;
  const int aa = lane_ok ? a : 0;
  const int per_step = W * CPW;

  // ---- tables and inflow faces (from the upstream tile, the ghost array,
  // or vacuum) into LDS
  for (int idx = threadIdx.x; idx < A.ncells_tile; idx += blockDim.x) tcells[idx] = A.cells[idx];
  for (int idx = threadIdx.x; idx <= ncp; idx += blockDim.x) toff[idx] = A.cell_off[idx];
  for (int idx = threadIdx.x; idx < cmom * nang; idx += blockDim.x) {
    const double v = A.ec[(size_t)corner * cmom * nang + idx];
    const int l = idx / nang, b = idx - l * nang;
    ecs[idx] = v;
    ect[b * cmom + l] = v;
  }
  for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) {
    const int a2 = idx % nang, slot = idx / nang, j = slot % TY, k = slot / TY;
    Real v = Real(0);
    if (j < ey && k < ez) {
      if (tx > 0) {
        v = faces_s[(size_t)(tile_id - 1) * per_tile + idx];
      } else if (GEN && G.gx.base) {
        const int py = syp ? y0 + j : ny - 1 - (y0 + j), pz = szp ? z0 + k : nz - 1 - (z0 + k);
        v = (Real)((const double *)G.gx.at(py, pz, 0))[a2];
      }
    }
    fx[idx] = v;
  }
  for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) {
    const int a2 = idx % nang, slot = idx / nang, i = slot % TX, k = slot / TX;
    Real v = Real(0);
    if (i < ex && k < ez) {
      if (ty > 0) {
        v = faces_s[(size_t)(tile_id - A.nTx) * per_tile + fxn + idx];
      } else if (GEN && G.gy.base) {
        const int px = sxp ? x0 + i : nx - 1 - (x0 + i), pz = szp ? z0 + k : nz - 1 - (z0 + k);
        v = (Real)((const double *)G.gy.at(px, pz, 0))[a2];
      }
    }
    fy[idx] = v;
  }
  for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) {
    const int a2 = idx % nang, slot = idx / nang, i = slot % TX, j = slot / TX;
    Real v = Real(0);
    if (i < ex && j < ey) {
      if (tz > 0) {
        v = faces_s[(size_t)(tile_id - A.nTx * A.nTy) * per_tile + fxn + fyn + idx];
      } else if (GEN && G.gz.base) {
        const int px = sxp ? x0 + i : nx - 1 - (x0 + i), py = syp ? y0 + j : ny - 1 - (y0 + j);
        v = (Real)((const double *)G.gz.at(px, py, 0))[a2];
      }
    }
    fz[idx] = v;
  }

  const double mu_d = A.mu[aa], eta_d = A.eta[aa], xi_d = A.xi[aa], w_d = A.w[aa];
  // EXACT keeps (psi*mu)*hi as the reference writes it; the others fold mu*hi.
  const Real mu_a = (Real)mu_d, eta_a = (Real)eta_d, xi_a = (Real)xi_d;
  const Real hi = (Real)A.hi, hj = (Real)A.hj, hk = (Real)A.hk;
  const Real muh = (Real)(mu_d * A.hi), etah = (Real)(eta_d * A.hj), xih = (Real)(xi_d * A.hk);
  // dinv recomputed as CalculateGeometryParam forms it (expxs.cc:378-381)
  const double g_hm = A.hi * mu_d, g_he = A.hj * eta_d, g_hx = A.hk * xi_d;
  const float g_sum32 = (float)(g_hm + g_he + g_hx);
  const Real tolr = (Real)1.0e-12;
  const int nm1 = cmom - 1;
  const bool q_in_lanes = cmom <= nang;
  __syncthreads();

  // ---- register prefetch of this wave's cells on the next diagonal: dinv
  // (lane a holds angle a), the source moments (lane l of the cell's segment
  // holds moment l when cmom <= nang) and t_xs.
  double n_dv[kMaxSteps], n_q[kMaxSteps], n_tx[kMaxSteps];
  int n_cell[kMaxSteps];
  auto cell_of = [&](int idx, int &i, int &j, int &k) -> bool {
    const int pk = tcells[idx];
    i = pk & 31;
    j = (pk >> 5) & 31;
    k = (pk >> 10) & 31;
    return (i < ex) && (j < ey) && (k < ez);
  };
  auto prefetch = [&](int dd) {
    const int c0 = toff[dd], c1 = toff[dd + 1];
#pragma unroll
    for (int st = 0; st < kMaxSteps; st++) {
      const int my = c0 + st * per_step + wave * CPW + cl;
      int i, j, k;
      n_dv[st] = 0.0;
      n_q[st] = 0.0;
      n_tx[st] = 0.0;
      n_cell[st] = -1;
      if (lane_ok && my < c1 && cell_of(my, i, j, k)) {
        const int X = x0 + i, Y = y0 + j, Z = z0 + k;
        const int px = sxp ? X : nx - 1 - X, py = syp ? Y : ny - 1 - Y, pz = szp ? Z : nz - 1 - Z;
        const int cell = (pz * ny + py) * nx + px;
        n_cell[st] = cell;
        if constexpr (GEN) {
          if (!recompute) n_dv[st] = (double)((const Real *)G.dinv.at(px, py, pz, 0))[a];
          if (q_in_lanes && a < cmom) n_q[st] = ((const double *)G.qtot.at(px, py, pz, 0))[a];
          n_tx[st] = (double)*(const Real *)G.txs.at(px, py, pz, 0);
        } else {
          if (!recompute) n_dv[st] = (double)dg[(size_t)cell * nang + a];
          if (q_in_lanes && a < cmom) n_q[st] = qg[(size_t)cell * cmom + a];
          n_tx[st] = (double)tg[cell];
        }
      }
    }
  };

  int nbuf = 0;  // cells in this wave's moment batch (wave-uniform)
  auto flush = [&]() {
    // per-cell flux moments over angles in ascending angle order
    // (sweep.cc:568-591): phi_l = sum_a ec[l][a] * (w[a]*pc[a])
    wave_sync();
    for (int o = lane; o < nbuf * cmom; o += 64) {
      const int kk = o / cmom, l = o - kk * cmom;
      const int cell = cix[kk];
      if (cell < 0) continue;
      const double *v = pwb + kk * nang;
      double sum = 0.0;
      for (int b = 0; b < nang; b++) sum = madd<EXACT>(ect[b * cmom + l], v[b], sum);
      if constexpr (GEN) {
        const int px = cell % nx, py = (cell / nx) % ny, pz = cell / (nx * ny);
        double *dst = (l == 0) ? (double *)G.flux.at(px, py, pz, 0)
                               : (double *)G.fluxm.at(px, py, pz, 0) + (l - 1);
        atomicAdd(dst, sum);
      } else {
        if (l == 0)
          pfg[cell] = sum;
        else
          pmg[(size_t)cell * nm1 + (l - 1)] = sum;
      }
    }
    wave_sync();
    nbuf = 0;
  };

  prefetch(0);
  for (int d = 0; d < ncp; ++d) {
    const int c0 = toff[d], c1 = toff[d + 1];
    const int nsteps = (c1 - c0 + per_step - 1) / per_step;
    double c_dv[kMaxSteps], c_tx[kMaxSteps];
    int c_cell[kMaxSteps];
    // source moments of this diagonal's cells to the wave's LDS slots
#pragma unroll
    for (int st = 0; st < kMaxSteps; st++) {
      c_dv[st] = n_dv[st];
      c_tx[st] = n_tx[st];
      c_cell[st] = n_cell[st];
      if (q_in_lanes && lane_ok && a < cmom) qsl[(st * CPW + cl) * cmom + a] = n_q[st];
    }
    if (d + 1 < ncp) prefetch(d + 1);
    wave_sync();
#pragma unroll
    for (int st = 0; st < kMaxSteps; ++st) {
      if (st >= nsteps) break;
      const int my = c0 + st * per_step + wave * CPW + cl;
      const int cell = c_cell[st];
      const bool valid = cell >= 0;
      double pcw = 0.0;
      if (valid) {
        const int pk = tcells[my];
        const int i = pk & 31, j = (pk >> 5) & 31, k = (pk >> 10) & 31;
        // angular source (sweep.cc:331-345)
        const double *q;
        if (q_in_lanes) {
          q = qsl + (st * CPW + cl) * cmom;
        } else if constexpr (GEN) {
          const int px = cell % nx, py = (cell / nx) % ny, pz = cell / (nx * ny);
          q = (const double *)G.qtot.at(px, py, pz, 0);
        } else {
          q = qg + (size_t)cell * cmom;
        }
        Real psi;
        if constexpr (kFP64) {
          double ps = q[0];
          for (int l = 1; l < cmom; l++) ps = madd<EXACT>(ecs[l * nang + a], q[l], ps);
          psi = ps;
        } else {
          float pf = (float)q[0];
          for (int l = 1; l < cmom; l++) pf = maddf((float)ecs[l * nang + a], (float)q[l], pf);
          psi = pf;
        }
        const Real txs = (Real)c_tx[st];
        Real dv;
        if (recompute) {
          if constexpr (kFP64)
            dv = 1.0 / ((((double)txs + g_hm) + g_he) + g_hx);
          else
            dv = 1.0f / ((float)txs + g_sum32);
        } else {
          dv = (Real)c_dv[st];
        }
        Real *sx = fx + (k * TY + j) * nang + a;
        Real *sy = fy + (k * TX + i) * nang + a;
        Real *sz = fz + (j * TX + i) * nang + a;
        const Real psii = *sx, psij = *sy, psik = *sz;
        // initial diamond-difference solution (sweep.cc:352-400)
        Real pc = psi;
        if constexpr (kFP64 && EXACT) {
          pc += psii * mu_a * hi;
          pc += psij * eta_a * hj;
          pc += psik * xi_a * hk;
        } else if constexpr (kFP64) {
          pc = __builtin_fma(psii, muh, pc);
          pc = __builtin_fma(psij, etah, pc);
          pc = __builtin_fma(psik, xih, pc);
        } else {
          pc = maddf(psii, muh, pc);
          pc = maddf(psij, etah, pc);
          pc = maddf(psik, xih, pc);
        }
        pc *= dv;
        Real ox, oy, oz;
        if (A.fixup) {
          // negative-flux fixup (sweep.cc:402-505; gpu_sweep.cu:1556-1630).
          // Each pass that continues clears at least one of the 3*nang
          // flags or ends the loop, so 3*nang+2 passes always suffice.
          Real hvx = Real(1), hvy = Real(1), hvz = Real(1);
          Real fxx = Real(0), fxy = Real(0), fxz = Real(0);
          unsigned old_neg = 0;
          for (int it = 0; it < 3 * nang + 2; ++it) {
            fxx = Real(2) * pc - psii;
            fxy = Real(2) * pc - psij;
            fxz = Real(2) * pc - psik;
            const bool nxn = fxx < Real(0), nyn = fxy < Real(0), nzn = fxz < Real(0);
            if (nxn) hvx = Real(0);
            if (nyn) hvy = Real(0);
            if (nzn) hvz = Real(0);
            const unsigned neg = (unsigned)(__popcll(__ballot(nxn) & seg) +
                                            __popcll(__ballot(nyn) & seg) +
                                            __popcll(__ballot(nzn) & seg));
            if (neg == old_neg) break;
            old_neg = neg;
            Real den;
            if constexpr (kFP64 && EXACT) {
              pc = psi + 0.5 * (psii * mu_a * hi * (1.0 + hvx) + psij * eta_a * hj * (1.0 + hvy) +
                                psik * xi_a * hk * (1.0 + hvz));
              den = (pc <= 0.0) ? 0.0 : (txs + mu_a * hi * hvx + eta_a * hj * hvy + xi_a * hk * hvz);
            } else {
              const Real t = psii * muh * (Real(1) + hvx) + psij * etah * (Real(1) + hvy) +
                             psik * xih * (Real(1) + hvz);
              pc = psi + Real(0.5) * t;
              den = (pc <= Real(0)) ? Real(0) : (txs + muh * hvx + etah * hvy + xih * hvz);
            }
            pc = (den < tolr) ? Real(0) : pc / den;
          }
          ox = fxx * hvx;
          oy = fxy * hvy;
          oz = fxz * hvz;
        } else {
          ox = Real(2) * pc - psii;
          oy = Real(2) * pc - psij;
          oz = Real(2) * pc - psik;
        }
        *sx = ox;
        *sy = oy;
        *sz = oz;
        pcw = w_d * (double)pc;
      }
      // ---- batch this step's cells for the moment sums
      if (lane_ok) {
        pwb[(nbuf + cl) * nang + a] = pcw;
        if (a == 0) cix[nbuf + cl] = (my < c1) ? cell : -1;
      }
      nbuf += CPW;
      if (nbuf + CPW > SB) flush();
    }
    __syncthreads();
  }
  if (nbuf > 0) flush();

  // ---- tile outflow faces: to the face buffer, and at the chunk's far side
  // to the ghost arrays
  Real *out = faces_s + (size_t)tile_id * per_tile;
  for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) {
    out[idx] = fx[idx];
    if (GEN && tx == A.nTx - 1 && G.gx.base) {
      const int a2 = idx % nang, slot = idx / nang, j = slot % TY, k = slot / TY;
      if (j < ey && k < ez) {
        const int py = syp ? y0 + j : ny - 1 - (y0 + j), pz = szp ? z0 + k : nz - 1 - (z0 + k);
        ((double *)G.gx.at(py, pz, 0))[a2] = (double)fx[idx];
      }
    }
  }
  for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) {
    out[fxn + idx] = fy[idx];
    if (GEN && ty == A.nTy - 1 && G.gy.base) {
      const int a2 = idx % nang, slot = idx / nang, i = slot % TX, k = slot / TX;
      if (i < ex && k < ez) {
        const int px = sxp ? x0 + i : nx - 1 - (x0 + i), pz = szp ? z0 + k : nz - 1 - (z0 + k);
        ((double *)G.gy.at(px, pz, 0))[a2] = (double)fy[idx];
      }
    }
  }
  for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) {
    out[fxn + fyn + idx] = fz[idx];
    if (GEN && tz == A.nTz - 1 && G.gz.base) {
      const int a2 = idx % nang, slot = idx / nang, i = slot % TX, j = slot / TX;
      if (i < ex && j < ey) {
        const int px = sxp ? x0 + i : nx - 1 - (x0 + i), py = syp ? y0 + j : ny - 1 - (y0 + j);
        ((double *)G.gz.at(px, py, 0))[a2] = (double)fz[idx];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
typedef void (*SweepKernel)(SweepArgs, SweepGeneral, int);

template <typename Real, bool EXACT, bool GEN>
static SweepKernel kernel_for(int cpw) {
  switch (cpw) {
    case 1: return k_sweep_tiles<Real, 1, EXACT, GEN>;
    case 2: return k_sweep_tiles<Real, 2, EXACT, GEN>;
    case 3: return k_sweep_tiles<Real, 3, EXACT, GEN>;
    case 4: return k_sweep_tiles<Real, 4, EXACT, GEN>;
    case 5: return k_sweep_tiles<Real, 5, EXACT, GEN>;
    case 6: return k_sweep_tiles<Real, 6, EXACT, GEN>;
    case 8: return k_sweep_tiles<Real, 8, EXACT, GEN>;
    default: return nullptr;
  }
}

static int ensure_lds_attr(SweepKernel k, size_t bytes) {
  // dynamic LDS above 64 KiB must be requested per kernel
  static std::mutex mu;
  static std::vector<std::pair<SweepKernel, size_t>> done;
  std::lock_guard<std::mutex> lock(mu);
  for (auto &d : done)
    if (d.first == k && d.second >= bytes) return SNAP_OK;
  SNAP_HIP_TRY(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)bytes));
  done.emplace_back(k, bytes);
  return SNAP_OK;
}

int sweep_enqueue(const SweepLaunch &L, hipStream_t st, hipEvent_t *events) {
  const SweepPlan &p = *L.plan;
  if (L.nsweeps < 1 || L.nsweeps > 65535) return fail(SNAP_ERR_INVALID, "bad sweep count");
  if (L.quad->nang != p.nang || L.quad->cmom != p.cmom)
    return fail(SNAP_ERR_INVALID, "quadrature does not match the sweep plan");
  if (L.general && (L.nsweeps != 1 || L.ng_job != 1))
    return fail(SNAP_ERR_INVALID, "the general sweep handles one (corner, group) per call");
  if ((int64_t)p.ext.nx * p.ext.ny * p.ext.nz >= (int64_t)1 << 31)
    return fail(SNAP_ERR_UNSUPPORTED, "chunks of 2^31 cells or more");
  SweepArgs A;
  memset(&A, 0, sizeof(A));
  SweepGeneral G;
  memset(&G, 0, sizeof(G));
  A.ext = p.ext;
  A.TX = p.TX;
  A.TY = p.TY;
  A.TZ = p.TZ;
  A.nTx = p.nTx;
  A.nTy = p.nTy;
  A.nTz = p.nTz;
  A.ntiles = p.ntiles;
  A.nang = p.nang;
  A.cmom = p.cmom;
  A.ncellplanes = p.ncellplanes;
  A.sb = p.sb;
  A.ncells_tile = p.TX * p.TY * p.TZ;
  A.recompute_dinv = L.recompute_dinv ? 1 : 0;
  A.ng_job = L.ng_job;
  A.corner0 = L.corner0;
  A.fixup = L.fixup;
  A.hi = L.hi;
  A.hj = L.hj;
  A.hk = L.hk;
  A.mu = L.quad->mu;
  A.eta = L.quad->eta;
  A.xi = L.quad->xi;
  A.w = L.quad->w;
  A.ec = L.quad->ec;
  A.q = L.q;
  A.dinv = L.dinv_dense;
  A.txs = L.txs_dense;
  A.part_flux = L.part_flux;
  A.part_fluxm = L.part_fluxm;
  A.faces = L.faces;
  A.plane_tiles = p.plane_tiles;
  A.cell_off = p.cell_off;
  A.cells = p.cells;
  if (L.general) {
    G.qtot = L.qtot;
    G.dinv = L.dinv;
    G.txs = L.txs;
    G.flux = L.flux;
    G.fluxm = L.fluxm;
    G.gx = L.gx;
    G.gy = L.gy;
    G.gz = L.gz;
    if (!L.recompute_dinv && !G.dinv.base) return fail(SNAP_ERR_INVALID, "dinv is required");
  } else {
    if (!A.q || !A.txs || !A.part_flux || (p.cmom > 1 && !A.part_fluxm))
      return fail(SNAP_ERR_INVALID, "dense sweep inputs missing");
    if (!L.recompute_dinv && !A.dinv) return fail(SNAP_ERR_INVALID, "dinv is required");
  }
  const size_t lds = p.lds_bytes(L.fp32 ? 4 : 8, L.recompute_dinv);
  SweepKernel k;
  if (L.general)
    k = L.fp32 ? kernel_for<float, false, true>(p.cpw)
               : (L.exact ? kernel_for<double, true, true>(p.cpw) : kernel_for<double, false, true>(p.cpw));
  else
    k = L.fp32 ? kernel_for<float, false, false>(p.cpw)
               : (L.exact ? kernel_for<double, true, false>(p.cpw) : kernel_for<double, false, false>(p.cpw));
  if (!k) return fail(SNAP_ERR_INVALID, "no sweep kernel for %d cells per wave", p.cpw);
  SNAP_TRY(ensure_lds_attr(k, lds));
  const dim3 block(p.waves * 64);
  for (int D = 0; D < p.nplanes; D++) {
    const int count = p.plane_off[D + 1] - p.plane_off[D];
    if (events) SNAP_HIP_TRY(hipEventRecord(events[2 * D], st));
    hipLaunchKernelGGL(k, dim3(count, L.nsweeps), block, lds, st, A, G, p.plane_off[D]);
    if (events) SNAP_HIP_TRY(hipEventRecord(events[2 * D + 1], st));
  }
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// ---------------------------------------------------------------------------
// the drop-in context (initialize_gpu_context) and run_gpu_sweep
// ---------------------------------------------------------------------------
namespace {
struct StreamWork {
  hipStream_t stream;
  void *faces;
  size_t bytes;
};
struct DeviceContext {
  bool valid = false;
  QuadDev quad;
  Extent3 chunk{};
  std::vector<std::pair<Extent3, SweepPlan *>> plans;
  std::vector<StreamWork> work;
};
constexpr int kMaxDevices = 64;
DeviceContext g_ctx[kMaxDevices];
std::mutex g_ctx_mutex;

int current_device(int *dev) {
  SNAP_HIP_TRY(hipGetDevice(dev));
  if (*dev < 0 || *dev >= kMaxDevices) return fail(SNAP_ERR_INVALID, "device %d out of range", *dev);
  return SNAP_OK;
}

void context_release(DeviceContext &c) {
  quad_free(&c.quad);
  for (auto &pp : c.plans) {
    plan_free(pp.second);
    delete pp.second;
  }
  c.plans.clear();
  for (auto &w : c.work)
    if (w.faces) (void)hipFree(w.faces);
  c.work.clear();
  c.valid = false;
}
}  // namespace

const QuadDev *context_quad(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  return g_ctx[dev].valid ? &g_ctx[dev].quad : nullptr;
}

}  // namespace snap

using namespace snap;

extern "C" int snap_hip_initialize_gpu_context(const double *ec, const double *mu,
                                               const double *eta, const double *xi,
                                               const double *w, int num_angles, int num_moments,
                                               int num_octants, int nx_per_chunk,
                                               int ny_per_chunk, int nz_per_chunk) {
  if (!ec || !mu || !eta || !xi || !w) return fail(SNAP_ERR_INVALID, "NULL quadrature array");
  if (num_octants != 8)
    return fail(SNAP_ERR_UNSUPPORTED, "the mini-KBA sweep is 3-D: num_octants must be 8");
  if (num_moments < 1 || num_moments > 4) return fail(SNAP_ERR_INVALID, "num_moments must be 1..4");
  if (nx_per_chunk < 1 || ny_per_chunk < 1 || nz_per_chunk < 1)
    return fail(SNAP_ERR_INVALID, "empty chunk");
  int dev;
  SNAP_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  DeviceContext &c = g_ctx[dev];
  context_release(c);
  SNAP_TRY(quad_upload(&c.quad, ec, mu, eta, xi, w, num_angles, num_moments));
  c.chunk = Extent3{nx_per_chunk, ny_per_chunk, nz_per_chunk};
  c.valid = true;
  return SNAP_OK;
}

extern "C" int snap_hip_run_gpu_sweep(const snap_sweep_args_t *args, snap_stream_t stream) {
  if (!args) return fail(SNAP_ERR_INVALID, "NULL sweep arguments");
  const snap_sweep_args_t &a = *args;
  if (a.mms_source) return fail(SNAP_ERR_UNSUPPORTED, "MMS sources are not built (steady fixed-source path)");
  if (a.vdelt != 0.0) return fail(SNAP_ERR_UNSUPPORTED, "time-dependent sweeps (vdelt != 0) are not built");
  if (a.stride_x_positive != ((a.corner & 1) != 0) || a.stride_y_positive != ((a.corner & 2) != 0) ||
      a.stride_z_positive != ((a.corner & 4) != 0) || a.corner < 0 || a.corner > 7)
    return fail(SNAP_ERR_INVALID, "corner %d does not match the stride directions", a.corner);
  int dev;
  SNAP_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  DeviceContext &c = g_ctx[dev];
  if (!c.valid) return fail(SNAP_ERR_NO_CONTEXT, "initialize_gpu_context was not called on device %d", dev);
  if (a.num_angles != c.quad.nang || a.num_moments != c.quad.nmom)
    return fail(SNAP_ERR_INVALID, "angles/moments differ from the context");
  const Extent3 e{a.x_range, a.y_range, a.z_range};
  if (e.nx < 1 || e.ny < 1 || e.nz < 1) return fail(SNAP_ERR_INVALID, "empty chunk");
  // plan per chunk shape
  SweepPlan *plan = nullptr;
  for (auto &pp : c.plans)
    if (pp.first.nx == e.nx && pp.first.ny == e.ny && pp.first.nz == e.nz) plan = pp.second;
  if (!plan) {
    int TX, TY, TZ, waves;
    plan_default_tile(e, c.quad.nang, c.quad.cmom, &TX, &TY, &TZ, &waves);
    plan = new SweepPlan();
    int rc = plan_create(plan, e, c.quad.nang, c.quad.cmom, TX, TY, TZ, waves);
    if (rc != SNAP_OK) {
      delete plan;
      return rc;
    }
    c.plans.emplace_back(e, plan);
  }
  // per-stream face workspace (gpu_sweep.cu keeps 12 per-stream scratch slots)
  const size_t need = plan->face_doubles_per_sweep() * sizeof(double);
  hipStream_t st = (hipStream_t)stream;
  StreamWork *wk = nullptr;
  for (auto &w : c.work)
    if (w.stream == st) wk = &w;
  if (!wk) {
    c.work.push_back(StreamWork{st, nullptr, 0});
    wk = &c.work.back();
  }
  if (wk->bytes < need) {
    if (wk->faces) {
      SNAP_HIP_TRY(hipStreamSynchronize(st));
      SNAP_HIP_TRY(hipFree(wk->faces));
      wk->faces = nullptr;
      wk->bytes = 0;
    }
    if (hipMalloc(&wk->faces, need) != hipSuccess)
      return fail(SNAP_ERR_NO_MEMORY, "face workspace of %zu bytes", need);
    wk->bytes = need;
  }
  snap_rect3_t b;
  for (int i = 0; i < 3; i++) b.lo[i] = a.origin[i];
  b.hi[0] = a.origin[0] + e.nx - 1;
  b.hi[1] = a.origin[1] + e.ny - 1;
  b.hi[2] = a.origin[2] + e.nz - 1;
  SweepLaunch L;
  memset(&L, 0, sizeof(L));
  L.plan = plan;
  L.quad = &c.quad;
  L.nsweeps = 1;
  L.ng_job = 1;
  L.corner0 = a.corner;
  L.fixup = a.fixup ? 1 : 0;
  L.hi = a.hi;
  L.hj = a.hj;
  L.hk = a.hk;
  SNAP_TRY(single_acc3(a.qtot, b, &L.qtot, "qtot"));
  SNAP_TRY(single_acc3(a.dinv, b, &L.dinv, "dinv"));
  SNAP_TRY(single_acc3(a.t_xs, b, &L.txs, "t_xs"));
  SNAP_TRY(single_acc3(a.flux, b, &L.flux, "flux"));
  SNAP_TRY(single_acc3(a.fluxm, b, &L.fluxm, "fluxm", c.quad.cmom == 1));
  if (c.quad.cmom > 1 && !L.fluxm.base) return fail(SNAP_ERR_INVALID, "fluxm is required when cmom > 1");
  L.general = true;
  // ghost points are absolute (y,z) / (x,z) / (x,y): fold the origin in
  auto ghost = [](const snap_accessor2_t &gacc, int64_t o0, int64_t o1) {
    Acc2 r;
    r.base = gacc.base ? (char *)gacc.base + o0 * gacc.stride[0] + o1 * gacc.stride[1] : nullptr;
    r.s0 = gacc.stride[0];
    r.s1 = gacc.stride[1];
    r.sg = 0;
    return r;
  };
  L.gx = ghost(a.ghostx, a.origin[1], a.origin[2]);
  L.gy = ghost(a.ghosty, a.origin[0], a.origin[2]);
  L.gz = ghost(a.ghostz, a.origin[0], a.origin[1]);
  L.faces = wk->faces;
  L.fp32 = false;
  L.exact = true;
  return sweep_enqueue(L, st);
}
