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
// Lanes carry angles: each lane holds APL (1 or 2) consecutive angles of one
// cell, a cell spans LPC = ceil(nang/APL) lanes and a 64-lane wavefront holds
// CPW = 64/LPC cells of the current diagonal.  The fixup's negative-flux count
// is cell-global over angles (sweep.cc:437-505), so each pass counts with wave
// ballots masked to the cell's lane segment.  Flux moments are reduced per
// cell through LDS in ascending angle order, as the CPU variant sums them
// (sweep.cc:568-591).
//
// The EXACT FP64 instantiation keeps the reference's operation order
// everywhere and the library is built with -ffp-contract=off, so its results
// are the CPU arithmetic bit for bit.  The fused FP64 instantiation uses
// multiply-adds; the FP32 one (mixed precision) runs the per-angle arithmetic
// and faces in float and the moment sums in double.
#include <algorithm>
#include <cstring>
#include <memory>
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

// Most cell groups (steps) one wavefront owns on a cell diagonal; the planner
// picks the wave count so that every diagonal fits.
__host__ __device__ constexpr int max_steps(int apl) { return apl == 1 ? 4 : 2; }
// Workgroup size bound: two angles per lane need more than the 128 registers
// a 16-wave workgroup leaves each lane, so those workgroups stop at 8 waves.
__host__ __device__ constexpr int max_threads(int apl) { return apl == 1 ? 1024 : 512; }
// The time-dependent kernels carry the time-edge inflow and MMS source of
// every prefetched cell as well: they stop at 8 waves (256 registers a lane).
__host__ __device__ constexpr int max_threads(int apl, bool td) {
  return td ? 512 : max_threads(apl);
}

// Row strides of the two ec tables: the per-angle table is read by lane
// (angle) with the moment as the row, the per-moment table by (cell, moment)
// lanes with the angle as the row; fixed strides keep both reads at
// immediate offsets.
__host__ __device__ constexpr int ec_row(int apl) { return 64 * apl; }
constexpr int kEctRow = 16;  // >= cmom (P3 at most)

// faces [x|y|z] (Real) | ec of the corner [cmom][ec_row] and [nang][kEctRow]
// | per-wave w*pc batches [W][SB][nang] | per-wave source moments
// [W][steps][CPW][cmom] | batch cell indices [W][SB] | per tile cell
// {chunk cell, face offsets} [ncells] (int4) | diagonal offsets [ncp+1]
__host__ __device__ inline SweepLds sweep_lds(int TX, int TY, int TZ, int nang, int cmom, int W,
                                              int SB, int ncells_tile, int ncp, int cpw, int apl,
                                              int rb) {
  SweepLds L;
  size_t o = a16((size_t)(TY * TZ + TX * TZ + TX * TY) * nang * rb);
  L.ecs = o;
  o += (size_t)cmom * ec_row(apl) * rb;
  L.ect = o;
  o += (size_t)nang * kEctRow * 8;
  L.pwb = o;
  o += (size_t)W * SB * nang * 8;
  L.qsl = o;
  o += (size_t)W * max_steps(apl) * cpw * cmom * rb;
  L.cix = o;
  o = a16(o + (size_t)W * SB * 4);
  L.cells = o;
  o = a16(o + (size_t)ncells_tile * 16);
  L.coff = o;
  o = a16(o + (size_t)(ncp + 1) * 4);
  L.total = o;
  return L;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
// LDS of the one-cell workgroup: fixup counts [2][16] | per-wave moment
// partials [16][16] | w*pc of every angle (the ordered moment sums)
__host__ __device__ inline size_t wide_lds(int nang) {
  return 2 * 16 * 4 + 16 * 16 * 8 + (size_t)nang * 8;
}

size_t SweepPlan::lds_bytes(int real_bytes, bool recompute_dinv) const {
  (void)recompute_dinv;
  if (wide) return wide_lds(nang);
  return sweep_lds(TX, TY, TZ, nang, cmom, waves, sb, TX * TY * TZ, ncellplanes, cpw, apl,
                   real_bytes).total;
}

size_t SweepPlan::face_doubles_per_sweep() const {
  if (seq)  // in-place pencils: x faces per (ty, tz), y per (tx, tz), z per (tx, ty)
    return ((size_t)nTy * nTz * TY * TZ + (size_t)nTx * nTz * TX * TZ +
            (size_t)nTx * nTy * TX * TY) * nang;
  if (wide)  // in-place x, y and z face planes
    return ((size_t)ext.ny * ext.nz + (size_t)ext.nx * ext.nz + (size_t)ext.nx * ext.ny) * nang;
  return (size_t)ntiles * (TY * TZ + TX * TZ + TX * TY) * nang;
}

static const int kCpwChoices[] = {1, 2, 3, 4, 5, 6, 8};

static int pick_cpw(int lpc) {
  int best = 0;
  for (int c : kCpwChoices)
    if (c * lpc <= 64) best = c;
  return best;
}

// angles per lane: the choice that uses the most of the wavefront's
// angle slots (ties keep one angle per lane)
static int pick_apl(int nang) {
  int best = 0;
  double best_use = -1.0;
  for (int apl = 1; apl <= 2; apl++) {
    const int lpc = (nang + apl - 1) / apl;
    const int cpw = pick_cpw(lpc);
    if (cpw == 0) continue;
    const double use = double(cpw * nang) / double(64 * apl);
    if (use > best_use + 1e-9) {
      best_use = use;
      best = apl;
    }
  }
  return best;
}

// most cells on one cell diagonal i+j+k = d of a TX x TY x TZ tile
static int tile_maxdiag(int TX, int TY, int TZ) {
  int best = 0;
  for (int d = 0; d < TX + TY + TZ - 2; d++) {
    int n = 0;
    for (int k = 0; k < TZ; k++)
      for (int j = 0; j < TY; j++) {
        const int i = d - j - k;
        if (i >= 0 && i < TX) n++;
      }
    best = std::max(best, n);
  }
  return best;
}

void plan_choose(Extent3 e, int nang, int cmom, int real_bytes, int *TX, int *TY, int *TZ,
                 int *waves, int *apl, bool td) {
  // angles per lane: one at FP64 (two would exceed 128 registers per lane
  // and halve the resident waves), the better-filled wavefront at FP32, two
  // whenever one cell's angles exceed a wavefront
  if (*apl <= 0) *apl = (nang > 64) ? 2 : (real_bytes == 4 ? pick_apl(nang) : 1);
  if (*apl < 1 || *apl > 2) return;  // plan_create reports it
  const int lpc = (nang + *apl - 1) / *apl;
  const int cpw = std::max(1, pick_cpw(lpc));
  const int ms = max_steps(*apl), wmax = max_threads(*apl, td) / 64;
  if (*TX <= 0 || *TY <= 0 || *TZ <= 0) {
    // largest candidate tile whose biggest cell diagonal the workgroup
    // covers and whose LDS fits; measured best first: 8x8x8 with FP32 faces
    // and at FP64 up to 24 angles, else 8x8x4 (ms per FP64 sweep, 8x8x8
    // against 8x8x4: 48^3 16 groups 24 angles 4.13 / 5.18, 48^3 16 g 16 a
    // 3.78 / 4.88, 64^3 8 g 24 a 6.81 / 8.15, but 32^3 16 g 32 a 2.91 / 2.47
    // and 64^3 8 g 36 a 12.5 / 10.8; FP32 8x8x8 ahead on all five)
    static const int cand[][3] = {{8, 8, 8}, {8, 8, 4}, {8, 4, 4}, {4, 4, 4}, {4, 4, 2},
                                  {4, 2, 2}, {2, 2, 2}, {2, 2, 1}, {2, 1, 1}, {1, 1, 1}};
    const int first = (real_bytes == 4 || nang <= 24) ? 0 : 1;
    for (int c = first; c < (int)(sizeof(cand) / sizeof(cand[0])); c++) {
      const int tx = std::min(cand[c][0], e.nx), ty = std::min(cand[c][1], e.ny),
                tz = std::min(cand[c][2], e.nz);
      const int need = (tile_maxdiag(tx, ty, tz) + ms * cpw - 1) / (ms * cpw);
      const int w = std::max(need, *waves > 0 ? *waves : (*apl == 1 ? 8 : 6));
      if (w > wmax) continue;
      const int sb = cpw * std::min(std::max(1, 64 / (((cmom + 1) / 2) * cpw)), std::max(1, 8 / cpw));
      const size_t lds = sweep_lds(tx, ty, tz, nang, cmom, w, sb, tx * ty * tz, tx + ty + tz - 2,
                                   cpw, *apl, 8).total;
      if (lds > 160 * 1024) continue;
      *TX = tx;
      *TY = ty;
      *TZ = tz;
      break;
    }
    if (*TX <= 0) *TX = *TY = *TZ = 1;
  }
  if (*waves <= 0) *waves = (*apl == 1) ? 8 : 6;  // plan_create raises it to what the tile needs
}

void plan_free(SweepPlan *p) {
  if (p->plane_tiles) (void)hipFree(p->plane_tiles);
  if (p->cell_off) (void)hipFree(p->cell_off);
  if (p->cells) (void)hipFree(p->cells);
  *p = SweepPlan();
}

// Host part of plan_create: every field of the plan and its three tables
// (tiles by tile diagonal, cell-diagonal offsets and cells inside one tile),
// without touching device memory.
int plan_layout(SweepPlan *p, Extent3 e, int nang, int cmom, int TX, int TY, int TZ, int waves,
                int apl, bool td, std::vector<int> &tiles, std::vector<int> &coff,
                std::vector<int16_t> &cells) {
  *p = SweepPlan();
  if (e.nx < 1 || e.ny < 1 || e.nz < 1) return fail(SNAP_ERR_INVALID, "empty chunk");
  if (nang < 1 || nang > 128)
    return fail(SNAP_ERR_UNSUPPORTED, "the tiled sweep handles 1..128 angles per octant, got %d",
                nang);
  if (cmom < 1 || cmom > 16) return fail(SNAP_ERR_INVALID, "cmom must be 1..16, got %d", cmom);
  if (TX < 1 || TY < 1 || TZ < 1 || TX > 32 || TY > 32 || TZ > 32)
    return fail(SNAP_ERR_INVALID, "tile edges must be 1..32");
  if (waves < 0 || waves > 16) return fail(SNAP_ERR_INVALID, "waves per tile must be 1..16");
  if (apl == 0) apl = pick_apl(nang);
  if (apl < 1 || apl > 2 || (nang + apl - 1) / apl > 64)
    return fail(SNAP_ERR_INVALID, "%d angles per lane cannot hold %d angles", apl, nang);
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
  // default: 8 waves (two workgroups per CU at 128 registers per lane); 6
  // with two angles per lane, where ~150 registers leave 12 waves per CU
  if (waves == 0) waves = (apl == 1) ? 8 : 6;
  p->waves = waves;
  p->td = td;
  p->apl = apl;
  p->lpc = (nang + apl - 1) / apl;
  p->cpw = pick_cpw(p->lpc);
  // cells per moment batch: enough (cell, moment) sums to fill a wavefront
  {
    int per = 64 / (((cmom + 1) / 2) * p->cpw);  // lanes sum moment pairs
    if (per < 1) per = 1;
    p->sb = p->cpw * std::min(per, std::max(1, 8 / p->cpw));
  }
  // tile diagonals in sweep order
  p->nplanes = p->nTx + p->nTy + p->nTz - 2;
  tiles.clear();
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
  cells.clear();
  coff.assign(p->ncellplanes + 1, 0);
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
  // every wavefront owns at most max_steps cell groups of a diagonal
  const int ms = max_steps(apl);
  const int need_waves = (p->maxdiag + ms * p->cpw - 1) / (ms * p->cpw);
  if (p->waves < need_waves) p->waves = need_waves;
  const int mt = max_threads(apl, td);
  if (p->waves * 64 > mt) {
    if (waves * 64 <= mt && need_waves * 64 <= mt)
      p->waves = mt / 64;
    else
      return fail(SNAP_ERR_INVALID, "tile %dx%dx%d needs %d waves per workgroup (> %d)", TX, TY,
                  TZ, std::max(waves, need_waves), mt / 64);
  }
  if (p->lds_bytes(8, false) > 160 * 1024)
    return fail(SNAP_ERR_INVALID, "tile %dx%dx%d with %d angles needs %zu B of LDS (> 160 KiB)",
                TX, TY, TZ, nang, p->lds_bytes(8, false));
  return SNAP_OK;
}

int plan_create(SweepPlan *p, Extent3 e, int nang, int cmom, int TX, int TY, int TZ, int waves,
                int apl, bool td) {
  plan_free(p);
  std::vector<int> tiles, coff;
  std::vector<int16_t> cells;
  SNAP_TRY(plan_layout(p, e, nang, cmom, TX, TY, TZ, waves, apl, td, tiles, coff, cells));
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

// Most angles the one-cell workgroup holds: 1024 lanes with two angles each.
constexpr int kWideMaxAngles = 2048;

int plan_layout_wide(SweepPlan *p, Extent3 e, int nang, int cmom, bool td,
                     std::vector<int> &cells) {
  *p = SweepPlan();
  if (e.nx < 1 || e.ny < 1 || e.nz < 1) return fail(SNAP_ERR_INVALID, "empty chunk");
  if (nang < 1 || nang > kWideMaxAngles)
    return fail(SNAP_ERR_UNSUPPORTED, "the sweep handles 1..%d angles per octant, got %d",
                kWideMaxAngles, nang);
  if (cmom < 1 || cmom > 16) return fail(SNAP_ERR_INVALID, "cmom must be 1..16, got %d", cmom);
  if (e.nx > 1023 || e.ny > 1023 || e.nz > 1023)
    return fail(SNAP_ERR_INVALID, "chunk edges above 1023 cells");
  p->ext = e;
  p->TX = p->TY = p->TZ = 1;
  p->nTx = e.nx;
  p->nTy = e.ny;
  p->nTz = e.nz;
  p->ntiles = e.nx * e.ny * e.nz;
  p->nang = nang;
  p->cmom = cmom;
  p->wide = true;
  p->td = td;
  // one angle per lane up to 1024 angles, two beyond; whole wavefronts
  p->apl = nang > 1024 ? 2 : 1;
  const int lanes = (nang + p->apl - 1) / p->apl;
  p->waves = (lanes + 63) / 64;
  p->lpc = lanes;
  p->cpw = 0;
  p->sb = 0;
  p->maxdiag = 0;
  p->ncellplanes = 1;
  p->nplanes = e.nx + e.ny + e.nz - 2;
  cells.clear();
  p->plane_off.assign(p->nplanes + 1, 0);
  for (int D = 0; D < p->nplanes; D++) {
    p->plane_off[D] = (int)cells.size();
    for (int k = 0; k < e.nz; k++)
      for (int j = 0; j < e.ny; j++) {
        const int i = D - j - k;
        if (i >= 0 && i < e.nx) cells.push_back(i | (j << 10) | (k << 20));
      }
    p->maxdiag = std::max(p->maxdiag, (int)cells.size() - p->plane_off[D]);
  }
  p->plane_off[p->nplanes] = (int)cells.size();
  return SNAP_OK;
}

int plan_create_wide(SweepPlan *p, Extent3 e, int nang, int cmom, bool td) {
  plan_free(p);
  std::vector<int> cells;
  SNAP_TRY(plan_layout_wide(p, e, nang, cmom, td, cells));
  SNAP_HIP_TRY(hipMalloc((void **)&p->plane_tiles, sizeof(int) * cells.size()));
  SNAP_HIP_TRY(hipMemcpy(p->plane_tiles, cells.data(), sizeof(int) * cells.size(),
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
  int nang, cmom, ncellplanes, ncells_tile, sb, lpc;
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
  // batch of problems (BATCH kernels): per-launch (problem, tile) entries,
  // per-problem chunks, and the predicates that skip finished problems
  const int2 *batch_tiles;
  const BatchDesc *descs;
  const int *active, *done;
  const int *prec;  // with prec, only problems whose prec[p] == prec_want
  int prec_want;
  int64_t total_cells;
  // time-dependent kernels (TD): vdelt per job group, or vdelt0 for the
  // general variant; time-edge fluxes and MMS source [8][ng_job][cells][nang]
  const double *vdelt;
  double vdelt0;
  const double *tin;
  double *tout;
  const double *qim;
  const int32_t *skip;  // non-NULL and set: every workgroup returns at once
};

// Arguments of the general variant (the drop-in entry point): Realm affine
// accessors, ghost faces, and flux accumulation by device atomics.
struct SweepGeneral {
  Acc3 qtot, dinv, txs, flux, fluxm;
  Acc2 gx, gy, gz;
  Acc3 tin, tout, qim;  // TD: nang doubles per cell (base NULL = none)
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

// Number of set bits of the ballots b[] inside the lane's cell segment,
// summed: the cell-global count of negative outflows.  With one cell per
// wavefront only that cell's lanes can vote; otherwise each segment is
// counted with scalar operations and the lane picks its own.
template <int CPW, int NB>
__device__ inline unsigned seg_count(const unsigned long long (&b)[NB], unsigned long long m0,
                                     int LPC, int cl) {
  if constexpr (CPW == 1) {
    unsigned n = 0;
#pragma unroll
    for (int k = 0; k < NB; k++) n += (unsigned)__popcll(b[k]);
    return n;
  } else {
    unsigned res = 0;
#pragma unroll
    for (int cs = 0; cs < CPW; cs++) {
      const unsigned long long m = m0 << (cs * LPC);
      unsigned n = 0;
#pragma unroll
      for (int k = 0; k < NB; k++) n += (unsigned)__popcll(b[k] & m);
      res = (cl == cs) ? n : res;
    }
    return res;
  }
}

// Tile sweep.  Real: arithmetic and face type.  CPW: cells per wavefront.
// APL: angles per lane.  EXACT: the reference's operation order with separate
// roundings (bit-equal to the CPU oracle); otherwise fused multiply-adds.
// REC: form dinv from t_xs instead of reading it.  GEN: general accessors,
// ghost faces and atomic accumulation (the drop-in entry point).  BATCH:
// each workgroup's (problem, tile) comes from the launch's table, arrays are
// the batch's concatenation ([g][all cells]), and a workgroup whose problem
// is finished or has converged its inner iteration exits at once.  TD: the
// time-dependent terms (sweep.cc:414-534; gpu_sweep.cu:738-1113) -- the
// time-edge inflow vdelt*psi_t,in, a fourth outflow edge in the fixup and
// psi_t,out -- and the MMS angular source (sweep.cc:347-351), each switched
// on at run time by vdelt != 0 and a non-NULL source.
template <typename Real, int CPW, int APL, bool EXACT, bool REC, bool GEN, bool BATCH = false,
          bool TD = false>
__global__ void __launch_bounds__(max_threads(APL, TD)) k_sweep_tiles(SweepArgs A, SweepGeneral G, int plane_off) {
  constexpr bool kFP64 = std::is_same<Real, double>::value;
  if (A.skip && *A.skip) return;  // the whole workgroup: an inner iteration enqueued ahead
  constexpr int MS = max_steps(APL);  // cell groups per wave and diagonal
  constexpr int QS = APL;             // source moments held per lane
  constexpr int ECR = ec_row(APL);    // row stride of the ec table
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int TX = A.TX, TY = A.TY, TZ = A.TZ, nang = A.nang, cmom = A.cmom, SB = A.sb;
  const int LPC = A.lpc;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const int ncp = A.ncellplanes;
  const SweepLds lay =
      sweep_lds(TX, TY, TZ, nang, cmom, W, SB, A.ncells_tile, ncp, CPW, APL, (int)sizeof(Real));
  const int fxn = TY * TZ * nang, fyn = TX * TZ * nang, fzn = TX * TY * nang;
  Real *fx = (Real *)smem_raw;
  Real *fy = fx + fxn;
  Real *fz = fy + fyn;
  Real *ecs = (Real *)(smem_raw + lay.ecs);       // [cmom][ECR]: ec[l][a] at l*ECR + a
  double *ect = (double *)(smem_raw + lay.ect);   // [nang][kEctRow]: ec[l][a] at a*kEctRow + l
  double *pwb = (double *)(smem_raw + lay.pwb) + wave * SB * nang;  // this wave's w*pc batch
  int *cix = (int *)(smem_raw + lay.cix) + wave * SB;               // batch cell indices
  Real *qsl = (Real *)(smem_raw + lay.qsl) + wave * MS * CPW * cmom;  // source moments
  int4 *tcr = (int4 *)(smem_raw + lay.cells);  // per tile cell: chunk cell (or -1), face offsets
  int *toff = (int *)(smem_raw + lay.coff);    // offsets of the cell diagonals in tcr

  const int s = blockIdx.y;
  const int corner = A.corner0 + s / A.ng_job;
  const int g = s % A.ng_job;
  // the chunk: the launch's own, or one problem of a batch
  Extent3 ext = A.ext;
  int nTx = A.nTx, nTy = A.nTy, nTz = A.nTz, ntiles = A.ntiles;
  double dhi = A.hi, dhj = A.hj, dhk = A.hk;
  size_t gstride, coff = 0, face_off = 0;
  int tp;
  if constexpr (BATCH) {
    const int2 bt = A.batch_tiles[plane_off + blockIdx.x];
    if (!A.active[bt.x] || A.done[bt.x]) return;  // the whole workgroup
    if (A.prec && A.prec[bt.x] != A.prec_want) return;
    const BatchDesc &P = A.descs[bt.x];
    ext = P.ext;
    nTx = P.nTx;
    nTy = P.nTy;
    nTz = P.nTz;
    ntiles = P.ntiles;
    dhi = P.hi;
    dhj = P.hj;
    dhk = P.hk;
    gstride = (size_t)A.total_cells;
    coff = (size_t)P.cell_off;
    face_off = (size_t)P.face_off;
    tp = bt.y;
  } else {
    gstride = (size_t)ext.nx * ext.ny * ext.nz;
    tp = A.plane_tiles[plane_off + blockIdx.x];
  }
  const int tx = tp & 1023, ty = (tp >> 10) & 1023, tz = (tp >> 20) & 1023;
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int nx = ext.nx, ny = ext.ny, nz = ext.nz;
  const int ex = min(TX, nx - x0), ey = min(TY, ny - y0), ez = min(TZ, nz - z0);
  const bool sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
  const int tile_id = (tz * nTy + ty) * nTx + tx;
  const int per_tile = fxn + fyn + fzn;
  Real *faces_s = (Real *)A.faces + face_off + (size_t)s * ntiles * per_tile;
  // dense per-sweep bases
  const double *qg = A.q + ((size_t)g * gstride + coff) * cmom;
  const Real *dg = (const Real *)A.dinv + ((size_t)g * gstride + coff) * nang;
  const Real *tg = (const Real *)A.txs + (size_t)g * gstride + coff;
  double *pfg = A.part_flux + ((size_t)corner * A.ng_job + g) * gstride + coff;
  double *pmg =
      A.part_fluxm + (((size_t)corner * A.ng_job + g) * gstride + coff) * (size_t)(cmom - 1);
  // time-dependent terms of this sweep (uniform over the workgroup)
  double vd = 0.0;
  bool has_time = false, has_qim = false;
  const double *tig = nullptr, *qmg = nullptr;
  double *tog = nullptr;
  if constexpr (TD) {
    vd = GEN ? A.vdelt0 : A.vdelt[g];
    has_time = vd != 0.0;
    if constexpr (GEN) {
      has_qim = G.qim.base != nullptr;
    } else {
      const size_t tb = (((size_t)corner * A.ng_job + g) * gstride + coff) * (size_t)nang;
      has_qim = A.qim != nullptr;
      if (has_time) {
        tig = A.tin + tb;
        tog = A.tout + tb;
      }
      if (has_qim) qmg = A.qim + tb;
    }
  }

  // ---- per-lane constants: lane = cell slot * LPC + la, angles la + r*LPC
  const int cl = (CPW == 1) ? 0 : lane / LPC;
  const int la = lane - cl * LPC;
  const bool lane_ok = cl < CPW;
  const unsigned long long seg0 = (LPC >= 64) ? ~0ull : ((1ull << LPC) - 1ull);
  int an[APL];     // angle of slot r (clamped into range)
  bool aok[APL];   // slot r holds a real angle of a real cell slot
#pragma unroll
  for (int r = 0; r < APL; r++) {
    const int a = la + r * LPC;
    aok[r] = lane_ok && a < nang;
    an[r] = aok[r] ? a : 0;
  }
  const int per_step = W * CPW;

  // ---- tables and inflow faces (from the upstream tile, the ghost array,
  // or vacuum) into LDS
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
  for (int idx = threadIdx.x; idx < cmom * nang; idx += blockDim.x) {
    const double v = A.ec[(size_t)corner * cmom * nang + idx];
    const int l = idx / nang, b = idx - l * nang;
    ecs[l * ECR + b] = (Real)v;
    ect[b * kEctRow + l] = v;
  }
  // Inflow slots of cells outside a ragged chunk edge are copied like the
  // others; no cell reads them.
  {
    const Real *up = faces_s + (size_t)(tile_id - 1) * per_tile;
    if (tx > 0) {
      for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) fx[idx] = up[idx];
    } else if (GEN && G.gx.base) {
      for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) {
        const int a2 = idx % nang, slot = idx / nang, j = slot % TY, k = slot / TY;
        Real v = Real(0);
        if (j < ey && k < ez) {
          const int py = syp ? y0 + j : ny - 1 - (y0 + j), pz = szp ? z0 + k : nz - 1 - (z0 + k);
          v = (Real)((const double *)G.gx.at(py, pz, g))[a2];
        }
        fx[idx] = v;
      }
    } else {
      for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) fx[idx] = Real(0);
    }
  }
  {
    const Real *up = faces_s + (size_t)(tile_id - nTx) * per_tile + fxn;
    if (ty > 0) {
      for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) fy[idx] = up[idx];
    } else if (GEN && G.gy.base) {
      for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) {
        const int a2 = idx % nang, slot = idx / nang, i = slot % TX, k = slot / TX;
        Real v = Real(0);
        if (i < ex && k < ez) {
          const int px = sxp ? x0 + i : nx - 1 - (x0 + i), pz = szp ? z0 + k : nz - 1 - (z0 + k);
          v = (Real)((const double *)G.gy.at(px, pz, g))[a2];
        }
        fy[idx] = v;
      }
    } else {
      for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) fy[idx] = Real(0);
    }
  }
  {
    const Real *up = faces_s + (size_t)(tile_id - nTx * nTy) * per_tile + fxn + fyn;
    if (tz > 0) {
      for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) fz[idx] = up[idx];
    } else if (GEN && G.gz.base) {
      for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) {
        const int a2 = idx % nang, slot = idx / nang, i = slot % TX, j = slot / TX;
        Real v = Real(0);
        if (i < ex && j < ey) {
          const int px = sxp ? x0 + i : nx - 1 - (x0 + i), py = syp ? y0 + j : ny - 1 - (y0 + j);
          v = (Real)((const double *)G.gz.at(px, py, g))[a2];
        }
        fz[idx] = v;
      }
    } else {
      for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) fz[idx] = Real(0);
    }
  }

  // per-angle constants.  EXACT keeps (psi*mu)*hi as the reference writes
  // it (cx,cy,cz = mu,eta,xi); the others fold them (cx = mu*hi, ...).
  // hi*mu is the same double either way, so the recomputed dinv is formed
  // from these too.
  Real cx[APL], cy[APL], cz[APL];
  double w_d[APL];
  float g_sum32[APL];
#pragma unroll
  for (int r = 0; r < APL; r++) {
    const double mu_d = A.mu[an[r]], eta_d = A.eta[an[r]], xi_d = A.xi[an[r]];
    w_d[r] = A.w[an[r]];
    if constexpr (EXACT) {
      cx[r] = (Real)mu_d;
      cy[r] = (Real)eta_d;
      cz[r] = (Real)xi_d;
    } else {
      cx[r] = (Real)(mu_d * dhi);
      cy[r] = (Real)(eta_d * dhj);
      cz[r] = (Real)(xi_d * dhk);
    }
    g_sum32[r] = (float)((dhi * mu_d + dhj * eta_d) + dhk * xi_d);
  }
  const Real hi = (Real)dhi, hj = (Real)dhj, hk = (Real)dhk;
  const Real tolr = (Real)1.0e-12;
  const int nm1 = cmom - 1;
  // source moments travel in lanes (lane la of a cell's segment holds moments
  // la + m*LPC) when they fit; otherwise the wave copies them to LDS itself
  const bool q_in_lanes = cmom <= QS * LPC;
  // moment sums: lane -> (batched cell fk, moment fl) for the first 64 sums
  const int pairs = (cmom + 1) / 2;  // moment pairs summed per lane in a flush
  __syncthreads();

  // ---- register prefetch of this wave's cells on the next diagonal: dinv of
  // the lane's angles, its source moments and t_xs (TD: and the lane's
  // time-edge inflows and MS sources).
  // (dinv and t_xs in the sweep's type; moments in FP64 until they are staged)
  Real n_dv[MS][APL], n_tx[MS];
  double n_q[MS][QS], n_ti[MS][APL], n_qm[MS][APL];
  int n_cell[MS];
  auto load_cell = [&](int cell, int st) {
    if constexpr (TD) {
      const double *tp = nullptr, *qp = nullptr;
      if constexpr (GEN) {
        const int px = cell % nx, py = (cell / nx) % ny, pz = cell / (nx * ny);
        if (has_time) tp = (const double *)G.tin.at(px, py, pz, g);
        if (has_qim) qp = (const double *)G.qim.at(px, py, pz, g);
      } else {
        if (has_time) tp = tig + (size_t)cell * nang;
        if (has_qim) qp = qmg + (size_t)cell * nang;
      }
#pragma unroll
      for (int r = 0; r < APL; r++) {
        if (has_time) n_ti[st][r] = tp[an[r]];
        if (has_qim) n_qm[st][r] = qp[an[r]];
      }
    }
    if constexpr (GEN) {
      const int px = cell % nx, py = (cell / nx) % ny, pz = cell / (nx * ny);
      const Real *dv = (const Real *)G.dinv.at(px, py, pz, g);
      const double *qc = (const double *)G.qtot.at(px, py, pz, g);
#pragma unroll
      for (int r = 0; r < APL; r++) n_dv[st][r] = dv[an[r]];
#pragma unroll
      for (int m = 0; m < QS; m++)
        if (q_in_lanes && la + m * LPC < cmom) n_q[st][m] = qc[la + m * LPC];
      n_tx[st] = *(const Real *)G.txs.at(px, py, pz, g);
    } else {
      if constexpr (!REC) {
        const Real *dv = dg + (size_t)cell * nang;
#pragma unroll
        for (int r = 0; r < APL; r++) n_dv[st][r] = dv[an[r]];
      }
      const double *qc = qg + (size_t)cell * cmom;
#pragma unroll
      for (int m = 0; m < QS; m++)
        if (q_in_lanes && la + m * LPC < cmom) n_q[st][m] = qc[la + m * LPC];
      n_tx[st] = tg[cell];
    }
  };
  auto prefetch = [&](int dd) {
    const int c0 = toff[dd], c1 = toff[dd + 1];
#pragma unroll
    for (int st = 0; st < MS; st++) {
      const int my = c0 + st * per_step + wave * CPW + cl;
#pragma unroll
      for (int r = 0; r < APL; r++) n_dv[st][r] = Real(0);
#pragma unroll
      for (int m = 0; m < QS; m++) n_q[st][m] = 0.0;
      if constexpr (TD) {
#pragma unroll
        for (int r = 0; r < APL; r++) n_ti[st][r] = n_qm[st][r] = 0.0;
      }
      n_tx[st] = Real(0);
      n_cell[st] = -1;
      if (lane_ok && my < c1) {
        const int cell = tcr[my].x;
        n_cell[st] = cell;
        if (cell >= 0) load_cell(cell, st);
      }
    }
  };

  int nbuf = 0;  // cells in this wave's moment batch (wave-uniform)
  auto flush = [&]() {
    // per-cell flux moments over angles in ascending angle order
    // (sweep.cc:568-591): phi_l = sum_a ec[l][a] * (w[a]*pc[a]).  Each lane
    // sums two adjacent moments of one cell, reading both coefficients of an
    // angle with one 16-byte load from the angle-major table.
    wave_sync();
    for (int o = lane; o < nbuf * pairs; o += 64) {
      const int kk = o / pairs, l0 = 2 * (o - kk * pairs);
      const int cell = cix[kk];
      if (cell < 0) continue;
      const double *v = pwb + kk * nang;
      const double *e = ect + l0;  // 16-byte aligned: l0 is even, rows are 128 B
      double s0 = 0.0, s1 = 0.0;
#pragma unroll 4
      for (int b = 0; b < nang; b++) {
        const double vb = v[b];
        const double2 ee = *(const double2 *)(e + b * kEctRow);
        s0 = madd<EXACT>(ee.x, vb, s0);
        s1 = madd<EXACT>(ee.y, vb, s1);
      }
      auto store = [&](int l, double sum) {
        if constexpr (GEN) {
          const int px = cell % nx, py = (cell / nx) % ny, pz = cell / (nx * ny);
          double *dst = (l == 0) ? (double *)G.flux.at(px, py, pz, g)
                                 : (double *)G.fluxm.at(px, py, pz, g) + (l - 1);
          atomicAdd(dst, sum);
        } else {
          if (l == 0)
            pfg[cell] = sum;
          else
            pmg[(size_t)cell * nm1 + (l - 1)] = sum;
        }
      };
      store(l0, s0);
      if (l0 + 1 < cmom) store(l0 + 1, s1);
    }
    wave_sync();
    nbuf = 0;
  };

  prefetch(0);
  for (int d = 0; d < ncp; ++d) {
    const int c0 = toff[d], c1 = toff[d + 1];
    const int nsteps = (c1 - c0 + per_step - 1) / per_step;
    Real c_dv[MS][APL], c_tx[MS];
    double c_ti[MS][APL], c_qm[MS][APL];
    int c_cell[MS];
    // source moments of this diagonal's cells to the wave's LDS slots
#pragma unroll
    for (int st = 0; st < MS; st++) {
#pragma unroll
      for (int r = 0; r < APL; r++) c_dv[st][r] = n_dv[st][r];
      if constexpr (TD) {
#pragma unroll
        for (int r = 0; r < APL; r++) {
          c_ti[st][r] = n_ti[st][r];
          c_qm[st][r] = n_qm[st][r];
        }
      }
      c_tx[st] = n_tx[st];
      c_cell[st] = n_cell[st];
#pragma unroll
      for (int m = 0; m < QS; m++)
        if (q_in_lanes && lane_ok && la + m * LPC < cmom)
          qsl[(st * CPW + cl) * cmom + la + m * LPC] = (Real)n_q[st][m];
    }
    if (!q_in_lanes) {
      // moments do not fit the cell's lanes: the wave copies its cells'
      // records to its LDS slots directly (not prefetched)
      wave_sync();
      for (int o = lane; o < MS * CPW * cmom; o += 64) {
        const int sc = o / cmom, l = o - sc * cmom;  // sc = st * CPW + slot
        const int st = sc / CPW, slot = sc - st * CPW;
        const int my = c0 + st * per_step + wave * CPW + slot;
        double v = 0.0;
        if (my < c1) {
          const int cell = tcr[my].x;
          if (cell >= 0) {
            if constexpr (GEN) {
              const int px = cell % nx, py = (cell / nx) % ny, pz = cell / (nx * ny);
              v = ((const double *)G.qtot.at(px, py, pz, g))[l];
            } else {
              v = qg[(size_t)cell * cmom + l];
            }
          }
        }
        qsl[sc * cmom + l] = (Real)v;
      }
    }
    if (d + 1 < ncp) prefetch(d + 1);
    wave_sync();
#pragma unroll
    for (int st = 0; st < MS; ++st) {
      if (st >= nsteps) break;
      const int my = c0 + st * per_step + wave * CPW + cl;
      const int cell = c_cell[st];
      double pcw[APL];
#pragma unroll
      for (int r = 0; r < APL; r++) pcw[r] = 0.0;
      if (cell >= 0) {
        const int4 rec = tcr[my];
        // angular source (sweep.cc:331-345), moments from the LDS slots
        const Real *q = qsl + (st * CPW + cl) * cmom;
        Real psi[APL];
        if constexpr (kFP64) {
          double ps[APL];
#pragma unroll
          for (int r = 0; r < APL; r++) ps[r] = q[0];
          if constexpr (EXACT) {
            for (int l = 1; l < cmom; l++) {
              const double ql = q[l];
#pragma unroll
              for (int r = 0; r < APL; r++) ps[r] = madd<EXACT>(ecs[l * ECR + an[r]], ql, ps[r]);
            }
          } else {
            // two independent chains halve the dependent-FMA latency
            double p1[APL];
#pragma unroll
            for (int r = 0; r < APL; r++) p1[r] = 0.0;
            int l = 1;
            for (; l + 1 < cmom; l += 2) {
              const double q0 = q[l], q1 = q[l + 1];
#pragma unroll
              for (int r = 0; r < APL; r++) {
                ps[r] = __builtin_fma(ecs[l * ECR + an[r]], q0, ps[r]);
                p1[r] = __builtin_fma(ecs[(l + 1) * ECR + an[r]], q1, p1[r]);
              }
            }
            if (l < cmom) {
              const double q0 = q[l];
#pragma unroll
              for (int r = 0; r < APL; r++) ps[r] = __builtin_fma(ecs[l * ECR + an[r]], q0, ps[r]);
            }
#pragma unroll
            for (int r = 0; r < APL; r++) ps[r] += p1[r];
          }
          if constexpr (TD) {
            if (has_qim) {
#pragma unroll
              for (int r = 0; r < APL; r++) ps[r] += c_qm[st][r];
            }
          }
#pragma unroll
          for (int r = 0; r < APL; r++) psi[r] = ps[r];
        } else {
          float pf[APL], p1[APL];
#pragma unroll
          for (int r = 0; r < APL; r++) {
            pf[r] = (float)q[0];
            p1[r] = 0.0f;
          }
          int l = 1;
          for (; l + 1 < cmom; l += 2) {
            const float q0 = (float)q[l], q1 = (float)q[l + 1];
#pragma unroll
            for (int r = 0; r < APL; r++) {
              pf[r] = maddf((float)ecs[l * ECR + an[r]], q0, pf[r]);
              p1[r] = maddf((float)ecs[(l + 1) * ECR + an[r]], q1, p1[r]);
            }
          }
          if (l < cmom) {
            const float q0 = (float)q[l];
#pragma unroll
            for (int r = 0; r < APL; r++) pf[r] = maddf((float)ecs[l * ECR + an[r]], q0, pf[r]);
          }
#pragma unroll
          for (int r = 0; r < APL; r++) pf[r] += p1[r];
          if constexpr (TD) {
            if (has_qim) {
#pragma unroll
              for (int r = 0; r < APL; r++) pf[r] += (float)c_qm[st][r];
            }
          }
#pragma unroll
          for (int r = 0; r < APL; r++) psi[r] = pf[r];
        }
        const Real txs = c_tx[st];
        Real dv[APL];
#pragma unroll
        for (int r = 0; r < APL; r++) {
          // dinv as CalculateGeometryParam forms it (expxs.cc:378-381)
          if constexpr (REC && kFP64 && EXACT)
            dv[r] = 1.0 / ((((double)txs + hi * cx[r]) + hj * cy[r]) + hk * cz[r]);
          else if constexpr (REC && kFP64)
            dv[r] = 1.0 / ((((double)txs + cx[r]) + cy[r]) + cz[r]);
          else if constexpr (REC)
            dv[r] = 1.0f / ((float)txs + g_sum32[r]);
          else
            dv[r] = c_dv[st][r];
        }
        // inflow faces; slots past the last angle read angle 0's entry and
        // are never stored
        Real *sx = fx + rec.y;
        Real *sy = fy + rec.z;
        Real *sz = fz + rec.w;
        Real psii[APL], psij[APL], psik[APL], pc[APL], ti[APL];
        const Real vdr = (Real)vd;
#pragma unroll
        for (int r = 0; r < APL; r++) {
          psii[r] = sx[an[r]];
          psij[r] = sy[an[r]];
          psik[r] = sz[an[r]];
          ti[r] = TD ? (Real)c_ti[st][r] : Real(0);
          // initial diamond-difference solution (sweep.cc:352-400)
          Real p = psi[r];
          if constexpr (kFP64 && EXACT) {
            p += psii[r] * cx[r] * hi;
            p += psij[r] * cy[r] * hj;
            p += psik[r] * cz[r] * hk;
          } else if constexpr (kFP64) {
            p = __builtin_fma(psii[r], cx[r], p);
            p = __builtin_fma(psij[r], cy[r], p);
            p = __builtin_fma(psik[r], cz[r], p);
          } else {
            p = maddf(psii[r], cx[r], p);
            p = maddf(psij[r], cy[r], p);
            p = maddf(psik[r], cz[r], p);
          }
          // time-edge inflow (sweep.cc:414-419)
          if constexpr (TD) {
            if (has_time) {
              if constexpr (kFP64 && EXACT)
                p += vdr * ti[r];
              else if constexpr (kFP64)
                p = __builtin_fma(vdr, ti[r], p);
              else
                p = maddf(vdr, ti[r], p);
            }
          }
          pc[r] = p * dv[r];
        }
        // outflow factors: 1 while the face keeps its diamond-difference
        // value, 0 once the fixup has set it to zero
        bool hx[APL], hy[APL], hz[APL], ht[APL];
#pragma unroll
        for (int r = 0; r < APL; r++) hx[r] = hy[r] = hz[r] = ht[r] = true;
        if (A.fixup) {
          // negative-flux fixup (sweep.cc:402-505; gpu_sweep.cu:1556-1630).
          // The count of negative outflows is cell-global over angles.  A
          // pass that clears no new flag leaves the next pc equal to this
          // one, so the pass after it repeats the count and ends the loop:
          // with 3*nang flags (4*nang with the time edge), 3*nang+2 passes
          // always reach the break, and the outflows are 2*pc - psi_in of
          // the final pc.
          constexpr int NE = TD ? 4 : 3;
          auto count_negative = [&]() -> unsigned {
            unsigned long long bal[NE * APL];
#pragma unroll
            for (int r = 0; r < APL; r++) {
              const Real p2 = Real(2) * pc[r];
              const bool nxn = aok[r] && p2 - psii[r] < Real(0);
              const bool nyn = aok[r] && p2 - psij[r] < Real(0);
              const bool nzn = aok[r] && p2 - psik[r] < Real(0);
              hx[r] = hx[r] && !nxn;
              hy[r] = hy[r] && !nyn;
              hz[r] = hz[r] && !nzn;
              bal[NE * r] = __builtin_amdgcn_ballot_w64(nxn);
              bal[NE * r + 1] = __builtin_amdgcn_ballot_w64(nyn);
              bal[NE * r + 2] = __builtin_amdgcn_ballot_w64(nzn);
              if constexpr (TD) {
                const bool ntn = has_time && aok[r] && p2 - ti[r] < Real(0);
                ht[r] = ht[r] && !ntn;
                bal[NE * r + 3] = __builtin_amdgcn_ballot_w64(ntn);
              }
            }
            return seg_count<CPW>(bal, seg0, LPC, cl);
          };
          unsigned old_neg = count_negative();
          if (old_neg != 0) {
            for (int it = 1; it < NE * nang + 2; ++it) {
#pragma unroll
              for (int r = 0; r < APL; r++) {
                const Real vx = hx[r] ? Real(1) : Real(0), vy = hy[r] ? Real(1) : Real(0),
                           vz = hz[r] ? Real(1) : Real(0), vt = ht[r] ? Real(1) : Real(0);
                Real den, p;
                if constexpr (kFP64 && EXACT) {
                  if (TD && has_time) {
                    // sweep.cc:481-494
                    p = psi[r] + 0.5 * (psii[r] * cx[r] * hi * (1.0 + vx) +
                                        psij[r] * cy[r] * hj * (1.0 + vy) +
                                        psik[r] * cz[r] * hk * (1.0 + vz) +
                                        ti[r] * vdr * (1.0 + vt));
                    den = (p <= 0.0) ? 0.0
                                     : (txs + cx[r] * hi * vx + cy[r] * hj * vy + cz[r] * hk * vz +
                                        vdr * vt);
                  } else {
                    p = psi[r] + 0.5 * (psii[r] * cx[r] * hi * (1.0 + vx) +
                                        psij[r] * cy[r] * hj * (1.0 + vy) +
                                        psik[r] * cz[r] * hk * (1.0 + vz));
                    den = (p <= 0.0) ? 0.0
                                     : (txs + cx[r] * hi * vx + cy[r] * hj * vy + cz[r] * hk * vz);
                  }
                } else {
                  Real t = psii[r] * cx[r] * (Real(1) + vx) +
                           psij[r] * cy[r] * (Real(1) + vy) + psik[r] * cz[r] * (Real(1) + vz);
                  Real dd = txs + cx[r] * vx + cy[r] * vy + cz[r] * vz;
                  if constexpr (TD) {
                    if (has_time) {
                      t += ti[r] * vdr * (Real(1) + vt);
                      dd += vdr * vt;
                    }
                  }
                  p = psi[r] + Real(0.5) * t;
                  den = (p <= Real(0)) ? Real(0) : dd;
                }
                pc[r] = (den < tolr) ? Real(0) : p / den;
              }
              const unsigned neg = count_negative();
              if (neg == old_neg) break;
              old_neg = neg;
            }
          }
        }
#pragma unroll
        for (int r = 0; r < APL; r++) {
          const Real p2 = Real(2) * pc[r];
          const Real ox = (p2 - psii[r]) * (hx[r] ? Real(1) : Real(0));
          const Real oy = (p2 - psij[r]) * (hy[r] ? Real(1) : Real(0));
          const Real oz = (p2 - psik[r]) * (hz[r] ? Real(1) : Real(0));
          if (aok[r]) {
            sx[an[r]] = ox;
            sy[an[r]] = oy;
            sz[an[r]] = oz;
          }
          pcw[r] = w_d[r] * (double)pc[r];
        }
        // time-edge outflow (sweep.cc:512-519, 529-533)
        if constexpr (TD) {
          if (has_time) {
            double *to;
            if constexpr (GEN) {
              const int px = cell % nx, py = (cell / nx) % ny, pz = cell / (nx * ny);
              to = (double *)G.tout.at(px, py, pz, g);
            } else {
              to = tog + (size_t)cell * nang;
            }
#pragma unroll
            for (int r = 0; r < APL; r++)
              if (aok[r]) to[an[r]] = (double)((Real(2) * pc[r] - ti[r]) * (ht[r] ? Real(1) : Real(0)));
          }
        }
      }
      // ---- batch this step's cells for the moment sums
      if (lane_ok) {
#pragma unroll
        for (int r = 0; r < APL; r++)
          if (aok[r]) pwb[(nbuf + cl) * nang + an[r]] = pcw[r];
        if (la == 0) cix[nbuf + cl] = cell;
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
  for (int idx = threadIdx.x; idx < per_tile; idx += blockDim.x) out[idx] = fx[idx];
  if constexpr (GEN) {
    if (tx == nTx - 1 && G.gx.base)
      for (int idx = threadIdx.x; idx < fxn; idx += blockDim.x) {
        const int a2 = idx % nang, slot = idx / nang, j = slot % TY, k = slot / TY;
        if (j < ey && k < ez) {
          const int py = syp ? y0 + j : ny - 1 - (y0 + j), pz = szp ? z0 + k : nz - 1 - (z0 + k);
          ((double *)G.gx.at(py, pz, g))[a2] = (double)fx[idx];
        }
      }
    if (ty == nTy - 1 && G.gy.base)
      for (int idx = threadIdx.x; idx < fyn; idx += blockDim.x) {
        const int a2 = idx % nang, slot = idx / nang, i = slot % TX, k = slot / TX;
        if (i < ex && k < ez) {
          const int px = sxp ? x0 + i : nx - 1 - (x0 + i), pz = szp ? z0 + k : nz - 1 - (z0 + k);
          ((double *)G.gy.at(px, pz, g))[a2] = (double)fy[idx];
        }
      }
    if (tz == nTz - 1 && G.gz.base)
      for (int idx = threadIdx.x; idx < fzn; idx += blockDim.x) {
        const int a2 = idx % nang, slot = idx / nang, i = slot % TX, j = slot / TX;
        if (i < ex && j < ey) {
          const int px = sxp ? x0 + i : nx - 1 - (x0 + i), py = syp ? y0 + j : ny - 1 - (y0 + j);
          ((double *)G.gz.at(px, py, g))[a2] = (double)fz[idx];
        }
      }
  }
}

// ---------------------------------------------------------------------------
// one cell per workgroup (any angle count)
// ---------------------------------------------------------------------------
// Each workgroup sweeps one cell of one (corner, group) sweep; launch D covers
// the cells on chunk diagonal i+j+k = D of every sweep of the job.  Thread t
// holds angles t + r*blockDim (r < APL), so per-angle arrays (dinv, time-edge
// fluxes, MMS source, faces) are read and written in contiguous runs.  The
// inflow faces come from per-sweep face planes updated in place, as the
// reference's GPU sweep keeps its x/y/z flux planes (gpu_sweep.cu:1470-1555):
// the cell at (i,j,k) reads the x plane at (j,k) written by (i-1,j,k) on the
// previous diagonal and overwrites it with its outflow.  The fixup count is
// cell-global: wave ballots, then a sum over the workgroup's waves in LDS
// (gpu_sweep.cu:1596-1614 reduces over the CTA the same way).
template <typename Real, int APL, bool EXACT, bool GEN, bool TD>
__global__ void __launch_bounds__(1024) k_sweep_wide(SweepArgs A, SweepGeneral G, int plane_off) {
  constexpr bool kFP64 = std::is_same<Real, double>::value;
  if (A.skip && *A.skip) return;  // the whole workgroup: an inner iteration enqueued ahead
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int *cnt = (int *)smem_raw;                       // [2][16] per-wave negative counts
  double *part = (double *)(smem_raw + 2 * 16 * 4);  // [16 waves][16 moments]
  double *wpc = part + 16 * 16;                      // [nang] w*pc
  const int nang = A.nang, cmom = A.cmom;
  const int tid = threadIdx.x, T = blockDim.x, lane = tid & 63, wave = tid >> 6, W = T >> 6;
  const int s = blockIdx.y;
  const int corner = A.corner0 + s / A.ng_job;
  const int g = s % A.ng_job;
  const int tp = A.plane_tiles[plane_off + blockIdx.x];
  const int i = tp & 1023, j = (tp >> 10) & 1023, k = (tp >> 20) & 1023;
  const int nx = A.ext.nx, ny = A.ext.ny, nz = A.ext.nz;
  const bool sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
  const int px = sxp ? i : nx - 1 - i, py = syp ? j : ny - 1 - j, pz = szp ? k : nz - 1 - k;
  const int cell = (pz * ny + py) * nx + px;
  const size_t cells = (size_t)nx * ny * nz;
  Real *fxp = (Real *)A.faces +
              (size_t)s * ((size_t)ny * nz + (size_t)nx * nz + (size_t)nx * ny) * nang;
  Real *fyp = fxp + (size_t)ny * nz * nang;
  Real *fzp = fyp + (size_t)nx * nz * nang;
  Real *sxf = fxp + (size_t)(k * ny + j) * nang;
  Real *syf = fyp + (size_t)(k * nx + i) * nang;
  Real *szf = fzp + (size_t)(j * nx + i) * nang;

  // per-angle constants of this thread's angles
  int an[APL];
  bool aok[APL];
#pragma unroll
  for (int r = 0; r < APL; r++) {
    const int a = tid + r * T;
    aok[r] = a < nang;
    an[r] = aok[r] ? a : 0;
  }
  const double dhi = A.hi, dhj = A.hj, dhk = A.hk;
  const Real hi = (Real)dhi, hj = (Real)dhj, hk = (Real)dhk;
  const Real tolr = (Real)1.0e-12;
  const double *ecc = A.ec + (size_t)corner * cmom * nang;

  // the cell's source moments, t_xs and (TD) vdelt are uniform over the workgroup
  const double *qc;
  Real txs;
  if constexpr (GEN) {
    qc = (const double *)G.qtot.at(px, py, pz, g);
    txs = *(const Real *)G.txs.at(px, py, pz, g);
  } else {
    qc = A.q + ((size_t)g * cells + cell) * cmom;
    txs = ((const Real *)A.txs)[(size_t)g * cells + cell];
  }
  double vd = 0.0;
  bool has_time = false, has_qim = false;
  if constexpr (TD) {
    vd = GEN ? A.vdelt0 : A.vdelt[g];
    has_time = vd != 0.0;
    has_qim = GEN ? G.qim.base != nullptr : A.qim != nullptr;
  }
  const size_t ab = (((size_t)corner * A.ng_job + g) * cells + cell) * (size_t)nang;  // dense angle base
  const Real vdr = (Real)vd;

  Real psi[APL], psii[APL], psij[APL], psik[APL], pc[APL], ti[APL], cx[APL], cy[APL], cz[APL];
  double w_d[APL];
#pragma unroll
  for (int r = 0; r < APL; r++) {
    const int a = an[r];
    const double mu_d = A.mu[a], eta_d = A.eta[a], xi_d = A.xi[a];
    w_d[r] = A.w[a];
    if constexpr (EXACT) {
      cx[r] = (Real)mu_d;
      cy[r] = (Real)eta_d;
      cz[r] = (Real)xi_d;
    } else {
      cx[r] = (Real)(mu_d * dhi);
      cy[r] = (Real)(eta_d * dhj);
      cz[r] = (Real)(xi_d * dhk);
    }
    // angular source (sweep.cc:331-351)
    if constexpr (kFP64) {
      double ps = qc[0];
      for (int l = 1; l < cmom; l++) ps = madd<EXACT>(ecc[(size_t)l * nang + a], qc[l], ps);
      if constexpr (TD) {
        if (has_qim)
          ps += GEN ? ((const double *)G.qim.at(px, py, pz, g))[a] : A.qim[ab + a];
      }
      psi[r] = ps;
    } else {
      float pf = (float)qc[0];
      for (int l = 1; l < cmom; l++) pf = maddf((float)ecc[(size_t)l * nang + a], (float)qc[l], pf);
      if constexpr (TD) {
        if (has_qim)
          pf += (float)(GEN ? ((const double *)G.qim.at(px, py, pz, g))[a] : A.qim[ab + a]);
      }
      psi[r] = pf;
    }
    ti[r] = Real(0);
    if constexpr (TD) {
      if (has_time)
        ti[r] = (Real)(GEN ? ((const double *)G.tin.at(px, py, pz, g))[a] : A.tin[ab + a]);
    }
    // inflow faces: the upstream cell's outflow, the ghost array at the
    // chunk edge, or vacuum
    Real vx = Real(0), vy = Real(0), vz = Real(0);
    if (i > 0)
      vx = sxf[a];
    else if (GEN && G.gx.base)
      vx = (Real)((const double *)G.gx.at(py, pz, g))[a];
    if (j > 0)
      vy = syf[a];
    else if (GEN && G.gy.base)
      vy = (Real)((const double *)G.gy.at(px, pz, g))[a];
    if (k > 0)
      vz = szf[a];
    else if (GEN && G.gz.base)
      vz = (Real)((const double *)G.gz.at(px, py, g))[a];
    psii[r] = vx;
    psij[r] = vy;
    psik[r] = vz;
    const Real dv = GEN ? ((const Real *)G.dinv.at(px, py, pz, g))[a]
                        : ((const Real *)A.dinv)[((size_t)g * cells + cell) * nang + a];
    // initial diamond-difference solution (sweep.cc:352-423)
    Real p = psi[r];
    if constexpr (kFP64 && EXACT) {
      p += psii[r] * cx[r] * hi;
      p += psij[r] * cy[r] * hj;
      p += psik[r] * cz[r] * hk;
      if (TD && has_time) p += vdr * ti[r];
    } else if constexpr (kFP64) {
      p = __builtin_fma(psii[r], cx[r], p);
      p = __builtin_fma(psij[r], cy[r], p);
      p = __builtin_fma(psik[r], cz[r], p);
      if (TD && has_time) p = __builtin_fma(vdr, ti[r], p);
    } else {
      p = maddf(psii[r], cx[r], p);
      p = maddf(psij[r], cy[r], p);
      p = maddf(psik[r], cz[r], p);
      if (TD && has_time) p = maddf(vdr, ti[r], p);
    }
    pc[r] = p * dv;
  }

  bool hx[APL], hy[APL], hz[APL], ht[APL];
#pragma unroll
  for (int r = 0; r < APL; r++) hx[r] = hy[r] = hz[r] = ht[r] = true;
  if (A.fixup) {
    // negative-flux fixup (sweep.cc:424-520), the count summed over the
    // workgroup's waves; counts alternate between two LDS rows so one
    // barrier per pass orders every write before every read
    int buf = 0;
    auto count_negative = [&]() -> int {
      unsigned n = 0;
#pragma unroll
      for (int r = 0; r < APL; r++) {
        const Real p2 = Real(2) * pc[r];
        const bool nxn = aok[r] && p2 - psii[r] < Real(0);
        const bool nyn = aok[r] && p2 - psij[r] < Real(0);
        const bool nzn = aok[r] && p2 - psik[r] < Real(0);
        const bool ntn = TD && has_time && aok[r] && p2 - ti[r] < Real(0);
        hx[r] = hx[r] && !nxn;
        hy[r] = hy[r] && !nyn;
        hz[r] = hz[r] && !nzn;
        ht[r] = ht[r] && !ntn;
        n += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(nxn)) +
             (unsigned)__popcll(__builtin_amdgcn_ballot_w64(nyn)) +
             (unsigned)__popcll(__builtin_amdgcn_ballot_w64(nzn)) +
             (unsigned)__popcll(__builtin_amdgcn_ballot_w64(ntn));
      }
      if (lane == 0) cnt[buf * 16 + wave] = (int)n;
      __syncthreads();
      int total = 0;
      for (int v = 0; v < W; v++) total += cnt[buf * 16 + v];
      buf ^= 1;
      return total;
    };
    int old_neg = count_negative();
    if (old_neg != 0) {
      const int ne = (TD && has_time) ? 4 : 3;
      for (int it = 1; it < ne * nang + 2; ++it) {
#pragma unroll
        for (int r = 0; r < APL; r++) {
          const Real vx = hx[r] ? Real(1) : Real(0), vy = hy[r] ? Real(1) : Real(0),
                     vz = hz[r] ? Real(1) : Real(0), vt = ht[r] ? Real(1) : Real(0);
          Real den, p;
          if constexpr (kFP64 && EXACT) {
            if (TD && has_time) {
              p = psi[r] + 0.5 * (psii[r] * cx[r] * hi * (1.0 + vx) +
                                  psij[r] * cy[r] * hj * (1.0 + vy) +
                                  psik[r] * cz[r] * hk * (1.0 + vz) + ti[r] * vdr * (1.0 + vt));
              den = (p <= 0.0) ? 0.0
                               : (txs + cx[r] * hi * vx + cy[r] * hj * vy + cz[r] * hk * vz +
                                  vdr * vt);
            } else {
              p = psi[r] + 0.5 * (psii[r] * cx[r] * hi * (1.0 + vx) +
                                  psij[r] * cy[r] * hj * (1.0 + vy) +
                                  psik[r] * cz[r] * hk * (1.0 + vz));
              den = (p <= 0.0) ? 0.0 : (txs + cx[r] * hi * vx + cy[r] * hj * vy + cz[r] * hk * vz);
            }
          } else {
            Real t = psii[r] * cx[r] * (Real(1) + vx) + psij[r] * cy[r] * (Real(1) + vy) +
                     psik[r] * cz[r] * (Real(1) + vz);
            Real dd = txs + cx[r] * vx + cy[r] * vy + cz[r] * vz;
            if (TD && has_time) {
              t += ti[r] * vdr * (Real(1) + vt);
              dd += vdr * vt;
            }
            p = psi[r] + Real(0.5) * t;
            den = (p <= Real(0)) ? Real(0) : dd;
          }
          pc[r] = (den < tolr) ? Real(0) : p / den;
        }
        const int neg = count_negative();
        if (neg == old_neg) break;
        old_neg = neg;
      }
    }
  }

  // outflows: the face planes in place, the ghost arrays at the chunk's far
  // side, the time edge (sweep.cc:506-566)
  double pcw[APL];
#pragma unroll
  for (int r = 0; r < APL; r++) {
    const int a = an[r];
    const Real p2 = Real(2) * pc[r];
    const Real ox = (p2 - psii[r]) * (hx[r] ? Real(1) : Real(0));
    const Real oy = (p2 - psij[r]) * (hy[r] ? Real(1) : Real(0));
    const Real oz = (p2 - psik[r]) * (hz[r] ? Real(1) : Real(0));
    pcw[r] = w_d[r] * (double)pc[r];
    if (!aok[r]) continue;
    if (GEN && i == nx - 1 && G.gx.base)
      ((double *)G.gx.at(py, pz, g))[a] = (double)ox;
    else
      sxf[a] = ox;
    if (GEN && j == ny - 1 && G.gy.base)
      ((double *)G.gy.at(px, pz, g))[a] = (double)oy;
    else
      syf[a] = oy;
    if (GEN && k == nz - 1 && G.gz.base)
      ((double *)G.gz.at(px, py, g))[a] = (double)oz;
    else
      szf[a] = oz;
    if constexpr (TD) {
      if (has_time) {
        const double ot = (double)((p2 - ti[r]) * (ht[r] ? Real(1) : Real(0)));
        if constexpr (GEN)
          ((double *)G.tout.at(px, py, pz, g))[a] = ot;
        else
          A.tout[ab + a] = ot;
      }
    }
  }

  // flux moments phi_l = sum_a ec[l][a] * w[a] pc[a] (sweep.cc:568-591)
  auto store = [&](int l, double sum) {
    if constexpr (GEN) {
      double *dst = (l == 0) ? (double *)G.flux.at(px, py, pz, g)
                             : (double *)G.fluxm.at(px, py, pz, g) + (l - 1);
      atomicAdd(dst, sum);
    } else {
      const size_t sc = ((size_t)corner * A.ng_job + g) * cells + cell;
      if (l == 0)
        A.part_flux[sc] = sum;
      else
        A.part_fluxm[sc * (size_t)(cmom - 1) + (l - 1)] = sum;
    }
  };
  if constexpr (EXACT) {
    // the CPU's ascending angle order: one lane per moment sums them all
#pragma unroll
    for (int r = 0; r < APL; r++)
      if (aok[r]) wpc[an[r]] = pcw[r];
    __syncthreads();
    if (tid < cmom) {
      const int l = tid;
      const double *e = ecc + (size_t)l * nang;
      double sum = 0.0;
      for (int a = 0; a < nang; a++) sum = (l == 0) ? sum + wpc[a] : madd<true>(e[a], wpc[a], sum);
      store(l, sum);
    }
  } else {
    // lane partials, wave sums, then the waves in order
    for (int l = 0; l < cmom; l++) {
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < APL; r++)
        if (aok[r]) v = (l == 0) ? v + pcw[r] : __builtin_fma(ecc[(size_t)l * nang + an[r]], pcw[r], v);
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) part[wave * 16 + l] = v;
    }
    __syncthreads();
    if (tid < cmom) {
      double sum = 0.0;
      for (int v = 0; v < W; v++) sum += part[v * 16 + tid];
      store(tid, sum);
    }
  }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
typedef void (*SweepKernel)(SweepArgs, SweepGeneral, int);

template <typename Real, int APL, bool EXACT, bool REC, bool GEN>
static SweepKernel kernel_for_apl(int cpw) {
  switch (cpw) {
    case 1: return k_sweep_tiles<Real, 1, APL, EXACT, REC, GEN>;
    case 2: return k_sweep_tiles<Real, 2, APL, EXACT, REC, GEN>;
    case 3: return k_sweep_tiles<Real, 3, APL, EXACT, REC, GEN>;
    case 4: return k_sweep_tiles<Real, 4, APL, EXACT, REC, GEN>;
    case 5: return k_sweep_tiles<Real, 5, APL, EXACT, REC, GEN>;
    case 6: return k_sweep_tiles<Real, 6, APL, EXACT, REC, GEN>;
    case 8: return k_sweep_tiles<Real, 8, APL, EXACT, REC, GEN>;
    default: return nullptr;
  }
}

// The drop-in (general) variant reads dinv as the reference does, and runs
// two angles per lane only where one cell's angles exceed a wavefront.
template <typename Real, bool EXACT>
static SweepKernel kernel_for(int cpw, int apl, bool rec, bool gen) {
  if (gen) {
    if (rec) return nullptr;
    if (apl == 1) return kernel_for_apl<Real, 1, EXACT, false, true>(cpw);
    if (apl == 2 && cpw == 1) return k_sweep_tiles<Real, 1, 2, EXACT, false, true>;
    return nullptr;
  }
  if (apl == 1)
    return rec ? kernel_for_apl<Real, 1, EXACT, true, false>(cpw)
               : kernel_for_apl<Real, 1, EXACT, false, false>(cpw);
  if (apl == 2)
    return rec ? kernel_for_apl<Real, 2, EXACT, true, false>(cpw)
               : kernel_for_apl<Real, 2, EXACT, false, false>(cpw);
  return nullptr;
}

// Batch kernels: dense, dinv streamed, FP64 (fused or exact) or FP32.
template <typename Real, bool EXACT>
static SweepKernel kernel_for_batch(int cpw, int apl) {
#define SNAP_BATCH_CASE(C, P) \
  if (cpw == C && apl == P) return k_sweep_tiles<Real, C, P, EXACT, false, false, true>;
  SNAP_BATCH_CASE(1, 1) SNAP_BATCH_CASE(2, 1) SNAP_BATCH_CASE(3, 1) SNAP_BATCH_CASE(4, 1)
  SNAP_BATCH_CASE(5, 1) SNAP_BATCH_CASE(6, 1) SNAP_BATCH_CASE(8, 1)
  SNAP_BATCH_CASE(1, 2) SNAP_BATCH_CASE(2, 2) SNAP_BATCH_CASE(3, 2) SNAP_BATCH_CASE(4, 2)
  SNAP_BATCH_CASE(5, 2) SNAP_BATCH_CASE(6, 2) SNAP_BATCH_CASE(8, 2)
#undef SNAP_BATCH_CASE
  return nullptr;
}

// Time-dependent kernels: dense with dinv streamed (FP64 fused or exact, or
// FP32), and the general variant in the exact order for the drop-in.
template <typename Real, bool EXACT>
static SweepKernel kernel_for_td(int cpw, int apl, bool gen) {
#define SNAP_TD_CASE(C, P) \
  if (cpw == C && apl == P) return k_sweep_tiles<Real, C, P, EXACT, false, false, false, true>;
#define SNAP_TDG_CASE(C, P) \
  if (cpw == C && apl == P) return k_sweep_tiles<Real, C, P, EXACT, false, true, false, true>;
  if (gen) {
    if constexpr (EXACT && std::is_same<Real, double>::value) {
      SNAP_TDG_CASE(1, 1) SNAP_TDG_CASE(2, 1) SNAP_TDG_CASE(3, 1) SNAP_TDG_CASE(4, 1)
      SNAP_TDG_CASE(5, 1) SNAP_TDG_CASE(6, 1) SNAP_TDG_CASE(8, 1) SNAP_TDG_CASE(1, 2)
    }
    return nullptr;
  }
  SNAP_TD_CASE(1, 1) SNAP_TD_CASE(2, 1) SNAP_TD_CASE(3, 1) SNAP_TD_CASE(4, 1)
  SNAP_TD_CASE(5, 1) SNAP_TD_CASE(6, 1) SNAP_TD_CASE(8, 1)
  SNAP_TD_CASE(1, 2) SNAP_TD_CASE(2, 2) SNAP_TD_CASE(3, 2) SNAP_TD_CASE(4, 2)
  SNAP_TD_CASE(5, 2) SNAP_TD_CASE(6, 2) SNAP_TD_CASE(8, 2)
#undef SNAP_TD_CASE
#undef SNAP_TDG_CASE
  return nullptr;
}

// One-cell workgroup kernels: FP64 (fused or exact) or FP32, steady or
// time-dependent, dense or general (the general ones in the exact order).
template <typename Real, bool EXACT>
static SweepKernel kernel_for_wide(int apl, bool gen, bool td) {
#define SNAP_WIDE(P, GN, TDV) \
  if (apl == P && gen == GN && td == TDV) return k_sweep_wide<Real, P, EXACT, GN, TDV>;
  SNAP_WIDE(1, false, false) SNAP_WIDE(2, false, false) SNAP_WIDE(1, false, true)
  SNAP_WIDE(2, false, true)
  if constexpr (EXACT && std::is_same<Real, double>::value) {
    SNAP_WIDE(1, true, false) SNAP_WIDE(2, true, false) SNAP_WIDE(1, true, true)
    SNAP_WIDE(2, true, true)
  }
#undef SNAP_WIDE
  return nullptr;
}

static int ensure_lds_attr(SweepKernel k, size_t bytes) {
  return ensure_dynamic_lds((const void *)k, bytes);
}

int sweep_enqueue(const SweepLaunch &L, hipStream_t st, hipEvent_t *events) {
  const SweepPlan &p = *L.plan;
  if (L.nsweeps < 1 || L.nsweeps > 65535) return fail(SNAP_ERR_INVALID, "bad sweep count");
  if (L.quad->nang != p.nang || L.quad->cmom != p.cmom)
    return fail(SNAP_ERR_INVALID, "quadrature does not match the sweep plan");
  if (L.general && L.nsweeps != L.ng_job)
    return fail(SNAP_ERR_INVALID, "the general sweep handles one corner per call");
  if ((int64_t)p.ext.nx * p.ext.ny * p.ext.nz >= (int64_t)1 << 31)
    return fail(SNAP_ERR_UNSUPPORTED, "chunks of 2^31 cells or more");
  if (L.batch && (L.general || L.recompute_dinv))
    return fail(SNAP_ERR_INVALID, "batch sweeps are dense with dinv streamed");
  if (L.td && (L.batch || L.recompute_dinv))
    return fail(SNAP_ERR_INVALID, "time-dependent sweeps read dinv and run one problem");
  if (L.td && !p.td && !p.wide)
    return fail(SNAP_ERR_INVALID, "the plan is not one for time-dependent sweeps");
  if (p.wide && (L.batch || L.recompute_dinv))
    return fail(SNAP_ERR_INVALID, "one-cell sweeps read dinv and run one problem");
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
  A.lpc = p.lpc;
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
  A.skip = L.skip;
  A.plane_tiles = p.plane_tiles;
  A.cell_off = p.cell_off;
  A.cells = p.cells;
  if (L.batch) {
    A.batch_tiles = L.batch->batch_tiles;
    A.descs = L.batch->descs;
    A.active = L.batch->active;
    A.done = L.batch->done;
    A.prec = L.batch->prec;
    A.prec_want = L.batch->prec_want;
    A.total_cells = L.batch->total_cells;
  }
  if (L.td) {
    A.vdelt = L.vdelt;
    A.vdelt0 = L.vdelt0;
    A.tin = L.tin;
    A.tout = L.tout;
    A.qim = L.qim;
    G.tin = L.tin_acc;
    G.tout = L.tout_acc;
    G.qim = L.qim_acc;
    // dense: the caller gives tin/tout whenever a vdelt it passes is nonzero
    if (L.general ? (L.vdelt0 != 0.0 && (!G.tin.base || !G.tout.base)) : !L.vdelt)
      return fail(SNAP_ERR_INVALID, "time-dependent sweep without vdelt or time-edge fluxes");
  }
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
  const SweepKernel k =
      p.wide ? (L.fp32 ? kernel_for_wide<float, false>(p.apl, L.general, L.td)
                       : (L.exact ? kernel_for_wide<double, true>(p.apl, L.general, L.td)
                                  : kernel_for_wide<double, false>(p.apl, L.general, L.td)))
      : L.td ? (L.fp32 ? kernel_for_td<float, false>(p.cpw, p.apl, L.general)
                     : (L.exact ? kernel_for_td<double, true>(p.cpw, p.apl, L.general)
                                : kernel_for_td<double, false>(p.cpw, p.apl, L.general)))
      : L.batch ? (L.fp32 ? kernel_for_batch<float, false>(p.cpw, p.apl)
                        : (L.exact ? kernel_for_batch<double, true>(p.cpw, p.apl)
                                   : kernel_for_batch<double, false>(p.cpw, p.apl)))
      : L.fp32 ? kernel_for<float, false>(p.cpw, p.apl, L.recompute_dinv, L.general)
               : (L.exact ? kernel_for<double, true>(p.cpw, p.apl, L.recompute_dinv, L.general)
                          : kernel_for<double, false>(p.cpw, p.apl, L.recompute_dinv, L.general));
  if (!k)
    return fail(SNAP_ERR_INVALID,
                "no sweep kernel for %d cells per wave, %d angles per lane%s%s", p.cpw, p.apl,
                L.recompute_dinv ? ", dinv recomputed" : "", L.general ? ", general" : "");
  SNAP_TRY(ensure_lds_attr(k, lds));
  const dim3 block(p.waves * 64);
  const int nplanes = L.batch ? L.batch->nplanes : p.nplanes;
  const std::vector<int> &poff = L.batch ? L.batch->plane_off : p.plane_off;
  for (int D = 0; D < nplanes; D++) {
    const int count = poff[D + 1] - poff[D];
    if (count == 0) continue;
    if (events) SNAP_HIP_TRY(hipEventRecord(events[2 * D], st));
    hipLaunchKernelGGL(k, dim3(count, L.nsweeps), block, lds, st, A, G, poff[D]);
    if (events) SNAP_HIP_TRY(hipEventRecord(events[2 * D + 1], st));
  }
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

// ---------------------------------------------------------------------------
// the drop-in context (initialize_gpu_context) and run_gpu_sweep
// ---------------------------------------------------------------------------
namespace {
// Face workspace of one stream.  Its own lock orders a growth (which waits
// for the stream's earlier sweeps before freeing) against enqueues from other
// threads on the same stream; sweeps on different streams never share it.
struct StreamWork {
  hipStream_t stream = nullptr;
  void *faces = nullptr;
  size_t bytes = 0;
  std::mutex lock;
};
struct DeviceContext {
  bool valid = false;
  QuadDev quad;
  Extent3 chunk{};
  // entries are never moved or freed until the context is released, so
  // pointers to them stay valid outside the context lock
  std::vector<std::pair<Extent3, SweepPlan *>> plans;
  std::vector<std::unique_ptr<StreamWork>> work;
};
constexpr int kMaxDevices = 64;
DeviceContext g_ctx[kMaxDevices];
std::mutex g_ctx_mutex;

// Scratch memory of the allocation-free sweep: the plan tables and the tile
// face workspace of one call, on every device the library runs on, held
// until the process ends.  Host code never uses this name as an address; it
// uses the per-device pointer that hipGetSymbolAddress returns.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
void *g_scratch[kMaxDevices];

inline size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

int current_device(int *dev) {
  SNAP_HIP_TRY(hipGetDevice(dev));
  if (*dev < 0 || *dev >= kMaxDevices) return fail(SNAP_ERR_INVALID, "device %d out of range", *dev);
  return SNAP_OK;
}

void context_release(DeviceContext &c) {
  // sweeps already enqueued on any stream may still read the tables and
  // face workspaces freed here
  if (c.valid || !c.work.empty()) (void)hipDeviceSynchronize();
  quad_free(&c.quad);
  for (auto &pp : c.plans) {
    plan_free(pp.second);
    delete pp.second;
  }
  c.plans.clear();
  for (auto &w : c.work)
    if (w->faces) (void)hipFree(w->faces);
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
  // the device address of this device's copy of the allocation-free sweep's
  // scratch array (each device holds its own)
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    SNAP_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[dev] = ptr;
  }
  c.chunk = Extent3{nx_per_chunk, ny_per_chunk, nz_per_chunk};
  c.valid = true;
  return SNAP_OK;
}

namespace snap {
namespace {

int check_sweep_args(const snap_sweep_args_t &a) {
  if (a.mms_source && !a.qim.base) return fail(SNAP_ERR_INVALID, "mms_source set without qim");
  if (a.vdelt != 0.0 && (!a.time_flux_in.base || !a.time_flux_out.base))
    return fail(SNAP_ERR_INVALID, "vdelt != 0 needs time_flux_in and time_flux_out");
  if (a.stride_x_positive != ((a.corner & 1) != 0) || a.stride_y_positive != ((a.corner & 2) != 0) ||
      a.stride_z_positive != ((a.corner & 4) != 0) || a.corner < 0 || a.corner > 7)
    return fail(SNAP_ERR_INVALID, "corner %d does not match the stride directions", a.corner);
  if (a.x_range < 1 || a.y_range < 1 || a.z_range < 1) return fail(SNAP_ERR_INVALID, "empty chunk");
  return SNAP_OK;
}

// Byte distance between consecutive groups' fields of a group-range sweep
// (0 for a single group).
struct DropinGroupStrides {
  int64_t qtot = 0, flux = 0, fluxm = 0, dinv = 0, txs = 0, gx = 0, gy = 0, gz = 0;
};

// The launch of drop-in sweeps on `plan` with the given face workspace,
// from the reference's run_gpu_sweep arguments (accessors relative to the
// chunk origin, ghost points absolute): one (corner, group), or with
// ngroups > 1 the groups ngroups of one corner, group g's fields at g times
// the strides in gs from the first group's.  exact: the reference's CPU
// operation order (bit-equal); else fused multiply-adds.
int dropin_launch(const snap_sweep_args_t &a, const SweepPlan *plan, const QuadDev *quad,
                  void *faces, hipStream_t st, int ngroups = 1,
                  const DropinGroupStrides *gs = nullptr, bool exact = true) {
  const Extent3 e{a.x_range, a.y_range, a.z_range};
  const bool td = a.vdelt != 0.0 || a.mms_source;
  snap_rect3_t b;
  for (int i = 0; i < 3; i++) b.lo[i] = a.origin[i];
  b.hi[0] = a.origin[0] + e.nx - 1;
  b.hi[1] = a.origin[1] + e.ny - 1;
  b.hi[2] = a.origin[2] + e.nz - 1;
  SweepLaunch L;
  memset(&L, 0, sizeof(L));
  L.plan = plan;
  L.quad = quad;
  L.nsweeps = ngroups;
  L.ng_job = ngroups;
  L.corner0 = a.corner;
  L.fixup = a.fixup ? 1 : 0;
  L.hi = a.hi;
  L.hj = a.hj;
  L.hk = a.hk;
  SNAP_TRY(single_acc3(a.qtot, b, &L.qtot, "qtot"));
  SNAP_TRY(single_acc3(a.dinv, b, &L.dinv, "dinv"));
  SNAP_TRY(single_acc3(a.t_xs, b, &L.txs, "t_xs"));
  SNAP_TRY(single_acc3(a.flux, b, &L.flux, "flux"));
  SNAP_TRY(single_acc3(a.fluxm, b, &L.fluxm, "fluxm", quad->cmom == 1));
  if (quad->cmom > 1 && !L.fluxm.base) return fail(SNAP_ERR_INVALID, "fluxm is required when cmom > 1");
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
  if (gs) {
    L.qtot.sg = gs->qtot;
    L.flux.sg = gs->flux;
    L.fluxm.sg = gs->fluxm;
    L.dinv.sg = gs->dinv;
    L.txs.sg = gs->txs;
    L.gx.sg = gs->gx;
    L.gy.sg = gs->gy;
    L.gz.sg = gs->gz;
  }
  L.faces = faces;
  L.fp32 = false;
  L.exact = exact;
  if (plan->seq) {
    // the fused steady form on the one-wavefront tiles (dinv formed in the
    // kernel from t_xs, as CalculateGeometryParam forms the caller's dinv)
    if (exact || td) return fail(SNAP_ERR_INVALID, "one-wavefront drop-in plans are fused and steady");
    L.recompute_dinv = true;
    return sweep_seq_enqueue(L, st, nullptr);
  }
  // time-dependent terms and the MMS source (gpu_sweep.cu:2102-2140 picks the
  // time-dependent kernel when vdelt != 0)
  if (td) {
    L.td = true;
    L.vdelt0 = a.vdelt;
    if (a.vdelt != 0.0) {
      SNAP_TRY(single_acc3(a.time_flux_in, b, &L.tin_acc, "time_flux_in"));
      SNAP_TRY(single_acc3(a.time_flux_out, b, &L.tout_acc, "time_flux_out"));
    }
    if (a.mms_source) SNAP_TRY(single_acc3(a.qim, b, &L.qim_acc, "qim"));
  }
  return sweep_enqueue(L, st);
}

// Tables of a plan built in device memory that the caller provides (the
// allocation-free entry point): tiles by tile diagonal, in plan_layout's
// order (diagonal, then tz, then ty), and for tiled plans the cell-diagonal
// offsets and cells of one tile.  One thread writes them: a few thousand
// entries at most for a chunk that fits the scratch array.
__global__ void k_plan_tables(int nTx, int nTy, int nTz, int TX, int TY, int TZ, int tiled,
                              int *plane_tiles, int *cell_off, int16_t *cells) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int n = 0;
  for (int D = 0; D < nTx + nTy + nTz - 2; D++)
    for (int tz = 0; tz < nTz; tz++)
      for (int ty = 0; ty < nTy; ty++) {
        const int tx = D - ty - tz;
        if (tx >= 0 && tx < nTx) plane_tiles[n++] = tx | (ty << 10) | (tz << 20);
      }
  if (!tiled) return;
  int m = 0;
  const int ncp = TX + TY + TZ - 2;
  for (int d = 0; d < ncp; d++) {
    cell_off[d] = m;
    for (int k = 0; k < TZ; k++)
      for (int j = 0; j < TY; j++) {
        const int i = d - j - k;
        if (i >= 0 && i < TX) cells[m++] = (int16_t)(i | (j << 5) | (k << 10));
      }
  }
  cell_off[ncp] = m;
}


}  // namespace
}  // namespace snap

namespace snap {
namespace {
// run_gpu_sweep's body: the plan of the chunk shape and the stream's face
// workspace (sized for ngroups sweeps), then the launches
int dropin_run(const snap_sweep_args_t &a, snap_stream_t stream, int ngroups,
               const DropinGroupStrides *gs, bool exact) {
  int dev;
  SNAP_TRY(current_device(&dev));
  const Extent3 e{a.x_range, a.y_range, a.z_range};
  // plan per chunk shape, one for steady and one for time-dependent sweeps;
  // fused steady sweeps of up to 64 angles take the one-wavefront tiles
  const bool td = a.vdelt != 0.0 || a.mms_source;
  hipStream_t st = (hipStream_t)stream;
  SweepPlan *plan = nullptr;
  StreamWork *wk = nullptr;
  const QuadDev *quad = nullptr;
  {
    // the context lock covers the table lookups only; workspace growth and
    // the enqueue run under the stream's own lock, so concurrent corner
    // tasks on different streams do not wait on each other
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    DeviceContext &c = g_ctx[dev];
    if (!c.valid)
      return fail(SNAP_ERR_NO_CONTEXT, "initialize_gpu_context was not called on device %d", dev);
    if (a.num_angles != c.quad.nang || a.num_moments != c.quad.nmom)
      return fail(SNAP_ERR_INVALID, "angles/moments differ from the context");
    quad = &c.quad;
    const bool seq = !exact && !td && c.quad.nang <= 64;
    for (auto &pp : c.plans)
      if (pp.first.nx == e.nx && pp.first.ny == e.ny && pp.first.nz == e.nz &&
          pp.second->td == td && pp.second->seq == seq)
        plan = pp.second;
    if (!plan) {
      int TX = 0, TY = 0, TZ = 0, waves = 0, apl = 0;
      plan = new SweepPlan();
      int rc;
      if (seq) {
        rc = plan_create_seq(plan, e, c.quad.nang, c.quad.cmom, 8);
      } else if (c.quad.nang > 128) {  // past two angles per lane of a wavefront
        rc = plan_create_wide(plan, e, c.quad.nang, c.quad.cmom, td);
      } else {
        plan_choose(e, c.quad.nang, c.quad.cmom, 8, &TX, &TY, &TZ, &waves, &apl, td);
        rc = plan_create(plan, e, c.quad.nang, c.quad.cmom, TX, TY, TZ, waves, apl, td);
      }
      if (rc != SNAP_OK) {
        delete plan;
        return rc;
      }
      c.plans.emplace_back(e, plan);
    }
    // per-stream face workspace (gpu_sweep.cu keeps 12 per-stream scratch slots)
    for (auto &w : c.work)
      if (w->stream == st) wk = w.get();
    if (!wk) {
      c.work.emplace_back(new StreamWork());
      wk = c.work.back().get();
      wk->stream = st;
    }
  }
  const size_t need = plan->face_doubles_per_sweep() * sizeof(double) * (size_t)ngroups;
  std::lock_guard<std::mutex> wlock(wk->lock);
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
  return dropin_launch(a, plan, quad, wk->faces, st, ngroups, gs, exact);
}

// the byte distance between consecutive groups' instances of one accessor
// field: equal strides, equally spaced bases (or NULL for every group where
// the field is optional)
int group_step(const void *const *bases, const int64_t *const *strides, int nstr, int n,
               bool allow_null, const char *what, int64_t *sg) {
  *sg = 0;
  const bool null0 = bases[0] == nullptr;
  if (null0 && !allow_null) return fail(SNAP_ERR_INVALID, "%s accessor has a NULL base", what);
  if (n > 1 && !null0) *sg = (const char *)bases[1] - (const char *)bases[0];
  for (int g = 1; g < n; g++) {
    if ((bases[g] == nullptr) != null0)
      return fail(SNAP_ERR_INVALID, "%s: group %d is NULL where group 0 is not (or the reverse)",
                  what, g);
    if (null0) continue;
    for (int i = 0; i < nstr; i++)
      if (strides[g][i] != strides[0][i])
        return fail(SNAP_ERR_UNSUPPORTED, "%s: group %d strides differ from group 0", what, g);
    if ((const char *)bases[g] - (const char *)bases[0] != *sg * g)
      return fail(SNAP_ERR_UNSUPPORTED,
                  "%s: the groups' fields are not equally spaced in one instance (sweep them "
                  "one group at a time)", what);
  }
  return SNAP_OK;
}
}  // namespace
}  // namespace snap

extern "C" int snap_hip_run_gpu_sweep(const snap_sweep_args_t *args, snap_stream_t stream) {
  if (!args) return fail(SNAP_ERR_INVALID, "NULL sweep arguments");
  SNAP_TRY(check_sweep_args(*args));
  return dropin_run(*args, stream, 1, nullptr, true);
}

extern "C" int snap_hip_run_gpu_sweep_groups(const snap_sweep_args_t *groups, int num_groups,
                                             int precision, snap_stream_t stream) {
  if (!groups || num_groups < 1) return fail(SNAP_ERR_INVALID, "no groups to sweep");
  if (num_groups > 65535) return fail(SNAP_ERR_INVALID, "at most 65535 groups per call");
  if (precision != SNAP_PRECISION_FP64 && precision != SNAP_PRECISION_FP64_EXACT)
    return fail(SNAP_ERR_INVALID, "group-range sweeps are FP64 (fused or exact)");
  const snap_sweep_args_t &a = groups[0];
  for (int g = 0; g < num_groups; g++) {
    const snap_sweep_args_t &b = groups[g];
    SNAP_TRY(check_sweep_args(b));
    if (b.vdelt != 0.0 || b.mms_source)
      return fail(SNAP_ERR_UNSUPPORTED,
                  "group-range sweeps are steady (call run_gpu_sweep per group for vdelt != 0 "
                  "or an MMS source)");
    if (b.origin[0] != a.origin[0] || b.origin[1] != a.origin[1] || b.origin[2] != a.origin[2] ||
        b.x_range != a.x_range || b.y_range != a.y_range || b.z_range != a.z_range ||
        b.corner != a.corner || b.num_moments != a.num_moments || b.num_angles != a.num_angles ||
        b.fixup != a.fixup || b.hi != a.hi || b.hj != a.hj || b.hk != a.hk)
      return fail(SNAP_ERR_INVALID, "group %d differs from group 0 in its chunk, corner or controls",
                  g);
  }
  DropinGroupStrides gs;
  std::vector<const void *> bases(num_groups);
  std::vector<const int64_t *> strides(num_groups);
  auto step3 = [&](snap_accessor3_t snap_sweep_args_t::*f, bool allow_null, const char *what,
                   int64_t *out) -> int {
    for (int g = 0; g < num_groups; g++) {
      bases[g] = (groups[g].*f).base;
      strides[g] = (groups[g].*f).stride;
    }
    return group_step(bases.data(), strides.data(), 3, num_groups, allow_null, what, out);
  };
  auto step2 = [&](snap_accessor2_t snap_sweep_args_t::*f, const char *what, int64_t *out) -> int {
    for (int g = 0; g < num_groups; g++) {
      bases[g] = (groups[g].*f).base;
      strides[g] = (groups[g].*f).stride;
    }
    return group_step(bases.data(), strides.data(), 2, num_groups, true, what, out);
  };
  SNAP_TRY(step3(&snap_sweep_args_t::qtot, false, "qtot", &gs.qtot));
  SNAP_TRY(step3(&snap_sweep_args_t::flux, false, "flux", &gs.flux));
  SNAP_TRY(step3(&snap_sweep_args_t::fluxm, true, "fluxm", &gs.fluxm));
  SNAP_TRY(step3(&snap_sweep_args_t::dinv, false, "dinv", &gs.dinv));
  SNAP_TRY(step3(&snap_sweep_args_t::t_xs, false, "t_xs", &gs.txs));
  SNAP_TRY(step2(&snap_sweep_args_t::ghostx, "ghostx", &gs.gx));
  SNAP_TRY(step2(&snap_sweep_args_t::ghosty, "ghosty", &gs.gy));
  SNAP_TRY(step2(&snap_sweep_args_t::ghostz, "ghostz", &gs.gz));
  return dropin_run(a, stream, num_groups, &gs, precision == SNAP_PRECISION_FP64_EXACT);
}

extern "C" int snap_hip_run_gpu_sweep_noalloc(const snap_sweep_args_t *args,
                                              snap_stream_t stream) {
  if (!args) return fail(SNAP_ERR_INVALID, "NULL sweep arguments");
  const snap_sweep_args_t &a = *args;
  SNAP_TRY(check_sweep_args(a));
  int dev;
  SNAP_TRY(current_device(&dev));
  const Extent3 e{a.x_range, a.y_range, a.z_range};
  const bool td = a.vdelt != 0.0 || a.mms_source;
  const QuadDev *quad = nullptr;
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    DeviceContext &c = g_ctx[dev];
    if (!c.valid)
      return fail(SNAP_ERR_NO_CONTEXT, "initialize_gpu_context was not called on device %d", dev);
    if (a.num_angles != c.quad.nang || a.num_moments != c.quad.nmom)
      return fail(SNAP_ERR_INVALID, "angles/moments differ from the context");
    quad = &c.quad;
    base = (char *)g_scratch[dev];
  }
  if (!base) return fail(SNAP_ERR_NO_CONTEXT, "no scratch address for device %d", dev);
  // the plan of this chunk shape, laid out on the host (no device memory)
  SweepPlan plan;
  std::vector<int> tiles, coff;
  std::vector<int16_t> cells;
  const bool wide = quad->nang > 128;
  if (wide) {
    SNAP_TRY(plan_layout_wide(&plan, e, quad->nang, quad->cmom, td, tiles));
  } else {
    int TX = 0, TY = 0, TZ = 0, waves = 0, apl = 0;
    plan_choose(e, quad->nang, quad->cmom, 8, &TX, &TY, &TZ, &waves, &apl, td);
    SNAP_TRY(plan_layout(&plan, e, quad->nang, quad->cmom, TX, TY, TZ, waves, apl, td, tiles, coff,
                         cells));
  }
  // scratch: plan tables, then the face workspace
  const size_t o_tiles = 0;
  const size_t o_coff = a256(o_tiles + sizeof(int) * tiles.size());
  const size_t o_cells = a256(o_coff + sizeof(int) * coff.size());
  const size_t o_faces = a256(o_cells + sizeof(int16_t) * cells.size());
  const size_t total = o_faces + plan.face_doubles_per_sweep() * sizeof(double);
  if (total > kScratchBytes)
    return fail(SNAP_ERR_UNSUPPORTED,
                "a %dx%dx%d chunk with %d angles needs %zu bytes of sweep scratch; the "
                "allocation-free sweep holds %zu (use snap_hip_run_gpu_sweep)",
                e.nx, e.ny, e.nz, quad->nang, total, kScratchBytes);
  hipStream_t st = (hipStream_t)stream;
  plan.plane_tiles = (int *)(base + o_tiles);
  plan.cell_off = wide ? nullptr : (int *)(base + o_coff);
  plan.cells = wide ? nullptr : (int16_t *)(base + o_cells);
  hipLaunchKernelGGL(k_plan_tables, dim3(1), dim3(64), 0, st, plan.nTx, plan.nTy, plan.nTz,
                     plan.TX, plan.TY, plan.TZ, wide ? 0 : 1, plan.plane_tiles, plan.cell_off,
                     plan.cells);
  SNAP_HIP_TRY(hipGetLastError());
  const int rc = dropin_launch(a, &plan, quad, base + o_faces, st);
  // the tables belong to the scratch array: nothing for plan_free to release
  plan.plane_tiles = nullptr;
  plan.cell_off = nullptr;
  plan.cells = nullptr;
  return rc;
}
