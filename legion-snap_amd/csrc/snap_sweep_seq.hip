// snap_sweep_seq.hip -- the mini-KBA sweep with one wavefront per tile
// (reference: MiniKBATask, sweep.cc:242-616 CPU, gpu_sweep.cu:1395-1742 GPU).
//
// Decomposition.  As in snap_sweep.hip, a chunk is cut into tiles and every
// launch covers one tile diagonal of every (octant, group) sweep of the job,
// so launch order is dependency order.  Here a tile is 4 x 4 x TZ cells and
// belongs to ONE wavefront, which walks its z planes in sweep order: every
// cell's upstream neighbours inside the tile were done by the same wavefront
// before it, so no barrier of any kind is needed.  The tile's internal faces
// never leave registers: the x inflows are one carry per row, the y inflows
// a 4-entry pencil and the z inflows a 16-entry plane, all indexed at compile
// time (the loops over a plane are unrolled).  Faces reach HBM only at tile
// entry and exit, in place: one slot per tile row, column and pencil.
// Lanes carry angles (up to 64 per octant, one per lane), so the fixup's
// negative count is a wavefront popcount.
//
// Full tiles of the fused and FP32 sweeps take the fast path, per z plane:
//  1. the angular sources of the plane's 16 cells, psi[cell][a] =
//     sum_l q[cell][l] ec[l][a], as FP64 matrix products
//     (v_mfma_f64_16x16x4f64) into LDS;
//  2. the diamond-difference update of all 16 cells with no branch at all,
//     the negative-outflow tests of every cell OR-ed into one flag, so the
//     scheduler overlaps independent cells (those on one i+j diagonal);
//  3. only when some outflow of the plane came out negative, the plane is
//     swept again cell by cell with the reference's fixup loop (from the
//     plane's saved z inflows) -- about a third of the planes of the 64^3
//     test workload, whose vacuum boundaries and source-free regions make
//     the fixup common;
//  4. the flux moments phi[l][cell] = sum_a ec[l][a] (w[a] pc[cell][a]) as
//     FP64 matrix products.
// The EXACT form (reference operation order, separate roundings, bit-equal
// to the CPU arithmetic) and ragged tiles take the general path: per cell,
// the source expansion on the vector units in the reference's order and the
// fixup test right after the cell (sweep.cc:331-591).
//
// Forms: dense (the solver: [g][cell] arrays, per-octant partial moments),
// BATCH (many problems per launch) and GEN (the drop-in group range: the
// caller's accessors, moments added by atomics, ghost faces staged through
// the tile-face slots by k_seq_ghosts).  The host may split a launch's
// sweeps over two streams (sweep offset s0).  Up to 32 angles, the dense and
// BATCH forms can run two groups per wavefront (PAIR: lanes 0-31 and 32-63),
// which fills the lanes that one group of 24 angles leaves idle.
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "snap_internal.h"

namespace snap {

namespace {

constexpr int kSX = 4, kSY = 4, kNP = kSX * kSY;  // tile cross-section, cells per z plane
#ifndef SEQ_WAVES
#define SEQ_WAVES 2  // waves per SIMD the register allocation aims for
#endif
// (FP32 sweeps of two groups per wavefront aim for 3: at P3 the batch form
// then spills a few registers and still runs the 48-problem mixed batch
// 2.5% faster than at 2 waves per SIMD)
// Wavefronts per workgroup: each takes its own tile of the same sweep, so
// they share the sweep's matrix-core operand tables in LDS.  Two, not four:
// a workgroup's LDS is freed only when its slowest tile (the one with the
// most fixup re-sweeps) ends, and pairs hold less of it idle (the 64^3
// FP64 solve 3% faster); one per workgroup repeats the tables too often.
constexpr int kWGW = 2;

// Per-wavefront LDS, in bytes, for row stride rs and a sweep type of rsz
// bytes: W (the moment product's cell operand, 16 rows plus a junk row of 64
// slots: w*pc in FP64 for the exact kernels, pc in the sweep type for the
// others, whose moment table carries the weights), P (the plane's angular
// sources, or its 16 x 16 source moments on the general path, in the sweep
// type) and Zb (the plane's z inflows, in the sweep type).  An FP32 sweep
// thus takes half the per-wavefront LDS of an FP64 one.
__host__ __device__ constexpr size_t seq_w_bytes(int rs, int wsz, bool junk = true) {
  return (size_t)wsz * ((size_t)kNP * rs + (junk ? 64 : 0));
}
__host__ __device__ constexpr size_t seq_p_bytes(int rs, int rsz) {
  return (size_t)rsz * (kNP * rs > kNP * 16 ? kNP * rs : kNP * 16);
}
__host__ __device__ constexpr size_t seq_wave_bytes(int rs, int rsz, int wsz, bool junk = true) {
  return seq_w_bytes(rs, wsz, junk) + seq_p_bytes(rs, rsz) + (size_t)rsz * kNP * rs;
}

// a*b + c with one rounding in the sweep's type
__device__ inline double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

typedef double f64x4 __attribute__((ext_vector_type(4)));

// lane n's value for the whole wavefront (n is a constant once the plane's
// loops are unrolled, so this is one or two v_readlane)
__device__ inline float bcast(float v, int n) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), n));
}
__device__ inline double bcast(double v, int n) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, n);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), n);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

struct SeqArgs {
  Extent3 ext;
  int TZ, nTx, nTy, nTz, ntiles;
  int nang, cmom, ng_job, corner0, fixup;
  int s0;               // sweep of blockIdx.y = 0 (corner = corner0 + s / ng_job)
  const int32_t *skip;  // non-NULL and set: every workgroup returns at once
  int rs;               // LDS row stride (elements) of the per-plane [cell][angle] tables
  double hi, hj, hk;
  const double *mu, *eta, *xi, *w, *ec;
  const double *q;      // [ng_job][cells][cmom]
  const void *dinv;     // [ng_job][cells][nang] of Real
  const void *txs;      // [ng_job][cells] of Real
  double *part_flux;    // [8][ng_job][cells]
  double *part_fluxm;   // [8][ng_job][cells][cmom-1]
  void *faces;          // [nsweeps][x faces per (ty,tz) | y per (tx,tz) | z per (tx,ty)] of Real
  const int *plane_tiles;
  // a batch of problems (BATCH kernels): per-launch (problem, tile) entries,
  // per-problem chunks (their faces in [nsweeps][pencils] from face_off),
  // and the predicates that skip finished problems or the other precision
  const int2 *batch_tiles;
  const BatchDesc *descs;
  const int *active, *done, *prec;
  int prec_want;
  int64_t total_cells;
  // the drop-in sweep of one chunk (GEN kernels): accessors of the source
  // moments, t_xs and the fluxes the sweep adds to (chunk-local cells,
  // group g at g * sg), and the ghost faces (base NULL = vacuum)
  Acc3 gq, gtxs, gflux, gfluxm;
  Acc2 ggx, ggy, ggz;
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// chunk cell of tile-local (i, j, k) in sweep order, or -1 outside a
// ragged chunk edge: the tile origin's cell plus signed strides (a sweep in
// the -x direction steps through the cells backwards)
struct TileGeom {
  int x0, y0, z0, ex, ey, ez, nx, ny, nz;
  int c0, sx, sy, sz;
  int px0, py0, pz0, dx, dy, dz;  // chunk coordinates of tile cell (0,0,0), steps
  __device__ inline void init(int corner) {
    const bool sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
    px0 = sxp ? x0 : nx - 1 - x0;
    py0 = syp ? y0 : ny - 1 - y0;
    pz0 = szp ? z0 : nz - 1 - z0;
    dx = sxp ? 1 : -1;
    dy = syp ? 1 : -1;
    dz = szp ? 1 : -1;
    c0 = (pz0 * ny + py0) * nx + px0;
    sx = dx;
    sy = dy * nx;
    sz = dz * nx * ny;
  }
  __device__ inline int px(int i) const { return px0 + dx * i; }
  __device__ inline int py(int j) const { return py0 + dy * j; }
  __device__ inline int pz(int k) const { return pz0 + dz * k; }
  __device__ inline int cell(int i, int j, int k) const {
    if (i >= ex || j >= ey || k >= ez) return -1;
    return c0 + i * sx + j * sy + k * sz;
  }
};

}  // namespace

// CMOM: flux moments (1, 4, 9 or 16), a compile-time count so that every loop
// over moments unrolls with register operands at fixed places.  GEN: the
// drop-in form (fused FP64, dinv formed in the kernel): sources and t_xs
// through accessors, ghost faces at the chunk's sides, and the moments added
// to the caller's fluxes by device atomics (the caller's other corner tasks
// add to the same cells concurrently).
// PAIR: two groups per wavefront for up to 32 angles: lanes 0-31 sweep group
// g and lanes 32-63 group g + 1 of the same octant and tile (blockIdx.y
// counts group pairs), each half with its own LDS tables, faces, sources,
// fixup counts and moments; the matrix products run once per group.  Full
// 4 x 4 tiles only (the caller's x and y extents are multiples of 4), fused,
// dinv formed in the kernel.  Every group's results are bit-equal to the
// one-group form: a plane swept again cell by cell because of the other
// group's fixup gives the values of the branch-free pass where no outflow is
// negative.
template <typename Real, bool EXACT, bool REC, int CMOM, bool BATCH = false, bool GEN = false,
          bool PAIR = false>
__global__ void __launch_bounds__(64 * kWGW)
__attribute__((amdgpu_waves_per_eu((PAIR && !std::is_same<Real, double>::value) ? 3 : SEQ_WAVES, 8)))
k_sweep_seq(SeqArgs A, int plane_off, int plane_count) {
  constexpr bool kFP64 = std::is_same<Real, double>::value;
  static_assert(!GEN || (kFP64 && !EXACT && REC && !BATCH), "GEN is the fused FP64 drop-in form");
  static_assert(!PAIR || (!EXACT && REC && !GEN), "PAIR is a fused form that forms dinv");
  constexpr int NGW = PAIR ? 2 : 1;  // groups per wavefront
  // FP64 pairs hold the operand tables in registers (two 16-angle blocks,
  // eight four-angle steps: up to 32 angles) and no junk W row, so that
  // four two-wave workgroups fit a CU's LDS at 24 angles
  constexpr bool kRT = PAIR && kFP64;
  if (A.skip && *A.skip) return;  // the whole workgroup: an inner iteration enqueued ahead
  constexpr int KS = (CMOM + 3) / 4;  // four-moment steps of the source product
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int RS = A.rs;
  const int nang = A.nang;
  const int NB = (nang + 15) >> 4;  // 16-angle blocks of the source product
  const int NST = (nang + 3) >> 2;  // four-angle steps of the moment product
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the lane's group of the pair and its angle slot
  const int hh = PAIR ? (lane >> 5) : 0, la = PAIR ? (lane & 31) : lane;
  const int s = A.s0 + NGW * (int)blockIdx.y;
  const int corner = A.corner0 + s / A.ng_job, g = s % A.ng_job;
  // the sweep's operand tables, shared by the workgroup's wavefronts:
  // EB[nb][kk][lane] = ec[4kk + lane/16][16nb + lane%16] (the source
  // product's B, ec[0] = 1) and EA[st][lane] = ec[lane%16][4st + lane/16]
  // (the moment product's A, times the angle's weight unless EXACT), zero
  // outside the moments and angles
  double *EB = (double *)smem_raw;
  double *EA = EB + (kRT ? 0 : (size_t)NB * KS * 64);
  if constexpr (!kRT) {
    const double *ecc0 = A.ec + (size_t)(A.corner0 + s / A.ng_job) * CMOM * nang;
    for (int idx = threadIdx.x; idx < NB * KS * 64; idx += 64 * kWGW) {
      const int L = idx & 63, blk = idx >> 6, kk = blk % KS, nb = blk / KS;
      const int l = 4 * kk + (L >> 4), ang = 16 * nb + (L & 15);
      EB[idx] = (l < CMOM && ang < nang) ? ((l == 0) ? 1.0 : ecc0[(size_t)l * nang + ang]) : 0.0;
    }
    for (int idx = threadIdx.x; idx < NST * 64; idx += 64 * kWGW) {
      const int L = idx & 63, st = idx >> 6;
      const int l = L & 15, ang = 4 * st + (L >> 4);
      const double e = (l < CMOM && ang < nang) ? ((l == 0) ? 1.0 : ecc0[(size_t)l * nang + ang]) : 0.0;
      EA[idx] = (EXACT || ang >= nang) ? e : e * A.w[ang];
    }
  }
  __syncthreads();  // the only workgroup barrier: before any wavefront can leave
  const int tix = blockIdx.x * kWGW + wave;
  if (tix >= plane_count) return;
  // per wavefront: W: (w*)pc of the plane's cells [slot][angle]; P: their
  // angular sources (fast path) or their source moments [slot][16]
  // (general); Z: the plane's z inflows [slot][angle] while the fixup may
  // sweep it again (fast path)
  using WT = typename std::conditional<EXACT, double, Real>::type;
  // (PAIR: one set of the three tables per group, the lane's own at W, P, Zb)
  const size_t gby = seq_wave_bytes(RS, sizeof(Real), sizeof(WT), !PAIR);
  unsigned char *wbase = (unsigned char *)(EA + (kRT ? 0 : (size_t)NST * 64)) + (size_t)wave * NGW * gby;
  auto W_of = [&](int h) { return (WT *)(wbase + h * gby); };
  auto P_of = [&](int h) {
    return (Real *)(wbase + h * gby + seq_w_bytes(RS, sizeof(WT), !PAIR));
  };
  WT *W = W_of(hh);
  Real *P = P_of(hh);
  Real *Zb = (Real *)(wbase + hh * gby + seq_w_bytes(RS, sizeof(WT), !PAIR) +
                      seq_p_bytes(RS, sizeof(Real)));
  Real *qs = P;
  // the chunk: the launch's own, or one problem of a batch
  Extent3 ext = A.ext;
  int nTx = A.nTx, nTy = A.nTy, nTz = A.nTz;
  double dhi = A.hi, dhj = A.hj, dhk = A.hk;
  size_t coff = 0, face_base = 0, gstride;
  int tp;
  if constexpr (BATCH) {
    const int2 bt = A.batch_tiles[plane_off + tix];
    if (!A.active[bt.x] || A.done[bt.x]) return;  // the whole wavefront
    if (A.prec && A.prec[bt.x] != A.prec_want) return;
    const BatchDesc &Pd = A.descs[bt.x];
    ext = Pd.ext;
    nTx = Pd.nTx;
    nTy = Pd.nTy;
    nTz = Pd.nTz;
    dhi = Pd.hi;
    dhj = Pd.hj;
    dhk = Pd.hk;
    coff = (size_t)Pd.cell_off;
    face_base = (size_t)Pd.face_off;
    gstride = (size_t)A.total_cells;
    tp = bt.y;
  } else {
    gstride = (size_t)ext.nx * ext.ny * ext.nz;
    tp = A.plane_tiles[plane_off + tix];
  }
  const int tx = tp & 1023, ty = (tp >> 10) & 1023, tz = (tp >> 20) & 1023;
  constexpr int cmom = CMOM;
  const int TZ = A.TZ;
  TileGeom T;
  T.nx = ext.nx;
  T.ny = ext.ny;
  T.nz = ext.nz;
  T.x0 = tx * kSX;
  T.y0 = ty * kSY;
  T.z0 = tz * TZ;
  T.ex = min(kSX, T.nx - T.x0);
  T.ey = min(kSY, T.ny - T.y0);
  T.ez = min(TZ, T.nz - T.z0);
  T.init(corner);

  // tile faces, in place: the x faces of a tile row (ty, tz) share one slot,
  // which the upstream tile (tx-1) wrote one launch earlier and this tile
  // reads before it overwrites it with its own outflow (only one tile of the
  // row is on a tile diagonal); likewise the y faces per (tx, tz) and the z
  // faces per (tx, ty)
  const int fxn = kSY * TZ * nang, fyn = kSX * TZ * nang, fzn = kSX * kSY * nang;
  const size_t nfx = (size_t)nTy * nTz * fxn, nfy = (size_t)nTx * nTz * fyn,
               nfz = (size_t)nTx * nTy * fzn;
  Real *faces_s = (Real *)A.faces + face_base + (size_t)s * (nfx + nfy + nfz);
  Real *out_x = faces_s + (size_t)(tz * nTy + ty) * fxn;
  Real *out_y = faces_s + nfx + (size_t)(tz * nTx + tx) * fyn;
  Real *out_z = faces_s + nfx + nfy + (size_t)(ty * nTx + tx) * fzn;
  // (GEN: with a ghost array on the chunk's upstream side, the tiles on that
  // side read their slot too, which the ghost staging filled before the
  // sweep; the last tiles' outflows stay in their slots for the staging
  // back to the ghost arrays)
  const Real *in_x = (tx > 0 || (GEN && A.ggx.base)) ? out_x : nullptr;
  const Real *in_y = (ty > 0 || (GEN && A.ggy.base)) ? out_y : nullptr;
  const Real *in_z = (tz > 0 || (GEN && A.ggz.base)) ? out_z : nullptr;
  // (PAIR: group g + 1's faces are sweep s + 1's, one sweep's faces later)
  const int fs = PAIR ? hh * (int)(nfx + nfy + nfz) : 0;

  // the dense arrays (the GEN form goes through the accessors instead)
  const double *qg = GEN ? nullptr : A.q + ((size_t)g * gstride + coff) * cmom;
  const Real *dg = GEN ? nullptr : (const Real *)A.dinv + ((size_t)g * gstride + coff) * nang;
  const Real *tg = GEN ? nullptr : (const Real *)A.txs + (size_t)g * gstride + coff;
  double *pfg = GEN ? nullptr : A.part_flux + ((size_t)corner * A.ng_job + g) * gstride + coff;
  double *pmg = GEN ? nullptr
                    : A.part_fluxm + (((size_t)corner * A.ng_job + g) * gstride + coff) *
                                         (size_t)(cmom - 1);

  // ---- per-lane constants (lane = angle; lanes past nang carry angle 0
  // with zero weight and never store)
  const bool aok = la < nang;
  const int a = aok ? la : 0;
  const int af = a + fs;  // the lane's angle in its group's faces
  // GEN: a cell's record in an accessor
  auto acc_cell = [&](const Acc3 &X, int i, int j, int k) -> const char * {
    return X.at(T.px(i), T.py(j), T.pz(k), g);
  };
  const double mu_d = A.mu[a], eta_d = A.eta[a], xi_d = A.xi[a];
  const double w_d = aok ? A.w[a] : 0.0;
  // EXACT keeps (psi*mu)*hi as the reference writes it; the others fold the
  // spacing into the direction cosine
  const Real cx = EXACT ? (Real)mu_d : (Real)(mu_d * dhi);
  const Real cy = EXACT ? (Real)eta_d : (Real)(eta_d * dhj);
  const Real cz = EXACT ? (Real)xi_d : (Real)(xi_d * dhk);
  const Real hi = (Real)dhi, hj = (Real)dhj, hk = (Real)dhk;
  const double *ecc = A.ec + (size_t)corner * cmom * nang;
  // dinv formed in the kernel: CalculateGeometryParam's sum (expxs.cc:378-381)
  const double dsum_x = dhi * mu_d, dsum_y = dhj * eta_d, dsum_z = dhk * xi_d;
  const float gsum32 = (float)((dhi * mu_d + dhj * eta_d) + dhk * xi_d);
  const Real tolr = (Real)1.0e-12;
  auto dinv_of = [&](Real t) -> Real {
    if constexpr (kFP64 && EXACT)
      return 1.0 / ((((double)t + dsum_x) + dsum_y) + dsum_z);
    else if constexpr (kFP64) {
      // the fused form: a hardware reciprocal refined by two Newton steps
      // (within an ulp or two of the division)
      const double den = (((double)t + (double)cx) + (double)cy) + (double)cz;
      double r = __builtin_amdgcn_rcp(den);
      double e = __builtin_fma(-den, r, 1.0);
      r = __builtin_fma(r, e, r);
      e = __builtin_fma(-den, r, 1.0);
      return __builtin_fma(r, e, r);
    }
    else {
      // a hardware reciprocal refined by one Newton step
      const float den = (float)t + gsum32;
      const float r = __builtin_amdgcn_rcpf(den);
      return __builtin_fmaf(r, __builtin_fmaf(-den, r, 1.0f), r);
    }
  };
  // 1/x in FP64 to within an ulp or two: the hardware reciprocal refined by
  // two Newton steps
  auto recip = [&](double x) -> double {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
  };
  const int r16 = lane & 15, q4 = lane >> 4;  // matrix-core operand coordinates
  // the register form of the operand tables (kRT): ebr[nb][kk] is EB's entry
  // of this lane and ear[st] EA's
  double ebr[kRT ? 2 : 1][KS], ear[kRT ? 8 : 1];
  if constexpr (kRT) {
    const double *ecc0 = A.ec + (size_t)corner * CMOM * nang;
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
      for (int kk = 0; kk < KS; kk++) {
        const int l = 4 * kk + q4, ang = 16 * nb + r16;
        ebr[nb][kk] =
            (l < CMOM && ang < nang) ? ((l == 0) ? 1.0 : ecc0[(size_t)l * nang + ang]) : 0.0;
      }
#pragma unroll
    for (int st = 0; st < 8; st++) {
      const int l = r16, ang = 4 * st + q4;
      const double e =
          (l < CMOM && ang < nang) ? ((l == 0) ? 1.0 : ecc0[(size_t)l * nang + ang]) : 0.0;
      ear[st] = ang >= nang ? e : e * A.w[ang];
    }
  }
  // W slot of the lane in row n: its angle's column, or (lanes past the
  // row) its own slot of a junk row after the 16 rows, which nothing reads
  // (PAIR: lanes past the row write their zeros to the row's last slot,
  // which is past nang and which nothing reads)
  auto wslot = [&](int n) {
    if constexpr (PAIR)
      return n * RS + (la < RS ? la : RS - 1);
    else
      return la < RS ? n * RS + la : kNP * RS + la;
  };
  // t_xs of the lane's group in cell slot n: lane n's (group g) or lane
  // 32 + n's (group g + 1), which hold slot n of their group
  auto tbcast = [&](Real v, int n) -> Real {
    if constexpr (PAIR) {
      const Real v0 = bcast(v, n), v1 = bcast(v, 32 + n);
      return hh ? v1 : v0;
    } else {
      return bcast(v, n);
    }
  };
  // Z slot of cell n: by angle, so that lanes past nang (which mirror angle
  // 0 and store its faces again) read and write angle 0's slot, with the
  // values angle 0 has
  auto zslot = [&](int n) { return n * RS + a; };
  // a cell's W entry: w*pc (EXACT: summed in the reference's order) or pc
  // (the moment table carries the weights); zero for lanes past nang
  auto w_of = [&](Real pc) -> WT {
    if constexpr (EXACT)
      return aok ? w_d * (double)pc : 0.0;
    else
      return aok ? pc : Real(0);
  };

  // ---- one cell of the general path: DD update, the reference's fixup loop
  // (sweep.cc:402-520) and the outflows; returns pc (fixed: some outflow of
  // the cell came out negative, so the fixup loop ran)
  auto cell_step = [&](Real psi, Real d, Real txs, Real pi_, Real pj_, Real pk_, Real &ox,
                       Real &oy, Real &oz, bool &fixed) -> Real {
    Real pc;
    if constexpr (EXACT) {
      Real p = psi;
      p += pi_ * cx * hi;
      p += pj_ * cy * hj;
      p += pk_ * cz * hk;
      pc = p * d;
    } else {
      Real p = fmar(pi_, cx, psi);
      p = fmar(pj_, cy, p);
      p = fmar(pk_, cz, p);
      pc = p * d;
    }
    bool hx = true, hy = true, hz = true;
    if (A.fixup) {
      // lanes past nang hold angle 0's values and flags (they mirror lane 0
      // exactly) but are not counted; PAIR: each half counts its own group,
      // and a half whose loop has ended (act false) keeps pc and its flags
      // while the other half's passes go on
      bool act = true;
      auto popc = [&](unsigned long long b) -> unsigned {
        if constexpr (PAIR)
          return (unsigned)__popc((unsigned)(b >> (32 * hh)));
        else
          return (unsigned)__popcll(b);
      };
      auto count_negative = [&]() -> unsigned {
        const Real p2 = Real(2) * pc;
        const bool nxn = p2 - pi_ < Real(0);
        const bool nyn = p2 - pj_ < Real(0);
        const bool nzn = p2 - pk_ < Real(0);
        if constexpr (PAIR) {
          hx = hx & !(act & nxn);
          hy = hy & !(act & nyn);
          hz = hz & !(act & nzn);
        } else {
          hx = hx & !nxn;
          hy = hy & !nyn;
          hz = hz & !nzn;
        }
        return popc(__builtin_amdgcn_ballot_w64(aok & nxn)) +
               popc(__builtin_amdgcn_ballot_w64(aok & nyn)) +
               popc(__builtin_amdgcn_ballot_w64(aok & nzn));
      };
      unsigned old_neg = count_negative();
      fixed = old_neg != 0;
      if constexpr (PAIR) {
        act = fixed;
        if (__builtin_amdgcn_ballot_w64(act) != 0) {
          for (int it = 1; it < 3 * nang + 2; ++it) {
            const Real vx = hx ? Real(1) : Real(0), vy = hy ? Real(1) : Real(0),
                       vz = hz ? Real(1) : Real(0);
            const Real t = pi_ * cx * (Real(1) + vx) + pj_ * cy * (Real(1) + vy) +
                           pk_ * cz * (Real(1) + vz);
            const Real p = psi + Real(0.5) * t;
            const Real den = (p <= Real(0)) ? Real(0) : (txs + cx * vx + cy * vy + cz * vz);
            Real pn;
            if constexpr (kFP64)
              pn = (den < tolr) ? Real(0) : p * recip(den);
            else
              pn = (den < tolr) ? Real(0) : p / den;
            pc = act ? pn : pc;
            const unsigned neg = count_negative();
            act = act && neg != old_neg;
            old_neg = neg;
            if (__builtin_amdgcn_ballot_w64(act) == 0) break;
          }
        }
      } else if (old_neg != 0) {
        // a pass that clears no new flag leaves pc unchanged, so the pass
        // after it repeats the count: 3*nang+2 passes always reach the break
        for (int it = 1; it < 3 * nang + 2; ++it) {
          const Real vx = hx ? Real(1) : Real(0), vy = hy ? Real(1) : Real(0),
                     vz = hz ? Real(1) : Real(0);
          Real den, p;
          if constexpr (EXACT) {
            p = psi + 0.5 * (pi_ * cx * hi * (1.0 + vx) + pj_ * cy * hj * (1.0 + vy) +
                             pk_ * cz * hk * (1.0 + vz));
            den = (p <= 0.0) ? 0.0 : (txs + cx * hi * vx + cy * hj * vy + cz * hk * vz);
          } else {
            const Real t = pi_ * cx * (Real(1) + vx) + pj_ * cy * (Real(1) + vy) +
                           pk_ * cz * (Real(1) + vz);
            p = psi + Real(0.5) * t;
            den = (p <= Real(0)) ? Real(0) : (txs + cx * vx + cy * vy + cz * vz);
          }
          if constexpr (EXACT || !kFP64) {
            pc = (den < tolr) ? Real(0) : p / den;
          } else {
            // fused FP64: a refined hardware reciprocal (as for dinv) instead
            // of the division; a zero den is never used.  (FP32 keeps the
            // division: one Newton step's last-bit errors cost the mixed
            // solve measurable accuracy.)
            pc = (den < tolr) ? Real(0) : p * recip(den);
          }
          const unsigned neg = count_negative();
          if (neg == old_neg) break;
          old_neg = neg;
        }
      }
    }
    const Real p2 = Real(2) * pc;
    ox = (p2 - pi_) * (hx ? Real(1) : Real(0));
    oy = (p2 - pj_) * (hy ? Real(1) : Real(0));
    oz = (p2 - pk_) * (hz ? Real(1) : Real(0));
    return pc;
  };

  // ---- flux moments of plane k from W (sweep.cc:568-591)
  auto moments = [&](int k) {
    if constexpr (EXACT) {
      // lane -> cell slot lane % 16, moments lane / 16 + 4 t, each summed in
      // ascending angle order with separate roundings
      const int n = r16;
      const int c = T.cell(n & 3, n >> 2, k);
      if (c >= 0) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int l = q4 + 4 * t;
          if (l >= cmom) break;
          double sum = 0.0;
          if (l == 0) {
            for (int b = 0; b < nang; b++) sum = sum + W[n * RS + b];
            pfg[c] = sum;
          } else {
            const double *e = ecc + (size_t)l * nang;
            for (int b = 0; b < nang; b++) sum = sum + e[b] * W[n * RS + b];
            pmg[(size_t)c * (cmom - 1) + (l - 1)] = sum;
          }
        }
      }
    } else {
      // phi[l][cell] over four angles per product: A[l][kk] = ec[l][4s+kk]
      // (ec[0] = 1) times w in lane l + 16 kk, B[kk][cell] = pc in lane cell + 16 kk,
      // D[l][cell] in lane cell + 16 (l % 4), register l / 4
      // two accumulators halve the dependent-product chain (PAIR: one
      // product per group, from its own W rows, into its own partials)
#pragma unroll
      for (int h = 0; h < NGW; h++) {
        const WT *Wh = W_of(h);
        f64x4 acc = {0.0, 0.0, 0.0, 0.0}, acc2 = {0.0, 0.0, 0.0, 0.0};
        if constexpr (kRT) {
#pragma unroll
          for (int st = 0; st < 8; st++) {
            if (st >= NST) break;
            if (st & 1)
              acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(
                  ear[st], (double)Wh[r16 * RS + 4 * st + q4], acc2, 0, 0, 0);
            else
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(
                  ear[st], (double)Wh[r16 * RS + 4 * st + q4], acc, 0, 0, 0);
          }
        } else {
        int st = 0;
        for (; st + 1 < NST; st += 2) {
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(
              EA[st * 64 + lane], (double)Wh[r16 * RS + 4 * st + q4], acc, 0, 0, 0);
          acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(
              EA[(st + 1) * 64 + lane], (double)Wh[r16 * RS + 4 * st + 4 + q4], acc2, 0, 0, 0);
        }
        if (st < NST)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(
              EA[st * 64 + lane], (double)Wh[r16 * RS + 4 * st + q4], acc, 0, 0, 0);
        }
        acc += acc2;
        const int n = r16;
        const int c = T.cell(n & 3, n >> 2, k);
        if (c >= 0) {
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int l = 4 * r + q4;
            if (l >= cmom) break;
            if constexpr (GEN) {
              double *dst = l == 0 ? (double *)acc_cell(A.gflux, n & 3, n >> 2, k)
                                   : (double *)acc_cell(A.gfluxm, n & 3, n >> 2, k) + (l - 1);
              atomicAdd(dst, acc[r]);
            } else if (l == 0) {
              pfg[(size_t)h * gstride + c] = acc[r];
            } else {
              pmg[((size_t)h * gstride + c) * (cmom - 1) + (l - 1)] = acc[r];
            }
          }
        }
      }
    }
  };

  // ---- faces inside the tile
  Real zf[kNP], yf[kSX], xf[kSY];
#pragma unroll
  for (int n = 0; n < kNP; n++) {
    const int i = n & 3, j = n >> 2;
    zf[n] = (in_z && i < T.ex && j < T.ey) ? in_z[(size_t)(j * kSX + i) * nang + af] : Real(0);
  }
#pragma unroll
  for (int i = 0; i < kSX; i++) yf[i] = Real(0);
#pragma unroll
  for (int j = 0; j < kSY; j++) xf[j] = Real(0);
  // upstream faces of a vacuum boundary read this tile's own (valid) face
  // memory and are replaced by zeros: loads without branches
  const bool has_x = in_x != nullptr, has_y = in_y != nullptr;
  const Real *in_x_ld = has_x ? in_x : out_x;
  const Real *in_y_ld = has_y ? in_y : out_y;
  auto load_xy_inflows = [&](int k, Real(&xin)[kSY], Real(&yin)[kSX]) {
#pragma unroll
    for (int j = 0; j < kSY; j++) {
      const Real v = in_x_ld[(size_t)(k * kSY + min(j, T.ey - 1)) * nang + af];
      xin[j] = (has_x && j < T.ey) ? v : Real(0);
    }
#pragma unroll
    for (int i = 0; i < kSX; i++) {
      const Real v = in_y_ld[(size_t)(k * kSX + min(i, T.ex - 1)) * nang + af];
      yin[i] = (has_y && i < T.ex) ? v : Real(0);
    }
  };

  const bool fast = !EXACT && T.ex == kSX && T.ey == kSY;
  // (PAIR launches cover full tiles only: their callers' x and y extents are
  // multiples of 4)
  if (PAIR && !fast) return;
  if (fast) {
    // ================= fast path: full tile, fused or FP32 =================
    // source operands of plane k: lane L holds q[slot L%16][4 kk + L/16]
    // (PAIR: of each group)
    double qa[NGW][KS];
    auto qa_load = [&](int k) {
      const int c = T.c0 + (r16 & 3) * T.sx + (r16 >> 2) * T.sy + k * T.sz;
#pragma unroll
      for (int h = 0; h < NGW; h++)
#pragma unroll
        for (int kk = 0; kk < KS; kk++) {
          const int l = 4 * kk + q4;
          if constexpr (GEN)
            qa[h][kk] =
                (l < CMOM) ? ((const double *)acc_cell(A.gq, r16 & 3, r16 >> 2, k))[l] : 0.0;
          else
            qa[h][kk] = (l < CMOM) ? qg[((size_t)h * gstride + c) * cmom + l] : 0.0;
        }
    };
    // t_xs of slot L%16 in lane L (read by the cells with a lane broadcast;
    // PAIR: lanes 32-63 hold group g + 1's)
    Real tv = Real(0);
    auto tv_load = [&](int k) {
      const int c = T.c0 + (r16 & 3) * T.sx + (r16 >> 2) * T.sy + k * T.sz;
      if constexpr (GEN)
        tv = *(const Real *)acc_cell(A.gtxs, r16 & 3, r16 >> 2, k);
      else
        tv = tg[(size_t)hh * gstride + c];
    };
    Real dv[kNP];
    auto dv_load = [&](int k) {
#pragma unroll
      for (int n = 0; n < kNP; n++) {
        if constexpr (!REC)
          dv[n] = dg[(size_t)(T.c0 + (n & 3) * T.sx + (n >> 2) * T.sy + k * T.sz) * nang + a];
        else
          dv[n] = Real(0);
      }
    };
    Real xin[kSY], yin[kSX];
    qa_load(0);
    tv_load(0);
    dv_load(0);
    load_xy_inflows(0, xin, yin);
    // the planes that need the fixup cluster (near-zero flux regions): after
    // a plane swept again with the fixup, the next plane goes straight to the
    // cell-by-cell pass (the same arithmetic, without the discarded
    // branch-free pass); after one without any fixup, back to step (2)
    bool predict_fixup = false;
    for (int k = 0; k < T.ez; k++) {
      // (1) psi[slot][angle] = sum_l q[slot][l] ec[l][angle] for each block of
      // 16 angles: A = q (cells x moments), B = ec (moments x angles), D[cell]
      // [angle] in lane angle + 16 (cell % 4), register cell / 4
#pragma unroll
      for (int h = 0; h < NGW; h++) {
        Real *Ph = P_of(h);
        if constexpr (kRT) {
#pragma unroll
          for (int nb = 0; nb < 2; nb++) {
            if (nb >= NB) break;
            f64x4 acc = {0.0, 0.0, 0.0, 0.0};
            const int ang = 16 * nb + r16;
#pragma unroll
            for (int kk = 0; kk < KS; kk++)
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(qa[h][kk], ebr[nb][kk], acc, 0, 0, 0);
            if (ang < RS) {
#pragma unroll
              for (int r = 0; r < 4; r++) Ph[(4 * r + q4) * RS + ang] = (Real)acc[r];
            }
          }
        } else {
          for (int nb = 0; nb < NB; nb++) {
            f64x4 acc = {0.0, 0.0, 0.0, 0.0};
            const int ang = 16 * nb + r16;
#pragma unroll
            for (int kk = 0; kk < KS; kk++)
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(qa[h][kk], EB[(nb * KS + kk) * 64 + lane],
                                                         acc, 0, 0, 0);
            if (ang < RS) {
#pragma unroll
              for (int r = 0; r < 4; r++) Ph[(4 * r + q4) * RS + ang] = (Real)acc[r];
            }
          }
        }
      }
      // the next plane's operands, in flight while this plane is swept
      Real xin_n[kSY], yin_n[kSX], dv_c[kNP];
      const Real tv_c = tv;
#pragma unroll
      for (int n = 0; n < kNP; n++) dv_c[n] = dv[n];
#pragma unroll
      for (int j = 0; j < kSY; j++) xin_n[j] = Real(0);
#pragma unroll
      for (int i = 0; i < kSX; i++) yin_n[i] = Real(0);
      if (k + 1 < T.ez) {
        qa_load(k + 1);
        tv_load(k + 1);
        dv_load(k + 1);
        load_xy_inflows(k + 1, xin_n, yin_n);
      }
      wave_sync();

      // (2) every cell of the plane without a branch, in i+j diagonal order
      // with a scheduling barrier between diagonals, so that at most four
      // independent cells are in flight (the z inflows are kept in LDS for
      // step 3)
      if (A.fixup) {
#pragma unroll
        for (int n = 0; n < kNP; n++) Zb[zslot(n)] = zf[n];
      }
      bool redo = predict_fixup;
      int r0 = 0;  // the first row the cell-by-cell pass sweeps again
      if (!predict_fixup) {
        bool neg[kSY] = {};  // per row of the plane
#pragma unroll
        for (int dg2 = 0; dg2 < kSX + kSY - 1; dg2++) {
#pragma unroll
          for (int i = 0; i < kSX; i++) {
            const int j = dg2 - i;
            if (j < 0 || j >= kSY) continue;  // resolved at compile time
            const int n = j * kSX + i;
            const Real psi = P[n * RS + a];
            Real d;
            if constexpr (REC)
              d = dinv_of(tbcast(tv_c, n));
            else
              d = dv_c[n];
            const Real pi_ = (i == 0) ? xin[j] : xf[j];
            const Real pj_ = (j == 0) ? yin[i] : yf[i];
            const Real pk_ = zf[n];
            Real p = fmar(pi_, cx, psi);
            p = fmar(pj_, cy, p);
            p = fmar(pk_, cz, p);
            const Real pc = p * d;
            const Real p2 = Real(2) * pc;
            const Real ox = p2 - pi_, oy = p2 - pj_, oz = p2 - pk_;
            neg[j] = neg[j] | (ox < Real(0)) | (oy < Real(0)) | (oz < Real(0));
            xf[j] = ox;
            yf[i] = oy;
            zf[n] = oz;
            // lanes past nang mirror lane 0 and store its values again
            if (i == kSX - 1) out_x[(size_t)(k * kSY + j) * nang + af] = ox;
            if (j == kSY - 1) out_y[(size_t)(k * kSX + i) * nang + af] = oy;
            W[wslot(n)] = w_of(pc);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        // the rows above the first one with a negative outflow in any lane
        // are final: no cell of theirs is fixed or downstream of a fixed cell
        r0 = kSY;
#pragma unroll
        for (int j = kSY - 1; j >= 0; j--)
          if (__builtin_amdgcn_ballot_w64(aok & neg[j]) != 0) r0 = j;
        redo = A.fixup && r0 < kSY;
      }
      // (3) a negative outflow anywhere in the plane: sweep it again cell by
      // cell with the fixup, from the plane's z inflows (zf) and tile inflows
      predict_fixup = false;
      if (redo) {
        // A rolled loop, so that this path does not set the register
        // budget: the y inflows of a row are the outflows of four cells
        // earlier (a four-entry queue), the z inflows come from the LDS
        // backup and each z outflow overwrites its own, now consumed, backup
        // slot; the tile inflows are picked from their registers by uniform
        // selects and dinv is formed again.
        // It starts at row r0: the y inflows of that row are the branch-free
        // y outflows of the row above, formed again from the rows' pc
        // exactly as that pass formed them (pc read from angle a's slot,
        // which lanes past nang share with lane 0).
        Real xc = Real(0), yq0 = Real(0), yq1 = Real(0), yq2 = Real(0), yq3 = Real(0);
        if (r0 > 0) {
          wave_sync();
          Real yo[kSX];
#pragma unroll
          for (int i = 0; i < kSX; i++) yo[i] = yin[i];
#pragma unroll 1
          for (int j = 0; j < r0; j++) {
#pragma unroll
            for (int i = 0; i < kSX; i++) yo[i] = Real(2) * (Real)W[(j * kSX + i) * RS + a] - yo[i];
          }
          yq0 = yo[0];
          yq1 = yo[1];
          yq2 = yo[2];
          yq3 = yo[3];
        }
        const int n0 = kSX * r0;
        // each cell's LDS operands are read one cell ahead
        Real psi_nx = P[n0 * RS + a], pk_nx = Zb[zslot(n0)];
#pragma unroll 1
        for (int n = n0; n < kNP; n++) {
          const int i = n & 3, j = n >> 2;
          const int c = T.c0 + i * T.sx + j * T.sy + k * T.sz;
          const Real psi = psi_nx, pk_ = pk_nx;
          if (n + 1 < kNP) {
            psi_nx = P[(n + 1) * RS + a];
            pk_nx = Zb[zslot(n + 1)];
          }
          const Real txs = tbcast(tv_c, n);
          Real d;
          if constexpr (REC)
            d = dinv_of(txs);
          else
            d = dg[(size_t)c * nang + a];
          const Real xj = j == 0 ? xin[0] : j == 1 ? xin[1] : j == 2 ? xin[2] : xin[3];
          const Real yi = i == 0 ? yin[0] : i == 1 ? yin[1] : i == 2 ? yin[2] : yin[3];
          const Real pi_ = i == 0 ? xj : xc;
          const Real pj_ = j == 0 ? yi : yq0;
          Real ox, oy, oz;
          bool fixed = false;
          const Real pc = cell_step(psi, d, txs, pi_, pj_, pk_, ox, oy, oz, fixed);
          predict_fixup = predict_fixup || fixed;
          xc = ox;
          yq0 = yq1;
          yq1 = yq2;
          yq2 = yq3;
          yq3 = oy;
          Zb[zslot(n)] = oz;
          if (i == kSX - 1) out_x[(size_t)(k * kSY + j) * nang + af] = ox;
          if (j == kSY - 1) out_y[(size_t)(k * kSX + i) * nang + af] = oy;
          W[wslot(n)] = w_of(pc);
        }
#pragma unroll
        for (int n = 0; n < kNP; n++)
          if (n >= n0) zf[n] = Zb[zslot(n)];
        // (PAIR: either group's fixup sends both to the cell-by-cell pass)
        if constexpr (PAIR) predict_fixup = __builtin_amdgcn_ballot_w64(predict_fixup) != 0;
      }
      wave_sync();
      // (4) flux moments
      moments(k);
      wave_sync();
#pragma unroll
      for (int j = 0; j < kSY; j++) xin[j] = xin_n[j];
#pragma unroll
      for (int i = 0; i < kSX; i++) yin[i] = yin_n[i];
    }
  } else if constexpr (!PAIR) {
    // ================= general path: exact form, ragged tiles ==============
    Real ecl[CMOM];  // ecl[l] = ec[l][a], l >= 1
#pragma unroll
    for (int l = 1; l < CMOM; l++) ecl[l] = (Real)ecc[(size_t)l * nang + a];
    // source moments of plane k in LDS: 16 cells x cmom values, four per
    // lane (the loads of the next plane are issued while the current one is
    // swept, and stored once its moments are done)
    double qn[4];
    auto q_load = [&](int k) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int idx = lane + 64 * r;  // = slot * 16 + l
        const int slot = idx >> 4, l = idx & 15;
        const int c = (k < T.ez && l < cmom) ? T.cell(slot & 3, slot >> 2, k) : -1;
        if constexpr (GEN)
          qn[r] = c >= 0 ? ((const double *)acc_cell(A.gq, slot & 3, slot >> 2, k))[l] : 0.0;
        else
          qn[r] = c >= 0 ? qg[(size_t)c * cmom + l] : 0.0;
      }
    };
    auto q_store = [&]() {
#pragma unroll
      for (int r = 0; r < 4; r++) qs[lane + 64 * r] = (Real)qn[r];
    };
    q_load(0);
    q_store();
    for (int k = 0; k < T.ez; k++) {
      Real xin[kSY], yin[kSX];
      load_xy_inflows(k, xin, yin);
      if (k + 1 < T.ez) q_load(k + 1);
      wave_sync();
#pragma unroll
      for (int n = 0; n < kNP; n++) {
        const int i = n & 3, j = n >> 2;
        if (i >= T.ex || j >= T.ey) continue;  // uniform: outside a ragged edge
        const int c = T.cell(i, j, k);
        const Real *q = qs + n * 16;
        // angular source (sweep.cc:331-345)
        Real psi;
        if constexpr (EXACT) {
          double p = q[0];
#pragma unroll
          for (int l = 1; l < CMOM; l++) p += (double)ecl[l] * (double)q[l];
          psi = (Real)p;
        } else {
          // two independent chains halve the dependent-FMA latency
          Real p0 = q[0], p1 = Real(0);
#pragma unroll
          for (int l = 1; l < CMOM; l++) {
            if (l & 1)
              p0 = fmar(ecl[l], q[l], p0);
            else
              p1 = fmar(ecl[l], q[l], p1);
          }
          psi = CMOM > 2 ? p0 + p1 : p0;
        }
        const Real txs = GEN ? *(const Real *)acc_cell(A.gtxs, i, j, k) : tg[c];
        const Real d = REC ? dinv_of(txs) : dg[(size_t)c * nang + a];
        const Real pi_ = (i == 0) ? xin[j] : xf[j];
        const Real pj_ = (j == 0) ? yin[i] : yf[i];
        Real ox, oy, oz;
        bool fixed = false;
        const Real pc = cell_step(psi, d, txs, pi_, pj_, zf[n], ox, oy, oz, fixed);
        xf[j] = ox;
        yf[i] = oy;
        zf[n] = oz;
        if (aok) {
          if (i == T.ex - 1) out_x[(size_t)(k * kSY + j) * nang + af] = ox;
          if (j == T.ey - 1) out_y[(size_t)(k * kSX + i) * nang + af] = oy;
        }
        W[wslot(n)] = w_of(pc);
      }
      wave_sync();
      moments(k);
      wave_sync();
      if (k + 1 < T.ez) {
        q_store();
        wave_sync();
      }
    }
  }
  // ---- the tile's z outflow face
  if (aok) {
#pragma unroll
    for (int n = 0; n < kNP; n++) {
      const int i = n & 3, j = n >> 2;
      if (i < T.ex && j < T.ey) {
        out_z[(size_t)(j * kSX + i) * nang + af] = zf[n];
      }
    }
  }
}

// Ghost faces of the drop-in form, through the pencil slots: before the
// sweep an upstream ghost array fills the slots its side's tiles read, and
// after it the downstream side's slots, which then hold the last tiles'
// outflows, go back to the ghost array -- read at the upstream face, the
// outflow written over it, as the reference keeps its ghost regions
// (sweep.cc:536-566).  One thread per (face point, angle) of sweep
// blockIdx.y; axis 0: the x faces over (y, z), 1: y over (x, z), 2: z over
// (x, y), in sweep order (p, q).
__global__ void k_seq_ghosts(SeqArgs A, int axis, int to_ghost) {
  const int nang = A.nang, TZ = A.TZ;
  const Extent3 e = A.ext;
  const int np = axis == 0 ? e.ny : e.nx;
  const int nq = axis == 2 ? e.ny : e.nz;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)np * nq * nang) return;
  const int a = (int)(idx % nang);
  const int pq = (int)(idx / nang), p = pq % np, q = pq / np;
  const int s = A.s0 + blockIdx.y;
  const int corner = A.corner0 + s / A.ng_job, g = s % A.ng_job;
  const bool sxp = (corner & 1) != 0, syp = (corner & 2) != 0, szp = (corner & 4) != 0;
  const size_t fxn = (size_t)kSY * TZ * nang, fyn = (size_t)kSX * TZ * nang,
               fzn = (size_t)kSX * kSY * nang;
  const size_t nfx = (size_t)A.nTy * A.nTz * fxn, nfy = (size_t)A.nTx * A.nTz * fyn,
               nfz = (size_t)A.nTx * A.nTy * fzn;
  double *faces_s = (double *)A.faces + (size_t)s * (nfx + nfy + nfz);
  double *slot, *gp;
  if (axis == 0) {  // (y', z')
    const int ty = p / kSY, j = p % kSY, tz = q / TZ, k = q % TZ;
    slot = faces_s + (size_t)(tz * A.nTy + ty) * fxn + (size_t)(k * kSY + j) * nang + a;
    gp = (double *)A.ggx.at(syp ? p : e.ny - 1 - p, szp ? q : e.nz - 1 - q, g) + a;
  } else if (axis == 1) {  // (x', z')
    const int tx = p / kSX, i = p % kSX, tz = q / TZ, k = q % TZ;
    slot = faces_s + nfx + (size_t)(tz * A.nTx + tx) * fyn + (size_t)(k * kSX + i) * nang + a;
    gp = (double *)A.ggy.at(sxp ? p : e.nx - 1 - p, szp ? q : e.nz - 1 - q, g) + a;
  } else {  // (x', y')
    const int tx = p / kSX, i = p % kSX, ty = q / kSY, j = q % kSY;
    slot = faces_s + nfx + nfy + (size_t)(ty * A.nTx + tx) * fzn + (size_t)(j * kSX + i) * nang + a;
    gp = (double *)A.ggz.at(sxp ? p : e.nx - 1 - p, syp ? q : e.ny - 1 - q, g) + a;
  }
  if (to_ghost)
    *gp = *slot;
  else
    *slot = *gp;
}

// ---------------------------------------------------------------------------
// plan and launch
// ---------------------------------------------------------------------------
namespace {
// LDS row stride of the [cell][angle] tables: the angles rounded up to the
// four-angle matrix steps, plus one so the 16 rows start on different banks
int seq_row_stride(int nang) { return ((nang + 3) & ~3) + 1; }
size_t seq_lds_bytes(int nang, int cmom, bool fp32, bool exact, int ngw) {
  const int rs = seq_row_stride(nang);
  // (FP64 pairs keep the operand tables in registers)
  const size_t tables = (ngw == 2 && !fp32)
                            ? 0
                            : (size_t)((nang + 15) / 16) * ((cmom + 3) / 4) * 64 +
                                  (size_t)((nang + 3) / 4) * 64;
  const int rsz = fp32 ? 4 : 8;
  return sizeof(double) * tables +
         (size_t)kWGW * ngw * seq_wave_bytes(rs, rsz, exact ? 8 : rsz, ngw == 1);
}
}  // namespace

int plan_create_seq(SweepPlan *p, Extent3 e, int nang, int cmom, int TZ) {
  plan_free(p);
  if (e.nx < 1 || e.ny < 1 || e.nz < 1) return fail(SNAP_ERR_INVALID, "empty chunk");
  if (nang < 1 || nang > 64)
    return fail(SNAP_ERR_UNSUPPORTED, "the one-wavefront tile sweep handles 1..64 angles, got %d",
                nang);
  if (cmom != 1 && cmom != 4 && cmom != 9 && cmom != 16)
    return fail(SNAP_ERR_INVALID, "cmom must be 1, 4, 9 or 16, got %d", cmom);
  if (TZ < 1 || TZ > 1023) return fail(SNAP_ERR_INVALID, "tile depth must be 1..1023");
  p->ext = e;
  p->seq = true;
  p->TX = kSX;
  p->TY = kSY;
  p->TZ = std::min(TZ, e.nz);
  p->nTx = (e.nx + kSX - 1) / kSX;
  p->nTy = (e.ny + kSY - 1) / kSY;
  p->nTz = (e.nz + p->TZ - 1) / p->TZ;
  if (p->nTx > 1023 || p->nTy > 1023 || p->nTz > 1023)
    return fail(SNAP_ERR_INVALID, "too many tiles along one axis");
  p->ntiles = p->nTx * p->nTy * p->nTz;
  p->nang = nang;
  p->cmom = cmom;
  p->waves = 1;
  p->apl = 1;
  p->lpc = nang;
  p->cpw = 1;
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
  SNAP_HIP_TRY(hipMalloc((void **)&p->plane_tiles, sizeof(int) * tiles.size()));
  SNAP_HIP_TRY(hipMemcpy(p->plane_tiles, tiles.data(), sizeof(int) * tiles.size(),
                         hipMemcpyHostToDevice));
  return SNAP_OK;
}

typedef void (*SeqKernel)(SeqArgs, int, int);

template <typename Real, bool EXACT>
static SeqKernel seq_kernel(int cmom, bool rec) {
  switch (cmom) {
    case 1: return rec ? k_sweep_seq<Real, EXACT, true, 1> : k_sweep_seq<Real, EXACT, false, 1>;
    case 4: return rec ? k_sweep_seq<Real, EXACT, true, 4> : k_sweep_seq<Real, EXACT, false, 4>;
    case 9: return rec ? k_sweep_seq<Real, EXACT, true, 9> : k_sweep_seq<Real, EXACT, false, 9>;
    case 16: return rec ? k_sweep_seq<Real, EXACT, true, 16> : k_sweep_seq<Real, EXACT, false, 16>;
    default: return nullptr;
  }
}

// Drop-in kernels: fused FP64 through accessors, dinv formed in the kernel.
static SeqKernel seq_kernel_gen(int cmom) {
  switch (cmom) {
    case 1: return k_sweep_seq<double, false, true, 1, false, true>;
    case 4: return k_sweep_seq<double, false, true, 4, false, true>;
    case 9: return k_sweep_seq<double, false, true, 9, false, true>;
    case 16: return k_sweep_seq<double, false, true, 16, false, true>;
    default: return nullptr;
  }
}

// Batch kernels: fused FP64 or FP32, dinv formed in the kernel.
template <typename Real>
static SeqKernel seq_kernel_batch(int cmom) {
  switch (cmom) {
    case 1: return k_sweep_seq<Real, false, true, 1, true>;
    case 4: return k_sweep_seq<Real, false, true, 4, true>;
    case 9: return k_sweep_seq<Real, false, true, 9, true>;
    case 16: return k_sweep_seq<Real, false, true, 16, true>;
    default: return nullptr;
  }
}

// Two groups per wavefront: fused FP64 or FP32, dense or batch.
template <typename Real, bool BATCH>
static SeqKernel seq_kernel_pair(int cmom) {
  switch (cmom) {
    case 1: return k_sweep_seq<Real, false, true, 1, BATCH, false, true>;
    case 4: return k_sweep_seq<Real, false, true, 4, BATCH, false, true>;
    case 9: return k_sweep_seq<Real, false, true, 9, BATCH, false, true>;
    case 16: return k_sweep_seq<Real, false, true, 16, BATCH, false, true>;
    default: return nullptr;
  }
}

int sweep_seq_streams(const SweepLaunch &L) {
  const int ncorner = L.ng_job > 0 ? L.nsweeps / L.ng_job : 0;
  return (L.side && L.fork && L.join && ncorner * L.ng_job == L.nsweeps && ncorner % 2 == 0)
             ? 2
             : 1;
}

// Two groups per wavefront (lanes 0-31 and 32-63): asked for by the caller
// and taken for fused sweeps that form dinv, of up to 32 angles, with group
// pairs inside every stream's part of the sweeps, on full 4 x 4 tiles (x and
// y extents multiples of 4: the chunk's, or every batch problem's)
bool sweep_seq_pair(const SweepLaunch &L) {
  const SweepPlan &p = *L.plan;
  if (!L.pair || !p.seq || L.general || L.exact || L.td || !L.recompute_dinv || p.nang > 32)
    return false;
  if (L.ng_job < 2 || L.ng_job % 2 != 0 || (L.nsweeps / sweep_seq_streams(L)) % L.ng_job != 0)
    return false;
  if (L.batch ? !L.batch->full_xy : (p.ext.nx % kSX != 0 || p.ext.ny % kSY != 0)) return false;
  // the second group's face offset is a 32-bit lane value
  return p.face_doubles_per_sweep() <= (size_t)INT32_MAX / 2;
}

int sweep_seq_enqueue(const SweepLaunch &L, hipStream_t st, hipEvent_t *events) {
  const SweepPlan &p = *L.plan;
  if (!p.seq) return fail(SNAP_ERR_INVALID, "not a one-wavefront tile plan");
  if (L.td) return fail(SNAP_ERR_INVALID, "the one-wavefront tile sweep is steady");
  if (L.batch && (L.exact || !L.recompute_dinv))
    return fail(SNAP_ERR_INVALID, "one-wavefront batch sweeps are fused and form dinv in the kernel");
  if (L.general) {
    if (L.exact || L.fp32 || L.batch || !L.recompute_dinv)
      return fail(SNAP_ERR_INVALID,
                  "the one-wavefront drop-in sweep is fused FP64 and forms dinv in the kernel");
    if (!L.qtot.base || !L.txs.base || !L.flux.base || (p.cmom > 1 && !L.fluxm.base))
      return fail(SNAP_ERR_INVALID, "drop-in sweep accessors missing");
  } else if (!L.q || !L.txs_dense || !L.part_flux || (p.cmom > 1 && !L.part_fluxm) ||
             (!L.recompute_dinv && !L.dinv_dense)) {
    return fail(SNAP_ERR_INVALID, "dense sweep inputs missing");
  }
  if (L.nsweeps < 1 || L.nsweeps > 65535) return fail(SNAP_ERR_INVALID, "bad sweep count");
  SeqArgs A;
  memset(&A, 0, sizeof(A));
  A.ext = p.ext;
  A.TZ = p.TZ;
  A.nTx = p.nTx;
  A.nTy = p.nTy;
  A.nTz = p.nTz;
  A.ntiles = p.ntiles;
  A.nang = p.nang;
  A.cmom = p.cmom;
  A.rs = seq_row_stride(p.nang);
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
  A.skip = L.skip;
  if (L.general) {
    A.gq = L.qtot;
    A.gtxs = L.txs;
    A.gflux = L.flux;
    A.gfluxm = L.fluxm;
    A.ggx = L.gx;
    A.ggy = L.gy;
    A.ggz = L.gz;
  }
  if (L.batch) {
    A.batch_tiles = L.batch->batch_tiles;
    A.descs = L.batch->descs;
    A.active = L.batch->active;
    A.done = L.batch->done;
    A.prec = L.batch->prec;
    A.prec_want = L.batch->prec_want;
    A.total_cells = L.batch->total_cells;
  }
  const int nstr = sweep_seq_streams(L);
  const int part = L.nsweeps / nstr;
  const int ngw = sweep_seq_pair(L) ? 2 : 1;
  const SeqKernel k =
      ngw == 2 ? (L.batch ? (L.fp32 ? seq_kernel_pair<float, true>(p.cmom)
                                    : seq_kernel_pair<double, true>(p.cmom))
                          : (L.fp32 ? seq_kernel_pair<float, false>(p.cmom)
                                    : seq_kernel_pair<double, false>(p.cmom)))
      : L.general ? seq_kernel_gen(p.cmom)
      : L.batch   ? (L.fp32 ? seq_kernel_batch<float>(p.cmom) : seq_kernel_batch<double>(p.cmom))
      : L.fp32    ? seq_kernel<float, false>(p.cmom, L.recompute_dinv)
      : L.exact   ? seq_kernel<double, true>(p.cmom, L.recompute_dinv)
                  : seq_kernel<double, false>(p.cmom, L.recompute_dinv);
  if (!k) return fail(SNAP_ERR_UNSUPPORTED, "no one-wavefront tile sweep for %d moments", p.cmom);
  const size_t lds = seq_lds_bytes(p.nang, p.cmom, L.fp32, L.exact && !L.fp32, ngw);
  SNAP_TRY(ensure_dynamic_lds((const void *)k, lds));
  const int nplanes = L.batch ? L.batch->nplanes : p.nplanes;
  const std::vector<int> &poff = L.batch ? L.batch->plane_off : p.plane_off;
  // Two streams, when the caller gives one: each takes half of the corners
  // (sweeps s0 .. s0 + part - 1, their own faces and partials), and the
  // launches alternate between them, so that the drain of one stream's
  // launch overlaps the other's; measured 5% off the 64^3 FP64 solve (more
  // streams were slower).  The side stream starts after, and the caller's
  // stream resumes after, everything enqueued on the other.
  const hipStream_t ss[2] = {st, nstr > 1 ? L.side : st};
  // the drop-in form's ghost faces (the drop-in runs on one stream)
  const Acc2 *ghosts[3] = {&L.gx, &L.gy, &L.gz};
  auto stage_ghosts = [&](int to_ghost) {
    if (!L.general) return;
    A.s0 = 0;
    for (int axis = 0; axis < 3; axis++) {
      if (!ghosts[axis]->base) continue;
      const int np = axis == 0 ? p.ext.ny : p.ext.nx, nq = axis == 2 ? p.ext.ny : p.ext.nz;
      const int64_t n = (int64_t)np * nq * p.nang;
      hipLaunchKernelGGL(k_seq_ghosts, dim3((unsigned)((n + 255) / 256), L.nsweeps), dim3(256), 0,
                         st, A, axis, to_ghost);
    }
  };
  stage_ghosts(0);
  if (nstr > 1) {
    SNAP_HIP_TRY(hipEventRecord(L.fork, st));
    SNAP_HIP_TRY(hipStreamWaitEvent(L.side, L.fork, 0));
  }
  for (int D = 0; D < nplanes; D++) {
    const int count = poff[D + 1] - poff[D];
    if (count == 0) continue;
    for (int q = 0; q < nstr; q++) {
      A.s0 = q * part;
      hipEvent_t *ev = events ? events + 2 * ((size_t)D * nstr + q) : nullptr;
      if (ev) SNAP_HIP_TRY(hipEventRecord(ev[0], ss[q]));
      hipLaunchKernelGGL(k, dim3((count + kWGW - 1) / kWGW, part / ngw), dim3(64 * kWGW), lds,
                         ss[q], A, poff[D], count);
      if (ev) SNAP_HIP_TRY(hipEventRecord(ev[1], ss[q]));
    }
  }
  if (nstr > 1) {
    SNAP_HIP_TRY(hipEventRecord(L.join, L.side));
    SNAP_HIP_TRY(hipStreamWaitEvent(st, L.join, 0));
  }
  stage_ghosts(1);
  SNAP_HIP_TRY(hipGetLastError());
  return SNAP_OK;
}

}  // namespace snap
