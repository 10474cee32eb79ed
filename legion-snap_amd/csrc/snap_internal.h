// snap_internal.h -- shared internals of libsnap_hip (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "snap_hip.h"

namespace snap {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// raise a kernel's dynamic LDS limit to bytes on the current device (once
// per device, kernel and size)
int ensure_dynamic_lds(const void *kernel, size_t bytes);

#define SNAP_HIP_TRY(expr)                                                     \
  do {                                                                         \
    hipError_t snap_e_ = (expr);                                               \
    if (snap_e_ != hipSuccess)                                                 \
      return ::snap::fail(SNAP_ERR_HIP, "%s: %s (%s:%d)", #expr,               \
                          hipGetErrorString(snap_e_), __FILE__, __LINE__);     \
  } while (0)

#define SNAP_TRY(expr)                                                         \
  do {                                                                         \
    int snap_rc_ = (expr);                                                     \
    if (snap_rc_ != SNAP_OK) return snap_rc_;                                  \
  } while (0)

// ---------------------------------------------------------------------------
// accessors: Realm's affine form with local coordinates and a group stride
// ---------------------------------------------------------------------------
// element (x, y, z) of group g is at base + x*s0 + y*s1 + z*s2 + g*sg bytes;
// x, y, z are relative to the subgrid's lower bound (folded into base).
struct Acc3 {
  char *base;
  int64_t s0, s1, s2, sg;
  __host__ __device__ inline char *at(int x, int y, int z, int g) const {
    return base + (int64_t)x * s0 + (int64_t)y * s1 + (int64_t)z * s2 + (int64_t)g * sg;
  }
};

// element (p0, p1) of group g: base + p0*s0 + p1*s1 + g*sg
struct Acc2 {
  char *base;
  int64_t s0, s1, sg;
  __host__ __device__ inline char *at(int p0, int p1, int g) const {
    return base + (int64_t)p0 * s0 + (int64_t)p1 * s1 + (int64_t)g * sg;
  }
};

struct Extent3 {
  int nx, ny, nz;
  __host__ __device__ inline int64_t cells() const { return (int64_t)nx * ny * nz; }
};

// Dense [g][z][y][x][rec] layout of `rec` elements of `elem` bytes.
inline Acc3 dense_acc3(void *base, Extent3 e, int rec, int elem) {
  Acc3 a;
  a.base = (char *)base;
  a.s0 = (int64_t)rec * elem;
  a.s1 = a.s0 * e.nx;
  a.s2 = a.s1 * e.ny;
  a.sg = a.s2 * e.nz;
  return a;
}

// Convert `n` per-group ABI accessors over `bounds` into one Acc3 when they
// share strides and their bases are equally spaced (one SOA instance).
int group_acc3(const snap_accessor3_t *acc, int n, const snap_rect3_t &bounds, Acc3 *out,
               const char *what);
int group_acc2(const snap_accessor2_t *acc, int n, Acc2 *out, const char *what);
int single_acc3(const snap_accessor3_t &acc, const snap_rect3_t &bounds, Acc3 *out,
                const char *what, bool allow_null = false);
int rect_extent(const snap_rect3_t &bounds, Extent3 *e);

// Legendre order of flux moment m >= 1 in 3-D (lma = {1,3,5,7}).
__host__ __device__ inline int moment_order(int m) {
  int l = 1, start = 1;
  while (m >= start + 2 * l + 1) {
    start += 2 * l + 1;
    l++;
  }
  return l;
}

inline int grid_for(int64_t n, int block, int max_blocks = 65536 * 8) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

// ---------------------------------------------------------------------------
// quadrature on the device
// ---------------------------------------------------------------------------
struct QuadDev {
  int nang = 0, nmom = 0, cmom = 0;
  double *mu = nullptr, *eta = nullptr, *xi = nullptr, *w = nullptr;
  double *ec = nullptr;  // [8][cmom][nang]
};
int quad_upload(QuadDev *q, const double *ec, const double *mu, const double *eta,
                const double *xi, const double *w, int nang, int nmom);
void quad_free(QuadDev *q);
// compute_derived_globals (snap.cc:1171-1342), host side, 3-D branch.
int quad_compute(int nang, int nmom, std::vector<double> &mu, std::vector<double> &eta,
                 std::vector<double> &xi, std::vector<double> &w, std::vector<double> &ec);

// ---------------------------------------------------------------------------
// element-wise task kernels (snap_xs.hip, snap_source.hip)
// ---------------------------------------------------------------------------
int launch_expand_xs(Acc2 sig, Acc3 mat, Acc3 xs, Extent3 e, int ng, hipStream_t st);
int launch_expand_scat(Acc2 slgg, int group_start, Acc3 mat, Acc3 xs, Extent3 e, int ng,
                       int rec, hipStream_t st);
int launch_geometry(Acc3 xs, Acc3 dinv, const double *vdelt_host, Extent3 e, int ng, int nang,
                    double hi, double hj, double hk, const QuadDev &q, bool fp32_out,
                    hipStream_t st);
int launch_outer_source(Acc3 qi, Acc3 flux0, Acc2 slgg, Acc3 qo0, Acc3 mat, Extent3 e, int ng,
                        hipStream_t st);
int launch_outer_source_m(Acc3 fluxm, Acc2 slgg, Acc3 qom, Acc3 mat, Extent3 e, int ng,
                          int cmom, const int *lma, int nmom, hipStream_t st);
// CalcInnerSource over dense [g][cell][...] arrays (cg = ng * cells
// cell-groups, the Legendre moment counts 1, 3, 5, 7)
// Kernels of the solver's inner iteration take `skip`: when it is given and
// *skip != 0 (the previous inner test passed; the host enqueues one inner
// iteration ahead of its reading of the test) every block returns at once.
int launch_inner_source_dense(const double *sxs, const double *flux0, const double *q2grp0,
                              const double *fluxm, const double *q2grpm, double *qtot, int64_t cg,
                              int nmom, hipStream_t st, const int32_t *skip = nullptr);
int launch_inner_source(Acc3 sxs, Acc3 flux0, Acc3 q2grp0, Acc3 fluxm, Acc3 q2grpm, Acc3 qtot,
                        Extent3 e, int ng, int cmom, const int *lma, int nmom, hipStream_t st);
int launch_convergence(Acc3 flux0, Acc3 flux0p, Extent3 e, int ng, double epsi, int32_t *result,
                       hipStream_t st, const int32_t *skip = nullptr);
// Outer source of the solver's dense layout: ordered = the CPU's summation
// order (one thread per cell and moment), else FP64 matrix cores.
int launch_outer_dense(const double *qi, const double *flux0, const double *fluxm,
                       const double *slgg, const int32_t *mat, double *q0, double *qm,
                       int64_t cells, int ng, int nmat, int nmom, bool ordered, hipStream_t st);
size_t outer_mfma_lds(int ng, int cmom, int sgt);
// zeta_new = r + S^T u and norms[0] = max|zeta_new - zeta|, norms[1] = max|zeta_new|
// (as the bit patterns of the non-negative doubles)
int launch_adjoint_update(const double *r, const double *u, const double *slgg,
                          const int32_t *mat, const double *zeta, double *zn,
                          unsigned long long *norms, int64_t cells, int ng, int nmat, int nmom,
                          hipStream_t st);
int launch_to_float(const double *a, float *b, int64_t n, hipStream_t st);
// out = (base ? base : 0) + sum over the 8 octant partials, octants in order
int launch_octant_reduce(const double *part, double *out, int64_t n, hipStream_t st,
                         const double *base = nullptr, const int32_t *skip = nullptr);
// dst = src over n doubles
int launch_copy(double *dst, const double *src, int64_t n, hipStream_t st,
                const int32_t *skip = nullptr);
// *out = bit pattern of max |f/fp - 1| over n values (|f| where |fp| < 1e-12)
int launch_max_change(const double *f, const double *fp, int64_t n, unsigned long long *out,
                      hipStream_t st);
// y -= x
int launch_subtract(double *y, const double *x, int64_t n, hipStream_t st);
int launch_fill_int(int32_t *p, int32_t v, hipStream_t st, const int32_t *skip = nullptr);

// ---------------------------------------------------------------------------
// tiled wavefront sweep (snap_sweep.hip)
// ---------------------------------------------------------------------------
struct SweepPlan {
  Extent3 ext{};
  int TX = 0, TY = 0, TZ = 0;
  int nTx = 0, nTy = 0, nTz = 0, ntiles = 0;
  int nang = 0, cmom = 0, waves = 0;
  int apl = 1;                 // angles per lane (1 or 2)
  int lpc = 0;                 // lanes per cell, ceil(nang / apl)
  int cpw = 0;                 // cells per 64-lane wavefront
  int sb = 0;                  // cells per per-wave moment batch
  int maxdiag = 0;             // most cells on one cell diagonal of a tile
  bool td = false;             // for the time-dependent kernels (512 threads at most)
  // one workgroup per cell (any angle count): tiles are single cells, the
  // faces are in-place planes per sweep and `waves` spans the cell's angles
  bool wide = false;
  // one wavefront per 4 x 4 x TZ tile walking its cells in sequence
  // (snap_sweep_seq.hip): no cell tables, waves = 1
  bool seq = false;
  int nplanes = 0;             // tile diagonals
  int ncellplanes = 0;         // cell diagonals inside one tile
  std::vector<int> plane_off;  // host: offset of each tile diagonal in plane_tiles
  int *plane_tiles = nullptr;  // device: packed tile coordinates, by diagonal
  int *cell_off = nullptr;     // device: offsets of cell diagonals in cells
  int16_t *cells = nullptr;    // device: packed (i | j<<5 | k<<10), by diagonal
  size_t lds_bytes(int real_bytes, bool recompute_dinv) const;
  size_t face_doubles_per_sweep() const;  // x + y + z faces, all tiles
};
// apl = 0 picks the angles-per-lane count that fills the wavefront best.
// td: a plan for the time-dependent kernels, whose workgroups stop at 512
// threads (their extra per-lane state needs more than 128 registers).
int plan_create(SweepPlan *p, Extent3 e, int nang, int cmom, int TX, int TY, int TZ, int waves,
                int apl = 0, bool td = false);
void plan_free(SweepPlan *p);
// The plan of the one-cell-per-workgroup sweep (angle counts past 128, or on
// request): cell diagonals of the chunk, angles across up to 1024 lanes.
int plan_create_wide(SweepPlan *p, Extent3 e, int nang, int cmom, bool td = false);
// Host layout of plan_create / plan_create_wide without device memory: the
// plan's fields and the tables plan_create would upload.
int plan_layout(SweepPlan *p, Extent3 e, int nang, int cmom, int TX, int TY, int TZ, int waves,
                int apl, bool td, std::vector<int> &tiles, std::vector<int> &coff,
                std::vector<int16_t> &cells);
int plan_layout_wide(SweepPlan *p, Extent3 e, int nang, int cmom, bool td,
                     std::vector<int> &cells);
// Fill in the tile edges, waves per workgroup and angles per lane left at 0
// for a sweep whose faces are real_bytes wide (8 FP64, 4 FP32).
void plan_choose(Extent3 e, int nang, int cmom, int real_bytes, int *TX, int *TY, int *TZ,
                 int *waves, int *apl, bool td = false);

// One problem of a batch sweep: its chunk, tile grid, offsets into the
// batch's concatenated arrays and face buffer, and cell spacing.
struct BatchDesc {
  Extent3 ext;
  int nTx, nTy, nTz, ntiles;
  int64_t cell_off;  // first cell in the [g][all cells] arrays
  int64_t face_off;  // first face element (of the sweep's real type)
  double hi, hj, hk;
};

// Tile-diagonal launches of a batch: launch D covers diagonal D of every
// problem; batch_tiles holds (problem, packed tile) entries by diagonal.
// Mixed-precision FP64 polish: an inner iteration sweeps in FP64 once the
// previous inner of its outer changed the flux by at most this much (the
// largest |flux0/flux0p - 1|).  Near convergence FP32 sweeps move the flux by
// their own rounding, a few 1e-6 at the worst cell, so below the 1e-5 floor
// they make no progress whatever the tolerance asks for.
constexpr double kFp32ChangeFloor = 1.0e-5;
inline double polish_threshold(double fp64_polish, double epsi) {
  return fp64_polish * epsi > kFp32ChangeFloor ? fp64_polish * epsi : kFp32ChangeFloor;
}

struct BatchSweep {
  int nplanes = 0;
  std::vector<int> plane_off;       // host: offsets of each diagonal in batch_tiles
  const int2 *batch_tiles = nullptr;
  const BatchDesc *descs = nullptr;
  const int *active = nullptr;      // device flags: problem still iterating
  const int *done = nullptr;        // device flags: inner iteration converged this outer
  const int *prec = nullptr;        // device flags: sweep precision (1 = FP64); NULL = all
  int prec_want = 0;                // this pass sweeps the problems with prec[p] == prec_want
  int64_t total_cells = 0;
  bool full_xy = false;             // every problem's x and y extents are multiples of 4
};

struct SweepLaunch {
  const SweepPlan *plan;
  const QuadDev *quad;
  int nsweeps;       // grid.y
  int ng_job;        // sweeps s: corner = corner0 + s / ng_job, group = s % ng_job
  int corner0;
  int fixup;
  double hi, hj, hk;
  bool general;      // true: accessors, ghosts, atomic sink (one sweep per call)
  // dense inputs and partial outputs (general == false)
  const double *q;             // [ng_job][cells][cmom]
  const void *dinv_dense;      // [ng_job][cells][nang] of the sweep's real type
  const void *txs_dense;       // [ng_job][cells] of the sweep's real type
  double *part_flux, *part_fluxm;  // [8][ng_job][cells], [8][ng_job][cells][cmom-1]
  // accessors (general == true)
  Acc3 qtot, dinv, txs, flux, fluxm;
  Acc2 gx, gy, gz;             // ghost faces in physical coordinates (base NULL = vacuum)
  void *faces;                 // [nsweeps][face_doubles_per_sweep] elements of Real
  bool fp32;                   // FP32 sweep arithmetic and faces
  bool exact;                  // FP64 in the reference's operation order, no fused ops
  bool recompute_dinv;         // form dinv from t_xs in the kernel instead of reading it
  const BatchSweep *batch;     // non-NULL: a batch of problems (plan gives the tile shape)
  // one-wavefront tile sweeps of one problem: a second stream that takes half
  // of the corners (used when side, fork and join are all set), so that the
  // tail of each launch on one stream overlaps work on the other
  hipStream_t side;
  hipEvent_t fork, join;
  const int32_t *skip;         // solver inner iterations enqueued ahead (see launch_inner_source_dense)
  // one-wavefront tile sweeps: two groups per wavefront where that form
  // applies (sweep_seq_pair)
  bool pair;
  // time-dependent terms (td): vdelt per job group (dense) or vdelt0 (general);
  // time-edge fluxes in / out and the MMS angular source (NULL = none), dense
  // [8][ng_job][cells][nang] doubles or per-cell accessors of nang doubles
  bool td;
  const double *vdelt;
  double vdelt0;
  const double *tin;
  double *tout;
  const double *qim;
  Acc3 tin_acc, tout_acc, qim_acc;
};
// Enqueue every tile diagonal of the sweep (plan->nplanes launches).  When
// `events` is given (2*nplanes events) each launch is bracketed by a pair.
int sweep_enqueue(const SweepLaunch &L, hipStream_t st, hipEvent_t *events = nullptr);
// The plan of the one-wavefront tile sweep (snap_sweep_seq.hip): 4 x 4 x TZ
// tiles, up to 64 angles, steady dense sweeps.
int plan_create_seq(SweepPlan *p, Extent3 e, int nang, int cmom, int TZ);
// Enqueue every tile diagonal of a one-wavefront tile sweep (dense, steady):
// nplanes launches per stream it uses (sweep_seq_streams).  When `events` is
// given (2 * nplanes * streams events) launch q of diagonal D is bracketed by
// events[2 (D streams + q)] and the next one, on its own stream.
int sweep_seq_enqueue(const SweepLaunch &L, hipStream_t st, hipEvent_t *events = nullptr);
int sweep_seq_streams(const SweepLaunch &L);
// True when the launch asks for, and can take, two groups per wavefront:
// fused, dinv formed in the kernel, up to 32 angles, an even group count,
// full 4 x 4 tiles (bit-equal results to one group per wavefront)
bool sweep_seq_pair(const SweepLaunch &L);
// Quadrature of the drop-in context on device `dev` (NULL before init).
const QuadDev *context_quad(int dev);

// Reverse sweep of all 8*ng (corner, group) sweeps on plan p (see
// snap_adjoint.hip); faces = the forward sweep's tile faces on the same plan.
int adjoint_enqueue(const SweepPlan &p, const QuadDev &quad, int ng, int fixup, double hi,
                    double hj, double hk, const double *q, const double *txs, const double *zeta,
                    double *part_u, double *part_grad, const void *faces, void *adj_faces,
                    bool fp32, hipStream_t st);
int adjoint_plan_fits(const SweepPlan &p, int nang, int real_bytes, int *cpw, int *waves, int *sb);

// Method of manufactured solutions (snap_mms.hip): host tables, reference
// flux and fixed source; the angular source on the device.
struct MmsState {
  std::vector<double> ref;   // host [ng][cells], at tf - dt/2 when time dependent
  std::vector<double> qi;    // host [ng][cells]: phi / v_g when time dependent, else 0
  double *tab = nullptr;     // device: per-axis cell averages cx|sx|cy|sy|cz|sz
  double *ref_dev = nullptr; // device: the unscaled manufactured flux [ng][cells]
};
int mms_prepare(const snap_problem_t &p, MmsState *m);
void mms_free(MmsState *m);
// qim [8][ng][cells][nang] = the unscaled angular source (MMSInitSource)
int mms_source(const MmsState &m, const QuadDev &q, Extent3 e, int ng, int nmat, int nmom,
               const int32_t *mat, const double *sigt, const double *slgg, double *qim,
               hipStream_t st);
int launch_scale(double *x, int64_t n, double f, hipStream_t st);
// MMSCompare (mms.cc:497-541): {max, min, mean} of |phi/ref - 1|
void mms_compare(const double *flux, const double *ref, size_t n, double result[3]);

}  // namespace snap
