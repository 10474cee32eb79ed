// snap_batch.hip -- many independent fixed-source problems of uneven grid
// sizes solved together.
//
// No counterpart in the reference (which runs one problem per process).
// The problems share ng, nang, nmom and the iteration controls; their
// meshes, material maps, material tables and sources differ.  Every per-cell
// array is the concatenation of the problems' arrays ([g][all cells][...]),
// so the outer source, inner source, octant reduction and copies run once
// over the whole batch.  The sweep runs one launch per tile diagonal D
// covering diagonal D of every problem (a (problem, tile) table per launch
// and a descriptor per problem), with one tile shape for the batch picked by
// an autotuner that times candidate shapes on the batch itself.
//
// Each problem keeps its own iteration control (snap.cc:468-540) through
// device predicates: active[p] (not yet converged) and done[p] (its inner
// iteration converged in this outer).  Workgroups of a finished problem exit
// at once, the elementwise steps leave its fluxes unchanged, and its
// counters stop.  Because every decision is on the device, one whole outer
// iteration -- outer source, iitm predicated inner iterations, outer test --
// is captured once in a hipGraph and replayed; the host only reads the flags
// between outer iterations.
//
// Mixed precision follows the single-problem solve's FP64 polish: prec[p]
// says whether problem p's next inner sweep runs in FP64 (its previous inner
// of the same outer changed its flux by at most fp64_polish * epsi).  Each
// sweep is then two passes over the diagonals, an FP32 kernel whose
// workgroups skip problems with prec[p] = 1 and an FP64 kernel that skips
// the others.  Completion times are read from the device's constant-rate
// clock when a problem's outer test passes.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "snap_internal.h"

namespace snap {

namespace {

struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = n ? n : 16;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      return fail(SNAP_ERR_NO_MEMORY, "device allocation of %zu bytes failed", bytes);
    }
    return SNAP_OK;
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fill(double *__restrict__ dst, double v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}

__global__ void k_fill_ints(int *__restrict__ dst, int v, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = v;
}

// Per-problem convergence (inner.cc:283-324 / outer.cc:354-403 test body):
// conv[p] stays 1 unless some group and cell of problem p fails
// |flux0/flux0p - 1| <= eps (|flux0| when |flux0p| < 1e-12).  Block
// (x, p) covers a slice of problem p's cells x groups.
// With `change`, also change[p] = max |flux0/flux0p - 1| over the problem
// as the bit pattern of the non-negative double (NaN above every number).
__global__ void k_problem_test(const double *__restrict__ f, const double *__restrict__ fp,
                               const int64_t *__restrict__ cell_off, int64_t total, int ng,
                               double eps, int *__restrict__ conv,
                               unsigned long long *__restrict__ change) {
  const int p = blockIdx.y;
  const int64_t c0 = cell_off[p], nc = cell_off[p + 1] - c0;
  const int64_t n = nc * ng;
  bool ok = true;
  double m = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(t / nc);
    const int64_t i = (int64_t)g * total + c0 + (t - (int64_t)g * nc);
    double a = fp[i], df = 1.0;
    if (fabs(a) < 1.0e-12) {
      a = 1.0;
      df = 0.0;
    }
    df = fabs(f[i] / a - df);
    if (df > eps) ok = false;  // NaN compares false: counts as converged, as in the reference
    m = (df > m || df != df) ? df : m;
  }
  const int all_ok = __syncthreads_and(ok ? 1 : 0);
  if (threadIdx.x == 0 && !all_ok) atomicAnd(&conv[p], 0);
  if (change) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double v = __shfl_xor(m, o, 64);
      m = (v > m || v != v) ? v : m;
    }
    if ((threadIdx.x & 63) == 0)
      atomicMax(&change[p], (unsigned long long)__double_as_longlong(fabs(m)));
  }
}

__global__ void k_clear_test(int *__restrict__ conv, unsigned long long *__restrict__ change,
                             int np) {
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    conv[p] = 1;
    if (change) change[p] = 0ull;
  }
}

// the device's constant-rate clock into t[i]
__global__ void k_stamp(unsigned long long *t, int i) {
  if (threadIdx.x == 0) t[i] = wall_clock64();
}

// Flag updates between the steps of the predicated outer iteration.  An
// outer's first inner sweep is FP32 (prec[p] = 0).
__global__ void k_begin_outer(const int *__restrict__ active, int *__restrict__ done,
                              int *__restrict__ outers, int *__restrict__ prec, int np) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np && active[p]) {
    done[p] = 0;
    outers[p] += 1;
    if (prec) prec[p] = 0;
  }
}

// After an inner test: count the inner (and its FP64 sweep), and pick the
// precision of the problem's next inner sweep from the change just measured.
__global__ void k_after_inner(const int *__restrict__ active, int *__restrict__ done,
                              const int *__restrict__ conv, int *__restrict__ inners,
                              int *__restrict__ prec, const unsigned long long *__restrict__ change,
                              double polish_at, int *__restrict__ n64, int np) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np && active[p] && !done[p]) {
    inners[p] += 1;
    if (prec) {
      n64[p] += prec[p];
      prec[p] = __longlong_as_double((long long)change[p]) <= polish_at ? 1 : 0;
    } else if (n64) {
      n64[p] += 1;  // an FP64 batch
    }
    done[p] = conv[p];
  }
}

// snap.cc:519-538: no outer test on the first outer iteration; it fails
// unless the last inner iteration converged
__global__ void k_end_outer(int *__restrict__ active, const int *__restrict__ done,
                            const int *__restrict__ conv, const int *__restrict__ otno,
                            int *__restrict__ converged, unsigned long long *__restrict__ t_end,
                            int np) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np && active[p] && *otno > 0 && done[p] && conv[p]) {
    active[p] = 0;
    converged[p] = 1;
    t_end[p] = wall_clock64();
  }
}

int64_t blocks_for(int64_t n) { return std::min<int64_t>((n + 255) / 256, 65536); }

// Tile shapes the autotuner picked, per batch signature (precision, ng,
// nang, nmom, every problem's extent), for the life of the process.
struct TunedShape {
  int TX, TY, TZ, waves, apl;
  bool seq;  // the one-wavefront tile sweep (4 x 4 x TZ)
};
std::mutex g_tuned_lock;
std::map<std::string, TunedShape> g_tuned;

}  // namespace

}  // namespace snap

using namespace snap;

struct snap_batch {
  int count = 0, ng = 0, nang = 0, nmom = 0, cmom = 0, nmat = 0;  // nmat: all problems' materials
  double epsi = 0;
  int iitm = 0, oitm = 0, fixup = 0, precision = 0;
  bool polish = false;  // MIXED with FP64 polish sweeps
  snap_batch_options_t opt{};
  int64_t total = 0;
  std::vector<int64_t> cell_off;  // host [count + 1]
  std::vector<Extent3> ext;
  std::vector<double> hi, hj, hk;
  QuadDev quad;
  SweepPlan plan;        // tile shape and lane layout (its extent: the largest problem)
  BatchSweep bs;
  int groups_per_wave = 1;  // of the last sweep pass enqueued
  hipStream_t stream = nullptr;
  // the one-wavefront sweep's second stream (half of the corners) and its
  // fork / join events; a captured outer iteration holds both branches
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  Buf mat, slgg, qi, txs, txs32, sxs, dinv, dinv32, flux0, flux0pi, flux0po, fluxm, q2grp0,
      q2grpm, qtot, part_flux, part_fluxm, faces, descs, tiles, d_cell_off;
  Buf active, done, conv, inners, outers, converged, otno;
  Buf prec, change, n64, stamps;  // stamps: [count] ends, then the run's start and end
  double autotune_ms = 0;
  bool plan_cached = false;

  ~snap_batch() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    plan_free(&plan);
    quad_free(&quad);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  Acc3 acc(const Buf &b, int rec, int elem) const {
    return dense_acc3(b.p, Extent3{(int)total, 1, 1}, rec, elem);
  }
  // problem p's cells of a [g][all cells][rec] array as a 3-D accessor
  Acc3 sub(const Buf &b, int p, int rec, int elem) const {
    Acc3 a;
    a.base = (char *)b.p + (size_t)cell_off[p] * rec * elem;
    a.s0 = (int64_t)rec * elem;
    a.s1 = a.s0 * ext[p].nx;
    a.s2 = a.s1 * ext[p].ny;
    a.sg = (int64_t)total * rec * elem;
    return a;
  }

  // tile shape (TX, TY, TZ, waves, apl), or the one-wavefront sweep's
  // 4 x 4 x TZ tiles (seq) -> plan, per-problem descriptors, (problem, tile)
  // table by diagonal, face buffer
  int build_plan(int TX, int TY, int TZ, int waves, int apl, bool seq = false) {
    Extent3 big{1, 1, 1};
    for (auto &e : ext) {
      big.nx = std::max(big.nx, e.nx);
      big.ny = std::max(big.ny, e.ny);
      big.nz = std::max(big.nz, e.nz);
    }
    if (seq) {
      SNAP_TRY(plan_create_seq(&plan, big, nang, cmom, TZ));
      TX = plan.TX;
      TY = plan.TY;
      TZ = plan.TZ;
    } else {
      SNAP_TRY(plan_create(&plan, big, nang, cmom, TX, TY, TZ, waves, apl));
    }
    const int per_tile = (TY * TZ + TX * TZ + TX * TY) * nang;
    std::vector<BatchDesc> d(count);
    int64_t face_off = 0;
    int maxd = 0;
    for (int p = 0; p < count; p++) {
      BatchDesc &b = d[p];
      b.ext = ext[p];
      b.nTx = (ext[p].nx + TX - 1) / TX;
      b.nTy = (ext[p].ny + TY - 1) / TY;
      b.nTz = (ext[p].nz + TZ - 1) / TZ;
      b.ntiles = b.nTx * b.nTy * b.nTz;
      b.cell_off = cell_off[p];
      b.face_off = face_off;
      b.hi = hi[p];
      b.hj = hj[p];
      b.hk = hk[p];
      // the one-wavefront sweep keeps its faces in place, one slot per tile
      // row, column and pencil
      face_off += seq ? (int64_t)8 * ng * nang *
                            ((int64_t)b.nTy * b.nTz * TY * TZ + (int64_t)b.nTx * b.nTz * TX * TZ +
                             (int64_t)b.nTx * b.nTy * TX * TY)
                      : (int64_t)8 * ng * b.ntiles * per_tile;
      maxd = std::max(maxd, b.nTx + b.nTy + b.nTz - 2);
    }
    std::vector<int2> tab;
    bs.plane_off.assign(maxd + 1, 0);
    for (int D = 0; D < maxd; D++) {
      bs.plane_off[D] = (int)tab.size();
      for (int p = 0; p < count; p++)
        for (int tz = 0; tz < d[p].nTz; tz++)
          for (int ty = 0; ty < d[p].nTy; ty++) {
            const int tx = D - ty - tz;
            if (tx >= 0 && tx < d[p].nTx) tab.push_back(make_int2(p, tx | (ty << 10) | (tz << 20)));
          }
    }
    bs.plane_off[maxd] = (int)tab.size();
    bs.nplanes = maxd;
    SNAP_TRY(descs.alloc(sizeof(BatchDesc) * count));
    SNAP_TRY(tiles.alloc(sizeof(int2) * tab.size()));
    SNAP_HIP_TRY(hipMemcpy(descs.p, d.data(), sizeof(BatchDesc) * count, hipMemcpyHostToDevice));
    SNAP_HIP_TRY(hipMemcpy(tiles.p, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice));
    const size_t face_bytes = (size_t)face_off * sizeof(double);
    // (a much larger buffer left by an autotuner candidate is given back)
    if (faces.bytes < face_bytes || faces.bytes > 2 * face_bytes) {
      SNAP_HIP_TRY(hipStreamSynchronize(stream));
      SNAP_TRY(faces.alloc(face_bytes));
    }
    bs.batch_tiles = tiles.as<int2>();
    bs.descs = descs.as<BatchDesc>();
    bs.active = active.as<int>();
    bs.done = done.as<int>();
    bs.total_cells = total;
    bs.full_xy = std::all_of(ext.begin(), ext.end(),
                             [](const Extent3 &e) { return e.nx % 4 == 0 && e.ny % 4 == 0; });
    if (graph_exec) {  // a captured outer iteration holds the old launches
      (void)hipGraphExecDestroy(graph_exec);
      (void)hipGraphDestroy(graph);
      graph_exec = nullptr;
      graph = nullptr;
    }
    return SNAP_OK;
  }

  // one sweep pass over every diagonal in one precision; with prec, only the
  // problems whose prec[p] equals want
  int sweep_pass(bool fp32, const int *prec, int want) {
    BatchSweep b = bs;
    b.prec = prec;
    b.prec_want = want;
    const bool seq = plan.seq;
    SweepLaunch L;
    memset(&L, 0, sizeof(L));
    L.plan = &plan;
    L.quad = &quad;
    L.nsweeps = 8 * ng;
    L.ng_job = ng;
    L.fixup = fixup;
    L.q = qtot.as<double>();
    L.dinv_dense = fp32 ? dinv32.p : dinv.p;
    L.txs_dense = fp32 ? txs32.p : txs.p;
    L.part_flux = part_flux.as<double>();
    L.part_fluxm = cmom > 1 ? part_fluxm.as<double>() : nullptr;
    L.faces = faces.p;
    L.fp32 = fp32;
    L.batch = &b;
    if (seq) {
      L.recompute_dinv = true;  // dinv formed from t_xs in the kernel
      L.side = side;
      L.fork = fork_ev;
      L.join = join_ev;
      L.pair = opt.groups_per_wave != 1;
      groups_per_wave = sweep_seq_pair(L) ? 2 : 1;
      return sweep_seq_enqueue(L, stream, nullptr);
    }
    groups_per_wave = 1;
    return sweep_enqueue(L, stream);
  }

  int sweep() {
    if (!polish) return sweep_pass(precision == SNAP_PRECISION_MIXED, nullptr, 0);
    SNAP_TRY(sweep_pass(true, prec.as<int>(), 0));
    return sweep_pass(false, prec.as<int>(), 1);
  }

  int copy(Buf &dst, const Buf &src, int64_t n) {
    hipLaunchKernelGGL(k_copy, dim3(blocks_for(n)), dim3(256), 0, stream, dst.as<double>(),
                       src.as<double>(), n);
    SNAP_HIP_TRY(hipGetLastError());
    return SNAP_OK;
  }

  // conv[p] (and change[p] when asked) of the convergence test against prev
  int test(const Buf &prev, double eps, bool with_change) {
    unsigned long long *ch = with_change ? change.as<unsigned long long>() : nullptr;
    hipLaunchKernelGGL(k_clear_test, dim3(1), dim3(256), 0, stream, conv.as<int>(), ch, count);
    int64_t maxn = 0;
    for (int p = 0; p < count; p++) maxn = std::max(maxn, (cell_off[p + 1] - cell_off[p]) * ng);
    const int bx = (int)std::min<int64_t>((maxn + 255) / 256, 1024);
    hipLaunchKernelGGL(k_problem_test, dim3(bx, count), dim3(256), 0, stream, flux0.as<double>(),
                       prev.as<double>(), d_cell_off.as<int64_t>(), total, ng, eps, conv.as<int>(),
                       ch);
    SNAP_HIP_TRY(hipGetLastError());
    return SNAP_OK;
  }

  // one outer iteration, every step predicated on the device flags
  int outer_iteration() {
    const int gb = (count + 255) / 256;
    const int64_t gc = (int64_t)ng * total;
    const int nm1 = cmom - 1;
    int *pr = polish ? prec.as<int>() : nullptr;
    int *n64p = (polish || precision == SNAP_PRECISION_FP64) ? n64.as<int>() : nullptr;
    hipLaunchKernelGGL(k_begin_outer, dim3(gb), dim3(256), 0, stream, active.as<int>(),
                       done.as<int>(), outers.as<int>(), pr, count);
    SNAP_TRY(launch_outer_dense(qi.as<double>(), flux0.as<double>(),
                                cmom > 1 ? fluxm.as<double>() : nullptr, slgg.as<double>(),
                                mat.as<int32_t>(), q2grp0.as<double>(),
                                cmom > 1 ? q2grpm.as<double>() : nullptr, total, ng, nmat, nmom,
                                false, stream));
    SNAP_TRY(copy(flux0po, flux0, gc));
    for (int inno = 0; inno < iitm; inno++) {
      SNAP_TRY(launch_inner_source_dense(sxs.as<double>(), flux0.as<double>(), q2grp0.as<double>(),
                                         cmom > 1 ? fluxm.as<double>() : nullptr,
                                         cmom > 1 ? q2grpm.as<double>() : nullptr, qtot.as<double>(),
                                         gc, nmom, stream));
      SNAP_TRY(copy(flux0pi, flux0, gc));
      SNAP_TRY(sweep());
      SNAP_TRY(launch_octant_reduce(part_flux.as<double>(), flux0.as<double>(), gc, stream));
      if (cmom > 1)
        SNAP_TRY(launch_octant_reduce(part_fluxm.as<double>(), fluxm.as<double>(), gc * nm1, stream));
      SNAP_TRY(test(flux0pi, epsi, polish));
      hipLaunchKernelGGL(k_after_inner, dim3(gb), dim3(256), 0, stream, active.as<int>(),
                         done.as<int>(), conv.as<int>(), inners.as<int>(), pr,
                         change.as<unsigned long long>(), polish_threshold(opt.fp64_polish, epsi),
                         n64p, count);
    }
    SNAP_TRY(test(flux0po, 100.0 * epsi, false));
    hipLaunchKernelGGL(k_end_outer, dim3(gb), dim3(256), 0, stream, active.as<int>(),
                       done.as<int>(), conv.as<int>(), otno.as<int>(), converged.as<int>(),
                       stamps.as<unsigned long long>(), count);
    SNAP_HIP_TRY(hipGetLastError());
    return SNAP_OK;
  }

  // the initial state of a run: zero fluxes, every problem active
  int reset() {
    const int64_t gc = (int64_t)ng * total;
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(gc)), dim3(256), 0, stream, flux0.as<double>(), 0.0, gc);
    if (cmom > 1)
      hipLaunchKernelGGL(k_fill, dim3(blocks_for(gc * (cmom - 1))), dim3(256), 0, stream,
                         fluxm.as<double>(), 0.0, gc * (cmom - 1));
    hipLaunchKernelGGL(k_fill_ints, dim3(1), dim3(256), 0, stream, active.as<int>(), 1, count);
    for (Buf *b : {&done, &inners, &outers, &converged, &prec, &n64})
      hipLaunchKernelGGL(k_fill_ints, dim3(1), dim3(256), 0, stream, b->as<int>(), 0, count);
    SNAP_HIP_TRY(hipGetLastError());
    return SNAP_OK;
  }

  // time `reps` sweeps of the current plan with every problem active
  // (warm: one untimed sweep first, for a kernel's first launches, which load
  // its code object)
  int time_plan(int reps, float *ms, bool warm) {
    SNAP_TRY(reset());
    hipEvent_t e0, e1;
    SNAP_HIP_TRY(hipEventCreate(&e0));
    SNAP_HIP_TRY(hipEventCreate(&e1));
    if (warm) SNAP_TRY(sweep());
    SNAP_HIP_TRY(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; r++) SNAP_TRY(sweep());
    SNAP_HIP_TRY(hipEventRecord(e1, stream));
    SNAP_HIP_TRY(hipEventSynchronize(e1));
    SNAP_HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    *ms /= reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return SNAP_OK;
  }
};

extern "C" int snap_hip_batch_create(const snap_problem_t *probs, int count,
                                     const snap_batch_options_t *o, snap_batch_t **out) {
  if (!probs || !out || count < 1) return fail(SNAP_ERR_INVALID, "empty batch");
  *out = nullptr;
  std::unique_ptr<snap_batch> b(new snap_batch());
  if (o) b->opt = *o;
  const snap_problem_t &p0 = probs[0];
  b->count = count;
  b->ng = p0.ng;
  b->nang = p0.nang;
  b->nmom = p0.nmom;
  b->cmom = p0.nmom * p0.nmom;
  b->epsi = p0.epsi;
  b->iitm = p0.iitm;
  b->oitm = p0.oitm;
  b->fixup = p0.fixup ? 1 : 0;
  b->precision = b->opt.precision;
  if (b->precision != SNAP_PRECISION_FP64 && b->precision != SNAP_PRECISION_MIXED)
    return fail(SNAP_ERR_INVALID, "batch precision must be FP64 or MIXED");
  if (!(b->opt.fp64_polish >= 0.0)) return fail(SNAP_ERR_INVALID, "fp64_polish must be >= 0");
  if (b->opt.sweep_kernel < SNAP_SWEEP_AUTO || b->opt.sweep_kernel > SNAP_SWEEP_WAVE_TILE)
    return fail(SNAP_ERR_INVALID, "unknown sweep kernel %d", b->opt.sweep_kernel);
  if (b->opt.groups_per_wave != 0 && b->opt.groups_per_wave != 1)
    return fail(SNAP_ERR_INVALID, "groups_per_wave must be 0 (choose) or 1, got %d",
                b->opt.groups_per_wave);
  if (b->opt.sweep_kernel == SNAP_SWEEP_WAVE_TILE && p0.nang > 64)
    return fail(SNAP_ERR_UNSUPPORTED, "the one-wavefront tile sweep handles up to 64 angles");
  b->polish = b->precision == SNAP_PRECISION_MIXED && b->opt.fp64_polish > 0.0;
  if (b->nmom < 1 || b->nmom > 4 || b->ng < 1 || b->nang < 1)
    return fail(SNAP_ERR_INVALID, "bad ng / nang / nmom");
  if (b->nang > 128)
    return fail(SNAP_ERR_UNSUPPORTED,
                "batches use the tiled sweep: at most 128 angles per octant, got %d", b->nang);
  b->cell_off.assign(count + 1, 0);
  int nmat_all = 0;
  for (int p = 0; p < count; p++) {
    const snap_problem_t &q = probs[p];
    if (q.ng != b->ng || q.nang != b->nang || q.nmom != b->nmom || q.epsi != b->epsi ||
        q.iitm != b->iitm || q.oitm != b->oitm || (q.fixup != 0) != (b->fixup != 0))
      return fail(SNAP_ERR_INVALID, "problem %d differs in ng, nang, nmom or iteration controls", p);
    if (q.nx < 1 || q.ny < 1 || q.nz < 1 || q.nmat < 1 || !q.mat || !q.sigt || !q.slgg || !q.qi)
      return fail(SNAP_ERR_INVALID, "problem %d is incomplete", p);
    if (!(q.lx > 0 && q.ly > 0 && q.lz > 0)) return fail(SNAP_ERR_INVALID, "problem %d: bad lengths", p);
    if (q.timedep || q.mms)
      return fail(SNAP_ERR_UNSUPPORTED, "problem %d: batches are of steady fixed-source problems", p);
    b->ext.push_back(Extent3{q.nx, q.ny, q.nz});
    b->hi.push_back(2.0 / (q.lx / double(q.nx)));
    b->hj.push_back(2.0 / (q.ly / double(q.ny)));
    b->hk.push_back(2.0 / (q.lz / double(q.nz)));
    b->cell_off[p + 1] = b->cell_off[p] + (int64_t)q.nx * q.ny * q.nz;
    nmat_all += q.nmat;
  }
  b->total = b->cell_off[count];
  if (b->total >= (int64_t)1 << 31) return fail(SNAP_ERR_UNSUPPORTED, "batches of 2^31 cells or more");
  b->nmat = nmat_all;
  const int ng = b->ng, nmom = b->nmom, cmom = b->cmom, nang = b->nang;
  const int64_t T = b->total;
  const size_t gc = (size_t)ng * T, nm1 = cmom > 1 ? (size_t)(cmom - 1) : 0;

  // ---- host assembly of the concatenated inputs: material ids offset per
  // problem, tables stacked along the material axis, t_xs per cell
  std::vector<int32_t> mat(T);
  std::vector<double> slgg((size_t)ng * nmat_all * ng * nmom), qi(gc), txs(gc);
  int moff = 0;
  for (int p = 0; p < count; p++) {
    const snap_problem_t &q = probs[p];
    const int64_t nc = b->cell_off[p + 1] - b->cell_off[p], c0 = b->cell_off[p];
    for (int64_t c = 0; c < nc; c++) {
      if (q.mat[c] < 0 || q.mat[c] >= q.nmat)
        return fail(SNAP_ERR_INVALID, "problem %d: material index out of range", p);
      mat[c0 + c] = q.mat[c] + moff;
    }
    for (int g1 = 0; g1 < ng; g1++)
      for (int m = 0; m < q.nmat; m++)
        memcpy(&slgg[(((size_t)g1 * nmat_all + moff + m) * ng) * nmom],
               &q.slgg[(((size_t)g1 * q.nmat + m) * ng) * nmom], sizeof(double) * ng * nmom);
    for (int g = 0; g < ng; g++)
      for (int64_t c = 0; c < nc; c++) {
        qi[(size_t)g * T + c0 + c] = q.qi[(size_t)g * nc + c];
        txs[(size_t)g * T + c0 + c] =
            q.sigt_cell ? q.sigt_cell[(size_t)g * nc + c] : q.sigt[(size_t)q.mat[c] * ng + g];
      }
    moff += q.nmat;
  }

  std::vector<double> mu, eta, xi, w, ec;
  SNAP_TRY(quad_compute(nang, nmom, mu, eta, xi, w, ec));
  SNAP_TRY(quad_upload(&b->quad, ec.data(), mu.data(), eta.data(), xi.data(), w.data(), nang, nmom));
  SNAP_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  SNAP_HIP_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  SNAP_HIP_TRY(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
  SNAP_HIP_TRY(hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming));
  const hipStream_t st = b->stream;
  auto up = [&](Buf &d, const void *h, size_t n) -> int {
    SNAP_TRY(d.alloc(n));
    SNAP_HIP_TRY(hipMemcpy(d.p, h, n, hipMemcpyHostToDevice));
    return SNAP_OK;
  };
  SNAP_TRY(up(b->mat, mat.data(), sizeof(int32_t) * T));
  SNAP_TRY(up(b->slgg, slgg.data(), sizeof(double) * slgg.size()));
  SNAP_TRY(up(b->qi, qi.data(), sizeof(double) * gc));
  SNAP_TRY(up(b->txs, txs.data(), sizeof(double) * gc));
  SNAP_TRY(up(b->d_cell_off, b->cell_off.data(), sizeof(int64_t) * (count + 1)));
  SNAP_TRY(b->sxs.alloc(sizeof(double) * gc * nmom));
  SNAP_TRY(b->dinv.alloc(sizeof(double) * gc * nang));
  for (Buf *x : {&b->flux0, &b->flux0pi, &b->flux0po, &b->q2grp0}) SNAP_TRY(x->alloc(sizeof(double) * gc));
  SNAP_TRY(b->qtot.alloc(sizeof(double) * gc * cmom));
  SNAP_TRY(b->fluxm.alloc(sizeof(double) * gc * nm1));
  SNAP_TRY(b->q2grpm.alloc(sizeof(double) * gc * nm1));
  SNAP_TRY(b->part_flux.alloc(sizeof(double) * 8 * gc));
  SNAP_TRY(b->part_fluxm.alloc(sizeof(double) * 8 * gc * nm1));
  for (Buf *x : {&b->active, &b->done, &b->conv, &b->inners, &b->outers, &b->converged, &b->prec,
                 &b->n64})
    SNAP_TRY(x->alloc(sizeof(int) * count));
  SNAP_TRY(b->change.alloc(sizeof(unsigned long long) * count));
  SNAP_TRY(b->stamps.alloc(sizeof(unsigned long long) * (count + 2)));
  SNAP_TRY(b->otno.alloc(sizeof(int)));

  // ExpandScatteringCrossSection over the batch, geometry per problem
  {
    Acc2 sl;
    sl.base = b->slgg.as<char>();
    sl.s1 = (int64_t)nmom * 8;
    sl.s0 = sl.s1 * ng;
    sl.sg = sl.s0 * nmat_all;
    SNAP_TRY(launch_expand_scat(sl, 0, b->acc(b->mat, 1, 4), b->acc(b->sxs, nmom, 8),
                                Extent3{(int)T, 1, 1}, ng, nmom, st));
  }
  const bool fp32 = b->precision == SNAP_PRECISION_MIXED;
  if (fp32) {
    SNAP_TRY(b->dinv32.alloc(sizeof(float) * gc * nang));
    SNAP_TRY(b->txs32.alloc(sizeof(float) * gc));
    SNAP_TRY(launch_to_float(b->txs.as<double>(), b->txs32.as<float>(), (int64_t)gc, st));
  }
  for (int p = 0; p < count; p++) {
    SNAP_TRY(launch_geometry(b->sub(b->txs, p, 1, 8), b->sub(b->dinv, p, nang, 8), nullptr,
                             b->ext[p], ng, nang, b->hi[p], b->hj[p], b->hk[p], b->quad, false, st));
    if (fp32)
      SNAP_TRY(launch_geometry(b->sub(b->txs, p, 1, 8), b->sub(b->dinv32, p, nang, 4), nullptr,
                               b->ext[p], ng, nang, b->hi[p], b->hj[p], b->hk[p], b->quad, true, st));
  }
  // the first inner's source of every problem (flux zero): qtot[0] = qi
  {
    const int lma[4] = {1, 3, 5, 7};
    Acc3 none{};
    SNAP_TRY(b->reset());
    SNAP_TRY(launch_outer_dense(b->qi.as<double>(), b->flux0.as<double>(),
                                cmom > 1 ? b->fluxm.as<double>() : nullptr, b->slgg.as<double>(),
                                b->mat.as<int32_t>(), b->q2grp0.as<double>(),
                                cmom > 1 ? b->q2grpm.as<double>() : nullptr, T, ng, nmat_all, nmom,
                                false, st));
    SNAP_TRY(launch_inner_source(b->acc(b->sxs, nmom, 8), b->acc(b->flux0, 1, 8),
                                 b->acc(b->q2grp0, 1, 8), cmom > 1 ? b->acc(b->fluxm, cmom - 1, 8) : none,
                                 cmom > 1 ? b->acc(b->q2grpm, cmom - 1, 8) : none,
                                 b->acc(b->qtot, cmom, 8), Extent3{(int)T, 1, 1}, ng, cmom, lma,
                                 nmom, st));
  }

  // ---- tile shape: the caller's, an earlier batch's of the same shapes, or
  // the fastest candidate on this batch
  int TX = b->opt.tile[0], TY = b->opt.tile[1], TZ = b->opt.tile[2];
  int waves = b->opt.waves_per_tile, apl = b->opt.angles_per_lane;
  const int kern = b->opt.sweep_kernel;
  std::string sig = std::to_string(b->precision) + ":" + std::to_string(kern) + ":" +
                    std::to_string(ng) + ":" +
                    std::to_string(nang) + ":" + std::to_string(nmom) + ":" +
                    std::to_string(waves) + ":" + std::to_string(apl);
  for (const Extent3 &e : b->ext)
    sig += ":" + std::to_string(e.nx) + "x" + std::to_string(e.ny) + "x" + std::to_string(e.nz);
  bool cached = false;
  TunedShape hit{};
  const bool fixed_seq = kern == SNAP_SWEEP_WAVE_TILE && TZ > 0;
  if (!(TX > 0 && TY > 0 && TZ > 0) && !fixed_seq) {
    std::lock_guard<std::mutex> g(g_tuned_lock);
    auto it = g_tuned.find(sig);
    if (it != g_tuned.end()) {
      cached = true;
      hit = it->second;
    }
  }
  if (fixed_seq) {
    SNAP_TRY(b->build_plan(4, 4, TZ, 1, 1, true));
  } else if (TX > 0 && TY > 0 && TZ > 0 && kern != SNAP_SWEEP_WAVE_TILE) {
    SNAP_TRY(b->build_plan(TX, TY, TZ, waves, apl));
  } else if (cached) {
    SNAP_TRY(b->build_plan(hit.TX, hit.TY, hit.TZ, hit.waves, hit.apl, hit.seq));
    b->plan_cached = true;
  } else {
    const auto t0 = std::chrono::steady_clock::now();
    static const int cand[][3] = {{8, 8, 4}, {8, 4, 4}, {4, 4, 4}, {8, 8, 8}, {16, 8, 4}};
    float best = 1e30f;
    int bi = -1, bapl = 0;
    bool warmed[3] = {false, false, false};  // per lane layout (one kernel each)
    // the one-wavefront tile sweep first (steady problems of up to 64 angles,
    // any options that only the multi-wave tiles have left unset); when it
    // takes the batch, the multi-wave shapes are not timed: it beat them on
    // every batch and problem measured from 8 to 64 angles, and timing them
    // cost more than the solve (3.6 s FP64, 7.2 s mixed on the 48-problem
    // workload)
    int best_seq_tz = 0;
    if (kern != SNAP_SWEEP_TILED && nang <= 64 && apl <= 1 && waves <= 1) {
      bool warm = true;
      for (int tz : {4, 8, 16}) {
        if (b->build_plan(4, 4, tz, 1, 1, true) != SNAP_OK) continue;
        float ms = 0.f;
        SNAP_TRY(b->time_plan(1, &ms, warm));
        warm = false;
        if (ms < best) {
          best = ms;
          best_seq_tz = tz;
        }
      }
    }
    const int ncand = (kern == SNAP_SWEEP_WAVE_TILE || best_seq_tz > 0)
                          ? 0
                          : (int)(sizeof(cand) / sizeof(cand[0]));
    for (int c = 0; c < ncand; c++) {
      for (int a = 1; a <= 2; a++) {
        if (apl > 0 && a != apl) continue;
        if (a == 2 && !fp32 && nang <= 64) continue;  // FP64 two-angle lanes halve occupancy
        int w = waves;
        const int rc = b->build_plan(cand[c][0], cand[c][1], cand[c][2], w, a);
        if (rc != SNAP_OK) continue;  // shape does not fit (LDS, waves)
        float ms = 0.f;
        SNAP_TRY(b->time_plan(1, &ms, !warmed[a]));
        warmed[a] = true;
        if (ms < best) {
          best = ms;
          bi = c;
          bapl = a;
        }
      }
    }
    if (bi >= 0) {  // a multi-wave shape beat the one-wavefront sweep
      best_seq_tz = 0;
    } else if (best_seq_tz == 0) {
      return fail(SNAP_ERR_INVALID, "no candidate tile fits this batch");
    }
    TunedShape t;
    if (best_seq_tz)
      t = TunedShape{4, 4, best_seq_tz, 1, 1, true};
    else
      t = TunedShape{cand[bi][0], cand[bi][1], cand[bi][2], waves, bapl, false};
    SNAP_TRY(b->build_plan(t.TX, t.TY, t.TZ, t.waves, t.apl, t.seq));
    b->autotune_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::lock_guard<std::mutex> g(g_tuned_lock);
    g_tuned[sig] = t;
  }
  SNAP_HIP_TRY(hipStreamSynchronize(st));
  *out = b.release();
  return SNAP_OK;
}

extern "C" int snap_hip_batch_run(snap_batch_t *b, snap_batch_stats_t *stats,
                                  snap_problem_stats_t *per_problem) {
  if (!b) return fail(SNAP_ERR_INVALID, "NULL batch");
  const hipStream_t st = b->stream;
  const int n = b->count;
  std::vector<int> active(n, 1), inners(n), outers(n), conv(n), n64(n);
  std::vector<unsigned long long> stamps(n + 2);
  int dev = 0, khz = 0;
  SNAP_HIP_TRY(hipGetDevice(&dev));
  SNAP_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) return fail(SNAP_ERR_INVALID, "the device reports no wall-clock rate");
  SNAP_HIP_TRY(hipStreamSynchronize(st));
  const auto t0 = std::chrono::steady_clock::now();
  SNAP_TRY(b->reset());
  unsigned long long *sp = b->stamps.as<unsigned long long>();
  hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, sp, n);
  int otno = 0;
  for (; otno < b->oitm; otno++) {
    SNAP_HIP_TRY(hipMemcpyAsync(b->otno.p, &otno, sizeof(int), hipMemcpyHostToDevice, st));
    if (b->opt.use_graph) {
      if (!b->graph_exec) {
        SNAP_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = b->outer_iteration();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc != SNAP_OK) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        SNAP_HIP_TRY(e);
        b->graph = g;
        SNAP_HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
      }
      SNAP_HIP_TRY(hipGraphLaunch(b->graph_exec, st));
    } else {
      SNAP_TRY(b->outer_iteration());
    }
    SNAP_HIP_TRY(hipMemcpyAsync(active.data(), b->active.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    SNAP_HIP_TRY(hipStreamSynchronize(st));
    int left = 0;
    for (int p = 0; p < n; p++) left += active[p];
    if (!left) {
      otno++;
      break;
    }
  }
  hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, sp, n + 1);
  SNAP_HIP_TRY(hipStreamSynchronize(st));
  const double solve_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  SNAP_HIP_TRY(hipMemcpy(inners.data(), b->inners.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  SNAP_HIP_TRY(hipMemcpy(outers.data(), b->outers.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  SNAP_HIP_TRY(hipMemcpy(conv.data(), b->converged.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  SNAP_HIP_TRY(hipMemcpy(n64.data(), b->n64.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  SNAP_HIP_TRY(hipMemcpy(stamps.data(), sp, sizeof(unsigned long long) * (n + 2),
                         hipMemcpyDeviceToHost));
  const double tick_ms = 1.0 / (double)khz;  // khz ticks per ms
  double units = 0.0;
  int nconv = 0;
  for (int p = 0; p < n; p++) {
    const double cells = (double)(b->cell_off[p + 1] - b->cell_off[p]);
    units += cells * 8.0 * b->nang * b->ng * inners[p];
    nconv += conv[p];
    if (per_problem) {
      per_problem[p].inner_total = inners[p];
      per_problem[p].outer_total = outers[p];
      per_problem[p].converged = conv[p];
      per_problem[p].sweeps_fp64 = n64[p];
      per_problem[p].time_ms = (double)((conv[p] ? stamps[p] : stamps[n + 1]) - stamps[n]) * tick_ms;
    }
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->outer_iterations = otno;
    stats->converged = nconv;
    stats->solve_ms = solve_ms;
    stats->units = units;
    stats->grind_ns = units > 0 ? solve_ms * 1e6 / units : 0.0;
    stats->tile[0] = b->plan.TX;
    stats->tile[1] = b->plan.TY;
    stats->tile[2] = b->plan.TZ;
    stats->waves = b->plan.waves;
    stats->angles_per_lane = b->plan.apl;
    stats->groups_per_wave = b->groups_per_wave;
    stats->autotune_ms = b->autotune_ms;
    stats->plan_cached = b->plan_cached ? 1 : 0;
  }
  return SNAP_OK;
}

extern "C" int snap_hip_batch_get_flux(snap_batch_t *b, int problem, double *flux0, double *fluxm) {
  if (!b || problem < 0 || problem >= b->count) return fail(SNAP_ERR_INVALID, "bad problem index");
  SNAP_HIP_TRY(hipStreamSynchronize(b->stream));
  const int64_t c0 = b->cell_off[problem], nc = b->cell_off[problem + 1] - c0;
  const int nm1 = b->cmom - 1;
  for (int g = 0; g < b->ng; g++) {
    if (flux0)
      SNAP_HIP_TRY(hipMemcpy(flux0 + (size_t)g * nc, b->flux0.as<double>() + (size_t)g * b->total + c0,
                             sizeof(double) * nc, hipMemcpyDeviceToHost));
    if (fluxm && nm1 > 0)
      SNAP_HIP_TRY(hipMemcpy(fluxm + (size_t)g * nc * nm1,
                             b->fluxm.as<double>() + ((size_t)g * b->total + c0) * nm1,
                             sizeof(double) * nc * nm1, hipMemcpyDeviceToHost));
  }
  return SNAP_OK;
}

extern "C" void snap_hip_batch_destroy(snap_batch_t *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
}
