// snap_solver.hip -- Legion-free host driver of the steady-state solve.
//
// Reproduces Snap::transport_solve (snap.cc:225-568) for one time step on one
// GPU: expand cross sections and geometry once (snap.cc:419-450), then outer
// iterations of {outer source, save flux0 -> flux0po, up to iitm inner
// iterations of {inner source, save flux0 -> flux0pi, 8 corners x all groups
// of sweeps, inner test}, outer test}.  The reference issues iterations ahead
// under predicates and stops at the first converged test (runahead tunables,
// mapper.cc:109-123; predicate_false_future = true, inner.cc:241,
// outer.cc:312); breaking at the first converged test gives the same
// iteration counts.  The outer test is skipped on the first outer iteration
// and requires the last inner test to have passed (snap.cc:521-525,
// outer.cc:362-365).
//
// All 8*ng sweeps of one inner iteration are one job: each tile-diagonal
// launch covers every (corner, group) pair.  Sweeps write per-octant partial
// fluxes, which are summed in octant order, so results are deterministic.
//
// Time dependence (snap.cc:386-541) repeats the loop once per time step:
// dinv includes vdelt, the sweeps read the previous step's time-edge fluxes
// and write this step's (two arrays that swap roles each step), the flux
// carries over, and a manufactured-solution source is rescaled to the step's
// mid-time before its outer iterations.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "snap_internal.h"

namespace snap {

// The adjoint's tile edges (x, y, z).  On the 48^3, 24-angle, 16-group
// gradient (FP64 / FP32 reverse sweeps, seconds): 8x4x2 0.131 / 0.109
// (three runs against 4x4x4's 0.137 / 0.114 on one box), 4x8x2 0.134 /
// 0.111, 4x4x8 0.138 / 0.110, 4x4x2 0.150 / 0.138, 8x4x4 0.152 / 0.115, 8x8x2
// 0.168 / 0.129, 8x8x4 (the forward solve's) 0.32 / 0.17, 2x2x2 0.30 / 0.28.
constexpr int kAdjTile[3] = {8, 4, 2};

struct DeviceBuffer {
  void *p = nullptr;
  size_t bytes = 0;
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    bytes = n ? n : 16;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      return fail(SNAP_ERR_NO_MEMORY, "device allocation of %zu bytes failed", bytes);
    }
    return SNAP_OK;
  }
  // at least n bytes (the old contents are not kept; the stream's work on
  // them is finished first)
  int grow(size_t n, hipStream_t st) {
    if (p && bytes >= n) return SNAP_OK;
    if (p) {
      SNAP_HIP_TRY(hipStreamSynchronize(st));
      SNAP_HIP_TRY(hipFree(p));
      p = nullptr;
    }
    return alloc(n);
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};

}  // namespace snap

using namespace snap;

struct snap_solver {
  snap_problem_t prob{};
  snap_solve_options_t opt{};
  Extent3 e{};
  int ng = 0, nang = 0, nmom = 0, cmom = 0, nmat = 0;
  int64_t cells = 0;
  double hi = 0, hj = 0, hk = 0;
  QuadDev quad;
  SweepPlan plan;    // the sweep plan of the solver's precision
  SweepPlan plan64;  // MIXED: the plan of its FP64 sweeps
  // one-wavefront tile sweep (steady problems up to 64 angles): used for
  // every inner sweep when present; the adjoint keeps the multi-wave plan
  SweepPlan plan_seq;
  SweepPlan plan_adj;  // the adjoint's tiles (and its checkpoint sweep's), made at the first gradient
  hipStream_t stream = nullptr;
  // the one-wavefront tile sweep's second stream and its fork / join events
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // inner iterations enqueued one ahead (FP64 solves): two slots, each with
  // the test's device flag, its pinned host copy, and events around the
  // sweep and after the copy
  int groups_per_wave = 1;         // of the last sweep enqueued
  DeviceBuffer ra_flag;            // [2] int32
  int32_t *h_ra = nullptr;         // pinned [2]
  hipEvent_t ra_ev[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int32_t *h_flag = nullptr;  // pinned: [0] the test's flag, [2..3] max change bits
  DeviceBuffer mat, sigt, slgg, qi, txs, txs32, sxs, dinv, dinv32;
  DeviceBuffer flux0, flux0pi, flux0po, fluxm, q2grp0, q2grpm, qtot;
  DeviceBuffer part_flux, part_fluxm, faces, flag;
  DeviceBuffer qbase, flux0base, fluxmbase;  // defect correction: the outer's FP64 base
  DeviceBuffer adj_faces, part_grad;          // adjoint: reverse-sweep faces and partials
  bool have_fp32 = false;
  // time dependence and the manufactured solution
  bool td = false;                  // sweeps carry time terms or an angular source
  std::vector<double> vdelt_h;      // [ng]: 2 / (dt v_g), or 0 when steady
  DeviceBuffer vdelt_d, qim;        // [ng]; MMS angular source [8][ng][cells][nang]
  DeviceBuffer tflux[2];            // time-edge fluxes [8][ng][cells][nang], even / odd
  int step = 0;                     // time step in progress: tflux[step & 1] is the inflow
  MmsState mms;
  std::vector<snap_iteration_record_t> log;  // convergence tests of the last run

  ~snap_solver() {
    mms_free(&mms);
    plan_free(&plan);
    plan_free(&plan64);
    plan_free(&plan_seq);
    plan_free(&plan_adj);
    quad_free(&quad);
    for (auto &x : ev)
      if (x) (void)hipEventDestroy(x);
    for (auto &sl : ra_ev)
      for (auto &x : sl)
        if (x) (void)hipEventDestroy(x);
    if (h_ra) (void)hipHostFree(h_ra);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (h_flag) (void)hipHostFree(h_flag);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  Acc3 g_acc(const DeviceBuffer &b, int rec, int elem) const { return dense_acc3(b.p, e, rec, elem); }

  int upload(DeviceBuffer &b, const void *h, size_t n) {
    SNAP_TRY(b.alloc(n));
    SNAP_HIP_TRY(hipMemcpy(b.p, h, n, hipMemcpyHostToDevice));
    return SNAP_OK;
  }

  const SweepPlan &plan_for(int precision) const {
    return (precision != SNAP_PRECISION_MIXED && plan64.nplanes) ? plan64 : plan;
  }
  // the plan the inner sweeps run on
  const SweepPlan &sweep_plan(int precision) const {
    return plan_seq.nplanes ? plan_seq : plan_for(precision);
  }
  // streams an inner sweep's launches use (launches per tile diagonal)
  int sweep_streams() const { return (plan_seq.nplanes && side) ? 2 : 1; }

  // tiled: a multi-wave tile plan to sweep on instead of the solver's plans
  // (the adjoint's checkpoint faces are in its own plan's layout)
  int sweep(int precision, hipEvent_t *events, bool fixup = true, bool recompute = false,
            const SweepPlan *tiled = nullptr, const int32_t *skip = nullptr) {
    const bool fp32 = precision == SNAP_PRECISION_MIXED;
    SweepLaunch L;
    memset(&L, 0, sizeof(L));
    L.plan = tiled ? tiled : &plan_for(precision);
    L.quad = &quad;
    L.nsweeps = 8 * ng;
    L.ng_job = ng;
    L.corner0 = 0;
    L.fixup = (prob.fixup && fixup) ? 1 : 0;
    L.skip = skip;
    L.hi = hi;
    L.hj = hj;
    L.hk = hk;
    L.general = false;
    L.q = qtot.as<double>();
    L.dinv_dense = fp32 ? dinv32.p : dinv.p;
    L.txs_dense = fp32 ? txs32.p : txs.p;
    L.part_flux = part_flux.as<double>();
    L.part_fluxm = cmom > 1 ? part_fluxm.as<double>() : nullptr;
    L.faces = faces.p;
    L.fp32 = fp32;
    L.exact = precision == SNAP_PRECISION_FP64_EXACT;
    // the one-wavefront tile sweep forms dinv from t_xs unless told to
    // stream it (measured faster: it saves 8 bytes per unit of HBM traffic)
    L.recompute_dinv = recompute || opt.dinv_mode == SNAP_DINV_RECOMPUTE ||
                       (opt.dinv_mode == SNAP_DINV_DEFAULT && plan_seq.nplanes && !tiled);
    if (fp32 && !L.recompute_dinv && !dinv32.p) return fail(SNAP_ERR_INVALID, "no FP32 dinv");
    L.td = td;
    if (td) {
      L.vdelt = vdelt_d.as<double>();
      if (prob.timedep) {
        L.tin = tflux[step & 1].as<double>();
        L.tout = tflux[(step & 1) ^ 1].as<double>();
      }
      L.qim = prob.mms ? qim.as<double>() : nullptr;
    }
    if (plan_seq.nplanes && !tiled) {
      L.plan = &plan_seq;
      L.side = side;
      L.fork = fork_ev;
      L.join = join_ev;
      L.pair = opt.groups_per_wave != 1;
      groups_per_wave = sweep_seq_pair(L) ? 2 : 1;
      return sweep_seq_enqueue(L, stream, events);
    }
    groups_per_wave = 1;
    return sweep_enqueue(L, stream, events);
  }

  // flux = sum of the octant partials (+ the base fluxes of a correction sweep)
  int reduce_octants(bool add_base = false, const int32_t *skip = nullptr) {
    SNAP_TRY(launch_octant_reduce(part_flux.as<double>(), flux0.as<double>(), (int64_t)ng * cells,
                                  stream, add_base ? flux0base.as<double>() : nullptr, skip));
    if (cmom > 1)
      SNAP_TRY(launch_octant_reduce(part_fluxm.as<double>(), fluxm.as<double>(),
                                    (int64_t)ng * cells * (cmom - 1), stream,
                                    add_base ? fluxmbase.as<double>() : nullptr, skip));
    return SNAP_OK;
  }

  int outer_source() {
    // the exact solve keeps the CPU's summation order; the others contract on
    // the FP64 matrix cores
    const bool ordered = opt.precision == SNAP_PRECISION_FP64_EXACT || opt.outer_ordered;
    return launch_outer_dense(qi.as<double>(), flux0.as<double>(),
                              cmom > 1 ? fluxm.as<double>() : nullptr, slgg.as<double>(),
                              mat.as<int32_t>(), q2grp0.as<double>(),
                              cmom > 1 ? q2grpm.as<double>() : nullptr, cells, ng, nmat, nmom,
                              ordered, stream);
  }

  int inner_source(const int32_t *skip = nullptr) {
    return launch_inner_source_dense(sxs.as<double>(), flux0.as<double>(), q2grp0.as<double>(),
                                     cmom > 1 ? fluxm.as<double>() : nullptr,
                                     cmom > 1 ? q2grpm.as<double>() : nullptr, qtot.as<double>(),
                                     (int64_t)ng * cells, nmom, stream, skip);
  }

  // change: when not NULL, also the largest relative change of the test's
  // quantity (the mixed-precision FP64 polish reads it)
  int test(const DeviceBuffer &prev, double eps, bool *converged, double *change = nullptr) {
    SNAP_TRY(launch_convergence(g_acc(flux0, 1, 8), g_acc(prev, 1, 8), e, ng, eps, flag.as<int32_t>(),
                                stream));
    if (change)
      SNAP_TRY(launch_max_change(flux0.as<double>(), prev.as<double>(), (int64_t)ng * cells,
                                 (unsigned long long *)(flag.as<char>() + 8), stream));
    SNAP_HIP_TRY(hipMemcpyAsync(h_flag, flag.p, change ? 16 : sizeof(int32_t),
                                hipMemcpyDeviceToHost, stream));
    SNAP_HIP_TRY(hipStreamSynchronize(stream));
    *converged = (*h_flag != 0);
    if (change) memcpy(change, (const char *)h_flag + 8, sizeof(double));
    return SNAP_OK;
  }

  int copy(DeviceBuffer &dst, const DeviceBuffer &src, size_t n) {
    SNAP_HIP_TRY(hipMemcpyAsync(dst.p, src.p, n, hipMemcpyDeviceToDevice, stream));
    return SNAP_OK;
  }
};

extern "C" int snap_hip_solver_create(const snap_problem_t *p, const snap_solve_options_t *o,
                                      snap_solver_t **out) {
  if (!p || !out) return fail(SNAP_ERR_INVALID, "NULL problem or output");
  *out = nullptr;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(SNAP_ERR_INVALID, "empty grid");
  if (p->ng < 1 || p->nang < 1 || p->nmat < 1) return fail(SNAP_ERR_INVALID, "ng, nang, nmat must be >= 1");
  if (p->nmom < 1 || p->nmom > 4) return fail(SNAP_ERR_INVALID, "nmom must be 1..4");
  if (!(p->epsi > 0.0 && p->epsi < 1e-2)) return fail(SNAP_ERR_INVALID, "epsi must be in (0, 1e-2)");
  if (p->iitm < 1 || p->oitm < 1) return fail(SNAP_ERR_INVALID, "iitm and oitm must be >= 1");
  if (!(p->lx > 0.0 && p->ly > 0.0 && p->lz > 0.0)) return fail(SNAP_ERR_INVALID, "domain lengths must be positive");
  if (!p->mat || !p->sigt || !p->slgg || (!p->qi && !p->mms))
    return fail(SNAP_ERR_INVALID, "NULL problem table");
  if (p->timedep && !(p->tf > 0.0)) return fail(SNAP_ERR_INVALID, "a time-dependent problem needs tf > 0");
  if (p->nsteps < 0) return fail(SNAP_ERR_INVALID, "nsteps must be >= 1");
  std::unique_ptr<snap_solver> s(new snap_solver());
  s->prob = *p;
  if (o) s->opt = *o;
  if (s->opt.precision != SNAP_PRECISION_FP64 && s->opt.precision != SNAP_PRECISION_MIXED &&
      s->opt.precision != SNAP_PRECISION_FP64_EXACT)
    return fail(SNAP_ERR_INVALID, "unknown precision %d", s->opt.precision);
  if (s->prob.nsteps == 0) s->prob.nsteps = 1;
  s->td = p->timedep || p->mms;
  if (s->td && s->opt.refine)
    return fail(SNAP_ERR_UNSUPPORTED, "defect correction of time-dependent or MMS problems");
  if (s->td && s->opt.dinv_mode == SNAP_DINV_RECOMPUTE)
    return fail(SNAP_ERR_UNSUPPORTED, "time-dependent sweeps read dinv (it includes vdelt)");
  s->e = Extent3{p->nx, p->ny, p->nz};
  s->ng = p->ng;
  s->nang = p->nang;
  s->nmom = p->nmom;
  s->cmom = p->nmom * p->nmom;
  s->nmat = p->nmat;
  s->cells = s->e.cells();
  // snap.cc:1178-1180
  s->hi = 2.0 / (p->lx / double(p->nx));
  s->hj = 2.0 / (p->ly / double(p->ny));
  s->hk = 2.0 / (p->lz / double(p->nz));
  for (int64_t c = 0; c < s->cells; c++)
    if (p->mat[c] < 0 || p->mat[c] >= p->nmat)
      return fail(SNAP_ERR_INVALID, "material index %d at cell %lld out of range", p->mat[c], (long long)c);

  std::vector<double> mu, eta, xi, w, ec;
  SNAP_TRY(quad_compute(p->nang, p->nmom, mu, eta, xi, w, ec));
  SNAP_TRY(quad_upload(&s->quad, ec.data(), mu.data(), eta.data(), xi.data(), w.data(), p->nang, p->nmom));
  int TX = s->opt.tile[0], TY = s->opt.tile[1], TZ = s->opt.tile[2], waves = s->opt.waves_per_tile;
  int apl = s->opt.angles_per_lane;
  if (s->opt.groups_per_wave != 0 && s->opt.groups_per_wave != 1)
    return fail(SNAP_ERR_INVALID, "groups_per_wave must be 0 (choose) or 1, got %d",
                s->opt.groups_per_wave);
  // more angles than a wavefront's lanes hold twice over: one workgroup per cell
  const bool wide = s->nang > 128 || s->opt.wide_sweep;
  if (wide) {
    SNAP_TRY(plan_create_wide(&s->plan, s->e, s->nang, s->cmom, s->td));
    if (s->opt.precision == SNAP_PRECISION_MIXED)
      SNAP_TRY(plan_create_wide(&s->plan64, s->e, s->nang, s->cmom, s->td));
  } else {
    // the one-wavefront tile sweep: the default for steady problems of
    // 33..64 angles, asked for on any steady problem of up to 64 angles
    const bool seq_ok = !s->td && s->nang <= 64;
    if (s->opt.sweep_kernel == SNAP_SWEEP_WAVE_TILE && !seq_ok)
      return fail(SNAP_ERR_UNSUPPORTED,
                  "the one-wavefront tile sweep takes steady problems of up to 64 angles");
    if (s->opt.sweep_kernel < SNAP_SWEEP_AUTO || s->opt.sweep_kernel > SNAP_SWEEP_WAVE_TILE)
      return fail(SNAP_ERR_INVALID, "unknown sweep kernel %d", s->opt.sweep_kernel);
    // options that only the multi-wave tile sweep has keep AUTO on it
    const bool tiled_opts = s->opt.angles_per_lane != 0 || s->opt.waves_per_tile != 0 ||
                            s->opt.tile[0] != 0 || s->opt.tile[1] != 0;
    if (s->opt.sweep_kernel == SNAP_SWEEP_WAVE_TILE ||
        (s->opt.sweep_kernel == SNAP_SWEEP_AUTO && seq_ok && s->nang > 32 && !tiled_opts)) {
      // 16 planes per tile from 64 cells deep (half the z faces of 8: the 64^3
      // FP64 solve 1.4% faster), else 8
      const int tz = (s->opt.sweep_kernel == SNAP_SWEEP_WAVE_TILE && s->opt.tile[2] > 0)
                         ? s->opt.tile[2]
                         : (s->e.nz >= 64 ? 16 : 8);
      SNAP_TRY(plan_create_seq(&s->plan_seq, s->e, s->nang, s->cmom, tz));
    }
    plan_choose(s->e, s->nang, s->cmom, s->opt.precision == SNAP_PRECISION_MIXED ? 4 : 8, &TX,
                &TY, &TZ, &waves, &apl, s->td);
    SNAP_TRY(plan_create(&s->plan, s->e, s->nang, s->cmom, TX, TY, TZ, waves, apl, s->td));
    if (s->opt.precision == SNAP_PRECISION_MIXED) {
      // the FP64 sweeps of a mixed solve (first sweep, refinement bases) get
      // their own plan unless the caller fixed the tiling
      int tx = s->opt.tile[0], ty = s->opt.tile[1], tz = s->opt.tile[2];
      int w = s->opt.waves_per_tile, a = s->opt.angles_per_lane;
      plan_choose(s->e, s->nang, s->cmom, 8, &tx, &ty, &tz, &w, &a, s->td);
      SNAP_TRY(plan_create(&s->plan64, s->e, s->nang, s->cmom, tx, ty, tz, w, a, s->td));
    }
  }

  SNAP_HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  for (auto &x : s->ev) SNAP_HIP_TRY(hipEventCreate(&x));
  if (s->plan_seq.nplanes) {
    SNAP_HIP_TRY(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    SNAP_HIP_TRY(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming));
    SNAP_HIP_TRY(hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming));
  }
  SNAP_HIP_TRY(hipHostMalloc((void **)&s->h_flag, 16, hipHostMallocDefault));
  SNAP_HIP_TRY(hipHostMalloc((void **)&s->h_ra, 2 * sizeof(int32_t), hipHostMallocDefault));
  SNAP_TRY(s->ra_flag.alloc(2 * sizeof(int32_t)));
  for (auto &sl : s->ra_ev)
    for (auto &x : sl) SNAP_HIP_TRY(hipEventCreate(&x));

  const size_t gc = (size_t)s->ng * s->cells;
  const size_t nm1 = s->cmom > 1 ? (size_t)(s->cmom - 1) : 0;
  SNAP_TRY(s->upload(s->mat, p->mat, sizeof(int32_t) * s->cells));
  SNAP_TRY(s->upload(s->sigt, p->sigt, sizeof(double) * p->nmat * p->ng));
  SNAP_TRY(s->upload(s->slgg, p->slgg, sizeof(double) * (size_t)p->ng * p->nmat * p->ng * p->nmom));
  // Snap::initialize_velocity (snap.cc:716-738)
  s->vdelt_h.assign(s->ng, 0.0);
  if (p->timedep) {
    const double dt = p->tf / (double)s->prob.nsteps;
    for (int g = 0; g < s->ng; g++) s->vdelt_h[g] = 2.0 / (dt * (double)(s->ng - g));
  }
  SNAP_TRY(s->upload(s->vdelt_d, s->vdelt_h.data(), sizeof(double) * s->ng));
  if (p->mms) {
    SNAP_TRY(mms_prepare(s->prob, &s->mms));
    SNAP_TRY(s->upload(s->qi, s->mms.qi.data(), sizeof(double) * gc));
    SNAP_TRY(s->qim.alloc(sizeof(double) * 8 * gc * s->nang));
  } else {
    SNAP_TRY(s->upload(s->qi, p->qi, sizeof(double) * gc));
  }
  if (p->timedep) {
    SNAP_TRY(s->tflux[0].alloc(sizeof(double) * 8 * gc * s->nang));
    SNAP_TRY(s->tflux[1].alloc(sizeof(double) * 8 * gc * s->nang));
  }
  SNAP_TRY(s->txs.alloc(sizeof(double) * gc));
  SNAP_TRY(s->sxs.alloc(sizeof(double) * gc * s->nmom));
  SNAP_TRY(s->dinv.alloc(sizeof(double) * gc * s->nang));
  SNAP_TRY(s->flux0.alloc(sizeof(double) * gc));
  SNAP_TRY(s->flux0pi.alloc(sizeof(double) * gc));
  SNAP_TRY(s->flux0po.alloc(sizeof(double) * gc));
  SNAP_TRY(s->q2grp0.alloc(sizeof(double) * gc));
  SNAP_TRY(s->qtot.alloc(sizeof(double) * gc * s->cmom));
  SNAP_TRY(s->fluxm.alloc(sizeof(double) * gc * nm1));
  SNAP_TRY(s->q2grpm.alloc(sizeof(double) * gc * nm1));
  SNAP_TRY(s->part_flux.alloc(sizeof(double) * 8 * gc));
  SNAP_TRY(s->part_fluxm.alloc(sizeof(double) * 8 * gc * nm1));
  SNAP_TRY(s->faces.alloc(sizeof(double) * 8 * s->ng *
                          std::max({s->plan.face_doubles_per_sweep(),
                                    s->plan64.nplanes ? s->plan64.face_doubles_per_sweep() : 0,
                                    s->plan_seq.nplanes ? s->plan_seq.face_doubles_per_sweep()
                                                        : 0})));
  SNAP_TRY(s->flag.alloc(16));

  // ExpandCrossSection(sigt -> t_xs), ExpandScatteringCrossSection,
  // CalculateGeometryParam (snap.cc:419-450)
  const hipStream_t st = s->stream;
  if (p->sigt_cell) {
    SNAP_HIP_TRY(hipMemcpyAsync(s->txs.p, p->sigt_cell, sizeof(double) * gc, hipMemcpyHostToDevice, st));
  } else {
    Acc2 sig;
    sig.base = s->sigt.as<char>();
    sig.s0 = (int64_t)p->ng * 8;
    sig.s1 = 0;
    sig.sg = 8;
    SNAP_TRY(launch_expand_xs(sig, s->g_acc(s->mat, 1, 4), s->g_acc(s->txs, 1, 8), s->e, s->ng, st));
  }
  {
    Acc2 sl;
    sl.base = s->slgg.as<char>();
    sl.s1 = (int64_t)s->nmom * 8;
    sl.s0 = sl.s1 * s->ng;
    sl.sg = sl.s0 * s->nmat;
    SNAP_TRY(launch_expand_scat(sl, 0, s->g_acc(s->mat, 1, 4), s->g_acc(s->sxs, s->nmom, 8), s->e,
                                s->ng, s->nmom, st));
  }
  const double *vd = p->timedep ? s->vdelt_h.data() : nullptr;
  SNAP_TRY(launch_geometry(s->g_acc(s->txs, 1, 8), s->g_acc(s->dinv, s->nang, 8), vd, s->e,
                           s->ng, s->nang, s->hi, s->hj, s->hk, s->quad, false, st));
  if (s->opt.precision == SNAP_PRECISION_MIXED) {
    SNAP_TRY(s->dinv32.alloc(sizeof(float) * gc * s->nang));
    SNAP_TRY(s->txs32.alloc(sizeof(float) * gc));
    SNAP_TRY(launch_geometry(s->g_acc(s->txs, 1, 8), s->g_acc(s->dinv32, s->nang, 4), vd, s->e,
                             s->ng, s->nang, s->hi, s->hj, s->hk, s->quad, true, st));
    // t_xs in float for the FP32 fixup
    std::vector<double> h(gc);
    std::vector<float> hf(gc);
    SNAP_HIP_TRY(hipMemcpyAsync(h.data(), s->txs.p, sizeof(double) * gc, hipMemcpyDeviceToHost, st));
    SNAP_HIP_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < gc; i++) hf[i] = (float)h[i];
    SNAP_HIP_TRY(hipMemcpy(s->txs32.p, hf.data(), sizeof(float) * gc, hipMemcpyHostToDevice));
    s->have_fp32 = true;
    if (s->opt.refine) {
      SNAP_TRY(s->qbase.alloc(sizeof(double) * gc * s->cmom));
      SNAP_TRY(s->flux0base.alloc(sizeof(double) * gc));
      SNAP_TRY(s->fluxmbase.alloc(sizeof(double) * gc * nm1));
    }
  }
  SNAP_HIP_TRY(hipStreamSynchronize(st));
  *out = s.release();
  return SNAP_OK;
}

extern "C" int snap_hip_solver_run(snap_solver_t *s, snap_solve_stats_t *stats) {
  if (!s) return fail(SNAP_ERR_INVALID, "NULL solver");
  snap_solve_stats_t st;
  memset(&st, 0, sizeof(st));
  const size_t gc = (size_t)s->ng * s->cells;
  const size_t nm1 = s->cmom > 1 ? (size_t)(s->cmom - 1) : 0;
  const hipStream_t stream = s->stream;
  const bool mixed = s->opt.precision == SNAP_PRECISION_MIXED;
  SNAP_HIP_TRY(hipMemsetAsync(s->flux0.p, 0, sizeof(double) * gc, stream));
  if (nm1) SNAP_HIP_TRY(hipMemsetAsync(s->fluxm.p, 0, sizeof(double) * gc * nm1, stream));
  // the time-edge fluxes start at zero (snap.cc:332-335) and the angular
  // source unscaled (MMSInitSource)
  const size_t ta = 8 * gc * (size_t)s->nang;
  if (s->prob.timedep)
    for (auto &b : s->tflux) SNAP_HIP_TRY(hipMemsetAsync(b.p, 0, sizeof(double) * ta, stream));
  if (s->prob.mms)
    SNAP_TRY(mms_source(s->mms, s->quad, s->e, s->ng, s->nmat, s->nmom, s->mat.as<int32_t>(),
                        s->sigt.as<double>(), s->slgg.as<double>(), s->qim.as<double>(), stream));
  SNAP_HIP_TRY(hipEventRecord(s->ev[2], stream));
  float sweep_ms = 0.f;
  bool all_converged = true;
  const double dt = s->prob.tf / (double)s->prob.nsteps;
  s->log.clear();
  const auto t0 = std::chrono::steady_clock::now();
  auto record = [&](int kind, int cy, int otno, int inno, bool conv) {
    snap_iteration_record_t r;
    r.kind = kind;
    r.time_step = cy;
    r.outer = otno;
    r.inner = inno;
    r.converged = conv ? 1 : 0;
    r.t_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    s->log.push_back(r);
  };
  for (int cy = 0; cy < s->prob.nsteps; cy++) {
    s->step = cy;
    // MMSScale (snap.cc:452-467): the source at the step's mid-time
    if (s->prob.mms)
      SNAP_TRY(launch_scale(s->qim.as<double>(), (int64_t)ta,
                            cy == 0 ? dt * 0.5 : double(2 * cy + 1) / double(2 * cy - 1), stream));
    bool inner_converged = false, converged = false;
    const bool polish = mixed && s->opt.fp64_polish > 0.0;
    for (int otno = 0; otno < s->prob.oitm; otno++) {
      SNAP_TRY(s->outer_source());
      SNAP_TRY(s->copy(s->flux0po, s->flux0, sizeof(double) * gc));
      st.outer_total++;
      inner_converged = false;
      // the previous inner's largest relative change, within this outer
      double change = HUGE_VAL;
      if (!mixed) {
        // FP64: inner iteration i+1 is enqueued before the host reads
        // iteration i's test, every kernel of it predicated on that test's
        // device flag (set: the loop has ended, and the iteration does
        // nothing), so the device does not wait for the host between inner
        // iterations.  Two slots of flags and events alternate.
        auto enqueue_inner = [&](int inno, const int32_t *skip) -> int {
          const int sl = inno & 1;
          int32_t *conv = s->ra_flag.as<int32_t>() + sl;
          SNAP_TRY(s->inner_source(skip));
          SNAP_TRY(launch_copy(s->flux0pi.as<double>(), s->flux0.as<double>(), (int64_t)gc, stream,
                               skip));
          SNAP_HIP_TRY(hipEventRecord(s->ra_ev[sl][0], stream));
          SNAP_TRY(s->sweep(s->opt.precision, nullptr, true, false, nullptr, skip));
          // the octant sum is part of the sweep: the reference accumulates
          // the flux moments inside its sweep kernel (gpu_sweep.cu:1680-1733)
          SNAP_TRY(s->reduce_octants(false, skip));
          SNAP_HIP_TRY(hipEventRecord(s->ra_ev[sl][1], stream));
          SNAP_TRY(launch_convergence(s->g_acc(s->flux0, 1, 8), s->g_acc(s->flux0pi, 1, 8), s->e,
                                      s->ng, s->prob.epsi, conv, stream, skip));
          SNAP_HIP_TRY(hipMemcpyAsync(s->h_ra + sl, conv, sizeof(int32_t), hipMemcpyDeviceToHost,
                                      stream));
          SNAP_HIP_TRY(hipEventRecord(s->ra_ev[sl][2], stream));
          return SNAP_OK;
        };
        SNAP_TRY(enqueue_inner(0, nullptr));
        for (int inno = 0; inno < s->prob.iitm; inno++) {
          const int sl = inno & 1;
          if (inno + 1 < s->prob.iitm)
            SNAP_TRY(enqueue_inner(inno + 1, s->ra_flag.as<int32_t>() + sl));
          SNAP_HIP_TRY(hipEventSynchronize(s->ra_ev[sl][2]));
          inner_converged = s->h_ra[sl] != 0;
          st.sweeps_fp64++;
          st.inner_total++;
          record(0, cy, otno, inno, inner_converged);
          float ms = 0.f;
          SNAP_HIP_TRY(hipEventElapsedTime(&ms, s->ra_ev[sl][0], s->ra_ev[sl][1]));
          sweep_ms += ms;
          if (inner_converged) break;  // the iteration enqueued after it does nothing
        }
      } else {
        // mixed precision: the next inner sweep's precision depends on this
        // test, so the host reads every test before it enqueues more
        for (int inno = 0; inno < s->prob.iitm; inno++) {
          SNAP_TRY(s->inner_source());
          SNAP_TRY(s->copy(s->flux0pi, s->flux0, sizeof(double) * gc));
          const bool refine = mixed && s->opt.refine;
          // FP64 polish: an inner whose predecessor moved the flux by at most
          // fp64_polish * epsi is likely one of the outer's last (the change
          // contracts by the in-group scattering ratio each inner), so it runs
          // in FP64 and the outer ends on FP64 sweeps; so does every inner once
          // FP32 sweeps have reached their own rounding level
          const bool polish_now =
              polish && change <= polish_threshold(s->opt.fp64_polish, s->prob.epsi);
          const bool fp32 = mixed && !(s->opt.first_sweep_fp64 && st.inner_total == 0) &&
                            !(refine && inno == 0) && !polish_now;
          if (refine && fp32)  // FP64 defect of the source against the outer's base
            SNAP_TRY(launch_subtract(s->qtot.as<double>(), s->qbase.as<double>(),
                                     (int64_t)gc * s->cmom, stream));
          SNAP_HIP_TRY(hipEventRecord(s->ev[0], stream));
          SNAP_TRY(s->sweep(fp32 ? SNAP_PRECISION_MIXED
                                 : (mixed ? SNAP_PRECISION_FP64 : s->opt.precision),
                            nullptr, !(refine && fp32)));
          // the octant sum is part of the sweep: the reference accumulates the
          // flux moments inside its sweep kernel (gpu_sweep.cu:1680-1733)
          SNAP_TRY(s->reduce_octants(refine && fp32));
          SNAP_HIP_TRY(hipEventRecord(s->ev[1], stream));
          if (refine && !fp32) {  // keep the FP64 base of this outer
            SNAP_TRY(s->copy(s->qbase, s->qtot, sizeof(double) * gc * s->cmom));
            SNAP_TRY(s->copy(s->flux0base, s->flux0, sizeof(double) * gc));
            if (nm1) SNAP_TRY(s->copy(s->fluxmbase, s->fluxm, sizeof(double) * gc * nm1));
          }
          if (fp32)
            st.sweeps_fp32++;
          else
            st.sweeps_fp64++;
          st.inner_total++;
          SNAP_TRY(s->test(s->flux0pi, s->prob.epsi, &inner_converged, polish ? &change : nullptr));
          record(0, cy, otno, inno, inner_converged);
          float ms = 0.f;
          SNAP_HIP_TRY(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
          sweep_ms += ms;
          if (inner_converged) break;
        }
      }
      if (otno == 0) continue;
      bool outer_converged = false;
      if (inner_converged) SNAP_TRY(s->test(s->flux0po, 100.0 * s->prob.epsi, &outer_converged));
      record(1, cy, otno, -1, outer_converged);
      if (outer_converged) {
        converged = true;
        break;
      }
    }
    all_converged = all_converged && converged;
    st.time_steps++;
  }
  st.converged = all_converged ? 1 : 0;
  SNAP_HIP_TRY(hipEventRecord(s->ev[3], stream));
  SNAP_HIP_TRY(hipEventSynchronize(s->ev[3]));
  float solve_ms = 0.f;
  SNAP_HIP_TRY(hipEventElapsedTime(&solve_ms, s->ev[2], s->ev[3]));
  st.sweep_ms = sweep_ms;
  st.solve_ms = solve_ms;
  st.units_per_sweep = (double)s->cells * 8.0 * s->nang * s->ng;
  st.groups_per_wave = s->groups_per_wave;
  st.grind_ns = st.inner_total ? (sweep_ms * 1e6) / (st.units_per_sweep * st.inner_total) : 0.0;
  if (stats) *stats = st;
  return SNAP_OK;
}

extern "C" int snap_hip_solver_mms_compare(snap_solver_t *s, double result[3]) {
  if (!s || !result) return fail(SNAP_ERR_INVALID, "NULL argument");
  if (!s->prob.mms) return fail(SNAP_ERR_INVALID, "the problem has no manufactured solution");
  const size_t gc = (size_t)s->ng * s->cells;
  std::vector<double> f(gc);
  SNAP_HIP_TRY(hipStreamSynchronize(s->stream));
  SNAP_HIP_TRY(hipMemcpy(f.data(), s->flux0.p, sizeof(double) * gc, hipMemcpyDeviceToHost));
  mms_compare(f.data(), s->mms.ref.data(), gc, result);
  return SNAP_OK;
}

extern "C" int snap_hip_solver_get_log(snap_solver_t *s, snap_iteration_record_t *records,
                                       int capacity, int *count) {
  if (!s || !count || (capacity > 0 && !records)) return fail(SNAP_ERR_INVALID, "NULL argument");
  const int n = (int)s->log.size();
  *count = n;
  for (int i = 0; i < n && i < capacity; i++) records[i] = s->log[i];
  return SNAP_OK;
}

extern "C" int snap_hip_mapper_tunables(int num_groups, int max_inner_iters, int processors,
                                        int sms, int32_t out[4]) {
  if (!out || num_groups < 1 || max_inner_iters < 1 || processors < 1 || sms < 1)
    return fail(SNAP_ERR_INVALID, "bad tunable query");
  out[0] = 2;                // OUTER_RUNAHEAD_TUNABLE
  out[1] = max_inner_iters;  // INNER_RUNAHEAD_TUNABLE: the whole inner loop
  // SWEEP_ENERGY_CHUNKS_TUNABLE: 8 directions x groups over the processors,
  // a quarter of that when it is at least 4, at most all groups
  int chunk = 8 * num_groups / processors;
  if (chunk >= 4) chunk /= 4;
  if (chunk > num_groups) chunk = num_groups;
  out[2] = chunk;
  // GPU_SMS_PER_SWEEP_TUNABLE
  out[3] = sms <= 4 ? sms : (sms <= 16 ? 4 : 8);
  return SNAP_OK;
}

extern "C" int snap_hip_solver_get_mms_reference(snap_solver_t *s, double *ref_flux) {
  if (!s || !ref_flux) return fail(SNAP_ERR_INVALID, "NULL argument");
  if (!s->prob.mms) return fail(SNAP_ERR_INVALID, "the problem has no manufactured solution");
  memcpy(ref_flux, s->mms.ref.data(), sizeof(double) * s->mms.ref.size());
  return SNAP_OK;
}

extern "C" int snap_hip_solver_get_flux(snap_solver_t *s, double *flux0, double *fluxm) {
  if (!s) return fail(SNAP_ERR_INVALID, "NULL solver");
  const size_t gc = (size_t)s->ng * s->cells;
  SNAP_HIP_TRY(hipStreamSynchronize(s->stream));
  if (flux0) SNAP_HIP_TRY(hipMemcpy(flux0, s->flux0.p, sizeof(double) * gc, hipMemcpyDeviceToHost));
  if (fluxm && s->cmom > 1)
    SNAP_HIP_TRY(hipMemcpy(fluxm, s->fluxm.p, sizeof(double) * gc * (s->cmom - 1), hipMemcpyDeviceToHost));
  return SNAP_OK;
}

extern "C" int snap_hip_solver_time_sweeps(snap_solver_t *s, int count, int precision,
                                           double *ms_per_sweep, double *ms_per_launch,
                                           int *launches_per_sweep) {
  if (!s || count < 1) return fail(SNAP_ERR_INVALID, "bad timing request");
  const bool fp32 = precision == SNAP_PRECISION_MIXED;
  if (fp32 && !s->have_fp32) return fail(SNAP_ERR_INVALID, "solver was created without mixed precision");
  if (precision < 0 || precision > 2) return fail(SNAP_ERR_INVALID, "unknown precision %d", precision);
  const int np = s->sweep_plan(precision).nplanes * s->sweep_streams();
  std::vector<hipEvent_t> evs((size_t)2 * np * count);
  for (auto &x : evs) SNAP_HIP_TRY(hipEventCreate(&x));
  SNAP_TRY(s->sweep(precision, nullptr));  // first launches load the code object
  SNAP_HIP_TRY(hipEventRecord(s->ev[0], s->stream));
  for (int i = 0; i < count; i++) {
    SNAP_TRY(s->sweep(precision, evs.data() + (size_t)2 * np * i));
    SNAP_TRY(s->reduce_octants());
  }
  SNAP_HIP_TRY(hipEventRecord(s->ev[1], s->stream));
  SNAP_HIP_TRY(hipEventSynchronize(s->ev[1]));
  float total = 0.f;
  SNAP_HIP_TRY(hipEventElapsedTime(&total, s->ev[0], s->ev[1]));
  double launch_sum = 0.0;
  for (int i = 0; i < np * count; i++) {
    float ms = 0.f;
    SNAP_HIP_TRY(hipEventElapsedTime(&ms, evs[2 * i], evs[2 * i + 1]));
    launch_sum += ms;
  }
  for (auto &x : evs) (void)hipEventDestroy(x);
  if (ms_per_sweep) *ms_per_sweep = total / count;
  if (ms_per_launch) *ms_per_launch = launch_sum / (double)(np * count);
  if (launches_per_sweep) *launches_per_sweep = np;
  return SNAP_OK;
}

extern "C" int snap_hip_solver_gradient(snap_solver_t *s, const double *resp, int precision,
                                        double tol, int maxit, double *grad, double *response,
                                        int *iterations) {
  if (!s || !resp || !grad) return fail(SNAP_ERR_INVALID, "NULL argument");
  if (s->nmom != 1) return fail(SNAP_ERR_UNSUPPORTED, "the adjoint handles nmom = 1 (isotropic scattering)");
  if (s->td) return fail(SNAP_ERR_UNSUPPORTED, "the adjoint is of the steady-state solve");
  if (s->plan.wide) return fail(SNAP_ERR_UNSUPPORTED, "the adjoint sweep is tiled (up to 64 angles)");
  if (precision != SNAP_PRECISION_FP64 && precision != SNAP_PRECISION_MIXED)
    return fail(SNAP_ERR_INVALID, "adjoint precision must be FP64 or MIXED (FP32 sweeps)");
  if (!(tol > 0.0) || maxit < 1) return fail(SNAP_ERR_INVALID, "bad tolerance or iteration cap");
  const bool fp32 = precision == SNAP_PRECISION_MIXED;
  // the checkpoint and the reverse kernel's forward recompute are FP64 (the
  // fixup's decisions of the FP64 solve); `fp32` sets the reverse recurrence.
  // The reverse sweep takes 8 x 4 x 2 tiles of its own (kAdjTile: 0.13 s
  // against 0.32 s on the forward solve's 8 x 8 x 4 tiles, the same gradient
  // to the last bit); the forward plan is the fallback where that tile does
  // not fit.
  const hipStream_t st = s->stream;
  if (!s->plan_adj.nplanes) {
    int tx = std::min(kAdjTile[0], s->e.nx), ty = std::min(kAdjTile[1], s->e.ny),
        tz = std::min(kAdjTile[2], s->e.nz);
    int w = 0, a = 0, cpw, waves, sb;
    plan_choose(s->e, s->nang, s->cmom, 8, &tx, &ty, &tz, &w, &a, false);
    if (plan_create(&s->plan_adj, s->e, s->nang, s->cmom, tx, ty, tz, w, a, false) != SNAP_OK ||
        adjoint_plan_fits(s->plan_adj, s->nang, 8, &cpw, &waves, &sb) != SNAP_OK)
      plan_free(&s->plan_adj);
  }
  const SweepPlan &pl = s->plan_adj.nplanes ? s->plan_adj : s->plan_for(SNAP_PRECISION_FP64);
  {
    int cpw, waves, sb;
    SNAP_TRY(adjoint_plan_fits(pl, s->nang, fp32 ? 4 : 8, &cpw, &waves, &sb));
  }
  const size_t gc = (size_t)s->ng * s->cells;
  // checkpoint and adjoint faces in the plan's layout, every (octant, group)
  const size_t face_bytes = sizeof(double) * 8 * s->ng * pl.face_doubles_per_sweep();
  SNAP_TRY(s->faces.grow(face_bytes, st));
  DeviceBuffer r, zeta, zn, u, gsum, norms;
  SNAP_TRY(r.alloc(sizeof(double) * gc));
  SNAP_TRY(zeta.alloc(sizeof(double) * gc));
  SNAP_TRY(zn.alloc(sizeof(double) * gc));
  SNAP_TRY(u.alloc(sizeof(double) * gc));
  SNAP_TRY(gsum.alloc(sizeof(double) * gc));
  SNAP_TRY(norms.alloc(2 * sizeof(unsigned long long)));
  if (!s->part_grad.p) SNAP_TRY(s->part_grad.alloc(sizeof(double) * 8 * gc));
  SNAP_TRY(s->adj_faces.grow(face_bytes, st));
  SNAP_HIP_TRY(hipMemcpyAsync(r.p, resp, sizeof(double) * gc, hipMemcpyHostToDevice, st));
  // forward source of the current flux, q = qi + S flux0, and one forward
  // sweep of it: its tile faces are the checkpoint the reverse sweeps start from
  SNAP_TRY(s->outer_source());
  SNAP_TRY(s->inner_source());
  SNAP_TRY(s->sweep(SNAP_PRECISION_FP64, nullptr, true, true, &pl));
  SNAP_HIP_TRY(hipMemcpyAsync(zeta.p, r.p, sizeof(double) * gc, hipMemcpyDeviceToDevice, st));
  auto reverse = [&](const DeviceBuffer &z) -> int {
    SNAP_TRY(adjoint_enqueue(pl, s->quad, s->ng, s->prob.fixup ? 1 : 0, s->hi, s->hj, s->hk,
                             s->qtot.as<double>(), s->txs.as<double>(), z.as<double>(),
                             s->part_flux.as<double>(), s->part_grad.as<double>(), s->faces.p,
                             s->adj_faces.p, fp32, st));
    return launch_octant_reduce(s->part_flux.as<double>(), u.as<double>(), (int64_t)gc, st);
  };
  int it = 0;
  bool done = false;
  double prev = HUGE_VAL;
  while (!done && it < maxit) {
    SNAP_TRY(reverse(zeta));
    SNAP_TRY(launch_adjoint_update(r.as<double>(), u.as<double>(), s->slgg.as<double>(),
                                   s->mat.as<int32_t>(), zeta.as<double>(), zn.as<double>(),
                                   norms.as<unsigned long long>(), s->cells, s->ng, s->nmat,
                                   s->nmom, st));
    unsigned long long h[2];
    SNAP_HIP_TRY(hipMemcpyAsync(h, norms.p, sizeof(h), hipMemcpyDeviceToHost, st));
    SNAP_HIP_TRY(hipStreamSynchronize(st));
    double dmax, zmax;
    memcpy(&dmax, &h[0], sizeof(double));
    memcpy(&zmax, &h[1], sizeof(double));
    std::swap(zeta.p, zn.p);
    it++;
    done = dmax <= tol * zmax;
    // FP32 reverse sweeps: stop at their rounding floor, where the change
    // (below 1e-5 of max|zeta|) no longer shrinks
    if (!done && fp32 && dmax <= 1.0e-5 * zmax && dmax >= prev) done = true;
    prev = dmax;
  }
  // the gradient at the converged adjoint
  SNAP_TRY(reverse(zeta));
  SNAP_TRY(launch_octant_reduce(s->part_grad.as<double>(), gsum.as<double>(), (int64_t)gc, st));
  SNAP_HIP_TRY(hipMemcpyAsync(grad, gsum.p, sizeof(double) * gc, hipMemcpyDeviceToHost, st));
  if (response) {
    std::vector<double> f(gc);
    SNAP_HIP_TRY(hipMemcpyAsync(f.data(), s->flux0.p, sizeof(double) * gc, hipMemcpyDeviceToHost, st));
    SNAP_HIP_TRY(hipStreamSynchronize(st));
    double acc = 0.0;
    for (size_t i = 0; i < gc; i++) acc += resp[i] * f[i];
    *response = acc;
  }
  SNAP_HIP_TRY(hipStreamSynchronize(st));
  if (iterations) *iterations = it;
  if (!done) return fail(SNAP_ERR_INVALID, "adjoint iteration did not reach %g in %d steps", tol, maxit);
  return SNAP_OK;
}

// cells per wavefront of the one-wavefront tile sweep: one 4 x 4 plane
constexpr int kNP_wave_tile = 16;

extern "C" int snap_hip_solver_plan(snap_solver_t *s, int32_t plan[6]) {
  if (!s || !plan) return fail(SNAP_ERR_INVALID, "null argument");
  const SweepPlan &p = s->sweep_plan(s->opt.precision);
  plan[0] = p.TX;
  plan[1] = p.TY;
  plan[2] = p.TZ;
  plan[3] = p.waves;
  plan[4] = p.apl;
  plan[5] = p.seq ? kNP_wave_tile : p.cpw;
  return SNAP_OK;
}

extern "C" void snap_hip_solver_destroy(snap_solver_t *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  delete s;
}
