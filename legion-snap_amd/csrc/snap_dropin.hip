// snap_dropin.hip -- the snap_hip_run_* entry points that replace the
// reference's element-wise GPU task helpers (outer/inner source, convergence,
// cross-section expansion, geometry).  Each converts the ABI accessors into
// group-strided device accessors and enqueues one kernel on the caller's
// stream.
#include <mutex>

#include "snap_internal.h"

using namespace snap;

extern "C" int snap_hip_run_flux0_outer_source(snap_rect3_t b, const snap_accessor3_t *fa_qi0,
                                               const snap_accessor3_t *fa_flux0,
                                               const snap_accessor2_t *fa_slgg,
                                               const snap_accessor3_t *fa_qo0,
                                               snap_accessor3_t fa_mat, int num_groups,
                                               snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  if (num_groups < 1) return fail(SNAP_ERR_INVALID, "num_groups must be >= 1");
  Acc3 qi, f0, qo, mat;
  Acc2 sl;
  SNAP_TRY(group_acc3(fa_qi0, num_groups, b, &qi, "qi0"));
  SNAP_TRY(group_acc3(fa_flux0, num_groups, b, &f0, "flux0"));
  SNAP_TRY(group_acc3(fa_qo0, num_groups, b, &qo, "qo0"));
  SNAP_TRY(group_acc2(fa_slgg, num_groups, &sl, "slgg"));
  SNAP_TRY(single_acc3(fa_mat, b, &mat, "mat"));
  return launch_outer_source(qi, f0, sl, qo, mat, e, num_groups, (hipStream_t)stream);
}

extern "C" int snap_hip_run_fluxm_outer_source(snap_rect3_t b, const snap_accessor3_t *fa_fluxm,
                                               const snap_accessor2_t *fa_slgg,
                                               const snap_accessor3_t *fa_qom,
                                               snap_accessor3_t fa_mat, int num_groups,
                                               int num_moments, const int lma[4],
                                               snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  if (num_groups < 1) return fail(SNAP_ERR_INVALID, "num_groups must be >= 1");
  if (num_moments < 1 || num_moments > 4 || !lma)
    return fail(SNAP_ERR_INVALID, "num_moments must be 1..4 with an lma table");
  if (num_moments == 1) return SNAP_OK;
  Acc3 fm, qo, mat;
  Acc2 sl;
  SNAP_TRY(group_acc3(fa_fluxm, num_groups, b, &fm, "fluxm"));
  SNAP_TRY(group_acc3(fa_qom, num_groups, b, &qo, "qom"));
  SNAP_TRY(group_acc2(fa_slgg, num_groups, &sl, "slgg"));
  SNAP_TRY(single_acc3(fa_mat, b, &mat, "mat"));
  return launch_outer_source_m(fm, sl, qo, mat, e, num_groups, num_moments * num_moments, lma,
                               num_moments, (hipStream_t)stream);
}

extern "C" int snap_hip_run_inner_source_single_moment(snap_rect3_t b, snap_accessor3_t fa_sxs,
                                                       snap_accessor3_t fa_flux0,
                                                       snap_accessor3_t fa_q2grp0,
                                                       snap_accessor3_t fa_qtot,
                                                       snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  Acc3 sxs, f0, q0, qt, none{};
  SNAP_TRY(single_acc3(fa_sxs, b, &sxs, "s_xs"));
  SNAP_TRY(single_acc3(fa_flux0, b, &f0, "flux0"));
  SNAP_TRY(single_acc3(fa_q2grp0, b, &q0, "q2grp0"));
  SNAP_TRY(single_acc3(fa_qtot, b, &qt, "qtot"));
  const int lma[4] = {1, 3, 5, 7};
  return launch_inner_source(sxs, f0, q0, none, none, qt, e, 1, 1, lma, 1, (hipStream_t)stream);
}

extern "C" int snap_hip_run_inner_source_multi_moment(
    snap_rect3_t b, snap_accessor3_t fa_sxs, snap_accessor3_t fa_flux0,
    snap_accessor3_t fa_q2grp0, snap_accessor3_t fa_fluxm, snap_accessor3_t fa_q2grpm,
    snap_accessor3_t fa_qtot, int num_moments, const int lma[4], snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  if (num_moments < 1 || num_moments > 4 || !lma)
    return fail(SNAP_ERR_INVALID, "num_moments must be 1..4 with an lma table");
  Acc3 sxs, f0, q0, fm, qm, qt;
  SNAP_TRY(single_acc3(fa_sxs, b, &sxs, "s_xs"));
  SNAP_TRY(single_acc3(fa_flux0, b, &f0, "flux0"));
  SNAP_TRY(single_acc3(fa_q2grp0, b, &q0, "q2grp0"));
  SNAP_TRY(single_acc3(fa_fluxm, b, &fm, "fluxm", num_moments == 1));
  SNAP_TRY(single_acc3(fa_q2grpm, b, &qm, "q2grpm", num_moments == 1));
  SNAP_TRY(single_acc3(fa_qtot, b, &qt, "qtot"));
  return launch_inner_source(sxs, f0, q0, fm, qm, qt, e, 1, num_moments * num_moments, lma,
                             num_moments, (hipStream_t)stream);
}

static int run_convergence(snap_rect3_t b, int32_t *result, const snap_accessor3_t *fa_flux0,
                           const snap_accessor3_t *fa_flux0p, int num_groups, double epsi,
                           snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  if (!result) return fail(SNAP_ERR_INVALID, "NULL result pointer");
  if (num_groups < 1) return fail(SNAP_ERR_INVALID, "num_groups must be >= 1");
  Acc3 f0, fp;
  SNAP_TRY(group_acc3(fa_flux0, num_groups, b, &f0, "flux0"));
  SNAP_TRY(group_acc3(fa_flux0p, num_groups, b, &fp, "flux0 previous"));
  return launch_convergence(f0, fp, e, num_groups, epsi, result, (hipStream_t)stream);
}

extern "C" int snap_hip_run_inner_convergence(snap_rect3_t b, int32_t *result,
                                              const snap_accessor3_t *fa_flux0,
                                              const snap_accessor3_t *fa_flux0pi, int num_groups,
                                              double epsi, snap_stream_t stream) {
  return run_convergence(b, result, fa_flux0, fa_flux0pi, num_groups, epsi, stream);
}

extern "C" int snap_hip_run_outer_convergence(snap_rect3_t b, int32_t *result,
                                              const snap_accessor3_t *fa_flux0,
                                              const snap_accessor3_t *fa_flux0po, int num_groups,
                                              double epsi, snap_stream_t stream) {
  return run_convergence(b, result, fa_flux0, fa_flux0po, num_groups, epsi, stream);
}

extern "C" int snap_hip_run_expand_cross_section(const snap_accessor1_t *sig_ptrs,
                                                 snap_accessor3_t fa_mat,
                                                 const snap_accessor3_t *fa_xs, int num_groups,
                                                 snap_rect3_t b, snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  if (!sig_ptrs || num_groups < 1) return fail(SNAP_ERR_INVALID, "no cross-section accessors");
  // one 1-D accessor per group over materials -> a group-strided 2-D view
  Acc2 sig;
  sig.base = (char *)sig_ptrs[0].base;
  sig.s0 = sig_ptrs[0].stride[0];
  sig.s1 = 0;
  sig.sg = num_groups > 1 ? (char *)sig_ptrs[1].base - (char *)sig_ptrs[0].base : 0;
  for (int g = 0; g < num_groups; g++) {
    if (!sig_ptrs[g].base || sig_ptrs[g].stride[0] != sig.s0 ||
        (char *)sig_ptrs[g].base - (char *)sig_ptrs[0].base != sig.sg * g)
      return fail(SNAP_ERR_UNSUPPORTED, "cross-section group fields are not equally spaced");
  }
  Acc3 mat, xs;
  SNAP_TRY(single_acc3(fa_mat, b, &mat, "mat"));
  SNAP_TRY(group_acc3(fa_xs, num_groups, b, &xs, "xs"));
  return launch_expand_xs(sig, mat, xs, e, num_groups, (hipStream_t)stream);
}

extern "C" int snap_hip_run_expand_scattering_cross_section(
    const snap_accessor2_t *fa_slgg, snap_accessor3_t fa_mat, const snap_accessor3_t *fa_xs,
    int num_groups, snap_rect3_t b, int group_start, int record_doubles, snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  if (num_groups < 1 || group_start < 0) return fail(SNAP_ERR_INVALID, "bad group range");
  if (record_doubles < 1 || record_doubles > 16)
    return fail(SNAP_ERR_INVALID, "record_doubles must be 1..16");
  Acc2 sl;
  Acc3 mat, xs;
  SNAP_TRY(group_acc2(fa_slgg, num_groups, &sl, "slgg"));
  SNAP_TRY(single_acc3(fa_mat, b, &mat, "mat"));
  SNAP_TRY(group_acc3(fa_xs, num_groups, b, &xs, "s_xs"));
  return launch_expand_scat(sl, group_start, mat, xs, e, num_groups, record_doubles,
                            (hipStream_t)stream);
}

extern "C" int snap_hip_run_geometry_param(const snap_accessor3_t *fa_xs,
                                           const snap_accessor3_t *fa_dinv, const double *vdelts,
                                           int num_groups, double hi, double hj, double hk,
                                           snap_rect3_t b, int num_angles, snap_stream_t stream) {
  Extent3 e;
  SNAP_TRY(rect_extent(b, &e));
  if (num_groups < 1) return fail(SNAP_ERR_INVALID, "num_groups must be >= 1");
  int dev;
  SNAP_HIP_TRY(hipGetDevice(&dev));
  const QuadDev *q = context_quad(dev);
  if (!q) return fail(SNAP_ERR_NO_CONTEXT, "initialize_gpu_context was not called on device %d", dev);
  Acc3 xs, dinv;
  SNAP_TRY(group_acc3(fa_xs, num_groups, b, &xs, "xs"));
  SNAP_TRY(group_acc3(fa_dinv, num_groups, b, &dinv, "dinv"));
  return launch_geometry(xs, dinv, vdelts, e, num_groups, num_angles, hi, hj, hk, *q, false,
                         (hipStream_t)stream);
}
