// snap_common.hip -- errors, accessor conversion and the device quadrature.
#include <cmath>
#include <cstdarg>
#include <cstring>

#include <algorithm>
#include <mutex>
#include <tuple>
#include <vector>

#include "snap_internal.h"

namespace snap {

static thread_local char g_last_error[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

// Dynamic LDS above 64 KiB must be requested per kernel and per device (the
// attribute belongs to the kernel's code object on the current device); a
// process may drive several devices, so requests are kept per (device, kernel).
int ensure_dynamic_lds(const void *kernel, size_t bytes) {
  int dev = 0;
  SNAP_HIP_TRY(hipGetDevice(&dev));
  static std::mutex mu;
  static std::vector<std::tuple<int, const void *, size_t>> done;
  std::lock_guard<std::mutex> lock(mu);
  for (auto &d : done)
    if (std::get<0>(d) == dev && std::get<1>(d) == kernel && std::get<2>(d) >= bytes) return SNAP_OK;
  SNAP_HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  for (auto &d : done)
    if (std::get<0>(d) == dev && std::get<1>(d) == kernel) {
      std::get<2>(d) = bytes;
      return SNAP_OK;
    }
  done.emplace_back(dev, kernel, bytes);
  return SNAP_OK;
}

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int rect_extent(const snap_rect3_t &b, Extent3 *e) {
  for (int i = 0; i < 3; i++)
    if (b.hi[i] < b.lo[i]) return fail(SNAP_ERR_INVALID, "empty subgrid bounds on axis %d", i);
  e->nx = (int)(b.hi[0] - b.lo[0] + 1);
  e->ny = (int)(b.hi[1] - b.lo[1] + 1);
  e->nz = (int)(b.hi[2] - b.lo[2] + 1);
  return SNAP_OK;
}

static char *fold_origin(void *base, const int64_t *stride, const int64_t *lo) {
  return (char *)base + lo[0] * stride[0] + lo[1] * stride[1] + lo[2] * stride[2];
}

int single_acc3(const snap_accessor3_t &acc, const snap_rect3_t &bounds, Acc3 *out,
                const char *what, bool allow_null) {
  if (!acc.base && !allow_null) return fail(SNAP_ERR_INVALID, "%s accessor has a NULL base", what);
  out->base = acc.base ? fold_origin(acc.base, acc.stride, bounds.lo) : nullptr;
  out->s0 = acc.stride[0];
  out->s1 = acc.stride[1];
  out->s2 = acc.stride[2];
  out->sg = 0;
  return SNAP_OK;
}

int group_acc3(const snap_accessor3_t *acc, int n, const snap_rect3_t &bounds, Acc3 *out,
               const char *what) {
  if (!acc || n < 1) return fail(SNAP_ERR_INVALID, "%s: no accessors", what);
  SNAP_TRY(single_acc3(acc[0], bounds, out, what));
  if (n > 1) {
    const int64_t sg = (char *)acc[1].base - (char *)acc[0].base;
    for (int g = 1; g < n; g++) {
      if (!acc[g].base) return fail(SNAP_ERR_INVALID, "%s[%d] has a NULL base", what, g);
      for (int i = 0; i < 3; i++)
        if (acc[g].stride[i] != acc[0].stride[i])
          return fail(SNAP_ERR_INVALID, "%s: group %d strides differ from group 0", what, g);
      if ((char *)acc[g].base - (char *)acc[0].base != sg * g)
        return fail(SNAP_ERR_UNSUPPORTED,
                    "%s: group fields are not equally spaced in one instance", what);
    }
    out->sg = sg;
  }
  return SNAP_OK;
}

int group_acc2(const snap_accessor2_t *acc, int n, Acc2 *out, const char *what) {
  if (!acc || n < 1 || !acc[0].base) return fail(SNAP_ERR_INVALID, "%s: no accessors", what);
  out->base = (char *)acc[0].base;
  out->s0 = acc[0].stride[0];
  out->s1 = acc[0].stride[1];
  out->sg = 0;
  if (n > 1) {
    const int64_t sg = (char *)acc[1].base - (char *)acc[0].base;
    for (int g = 1; g < n; g++) {
      if (!acc[g].base) return fail(SNAP_ERR_INVALID, "%s[%d] has a NULL base", what, g);
      if (acc[g].stride[0] != acc[0].stride[0] || acc[g].stride[1] != acc[0].stride[1])
        return fail(SNAP_ERR_INVALID, "%s: group %d strides differ from group 0", what, g);
      if ((char *)acc[g].base - (char *)acc[0].base != sg * g)
        return fail(SNAP_ERR_UNSUPPORTED,
                    "%s: group fields are not equally spaced in one instance", what);
    }
    out->sg = sg;
  }
  return SNAP_OK;
}

// compute_derived_globals, snap.cc:1171-1342 (3-D branch): mu accumulates by
// dm, eta counts down from 1 - dm/2, xi = sqrt(1 - (mu^2 + eta^2)),
// w = 0.125/nang, ec[oct][mom][a] = (is mu)^(2l-1) (ks xi js eta)^m with the
// octant stride cmom*nang.
int quad_compute(int nang, int nmom, std::vector<double> &mu, std::vector<double> &eta,
                 std::vector<double> &xi, std::vector<double> &w, std::vector<double> &ec) {
  if (nang < 1 || nmom < 1 || nmom > 4)
    return fail(SNAP_ERR_INVALID, "quadrature needs nang >= 1 and 1 <= nmom <= 4");
  const int cmom = nmom * nmom;
  mu.assign(nang, 0.0);
  eta.assign(nang, 0.0);
  xi.assign(nang, 0.0);
  w.assign(nang, 0.0);
  ec.assign((size_t)8 * cmom * nang, 0.0);
  const double dm = 1.0 / double(nang);
  mu[0] = 0.5 * dm;
  for (int i = 1; i < nang; i++) mu[i] = mu[i - 1] + dm;
  eta[0] = 1.0 - 0.5 * dm;
  for (int i = 1; i < nang; i++) eta[i] = eta[i - 1] - dm;
  for (int i = 0; i < nang; i++) {
    const double t = mu[i] * mu[i] + eta[i] * eta[i];
    xi[i] = std::sqrt(1.0 - t);
  }
  for (int i = 0; i < nang; i++) w[i] = 0.125 / double(nang);
  for (int kd = 0; kd < 2; kd++) {
    const int ks = kd ? 1 : -1;
    for (int jd = 0; jd < 2; jd++) {
      const int js = jd ? 1 : -1;
      for (int id = 0; id < 2; id++) {
        const int is = id ? 1 : -1;
        const int oct = 4 * kd + 2 * jd + id;
        double *e = ec.data() + (size_t)oct * cmom * nang;
        for (int a = 0; a < nang; a++) e[a] = 1.0;
        int moment = 1;
        for (int l = 1; l < nmom; l++) {
          for (int m = 0; m < 2 * (l + 1) - 1; m++) {
            for (int a = 0; a < nang; a++)
              e[moment * nang + a] = std::pow(is * mu[a], double(2 * (l + 1) - 3)) *
                                     std::pow(ks * xi[a] * js * eta[a], double(m));
            moment++;
          }
        }
      }
    }
  }
  return SNAP_OK;
}

int quad_upload(QuadDev *q, const double *ec, const double *mu, const double *eta,
                const double *xi, const double *w, int nang, int nmom) {
  quad_free(q);
  q->nang = nang;
  q->nmom = nmom;
  q->cmom = nmom * nmom;
  const size_t an = sizeof(double) * nang;
  const size_t en = sizeof(double) * 8 * q->cmom * nang;
  SNAP_HIP_TRY(hipMalloc((void **)&q->mu, an));
  SNAP_HIP_TRY(hipMalloc((void **)&q->eta, an));
  SNAP_HIP_TRY(hipMalloc((void **)&q->xi, an));
  SNAP_HIP_TRY(hipMalloc((void **)&q->w, an));
  SNAP_HIP_TRY(hipMalloc((void **)&q->ec, en));
  SNAP_HIP_TRY(hipMemcpy(q->mu, mu, an, hipMemcpyHostToDevice));
  SNAP_HIP_TRY(hipMemcpy(q->eta, eta, an, hipMemcpyHostToDevice));
  SNAP_HIP_TRY(hipMemcpy(q->xi, xi, an, hipMemcpyHostToDevice));
  SNAP_HIP_TRY(hipMemcpy(q->w, w, an, hipMemcpyHostToDevice));
  SNAP_HIP_TRY(hipMemcpy(q->ec, ec, en, hipMemcpyHostToDevice));
  return SNAP_OK;
}

void quad_free(QuadDev *q) {
  if (q->mu) (void)hipFree(q->mu);
  if (q->eta) (void)hipFree(q->eta);
  if (q->xi) (void)hipFree(q->xi);
  if (q->w) (void)hipFree(q->w);
  if (q->ec) (void)hipFree(q->ec);
  *q = QuadDev();
}

}  // namespace snap

extern "C" const char *snap_hip_last_error(void) { return snap::g_last_error; }
extern "C" const char *snap_hip_version(void) { return "snap_hip 0.1 (gfx950)"; }

extern "C" int snap_hip_quadrature(int num_angles, int num_moments, double *mu, double *eta,
                                   double *xi, double *w, double *ec) {
  if (!mu || !eta || !xi || !w || !ec) return snap::fail(SNAP_ERR_INVALID, "NULL quadrature array");
  std::vector<double> m, e, x, ww, c;
  SNAP_TRY(snap::quad_compute(num_angles, num_moments, m, e, x, ww, c));
  std::copy(m.begin(), m.end(), mu);
  std::copy(e.begin(), e.end(), eta);
  std::copy(x.begin(), x.end(), xi);
  std::copy(ww.begin(), ww.end(), w);
  std::copy(c.begin(), c.end(), ec);
  return SNAP_OK;
}
