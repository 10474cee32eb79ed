"""Host-side mirror of the reference's GPU task bodies over the C ABI.

Each function here does what the matching ``gpu_implementation`` of the
reference does before its ``run_*`` call: build one affine accessor per
energy-group field and hand them to the library.  Device memory is owned by
the caller (torch tensors on the GPU standing in for Legion instances); each
tensor holds all group fields of one logical region in SOA order, shaped
``[ng, nz, ny, nx]`` or ``[ng, nz, ny, nx, record]``.

* initialize_gpu_context       InitGPUSweep::gpu_implementation      init.cc:224-236
* mini_kba                     MiniKBATask::gpu_implementation       sweep.cc:1550-1620
* calc_outer_source            CalcOuterSource::gpu_implementation   outer.cc:235-284
* calc_inner_source            CalcInnerSource::gpu_implementation   inner.cc:178-214
* inner_convergence            TestInnerConvergence::gpu_implementation inner.cc:335-367
* outer_convergence            TestOuterConvergence::gpu_implementation outer.cc:414-451
* expand_cross_section         ExpandCrossSection::gpu_implementation expxs.cc:125-156
* expand_scattering_cross_section  ExpandScatteringCrossSection::gpu_implementation expxs.cc:261-293
* calculate_geometry_param     CalculateGeometryParam::gpu_implementation expxs.cc:402-439
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib

LMA = (ctypes.c_int * 4)(1, 3, 5, 7)  # snap.cc:1307-1308 (3-D)


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _esize(t) -> int:
    return t.element_size()


def acc3(t, g: Optional[int] = None) -> _lib.Accessor3:
    """Accessor of group field g of t [ng, nz, ny, nx(, rec)], or of t
    [nz, ny, nx(, rec)] when g is None; points are (x, y, z)."""
    a = _lib.Accessor3()
    e = _esize(t)
    if g is None:
        a.base = t.data_ptr()
        a.stride[0], a.stride[1], a.stride[2] = t.stride(2) * e, t.stride(1) * e, t.stride(0) * e
    else:
        a.base = t.data_ptr() + g * t.stride(0) * e
        a.stride[0], a.stride[1], a.stride[2] = t.stride(3) * e, t.stride(2) * e, t.stride(1) * e
    return a


def acc3_groups(t) -> ctypes.Array:
    return (_lib.Accessor3 * t.shape[0])(*[acc3(t, g) for g in range(t.shape[0])])


def slgg_accessors(t) -> ctypes.Array:
    """slgg [ng_to, nmat, ng_from, nmom]: field g_to, point (mat, g_from)."""
    e = _esize(t)
    out = (_lib.Accessor2 * t.shape[0])()
    for g in range(t.shape[0]):
        out[g].base = t.data_ptr() + g * t.stride(0) * e
        out[g].stride[0], out[g].stride[1] = t.stride(1) * e, t.stride(2) * e
    return out


def rect(nx, ny, nz, origin=(0, 0, 0)) -> _lib.Rect3:
    r = _lib.Rect3()
    for i, n in enumerate((nx, ny, nz)):
        r.lo[i] = origin[i]
        r.hi[i] = origin[i] + n - 1
    return r


def initialize_gpu_context(quad, nang: int, nmom: int, chunk) -> None:
    mu, eta, xi, w, ec = (np.ascontiguousarray(x, dtype=np.float64) for x in quad)
    dp = ctypes.POINTER(ctypes.c_double)
    lib = _lib.load()
    _lib.check(lib.snap_hip_initialize_gpu_context(
        ec.ctypes.data_as(dp), mu.ctypes.data_as(dp), eta.ctypes.data_as(dp),
        xi.ctypes.data_as(dp), w.ctypes.data_as(dp), nang, nmom, 8, *chunk),
        "snap_hip_initialize_gpu_context")


def quadrature(nang: int, nmom: int):
    """compute_derived_globals (snap.cc:1171-1342): mu, eta, xi, w, ec[8][cmom][nang]."""
    cmom = nmom * nmom
    mu, eta, xi, w = (np.zeros(nang) for _ in range(4))
    ec = np.zeros((8, cmom, nang))
    dp = ctypes.POINTER(ctypes.c_double)
    _lib.check(_lib.load().snap_hip_quadrature(nang, nmom, *(a.ctypes.data_as(dp) for a in
                                                               (mu, eta, xi, w, ec))),
               "snap_hip_quadrature")
    return mu, eta, xi, w, ec


def mini_kba(corner: int, groups: Sequence[int], qtot, flux, fluxm, dinv, t_xs,
             ghosts, hijk, nang: int, nmom: int, fixup: bool = True, origin=(0, 0, 0),
             vdelt=None, time_flux_in=None, time_flux_out=None, qim=None, extent=None,
             noalloc: bool = False, group_range: Optional[str] = None) -> None:
    """One corner over a group range: run_gpu_sweep once per group.
    ghosts: (gx [ng, nz, ny, nang], gy [ng, nz, nx, nang], gz [ng, ny, nx, nang]) or None,
    indexed by absolute points.  extent: (nx, ny, nz) of the chunk at `origin`
    (default: the whole of flux), whose arrays are indexed by absolute points.
    Time dependence (MiniKBATask regions 5, 6, 11, sweep.cc:45-84): vdelt
    [ng] host values, time_flux_in / time_flux_out [ng, nz, ny, nx, nang] of
    this corner; qim [ng, nz, ny, nx, nang] is the corner's MMS source.
    noalloc: call the allocation-free entry point (scratch array, no device
    allocation) instead of run_gpu_sweep.  group_range: "exact" or "fused" --
    one snap_hip_run_gpu_sweep_groups call for the whole (contiguous) group
    range instead of one run_gpu_sweep per group."""
    lib = _lib.load()
    nx, ny, nz = extent if extent is not None else flux.shape[1:4][::-1]
    groups = list(groups)
    batch = (_lib.SweepArgs * len(groups))() if group_range else None
    for gi, g in enumerate(groups):
        a = _lib.SweepArgs()
        a.origin[0], a.origin[1], a.origin[2] = origin
        a.qtot = acc3(qtot, g)
        a.flux = acc3(flux, g)
        if fluxm is not None:
            a.fluxm = acc3(fluxm, g)
        a.dinv = acc3(dinv, g)
        a.t_xs = acc3(t_xs, g)
        if ghosts is not None:
            gx, gy, gz = ghosts
            e = 8
            # ghost x at point (y, z); y at (x, z); z at (x, y) (sweep.cc:218-240)
            for field, t in (("ghostx", gx), ("ghosty", gy), ("ghostz", gz)):
                acc = _lib.Accessor2()
                acc.base = t.data_ptr() + g * t.stride(0) * e
                acc.stride[0], acc.stride[1] = t.stride(2) * e, t.stride(1) * e
                setattr(a, field, acc)
        a.x_range, a.y_range, a.z_range = nx, ny, nz
        a.corner = corner
        a.stride_x_positive = int(bool(corner & 1))
        a.stride_y_positive = int(bool(corner & 2))
        a.stride_z_positive = int(bool(corner & 4))
        a.num_moments = nmom
        a.hi, a.hj, a.hk = hijk
        a.vdelt = 0.0 if vdelt is None else float(vdelt[g])
        if a.vdelt != 0.0:
            a.time_flux_in = acc3(time_flux_in, g)
            a.time_flux_out = acc3(time_flux_out, g)
        if qim is not None:
            a.qim = acc3(qim, g)
            a.mms_source = 1
        a.num_angles = nang
        a.fixup = int(bool(fixup))
        if batch is not None:
            batch[gi] = a
            continue
        if noalloc:
            _lib.check(lib.snap_hip_run_gpu_sweep_noalloc(ctypes.byref(a), _stream()),
                       "snap_hip_run_gpu_sweep_noalloc")
        else:
            _lib.check(lib.snap_hip_run_gpu_sweep(ctypes.byref(a), _stream()),
                       "snap_hip_run_gpu_sweep")
    if batch is not None:
        prec = {"exact": _lib.PRECISION_FP64_EXACT, "fused": _lib.PRECISION_FP64}[group_range]
        _lib.check(lib.snap_hip_run_gpu_sweep_groups(batch, len(groups), prec, _stream()),
                   "snap_hip_run_gpu_sweep_groups")


def calc_outer_source(qi, flux0, slgg, qo0, mat, fluxm=None, qom=None, nmom: int = 1) -> None:
    lib = _lib.load()
    ng, nz, ny, nx = flux0.shape[:4]
    r = rect(nx, ny, nz)
    sl = slgg_accessors(slgg)
    _lib.check(lib.snap_hip_run_flux0_outer_source(r, acc3_groups(qi), acc3_groups(flux0), sl,
                                                   acc3_groups(qo0), acc3(mat), ng, _stream()),
               "snap_hip_run_flux0_outer_source")
    if nmom > 1:
        _lib.check(lib.snap_hip_run_fluxm_outer_source(r, acc3_groups(fluxm), sl, acc3_groups(qom),
                                                       acc3(mat), ng, nmom, LMA, _stream()),
                   "snap_hip_run_fluxm_outer_source")


def calc_inner_source(s_xs, flux0, q2grp0, qtot, fluxm=None, q2grpm=None, nmom: int = 1) -> None:
    lib = _lib.load()
    ng, nz, ny, nx = flux0.shape[:4]
    r = rect(nx, ny, nz)
    for g in range(ng):
        if nmom > 1:
            rc = lib.snap_hip_run_inner_source_multi_moment(
                r, acc3(s_xs, g), acc3(flux0, g), acc3(q2grp0, g), acc3(fluxm, g),
                acc3(q2grpm, g), acc3(qtot, g), nmom, LMA, _stream())
        else:
            rc = lib.snap_hip_run_inner_source_single_moment(
                r, acc3(s_xs, g), acc3(flux0, g), acc3(q2grp0, g), acc3(qtot, g), _stream())
        _lib.check(rc, "snap_hip_run_inner_source")


def _convergence(fn, name, flux0, prev, epsi, result) -> None:
    ng, nz, ny, nx = flux0.shape[:4]
    _lib.check(fn(rect(nx, ny, nz), ctypes.c_void_p(result.data_ptr()), acc3_groups(flux0),
                  acc3_groups(prev), ng, epsi, _stream()), name)


def inner_convergence(flux0, flux0pi, epsi, result) -> None:
    lib = _lib.load()
    _convergence(lib.snap_hip_run_inner_convergence, "snap_hip_run_inner_convergence", flux0,
                 flux0pi, epsi, result)


def outer_convergence(flux0, flux0po, epsi, result) -> None:
    """epsi here is the outer tolerance, 100 * convergence_eps (outer.cc:439)."""
    lib = _lib.load()
    _convergence(lib.snap_hip_run_outer_convergence, "snap_hip_run_outer_convergence", flux0,
                 flux0po, epsi, result)


def expand_cross_section(sig, mat, xs) -> None:
    """sig [nmat, ng] on the device: field g, point (material)."""
    lib = _lib.load()
    ng, nz, ny, nx = xs.shape[:4]
    e = _esize(sig)
    accs = (_lib.Accessor1 * ng)()
    for g in range(ng):
        accs[g].base = sig.data_ptr() + g * sig.stride(1) * e
        accs[g].stride[0] = sig.stride(0) * e
    _lib.check(lib.snap_hip_run_expand_cross_section(accs, acc3(mat), acc3_groups(xs), ng,
                                                     rect(nx, ny, nz), _stream()),
               "snap_hip_run_expand_cross_section")


def expand_scattering_cross_section(slgg, mat, s_xs, group_start: int = 0) -> None:
    lib = _lib.load()
    ng, nz, ny, nx = s_xs.shape[:4]
    sl = slgg_accessors(slgg[group_start:group_start + ng])
    _lib.check(lib.snap_hip_run_expand_scattering_cross_section(
        sl, acc3(mat), acc3_groups(s_xs), ng, rect(nx, ny, nz), group_start, s_xs.shape[4],
        _stream()), "snap_hip_run_expand_scattering_cross_section")


def calculate_geometry_param(t_xs, dinv, hijk, vdelts=None) -> None:
    lib = _lib.load()
    ng, nz, ny, nx = t_xs.shape[:4]
    nang = dinv.shape[4]
    vd = np.zeros(ng) if vdelts is None else np.ascontiguousarray(vdelts, dtype=np.float64)
    _lib.check(lib.snap_hip_run_geometry_param(
        acc3_groups(t_xs), acc3_groups(dinv), vd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        ng, hijk[0], hijk[1], hijk[2], rect(nx, ny, nz), nang, _stream()),
        "snap_hip_run_geometry_param")
