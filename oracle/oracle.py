"""ctypes front end of the CPU oracle (libsnap_oracle.so).

TEST INFRASTRUCTURE ONLY: the HIP library's results are tested against it.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
Each wrapper names the reference task body its C function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsnap_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class OracleProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("ng", ctypes.c_int), ("nang", ctypes.c_int), ("nmom", ctypes.c_int),
        ("nmat", ctypes.c_int),
        ("mat", _ip), ("sigt", _dp), ("slgg", _dp), ("qi", _dp), ("sigt_cell", _dp),
        ("epsi", ctypes.c_double), ("iitm", ctypes.c_int), ("oitm", ctypes.c_int),
        ("fixup", ctypes.c_int), ("timedep", ctypes.c_int), ("tf", ctypes.c_double),
        ("nsteps", ctypes.c_int), ("mms", ctypes.c_int), ("sweep_order", ctypes.c_int),
    ]


# operation orders of the sweep body: the reference's generic CPU variant
# (sweep.cc:242-616) and its AVX variant (sweep.cc:1037-1520, nang % 4 == 0)
ORDERS = {"generic": 0, "avx": 1}


class OracleStats(ctypes.Structure):
    _fields_ = [
        ("inner_total", ctypes.c_int), ("outer_total", ctypes.c_int),
        ("converged", ctypes.c_int), ("fixup_cells", ctypes.c_int),
        ("sweep_seconds", ctypes.c_double), ("solve_seconds", ctypes.c_double),
        ("time_steps", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def build() -> None:
    """Compile the oracle with its committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.snap_oracle_quadrature.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        lib.snap_oracle_quadrature.restype = ctypes.c_int
        lib.snap_oracle_expand_xs.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp]
        lib.snap_oracle_expand_scat_xs.argtypes = [ctypes.c_int] * 4 + [_ip, _dp, _dp]
        lib.snap_oracle_geometry_param.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        lib.snap_oracle_outer_source.argtypes = [ctypes.c_int] * 4 + [_ip] + [_dp] * 6
        lib.snap_oracle_inner_source.argtypes = [ctypes.c_int] * 3 + [_dp] * 6
        lib.snap_oracle_converged.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double]
        lib.snap_oracle_converged.restype = ctypes.c_int
        lib.snap_oracle_sweep.argtypes = (
            [ctypes.c_int] * 6 + [ctypes.c_double] * 3 + [_dp] * 8 + [ctypes.c_int] + [_dp] * 8)
        lib.snap_oracle_sweep.restype = ctypes.c_int
        lib.snap_oracle_sweep_avx.argtypes = lib.snap_oracle_sweep.argtypes
        lib.snap_oracle_sweep_avx.restype = ctypes.c_int
        lib.snap_oracle_sweep_td.argtypes = (
            [ctypes.c_int] * 6 + [ctypes.c_double] * 3 + [_dp] * 8 + [ctypes.c_int]
            + [ctypes.c_double] + [_dp] * 8)
        lib.snap_oracle_sweep_td.restype = ctypes.c_int
        lib.snap_oracle_mms_setup.argtypes = [ctypes.POINTER(OracleProblem), _dp, _dp, _dp]
        lib.snap_oracle_mms_setup.restype = ctypes.c_int
        lib.snap_oracle_mms_compare.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        lib.snap_oracle_mms_compare.restype = None
        lib.snap_oracle_solve.argtypes = [ctypes.POINTER(OracleProblem), _dp, _dp,
                                          ctypes.POINTER(OracleStats), ctypes.c_int]
        lib.snap_oracle_solve.restype = ctypes.c_int
        lib.snap_oracle_adjoint_sweep.argtypes = (
            [ctypes.c_int] * 5 + [ctypes.c_double] * 3 + [_dp] * 6 + [ctypes.c_int] + [_dp] * 3)
        lib.snap_oracle_adjoint_sweep.restype = None
        lib.snap_oracle_adjoint.argtypes = [ctypes.POINTER(OracleProblem), _dp, _dp, ctypes.c_double,
                                            ctypes.c_int, _dp, _dp, ctypes.c_int]
        lib.snap_oracle_adjoint.restype = ctypes.c_int
        lib.snap_oracle_time_sweeps.argtypes = [ctypes.POINTER(OracleProblem), ctypes.c_int,
                                                ctypes.c_int]
        lib.snap_oracle_time_sweeps.restype = ctypes.c_double
        _lib = lib
    return _lib


def _p(a: Optional[np.ndarray], ptr=_dp):
    if a is None:
        return ctypes.cast(None, ptr)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ptr)


def quadrature(nang: int, nmom: int):
    """compute_derived_globals (snap.cc:1171-1342): mu, eta, xi, w, ec[8][cmom][nang]."""
    lib = load()
    cmom = nmom * nmom
    mu, eta, xi, w = (np.zeros(nang) for _ in range(4))
    ec = np.zeros((8, cmom, nang))
    rc = lib.snap_oracle_quadrature(nang, nmom, _p(mu), _p(eta), _p(xi), _p(w), _p(ec))
    if rc:
        raise ValueError("bad quadrature request")
    return mu, eta, xi, w, ec


def spacing(p):
    """hi, hj, hk = 2/dx (snap.cc:1178-1180)."""
    return 2.0 / (p.lx / p.nx), 2.0 / (p.ly / p.ny), 2.0 / (p.lz / p.nz)


def expand_xs(mat: np.ndarray, sig: np.ndarray) -> np.ndarray:
    lib = load()
    ng = sig.shape[1]
    xs = np.zeros((ng,) + mat.shape)
    lib.snap_oracle_expand_xs(mat.size, ng, _p(mat, _ip), _p(np.ascontiguousarray(sig)), _p(xs))
    return xs


def expand_scat_xs(mat: np.ndarray, slgg: np.ndarray) -> np.ndarray:
    lib = load()
    ng, nmat, _, nmom = slgg.shape
    sxs = np.zeros((ng,) + mat.shape + (nmom,))
    lib.snap_oracle_expand_scat_xs(mat.size, ng, nmat, nmom, _p(mat, _ip), _p(slgg), _p(sxs))
    return sxs


def geometry_param(t_xs: np.ndarray, nang: int, hijk, mu, eta, xi, vdelt=None) -> np.ndarray:
    lib = load()
    ng = t_xs.shape[0]
    dinv = np.zeros(t_xs.shape + (nang,))
    lib.snap_oracle_geometry_param(t_xs[0].size, ng, nang, _p(t_xs), _p(vdelt), hijk[0], hijk[1],
                                   hijk[2], _p(mu), _p(eta), _p(xi), _p(dinv))
    return dinv


def outer_source(mat, qi, slgg, flux0, fluxm=None):
    lib = load()
    ng, nmat, _, nmom = slgg.shape
    cmom = nmom * nmom
    ncells = mat.size
    q2grp0 = np.zeros_like(flux0)
    q2grpm = np.zeros(flux0.shape + (cmom - 1,)) if cmom > 1 else None
    if cmom > 1 and fluxm is None:
        fluxm = np.zeros(flux0.shape + (cmom - 1,))
    lib.snap_oracle_outer_source(ncells, ng, nmat, nmom, _p(mat, _ip), _p(qi), _p(slgg),
                                 _p(flux0), _p(fluxm), _p(q2grp0), _p(q2grpm))
    return q2grp0, q2grpm


def inner_source(s_xs, flux0, q2grp0, fluxm=None, q2grpm=None):
    lib = load()
    ng = flux0.shape[0]
    nmom = s_xs.shape[-1]
    cmom = nmom * nmom
    qtot = np.zeros(flux0.shape + (cmom,))
    lib.snap_oracle_inner_source(flux0[0].size, ng, nmom, _p(s_xs), _p(flux0), _p(q2grp0),
                                 _p(fluxm), _p(q2grpm), _p(qtot))
    return qtot


def converged(flux0, flux0p, epsi) -> bool:
    lib = load()
    return bool(lib.snap_oracle_converged(flux0.size, _p(flux0), _p(flux0p), epsi))


def sweep(nx, ny, nz, nang, cmom, corner, hijk, quad, qtot_g, dinv_g, t_xs_g, fixup=True,
          ghost_in=(None, None, None), want_ghost_out=False, order: str = "generic"):
    """One (corner, group) MiniKBA sweep (sweep.cc:242-616, or the AVX
    variant's operation order, sweep.cc:1037-1520, with order="avx").  Returns
    (part_flux [nz,ny,nx], part_fluxm [nz,ny,nx,cmom-1] or None, fixup_cells, ghost_out)."""
    lib = load()
    fn = lib.snap_oracle_sweep_avx if ORDERS[order] else lib.snap_oracle_sweep
    mu, eta, xi, w, ec = quad
    pf = np.zeros((nz, ny, nx))
    pm = np.zeros((nz, ny, nx, cmom - 1)) if cmom > 1 else None
    outs = (np.zeros((nz, ny, nang)), np.zeros((nz, nx, nang)), np.zeros((ny, nx, nang))) \
        if want_ghost_out else (None, None, None)
    n = fn(nx, ny, nz, nang, cmom, corner, hijk[0], hijk[1], hijk[2],
           _p(mu), _p(eta), _p(xi), _p(w), _p(ec), _p(qtot_g), _p(dinv_g),
           _p(t_xs_g), int(bool(fixup)), _p(ghost_in[0]), _p(ghost_in[1]),
           _p(ghost_in[2]), _p(outs[0]), _p(outs[1]), _p(outs[2]),
           _p(pf), _p(pm))
    if n < 0:
        raise ValueError("the AVX order needs nang % 4 == 0")
    return pf, pm, n, outs


def _problem_struct(p, keep, order: str = "generic"):
    q = p.contiguous()
    keep.append(q)
    s = OracleProblem()
    s.nx, s.ny, s.nz = q.nx, q.ny, q.nz
    s.lx, s.ly, s.lz = q.lx, q.ly, q.lz
    s.ng, s.nang, s.nmom, s.nmat = q.ng, q.nang, q.nmom, q.nmat
    s.mat = _p(q.mat, _ip)
    s.sigt = _p(q.sigt)
    s.slgg = _p(q.slgg)
    s.qi = _p(q.qi)
    s.sigt_cell = _p(q.sigt_cell)
    s.epsi, s.iitm, s.oitm, s.fixup = q.epsi, q.iitm, q.oitm, int(bool(q.fixup))
    s.timedep, s.tf, s.nsteps, s.mms = int(bool(q.timedep)), q.tf, q.nsteps, int(bool(q.mms))
    s.sweep_order = ORDERS[order]
    return s


def sweep_td(nx, ny, nz, nang, cmom, corner, hijk, quad, qtot_g, dinv_g, t_xs_g, fixup, vdelt,
             tflux_in, qim_g=None):
    """One (corner, group) sweep from vacuum inflow with the time-dependent
    terms (sweep.cc:414-534) and the optional MMS angular source.
    tflux_in / qim_g: [nz,ny,nx,nang].  Returns (part_flux, part_fluxm or
    None, fixup_cells, tflux_out, (ghost_x_out, ghost_y_out, ghost_z_out))."""
    lib = load()
    mu, eta, xi, w, ec = quad
    pf = np.zeros((nz, ny, nx))
    pm = np.zeros((nz, ny, nx, cmom - 1)) if cmom > 1 else None
    tout = np.zeros((nz, ny, nx, nang))
    tin = np.ascontiguousarray(tflux_in, dtype=np.float64)
    qm = None if qim_g is None else np.ascontiguousarray(qim_g, dtype=np.float64)
    outs = (np.zeros((nz, ny, nang)), np.zeros((nz, nx, nang)), np.zeros((ny, nx, nang)))
    n = lib.snap_oracle_sweep_td(nx, ny, nz, nang, cmom, corner, hijk[0], hijk[1], hijk[2],
                                 _p(mu), _p(eta), _p(xi), _p(w), _p(ec), _p(qtot_g), _p(dinv_g),
                                 _p(t_xs_g), int(bool(fixup)), float(vdelt), _p(tin), _p(tout),
                                 _p(qm), _p(outs[0]), _p(outs[1]), _p(outs[2]), _p(pf), _p(pm))
    return pf, pm, n, tout, outs


def mms_setup(p):
    """MMSInitFlux / MMSInitSource / MMSInitTimeDependent (mms.cc:81-409):
    (ref_flux [ng,nz,ny,nx], qim [8,ng,nz,ny,nx,nang] unscaled, qi [ng,nz,ny,nx])."""
    lib = load()
    keep = []
    s = _problem_struct(p, keep)
    ref = np.zeros((p.ng, p.nz, p.ny, p.nx))
    qim = np.zeros((8, p.ng, p.nz, p.ny, p.nx, p.nang))
    qi = np.zeros((p.ng, p.nz, p.ny, p.nx))
    if lib.snap_oracle_mms_setup(ctypes.byref(s), _p(ref), _p(qim), _p(qi)):
        raise ValueError("MMS needs nmom = 1")
    return ref, qim, qi


def mms_compare(flux0, ref):
    """MMSCompare (mms.cc:497-541): (max, min, mean) of |flux/ref - 1|."""
    lib = load()
    f = np.ascontiguousarray(flux0, dtype=np.float64)
    r = np.ascontiguousarray(ref, dtype=np.float64)
    out = np.zeros(3)
    lib.snap_oracle_mms_compare(f.size, _p(f), _p(r), _p(out))
    return float(out[0]), float(out[1]), float(out[2])


def solve(p, nthreads: int = 0, order: str = "generic"):
    """transport_solve (snap.cc:386-541), every time step, with the sweeps in
    the generic or the AVX variant's operation order.  Returns (flux0
    [ng,nz,ny,nx], fluxm [ng,nz,ny,nx,cmom-1] or None, stats dict)."""
    lib = load()
    p.validate()
    keep = []
    s = _problem_struct(p, keep, order)
    flux0 = np.zeros((p.ng, p.nz, p.ny, p.nx))
    fluxm = np.zeros((p.ng, p.nz, p.ny, p.nx, p.cmom - 1)) if p.cmom > 1 else None
    st = OracleStats()
    rc = lib.snap_oracle_solve(ctypes.byref(s), _p(flux0), _p(fluxm), ctypes.byref(st), nthreads)
    if rc:
        raise RuntimeError("oracle solve failed")
    return flux0, fluxm, st.as_dict()


def adjoint_sweep(nx, ny, nz, nang, corner, h, quad, qtot_g, t_xs_g, fixup, zeta_g):
    """One reverse sweep (nmom = 1): returns (u_g, grad_g) for the direct
    response w_a * zeta_g of pc.  quad = (mu, eta, xi, w)."""
    lib = load()
    mu, eta, xi, w = (np.ascontiguousarray(a, dtype=np.float64) for a in quad)
    q = np.ascontiguousarray(qtot_g, dtype=np.float64)
    t = np.ascontiguousarray(t_xs_g, dtype=np.float64)
    z = np.ascontiguousarray(zeta_g, dtype=np.float64)
    u = np.zeros(nx * ny * nz)
    g = np.zeros(nx * ny * nz)
    lib.snap_oracle_adjoint_sweep(nx, ny, nz, nang, corner, h[0], h[1], h[2], _p(mu), _p(eta),
                                  _p(xi), _p(w), _p(q), _p(t), int(fixup), _p(z), _p(u), _p(g))
    return u, g


def adjoint(p, flux0, resp, tol: float = 1e-13, maxit: int = 500, nthreads: int = 0):
    """FP64 reference gradient dR/dt_xs[g, z, y, x] of R = sum(resp * flux0) at
    the converged flux0 of p (nmom = 1), by adjoint source iteration.
    Returns (grad, zeta, iterations)."""
    lib = load()
    p.validate()
    keep = []
    s = _problem_struct(p, keep)
    f = np.ascontiguousarray(flux0, dtype=np.float64)
    r = np.ascontiguousarray(resp, dtype=np.float64)
    grad = np.zeros((p.ng, p.nz, p.ny, p.nx))
    zeta = np.zeros((p.ng, p.nz, p.ny, p.nx))
    it = lib.snap_oracle_adjoint(ctypes.byref(s), _p(f), _p(r), tol, maxit, _p(grad), _p(zeta),
                                 nthreads)
    if it < 0:
        raise RuntimeError("oracle adjoint did not converge")
    return grad, zeta, it


def time_sweeps(p, ninner: int = 1, nthreads: int = 0, order: str = "generic") -> float:
    """Seconds for `ninner` inner iterations' sweeps (all octants and groups)."""
    lib = load()
    keep = []
    s = _problem_struct(p, keep, order)
    t = lib.snap_oracle_time_sweeps(ctypes.byref(s), ninner, nthreads)
    if t < 0:
        raise RuntimeError("oracle timing failed")
    return t
