"""The reference's transport_solve (snap.cc:225-568) with its chunk
decomposition, driven on one GPU through the drop-in task entry points only.

Every step is the run_* call the reference's GPU task variant makes, on torch
tensors standing in for the Legion region instances (one tensor per logical
region, all group fields in SOA order):

* ExpandCrossSection, ExpandScatteringCrossSection, CalculateGeometryParam
  once (snap.cc:419-450; the reference repeats them each time step with the
  same result);
* per outer iteration CalcOuterSource and the flux0 -> flux0po save; per
  inner iteration CalcInnerSource, the flux0 -> flux0pi save, flux0 (and the
  flux moments, which the reference leaves accumulating) zeroed;
* perform_sweeps (snap.cc:849-887): the face regions flux_xy, flux_yz and
  flux_xz zeroed (vacuum at the first chunk), then for each corner one
  MiniKBA sweep per chunk and group (run_gpu_sweep), the chunks of each
  column taken in the order the flux projection functors give Legion
  (snap.cc:1711-1790: increasing along an axis whose corner bit is set,
  decreasing otherwise).  A chunk reads its column's face subregion at its
  upstream side and overwrites it with its downstream outflow, so the next
  chunk of the column reads it as inflow -- the face exchange a multi-GPU run
  does between devices;
* the inner and outer convergence tests (run_inner_convergence,
  run_outer_convergence) with the iteration control of snap.cc:468-540.

Face regions hold nang doubles per face point and (group, corner) field
(flux_fs, snap.cc:133-156), stored as [8][ng][n1][n2][nang] tensors.
Time-dependent problems keep the even/odd time-edge fluxes per corner
(snap.cc:288-298).  The manufactured-solution source is built by the reference
with CPU tasks and is not part of this driver (the Solver builds it).
"""
from __future__ import annotations

from typing import Tuple

import numpy as np

from . import tasks
from .problems import Problem


class KBASolve:
    """One problem split into chunks[0] x chunks[1] x chunks[2] chunks along x,
    y, z (the decks' ichunk, npey, npez), solved on the current GPU."""

    def __init__(self, problem: Problem, chunks: Tuple[int, int, int] = (1, 1, 1)):
        import torch
        problem.validate()
        if problem.mms:
            raise ValueError("the chunked driver takes fixed sources (the Solver builds MMS sources)")
        p = self.p = problem.contiguous()
        cx, cy, cz = chunks
        for c, n, name in ((cx, p.nx, "x"), (cy, p.ny, "y"), (cz, p.nz, "z")):
            if c < 1 or n % c:
                raise ValueError(f"{c} chunks do not divide the {name} edge {n}")
        self.chunks = chunks
        self.chunk = (p.nx // cx, p.ny // cy, p.nz // cz)
        self.h = (2.0 / (p.lx / p.nx), 2.0 / (p.ly / p.ny), 2.0 / (p.lz / p.nz))
        self.quad = tasks.quadrature(p.nang, p.nmom)
        tasks.initialize_gpu_context(self.quad, p.nang, p.nmom, self.chunk)
        dev = torch.device("cuda")
        f64 = dict(dtype=torch.float64, device=dev)
        T = lambda a, dt=torch.float64: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)  # noqa: E731
        ng, nz, ny, nx, nang, nmom, cmom = p.ng, p.nz, p.ny, p.nx, p.nang, p.nmom, p.cmom
        self.torch = torch
        self.mat = T(p.mat, torch.int32)
        self.sigt = T(p.sigt)
        self.slgg = T(p.slgg)
        self.qi = T(p.qi)
        grid = (ng, nz, ny, nx)
        self.t_xs = torch.zeros(grid, **f64)
        self.s_xs = torch.zeros(grid + (nmom,), **f64)
        self.dinv = torch.zeros(grid + (nang,), **f64)
        self.flux0 = torch.zeros(grid, **f64)
        self.flux0po = torch.zeros(grid, **f64)
        self.flux0pi = torch.zeros(grid, **f64)
        self.q2grp0 = torch.zeros(grid, **f64)
        self.qtot = torch.zeros(grid + (cmom,), **f64)
        m1 = max(cmom - 1, 1)
        self.fluxm = torch.zeros(grid + (m1,), **f64)
        self.q2grpm = torch.zeros(grid + (m1,), **f64)
        # face regions: yz (x faces), xz (y faces), xy (z faces) per corner
        self.fyz = torch.zeros((8, ng, nz, ny, nang), **f64)
        self.fxz = torch.zeros((8, ng, nz, nx, nang), **f64)
        self.fxy = torch.zeros((8, ng, ny, nx, nang), **f64)
        self.flag = torch.zeros(1, dtype=torch.int32, device=dev)
        # Snap::initialize_velocity (snap.cc:716-738)
        dt = p.tf / p.nsteps
        self.vdelt = np.array([2.0 / (dt * (ng - g)) if p.timedep else 0.0 for g in range(ng)])
        self.tflux = torch.zeros((2, 8) + grid + (nang,), **f64) if p.timedep else None
        # snap.cc:419-450
        if p.sigt_cell is not None:
            self.t_xs.copy_(T(p.sigt_cell))
        else:
            tasks.expand_cross_section(self.sigt, self.mat, self.t_xs)
        tasks.expand_scattering_cross_section(self.slgg, self.mat, self.s_xs)
        tasks.calculate_geometry_param(self.t_xs, self.dinv, self.h, self.vdelt)

    def _chunk_order(self, corner: int):
        """Chunk origins of one corner's sweep in dependence order: along each
        axis increasing when the corner's bit is set (the functors' forward
        order), decreasing otherwise."""
        cx, cy, cz = self.chunks
        ex, ey, ez = self.chunk
        xs = range(cx) if corner & 1 else range(cx - 1, -1, -1)
        ys = range(cy) if corner & 2 else range(cy - 1, -1, -1)
        zs = range(cz) if corner & 4 else range(cz - 1, -1, -1)
        return [(i * ex, j * ey, k * ez) for k in zs for j in ys for i in xs]

    def _perform_sweeps(self, tin, tout):
        p = self.p
        self.fyz.zero_()
        self.fxz.zero_()
        self.fxy.zero_()
        fluxm = self.fluxm if p.cmom > 1 else None
        for corner in range(8):
            ghosts = (self.fyz[corner], self.fxz[corner], self.fxy[corner])
            for origin in self._chunk_order(corner):
                tasks.mini_kba(corner, range(p.ng), self.qtot, self.flux0, fluxm, self.dinv,
                               self.t_xs, ghosts, self.h, p.nang, p.nmom, fixup=p.fixup,
                               origin=origin, extent=self.chunk,
                               vdelt=self.vdelt if p.timedep else None,
                               time_flux_in=None if tin is None else tin[corner],
                               time_flux_out=None if tout is None else tout[corner])

    def _test(self, fn, prev, eps) -> bool:
        fn(self.flux0, prev, eps, self.flag)
        self.torch.cuda.synchronize()
        return bool(int(self.flag.item()))

    def run(self) -> dict:
        p = self.p
        fluxm = self.fluxm if p.cmom > 1 else None
        q2grpm = self.q2grpm if p.cmom > 1 else None
        self.flux0.zero_()
        self.fluxm.zero_()
        if self.tflux is not None:
            self.tflux.zero_()
        inner_total = outer_total = 0
        all_converged = True
        for cy in range(p.nsteps):
            tin = tout = None
            if self.tflux is not None:
                tin, tout = self.tflux[cy & 1], self.tflux[(cy & 1) ^ 1]
            converged = False
            for otno in range(p.oitm):
                tasks.calc_outer_source(self.qi, self.flux0, self.slgg, self.q2grp0, self.mat,
                                        fluxm=fluxm, qom=q2grpm, nmom=p.nmom)
                self.flux0po.copy_(self.flux0)
                outer_total += 1
                inner_converged = False
                for _ in range(p.iitm):
                    tasks.calc_inner_source(self.s_xs, self.flux0, self.q2grp0, self.qtot,
                                            fluxm=fluxm, q2grpm=q2grpm, nmom=p.nmom)
                    self.flux0pi.copy_(self.flux0)
                    self.flux0.zero_()
                    self.fluxm.zero_()
                    self._perform_sweeps(tin, tout)
                    inner_total += 1
                    inner_converged = self._test(tasks.inner_convergence, self.flux0pi, p.epsi)
                    if inner_converged:
                        break
                if otno == 0:
                    continue
                if inner_converged and self._test(tasks.outer_convergence, self.flux0po,
                                                  100.0 * p.epsi):
                    converged = True
                    break
            all_converged = all_converged and converged
        return {"inner_total": inner_total, "outer_total": outer_total,
                "converged": int(all_converged), "time_steps": p.nsteps}

    def flux(self):
        f = self.flux0.cpu().numpy().copy()
        fm = self.fluxm.cpu().numpy().copy() if self.p.cmom > 1 else None
        return f, fm
