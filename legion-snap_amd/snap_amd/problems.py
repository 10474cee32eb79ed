"""Problem definitions for the steady-state mini-KBA solve.

A :class:`Problem` carries the same tables the reference builds before its
solve loop: the material map ``mat`` (InitMaterial, init.cc:49-104), the
per-material cross sections ``sigt/siga/sigs`` and the group-to-group
scattering matrix ``slgg`` (Snap::initialize_scattering, snap.cc:571-714), the
fixed source ``qi`` (InitSource, init.cc:131-191) and the deck scalars
(snap.cc:1032-1168).  The synthetic workloads below replace the reference's
built-in tables; the solver arithmetic that consumes them is unchanged.

Array layouts (numpy, C order, x fastest as in the reference's SOA layout
snap.cc:1456-1475):

* ``mat``        int32   [nz, ny, nx]          0-based material index
* ``sigt``       float64 [nmat, ng]
* ``slgg``       float64 [ng_to, nmat, ng_from, nmom]   (field g_to, point (mat, g_from), order l)
* ``qi``         float64 [ng, nz, ny, nx]
* ``sigt_cell``  float64 [ng, nz, ny, nx] or None  (per-cell total cross section)
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

__all__ = [
    "Problem", "config0", "config1", "config2", "config3", "config4",
    "scattering_split", "splitmix64_uniform", "standard_normal",
    "lognormal_absorption", "corner_block_mask", "named_config",
]


@dataclasses.dataclass
class Problem:
    """One fixed-source, steady-state, non-multiplying transport problem."""

    nx: int
    ny: int
    nz: int
    ng: int
    nang: int
    nmom: int
    mat: np.ndarray
    sigt: np.ndarray
    siga: np.ndarray
    sigs: np.ndarray
    slgg: np.ndarray
    qi: np.ndarray
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0
    epsi: float = 1.0e-4
    iitm: int = 5
    oitm: int = 100
    fixup: bool = True
    sigt_cell: Optional[np.ndarray] = None
    name: str = "problem"
    # time dependence (snap.cc:716-738, 400-541): nsteps steps of dt = tf/nsteps
    # with vdelt_g = 2/(dt (ng-g)); a steady problem runs one step with vdelt 0
    timedep: bool = False
    tf: float = 1.0
    nsteps: int = 1
    # manufactured-solution source (src_opt 3, mms.cc): the solver builds qi
    # and the angular source itself and qi is ignored
    mms: bool = False

    @property
    def nmat(self) -> int:
        return int(self.sigt.shape[0])

    @property
    def cmom(self) -> int:
        # 3-D flux moments: nmom**2 (snap.cc:1203-1206)
        return self.nmom * self.nmom

    @property
    def ncells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def units_per_sweep(self) -> int:
        """cells x directions (8*nang) x groups: the grind-time unit count."""
        return self.ncells * 8 * self.nang * self.ng

    def validate(self) -> None:
        """The deck checks of snap.cc:1139-1160 that apply to this path."""
        if min(self.nx, self.ny, self.nz) < 1:
            raise ValueError("grid edges must be >= 1")
        if not (1 <= self.nmom <= 4):
            raise ValueError("nmom must be in [1, 4]")
        if self.nang < 1 or self.ng < 1:
            raise ValueError("nang and ng must be >= 1")
        if not (0.0 < self.epsi < 1e-2):
            raise ValueError("epsi must be in (0, 1e-2)")
        if self.iitm < 1 or self.oitm < 1:
            raise ValueError("iitm and oitm must be >= 1")
        if min(self.lx, self.ly, self.lz) <= 0.0:
            raise ValueError("domain lengths must be positive")
        shp = (self.nz, self.ny, self.nx)
        if self.mat.shape != shp or self.mat.dtype != np.int32:
            raise ValueError("mat must be int32 [nz, ny, nx]")
        if self.mat.min() < 0 or self.mat.max() >= self.nmat:
            raise ValueError("mat index out of range")
        if self.sigt.shape != (self.nmat, self.ng):
            raise ValueError("sigt must be [nmat, ng]")
        if self.slgg.shape != (self.ng, self.nmat, self.ng, self.nmom):
            raise ValueError("slgg must be [ng, nmat, ng, nmom]")
        if self.qi.shape != (self.ng,) + shp:
            raise ValueError("qi must be [ng, nz, ny, nx]")
        if self.sigt_cell is not None and self.sigt_cell.shape != (self.ng,) + shp:
            raise ValueError("sigt_cell must be [ng, nz, ny, nx]")
        if self.nsteps < 1:
            raise ValueError("nsteps must be >= 1")
        if self.timedep and not self.tf > 0.0:
            raise ValueError("a time-dependent problem needs tf > 0")
        if self.mms and (self.nmom != 1 or self.sigt_cell is not None):
            raise ValueError("the manufactured solution is defined for nmom = 1 and material cross sections")

    def contiguous(self) -> "Problem":
        """Copy with every array C-contiguous in the dtype the ABIs expect."""
        p = dataclasses.replace(self)
        p.mat = np.ascontiguousarray(self.mat, dtype=np.int32)
        for f in ("sigt", "siga", "sigs", "slgg", "qi"):
            setattr(p, f, np.ascontiguousarray(getattr(self, f), dtype=np.float64))
        if self.sigt_cell is not None:
            p.sigt_cell = np.ascontiguousarray(self.sigt_cell, dtype=np.float64)
        return p


# --------------------------------------------------------------------------
# scattering tables
# --------------------------------------------------------------------------

def scattering_split(sigs: np.ndarray, nmom: int, order_scale: float = 0.1) -> np.ndarray:
    """Group-to-group matrix for one material, shape [ng_to, ng_from, nmom].

    Half of ``sigs[g]`` stays in group g; the other half is split equally
    over the lower-energy groups g' > g; the last group keeps all of its
    scattering.  Legendre order l is ``order_scale**l`` times order 0.
    """
    ng = int(sigs.shape[0])
    m = np.zeros((ng, ng, nmom), dtype=np.float64)
    for g in range(ng):
        if g == ng - 1:
            m[g, g, 0] = sigs[g]
        else:
            m[g, g, 0] = 0.5 * sigs[g]
            share = 0.5 * sigs[g] / float(ng - 1 - g)
            for g2 in range(g + 1, ng):
                m[g2, g, 0] = share
    for l in range(1, nmom):
        m[:, :, l] = (order_scale ** l) * m[:, :, 0]
    return m


def _slgg_from_materials(mats: List[np.ndarray]) -> np.ndarray:
    # [ng_to, nmat, ng_from, nmom]
    return np.ascontiguousarray(np.stack(mats, axis=1))


def _material1(ng: int):
    g = np.arange(ng, dtype=np.float64)
    sigt = 1.0 + 0.01 * g          # sigt_g = 1 + 0.01 (g-1) with 1-based g
    siga = 0.5 * sigt
    sigs = sigt - siga
    return sigt, siga, sigs


def centre_cube_mask(nx: int, ny: int, nz: int) -> np.ndarray:
    """Cells [n/4, 3n/4) on each axis: the reference's CENTER layout
    (init.cc:59-75: i1 = n/4+1 .. i2 = 3n/4, 1-based inclusive)."""
    m = np.zeros((nz, ny, nx), dtype=bool)
    m[nz // 4:(3 * nz) // 4, ny // 4:(3 * ny) // 4, nx // 4:(3 * nx) // 4] = True
    return m


def corner_block_mask(nx: int, ny: int, nz: int, edge: int, far: bool = False) -> np.ndarray:
    """An edge^3 block at the origin corner, or at the opposite corner."""
    m = np.zeros((nz, ny, nx), dtype=bool)
    if far:
        m[nz - edge:, ny - edge:, nx - edge:] = True
    else:
        m[:edge, :edge, :edge] = True
    return m


# --------------------------------------------------------------------------
# deterministic random numbers (splitmix64 + Box-Muller), numpy-version free
# --------------------------------------------------------------------------

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n uniforms in (0, 1) from the splitmix64 sequence of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def standard_normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    u1 = splitmix64_uniform(seed, n, 2 * stream)
    u2 = splitmix64_uniform(seed, n, 2 * stream + 1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def lognormal_absorption(nx: int, ny: int, nz: int, seed: int,
                         log_mean: float = math.log(0.5), sigma: float = 0.3) -> np.ndarray:
    """Per-cell absorption: exp(log_mean + sigma * z), z a unit-variance field
    made by smoothing white noise with a separable 3-tap Gaussian (sigma of one
    cell, taps at -1/0/+1, edges clamped) and dividing by the kernel's variance
    factor (sum of squared taps, cubed)."""
    zf = standard_normal(seed, nx * ny * nz).reshape(nz, ny, nx)
    k = np.array([math.exp(-0.5), 1.0, math.exp(-0.5)])
    k = k / k.sum()
    for axis in range(3):
        pad = [(0, 0)] * 3
        pad[axis] = (1, 1)
        zp = np.pad(zf, pad, mode="edge")
        sl = [slice(None)] * 3
        out = np.zeros_like(zf)
        n_ax = zf.shape[axis]
        for t in range(3):
            sl[axis] = slice(t, t + n_ax)
            out += k[t] * zp[tuple(sl)]
        zf = out
    zf = zf / math.sqrt(float(np.sum(k * k)) ** 3)
    return np.exp(log_mean + sigma * zf)


# --------------------------------------------------------------------------
# the five workloads
# --------------------------------------------------------------------------

def config0(n: int = 16, ng: int = 4, nang: int = 10) -> Problem:
    """Plain FP64 reference solve: one uniform material, unit source everywhere,
    vacuum boundaries, epsi 1e-4, 5 inners per outer, up to 100 outers."""
    sigt, siga, sigs = _material1(ng)
    slgg = _slgg_from_materials([scattering_split(sigs, 1)])
    return Problem(
        nx=n, ny=n, nz=n, ng=ng, nang=nang, nmom=1,
        mat=np.zeros((n, n, n), dtype=np.int32),
        sigt=sigt[None, :].copy(), siga=siga[None, :].copy(), sigs=sigs[None, :].copy(),
        slgg=slgg, qi=np.ones((ng, n, n, n)), name=f"config0_{n}c_{ng}g_{nang}a",
    )


def config1(n: int = 64, ng: int = 32, nang: int = 36, nmom: int = 4) -> Problem:
    """Two materials: material 2 fills the centred cube of half-width 0.25 with
    doubled absorption (scattering kept, so sigt2 = 2 siga1 + sigs1); the unit
    source sits only in that cube; P_l scattering = 0.1**l times P0."""
    sigt1, siga1, sigs1 = _material1(ng)
    siga2 = 2.0 * siga1
    sigs2 = sigs1.copy()
    sigt2 = siga2 + sigs2
    cube = centre_cube_mask(n, n, n)
    mat = np.where(cube, 1, 0).astype(np.int32)
    slgg = _slgg_from_materials([scattering_split(sigs1, nmom), scattering_split(sigs2, nmom)])
    qi = np.zeros((ng, n, n, n))
    qi[:, cube] = 1.0
    return Problem(
        nx=n, ny=n, nz=n, ng=ng, nang=nang, nmom=nmom, mat=mat,
        sigt=np.stack([sigt1, sigt2]), siga=np.stack([siga1, siga2]),
        sigs=np.stack([sigs1, sigs2]), slgg=slgg, qi=qi,
        name=f"config1_{n}c_{ng}g_{nang}a_p{nmom - 1}",
    )


def config2(**kw) -> Problem:
    """Mixed-precision form of config 1: the same problem."""
    p = config1(**kw)
    p.name = p.name.replace("config1", "config2")
    return p


def config3(n: int = 48, ng: int = 16, nang: int = 24, seed: int = 20240611,
            block: int = 8) -> Problem:
    """Per-cell log-normal absorption, sigs = 0.5 per group split as in config 0,
    unit source in an 8^3 corner block (the detector is the opposite block)."""
    siga_cell = lognormal_absorption(n, n, n, seed)
    sigs = np.full(ng, 0.5)
    slgg = _slgg_from_materials([scattering_split(sigs, 1)])
    sigt_cell = np.broadcast_to(siga_cell + 0.5, (ng, n, n, n)).copy()
    qi = np.zeros((ng, n, n, n))
    qi[:, corner_block_mask(n, n, n, block)] = 1.0
    mean_siga = float(np.mean(siga_cell))
    return Problem(
        nx=n, ny=n, nz=n, ng=ng, nang=nang, nmom=1,
        mat=np.zeros((n, n, n), dtype=np.int32),
        sigt=np.full((1, ng), mean_siga + 0.5), siga=np.full((1, ng), mean_siga),
        sigs=sigs[None, :].copy(), slgg=slgg, qi=qi, sigt_cell=sigt_cell,
        name=f"config3_{n}c_{ng}g_{nang}a",
    )


def config4(count: int = 48, seed: int = 4242, ng: int = 16, nang: int = 24,
            nmom: int = 4, edges=(24, 32, 40, 48, 56, 64)) -> List[Problem]:
    """Independent uneven two-material problems: edges per axis drawn from
    ``edges``, material-2 cube (and its unit source) centred and sized per
    problem from the seed."""
    edges = np.asarray(edges)
    u = splitmix64_uniform(seed, count * 7, stream=7).reshape(count, 7)
    probs = []
    sigt1, siga1, sigs1 = _material1(ng)
    siga2 = 2.0 * siga1
    sigs2 = sigs1.copy()
    sigt2 = siga2 + sigs2
    slgg = _slgg_from_materials([scattering_split(sigs1, nmom), scattering_split(sigs2, nmom)])
    for i in range(count):
        nx, ny, nz = (int(edges[min(int(u[i, a] * len(edges)), len(edges) - 1)]) for a in range(3))
        half = 0.15 + 0.15 * u[i, 3]
        centre = [half + (1.0 - 2.0 * half) * u[i, 4 + a] for a in range(3)]
        xc = (np.arange(nx) + 0.5) / nx
        yc = (np.arange(ny) + 0.5) / ny
        zc = (np.arange(nz) + 0.5) / nz
        inside = ((np.abs(zc - centre[2]) < half)[:, None, None]
                  & (np.abs(yc - centre[1]) < half)[None, :, None]
                  & (np.abs(xc - centre[0]) < half)[None, None, :])
        qi = np.zeros((ng, nz, ny, nx))
        qi[:, inside] = 1.0
        probs.append(Problem(
            nx=nx, ny=ny, nz=nz, ng=ng, nang=nang, nmom=nmom,
            mat=np.where(inside, 1, 0).astype(np.int32),
            sigt=np.stack([sigt1, sigt2]), siga=np.stack([siga1, siga2]),
            sigs=np.stack([sigs1, sigs2]), slgg=slgg, qi=qi,
            name=f"config4_{i}_{nx}x{ny}x{nz}",
        ))
    return probs


def named_config(name: str, **kw) -> Problem:
    return {"config0": config0, "config1": config1, "config2": config2,
            "config3": config3}[name](**kw)
