"""Reader for the reference's input decks (``input/*.in``): the namelist of
Snap::parse_arguments (snap.cc:1032-1168) turned into a Problem with the
reference's own material, scattering and source set-up:

* materials (InitMaterial, init.cc:49-111): material 1 everywhere; with
  mat_opt 1 material 2 fills cells [nx/4, 3nx/4) per axis (the centre), with
  mat_opt 2 cells [0, nx/2) (a corner);
* cross sections (Snap::initialize_scattering, snap.cc:571-714): material 1
  sigt/siga/sigs = 1.0/0.5/0.5 in group 0, each later group 0.01/0.005/0.005
  of the previous one; material 2 = 2.0/0.8/1.2 with the same ratios;
  group-to-group scattering splits sigs of each source group (material 1:
  0.2 in-group, 0.1 spread over lower groups, 0.7 over higher groups; 0.3 /
  0.9 in-group for the first / last group; material 2: 0.5, 0.1, 0.4, and 0.6
  / 0.9), higher Legendre orders scaled by 0.1 then 0.5 (material 2: 0.8,
  0.6) per order;
* fixed source (InitSource, init.cc:113-191): 1.0 in every group inside the
  source region -- everywhere (src_opt 0), the centre (1) or a corner (2).

Material indices here are 0-based (the reference's 1 and 2 become 0 and 1).
The decks' chunking (npey, npez, ichunk), thread counts and output flags are
read and kept in ``extra`` but do not change the problem: one GPU solves the
whole grid as one chunk.  timedep, tf and nsteps set the Problem's time
stepping and src_opt 3 its manufactured-solution source.
"""
from __future__ import annotations

import re
from typing import Dict, Tuple

import numpy as np

from .problems import Problem

# keys in the order Snap::parse_arguments reads them (snap.cc:1042-1110)
_KEYS = ["npey", "npez", "ichunk", "nthreads", "nnested", "ndimen", "nx", "lx", "ny", "ly",
         "nz", "lz", "nmom", "nang", "ng", "epsi", "iitm", "oitm", "timedep", "tf", "nsteps",
         "mat_opt", "src_opt", "scatp", "it_det", "fluxp", "fixup", "soloutp", "kplane",
         "popout", "swp_typ", "angcpy"]
_FLOATS = {"lx", "ly", "lz", "epsi", "tf"}


def parse_namelist(text: str) -> Dict[str, float]:
    """key=value pairs of the &invar namelist (comments after '!' and the
    '&invar' / '/' markers are skipped).  Unknown keys are an error."""
    vals: Dict[str, float] = {}
    for line in text.splitlines():
        line = line.split("!", 1)[0].strip()
        if not line or line.startswith("&") or line == "/":
            continue
        for m in re.finditer(r"([A-Za-z_]\w*)\s*=\s*([-+0-9.eEdD]+)", line):
            key, v = m.group(1).lower(), m.group(2).replace("d", "e").replace("D", "e")
            if key not in _KEYS:
                raise ValueError(f"unknown deck key {key!r}")
            vals[key] = float(v) if key in _FLOATS else int(float(v))
    return vals


def check_deck(vals: Dict[str, float]) -> None:
    """The deck conditions of snap.cc:1117-1160."""
    if int(vals.get("angcpy", 0)) == 1:
        raise ValueError("angcpy=1 is not supported (two copies of the angular flux are kept)")
    if int(vals.get("swp_typ", 1)) == 0:
        raise ValueError("swp_typ=0 is not supported (mini-KBA sweeps only)")
    nx, ny, nz = int(vals["nx"]), int(vals["ny"]), int(vals["nz"])
    for key, n in (("ichunk", nx), ("npey", ny), ("npez", nz)):
        c = int(vals.get(key, 1))
        if not (1 <= c <= n) or n % c:
            raise ValueError(f"{key}={c} must divide the grid edge {n}")
    if min(nx, ny) < 4:
        raise ValueError("nx and ny must be >= 4")
    if int(vals.get("timedep", 0)) and float(vals.get("tf", 0.0)) < 0.0:
        raise ValueError("tf must be >= 0")
    if int(vals.get("nsteps", 1)) < 1:
        raise ValueError("nsteps must be >= 1")


def reference_cross_sections(ng: int, nmom: int, two_materials: bool):
    """sigt, siga, sigs [nmat][ng] and slgg [ng_to][nmat][ng_from][nmom] of
    Snap::initialize_scattering (snap.cc:571-714)."""
    nmat = 2 if two_materials else 1
    base = [(1.0, 0.5, 0.5), (2.0, 0.8, 1.2)]
    sigt = np.zeros((nmat, ng))
    siga = np.zeros((nmat, ng))
    sigs = np.zeros((nmat, ng))
    for m in range(nmat):
        sigt[m, 0], siga[m, 0], sigs[m, 0] = base[m]
        for g in range(1, ng):
            sigt[m, g] = 0.01 * sigt[m, g - 1]
            siga[m, g] = 0.005 * siga[m, g - 1]
            sigs[m, g] = 0.005 * sigs[m, g - 1]
    slgg = np.zeros((ng, nmat, ng, nmom))
    if ng == 1:
        # snap.cc:629 stores material 2's one-group value at point (1, 1), a
        # group index past the table for ng = 1; it is kept here at (2, 0)
        for m in range(nmat):
            slgg[0, m, 0, 0] = sigs[m, 0]
    else:
        # (in-group, total upscatter, total downscatter, first-group in-group,
        # last-group in-group) fractions of sigs per material
        split = [(0.2, 0.1, 0.7, 0.3, 0.9), (0.5, 0.1, 0.4, 0.6, 0.9)]
        for m in range(nmat):
            fin, fup, fdown, ffirst, flast = split[m]
            for g in range(ng):
                # the split fractions multiply by a precomputed reciprocal of
                # the group count, (frac * sigs) * (1/n), as snap.cc:637-650
                # and 663-676 do (dividing by n can differ in the last bit)
                slgg[g, m, g, 0] = fin * sigs[m, g]
                if g > 0:
                    t = 1.0 / float(g)
                    for g2 in range(g):
                        slgg[g2, m, g, 0] = fup * sigs[m, g] * t
                else:
                    slgg[g, m, g, 0] = ffirst * sigs[m, g]
                if g < ng - 1:
                    t = 1.0 / float(ng - (g + 1))
                    for g2 in range(g + 1, ng):
                        slgg[g2, m, g, 0] = fdown * sigs[m, g] * t
                else:
                    slgg[g, m, g, 0] = flast * sigs[m, g]
    ratios = [(0.1, 0.5), (0.8, 0.6)]
    for m in range(nmat):
        for l in range(1, nmom):
            slgg[:, m, :, l] = (ratios[m][0] if l == 1 else ratios[m][1]) * slgg[:, m, :, l - 1]
    return sigt, siga, sigs, slgg


def reference_layouts(nx: int, ny: int, nz: int, ng: int, mat_opt: int, src_opt: int):
    """mat [nz][ny][nx] (0-based) and qi [ng][nz][ny][nx] of InitMaterial and
    InitSource (init.cc:49-191), for a 3-D grid."""
    mat = np.zeros((nz, ny, nx), dtype=np.int32)
    if mat_opt == 1:
        mat[nz // 4:3 * nz // 4, ny // 4:3 * ny // 4, nx // 4:3 * nx // 4] = 1
    elif mat_opt == 2:
        mat[:nz // 2, :ny // 2, :nx // 2] = 1
    elif mat_opt != 0:
        raise ValueError(f"mat_opt {mat_opt} is not 0, 1 or 2")
    qi = np.zeros((ng, nz, ny, nx))
    if src_opt == 0:
        qi[:] = 1.0
    elif src_opt == 1:
        qi[:, nz // 4:3 * nz // 4, ny // 4:3 * ny // 4, nx // 4:3 * nx // 4] = 1.0
    elif src_opt == 2:
        qi[:, :nz // 2, :ny // 2, :nx // 2] = 1.0
    elif src_opt != 3:
        raise ValueError(f"src_opt {src_opt} is not 0..3")
    return mat, qi


def read_deck(path: str) -> Tuple[Problem, Dict[str, float]]:
    """Problem of a reference input deck, and the deck's other settings."""
    with open(path) as f:
        vals = parse_namelist(f.read())
    missing = [k for k in ("nx", "ny", "nz", "lx", "ly", "lz", "nmom", "nang", "ng") if k not in vals]
    if missing:
        raise ValueError(f"{path}: missing {missing}")
    if int(vals.get("ndimen", 3)) != 3:
        raise ValueError("only 3-D decks (ndimen = 3) are supported")
    check_deck(vals)
    nx, ny, nz, ng = int(vals["nx"]), int(vals["ny"]), int(vals["nz"]), int(vals["ng"])
    nmom, nang = int(vals["nmom"]), int(vals["nang"])
    mat_opt, src_opt = int(vals.get("mat_opt", 0)), int(vals.get("src_opt", 0))
    sigt, siga, sigs, slgg = reference_cross_sections(ng, nmom, mat_opt != 0)
    mat, qi = reference_layouts(nx, ny, nz, ng, mat_opt, src_opt)
    p = Problem(nx=nx, ny=ny, nz=nz, lx=float(vals["lx"]), ly=float(vals["ly"]),
                lz=float(vals["lz"]), ng=ng, nang=nang, nmom=nmom, mat=mat, sigt=sigt, siga=siga,
                sigs=sigs, slgg=slgg, qi=qi, epsi=float(vals.get("epsi", 1e-4)),
                iitm=int(vals.get("iitm", 5)), oitm=int(vals.get("oitm", 100)),
                fixup=bool(int(vals.get("fixup", 1))), name=path.rsplit("/", 1)[-1],
                timedep=bool(int(vals.get("timedep", 0))), tf=float(vals.get("tf", 1.0)),
                nsteps=int(vals.get("nsteps", 1)), mms=src_opt == 3)
    extra = {k: v for k, v in vals.items()}
    return p, extra
