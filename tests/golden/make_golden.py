"""Write the golden fixtures under tests/golden/ with the CPU oracle.

The reference (Legion-SNAP) cannot be built or run here (Legion is not
vendored), and it ships no golden vectors, so these fixtures pin the
oracle's own output: a regression anchor for the oracle and a target for the
GPU parity tests.  Each fixture stores the problem parameters, the scalar
flux (full for config 0, sampled cells otherwise), flux-moment samples and
the iteration counts.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "legion-snap_amd"))
sys.path.insert(0, REPO)

from snap_amd import problems  # noqa: E402
from oracle import oracle  # noqa: E402

# name -> (generator, kwargs, store full flux?)
CASES = {
    "config0": (problems.config0, {}, True),
    "config1_small": (problems.config1, dict(n=16, ng=8, nang=12, nmom=4), False),
    "config1_ragged": (problems.config1, dict(n=20, ng=6, nang=36, nmom=2), False),
    "config3_small": (problems.config3, dict(n=16, ng=4, nang=24), False),
}


def sample_index(shape, count=128, seed=7):
    """Deterministic flat sample of cell indices (all groups kept)."""
    ncells = int(np.prod(shape))
    u = problems.splitmix64_uniform(seed, count)
    idx = np.unique((u * ncells).astype(np.int64))
    return idx


def make(name):
    gen, kw, full = CASES[name]
    p = gen(**kw)
    f, fm, st = oracle.solve(p)
    out = {
        "kwargs_keys": np.array(list(kw.keys()), dtype="U16"),
        "kwargs_vals": np.array([int(v) for v in kw.values()], dtype=np.int64),
        "inner_total": np.int64(st["inner_total"]),
        "outer_total": np.int64(st["outer_total"]),
        "converged": np.int64(st["converged"]),
        "fixup_cells": np.int64(st["fixup_cells"]),
    }
    if full:
        out["flux0"] = f
    else:
        idx = sample_index(f.shape[1:])
        out["cells"] = idx
        out["flux0_samples"] = f.reshape(f.shape[0], -1)[:, idx]
        if fm is not None:
            out["fluxm_samples"] = fm.reshape(fm.shape[0], -1, fm.shape[-1])[:, idx, :]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    return st


if __name__ == "__main__":
    for n in CASES:
        print(n, make(n))
