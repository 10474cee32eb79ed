"""Two groups per wavefront on the one-wavefront tiles (csrc/snap_sweep_seq.hip,
PAIR: lanes 0-31 sweep group g, lanes 32-63 group g + 1): bit-equal to one
group per wavefront in the solver and the batch, for FP64 and FP32 sweeps,
with the same iteration counts; within 1e-12 of the oracle; and not taken
where it does not apply (odd group counts, x or y extents that are not
multiples of 4, more than 32 angles)."""
import numpy as np
import pytest

from snap_amd import problems
from snap_amd.batch import Batch
from snap_amd.solver import Solver

pytestmark = pytest.mark.gpu

CASES = {
    "config0": lambda: problems.config0(),
    "config3_small": lambda: problems.config3(n=16, ng=4, nang=24),
    "p3_nang32": lambda: problems.config1(n=16, ng=6, nang=32, nmom=4),
    "p2_nang10": lambda: problems.config1(n=24, ng=4, nang=10, nmom=3),
    # 20 cells deep: the last tile of each column is 4 of its 8 planes
    "ragged_z": lambda: problems.config1(n=20, ng=4, nang=12, nmom=2),
}


def _solve(prob, precision, gpw, **kw):
    with Solver(prob, precision=precision, sweep_kernel="wave_tile", groups_per_wave=gpw,
                **kw) as s:
        st = s.run()
        f, fm = s.flux()
    return st, f, fm


@pytest.mark.parametrize("name", ["config3_small", "p3_nang32"])
def test_pair_defect_correction_bit_equal(gpu, name):
    # defect correction sweeps FP32 corrections without the fixup
    s2, f2, fm2 = _solve(CASES[name](), "mixed", 0, refine=True)
    s1, f1, fm1 = _solve(CASES[name](), "mixed", 1, refine=True)
    assert s2["groups_per_wave"] == 2 and s1["groups_per_wave"] == 1
    assert (s2["inner_total"], s2["sweeps_fp64"]) == (s1["inner_total"], s1["sweeps_fp64"])
    assert np.array_equal(f2, f1)
    if fm1 is not None:
        assert np.array_equal(fm2, fm1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_pair_bit_equal_to_one_group(gpu, name, precision):
    s2, f2, fm2 = _solve(CASES[name](), precision, 0)
    s1, f1, fm1 = _solve(CASES[name](), precision, 1)
    print(name, precision, s2["grind_ns"], s1["grind_ns"])
    assert s2["groups_per_wave"] == 2 and s1["groups_per_wave"] == 1
    keys = ("inner_total", "outer_total", "sweeps_fp64", "sweeps_fp32", "converged")
    assert tuple(s2[k] for k in keys) == tuple(s1[k] for k in keys)
    assert np.array_equal(f2, f1)
    if fm1 is not None:
        assert np.array_equal(fm2, fm1)


def test_pair_against_the_oracle(gpu, oracle_mod):
    p = CASES["p3_nang32"]()
    fo, fmo, sto = oracle_mod.solve(p)
    st, f, fm = _solve(p, "fp64", 0)
    assert st["groups_per_wave"] == 2
    assert (st["inner_total"], st["outer_total"]) == (sto["inner_total"], sto["outer_total"])
    assert np.max(np.abs(f - fo) / np.abs(fo)) <= 1e-12
    assert np.max(np.abs(fm - fmo)) <= 1e-12 * np.max(np.abs(fmo))


@pytest.mark.parametrize("prob", [
    lambda: problems.config1(n=16, ng=3, nang=12, nmom=1),   # odd group count
    lambda: problems.config1(n=10, ng=4, nang=12, nmom=2),   # x, y not multiples of 4
    lambda: problems.config1(n=8, ng=2, nang=36, nmom=1),    # more than 32 angles
])
def test_pair_only_where_it_applies(gpu, prob):
    st, _, _ = _solve(prob(), "fp64", 0)
    assert st["groups_per_wave"] == 1


@pytest.mark.parametrize("precision,polish", [("fp64", None), ("mixed", 0.0), ("mixed", None)])
def test_batch_pair_bit_equal_to_one_group(gpu, precision, polish):
    probs = problems.config4(count=4, seed=5, ng=4, nang=24, nmom=2, edges=(8, 12, 16))
    res = []
    for gpw in (0, 1):
        with Batch(probs, precision=precision, fp64_polish=polish, sweep_kernel="wave_tile",
                   tile=(0, 0, 8), groups_per_wave=gpw) as b:
            st, per = b.run()
            res.append((st, per, [b.flux(i) for i in range(len(probs))]))
    (st2, per2, fl2), (st1, per1, fl1) = res
    print(precision, polish, st2["grind_ns"], st1["grind_ns"])
    assert st2["groups_per_wave"] == 2 and st1["groups_per_wave"] == 1
    assert st2["converged"] == len(probs)
    for a, b_ in zip(per2, per1):
        assert (a["inner_total"], a["outer_total"], a["sweeps_fp64"]) == \
            (b_["inner_total"], b_["outer_total"], b_["sweeps_fp64"])
    for (fa, fma), (fb, fmb) in zip(fl2, fl1):
        assert np.array_equal(fa, fb) and np.array_equal(fma, fmb)


def test_groups_per_wave_option_checked(gpu):
    from snap_amd._lib import SnapError
    p = problems.config1(n=8, ng=2, nang=12, nmom=1)
    with pytest.raises(SnapError, match="groups_per_wave"):
        Solver(p, precision="fp64", sweep_kernel="wave_tile", groups_per_wave=2)
    with pytest.raises(SnapError, match="groups_per_wave"):
        Batch([p], precision="fp64", sweep_kernel="wave_tile", groups_per_wave=-1)
