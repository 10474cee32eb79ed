"""Batched solves: every problem of a batch gets the same fluxes and
iteration counts as its own solve (the batch changes launch structure and
predication, never a cell's arithmetic), with and without the captured
hipGraph, in FP64 and mixed precision."""
import numpy as np
import pytest

from snap_amd import problems
from snap_amd.batch import Batch
from snap_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _small_batch(count=5, seed=17):
    return problems.config4(count=count, seed=seed, ng=4, nang=10, nmom=2, edges=(6, 9, 12, 16))


@pytest.mark.parametrize("kernel", ["auto", "tiled"])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("precision,polish", [("fp64", None), ("mixed", 0.0), ("mixed", None)])
def test_batch_equals_single_solves(gpu, precision, polish, use_graph, kernel):
    probs = _small_batch()
    with Batch(probs, precision=precision, use_graph=use_graph, fp64_polish=polish,
               sweep_kernel=kernel) as b:
        st, per = b.run()
        fluxes = [b.flux(i) for i in range(len(probs))]
    print(st, [(x["sweeps_fp64"], x["inner_total"], round(x["time_ms"], 3)) for x in per])
    assert st["converged"] == len(probs)
    # the single solves on the kernel the batch ran: the autotuner takes the
    # one-wavefront tiles up to 64 angles (the solver's default here is the
    # multi-wave tiles, whose fused arithmetic differs in the last bits)
    seq = st["waves"] == 1 and tuple(st["tile"][:2]) == (4, 4)
    assert seq == (kernel == "auto")
    kw = dict(sweep_kernel="wave_tile", tile=(0, 0, st["tile"][2])) if seq else {}
    for i, p in enumerate(probs):
        # the same polish rule (FP32 throughout with 0) picks the same sweeps;
        # batches sweep the whole source (no defect correction)
        with Solver(p, precision=precision, fp64_polish=polish, refine=False, **kw) as s:
            ss = s.run()
            f, fm = s.flux()
        assert (per[i]["inner_total"], per[i]["outer_total"]) == (ss["inner_total"], ss["outer_total"])
        assert per[i]["sweeps_fp64"] == ss["sweeps_fp64"]
        assert np.array_equal(fluxes[i][0], f)
        if fm is not None:
            assert np.array_equal(fluxes[i][1], fm)
        assert 0.0 < per[i]["time_ms"] <= st["solve_ms"]
    if precision == "mixed" and polish is None:
        assert any(x["sweeps_fp64"] > 0 for x in per)


@pytest.mark.parametrize("precision,polish", [("fp64", None), ("mixed", 0.0), ("mixed", None)])
@pytest.mark.parametrize("tz", [4, 8])
def test_wave_tile_batch_equals_single_solves(gpu, precision, polish, tz):
    # the batch on one-wavefront 4 x 4 x TZ tiles against single solves on
    # the same kernel: the batch form changes addressing and predication only
    probs = _small_batch()
    with Batch(probs, precision=precision, fp64_polish=polish, sweep_kernel="wave_tile",
               tile=(0, 0, tz)) as b:
        st, per = b.run()
        fluxes = [b.flux(i) for i in range(len(probs))]
    print(st)
    assert st["waves"] == 1 and tuple(st["tile"][:2]) == (4, 4)
    assert st["converged"] == len(probs)
    for i, p in enumerate(probs):
        with Solver(p, precision=precision, fp64_polish=polish, refine=False,
                    sweep_kernel="wave_tile", tile=(0, 0, tz)) as s:
            ss = s.run()
            f, fm = s.flux()
        assert (per[i]["inner_total"], per[i]["outer_total"]) == (ss["inner_total"], ss["outer_total"])
        assert per[i]["sweeps_fp64"] == ss["sweeps_fp64"]
        assert np.array_equal(fluxes[i][0], f)
        if fm is not None:
            assert np.array_equal(fluxes[i][1], fm)


def test_batch_reuses_the_tuned_tile(gpu):
    probs = _small_batch(count=3, seed=9)
    with Batch(probs, precision="fp64") as b:
        st1, _ = b.run()
        f1 = b.flux(2)[0]
    with Batch(probs, precision="fp64") as b:
        st2, _ = b.run()
        f2 = b.flux(2)[0]
    assert st1["plan_cached"] == 0 and st1["autotune_ms"] > 0
    assert st2["plan_cached"] == 1 and st2["autotune_ms"] == 0
    assert st2["tile"] == st1["tile"] and st2["angles_per_lane"] == st1["angles_per_lane"]
    assert np.array_equal(f1, f2)


def test_batch_with_fixed_tile(gpu):
    probs = _small_batch(count=3, seed=5)
    with Batch(probs, precision="fp64", tile=(4, 4, 4)) as b:
        st, per = b.run()
        assert tuple(st["tile"]) == (4, 4, 4)
        f0, _ = b.flux(1)
    with Solver(probs[1], precision="fp64") as s:
        s.run()
        f, _ = s.flux()
    assert np.array_equal(f0, f)


def test_batch_two_angles_per_lane_p3(gpu):
    # 80 angles need two angles per lane in FP64 as well; P3 moments
    probs = problems.config4(count=3, seed=3, ng=3, nang=80, nmom=4, edges=(6, 8, 10))
    with Batch(probs, precision="fp64") as b:
        st, per = b.run()
        assert st["angles_per_lane"] == 2
        fluxes = [b.flux(i) for i in range(len(probs))]
    for i, p in enumerate(probs):
        with Solver(p, precision="fp64") as s:
            ss = s.run()
            f, fm = s.flux()
        assert (per[i]["inner_total"], per[i]["outer_total"]) == (ss["inner_total"], ss["outer_total"])
        assert np.array_equal(fluxes[i][0], f) and np.array_equal(fluxes[i][1], fm)


def test_batch_rejects_more_than_128_angles(gpu):
    from snap_amd import _lib
    probs = problems.config4(count=2, seed=1, ng=2, nang=130, nmom=1, edges=(4, 6))
    with pytest.raises(_lib.SnapError) as e:
        Batch(probs, precision="fp64")
    assert e.value.code == _lib.SNAP_ERR_UNSUPPORTED
