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


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_single_solves(gpu, precision, use_graph):
    probs = _small_batch()
    with Batch(probs, precision=precision, use_graph=use_graph) as b:
        st, per = b.run()
        fluxes = [b.flux(i) for i in range(len(probs))]
    print(st)
    assert st["converged"] == len(probs)
    for i, p in enumerate(probs):
        with Solver(p, precision=precision) as s:
            ss = s.run()
            f, fm = s.flux()
        assert (per[i]["inner_total"], per[i]["outer_total"]) == (ss["inner_total"], ss["outer_total"])
        assert np.array_equal(fluxes[i][0], f)
        if fm is not None:
            assert np.array_equal(fluxes[i][1], fm)


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
