"""Reference input decks run end to end on the GPU through the deck runner
(the `snap <deck>` equivalent): results against the oracle, the recorded
convergence tests against the iteration counts, the MMS acceptance."""
import os

import numpy as np
import pytest

from snap_amd import deck, run_deck
from snap_amd.solver import Solver

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


@pytest.mark.parametrize("name", ["base.in", "mms.in", "tiny_1K_1x1x1.in"])
def test_deck_runner(gpu, oracle_mod, name):
    path = os.path.join(DECKS, name)
    print("deck", name, flush=True)
    r = run_deck.run(path, precision="fp64_exact", quiet=True)
    print("\n".join(r["lines"][-12:]))
    assert r["ok"]
    p, _ = deck.read_deck(path)
    fo, _, sto = oracle_mod.solve(p)
    st = r["stats"]
    assert (st["inner_total"], st["outer_total"], st["time_steps"]) == \
        (sto["inner_total"], sto["outer_total"], sto["time_steps"])
    with Solver(p, precision="fp64_exact") as s:
        s.run()
        f, _ = s.flux()
        log = s.iteration_log()
    assert np.array_equal(f, fo)
    inner = [e for e in log if e["kind"] == "inner"]
    outer = [e for e in log if e["kind"] == "outer"]
    assert len(inner) == st["inner_total"]
    # no outer test in the first outer iteration of each step
    assert len(outer) == st["outer_total"] - st["time_steps"]
    assert sum(e["converged"] for e in outer) == st["time_steps"]
    assert all(b["t_us"] >= a["t_us"] for a, b in zip(log, log[1:]))
