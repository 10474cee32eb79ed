"""The reference's solve loop with its chunk decomposition, run through the
drop-in entry points only (snap_amd.kba), against the CPU oracle's
single-chunk solve: the chunks hand their faces on through the flux_xy / yz /
xz face regions, so the result is the single-chunk answer.  The scalar flux
collects each corner's contribution by device atomics in whatever order the
corners finish (as the reference's relaxed-coherence sweep does), so the bar
is 1e-13 relative with identical iteration counts."""
import dataclasses
import os

import numpy as np
import pytest

from snap_amd import deck, problems
from snap_amd.kba import KBASolve

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def _base_deck():
    p, extra = deck.read_deck(os.path.join(DECKS, "base.in"))
    return p, (extra["ichunk"], extra["npey"], extra["npez"])


CASES = {
    "base_deck_2x2x2": lambda: _base_deck(),
    "config1_p1_2x1x4": lambda: (problems.config1(n=8, ng=3, nang=12, nmom=2), (2, 1, 4)),
    "config3_3x3x1": lambda: (problems.config3(n=9, ng=2, nang=10, block=3), (3, 3, 1)),
    "td_config0_2x2x2": lambda: (dataclasses.replace(problems.config0(n=8, ng=2, nang=6),
                                                     timedep=True, nsteps=3, tf=1.5), (2, 2, 2)),
    "nang200_1x2x2": lambda: (problems.config1(n=4, ng=2, nang=200, nmom=1), (1, 2, 2)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_chunked_solve_matches_single_chunk(gpu, oracle_mod, name):
    p, chunks = CASES[name]()
    print("case", name, chunks, flush=True)
    fo, fmo, sto = oracle_mod.solve(p)
    k = KBASolve(p, chunks)
    st = k.run()
    f, fm = k.flux()
    print(name, st, sto)
    assert (st["inner_total"], st["outer_total"], st["converged"]) == \
        (sto["inner_total"], sto["outer_total"], sto["converged"])
    assert np.max(np.abs(f - fo) / np.abs(fo)) <= 1e-13
    if fmo is not None:
        assert np.max(np.abs(fm - fmo)) <= 1e-13 * np.max(np.abs(fmo))


def test_chunk_order_follows_the_projection_functors(gpu):
    p = problems.config0(n=4, ng=1, nang=4)
    k = KBASolve(p, (2, 2, 1))
    # corner 0 sweeps -x, -y: the chunk at the high x and y corner first
    assert k._chunk_order(0) == [(2, 2, 0), (0, 2, 0), (2, 0, 0), (0, 0, 0)]
    # corner 3 sweeps +x, +y
    assert k._chunk_order(3) == [(0, 0, 0), (2, 0, 0), (0, 2, 0), (2, 2, 0)]
    with pytest.raises(ValueError):
        KBASolve(p, (3, 1, 1))
