"""swm_debug_sync_cost: the probe behind the engine's launch-per-stage design
(a kernel boundary against a grid barrier in one cooperative launch).  Every
mode must carry a chain of values through all blocks round after round (the
probe checks the final values itself and fails otherwise), and report a
positive cost per round."""
from __future__ import annotations

import pytest

import swm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nblk", [8, 256, 1024])
def test_sync_cost_modes_carry_the_chain(nblk):
    launch, graph, barrier = swm.sync_cost(nblk, 200)
    print(f"sync cost, {nblk} blocks: launch {launch:.2f} us, graph {graph:.2f} us, "
          f"grid barrier {barrier:.2f} us per round", flush=True)
    assert launch > 0 and graph > 0 and barrier > 0


def test_sync_cost_rejects_bad_arguments():
    with pytest.raises(swm.SwmError):
        swm.sync_cost(0, 10)
    with pytest.raises(swm.SwmError):
        swm.sync_cost(1 << 22, 10)  # more blocks than can be resident at once
