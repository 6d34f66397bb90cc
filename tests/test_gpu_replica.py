"""Replica exchange between boxes of one batch (gomc_batch_exchange)
against the oracle's restatement of the neighbour exchange
(ParallelTemperingUtilities.cpp:108-240) over oracle systems that run the
same Philox streams: after every MC cycle and exchange attempt the accepted
pairs and the ladder are identical, and every box's energies (1e-10) and
coordinates (1e-9 A) equal the oracle system's.  An exchange moves the
temperature-bound state, so each box keeps its configuration and HBM
arrays and a swap costs no copy of coordinates.
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Batch
from oracle import pyoracle
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def test_replica_exchange_matches_oracle():
    temps = [280.0, 300.0, 320.0, 345.0, 370.0, 400.0]
    sys_ = [systems.water_box(260, temperature=t, seed=100 + i, place_seed=7, adj_steps=130)
            for i, t in enumerate(temps)]
    bt = Batch(sys_)
    orcs = [Oracle(s, threads=4) for s in sys_]
    ladder_d = np.arange(len(temps), dtype=np.int32)
    ladder_o = ladder_d.copy()
    draws, swaps = 0, 0
    n = sys_[0].nmols
    for attempt in range(12):
        bt.run([n] * len(temps))
        for o in orcs:
            o.run(n, start_step=attempt * n, mode=1, record=False)
        acc_d = bt.exchange(ladder_d, attempt, seed=2024)
        acc_o, d = pyoracle.exchange(orcs, ladder_o, attempt, 2024, draws)
        draws += d
        swaps += int(acc_o.sum())
        assert np.array_equal(acc_d, acc_o), (attempt, acc_d, acc_o)
        assert np.array_equal(ladder_d, ladder_o)
        e, _, _ = bt.stats()
        for b, o in enumerate(orcs):
            assert np.allclose(e[b], o.energies(), rtol=1e-10, atol=0), (attempt, b)
    assert swaps >= 3
    for b, o in enumerate(orcs):
        assert np.allclose(bt.coords(b), o.coords(), rtol=0, atol=1e-9)
