"""The oracle's replica exchange (ParallelTemperingUtilities.cpp:108-240,
neighbour branch) on cases the criterion fixes: replicas with equal
energies always exchange (exp(0) = 1), so each attempt swaps every pair of
its parity and draws once per pair; replicas at equal temperature likewise;
a pair whose acceptance underflows to zero never exchanges."""
import numpy as np

from gomc_amd import systems
from oracle import pyoracle
from oracle.pyoracle import Oracle


def test_equal_energies_always_exchange():
    temps = [280.0, 300.0, 330.0, 360.0, 400.0]
    orcs = [Oracle(systems.water_box(120, temperature=t, seed=5, place_seed=9)) for t in temps]
    ladder = np.arange(5, dtype=np.int32)
    acc, d = pyoracle.exchange(orcs, ladder, 0, 11, 0)  # parity 0: pairs (1,2), (3,4)
    assert list(acc) == [0, 0, 1, 0, 1] and d == 2
    assert list(ladder) == [0, 2, 1, 4, 3]
    acc, d = pyoracle.exchange(orcs, ladder, 1, 11, 2)  # parity 1: pairs (0,1), (2,3)
    assert list(acc) == [0, 1, 0, 1, 0] and d == 2
    assert list(ladder) == [2, 0, 4, 1, 3]


def test_far_apart_replicas_never_exchange():
    # a dilute box at 100 K below a dense (randomly placed, high-energy) box
    # at 5000 K: the energy gap times the beta gap makes exp(...) underflow
    cold = Oracle(systems.water_box(200, temperature=100.0, seed=4, density=0.2))
    hot = Oracle(systems.water_box(200, temperature=5000.0, seed=3))
    e_cold, e_hot = cold.energies()[6], hot.energies()[6]
    assert (1 / 5000.0 - 1 / 100.0) * (e_hot - e_cold) < -800
    ladder = np.array([0, 1], dtype=np.int32)
    acc, d = pyoracle.exchange([cold, hot], ladder, 1, 5, 0)
    assert list(acc) == [0, 0] and d == 1 and list(ladder) == [0, 1]
