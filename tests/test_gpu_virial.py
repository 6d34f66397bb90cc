"""Molecular virial and pressure on the device (gomc_box_virial) against the
oracle's VirialCalc / VirialReciprocal / VirialCorrection and pressure.

Bar: every virial component and the pressure within 1e-10 relative of the
oracle's (1e-9 of the largest component for each tensor diagonal term),
on water boxes (config 1 size included), a GCMC box with a reservoir (only
box-0 molecules count), and after a run of moves.
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def close(a, b, tol=1e-10):
    return abs(a - b) <= tol * max(abs(b), 1e-300)


def check(vd, vo):
    for key in ("inter", "tail", "real", "recip", "total", "pressure"):
        assert close(vd[key], vo[key]), (key, vd[key], vo[key])
    scale = np.abs(vo["tensor"]).max()
    assert np.all(np.abs(vd["tensor"] - vo["tensor"]) <= 1e-9 * scale)


@pytest.mark.parametrize("n,rcut", [(150, 5.5), (300, 10.0), (1000, 10.0)])
def test_water_virial_and_pressure(n, rcut):
    s = systems.water_box(n, seed=21, rcut=rcut)
    check(Box(s).virial(), Oracle(s, threads=8).virial())


def test_gcmc_box_virial_leaves_out_the_reservoir():
    s = systems.gcmc_mixture(80, 60, 26.0, fugacity_bar=(100.0, 40.0), reservoir=100, seed=5)
    check(Box(s).virial(), Oracle(s, threads=8).virial())


def test_virial_after_moves():
    s = systems.water_box(300, seed=7)
    b, o = Box(s), Oracle(s, threads=8)
    b.run(1500)
    o.run(1500, mode=0)
    check(b.virial(), o.virial())
    b.run(500, start_step=1500, device=True)
    o.run(500, start_step=1500, mode=0)
    check(b.virial(), o.virial())
