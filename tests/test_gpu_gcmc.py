"""Parity of the grand-canonical move set on the device path with the
oracle: MoleculeTransfer with CBMC growth of methane (DCSingle) and rigid
CO2 (DCFreeHedron), interleaved with translations and rotations.

Tolerances (FP64): box totals with a reservoir 1e-11 relative; the
move-kind, molecule, overlap and accept/reject sequences identical; per-move
energy changes within 1e-10 of the box totals; the final box membership
identical and coordinates within 1e-9 A; running energies within 1e-10.
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def dE_close(a, b, total, rtol=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), abs(total))))


def test_totals_exclude_the_reservoir():
    s = systems.gcmc_mixture(80, 60, 26.0, fugacity_bar=(100.0, 40.0), reservoir=100, seed=5)
    b, o = Box(s), Oracle(s, threads=8)
    assert np.allclose(b.system_total(), o.system_total(), rtol=1e-11, atol=1e-9)
    n, boxes = b.mol_boxes()
    no, oboxes = o.mol_boxes()
    assert n == no == 140 and np.array_equal(boxes, oboxes)


@pytest.mark.parametrize("nch4,nco2,L,fug,nsteps", [
    (60, 60, 25.0, (200.0, 60.0), 3000),
    (150, 150, 30.0, (400.0, 150.0), 1500),
])
def test_gcmc_moves_same_sequence_as_oracle(nch4, nco2, L, fug, nsteps):
    s = systems.gcmc_mixture(nch4, nco2, L, fugacity_bar=fug, reservoir=max(nch4, nco2) + 60,
                             seed=3, adj_steps=250)
    b, o = Box(s), Oracle(s, threads=8)
    e0 = o.energies()
    ng, rg = b.run(nsteps)
    no, ro = o.run(nsteps, mode=0)
    k = ro["kind"]
    assert (k == 10).sum() > nsteps // 5 and ro["acc"][k == 10].sum() > 20
    assert np.array_equal(rg["kind"], ro["kind"])
    assert np.array_equal(rg["mol"], ro["mol"])
    assert np.array_equal(rg["ovl"], ro["ovl"])
    assert np.array_equal(rg["acc"], ro["acc"]), np.nonzero(rg["acc"] != ro["acc"])[0][:5]
    assert ng == no
    assert dE_close(rg["dlj"], ro["dlj"], e0[0])
    assert dE_close(rg["dreal"], ro["dreal"], e0[2])
    ok = ro["ovl"] == 0
    assert dE_close(rg["drecip"][ok], ro["drecip"][ok], e0[3])
    n, boxes = b.mol_boxes()
    no2, oboxes = o.mol_boxes()
    assert n == no2 and np.array_equal(boxes, oboxes)
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)
