"""NPT volume moves on the device path (VolumeTransfer with one box,
src/moves/VolumeTransfer.h:74-310) against the oracle under the same
MT19937 stream: identical move-kind, molecule, overlap and accept
sequences over a mix of translations, rotations and volume moves; per-move
energy changes within 1e-10 of the box totals; the box volume, running
energies (1e-10), coordinates (1e-9 A) and the virial and pressure of the
final box equal the oracle's.
"""
from dataclasses import replace

import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box, GomcError
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def dE_close(a, b, total, rtol=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), abs(total))))


def npt_water(n=300, seed=11):
    return replace(systems.water_box(n, seed=seed, rcut=8.0), ensemble=2, pressure=1.0,
                   move_vol=0.05, adj_steps=150, _keep=[])


def test_npt_same_sequence_as_oracle():
    s = npt_water()
    b, o = Box(s), Oracle(s, threads=8)
    e0 = o.energies()
    n = 2500
    ng, rg = b.run(n)
    no, ro = o.run(n, mode=0)
    k = ro["kind"]
    assert (k == 9).sum() > 60 and ro["acc"][k == 9].sum() > 5
    for key in ("kind", "mol", "ovl", "acc"):
        assert np.array_equal(rg[key], ro[key]), (key, np.nonzero(rg[key] != ro[key])[0][:5])
    assert ng == no
    for key, idx in (("dlj", 0), ("dreal", 2), ("drecip", 3)):
        assert dE_close(rg[key], ro[key], e0[idx]), key
    assert b.dims()[3] == pytest.approx(o.volume(), rel=1e-13)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    vd, vo = b.virial(), o.virial()
    for key in ("inter", "tail", "real", "recip", "pressure"):
        assert vd[key] == pytest.approx(vo[key], rel=1e-9), key


def test_device_loop_refuses_volume_moves():
    b = Box(npt_water(150, seed=3))
    with pytest.raises(GomcError):
        b.run(10, device=True)
