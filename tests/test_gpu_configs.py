"""The plain workloads at their full sizes on the device path.

* 1000 waters in a 31.07 A box (plain NVT, rc 10 A + LRC, Ewald 1e-5):
  the accept sequence of 2000 moves equals the oracle's under the same
  MT19937 stream; energies within 1e-10 relative.
* 10,000 waters in a 66.9 A box (rc 12 A, Ewald 1e-6, about 10^4
  k-vectors; MultiParticle 1:1 with translate/rotate): the whole-box
  totals and the pair, reciprocal and molecule forces and torques against
  the oracle at full size (forces within 1e-9 of the largest component);
  after a run of moves that includes MultiParticle moves, the running
  energies equal a from-scratch recomputation (1e-10 relative) -- the
  size-independent check of the MP transform and acceptance bookkeeping.
* GCMC methane / CO2 in a 40 A box (1200 molecules seeded, fugacities for
  about 2000, a reservoir of 3000): totals with the reservoir excluded, and
  400 moves (transfers with CBMC growth,
  translations, rotations) with the oracle's kind, molecule, overlap and
  accept sequences; final membership identical.
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def rel_close(a, b, tol=1e-9):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max()) <= tol * max(float(np.abs(b).max()), 1e-300)


def test_plain_nvt_1000_waters_same_sequence_as_oracle():
    s = systems.config1()
    b, o = Box(s), Oracle(s, threads=8)
    assert np.allclose(b.system_total(), o.system_total(), rtol=1e-11, atol=0)
    n = 2000
    ng, rg = b.run(n)
    no, ro = o.run(n, mode=0)
    assert np.array_equal(rg["acc"], ro["acc"]), np.nonzero(rg["acc"] != ro["acc"])[0][:5]
    assert ng == no
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=0)
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)


def test_large_box_10000_waters_forces_and_mp_bookkeeping():
    s = systems.config2()
    b = Box(s)
    o = Oracle(s, threads=8)
    assert np.allclose(b.system_total(), o.system_total(), rtol=1e-11, atol=0)
    lj, real, aF, mF = b.box_force()
    olj, oreal, oaF, omF = o.box_force()
    assert abs(lj - olj) <= 1e-11 * abs(olj) and abs(real - oreal) <= 1e-11 * abs(oreal)
    assert rel_close(aF, oaF) and rel_close(mF, omF)
    aFR, mFR = b.box_force_recip()
    oaFR, omFR = o.box_force_recip()
    assert rel_close(aFR, oaFR) and rel_close(mFR, omFR)
    assert rel_close(b.torque(), o.torque())
    del o
    ng, rg = b.run(60)
    assert (rg["kind"] == 2).sum() >= 10 and ng > 0
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)


def test_gcmc_40A_box_same_sequence_as_oracle():
    s = systems.config3()
    b, o = Box(s), Oracle(s, threads=8)
    assert np.allclose(b.system_total(), o.system_total(), rtol=1e-11, atol=1e-9)
    n = 400
    ng, rg = b.run(n)
    no, ro = o.run(n, mode=0)
    assert (ro["kind"] == 10).sum() > n // 5
    for key in ("kind", "mol", "ovl", "acc"):
        assert np.array_equal(rg[key], ro[key]), key
    assert ng == no
    n1, box1 = b.mol_boxes()
    n2, box2 = o.mol_boxes()
    assert n1 == n2 and np.array_equal(box1, box2)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)
