"""Parity of the MultiParticle pieces of the device path with the oracle
(BoxForce, BoxForceReciprocal, CalculateTorque, and the whole
MultiParticle move inside GOMC's move loop).

Tolerances (FP64), also for Mie exponents other than 12 and Ewald
tolerances that take alpha r beyond 4: pair energies 1e-11 relative; every force and torque
component within 1e-9 of the largest component of its array (the sums run
in a different order on the device); the move loop with MultiParticle
moves keeps the identical move-kind, move-type, accept/reject and
transform-failure sequence, per-move energy changes within 1e-10 of the
box totals, and the final state within 1e-9 A.
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


def rel_close(a, b, tol=1e-9):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max()) <= tol * max(float(np.abs(b).max()), 1e-300)


CASES = [(300, 10.0, 11), (1000, 7.0, 3)]  # 3 and 4 cells per edge


@pytest.mark.parametrize("nmol,rcut,seed", CASES)
def test_box_force_matches_oracle(nmol, rcut, seed):
    s = systems.water_box(nmol, rcut=rcut, seed=seed)
    b, o = Box(s), Oracle(s, threads=8)
    lj, real, aF, mF = b.box_force()
    olj, oreal, oaF, omF = o.box_force()
    assert abs(lj - olj) <= 1e-11 * abs(olj)
    assert abs(real - oreal) <= 1e-11 * abs(oreal)
    assert rel_close(aF, oaF)
    assert rel_close(mF, omF)


@pytest.mark.parametrize("n_rep,tol", [(13.0, 1e-5), (16.0, 1e-5), (15.3, 1e-5), (12.0, 1e-9)])
def test_box_force_mie_exponents_and_tight_ewald(n_rep, tol):
    """The pair-force kernel's own forms: Mie repulsion (sigma^2/r^2)^(n/2)
    by products for whole and half-integer n/2 and by exp(n/2 log x)
    otherwise, against the oracle's pow; erfc(alpha r) beyond alpha r = 4
    (tolerance 1e-9: alpha rc = 4.55) by the continued fraction, against
    libm's erfc."""
    from dataclasses import replace
    s = systems.water_box(400, rcut=9.0, seed=17, tolerance=tol)
    s = replace(s, n_rep=[n_rep, 12.0], _keep=[])
    b, o = Box(s), Oracle(s, threads=8)
    lj, real, aF, mF = b.box_force()
    olj, oreal, oaF, omF = o.box_force()
    assert abs(lj - olj) <= 1e-11 * abs(olj)
    assert abs(real - oreal) <= 1e-11 * abs(oreal)
    assert rel_close(aF, oaF)
    assert rel_close(mF, omF)


@pytest.mark.parametrize("nmol,rcut,seed", CASES)
def test_box_force_recip_and_torque_match_oracle(nmol, rcut, seed):
    s = systems.water_box(nmol, rcut=rcut, seed=seed)
    b, o = Box(s), Oracle(s, threads=8)
    b.box_force()
    aFR, mFR = b.box_force_recip()
    oaFR, omFR = o.box_force_recip()
    assert rel_close(aFR, oaFR)
    assert rel_close(mFR, omFR)
    tq = b.torque()
    assert rel_close(tq, o.torque())


@pytest.mark.parametrize("nmol,nsteps", [(200, 600), (500, 300)])
def test_run_moves_with_multiparticle_same_sequence(nmol, nsteps):
    """MultiParticle interleaved 1:1 with translate/rotate (the large-box
    water workload's move mix) through the driver, against the oracle."""
    s = systems.water_box(nmol, seed=21, adj_steps=100, move_disp=0.25, move_rot=0.25,
                          move_mp=0.5)
    b, o = Box(s), Oracle(s, threads=8)
    e0 = o.energies()
    ng, rg = b.run(nsteps)
    no, ro = o.run(nsteps, mode=0)
    assert (ro["kind"] == 2).sum() > nsteps // 4
    assert np.array_equal(rg["kind"], ro["kind"])
    assert np.array_equal(rg["mol"], ro["mol"])
    assert np.array_equal(rg["ovl"], ro["ovl"])
    assert np.array_equal(rg["acc"], ro["acc"]), np.nonzero(rg["acc"] != ro["acc"])[0][:5]
    assert ng == no
    ok = ro["ovl"] == 0
    assert dE_close(rg["dlj"], ro["dlj"], e0[0])
    assert dE_close(rg["dreal"], ro["dreal"], e0[2])
    assert dE_close(rg["drecip"][ok], ro["drecip"][ok], e0[3])
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=0)
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    assert np.allclose(b.mp_settings(), o.mp_settings(), rtol=1e-12)
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)


def test_mp_entry_points_reject_and_restore():
    """gomc_mp_trial + gomc_mp_accept(reject) leave the box as it was."""
    s = systems.water_box(200, seed=4)
    b = Box(s)
    x0 = b.coords().copy()
    e0 = b.system_total()
    b.box_force()
    b.box_force_recip()
    b.torque()
    out = b.mp_trial(0, 7, 0.02)
    assert out[4] == 0.0 and np.isfinite(out[3]) and out[3] > 0
    b.mp_accept(False, out)
    assert np.array_equal(b.coords(), x0)
    assert np.allclose(b.system_total(), e0, rtol=1e-12, atol=0)
