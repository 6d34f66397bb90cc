"""Parity of the plain single-box device path with the CPU oracle.

Tolerances (FP64): whole-box energies agree to 1e-11 relative; a per-move
energy change agrees to 1e-10 relative to the larger of its own magnitude
and the magnitude of the matching box total (|E_inter|, |E_real|,
|E_recip|), the scale at which the Metropolis test reads it; the accept/reject sequence and the molecule/move-kind sequence
are identical; the cell list (CSR) is identical element for element.
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box
from oracle.pyoracle import Oracle


def dE_close(a, b, total, rtol=1e-10):
    """Per-move energy changes: |a-b| <= rtol * max(|b|, |box total|)."""
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), abs(total))))

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    s = systems.water_box(300, seed=11)
    return s


def _pair(s, threads=8):
    return Box(s), Oracle(s, threads=threads)


def test_cells_csr_identical(small):
    b, o = _pair(small)
    cv, cs, ac = b.build_cells()
    ocv, ocs, omap, edge = o.cell_csr()
    assert edge == b.edge
    assert np.array_equal(cs, ocs)
    assert np.array_equal(cv, ocv)
    assert np.array_equal(ac, omap)


def test_cells_csr_identical_multi_cell():
    s = systems.water_box(3000, seed=5)  # L = 44.8 A -> 4 cells per edge
    b, o = _pair(s)
    assert b.edge == (4, 4, 4)
    cv, cs, ac = b.build_cells()
    ocv, ocs, omap, _ = o.cell_csr()
    assert np.array_equal(cs, ocs) and np.array_equal(cv, ocv) and np.array_equal(ac, omap)


@pytest.mark.parametrize("nmol,seed", [(300, 11), (1000, 3), (3000, 5)])
def test_box_inter_and_totals(nmol, seed):
    s = systems.water_box(nmol, seed=seed)
    b, o = _pair(s)
    lj, real = b.box_inter()
    olj, oreal = o.box_inter()
    assert abs(lj - olj) <= 1e-11 * abs(olj)
    assert abs(real - oreal) <= 1e-11 * abs(oreal)
    tot, otot = b.system_total(), o.system_total()
    assert np.allclose(tot, otot, rtol=1e-11, atol=0.0), (tot, otot)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-11, atol=0.0)
    assert b.nkvec == o.nkvec
    kv, okv = b.recip_get(), o.kvectors()
    for key in ("kx", "ky", "kz", "prefact"):
        assert np.array_equal(kv[key], okv[key]), key
    sr, si = o.recip_sums()
    assert np.allclose(kv["sum_r"], sr, rtol=0, atol=1e-11 * np.abs(sr).max())
    assert np.allclose(kv["sum_i"], si, rtol=0, atol=1e-11 * np.abs(si).max())


def _trials(s, rng, n):
    """Random displaced / rotated molecule positions (wrapped)."""
    L = s.box[0]
    starts = s.mol_start()
    out = []
    for _ in range(n):
        m = int(rng.integers(s.nmols))
        st = starts[m]
        xyz = s.xyz[st:st + 3].copy()
        c = xyz[0]
        d = xyz - c
        d -= L * np.round(d / L)
        th = rng.uniform(-np.pi, np.pi)
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
        new = c + d @ R.T + rng.uniform(-1.5, 1.5, size=3)
        out.append((m, np.mod(new, L)))
    return out


def test_mol_trial_matches_oracle(small):
    b, o = _pair(small)
    e0 = o.energies()
    rng = np.random.default_rng(1)
    for m, new in _trials(small, rng, 40):
        lj, real, ov = b.mol_inter(m, new)
        olj, oreal, oov = o.mol_inter(m, new)
        assert ov == oov
        assert dE_close(lj, olj, e0[0])
        assert dE_close(real, oreal, e0[2])
        enew = b.mol_recip(m, new)
        od = o.mol_recip(m, new)
        assert dE_close(enew - e0[3], od, e0[3])
        fused = b.mol_trial(m, new)
        assert fused[0] == lj and fused[1] == real and fused[2] == enew and int(fused[3]) == ov


def test_swap_and_particle_inter(small):
    b, o = _pair(small)
    e0 = o.energies()
    rng = np.random.default_rng(2)
    L = small.box[0]
    kind = small.kinds[0]
    for _ in range(10):
        c = rng.uniform(0, L, size=3)
        xyz = np.mod(c + kind.geometry, L)
        for ins in (True, False):
            ge = b.swap_recip(0, ins, xyz)
            oe = o.swap_recip(0, ins, xyz)
            assert dE_close(ge - e0[3], oe, e0[3])
    trials = rng.uniform(0, L, size=(16, 3))
    for t, q in ((0, -0.8476), (1, 0.4238)):
        en, real, ov = b.particle_inter(t, q, trials)
        oen, oreal, oov = o.particle_inter(t, q, trials)
        assert np.array_equal(ov, oov)
        assert np.allclose(en, oen, rtol=1e-11, atol=1e-9)
        assert np.allclose(real, oreal, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("nmol,nsteps", [(300, 10000), (1000, 3000)])
def test_run_moves_same_sequence_as_oracle(nmol, nsteps):
    """GOMC's move loop with MT19937 draws: identical accept/reject,
    molecule and move-kind sequence; per-move energies and the final state
    within the stated tolerances."""
    s = systems.water_box(nmol, seed=21, adj_steps=500)
    b, o = _pair(s)
    e0 = o.energies()
    ng, rg = b.run(nsteps)
    no, ro = o.run(nsteps, mode=0)
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
    gd, gr = b.scales()
    od, orr = o.scales()
    assert np.allclose(gd, od, rtol=1e-12) and np.allclose(gr, orr, rtol=1e-12)
    # the running totals still equal a from-scratch recomputation
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)
