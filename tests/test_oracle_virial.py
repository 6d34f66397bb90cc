"""The oracle's virial and pressure (CalculateEnergy::VirialCalc,
Ewald::VirialReciprocal, VirialCorrection, OutputVars' pressure) pinned
without the reference:

* the molecular virial of each energy term is -3V dU/dV under a uniform
  scaling of the molecule centres with rigid molecules: checked against
  central differences of the oracle's own energies (LJ, real-space Coulomb,
  reciprocal) at V(1 +- 3e-7), to 1e-7 relative (they agree to 2e-10,
  4e-10 and 2e-8), on a box whose cutoff sphere plus a
  molecule's size stays inside half the box (so the centre and atom images
  agree) and where no pair and no k-vector crosses its cutoff between the
  two volumes;
* the tail virial equals the textbook LJ tail pressure
  (16/3) pi rho^2 eps sigma^3 [(2/3)(sigma/rc)^9 - (sigma/rc)^3] times 3V;
* the pressure is the ideal-gas term plus W / 3V in bar.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

from gomc_amd import systems
from oracle.pyoracle import Oracle


def _coms(s):
    """Centres as the reference computes them (atom 0 as the unwrap anchor)."""
    L = np.asarray(s.box)
    xyz = np.asarray(s.xyz)
    out, a = [], 0
    for k in s.mol_kind:
        n = len(s.kinds[k].atom_types)
        p = xyz[a:a + n]
        d = p - p[0]
        d -= L * np.round(d / L)
        out.append((p[0] + d.mean(0)) % L)
        a += n
    return np.array(out)


def _scaled(s, f):
    """The system with box L f and every molecule moved rigidly so that its
    centre scales by f."""
    L = np.asarray(s.box)
    com = _coms(s)
    xyz = np.asarray(s.xyz).copy()
    a = 0
    for m, k in enumerate(s.mol_kind):
        n = len(s.kinds[k].atom_types)
        d = xyz[a:a + n] - com[m]
        d -= L * np.round(d / L)
        xyz[a:a + n] = (com[m] * f + d) % (L * f)
        a += n
    return replace(s, box=tuple(L * f), xyz=xyz, _keep=[])


def _pairs_inside(s):
    xyz = np.asarray(s.xyz)
    L = np.asarray(s.box)
    d = xyz[:, None, :] - xyz[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(-1)
    return r2 < s.rcut * s.rcut


def _n_kvectors(s):
    alpha = math.sqrt(-math.log(s.tolerance)) / s.rcut_coulomb
    kc = -2.0 * math.log(s.tolerance) / s.rcut_coulomb
    L = s.box[0]
    c = 2 * math.pi / L
    nmax = int(kc * L / (2 * math.pi)) + 1
    n = np.arange(-nmax, nmax + 1)
    gx, gy, gz = np.meshgrid(n, n, n, indexing="ij")
    k2 = c * c * (gx * gx + gy * gy + gz * gz)
    half = (gx > 0) | ((gx == 0) & (gy > 0)) | ((gx == 0) & (gy == 0) & (gz > 0))
    return int(np.count_nonzero(half & (k2 < kc * kc))), alpha


@pytest.fixture(scope="module")
def water():
    return systems.water_box(150, seed=21, rcut=5.5)


def test_virial_is_minus_3V_dU_dV(water):
    s = water
    assert s.box[0] / 2 > s.rcut + 2.0
    h = 3e-7
    lo, hi = _scaled(s, 1 - h), _scaled(s, 1 + h)
    assert np.array_equal(_pairs_inside(lo), _pairs_inside(hi))
    assert _n_kvectors(lo)[0] == _n_kvectors(hi)[0] == _n_kvectors(s)[0]
    v = Oracle(s, threads=4).virial()
    e_lo, e_hi = Oracle(lo, threads=4).system_total(), Oracle(hi, threads=4).system_total()
    V, dV = s.box[0] ** 3, s.box[0] ** 3 * ((1 + h) ** 3 - (1 - h) ** 3)
    for name, idx in (("inter", 0), ("real", 2), ("recip", 3)):
        w_fd = -3.0 * V * (e_hi[idx] - e_lo[idx]) / dV
        assert abs(v[name] - w_fd) <= 1e-7 * max(abs(w_fd), 1.0), (name, v[name], w_fd)
    # self and intramolecular correction do not change with the centres
    assert abs(e_hi[4] - e_lo[4]) <= 1e-9 * abs(e_lo[4])
    assert abs(e_hi[5] - e_lo[5]) <= 1e-9 * abs(e_lo[5])


def test_tail_virial_is_the_textbook_lj_tail_pressure(water):
    s = water
    v = Oracle(s).virial()
    N, V = len(s.mol_kind), s.box[0] ** 3
    eps, sig, rc = s.epsilon[0], s.sigma[0], s.rcut
    x = sig / rc
    p_tail = 16.0 / 3.0 * math.pi * (N / V) ** 2 * eps * sig ** 3 * (2.0 / 3.0 * x ** 9 - x ** 3)
    assert v["tail"] == pytest.approx(3.0 * V * p_tail, rel=1e-12)


def test_pressure_is_ideal_plus_virial_over_3V(water):
    s = water
    v = Oracle(s).virial()
    N, V = len(s.mol_kind), s.box[0] ** 3
    p = (N / V * s.temperature + v["total"] / (3.0 * V)) * 138.0649
    assert v["pressure"] == pytest.approx(p, rel=1e-12)
    assert v["total"] == pytest.approx(v["inter"] + v["tail"] + v["real"] + v["recip"], rel=1e-14)
    assert v["tensor"].sum() == pytest.approx(v["inter"] + v["real"] + v["recip"], rel=1e-10)
