"""The oracle's forces and torques against finite differences of its own
energies, so that BoxForce / BoxForceReciprocal / CalculateTorque
(CalculateEnergy.cpp:268-406,1365-1406; Ewald.cpp:1496-1596) are pinned by
F = -dE/dr rather than by a second copy of their formulas.

The system is a cluster of rigid waters in a box much larger than the
cluster with a cutoff that holds every intermolecular pair of the cluster
and no periodic image, so the pair energy is smooth in every coordinate
the differences move.  The reciprocal force is checked against the
derivative of E_recip + E_correction (the reciprocal force subtracts the
intramolecular erf term, Ewald.cpp:1541-1555).
"""
import numpy as np
import pytest

from gomc_amd import systems
from oracle.pyoracle import Oracle, rotation

H = 1e-5  # A (or rad) central-difference step


def cluster(nmol=6, box=30.0, seed=3):
    rng = np.random.default_rng(seed)
    kind = systems.water_kind()
    geo = np.asarray(kind.geometry, dtype=np.float64)
    geo = geo - geo.mean(0)
    centres = []
    while len(centres) < nmol:
        c = box / 2 + rng.uniform(-4.5, 4.5, size=3)
        if all(np.linalg.norm(c - o) > 3.0 for o in centres):
            centres.append(c)
    xyz = []
    for c in centres:
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        R = rotation(rng.uniform(0, np.pi), ax)
        xyz.append(c + geo @ R.T)
    xyz = np.concatenate(xyz)
    return systems.System(box=(box, box, box), temperature=300.0, rcut=14.0, rcut_low=1.0,
                          rcut_coulomb=14.0, tolerance=1e-6, kinds=[kind],
                          mol_kind=np.zeros(nmol, np.int32), xyz=xyz, seed=1,
                          **systems.WATER_TYPES)


def energies(s, xyz):
    s2 = systems.System(**{**s.__dict__, "xyz": xyz, "_keep": []})
    o = Oracle(s2)
    e = o.system_total()  # inter, tc, real, recip, self, corr, total
    o.close()
    return e[0] + e[2], e[3] + e[5]


@pytest.fixture(scope="module")
def sys_and_forces():
    s = cluster()
    o = Oracle(s)
    lj, real, aF, mF = o.box_force()
    aFR, mFR = o.box_force_recip()
    tq = o.torque()
    com = o.com()
    e = o.system_total()
    o.close()
    return s, dict(lj=lj, real=real, aF=aF, mF=mF, aFR=aFR, mFR=mFR, tq=tq, com=com, e=e)


def test_box_force_energies_equal_box_inter(sys_and_forces):
    s, f = sys_and_forces
    assert f["lj"] == pytest.approx(f["e"][0], rel=1e-12)
    assert f["real"] == pytest.approx(f["e"][2], rel=1e-12)
    # Newton's third law over the whole box, and molecule sums of atom forces
    assert np.abs(f["aF"].sum(0)).max() < 1e-9 * np.abs(f["aF"]).max()
    ms = s.mol_start()
    for m in range(s.nmols):
        assert np.allclose(f["mF"][m], f["aF"][ms[m]:ms[m] + 3].sum(0), rtol=1e-12, atol=1e-9)
        assert np.allclose(f["mFR"][m], f["aFR"][ms[m]:ms[m] + 3].sum(0), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("atom", [0, 1, 2, 7, 16])
def test_atom_forces_are_minus_gradient(sys_and_forces, atom):
    s, f = sys_and_forces
    for d in range(3):
        xp, xm = s.xyz.copy(), s.xyz.copy()
        xp[atom, d] += H
        xm[atom, d] -= H
        (pp, ep), (pm, em) = energies(s, xp), energies(s, xm)
        fd_pair = -(pp - pm) / (2 * H)
        fd_ewald = -(ep - em) / (2 * H)
        scale = max(1.0, np.abs(f["aF"][atom]).max())
        assert abs(f["aF"][atom, d] - fd_pair) < 2e-5 * scale, (atom, d)
        scale = max(1.0, np.abs(f["aFR"][atom]).max())
        assert abs(f["aFR"][atom, d] - fd_ewald) < 2e-5 * scale, (atom, d)


@pytest.mark.parametrize("mol", [0, 3])
def test_torque_is_minus_angular_derivative(sys_and_forces, mol):
    s, f = sys_and_forces
    st = s.mol_start()[mol]
    c = f["com"][mol]
    for d in range(3):
        ax = np.zeros(3)
        ax[d] = 1.0
        vals = []
        for sgn in (1, -1):
            R = rotation(sgn * H, ax)
            x = s.xyz.copy()
            x[st:st + 3] = c + (x[st:st + 3] - c) @ R.T
            p, e = energies(s, x)
            vals.append(p + e)
        fd = -(vals[0] - vals[1]) / (2 * H)
        scale = max(1.0, np.abs(f["tq"][mol]).max())
        assert abs(f["tq"][mol, d] - fd) < 2e-5 * scale, (mol, d)


def test_molecule_force_is_minus_translation_derivative(sys_and_forces):
    s, f = sys_and_forces
    mol = 2
    st = s.mol_start()[mol]
    for d in range(3):
        vals = []
        for sgn in (1, -1):
            x = s.xyz.copy()
            x[st:st + 3, d] += sgn * H
            p, e = energies(s, x)
            vals.append(p + e)
        fd = -(vals[0] - vals[1]) / (2 * H)
        tot = f["mF"][mol, d] + f["mFR"][mol, d]
        assert abs(tot - fd) < 2e-5 * max(1.0, abs(tot)), d
