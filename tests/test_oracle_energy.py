"""Pins the oracle's energy functions against mathematics the reference's
formulas must reduce to, and against brute force:

* LJ: 4 eps [(s/r)^12 - (s/r)^6] (GOMC's Mie form at n = 12), its minimum
  -eps at 2^(1/6) s and zero at s; the virial is -dE/dr / r (finite
  difference).
* erfc-Coulomb: qq erfc(a r)/r with math.erfc; its virial by finite
  difference.
* LRC: the Allen-Tildesley tail (8/3) pi eps s^3 [(s/rc)^9/3 - (s/rc)^3].
* Cell list: the cell-list pair sum equals an O(N^2) sum over all pairs;
  the list is exhaustive; 3- and 4-cells-per-edge boxes.
* Ewald: the rock-salt lattice gives the Madelung constant 1.747565 and
  the total does not depend on the splitting (tolerance); self and
  intramolecular correction equal their closed forms; the k-vector count
  equals a direct count of half-space integer triples.
* Deltas: MoleculeInter / MolReciprocal equal the change of the full-box
  energies; SwapDestRecip / SwapSourceRecip equal the reciprocal energy with
  minus without the molecule; ParticleInter equals the change of BoxInter
  when one atom is added; the tail change equals the difference of tails.
* Running totals stay equal to a from-scratch recomputation (the
  reference's RecalculateAndCheck) over thousands of moves.
"""
import math

import numpy as np
import pytest

from gomc_amd import systems
from oracle import pyoracle as po
from oracle.pyoracle import Oracle

QQ = 167103.208067979


def _lj_system():
    return systems.water_box(50, box=25.0, seed=1)


def test_lj_closed_form_and_virial():
    s = _lj_system()
    sig, eps = 3.166, 78.2
    for r in [2.9, 3.166, 2 ** (1 / 6) * 3.166, 4.0, 7.5, 9.99]:
        t = po.pair_terms(s, 0, 0, r * r, 0.0)
        ref = 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
        assert abs(t[0] - ref) <= 1e-12 * max(abs(ref), 1e-3)
        h = 1e-6
        dEdr = (po.pair_terms(s, 0, 0, (r + h) ** 2, 0)[0] - po.pair_terms(s, 0, 0, (r - h) ** 2, 0)[0]) / (2 * h)
        assert abs(t[2] - (-dEdr / r)) <= 1e-6 * max(abs(dEdr / r), 1.0)
    assert abs(po.pair_terms(s, 0, 0, (2 ** (1 / 6) * sig) ** 2, 0)[0] + eps) < 1e-10
    assert abs(po.pair_terms(s, 0, 0, sig * sig, 0)[0]) < 1e-10
    assert po.pair_terms(s, 0, 0, 10.01 ** 2, 0)[0] == 0.0  # beyond Rcut
    # O-H and H-H carry no LJ (eps_H = 0)
    assert po.pair_terms(s, 0, 1, 4.0, 0)[0] == 0.0 and po.pair_terms(s, 1, 1, 4.0, 0)[0] == 0.0


def test_coulomb_closed_form_and_virial():
    s = _lj_system()
    alpha = math.sqrt(-math.log(1e-5)) / 10.0
    qq = -0.8476 * 0.4238 * QQ
    for r in [1.0, 2.7, 5.0, 9.9]:
        t = po.pair_terms(s, 0, 1, r * r, qq)
        ref = qq * math.erfc(alpha * r) / r
        assert abs(t[1] - ref) <= 1e-14 * abs(ref) + 1e-12
        h = 1e-6
        e = lambda rr: po.pair_terms(s, 0, 1, rr * rr, qq)[1]
        dEdr = (e(r + h) - e(r - h)) / (2 * h)
        assert abs(t[3] - (-dEdr / r)) <= 1e-6 * abs(dEdr / r)
    assert po.pair_terms(s, 0, 1, 10.01 ** 2, qq)[1] == 0.0


def test_lrc_allen_tildesley():
    s = _lj_system()
    sig, eps, rc = 3.166, 78.2, 10.0
    ref = (8.0 / 3.0) * math.pi * eps * sig ** 3 * ((sig / rc) ** 9 / 3 - (sig / rc) ** 3)
    t = po.pair_terms(s, 0, 0, 1.0, 0.0)[4]
    assert abs(t - ref) <= 1e-13 * abs(ref)
    o = Oracle(s)
    n, V = s.nmols, s.box[0] ** 3
    assert abs(o.self_corr_lrc()[2] - ref * n * n / V) <= 1e-12 * abs(ref * n * n / V)
    # adding one molecule changes the tail by (2N+1) ref / V
    assert abs(o.tail_change(0, True) - ref * (2 * n + 1) / V) <= 1e-12 * abs(ref * (2 * n + 1) / V)


@pytest.mark.parametrize("nmol,edge", [(300, 3), (3000, 4)])
def test_cell_list_equals_brute_force(nmol, edge):
    s = systems.water_box(nmol, seed=9)
    o = Oracle(s, threads=8)
    assert o.cell_csr()[3] == (edge,) * 3
    assert o.cells_exhaustive()
    lj, real = o.box_inter()
    blj, breal = o.box_inter_brute()
    assert abs(lj - blj) <= 1e-12 * abs(blj)
    assert abs(real - breal) <= 1e-12 * abs(breal)
    # CSR: each cell's atoms ascending, every atom once, cell index matches
    cv, cs, mp, _ = o.cell_csr()
    assert sorted(cv.tolist()) == list(range(o.natoms))
    for c in range(len(cs) - 1):
        seg = cv[cs[c]:cs[c + 1]]
        assert np.all(np.diff(seg) > 0) and np.all(mp[seg] == c)


def _rocksalt(n: int, a: float, tol: float) -> systems.System:
    idx = np.array([(i, j, k) for i in range(n) for j in range(n) for k in range(n)])
    xyz = (idx + 0.25) * a  # keep every ion off the box faces
    mk = (idx.sum(axis=1) % 2).astype(np.int32)
    kinds = [systems.MolKind("NA", [0], [1.0], np.zeros((1, 3))),
             systems.MolKind("CL", [0], [-1.0], np.zeros((1, 3)))]
    L = n * a
    return systems.System(box=(L, L, L), temperature=300.0, rcut=L / 2, rcut_low=0.5,
                          rcut_coulomb=L / 2, tolerance=tol, sigma=[1.0], epsilon=[0.0],
                          n_rep=[12.0], kinds=kinds, mol_kind=mk, xyz=xyz, lrc=False)


@pytest.mark.parametrize("tol", [1e-10, 1e-8])
def test_ewald_madelung_rocksalt(tol):
    n, a = 6, 2.8
    o = Oracle(_rocksalt(n, a, tol), threads=8)
    e = o.system_total()
    N = n ** 3
    madelung = -e[6] * 2 * a / (N * QQ)
    assert abs(madelung - 1.747565) < 2e-6, madelung
    assert e[0] == 0.0 and e[5] == 0.0  # no LJ, single-site "molecules"


def test_self_and_correction_closed_forms():
    s = systems.water_box(60, box=22.0, seed=2)
    o = Oracle(s)
    alpha = o.cutoffs()["alpha"]
    q = np.array([-0.8476, 0.4238, 0.4238])
    self_ref = -alpha / math.sqrt(math.pi) * QQ * s.nmols * float((q * q).sum())
    corr_ref = 0.0
    L = s.box[0]
    for m in range(s.nmols):
        xyz = s.xyz[3 * m:3 * m + 3]
        for i in range(3):
            for j in range(i + 1, 3):
                d = xyz[i] - xyz[j]
                d -= L * np.round(d / L)
                r = float(np.linalg.norm(d))
                corr_ref -= QQ * q[i] * q[j] * math.erf(alpha * r) / r
    got = o.self_corr_lrc()
    assert abs(got[0] - self_ref) <= 1e-13 * abs(self_ref)
    assert abs(got[1] - corr_ref) <= 1e-12 * abs(corr_ref)


@pytest.mark.parametrize("L,rc,tol,expect", [(31.07, 10, 1e-5, 3093), (66.9, 12, 1e-6, 30902),
                                             (40.0, 10, 1e-5, 6602), (20.0, 10, 1e-5, 823),
                                             (60.0, 10, 1e-5, 22197)])
def test_kvector_count(L, rc, tol, expect):
    kc = -2 * math.log(tol) / rc
    nmax = int(kc * L / (2 * math.pi)) + 1
    c = 2 * math.pi / L
    r = np.arange(-nmax, nmax + 1)
    X, Y, Z = np.meshgrid(np.arange(0, nmax + 1), r, r, indexing="ij")
    half = (X > 0) | ((X == 0) & (Y > 0)) | ((X == 0) & (Y == 0) & (Z > 0))
    direct = int(np.count_nonzero(half & ((c * c) * (X * X + Y * Y + Z * Z) < kc * kc)))
    assert direct == expect
    if L < 45:  # build the oracle's own list for the smaller boxes
        s = systems.water_box(20, box=L, rcut=rc, tolerance=tol, seed=3)
        assert Oracle(s).nkvec == expect


def _moved(s, m, new):
    xyz = s.xyz.copy()
    st = s.mol_start()[m]
    xyz[st:st + len(new)] = new
    return systems.System(**{**s.__dict__, "xyz": xyz, "_keep": []})


def test_deltas_equal_full_recomputation():
    s = systems.water_box(200, seed=4)
    o = Oracle(s, threads=8)
    e0 = o.energies()
    rng = np.random.default_rng(0)
    L = s.box[0]
    for _ in range(6):
        m = int(rng.integers(s.nmols))
        new = np.mod(s.xyz[3 * m:3 * m + 3] + rng.uniform(-0.7, 0.7, 3), L)
        lj, real, ov = o.mol_inter(m, new)
        d_rec = o.mol_recip(m, new)
        e1 = Oracle(_moved(s, m, new), threads=8).energies()
        assert abs(lj - (e1[0] - e0[0])) <= 1e-9 * abs(e0[0])
        assert abs(real - (e1[2] - e0[2])) <= 1e-9 * abs(e0[2])
        assert abs(d_rec - (e1[3] - e0[3])) <= 1e-10 * abs(e0[3])


def test_swap_recip_equals_with_minus_without():
    s = systems.water_box(120, seed=6)
    m = 17
    keep = np.array([i for i in range(s.nmols) if i != m])
    xyz_wo = np.concatenate([s.xyz[3 * i:3 * i + 3] for i in keep])
    s_wo = systems.System(**{**s.__dict__, "xyz": xyz_wo, "mol_kind": s.mol_kind[keep], "_keep": []})
    o_with, o_wo = Oracle(s, threads=8), Oracle(s_wo, threads=8)
    pos = s.xyz[3 * m:3 * m + 3]
    e_with, e_wo = o_with.energies()[3], o_wo.energies()[3]
    assert abs(o_wo.swap_recip(0, True, pos) - (e_with - e_wo)) <= 1e-10 * abs(e_with)
    # removal needs the molecule's own atoms in the box sums: delete m from s
    assert abs(o_with.swap_recip(0, False, pos) - (e_wo - e_with)) <= 1e-10 * abs(e_with)


def test_particle_inter_equals_box_inter_change():
    s = systems.mixture_box(60, 20, box=25.0, seed=8)
    o = Oracle(s, threads=8)
    lj0, real0 = o.box_inter()
    rng = np.random.default_rng(3)
    trials = rng.uniform(0, 25.0, size=(8, 3))
    en, real, ov = o.particle_inter(0, 0.0, trials)  # a methane atom
    for t in range(8):
        s2 = systems.System(**{**s.__dict__, "xyz": np.vstack([s.xyz, trials[t]]),
                               "mol_kind": np.append(s.mol_kind, 0).astype(np.int32), "_keep": []})
        lj1, real1 = Oracle(s2, threads=8).box_inter()
        assert abs(en[t] - (lj1 - lj0)) <= 1e-9 * max(abs(lj0), 1.0)
        assert abs(real[t] - (real1 - real0)) <= 1e-9 * max(abs(real0), 1.0)


@pytest.mark.parametrize("mode", [0, 1])
def test_running_totals_match_recomputation(mode):
    s = systems.water_box(250, seed=12, adj_steps=300)
    o = Oracle(s, threads=8)
    o.run(4000, mode=mode)
    run, ref = o.energies(), o.system_total()
    assert np.allclose(run, ref, rtol=1e-12, atol=1e-8)
