"""The oracle's GCMC move (MoleculeTransfer with CBMC growth) against what
statistical mechanics fixes, and its bookkeeping against recomputation.

* Ideal gas: with no interactions the grand-canonical mean occupation of
  each kind is <N_k> = beta f_k V for fugacity f_k.  A dropped volume or
  beta factor in the acceptance coefficient (MoleculeTransfer.h:143-165), or
  Rosenbluth weights of the new and old molecules normalised differently,
  moves <N> far outside the tolerance.
* Ideal gas at low occupation: the whole distribution of N_k is Poisson
  with mean beta f_k V.  With <N> near 1 an N -> N + 1 slip in the
  insertion coefficient (or a deletion coefficient off by one) changes
  P(0), P(1), P(2) by tens of percent, which the chi-square test sees at
  any seed (a planted slip gives p < 1e-100).
* CBMC weighting with interactions on: the Rosenbluth weights of
  DCSingle / DCFreeHedron (DCSingle.cpp:23-99, DCFreeHedron.cpp:85-238)
  and PRNG::PickWeighted (PRNG.h:167-203) change how insertions are
  proposed, not what is sampled, so <N_k> with 16 first-atom and 8
  later-atom trials equals <N_k> with one trial each (plain random
  insertion).  A builder that keeps the wrong trial (planted: the one
  after the weighted pick) inserts molecules where its weight does not
  apply, and <N_CO2> falls by a third.
* Interacting mixture: after thousands of insertions, deletions,
  translations and rotations the running energies equal a from-scratch
  recomputation of the final state.
"""
import numpy as np
from scipy import stats

from gomc_amd import systems
from oracle.pyoracle import Oracle, planted


def test_ideal_gas_mean_occupation():
    T, L = 300.0, 25.0
    target = np.array([40.0, 25.0])
    fug_bar = target / ((1.0 / T) * L**3) / systems.BAR_TO_K_PER_A3
    s = systems.gcmc_mixture(40, 25, L, temperature=T, fugacity_bar=tuple(fug_bar),
                             reservoir=120, seed=9, swap=0.6, disp=0.2, rot=0.2)
    s.epsilon = [0.0, 0.0, 0.0]
    s.electrostatic = False
    s.ewald = False
    s.rcut_low = 1e-3
    o = Oracle(s, threads=4)
    counts = []
    mk = s.mol_kind
    for _ in range(400):
        o.run(500, mode=0, record=False)
        n, b = o.mol_boxes()
        counts.append([np.sum((b == 0) & (mk == 0)), np.sum((b == 0) & (mk == 1))])
    c = np.array(counts[20:], dtype=float)
    mean, var = c.mean(0), c.var(0)
    # blocks of 500 moves are close to independent: standard error of the
    # mean about sqrt(<N> / samples) (0.3 here); 4.5 of them is the bar.  A
    # dropped beta, volume or fugacity factor moves <N> by orders of
    # magnitude; an (N + 1) -> N slip moves it by one molecule, about three
    # standard errors, so this check alone does not rule that slip out.
    sem = np.sqrt(target / len(c))
    assert np.all(np.abs(mean - target) < 4.5 * sem), (mean, target, sem)
    # and the occupation is Poisson: variance equals the mean
    assert np.all(np.abs(var / target - 1.0) < 0.3), (var, target)


def test_running_energy_equals_recompute_after_transfers():
    s = systems.gcmc_mixture(60, 60, 25.0, fugacity_bar=(200.0, 60.0), reservoir=120, seed=3)
    o = Oracle(s, threads=4)
    n, rec = o.run(4000, mode=0)
    k = rec["kind"]
    assert (k == 10).sum() > 1000
    assert rec["acc"][k == 10].sum() > 100
    e, tot = o.energies(), o.system_total()
    assert np.allclose(e, tot, rtol=1e-10, atol=1e-6), (e, tot)


def test_empty_box_fills_and_running_energy_equals_recompute():
    """GCMC from an empty box (every molecule in the reservoir): zero
    energies at the start, insertions accepted, and the running energies
    equal a from-scratch recomputation afterwards."""
    s = systems.gcmc_mixture(0, 0, 25.0, fugacity_bar=(200.0, 60.0), reservoir=120, seed=3,
                             adj_steps=250)
    o = Oracle(s, threads=4)
    assert o.mol_boxes()[0] == 0 and np.all(o.energies() == 0.0)
    n, rec = o.run(3000)
    k = rec["kind"]
    assert rec["acc"][k == 10].sum() > 100 and o.mol_boxes()[0] > 50
    assert np.allclose(o.energies(), o.system_total(), rtol=1e-10, atol=1e-6)


def _ideal_pn(lam, samples, every, seed):
    """N_k sampled every `every` moves in an ideal-gas GCMC box (20 A,
    300 K) whose fugacities give <N_k> = lam."""
    T, L = 300.0, 20.0
    target = np.asarray(lam)
    fug_bar = target / ((1.0 / T) * L**3) / systems.BAR_TO_K_PER_A3
    s = systems.gcmc_mixture(2, 1, L, temperature=T, fugacity_bar=tuple(fug_bar),
                             reservoir=30, seed=seed, swap=0.8, disp=0.1, rot=0.1)
    s.epsilon = [0.0, 0.0, 0.0]
    s.electrostatic = False
    s.ewald = False
    s.rcut_low = 1e-3
    o = Oracle(s, threads=1)
    mk = s.mol_kind
    o.run(2000, record=False)
    c = []
    for _ in range(samples):
        o.run(every, record=False)
        _, b = o.mol_boxes()
        c.append([np.sum((b == 0) & (mk == 0)), np.sum((b == 0) & (mk == 1))])
    return np.array(c)


def _poisson_chi2_p(counts, lam, nmax=6):
    obs = np.array([np.sum(counts == n) for n in range(nmax)] + [np.sum(counts >= nmax)])
    p = stats.poisson.pmf(np.arange(nmax), lam)
    exp = np.append(p, 1.0 - p.sum()) * len(counts)
    return stats.chi2.sf(((obs - exp) ** 2 / exp).sum(), len(obs) - 1)


def test_ideal_gas_occupation_is_poisson_near_one_molecule():
    lam = (1.5, 1.0)
    c = _ideal_pn(lam, samples=6000, every=150, seed=9)
    for k in range(2):
        p = _poisson_chi2_p(c[:, k], lam[k])
        assert p > 1e-4, (k, p, np.bincount(c[:, k]))
    # the check sees an N -> N + 1 slip in the insertion coefficient
    with planted(3):
        cp = _ideal_pn(lam, samples=1500, every=150, seed=9)
    assert max(_poisson_chi2_p(cp[:, k], lam[k]) for k in range(2)) < 1e-20


def _mixture_mean_n(first, nth, seed, blocks=40, per=4000):
    """<N_k> and its block standard error in an interacting methane/CO2
    GCMC box (20 A, 300 K, rc 8 A, Ewald) from 20 molecules."""
    s = systems.gcmc_mixture(10, 10, 20.0, temperature=300.0, fugacity_bar=(60.0, 20.0),
                             rcut=8.0, reservoir=150, seed=seed, swap=0.6, disp=0.2, rot=0.2,
                             cbmc_first=first, cbmc_nth=nth, cbmc_ang=100, adj_steps=500)
    o = Oracle(s, threads=4)
    mk = s.mol_kind
    o.run(10000, record=False)
    bm = []
    for _ in range(blocks):
        c = []
        for _ in range(per // 100):
            o.run(100, record=False)
            _, b = o.mol_boxes()
            c.append([np.sum((b == 0) & (mk == 0)), np.sum((b == 0) & (mk == 1))])
        bm.append(np.mean(c, 0))
    bm = np.array(bm)
    return bm.mean(0), bm.std(0, ddof=1) / np.sqrt(len(bm))


def test_cbmc_trials_do_not_change_what_is_sampled():
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(2) as ex:  # the oracle releases the GIL
        f16 = ex.submit(_mixture_mean_n, 16, 8, 10)
        f1 = ex.submit(_mixture_mean_n, 1, 1, 10)
        (m16, e16), (m1, e1) = f16.result(), f1.result()
    sig = np.sqrt(e16 ** 2 + e1 ** 2)
    assert np.all(np.abs(m16 - m1) < 3.5 * sig), (m16, m1, sig)
    assert m16[0] > 10 and m16[1] > 5   # an interacting, occupied box
    # a builder that keeps the trial after the weighted pick is seen
    with planted(2):
        mp, ep = _mixture_mean_n(16, 8, seed=10)
    assert abs(mp[1] - m16[1]) > 6.0 * np.sqrt(ep[1] ** 2 + e16[1] ** 2), (mp, m16)
