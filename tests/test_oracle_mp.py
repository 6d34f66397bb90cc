"""The oracle's MultiParticle move samples the Boltzmann distribution.

The force-biased proposal of MultiParticle::CalculateTrialDistRot
(src/moves/MultiParticle.h:546-611) draws each molecule's shift (or
rotation vector) k from p(k) = prod_d c_d exp(c_d k_d) / (2 sinh(c_d max))
with c = lambda beta F (or the torque), and GetCoeff / CalculateWRatio
(MultiParticle.h:443-511) weigh the move by the reverse proposal over the
forward one.  If that ratio is right, a run of MultiParticle moves only
samples the same canonical distribution as a run of single-molecule
translations and rotations (plain Metropolis, symmetric proposals), so the
mean energies agree within the spread of independent chains.  The ratio with the new and old
forces swapped (planted in the oracle) drives the mean energy of this box
up by about 2500 K per molecule.

The box is four SPC/E waters in a 12.4 A box at 400 K (rc 6 A, Ewald), so
that a chain decorrelates in a few hundred moves.  Step sizes are held
fixed (AdjSteps beyond the run): GOMC's adaptation of the step sizes to the
acceptance rate (MoveSettings.cpp:164-258) feeds the state back into the
proposal, and on a box this small it shifts <E> of both samplers by several
hundred K, differently for each.
"""
from concurrent.futures import ThreadPoolExecutor
from dataclasses import replace

import numpy as np

from gomc_amd import systems
from oracle.pyoracle import Oracle, planted

N, STEP = 4, 0.3


def mean_energy(mp: bool, seed: int, samples: int, gap: int = 10):
    """Mean total energy per molecule and its standard error (20 blocks)."""
    s = systems.water_box(N, box=12.4, seed=5, rcut=6.0, temperature=400.0, adj_steps=10**12)
    if mp:
        s = replace(s, seed=seed, move_disp=0.0, move_rot=0.0, move_mp=1.0, _keep=[])
    else:
        s = replace(s, seed=seed, move_disp=0.5, move_rot=0.5, move_mp=0.0, _keep=[])
        gap *= N  # one sample per gap cycles of N moves
    o = Oracle(s, threads=1)
    o.set_scales([STEP], [STEP], STEP, STEP)
    step = 200 * gap
    o.run(step, record=False)
    e = []
    for _ in range(samples):
        o.run(gap, start_step=step, record=False)
        step += gap
        e.append(o.energies()[6] / N)
    b = np.array(e).reshape(20, -1).mean(1)
    o.close()
    return b.mean(), b.std(ddof=1) / np.sqrt(20)


def test_multiparticle_samples_the_same_distribution_as_single_moves():
    with ThreadPoolExecutor(2) as ex:  # the oracle releases the GIL
        f_mp = ex.submit(mean_energy, True, 11, 2000)
        f_tr = ex.submit(mean_energy, False, 12, 2000)
        (m_mp, e_mp), (m_tr, e_tr) = f_mp.result(), f_tr.result()
    # Chains of different seeds spread by about 200 K per molecule, more
    # than the block errors (60-150 K) say, because the cluster changes shape
    # rarely; the bar is 800 K, a third of what the planted mistake moves.
    assert abs(m_mp - m_tr) < 800.0, (m_mp, e_mp, m_tr, e_tr)
    assert -3500 < m_tr < -1000  # a bound, interacting cluster
    # the check sees new and old forces swapped in the weight ratio
    with planted(1):
        m_bad, e_bad = mean_energy(True, 11, 1000)
    assert m_bad - m_mp > 1500.0, (m_bad, e_bad, m_mp, e_mp)
