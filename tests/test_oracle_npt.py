"""The oracle's NPT volume move (VolumeTransfer with one box), pinned
without the reference:

* an ideal gas (no LJ, no charges) samples V^N exp(-beta P V): the mean
  volume is (N + 1) / (beta P) (the gamma distribution's mean), checked
  against a block-averaged estimate;
* with interactions, the running energies after a run of translations,
  rotations and volume moves equal a from-scratch recomputation in the
  final box, and the volume stays at least 8 rc^3 (BoxDimensions minVol).
"""
import math
from dataclasses import replace

import numpy as np

from gomc_amd import systems
from gomc_amd.systems import MolKind, System
from oracle.pyoracle import Oracle


def ideal_gas(n=20, L=27.0, T=300.0, p_bar=40.0, seed=3):
    rng = np.random.default_rng(seed)
    kind = MolKind("IG", [0], [0.0], np.zeros((1, 3)))
    return System(box=(L, L, L), temperature=T, rcut=6.0, rcut_low=0.0, rcut_coulomb=6.0,
                  tolerance=1e-5, sigma=[1.0], epsilon=[0.0], n_rep=[12.0], kinds=[kind],
                  mol_kind=np.zeros(n, np.int32), xyz=rng.uniform(0, L, (n, 3)), seed=seed,
                  ewald=False, electrostatic=False, lrc=False, ensemble=2, pressure=p_bar,
                  move_disp=0.5, move_rot=0.0, move_vol=0.5, adj_steps=400)


def test_ideal_gas_mean_volume():
    s = ideal_gas()
    o = Oracle(s)
    o.run(4000, record=False)  # equilibrate the volume and its step size
    vols = []
    for i in range(200):
        o.run(200, start_step=4000 + 200 * i, record=False)
        vols.append(o.volume())
    vols = np.array(vols)
    blocks = vols.reshape(20, -1).mean(1)
    mean, err = blocks.mean(), blocks.std(ddof=1) / math.sqrt(len(blocks))
    expect = (len(s.mol_kind) + 1) * s.temperature / (s.pressure * 0.007242971)
    assert abs(mean - expect) < max(4 * err, 1e-3 * expect), (mean, err, expect)
    assert abs(mean - expect) < 0.08 * expect


def test_npt_water_bookkeeping():
    s = replace(systems.water_box(200, seed=11, rcut=8.0), ensemble=2, pressure=1.0,
                move_vol=0.05, adj_steps=150, _keep=[])
    o = Oracle(s, threads=4)
    v0 = o.volume()
    n, rec = o.run(2500)
    k = rec["kind"]
    assert (k == 9).sum() > 60 and rec["acc"][k == 9].sum() > 5
    assert o.volume() != v0
    assert o.volume() >= 8 * s.rcut ** 3
    assert np.allclose(o.energies(), o.system_total(), rtol=1e-10, atol=1e-6)
