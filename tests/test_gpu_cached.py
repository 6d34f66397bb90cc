"""EwaldCached mode (per-molecule structure factors kept in HBM,
src/EwaldCached.cpp) against the uncached path.  In the reference the two
modes give the same energies: a molecule's cached terms are exactly the
terms its current positions give.  Here the cache is built with the same
arithmetic the trial kernel uses for old positions, so the move sequences
must be identical and the energies equal to 1e-13 relative, through
translations, rotations, GCMC insertions/deletions and MultiParticle moves
(which refill the cache).
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box

pytestmark = pytest.mark.gpu


def _compare(s, nsteps):
    a, b = Box(s), Box(s)
    b.recip_cache(True)
    na, ra = a.run(nsteps)
    nb, rb = b.run(nsteps)
    for key in ("kind", "mol", "acc", "ovl"):
        assert np.array_equal(ra[key], rb[key]), key
    assert na == nb
    for key in ("dlj", "dreal", "drecip"):
        assert np.allclose(ra[key], rb[key], rtol=1e-12, atol=1e-9), key
    assert np.allclose(a.energies(), b.energies(), rtol=1e-13, atol=1e-9)
    assert np.array_equal(a.coords(), b.coords())
    # the cache still matches the final positions: totals agree
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)


def test_cached_translate_rotate():
    _compare(systems.water_box(300, seed=17), 4000)


def test_cached_gcmc():
    _compare(systems.gcmc_mixture(60, 60, 25.0, fugacity_bar=(200.0, 60.0), reservoir=120,
                                  seed=3), 2000)


def test_cached_multiparticle():
    _compare(systems.water_box(200, seed=21, adj_steps=100, move_disp=0.25, move_rot=0.25,
                               move_mp=0.5), 300)
