"""Checkpoint and restore of a box (gomc_box_checkpoint / gomc_box_restore),
tested as the reference's CheckpointTest tests its own: a run interrupted
by a checkpoint and continued in a new box equals the uninterrupted run.

Bar: after restoring into a freshly created box of the same system, the
next moves have the same kind, molecule, overlap and accept sequences and
the same per-move energy changes (bitwise), and end with the same
coordinates, energies and move settings as the box the checkpoint was
taken from -- for NVT water with Ewald (host driver), GCMC methane/CO2 on
the device move loop, NPT water whose volume changed before the checkpoint
(the restored box regrids and rebuilds its k-vectors), water with
MultiParticle moves and the LJ-only AR_KR example read from its GOMC
files; a GCMC box in EwaldCached mode rebuilds its cache on restore, so
there the energy changes agree within 1e-10 rather than bitwise.  A
checkpoint of another system, or a truncated one, is refused and leaves
the box as it was.
"""
from dataclasses import replace

import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box, GomcError

pytestmark = pytest.mark.gpu


def same(ra, rb, exact=True):
    for key in ("kind", "mol", "ovl", "acc", "dlj", "dreal", "drecip"):
        if exact or key in ("kind", "mol", "ovl", "acc"):
            assert np.array_equal(ra[key], rb[key]), (key, np.nonzero(ra[key] != rb[key])[0][:5])
        else:
            scale = max(np.abs(ra[key]).max(), 1.0)
            assert np.all(np.abs(ra[key] - rb[key]) <= 1e-10 * scale), key


def continued_equals_uninterrupted(s, n1, n2, device=False, cached=False):
    a = Box(s)
    if cached:
        a.recip_cache(True)
    a.run(n1, device=device)
    ck = a.checkpoint(step=n1)
    na, ra = a.run(n2, start_step=n1, device=device)
    b = Box(s)
    if cached:
        b.recip_cache(True)
    assert b.restore(ck) == n1
    nb, rb = b.run(n2, start_step=n1, device=device)
    same(ra, rb, exact=not cached)
    assert na == nb
    if cached:
        assert np.allclose(a.coords(), b.coords(), rtol=0, atol=1e-9)
        assert np.allclose(a.energies(), b.energies(), rtol=1e-10, atol=1e-6)
    else:
        assert np.array_equal(a.coords(), b.coords())
        assert np.array_equal(a.energies(), b.energies())
    assert np.array_equal(a.scales()[0], b.scales()[0])
    assert np.array_equal(a.dims(), b.dims())
    return a, b, ra


def test_nvt_water_host_driver():
    a, b, ra = continued_equals_uninterrupted(systems.water_box(300, seed=11), 1500, 1500)
    assert ra["acc"].sum() > 300


def test_gcmc_device_loop():
    s = systems.gcmc_mixture(60, 60, 25.0, fugacity_bar=(200.0, 60.0), reservoir=120, seed=3,
                             adj_steps=250)
    a, b, ra = continued_equals_uninterrupted(s, 1500, 1500, device=True)
    assert ra["acc"][ra["kind"] == 10].sum() > 10
    assert np.array_equal(a.mol_boxes()[1], b.mol_boxes()[1])


def test_npt_volume_changed_before_the_checkpoint():
    s = replace(systems.water_box(300, seed=11, rcut=8.0), ensemble=2, pressure=1.0,
                move_vol=0.05, adj_steps=150, _keep=[])
    v0 = Box(s).dims()[3]
    a, b, ra = continued_equals_uninterrupted(s, 1200, 1000)
    assert a.dims()[3] != v0 and (ra["kind"] == 9).sum() > 20


def test_multiparticle_moves():
    s = systems.water_box(200, seed=21, adj_steps=100, move_disp=0.25, move_rot=0.25,
                          move_mp=0.5)
    a, b, ra = continued_equals_uninterrupted(s, 300, 300)
    assert (ra["kind"] == 2).sum() > 50


def test_cached_ewald_gcmc():
    """EwaldCached mode: the restored box rebuilds its per-molecule cache
    from the restored coordinates, so its cached terms may differ from the
    interrupted box's by rounding: sequences identical, energy changes and
    running energies within 1e-10."""
    s = systems.gcmc_mixture(40, 40, 24.0, fugacity_bar=(150.0, 50.0), reservoir=80, seed=8,
                             adj_steps=200)
    a, b, ra = continued_equals_uninterrupted(s, 1200, 1200, cached=True)
    assert ra["acc"][ra["kind"] == 10].sum() > 10


def test_ewald_off_from_gomc_files(tmp_path):
    """No k-vectors (the AR_KR example: LJ only), GCMC on the device loop."""
    from gomc_amd import gomc_io
    from test_gpu_gomc_io import ar_kr_gcmc
    s = gomc_io.system_from_gomc(ar_kr_gcmc(tmp_path))
    a, b, ra = continued_equals_uninterrupted(s, 1000, 1500, device=True)
    assert ra["acc"][ra["kind"] == 10].sum() > 10


def test_foreign_and_truncated_checkpoints_are_refused():
    s = systems.water_box(200, seed=5)
    other = Box(systems.water_box(201, seed=5)).checkpoint()
    b = Box(s)
    e0, x0 = b.energies(), b.coords()
    ck = b.checkpoint(step=7)
    with pytest.raises(GomcError, match="different system"):
        b.restore(other)
    with pytest.raises(GomcError):
        b.restore(ck[:-9])
    with pytest.raises(GomcError, match="not a box checkpoint"):
        b.restore(b"x" * 64)
    assert np.array_equal(b.energies(), e0) and np.array_equal(b.coords(), x0)
    assert b.restore(ck) == 7
    b.run(200)


def test_other_hamiltonian_and_corrupt_lookup_are_refused_before_the_box_changes():
    """A checkpoint of the same molecules under another cutoff (another
    Hamiltonian: its running energies do not belong to this box) and one
    whose molecule lookup lists a molecule twice are both refused, and the
    box continues exactly as a box that never saw them."""
    s = systems.water_box(200, seed=5)
    b, ref = Box(s), Box(s)
    b.run(300)
    ref.run(300)
    e0, x0 = b.energies(), b.coords()
    other = Box(replace(s, rcut=8.0, _keep=[]))
    other.run(300)
    with pytest.raises(GomcError, match="another force field, cutoff"):
        b.restore(other.checkpoint(step=300))
    ck = bytearray(b.checkpoint(step=300))
    ident = np.arange(s.nmols, dtype=np.uint32).tobytes()
    at = bytes(ck).find(ident)
    assert at > 0 and bytes(ck).find(ident, at + 1) < 0
    bad = bytearray(ck)
    bad[at + 4:at + 8] = np.uint32(0).tobytes()   # molecule 0 listed twice, 1 never
    with pytest.raises(GomcError, match="molecule lookup"):
        b.restore(bytes(bad))
    assert np.array_equal(b.energies(), e0) and np.array_equal(b.coords(), x0)
    _, rb = b.run(500, start_step=300)
    _, rr = ref.run(500, start_step=300)
    same(rb, rr)
    assert np.array_equal(b.coords(), ref.coords())
