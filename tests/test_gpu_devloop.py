"""The device move loop (gomc_run_moves_device / _multi): whole moves --
Translate, Rotate, MoleculeTransfer with CBMC growth -- run inside one
launch, the box's MT19937 stream continued on the device.

Parity bar (FP64): against the oracle under the same MT19937 stream, the
move-kind, molecule, overlap and accept sequences are identical, per-move
energy changes agree within 1e-10 of the box totals, final membership is
identical, coordinates within 1e-9 A, running energies within 1e-10
relative.  Against the host driver (gomc_run_moves) on the same box the
records agree to the same bar, and a run may switch between the two
drivers at any move.  Several boxes in one launch end exactly as each run
alone.  Mixed precision: accept sequence equal to the FP64 device loop's
for the compared prefix, per-move error within 1e-6 of max(|dE|, |E|).
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box, run_device
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def dE_close(a, b, total, rtol=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), abs(total))))


def same_records(rg, ro, e0):
    for key in ("kind", "mol", "ovl", "acc"):
        assert np.array_equal(rg[key], ro[key]), (key, np.nonzero(rg[key] != ro[key])[0][:5])
    assert dE_close(rg["dlj"], ro["dlj"], e0[0])
    assert dE_close(rg["dreal"], ro["dreal"], e0[2])
    ok = ro["ovl"] == 0
    assert dE_close(rg["drecip"][ok], ro["drecip"][ok], e0[3])


def gcmc_small(seed=3, adj=250):
    return systems.gcmc_mixture(60, 60, 25.0, fugacity_bar=(200.0, 60.0), reservoir=120,
                                seed=seed, adj_steps=adj)


def test_nvt_water_same_sequence_as_oracle():
    s = systems.water_box(300, seed=11)
    b, o = Box(s), Oracle(s, threads=8)
    e0 = o.energies()
    n = 3000
    ng, rg = b.run(n, device=True)
    no, ro = o.run(n, mode=0)
    same_records(rg, ro, e0)
    assert ng == no
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=0)
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    assert np.allclose(b.scales()[0], o.scales()[0], rtol=1e-12)


def test_gcmc_same_sequence_as_oracle():
    s = gcmc_small()
    b, o = Box(s), Oracle(s, threads=8)
    e0 = o.energies()
    n = 3000
    ng, rg = b.run(n, device=True)
    no, ro = o.run(n, mode=0)
    k = ro["kind"]
    assert (k == 10).sum() > n // 5 and ro["acc"][k == 10].sum() > 20
    same_records(rg, ro, e0)
    assert ng == no
    n1, box1 = b.mol_boxes()
    n2, box2 = o.mol_boxes()
    assert n1 == n2 and np.array_equal(box1, box2)
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)


def test_gcmc_40A_box_same_sequence_as_oracle():
    s = systems.config3()
    b, o = Box(s), Oracle(s, threads=8)
    e0 = o.energies()
    n = 600
    ng, rg = b.run(n, device=True)
    no, ro = o.run(n, mode=0)
    same_records(rg, ro, e0)
    assert ng == no
    n1, box1 = b.mol_boxes()
    n2, box2 = o.mol_boxes()
    assert n1 == n2 and np.array_equal(box1, box2)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)


def test_device_and_host_drivers_interleave():
    """Host moves, device moves, host moves again: the same records as the
    host driver alone, so every piece of state crosses the boundary."""
    s = gcmc_small(seed=7, adj=100)
    a, h = Box(s), Box(s)
    e0 = h.energies()
    parts = [(0, 400, False), (400, 700, True), (1100, 300, False), (1400, 600, True)]
    recs = []
    for start, n, dev in parts:
        _, r = a.run(n, start_step=start, device=dev)
        recs.append(r)
    ra = {k: np.concatenate([r[k] for r in recs]) for k in recs[0]}
    _, rh = h.run(2000)
    same_records(ra, rh, e0)
    assert np.array_equal(a.mol_boxes()[1], h.mol_boxes()[1])
    assert np.allclose(a.energies(), h.energies(), rtol=1e-10, atol=1e-6)
    assert np.allclose(a.scales()[0], h.scales()[0], rtol=1e-12)


def test_multi_box_launch_equals_each_alone():
    specs = [gcmc_small(seed=s) for s in (21, 22)] + [
        systems.gcmc_mixture(100, 40, 28.0, fugacity_bar=(300.0, 30.0), reservoir=150, seed=23)]
    together = [Box(s) for s in specs]
    moves, acc, sums, samples = run_device(together, 5000, cycles=4)
    for i, s in enumerate(specs):
        alone = Box(s)
        m1, a1, s1, n1 = run_device([alone], 5000, cycles=4)
        assert moves[i] == m1[0] and acc[i] == a1[0] and samples[i] == n1[0] == 4
        assert np.array_equal(sums[i], s1[0])
        assert np.array_equal(together[i].mol_boxes()[1], alone.mol_boxes()[1])
        assert np.array_equal(together[i].coords(), alone.coords())
        assert np.array_equal(together[i].energies(), alone.energies())


def test_cycle_sampling_matches_host_chunks():
    """Cycles of N_box moves with the counts sampled after each, as the
    serving lane's host loop does it."""
    s = gcmc_small(seed=31)
    mk = np.asarray(s.mol_kind)
    h = Box(s)
    done, counts = 0, []
    for _ in range(5):
        n_in, _ = h.mol_boxes()
        chunk = max(1, n_in)
        h.run(chunk, start_step=done, record=False)
        done += chunk
        _, where = h.mol_boxes()
        counts.append([np.count_nonzero((where == 0) & (mk == k)) for k in (0, 1)])
    d = Box(s)
    moves, _, sums, samples = run_device([d], 10**6, cycles=5)
    assert moves[0] == done and samples[0] == 5
    assert np.array_equal(sums[0], np.sum(counts, axis=0))
    assert np.allclose(d.energies(), h.energies(), rtol=1e-10, atol=1e-6)


def test_mixed_precision_device_loop_error_against_fp64():
    s = gcmc_small(seed=41)
    f, m = Box(s), Box(s)
    m.set_precision(1)
    n = 1500
    _, rf = f.run(n, device=True)
    _, rm = m.run(n, device=True)
    flip = np.nonzero(rf["acc"] != rm["acc"])[0]
    upto = flip[0] if len(flip) else n
    assert upto >= 400, upto
    e = abs(f.energies()[6])
    for key in ("dlj", "dreal", "drecip"):
        a, b = rm[key][:upto], rf[key][:upto]
        ok = rf["ovl"][:upto] == 0
        assert np.all(np.abs(a[ok] - b[ok]) <= 1e-6 * np.maximum(np.abs(b[ok]), e)), key


@pytest.mark.parametrize("team", [2, 4, 8])
def test_team_of_workgroups_same_sequence_as_oracle(team):
    """A box on a team of workgroups (the k-vectors split among the members,
    every member running the whole move): the oracle's sequences, for
    translations, rotations and transfers with CBMC growth, and the running
    energies equal to a recomputation of the final state."""
    s = gcmc_small(seed=5)
    b, o = Box(s), Oracle(s, threads=8)
    b.set_device_team(team)
    e0 = o.energies()
    n = 2500
    ng, rg = b.run(n, device=True)
    no, ro = o.run(n, mode=0)
    assert ro["acc"][ro["kind"] == 10].sum() > 10
    same_records(rg, ro, e0)
    assert ng == no
    assert np.array_equal(b.mol_boxes()[1], o.mol_boxes()[1])
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)
    # the box continues on one workgroup: the state the team left is whole
    b.set_device_team(1)
    _, rg2 = b.run(500, start_step=n, device=True)
    _, ro2 = o.run(500, start_step=n, mode=0)
    same_records(rg2, ro2, e0)


def test_teamed_box_alone_equals_in_a_launch_with_others():
    """A 60 A box takes a team of 4 by its k-vector count; run beside boxes
    on one workgroup it ends bit for bit as when run alone, with cycle
    sampling."""
    big = systems.gcmc_mixture(200, 200, 60.0, fugacity_bar=(5.0, 5.0), reservoir=3300, seed=41)
    specs = [gcmc_small(seed=31), big, gcmc_small(seed=32)]
    together = [Box(s) for s in specs]
    assert together[1].nkvec >= 16384
    moves, acc, sums, samples = run_device(together, 20000, cycles=3)
    for i, s in enumerate(specs):
        alone = Box(s)
        m1, a1, s1, n1 = run_device([alone], 20000, cycles=3)
        assert moves[i] == m1[0] and acc[i] == a1[0] and samples[i] == n1[0] == 3
        assert np.array_equal(sums[i], s1[0])
        assert np.array_equal(together[i].coords(), alone.coords())
        assert np.array_equal(together[i].energies(), alone.energies())
    b = together[1]
    assert np.allclose(b.system_total(), b.energies(), rtol=1e-10, atol=1e-6)
