"""Batched persistent MC: each box follows its own Philox stream.

Parity: (1) a box run in the batch makes the same accept/reject decisions
as the CPU oracle running the same Philox stream, with per-move energy
changes within 1e-10 relative to the larger of their own magnitude and the
matching box total; (2) a box
gives bit-identical results alone and in a batch of other boxes; (3) the
running energies after a launch equal a from-scratch recomputation of the
final coordinates (1e-10 relative).
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Batch, Box
from oracle.pyoracle import Oracle


def dE_close(a, b, total, rtol=1e-10):
    """Per-move energy changes: |a-b| <= rtol * max(|b|, |box total|)."""
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), abs(total))))

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nmol,n", [(300, 6000), (1500, 4000)])
def test_batch_box_matches_oracle_philox(nmol, n):
    # 300 molecules: every other molecule is paired; 1500: the centre cull
    # lists only molecules within rc + 2 r_max of the moving one
    s = systems.water_box(nmol, seed=31, adj_steps=500)
    o = Oracle(s, threads=8)
    e0 = o.energies()
    bt = Batch([s])
    bt.set_trace(0, n)
    bt.run([n])
    tr = bt.trace()
    no, ro = o.run(n, mode=1)
    assert np.array_equal(tr["acc"], ro["acc"]), np.nonzero(tr["acc"] != ro["acc"])[0][:5]
    assert dE_close(tr["dlj"], ro["dlj"], e0[0])
    assert dE_close(tr["dreal"], ro["dreal"], e0[2])
    assert dE_close(tr["drecip"], ro["drecip"], e0[3])
    e, acc, done = bt.stats()
    assert done[0] == n and acc[0] == no
    assert np.allclose(e[0], o.energies(), rtol=1e-10, atol=0)
    assert np.allclose(bt.coords(0), o.coords(), rtol=0, atol=1e-9)


@pytest.mark.parametrize("nmol,team,n", [(2200, 2, 2500), (3000, 4, 2000), (3700, 8, 1500)])
def test_team_box_matches_oracle_philox(nmol, team, n):
    """A box run by a team of workgroups (split pair items and k-vector
    rounds, partial sums exchanged each move) makes the oracle's decisions."""
    s = systems.water_box(nmol, seed=61, adj_steps=500)
    bt = Batch([s], tuning="default")
    assert bt.layout(0)["team"] == team
    o = Oracle(s, threads=8)
    e0 = o.energies()
    bt.set_trace(0, n)
    bt.run([n])
    tr = bt.trace()
    no, ro = o.run(n, mode=1)
    assert np.array_equal(tr["acc"], ro["acc"]), np.nonzero(tr["acc"] != ro["acc"])[0][:5]
    assert dE_close(tr["dlj"], ro["dlj"], e0[0])
    assert dE_close(tr["dreal"], ro["dreal"], e0[2])
    assert dE_close(tr["drecip"], ro["drecip"], e0[3])
    e, acc, done = bt.stats()
    assert done[0] == n and acc[0] == no
    assert np.allclose(e[0], o.energies(), rtol=1e-10, atol=0)
    assert np.allclose(bt.coords(0), o.coords(), rtol=0, atol=1e-9)


def test_team_box_same_alone_and_in_mix():
    """Team sizes depend on the box alone, so a team box gives the same bits
    alone and beside other boxes (teams and single-workgroup boxes)."""
    big = systems.water_box(2600, seed=71)
    mix = [systems.water_box(300, seed=72), big, systems.water_box(3100, seed=73),
           systems.water_box(500, seed=74)]
    n = [600, 1200, 900, 800]
    full = Batch(mix, tuning="default")
    assert [full.layout(i)["team"] for i in range(4)] == [1, 2, 4, 1]
    full.set_trace(1, n[1])
    full.run(n)
    t_mix = full.trace()
    e_mix, _, _ = full.stats()
    alone = Batch([big], tuning="default")
    alone.set_trace(0, n[1])
    alone.run([n[1]])
    t_alone = alone.trace()
    e_alone, _, _ = alone.stats()
    for k in ("acc", "dlj", "dreal", "drecip"):
        assert np.array_equal(t_mix[k], t_alone[k]), k
    assert np.array_equal(e_mix[1], e_alone[0])
    assert np.array_equal(full.coords(1), alone.coords(0))
    # the running energies of the team boxes match a recomputation
    for i in (1, 2):
        s2 = systems.System(**{**mix[i].__dict__, "xyz": full.coords(i), "_keep": []})
        assert np.allclose(e_mix[i], Box(s2).system_total(), rtol=1e-10, atol=1e-6)


def test_batch_alone_equals_in_mix():
    mix = [systems.water_box(n, seed=100 + i, temperature=t)
           for i, (n, t) in enumerate([(250, 290.0), (600, 330.0), (400, 300.0), (900, 350.0)])]
    n = [1500, 2500, 2000, 1200]
    full = Batch(mix)
    full.set_trace(2, n[2])
    full.run(n)
    t_mix = full.trace()
    e_mix, _, _ = full.stats()
    alone = Batch([mix[2]])
    alone.set_trace(0, n[2])
    alone.run([n[2]])
    t_alone = alone.trace()
    e_alone, _, _ = alone.stats()
    for k in ("acc", "dlj", "dreal", "drecip"):
        assert np.array_equal(t_mix[k], t_alone[k]), k
    assert np.array_equal(e_mix[2], e_alone[0])
    assert np.array_equal(full.coords(2), alone.coords(0))


def test_batch_running_energy_matches_recompute():
    boxes = [systems.water_box(n, seed=200 + i) for i, n in enumerate([300, 700])]
    bt = Batch(boxes)
    bt.run([3000, 3000])
    e, _, _ = bt.stats()
    for i, s in enumerate(boxes):
        s2 = systems.System(**{**s.__dict__, "xyz": bt.coords(i), "_keep": []})
        ref = Box(s2).system_total()
        assert np.allclose(e[i], ref, rtol=1e-10, atol=1e-6), (e[i], ref)


def test_mixed_precision_error_against_fp64():
    """Mixed mode (FP32 pair terms and phase products, FP64 distances,
    accumulation, S(k), totals and Metropolis exponent): while the two runs
    have made the same decisions their configurations are identical, so the
    per-move energy changes can be compared directly.  Bar: |dE_mixed -
    dE_fp64| <= 1e-6 x max(|dE_fp64|, |E_total of the box|) (relative error
    at the box's energy scale; overlapping trials with |dE| >> |E_total| at
    their own scale).  A decision flips once the FP32 error straddles the
    acceptance draw, so the runs are compared up to their first differing
    decision, over four boxes, with at least 400 moves compared in all."""
    n = 3000
    compared = 0
    for seed, nmol in ((41, 400), (42, 400), (43, 600), (44, 300)):
        s = systems.water_box(nmol, seed=seed, adj_steps=500)
        runs = {}
        for prec in (0, 1):
            bt = Batch([s], precision=prec)
            bt.set_trace(0, n)
            bt.run([n])
            runs[prec] = bt.trace()
        etot = abs(Oracle(s, threads=4).energies()[6])
        a64, a32 = runs[0]["acc"], runs[1]["acc"]
        diff = np.nonzero(a64 != a32)[0]
        first = int(diff[0]) if len(diff) else n
        for k in ("dlj", "dreal", "drecip"):
            ref = runs[0][k][:first]
            err = np.abs(runs[1][k][:first] - ref)
            assert np.all(err <= 1e-6 * np.maximum(np.abs(ref), etot)), (seed, k, err.max(), etot)
        compared += first
    assert compared >= 400, compared


def test_graph_replayed_cycles_equal_eager_cycles():
    """One MC cycle captured as a HIP graph and replayed, with eager launches
    and host synchronisation between replays, ends bit for bit where the
    same cycles launched eagerly end."""
    boxes = [systems.water_box(n, seed=s, temperature=t)
             for n, s, t in ((210, 1, 290.0), (340, 2, 330.0), (500, 3, 300.0))]
    eager = Batch(boxes)
    for _ in range(5):
        eager.run_cycles(1, graph=False)
    e_ref, acc_ref, done_ref = eager.stats()
    g = Batch(boxes)
    g.run_cycles(2, graph=True)
    g.stats()                       # host work between replays
    g.run_cycles(1, graph=False)    # an eager cycle between replays
    g.run_cycles(2, graph=True)
    e, acc, done = g.stats()
    assert np.array_equal(done, done_ref)
    assert np.array_equal(acc, acc_ref)
    assert np.array_equal(e, e_ref)
    for i in range(len(boxes)):
        assert np.array_equal(g.coords(i), eager.coords(i))


def test_mixed_precision_observables_within_two_sigma():
    """FP32 pair terms / FP64 accumulation against FP64 as samplers: five
    boxes are equilibrated in FP64 (150 cycles), then each precision runs
    200 cycles from the same configuration on its own Philox stream, so the
    two runs are independent samples of one ensemble.  Per box the mean
    total energy of the last 160 cycles (8 blocks of 20) must agree within
    2 combined standard errors in at least 3 of the 5 boxes and within 4 in
    all of them (for unbiased samplers each bound fails with probability
    below 0.2%)."""
    base = [systems.water_box(n, seed=40 + i, temperature=t)
            for i, (n, t) in enumerate(((220, 300.0), (260, 320.0), (300, 290.0),
                                        (240, 340.0), (280, 310.0)))]
    eq = Batch(base)
    eq.run_cycles(150, graph=True)
    start = [systems.System(**{**s.__dict__, "xyz": eq.coords(i), "_keep": []})
             for i, s in enumerate(base)]
    eq.close()
    series = {}
    for prec in (0, 1):
        boxes = [systems.System(**{**s.__dict__, "seed": s.seed + 1000 * (prec + 1), "_keep": []})
                 for s in start]
        bt = Batch(boxes, precision=prec)
        es = []
        for c in range(200):
            bt.run_cycles(1, graph=False)
            e, _, _ = bt.stats()
            es.append(e[:, 6].copy())
        series[prec] = np.array(es[40:])
    z = []
    for b in range(len(base)):
        blocks = [series[p][:, b].reshape(8, -1).mean(1) for p in (0, 1)]
        se = np.sqrt(sum(x.var(ddof=1) / len(x) for x in blocks))
        z.append(abs(blocks[0].mean() - blocks[1].mean()) / se)
    z = np.array(z)
    assert np.sum(z < 2.0) >= 3 and np.all(z < 4.0), z


def test_autotuner_picks_a_timed_candidate():
    """The autotuner times the whole coarse grid of launch shapes on a small
    batch, then refines the team thresholds, and returns one of the shapes
    it timed; the per-box times it relies on are positive."""
    from gomc_amd import autotune
    boxes = [systems.water_box(n, seed=300 + i) for i, n in enumerate((250, 600, 1900, 2100))]
    rep = autotune.tune(boxes, cycles=1, log=lambda m: None)
    ms = [c["ms_per_cycle"] for c in rep["candidates"]]
    coarse = len(autotune.PAIR12) + len(autotune.TEAMS) - 1
    assert len(ms) >= coarse and all(m > 0 for m in ms)
    assert rep["tuning"] in [c["tuning"] for c in rep["candidates"]]
    assert rep["ms_per_cycle"] > 0
    bt = Batch(boxes, tuning=rep["tuning"])
    bt.run_cycles(1, graph=False)
    assert np.all(bt.box_times() > 0)
