"""The serving loop (continuous batching of resident jobs): a job's result
is the same alone, in a mix of jobs that arrive while it runs, and as one
batched launch; it matches the FP64 oracle under the same Random123 stream;
admission respects the memory budget.

Tolerances: alone / in the mix / one launch are bitwise identical (the
kernel is deterministic and a job's draws are indexed by its move number);
against the oracle the final energies agree to 1e-10 relative.
"""
import time

import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Batch
from gomc_amd.serve import Server
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _job(n, seed, t=300.0):
    return systems.water_box(n, temperature=t, seed=seed, place_seed=seed + 1)


def test_job_same_alone_in_mix_and_batched():
    a = _job(300, 11)
    moves = 3000
    sv = Server(slots=8, budget_bytes=64 << 20, tick_ms=2.0)
    ja = sv.submit(a, moves)
    sv.drain(120)
    alone = sv.poll(ja)
    sv.close()

    sv = Server(slots=8, budget_bytes=64 << 20, tick_ms=2.0)
    others = [sv.submit(_job(n, s, t), n * 10) for n, s, t in ((500, 3, 280.0), (220, 4, 340.0))]
    time.sleep(0.01)
    jb = sv.submit(a, moves)
    others.append(sv.submit(_job(800, 5, 310.0), 8000))
    sv.drain(300)
    mixed = sv.poll(jb)
    assert all(sv.poll(o)["status"] == 2 for o in others)
    sv.close()
    assert alone["status"] == mixed["status"] == 2
    assert np.array_equal(alone["energies"], mixed["energies"])
    assert alone["accepted"] == mixed["accepted"]

    bt = Batch([a])
    bt.run([moves])
    e, acc, _ = bt.stats()
    assert np.array_equal(e[0], mixed["energies"]) and acc[0] == mixed["accepted"]

    o = Oracle(a, threads=8)
    n_acc, _ = o.run(moves, mode=1, record=False)
    assert n_acc == mixed["accepted"]
    assert np.allclose(mixed["energies"], o.energies(), rtol=1e-10, atol=1e-6)


def test_admission_respects_memory_budget():
    jobs = [_job(400, 20 + i) for i in range(5)]
    # each job needs 1200 atoms x 24 B, 400 centres x 24 B and 2048
    # k-vector slots x 44 B; the k-vector share (0.665 of the budget) of
    # 280 kB holds two jobs at a time
    budget = 280_000
    sv = Server(slots=8, budget_bytes=budget, tick_ms=2.0)
    ids = [sv.submit(j, 800) for j in jobs]
    sv.drain(300)
    st = sv.stats()
    res = [sv.poll(i) for i in ids]
    sv.close()
    assert all(r["status"] == 2 for r in res)
    assert st["peak_bytes"] <= budget
    # later jobs waited for memory: admission times are not all equal
    adm = sorted(r["admitted"] for r in res)
    assert adm[-1] > adm[0]


def _gcmc_job(seed, nch4=12, nco2=12, L=24.0, fug=(80.0, 30.0)):
    # the reservoir must not run dry (the reference stops when it does)
    return systems.gcmc_mixture(nch4, nco2, L, fugacity_bar=fug, reservoir=250, seed=seed,
                                adj_steps=200)


def test_gcmc_job_same_alone_and_in_mix_and_matches_oracle():
    """GCMC lane: a methane/CO2 job gives the same energies, accepted count
    and mean molecule counts alone and while NVT and other GCMC jobs run
    beside it (each GCMC job has its own box, stream and MT19937 stream);
    its accepted count and final energies equal the FP64 oracle's under the
    same draws (1e-10 relative)."""
    g = _gcmc_job(5)
    moves = 1500
    sv = Server(slots=8, budget_bytes=256 << 20, tick_ms=2.0)
    ja = sv.submit(g, moves)
    sv.drain(300)
    alone = sv.poll(ja)
    sv.close()
    assert alone["status"] == 2 and alone["moves_done"] == moves

    sv = Server(slots=8, budget_bytes=256 << 20, tick_ms=2.0)
    others = [sv.submit(_job(400, 7, 300.0), 4000), sv.submit(_gcmc_job(9), 1200)]
    jb = sv.submit(g, moves)
    others.append(sv.submit(_gcmc_job(11, fug=(300.0, 10.0)), 900))
    sv.drain(300)
    mixed = sv.poll(jb)
    assert all(sv.poll(o)["status"] == 2 for o in others), (
        [sv.poll(o)["status"] for o in others], sv.last_error())
    sv.close()
    assert np.array_equal(alone["energies"], mixed["energies"])
    assert alone["accepted"] == mixed["accepted"]
    assert np.array_equal(alone["mean_n"], mixed["mean_n"])

    # the oracle under the same draws, sampled the same way (after every
    # cycle of N_box moves): identical sequences give identical counts
    o = Oracle(g, threads=8)
    done, n_acc, sums, samples = 0, 0, np.zeros(2), 0
    while done < moves:
        n_in, _ = o.mol_boxes()
        chunk = min(max(1, n_in), moves - done)
        acc, _ = o.run(chunk, start_step=done, mode=0, record=False)
        n_acc += acc
        done += chunk
        _, where = o.mol_boxes()
        for k in (0, 1):
            sums[k] += np.count_nonzero((where == 0) & (g.mol_kind == k))
        samples += 1
    assert n_acc == mixed["accepted"]
    assert np.allclose(mixed["energies"], o.energies(), rtol=1e-10, atol=1e-6)
    assert np.array_equal(mixed["mean_n"], sums / samples)


REPLICAS, BURN, SAMPLE = 8, 100, 300


def _oracle_replica(s, seed, burn, sample):
    """<N_k> of one FP64 oracle replica over `sample` cycles after `burn`
    (a cycle is N_box moves, counts taken after each, as the device loop
    samples)."""
    from dataclasses import replace
    o = Oracle(replace(s, seed=seed, _keep=[]), threads=1)
    mk = s.mol_kind

    def counts():
        _, b = o.mol_boxes()
        return np.array([np.sum((b == 0) & (mk == 0)), np.sum((b == 0) & (mk == 1))], float)

    step, acc = 0, np.zeros(2)
    for cyc in range(burn + sample):
        n = max(1, int(counts().sum()))
        o.run(n, start_step=step, record=False)
        step += n
        if cyc >= burn:
            acc += counts()
    return acc / sample


@pytest.mark.parametrize("L,T,fug", [(20.0, 300.0, (5.0, 2.0)), (22.0, 280.0, (60.0, 30.0)),
                                     (24.0, 320.0, (20.0, 80.0))])
def test_mixed_precision_gcmc_occupation_matches_independent_fp64_oracle(L, T, fug):
    """The served GCMC lane's mixed precision (FP32 pair terms, CBMC trial
    energies and phases; FP64 sums and Metropolis) samples the FP64
    Hamiltonian's grand-canonical <N_k>, at a dilute, a dense
    (high-fugacity) and a CO2-rich state point.

    Eight mixed-precision replicas run in one launch of the device move loop
    and eight FP64 oracle replicas run on independent seeds, all from the
    same start; each replica's <N_k> is averaged over 300 cycles after 100.
    In the dense adsorbed phase N_CO2 wanders over tens of molecules for
    hundreds of cycles, so the error of each side is taken from the scatter
    of its replica means (not from blocks of one run).  Bar: 4 standard
    errors of the difference per kind (eight replicas a side, so a t
    distribution with about 14 degrees of freedom)."""
    from concurrent.futures import ThreadPoolExecutor
    from dataclasses import replace

    from gomc_amd.api import Box, run_device
    s = systems.gcmc_mixture(6, 6, L, temperature=T, fugacity_bar=fug, reservoir=400, seed=23,
                             adj_steps=500)
    boxes = []
    for r in range(REPLICAS):
        b = Box(replace(s, seed=100 + r, _keep=[]))
        b.set_precision(1)
        boxes.append(b)
    cap = 10 ** 7
    moves, _, _, _ = run_device(boxes, cap, cycles=BURN)
    _, _, sums, samples = run_device(boxes, cap, start_steps=moves, cycles=SAMPLE)
    assert np.all(samples == SAMPLE)
    dev = sums / samples[:, None]
    for b in boxes:
        b.close()
    with ThreadPoolExecutor(REPLICAS) as ex:
        orc = np.array(list(ex.map(lambda r: _oracle_replica(s, 9001 + r, BURN, SAMPLE),
                                   range(REPLICAS))))
    mg, mo = dev.mean(0), orc.mean(0)
    eg = dev.std(0, ddof=1) / np.sqrt(REPLICAS)
    eo = orc.std(0, ddof=1) / np.sqrt(REPLICAS)
    z = (mg - mo) / np.maximum(np.hypot(eg, eo), 1e-12)
    print(f"L {L} T {T} fug {fug}: device mixed {mg} +- {eg}, oracle FP64 {mo} +- {eo}, z {z}")
    assert np.all(np.abs(z) < 4.0), (mg, eg, mo, eo, z)
