"""Autotuner of the batched kernel's launch shape on the device it runs on.

Two choices change how the persistent kernel uses the GPU (see
gomc_batch_tuning in include/gomc_hip.h): the molecule counts from which a
box runs on a team of 2, 4 or 8 workgroups, and the largest k-vector count
at which a single-workgroup box gives its pair phase 12 waves (beside a
4-wave reciprocal phase) instead of 8 + 8.  The tuner times MC cycles of a
workload (default: the 256-box water batch) replayed from a HIP graph,
first over the wave split with the default teams, then over the team
thresholds with the best split, then refines the thresholds by coordinate
descent (each moved by +-100 and +-200 molecules, kept in order, two
rounds); a refined choice replaces the coarse one only when a longer
re-measurement of both finds it at least 1 % faster.  The fastest choice
goes to gomc_amd/tuning.json, which `Batch` then uses by default.  The packing of
single-workgroup boxes onto workgroups is tuned online instead: each launch
measures every box's time and the next launch packs by it.

    python -m gomc_amd.autotune [--boxes 256] [--cycles 3] [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import systems
from .api import TUNING_FILE, Batch

PAIR12 = (0, 1024, 2048, 4096)
TEAMS = ((1800, 2900, 3500), (1600, 2600, 3300), (2000, 2800, 3600), (1500, 2400, 3200),
         (2200, 3100, 3800))


def cycle_ms(boxes, tuning: dict, cycles: int) -> float:
    bt = Batch(boxes, tuning=tuning)
    try:
        bt.run_cycles(1, graph=True)  # capture + warm-up
        bt.run_cycles(1, graph=True)  # task lists now packed by measured box times
        bt.run_cycles(cycles, graph=True)
        return bt.last_kernel_ms() / cycles
    finally:
        bt.close()


def tune(boxes, cycles: int = 3, log=print) -> dict:
    best = {"team_min": list(TEAMS[0]), "pair12_max_k": 2048}
    results = []
    best_ms = None
    for k in PAIR12:
        t = {"team_min": list(TEAMS[0]), "pair12_max_k": k}
        ms = cycle_ms(boxes, t, cycles)
        results.append((t, ms))
        log(f"# pair12_max_k {k:5d} teams {TEAMS[0]}: {ms:.2f} ms/cycle")
        if best_ms is None or ms < best_ms:
            best, best_ms = t, ms
    for tm in TEAMS[1:]:
        t = {"team_min": list(tm), "pair12_max_k": best["pair12_max_k"]}
        ms = cycle_ms(boxes, t, cycles)
        results.append((t, ms))
        log(f"# pair12_max_k {t['pair12_max_k']:5d} teams {tm}: {ms:.2f} ms/cycle")
        if ms < best_ms:
            best, best_ms = t, ms
    coarse = dict(best)
    seen = {(tuple(t["team_min"]), t["pair12_max_k"]) for t, _ in results}
    for _ in range(2):
        moved = False
        for i in range(3):
            for step in (-200, -100, 100, 200):
                tm = list(best["team_min"])
                tm[i] += step
                if not (0 < tm[0] < tm[1] < tm[2]):
                    continue
                t = {"team_min": tm, "pair12_max_k": best["pair12_max_k"]}
                key = (tuple(tm), t["pair12_max_k"])
                if key in seen:
                    continue
                seen.add(key)
                ms = cycle_ms(boxes, t, cycles)
                results.append((t, ms))
                log(f"# refine teams {tuple(tm)}: {ms:.2f} ms/cycle")
                if ms < best_ms:
                    best, best_ms, moved = t, ms, True
        if not moved:
            break
    if best != coarse:  # confirm against the coarse choice on longer runs
        a = cycle_ms(boxes, coarse, 3 * cycles)
        b = cycle_ms(boxes, best, 3 * cycles)
        log(f"# confirm coarse {coarse['team_min']} {a:.2f} vs refined {best['team_min']} {b:.2f}")
        if b < 0.99 * a:
            best_ms = b
        else:
            best, best_ms = coarse, a
    return {"tuning": best, "ms_per_cycle": best_ms,
            "candidates": [{"tuning": t, "ms_per_cycle": m} for t, m in results]}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--boxes", type=int, default=256)
    ap.add_argument("--cycles", type=int, default=3)
    ap.add_argument("--out", default=str(TUNING_FILE))
    args = ap.parse_args()
    boxes = systems.config4(args.boxes)
    t0 = time.time()
    rep = tune(boxes, args.cycles, log=lambda m: print(m, file=sys.stderr, flush=True))
    rep["workload"] = f"config4({args.boxes}): water boxes, N log-uniform 200-4000"
    rep["seconds"] = time.time() - t0
    with open(args.out, "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps({"tuning": rep["tuning"], "ms_per_cycle": rep["ms_per_cycle"]}))


if __name__ == "__main__":
    main()
