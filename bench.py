"""Benchmark: MC trial moves/s of the batched persistent MC kernel on one
MI355X, on 256 independent NVT water boxes (N log-uniform in [200, 4000]
molecules at 0.997 g/cm^3, T uniform in [280, 360] K, one Philox stream per
box).  One step = one MC cycle of the whole batch (box i makes N_i moves),
replayed from one captured HIP graph.  `value` is the FP64 batched rate.

Reported beside it:
  * the same cycles with FP32 pair terms / FP64 accumulation;
  * the plain FP64 GPU path (one fused launch per move, host Metropolis,
    MT19937 draws -- the reference's call pattern with state kept in HBM),
    on a bounded sample of moves per box;
  * the single-box workloads (1000 waters; 10,000 waters with MultiParticle
    moves 1:1; GCMC methane/CO2 in a 40 A box): GPU moves/s on the host
    driver and the device move loop, and the CPU oracle on all cores and on
    one core, on bounded samples;
  * serving: a Poisson stream of water state-point jobs through the
    continuous-batching server, and the GCMC adsorption stream (60 jobs of
    2000 MC cycles, 200 GB budget) through its GCMC lane (moves/s, p50/p99);
  * the CPU oracle (C++/OpenMP, one box per thread) on a bounded sample of
    the 256-box workload, on all cores and on one.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("MC trial moves/s (1 GPU), batched+served vs plain at ≤1e-6 rel ΔE; "
          "p99 latency")
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) peak
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
# HBM bytes per launch of McBatchKernel<0> on the default 256-box workload,
# from rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a known-size copy
# (profiles/McBatchKernel_water256_fp64_hbm.md); reported for that workload.
PROFILED_TRAFFIC = {"bytes": 167.5e9, "source": "profiles/McBatchKernel_water256_fp64_hbm.md"}
# Fixed per-term FLOP weights of the algorithmic work model: distance +
# minimum image 12, erfc-Coulomb 30, LJ 15, sincos + 2 FMA 24, per-k energy 6.
W_DIST, W_COUL, W_LJ, W_SINCOS, W_KENERGY = 12, 30, 15, 24, 6


def move_flops(s, nkvec: int) -> float:
    """Algorithmic FLOPs of one single-molecule trial in box `s`: the
    distance checks the reference makes over the 27 neighbour cells of its
    cell list (cells per edge max(floor(L / rc), 3), CellList.cpp:138-163),
    in-cutoff Coulomb and LJ terms, 2 a K sincos and K energy terms."""
    a = 3
    n_atoms = s.natoms
    L = s.box[0]
    frac = min(1.0, (4.0 / 3.0) * math.pi * s.rcut ** 3 / L ** 3)
    edge = max(int(L / max(s.rcut, s.rcut_coulomb)), 3)
    D = 2 * a * (n_atoms - a) * min(1.0, 27.0 / edge ** 3)   # distance checks
    C = 2 * a * ((n_atoms - a) * frac)           # charged pairs in cutoff
    Lj = 2 * 1 * ((s.nmols - 1) * frac)          # O-O pairs in cutoff
    S = 2 * a * nkvec                            # sincos (old + new)
    return D * W_DIST + C * W_COUL + Lj * W_LJ + S * W_SINCOS + nkvec * W_KENERGY


def move_bytes(s, nkvec: int) -> float:
    """Algorithmic (compulsory) HBM bytes of one single-molecule trial:
    64 K for the k-vector data (kx, ky, kz, prefactor and the two
    structure-factor sums read, the two new sums written, 8 B each) plus
    40 M for the M atoms of the 27-cell neighbourhood (x, y, z, kind,
    charge, molecule id)."""
    L = s.box[0]
    edge = max(int(L / max(s.rcut, s.rcut_coulomb)), 3)
    M = s.natoms * min(1.0, 27.0 / edge ** 3)
    return 64.0 * nkvec + 40.0 * M


def cpu_threads() -> tuple[int, int]:
    """(threads used, cores visible): every core this process may run on,
    limited by OMP_NUM_THREADS where the host sets it (a GPU box's share
    of its host)."""
    seen = len(os.sched_getaffinity(0))
    lim = os.environ.get("OMP_NUM_THREADS", "")
    return (min(seen, int(lim)) if lim.isdigit() and int(lim) > 0 else seen), seen


def copy_reference(nbytes: int) -> None:
    """Three device-to-device copies of `nbytes` through the HIP runtime (its
    blit kernel): a known-size reference for the HBM byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def next_rows(cpu_threads: int, with_cpu: bool) -> dict:
    """Timings of the rows around the batched path, each beside the oracle on
    the host: the molecular virial and pressure (config-1 and config-2
    water), MultiParticle moves (config-1 and config-2 water; the oracle on
    config 1 only, a config-2 move takes it tens of seconds), NPT water with
    volume moves, a replica-exchange attempt over a 6-temperature batch, and
    a checkpoint/restore round trip."""
    from dataclasses import replace

    from gomc_amd import systems
    from gomc_amd.api import Batch, Box
    from oracle.pyoracle import Oracle

    def best_ms(fn, reps=5):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t) * 1e3)
        return float(np.median(ts))

    out = {}
    for name, s in (("config1_1000_water", systems.config1()),
                    ("config2_10000_water", systems.config2())):
        b = Box(s)
        row = dict(gpu_ms=best_ms(b.virial), pressure_bar=b.virial()["pressure"])
        if with_cpu:
            o = Oracle(s, threads=cpu_threads)
            t = time.perf_counter()
            vo = o.virial()
            row["cpu_ms"] = (time.perf_counter() - t) * 1e3
            row["cpu_threads"] = cpu_threads
            row["rel_diff_pressure"] = abs(row["pressure_bar"] - vo["pressure"]) / abs(vo["pressure"])
            o.close()
        b.close()
        out[f"virial_{name}"] = row
        print(f"# virial {name}: {row}", file=sys.stderr, flush=True)

    # MultiParticle moves only: structure factors, reciprocal and pair forces,
    # torques and the force-biased transform of every molecule per move
    for name, s, n_gpu, n_cpu in (("config1_1000_water", systems.config1(), 20, 4),
                                  ("config2_10000_water", systems.config2(), 10, 0)):
        s = replace(s, move_disp=0.0, move_rot=0.0, move_mp=1.0, _keep=[])
        b = Box(s)
        b.run(2, record=False)
        t = time.perf_counter()
        b.run(n_gpu, start_step=2, record=False)
        row = dict(gpu_ms_per_move=(time.perf_counter() - t) / n_gpu * 1e3)
        b.close()
        if with_cpu and n_cpu:
            o = Oracle(s, threads=cpu_threads)
            t = time.perf_counter()
            o.run(n_cpu, mode=0, record=False)
            row["cpu_ms_per_move"] = (time.perf_counter() - t) / n_cpu * 1e3
            row["cpu_threads"] = cpu_threads
            o.close()
        out[f"multiparticle_{name}"] = row
        print(f"# multiparticle {name}: {row}", file=sys.stderr, flush=True)

    s = replace(systems.water_box(300, seed=11, rcut=8.0), ensemble=2, pressure=1.0,
                move_vol=0.05, adj_steps=150, _keep=[])
    b = Box(s)
    b.run(200, record=False)
    t = time.perf_counter()
    b.run(2000, start_step=200, record=False)
    row = dict(workload="300 waters, rc 8 A, 1 bar, VolFreq 0.05 of translate/rotate",
               gpu_moves_per_s=2000 / (time.perf_counter() - t))
    b.close()
    if with_cpu:
        o = Oracle(s, threads=cpu_threads)
        o.run(200, record=False)
        t = time.perf_counter()
        o.run(2000, start_step=200, mode=0, record=False)
        row["cpu_moves_per_s"] = 2000 / (time.perf_counter() - t)
        row["cpu_threads"] = cpu_threads
        o.close()
    out["npt_water"] = row
    print(f"# npt {row}", file=sys.stderr, flush=True)

    temps = [280.0, 300.0, 320.0, 345.0, 370.0, 400.0]
    bt = Batch([systems.water_box(260, temperature=tt, seed=100 + i, place_seed=7)
                for i, tt in enumerate(temps)])
    ladder = np.arange(len(temps), dtype=np.int32)
    bt.run([260] * len(temps))
    n_att = 200
    t = time.perf_counter()
    for a in range(n_att):
        bt.exchange(ladder, a, seed=3)
    out["replica_exchange"] = dict(replicas=len(temps),
                                   us_per_attempt=(time.perf_counter() - t) / n_att * 1e6)
    bt.close()

    b = Box(systems.config1())
    b.run(500, record=False)
    ck = b.checkpoint(500)
    b2 = Box(systems.config1())
    out["checkpoint_config1"] = dict(bytes=len(ck), save_ms=best_ms(lambda: b.checkpoint(500)),
                                     restore_ms=best_ms(lambda: b2.restore(ck)))
    b.close()
    b2.close()
    print(f"# replica/checkpoint {out['replica_exchange']} {out['checkpoint_config1']}",
          file=sys.stderr, flush=True)
    return out


def config_rows(threads: int, with_cpu: bool) -> dict:
    """Moves/s on the single-box workloads: 1000 waters (translate/rotate
    1:1), 10,000 waters (MultiParticle 1:1 with translate/rotate) and GCMC
    methane/CO2 in a 40 A box (40 % transfers with CBMC), on the GPU (host
    driver: one fused launch per energy call; device move loop: whole moves
    in one launch) and on the CPU oracle with `threads` threads and with one,
    each on a bounded sample of the same move stream."""
    from dataclasses import replace

    from gomc_amd import systems
    from gomc_amd.api import Box
    from oracle.pyoracle import Oracle

    def gpu_rate(s, n, device=False, warm=20):
        b = Box(s)
        b.run(warm, record=False, device=device)
        t = time.perf_counter()
        _, rec = b.run(n, start_step=warm, device=device)
        dt = time.perf_counter() - t
        b.close()
        kinds = {int(k): int((rec["kind"] == k).sum()) for k in np.unique(rec["kind"])}
        return n / dt, kinds

    def cpu_rate(s, n, th, warm=5):
        o = Oracle(s, threads=th)
        o.run(warm, record=False)
        t = time.perf_counter()
        o.run(n, start_step=warm, record=False)
        dt = time.perf_counter() - t
        o.close()
        return n / dt

    out = {}
    s1 = systems.config1()
    r, k = gpu_rate(s1, 10000)
    rd, _ = gpu_rate(s1, 10000, device=True)
    row = dict(workload="1000 SPC/E waters, L 31.07 A, rc 10 A + LRC, Ewald 1e-5, "
                        "translate/rotate 1:1", gpu_host_driver_moves_per_s=r,
               gpu_device_loop_moves_per_s=rd, moves_by_kind=k, gpu_sample="10000 moves")
    if with_cpu:
        row.update(cpu_moves_per_s=cpu_rate(s1, 2000, threads), cpu_threads=threads,
                   cpu_1thread_moves_per_s=cpu_rate(s1, 600, 1), cpu_sample="2000 / 600 moves")
    out["config1"] = row
    print(f"# config1 {row}", file=sys.stderr, flush=True)

    s2 = systems.config2()
    r, k = gpu_rate(s2, 200, warm=4)
    row = dict(workload="10,000 SPC/E waters, L 66.9 A, rc 12 A, Ewald 1e-6 (30,902 "
                        "k-vectors), MultiParticle 1:1 with translate/rotate",
               gpu_host_driver_moves_per_s=r, moves_by_kind=k, gpu_sample="200 moves")
    if with_cpu:
        # a MultiParticle move takes the oracle seconds: its time and the
        # single-move time are measured apart and combined 1:1
        single = cpu_rate(replace(s2, move_disp=0.5, move_rot=0.5, move_mp=0.0, _keep=[]), 100,
                          threads)
        o = Oracle(replace(s2, move_disp=0.0, move_rot=0.0, move_mp=1.0, _keep=[]),
                   threads=threads)
        t = time.perf_counter()
        o.run(1, record=False)
        t_mp = time.perf_counter() - t
        o.close()
        row.update(cpu_moves_per_s=1.0 / (0.5 / single + 0.5 * t_mp), cpu_threads=threads,
                   cpu_single_moves_per_s=single, cpu_s_per_mp_move=t_mp,
                   cpu_sample="100 single moves and 1 MultiParticle move, combined 1:1")
    out["config2"] = row
    print(f"# config2 {row}", file=sys.stderr, flush=True)

    s3 = systems.config3()
    r, k = gpu_rate(s3, 10000)
    rd, _ = gpu_rate(s3, 10000, device=True)
    row = dict(workload="GCMC methane/CO2, L 40 A, 300 K, rc 10 A, Ewald 1e-5, fugacities "
                        "960/400 bar (the box's densest state, about 980 molecules), 1000 "
                        "seeded, 40 % transfers (CBMC 16/8), 30 % translate, 30 % rotate",
               gpu_host_driver_moves_per_s=r,
               gpu_device_loop_moves_per_s=rd, moves_by_kind=k, gpu_sample="10000 moves")
    if with_cpu:
        row.update(cpu_moves_per_s=cpu_rate(s3, 3000, threads), cpu_threads=threads,
                   cpu_1thread_moves_per_s=cpu_rate(s3, 800, 1), cpu_sample="3000 / 800 moves")
    out["config3"] = row
    print(f"# config3 {row}", file=sys.stderr, flush=True)
    return out


def served_n_check(jobs, results, replicas: int = 4, burn: int = 100, sample: int = 200) -> dict:
    """<N> per kind of the stream's densest job in the served precision
    (mixed) against FP64: `replicas` runs a side from the job's initial
    state, every run with its own MT19937 seed (the two sides' seeds differ
    too: under one seed the two precisions take the same decisions for
    hundreds of cycles, which would compare nothing), `burn` cycles, then
    `sample` cycles sampled; standard errors from the replica scatter.  The FP64 side is the
    device move loop in FP64, which takes the FP64 oracle's decisions move
    for move (tests/test_gpu_devloop.py), so it stands for the oracle at a
    size the oracle cannot run here in seconds."""
    from dataclasses import replace
    from gomc_amd.api import Box, run_device
    j = int(np.argmax([r["mean_n"].sum() for r in results]))
    s0 = jobs[j].system
    big = 1 << 40
    out = dict(job=j, box_A=float(s0.box[0]), temperature_K=float(s0.temperature),
               fugacity_bar=[float(f) for f in s0.fugacity], replicas=replicas,
               burn_cycles=burn, sample_cycles=sample,
               served_mean_n=[float(v) for v in results[j]["mean_n"]])
    for prec, tag in ((1, "mixed"), (0, "fp64")):
        boxes = [Box(replace(s0, seed=int(s0.seed) * 7919 + 1000 * prec + r + 1, _keep=[]))
                 for r in range(replicas)]
        for b in boxes:
            b.set_precision(prec)
        moves, _, _, _ = run_device(boxes, big, cycles=burn)
        _, _, sums, samples = run_device(boxes, big, start_steps=moves, cycles=sample)
        n = sums / samples[:, None]
        for b in boxes:
            b.close()
        out[tag] = [float(v) for v in n.mean(0)]
        out[tag + "_se"] = [float(v) for v in n.std(0, ddof=1) / math.sqrt(replicas)]
    out["z"] = [float(abs(out["mixed"][k] - out["fp64"][k]) /
                      max(math.hypot(out["mixed_se"][k], out["fp64_se"][k]), 1e-12)) for k in range(2)]
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--boxes", type=int, default=256)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--plain-moves", type=int, default=400, help="plain-path sample per box")
    ap.add_argument("--plain-boxes", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the virial / NPT / replica-exchange / checkpoint timings")
    ap.add_argument("--serve-jobs", type=int, default=60, help="0 skips the serving run")
    ap.add_argument("--serve-gcmc-cycles", type=int, default=2000)
    ap.add_argument("--serve-gcmc-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the single-box workload rows (configs 1-3)")
    ap.add_argument("--serve-gcmc-jobs", type=int, default=60,
                    help="GCMC state-point jobs in the served run (0 skips)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (a known-size reference "
                         "kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from gomc_amd import systems
    from gomc_amd.api import Batch, Box

    t0 = time.time()
    boxes = systems.config4(args.boxes)
    nmol = np.array([b.nmols for b in boxes], dtype=np.int64)
    print(f"# generated {len(boxes)} boxes, {nmol.sum()} molecules in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    bt = Batch(boxes, precision=args.precision)
    # k-vector counts (the reference formula) for the work model
    from gomc_amd._lib import load  # noqa: F401
    nk = []
    for b in boxes:
        L = b.box[0]
        alpha = math.sqrt(-math.log(b.tolerance)) / b.rcut_coulomb
        kc = -2.0 * math.log(b.tolerance) / b.rcut_coulomb
        nmax = int(kc * L / (2 * math.pi)) + 1
        c = 2 * math.pi / L
        n = np.arange(-nmax, nmax + 1)
        X, Y, Z = np.meshgrid(np.arange(0, nmax + 1), n, n, indexing="ij")
        half = (X > 0) | ((X == 0) & (Y > 0)) | ((X == 0) & (Y == 0) & (Z > 0))
        ks = (c * c) * (X * X + Y * Y + Z * Z)
        nk.append(int(np.count_nonzero(half & (ks < kc * kc))))
    flops_step = float(sum(nmol[i] * move_flops(boxes[i], nk[i]) for i in range(len(boxes))))
    bytes_step = float(sum(nmol[i] * move_bytes(boxes[i], nk[i]) for i in range(len(boxes))))
    print(f"# batch ready ({time.time() - t0:.1f}s); warmup {args.warmup}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        bt.run_cycles(1, graph=True)   # the first call captures the cycle
    kernel_ms = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        bt.run_cycles(1, graph=True)
        kernel_ms.append(bt.last_kernel_ms())
        print(f"# step {i}: kernel {kernel_ms[-1]:.1f} ms", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    moves = float(nmol.sum() * args.steps)
    value = moves / elapsed
    avg_ms = float(np.mean(kernel_ms))
    achieved_tflops = flops_step / (avg_ms * 1e-3) / 1e12
    achieved_gbs = bytes_step / (avg_ms * 1e-3) / 1e9
    e, acc, done = bt.stats()
    accept_rate = float(acc.sum() / done.sum())

    mixed = None
    if not args.no_mixed and args.precision == 0:
        bm = Batch(boxes, precision=1)
        for _ in range(args.warmup):
            bm.run_cycles(1, graph=True)
        mk = []
        tm = time.perf_counter()
        for i in range(args.steps):
            bm.run_cycles(1, graph=True)
            mk.append(bm.last_kernel_ms())
        tm = time.perf_counter() - tm
        mixed = dict(moves_per_s=float(nmol.sum() * args.steps / tm),
                     kernel_ms_per_step=float(np.mean(mk)),
                     dtype="f32 pair terms and phase products, f64 accumulation")
        bm.close()
        print(f"# mixed precision {mixed['moves_per_s']:.0f} moves/s", file=sys.stderr, flush=True)
    bt.close()

    served = None
    if args.serve_jobs > 0:
        from gomc_amd.serve import Server, poisson_stream, run_stream
        jobs = poisson_stream(args.serve_jobs, rate=20.0, seed=11, nmin=200, nmax=2000,
                              cycles=20)
        sv = Server(precision=1, slots=256, budget_bytes=8 << 30, tick_ms=20.0)
        rep = run_stream(sv, jobs, timeout=300.0)
        sv.close()
        served = dict(workload=f"Poisson(20/s) stream of {rep['jobs']} NVT water state-point "
                               "jobs, N log-uniform 200-2000, T 250-350 K, 20 cycles each; "
                               "mixed precision (FP32 pair terms and phases, FP64 sums)",
                      moves_per_s=rep["moves_per_s"], p50_s=rep["p50_s"], p99_s=rep["p99_s"],
                      mean_queue_s=rep["mean_queue_s"], wall_s=rep["wall_s"],
                      moves=rep["moves"], budget_bytes=rep["stats"]["budget_bytes"],
                      peak_bytes=rep["stats"]["peak_bytes"])
        print(f"# served {served['moves_per_s']:.0f} moves/s p50 {served['p50_s']:.2f} s "
              f"p99 {served['p99_s']:.2f} s", file=sys.stderr, flush=True)

    served_gcmc = None
    if args.serve_gcmc_jobs > 0:
        from gomc_amd.serve import Server, poisson_stream_gcmc, run_stream
        jobs = poisson_stream_gcmc(args.serve_gcmc_jobs, rate=20.0, seed=17, lmin=20.0,
                                   lmax=60.0, cycles=args.serve_gcmc_cycles)
        sv = Server(precision=1, slots=64, budget_bytes=int(args.serve_gcmc_budget_gb * 2**30),
                    tick_ms=20.0)
        rep = run_stream(sv, jobs, timeout=480.0, progress=30.0)
        sv.close()
        served_gcmc = dict(
            workload=f"Poisson(20/s) stream of {rep['jobs']} GCMC methane/CO2 adsorption jobs, "
                     f"L 20-60 A, T 250-350 K, fugacities log-uniform 1e3-1e7 Pa, "
                     f"{args.serve_gcmc_cycles} cycles each; GCMC lane (device move loop in "
                     "mixed precision, every running job in each launch)",
            moves_per_s=rep["moves_per_s"], p50_s=rep["p50_s"], p99_s=rep["p99_s"],
            moves=rep["moves"], wall_s=rep["wall_s"],
            budget_bytes=rep["stats"]["budget_bytes"], peak_bytes=rep["stats"]["peak_bytes"],
            mean_n_range=[float(min(r["mean_n"].sum() for r in rep["results"])),
                          float(max(r["mean_n"].sum() for r in rep["results"]))])
        print(f"# served GCMC {served_gcmc['moves_per_s']:.0f} moves/s p50 "
              f"{served_gcmc['p50_s']:.2f} s p99 {served_gcmc['p99_s']:.2f} s; <N> per job "
              f"{served_gcmc['mean_n_range']}", file=sys.stderr, flush=True)
        served_gcmc["n_check"] = served_n_check(jobs, rep["results"])
        print(f"# served GCMC <N> check {served_gcmc['n_check']}", file=sys.stderr, flush=True)

    plain = None
    if not args.no_plain:
        idx = np.linspace(0, len(boxes) - 1, min(args.plain_boxes, len(boxes))).astype(int)
        tot_moves, tot_time = 0, 0.0
        for i in idx:
            bx = Box(boxes[i])
            bx.run(50, record=False)  # warm
            t = time.perf_counter()
            bx.run(args.plain_moves, start_step=50, record=False)
            tot_time += time.perf_counter() - t
            tot_moves += args.plain_moves
            bx.close()
        # the plain path advances one box at a time: its rate is moves over
        # the summed per-box times
        plain_rate = tot_moves / tot_time
        plain = dict(moves_per_s=plain_rate, sample=f"{len(idx)} boxes x {args.plain_moves} moves")
        print(f"# plain path {plain_rate:.0f} moves/s", file=sys.stderr, flush=True)

    cpu = None
    threads, seen = cpu_threads()
    if not args.no_cpu:
        from oracle.pyoracle import Oracle, batch_run
        oracles = [Oracle(b, threads=1) for b in boxes]
        # calibrate: a short run, then scale to ~cpu-seconds of work
        probe = np.maximum(1, nmol // 200)
        tp = batch_run(oracles, probe, threads)
        rate = probe.sum() / max(tp, 1e-6)
        frac = max(1e-3, min(1.0, args.cpu_seconds * rate / nmol.sum()))
        sample = np.maximum(1, (nmol * frac).astype(np.int64))
        tc = batch_run(oracles, sample, threads)
        # the same work model on one core: a smaller sample of the same boxes
        s1 = np.maximum(1, (sample // max(threads, 1)).astype(np.int64))
        t1 = batch_run(oracles, s1, 1)
        cpu = dict(value=float(sample.sum() / tc), unit="moves/s", cores=threads, kind="port",
                   sample=f"{int(sample.sum())} Philox-stream moves over the same {len(boxes)} "
                          f"boxes ({frac:.3f} of one MC cycle per box), one box per thread",
                   cores_visible=seen, value_1thread=float(s1.sum() / t1),
                   sample_1thread=f"{int(s1.sum())} moves over the same boxes on one thread")
        print(f"# cpu oracle {cpu['value']:.0f} moves/s on {threads} threads, "
              f"{cpu['value_1thread']:.0f} on one", file=sys.stderr, flush=True)

    cfg = None
    if not args.no_configs:
        cfg = config_rows(threads, not args.no_cpu)

    nxt = None
    if not args.no_next:
        nxt = next_rows(threads, not args.no_cpu)

    if args.calibrate_copy > 0:
        copy_reference(args.calibrate_copy << 20)
        print(f"# copy reference: {args.calibrate_copy} MiB read + written per copy, 3 copies",
              file=sys.stderr, flush=True)

    default_workload = args.boxes == 256 and args.precision == 0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32-pair/f64-accumulate",
        "data": "synthetic",
        "config": {
            "workload": "batched NVT: 256 rigid SPC/E water boxes, N log-uniform 200-4000, "
                        "0.997 g/cm3, T 280-360 K, rc min(10, L/2) A, Ewald tol 1e-5; "
                        "step = one MC cycle (N_i moves per box)",
            "boxes": len(boxes),
            "molecules": int(nmol.sum()),
            "moves_per_step": int(nmol.sum()),
            "precision": args.precision,
        },
        "kernel_ms_per_step": avg_ms,
        "accept_rate": accept_rate,
        "graph": "one HIP graph per MC cycle, replayed",
        "mixed_precision": mixed,
        "served": served,
        "served_gcmc": served_gcmc,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain": (value / plain["moves_per_s"]) if plain else None,
        "mixed_speedup_vs_plain": (mixed["moves_per_s"] / plain["moves_per_s"])
        if (plain and mixed) else None,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": PROFILED_TRAFFIC["bytes"] if default_workload else None,
            "traffic_source": PROFILED_TRAFFIC["source"] if default_workload else None,
            "kernel": "McBatchKernel",
            "bytes_per_launch": bytes_step,
            "byte_model": "per move 64 K (k-vector data) + 40 M (27-cell neighbourhood atoms)",
        },
        "roofline_valu": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "flops_per_launch": flops_step,
        },
        "cpu_baseline": cpu,
        "configs_1_3": cfg,
        "next_rows": nxt,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
