"""Serving front end over `gomc_server` (include/gomc_hip.h): a request
loop that keeps state-point jobs resident on the device and batches every
resident job into each launch of the persistent MC kernel.

`Server` wraps the C handle; `poisson_stream` draws a seeded job stream;
`run_stream` submits the jobs at their arrival times and reports moves/s
and the p50 / p99 job latency (submission to completion).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib, systems
from ._lib import ptr
from .api import GomcError


class Server:
    """One serving loop on device `device` (gomc_server)."""

    def __init__(self, device: int = 0, precision: int = 0, slots: int = 256,
                 budget_bytes: int = 2 << 30, tick_ms: float = 20.0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.gomc_server_create(device, precision, slots, int(budget_bytes), tick_ms,
                                         C.byref(h))
        if rc != 0:
            raise GomcError(f"gomc_server_create failed ({rc})")
        self.h = h
        self._keep = {}

    def _check(self, rc):
        if rc < 0:
            raise GomcError(self.lib.gomc_server_last_error(self.h).decode())
        return rc

    def submit(self, system: systems.System, moves: int) -> int:
        d = system.desc()
        jid = C.c_int64()
        self._check(self.lib.gomc_server_submit(self.h, C.byref(d), int(moves), C.byref(jid)))
        return jid.value

    def poll(self, job: int) -> dict:
        st = C.c_int32()
        times, en = np.zeros(3), np.zeros(7)
        acc, done = C.c_int64(), C.c_int64()
        self._check(self.lib.gomc_server_poll(self.h, job, C.byref(st), ptr(times), ptr(en),
                                              C.byref(acc), C.byref(done)))
        mean_n = np.zeros(2)
        self._check(self.lib.gomc_server_poll_counts(self.h, job, ptr(mean_n)))
        return dict(status=st.value, submitted=times[0], admitted=times[1], done=times[2],
                    energies=en, accepted=acc.value, moves_done=done.value, mean_n=mean_n)

    def last_error(self) -> str:
        return self.lib.gomc_server_last_error(self.h).decode()

    def drain(self, timeout: float = 600.0):
        self._check(self.lib.gomc_server_drain(self.h, timeout))

    def stats(self) -> dict:
        out = np.zeros(8)
        self._check(self.lib.gomc_server_stats(self.h, ptr(out)))
        keys = ("moves", "kernel_s", "launches", "peak_bytes", "budget_bytes", "elapsed_s",
                "queued", "outstanding")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.gomc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Job:
    arrival: float          # seconds after the stream starts
    system: systems.System
    moves: int


def poisson_stream(njobs: int, rate: float = 20.0, seed: int = 7, nmin: int = 200,
                   nmax: int = 2000, tmin: float = 250.0, tmax: float = 350.0,
                   cycles: int = 20) -> list[Job]:
    """Seeded stream of NVT water state-point jobs: exponential
    inter-arrival times of mean 1/rate, molecule counts log-uniform in
    [nmin, nmax] at 0.997 g/cm^3, temperatures uniform in [tmin, tmax],
    `cycles` MC cycles (N moves each) per job, one Random123 stream each."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    for i in range(njobs):
        t += rng.exponential(1.0 / rate)
        n = int(math.exp(rng.uniform(math.log(nmin), math.log(nmax))))
        temp = float(rng.uniform(tmin, tmax))
        s = systems.water_box(n, temperature=temp, seed=seed * 100003 + i,
                              place_seed=seed * 7919 + i)
        jobs.append(Job(t, s, cycles * n))
    return jobs


def poisson_stream_gcmc(njobs: int, rate: float = 20.0, seed: int = 17, lmin: float = 20.0,
                        lmax: float = 60.0, tmin: float = 250.0, tmax: float = 350.0,
                        pmin: float = 1e3, pmax: float = 1e7, cycles: int = 2000,
                        density: float = 0.004) -> list[Job]:
    """Seeded stream of GCMC adsorption state points: methane / CO2 boxes of
    edge uniform in [lmin, lmax] A, temperature uniform in [tmin, tmax] K,
    each kind's fugacity log-uniform in [pmin, pmax] Pa; the box starts with
    `density` molecules per A^3 split evenly between the kinds, beside a
    reservoir of 0.015 molecules per A^3 of each kind to insert from (more
    than the box holds at these state points: the reference stops when the
    reservoir runs dry); `cycles` MC cycles of the initial molecule count
    each.  40 % transfers, 30 % translations, 30 % rotations."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    for i in range(njobs):
        t += rng.exponential(1.0 / rate)
        L = float(rng.uniform(lmin, lmax))
        temp = float(rng.uniform(tmin, tmax))
        fug = tuple(float(math.exp(rng.uniform(math.log(pmin), math.log(pmax)))) * 1e-5
                    for _ in range(2))  # Pa -> bar
        n = max(2, int(density * L ** 3) // 2)
        s = systems.gcmc_mixture(n, n, L, temperature=temp, fugacity_bar=fug,
                                 reservoir=int(0.015 * L ** 3) + 32, seed=seed * 1009 + i)
        jobs.append(Job(t, s, cycles * 2 * n))
    return jobs


def run_stream(server: Server, jobs: list[Job], timeout: float = 600.0,
               progress: float = 0.0) -> dict:
    """Submits each job at its arrival time (relative to now), waits for all
    of them, and returns throughput and latency figures.  With progress > 0
    a line with the jobs still outstanding is printed to stderr every
    `progress` seconds while it waits."""
    import sys
    t0 = time.perf_counter()
    ids = []
    for j in jobs:
        delay = j.arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(j.system, j.moves))
    if progress > 0:
        while True:
            left = timeout - (time.perf_counter() - t0)
            if left <= 0:
                raise GomcError("served jobs did not finish in time")
            rc = server.lib.gomc_server_drain(server.h, min(progress, left))
            if rc == 0:
                break
            st = server.stats()
            print(f"# serving: {int(st['outstanding'])} jobs outstanding after "
                  f"{time.perf_counter() - t0:.0f} s, {st['moves']:.3g} moves done",
                  file=sys.stderr, flush=True)
    else:
        server.drain(timeout)
    wall = time.perf_counter() - t0
    res = [server.poll(i) for i in ids]
    if any(r["status"] != 2 for r in res):
        raise GomcError("a served job failed")
    lat = np.array([r["done"] - r["submitted"] for r in res])
    moves = sum(j.moves for j in jobs)
    return dict(jobs=len(jobs), moves=moves, wall_s=wall, moves_per_s=moves / wall,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_queue_s=float(np.mean([r["admitted"] - r["submitted"] for r in res])),
                results=res, stats=server.stats())
