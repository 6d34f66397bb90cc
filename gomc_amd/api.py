"""Python handles over the C ABI: `Box` (one box, the plain path) and
`Batch` (many boxes, one persistent launch).  Both call libgomc_hip.so for
every energy; nothing here computes an energy on the CPU."""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import ptr
from .systems import SystemDesc, System

ENERGY_NAMES = ("inter", "tail", "real", "recip", "self", "correction", "total")


class GomcError(RuntimeError):
    pass


class Box:
    """One simulation box resident on a gfx950 device (gomc_box)."""

    def __init__(self, system: System, device: int = 0):
        self.lib = _lib.load()
        self.system = system
        self._desc = system.desc()
        h = C.c_void_p()
        rc = self.lib.gomc_box_create(C.byref(self._desc), device, C.byref(h))
        if rc != 0:
            raise GomcError(f"gomc_box_create failed ({rc})")
        self.h = h
        info = np.zeros(7, dtype=np.int64)
        self._check(self.lib.gomc_box_info(self.h, ptr(info)))
        self.natoms, self.nmols, self.nkvec, self.ncells = (int(v) for v in info[:4])
        self.edge = tuple(int(v) for v in info[4:])
        self.mol_start = system.mol_start()

    def _check(self, rc):
        if rc < 0:
            raise GomcError(self.lib.gomc_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.gomc_box_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----
    def coords(self) -> np.ndarray:
        x, y, z = (np.zeros(self.natoms) for _ in range(3))
        self._check(self.lib.gomc_get_coords(self.h, ptr(x), ptr(y), ptr(z)))
        return np.stack([x, y, z], axis=1)

    def set_coords(self, xyz: np.ndarray):
        x, y, z = (np.ascontiguousarray(xyz[:, i], dtype=np.float64) for i in range(3))
        self._check(self.lib.gomc_set_coords(self.h, ptr(x), ptr(y), ptr(z)))

    def mol_len(self, m: int) -> int:
        return len(self.system.kinds[int(self.system.mol_kind[m])].atom_types)

    # ---- energies ----
    def box_inter(self) -> tuple[float, float]:
        lj, real = C.c_double(), C.c_double()
        self._check(self.lib.gomc_box_inter(self.h, C.byref(lj), C.byref(real)))
        return lj.value, real.value

    def system_total(self) -> np.ndarray:
        out = np.zeros(7)
        self._check(self.lib.gomc_system_total(self.h, ptr(out)))
        return out

    def energies(self) -> np.ndarray:
        out = np.zeros(7)
        self._check(self.lib.gomc_energies(self.h, ptr(out)))
        return out

    def _xyz3(self, xyz):
        xyz = np.asarray(xyz, dtype=np.float64)
        return tuple(np.ascontiguousarray(xyz[:, i]) for i in range(3))

    def mol_inter(self, m: int, new_xyz) -> tuple[float, float, int]:
        x, y, z = self._xyz3(new_xyz)
        lj, real, ov = C.c_double(), C.c_double(), C.c_int32()
        self._check(self.lib.gomc_mol_inter(self.h, m, ptr(x), ptr(y), ptr(z), C.byref(lj),
                                            C.byref(real), C.byref(ov)))
        return lj.value, real.value, ov.value

    def mol_recip(self, m: int, new_xyz) -> float:
        x, y, z = self._xyz3(new_xyz)
        e = C.c_double()
        self._check(self.lib.gomc_mol_recip(self.h, m, ptr(x), ptr(y), ptr(z), C.byref(e)))
        return e.value

    def mol_trial(self, m: int, new_xyz) -> np.ndarray:
        x, y, z = self._xyz3(new_xyz)
        out = np.zeros(4)
        self._check(self.lib.gomc_mol_trial(self.h, m, ptr(x), ptr(y), ptr(z), ptr(out)))
        return out

    def swap_recip(self, kind: int, insert: bool, xyz, mol: int = -1) -> float:
        x, y, z = self._xyz3(xyz)
        e = C.c_double()
        self._check(self.lib.gomc_swap_recip(self.h, kind, mol, int(insert), ptr(x), ptr(y),
                                             ptr(z), C.byref(e)))
        return e.value

    def set_precision(self, precision: int):
        """0 FP64; 1 FP32 pair terms and phases in trials, FP64 sums."""
        self._check(self.lib.gomc_set_precision(self.h, precision))

    def recip_cache(self, enable: bool = True):
        """EwaldCached mode: per-molecule structure factors kept in HBM."""
        self._check(self.lib.gomc_recip_cache(self.h, int(enable)))

    def particle_inter(self, atom_type: int, charge: float, trials, skip_mol: int = -1):
        x, y, z = self._xyz3(trials)
        n = len(x)
        en, real = np.zeros(n), np.zeros(n)
        ov = np.zeros(n, dtype=np.int32)
        self._check(self.lib.gomc_particle_inter(self.h, atom_type, charge, skip_mol, n, ptr(x),
                                                 ptr(y), ptr(z), ptr(en), ptr(real), ptr(ov)))
        return en, real, ov

    def recip_init(self) -> float:
        e = C.c_double()
        self._check(self.lib.gomc_recip_init(self.h, C.byref(e)))
        return e.value

    def recip_sums(self) -> float:
        e = C.c_double()
        self._check(self.lib.gomc_recip_sums(self.h, C.byref(e)))
        return e.value

    def set_device_team(self, team: int):
        """Workgroups per box in the device move loop (0: by k-vector count)."""
        self._check(self.lib.gomc_set_device_team(self.h, int(team)))

    def recip_sums_noalloc(self) -> float:
        """recip_sums with its scratch in the library's static device array
        (no device allocation in the call)."""
        e = C.c_double()
        self._check(self.lib.gomc_recip_sums_noalloc(self.h, C.byref(e)))
        return e.value

    def recip_accept(self):
        self._check(self.lib.gomc_recip_accept(self.h))

    def recip_get(self):
        arrs = [np.zeros(self.nkvec) for _ in range(6)]
        self._check(self.lib.gomc_recip_get(self.h, *(ptr(a) for a in arrs)))
        return dict(zip(("kx", "ky", "kz", "prefact", "sum_r", "sum_i"), arrs))

    def commit_mol(self, m: int, xyz):
        x, y, z = self._xyz3(xyz)
        self._check(self.lib.gomc_commit_mol(self.h, m, ptr(x), ptr(y), ptr(z)))

    def build_cells(self):
        cv = np.zeros(self.natoms, dtype=np.int32)
        cs = np.zeros(self.ncells + 1, dtype=np.int32)
        ac = np.zeros(self.natoms, dtype=np.int32)
        self._check(self.lib.gomc_build_cells(self.h, ptr(cv), ptr(cs), ptr(ac)))
        return cv, cs, ac

    def trial_cells(self):
        """The device's trial cell table: (count[n_cells], slots[n_cells, cap])."""
        info = np.zeros(7, dtype=np.int64)
        self._check(self.lib.gomc_box_info(self.h, ptr(info)))
        nc = int(info[3])  # a volume move may have changed the grid
        cap = C.c_int32()
        self._check(self.lib.gomc_trial_cells(self.h, C.byref(cap), None, None))
        cnt = np.zeros(nc, dtype=np.int32)
        slots = np.zeros(nc * cap.value, dtype=np.int32)
        self._check(self.lib.gomc_trial_cells(self.h, C.byref(cap), ptr(cnt), ptr(slots)))
        return cnt, slots.reshape(nc, cap.value)

    # ---- forces and the MultiParticle move ----
    def box_force(self):
        """(lj, real, atom forces (n_atoms, 3), molecule forces (n_mols, 3))."""
        lj, real = C.c_double(), C.c_double()
        a = [np.zeros(self.natoms) for _ in range(3)] + [np.zeros(self.nmols) for _ in range(3)]
        self._check(self.lib.gomc_box_force(self.h, C.byref(lj), C.byref(real),
                                            *(ptr(v) for v in a)))
        return lj.value, real.value, np.stack(a[:3], 1), np.stack(a[3:], 1)

    def box_force_recip(self, copy_ref_to_new: bool = True):
        if copy_ref_to_new:
            self._check(self.lib.gomc_recip_copy_ref_to_new(self.h))
        a = [np.zeros(self.natoms) for _ in range(3)] + [np.zeros(self.nmols) for _ in range(3)]
        self._check(self.lib.gomc_box_force_recip(self.h, *(ptr(v) for v in a)))
        return np.stack(a[:3], 1), np.stack(a[3:], 1)

    def torque(self) -> np.ndarray:
        t = [np.zeros(self.nmols) for _ in range(3)]
        self._check(self.lib.gomc_torque(self.h, *(ptr(v) for v in t)))
        return np.stack(t, 1)

    def dims(self) -> np.ndarray:
        """Current edge lengths and volume: [Lx, Ly, Lz, V]."""
        out = np.zeros(4)
        self._check(self.lib.gomc_box_dims(self.h, ptr(out)))
        return out

    def virial(self) -> dict:
        """Molecular virial (K) and pressure (bar) of the current configuration:
        {inter, tail, real, recip, total, pressure, tensor (xx, yy, zz)}."""
        out = np.zeros(9)
        self._check(self.lib.gomc_box_virial(self.h, ptr(out)))
        return dict(inter=out[0], tail=out[1], real=out[2], recip=out[3], total=out[4],
                    pressure=out[5], tensor=out[6:9].copy())

    def checkpoint(self, step: int = 0) -> bytes:
        """The box's Monte Carlo state (coordinates, MT19937 state, move
        settings, molecule lookup, running energies, structure factors) and
        `step`, as bytes that `restore` takes back."""
        n = self.lib.gomc_box_checkpoint(self.h, int(step), None, 0)
        if n < 0:
            self._check(int(n))
        buf = C.create_string_buffer(int(n))
        m = self.lib.gomc_box_checkpoint(self.h, int(step), C.cast(buf, C.c_void_p), int(n))
        if m != n:
            self._check(int(m) if m < 0 else -3)
        return buf.raw

    def restore(self, data: bytes) -> int:
        """Restore a checkpoint of a box of the same system; returns its step."""
        step = np.zeros(1, np.int64)
        buf = C.create_string_buffer(bytes(data), len(data))
        self._check(self.lib.gomc_box_restore(self.h, C.cast(buf, C.c_void_p), len(data),
                                              ptr(step)))
        return int(step[0])

    def save(self, path: str, step: int = 0) -> None:
        with open(path, "wb") as f:
            f.write(self.checkpoint(step))

    def load(self, path: str) -> int:
        with open(path, "rb") as f:
            return self.restore(f.read())

    def mp_trial(self, move_type: int, step: int, max_step: float) -> np.ndarray:
        out = np.zeros(5)
        self._check(self.lib.gomc_mp_trial(self.h, move_type, step, max_step, ptr(out)))
        return out

    def mp_accept(self, accept: bool, out: np.ndarray):
        self._check(self.lib.gomc_mp_accept(self.h, int(accept), ptr(out)))

    def mp_settings(self) -> np.ndarray:
        out = np.zeros(2)
        self._check(self.lib.gomc_mp_settings(self.h, ptr(out)))
        return out

    # ---- MC ----
    def run(self, nsteps: int, start_step: int = 0, record: bool = True, device: bool = False):
        """Run `nsteps` moves from `start_step`: on the host driver (one fused
        launch per energy call) or, with device=True, in one launch of the
        device move loop (same draws, same moves, same outcomes)."""
        fn = self.lib.gomc_run_moves_device if device else self.lib.gomc_run_moves
        if not record:
            return self._check(fn(self.h, start_step, nsteps, None, None, None,
                                  None, None, None, None)), None
        rec = dict(kind=np.zeros(nsteps, np.int32), mol=np.zeros(nsteps, np.int32),
                   acc=np.zeros(nsteps, np.int32), ovl=np.zeros(nsteps, np.int32),
                   dlj=np.zeros(nsteps), dreal=np.zeros(nsteps), drecip=np.zeros(nsteps))
        n = self._check(fn(self.h, start_step, nsteps, ptr(rec["kind"]),
                                                ptr(rec["mol"]), ptr(rec["acc"]), ptr(rec["ovl"]),
                                                ptr(rec["dlj"]), ptr(rec["dreal"]),
                                                ptr(rec["drecip"])))
        return n, rec

    def mol_boxes(self):
        """(molecules in box 0, box of every molecule: 0 box, 1 reservoir)."""
        b = np.zeros(self.nmols, dtype=np.int32)
        n = self._check(self.lib.gomc_mol_boxes(self.h, ptr(b)))
        return n, b

    def scales(self):
        nk = len(self.system.kinds)
        d, r = np.zeros(nk), np.zeros(nk)
        self._check(self.lib.gomc_get_scales(self.h, ptr(d), ptr(r)))
        return d, r


def run_device(boxes, nsteps, start_steps=None, cycles=None):
    """One launch of the device move loop over several boxes (one workgroup
    each).  `nsteps` and `cycles` are per box (scalars broadcast); with
    cycles, each box runs up to that many cycles of N_box moves and samples
    its kind-0/1 counts after each.  Returns (moves, accepted, count sums
    (n, 2), samples)."""
    n = len(boxes)
    lib = boxes[0].lib
    hs = (C.c_void_p * n)(*[b.h for b in boxes])
    ns = np.broadcast_to(np.asarray(nsteps, np.int64), (n,)).copy()
    st = (np.zeros(n, np.int64) if start_steps is None
          else np.broadcast_to(np.asarray(start_steps, np.int64), (n,)).copy())
    cy = (None if cycles is None
          else np.broadcast_to(np.asarray(cycles, np.int32), (n,)).copy())
    moves, acc, samples = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int64)
    sums = np.zeros((n, 2))
    status = np.zeros(n, np.int32)
    rc = lib.gomc_run_moves_device_multi(hs, n, ptr(st), ptr(ns), None if cy is None else ptr(cy),
                                         ptr(moves), ptr(acc), ptr(sums), ptr(samples),
                                         ptr(status))
    if rc != 0:
        bad = [i for i in range(n) if status[i] != 0] or [0]
        msg = lib.gomc_last_error(boxes[bad[0]].h).decode()
        raise GomcError(f"device move loop failed ({rc}) for box {bad[0]}: {msg}")
    return moves, acc, sums, samples


TUNING_FILE = Path(__file__).resolve().parent / "tuning.json"


def load_tuning(path=TUNING_FILE) -> dict | None:
    """The launch shape written by gomc_amd.autotune, if there is one."""
    try:
        return json.loads(Path(path).read_text())["tuning"]
    except (OSError, KeyError, ValueError):
        return None


class Batch:
    """Many independent boxes advanced by one persistent launch.

    tuning: {"team_min": [n2, n4, n8], "pair12_max_k": k} (see
    gomc_batch_tuning); None takes gomc_amd/tuning.json when present, else
    the library defaults; "default" forces the library defaults."""

    def __init__(self, systems: list[System], device: int = 0, precision: int = 0,
                 tuning: dict | str | None = None):
        self.lib = _lib.load()
        self.systems = systems
        arr = (SystemDesc * len(systems))()
        for i, s in enumerate(systems):
            arr[i] = s.desc()
        self._descs = arr
        if tuning is None:
            tuning = load_tuning()
        tp = None
        if isinstance(tuning, dict):
            t = _lib.BatchTuning()
            self.lib.gomc_batch_default_tuning(C.byref(t))
            for i, v in enumerate(tuning.get("team_min", list(t.team_min))):
                t.team_min[i] = int(v)
            t.pair12_max_k = int(tuning.get("pair12_max_k", t.pair12_max_k))
            tp = C.byref(t)
        t_used = _lib.BatchTuning()
        self.lib.gomc_batch_default_tuning(C.byref(t_used))
        if isinstance(tuning, dict):
            t_used = t
        self.tuning = {"team_min": [int(v) for v in t_used.team_min],
                       "pair12_max_k": int(t_used.pair12_max_k)}
        h = C.c_void_p()
        rc = self.lib.gomc_batch_create_tuned(arr, len(systems), device, precision, tp, C.byref(h))
        if rc != 0:
            raise GomcError(f"gomc_batch_create failed ({rc})")
        self.h = h
        self.nboxes = len(systems)

    def box_times(self) -> np.ndarray:
        """Per-box wall time of the last launch, microseconds."""
        out = np.zeros(self.nboxes)
        self._check(self.lib.gomc_batch_box_times(self.h, ptr(out)))
        return out

    def _check(self, rc):
        if rc < 0:
            raise GomcError(self.lib.gomc_batch_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.gomc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, moves=None):
        mv = None if moves is None else np.ascontiguousarray(moves, dtype=np.int64)
        self._check(self.lib.gomc_batch_run(self.h, ptr(mv) if mv is not None else None))

    def run_cycles(self, ncycles: int, graph: bool = True):
        """`ncycles` MC cycles, replaying one captured HIP graph per cycle."""
        self._check(self.lib.gomc_batch_run_cycles(self.h, ncycles, int(graph)))

    def stats(self):
        e = np.zeros((self.nboxes, 7))
        acc = np.zeros(self.nboxes, dtype=np.int64)
        done = np.zeros(self.nboxes, dtype=np.int64)
        self._check(self.lib.gomc_batch_stats(self.h, ptr(e), ptr(acc), ptr(done)))
        return e, acc, done

    def exchange(self, ladder: np.ndarray, attempt: int, seed: int = 0) -> np.ndarray:
        """One replica-exchange attempt over the boxes listed by temperature
        slot in `ladder` (int32, updated in place); returns accepted[s] for
        the pairs (s-1, s)."""
        assert ladder.dtype == np.int32 and ladder.flags.c_contiguous
        acc = np.zeros(len(ladder), np.int32)
        self._check(self.lib.gomc_batch_exchange(self.h, ptr(ladder), len(ladder), attempt, seed,
                                                 ptr(acc)))
        return acc

    def coords(self, box: int) -> np.ndarray:
        n = self.systems[box].natoms
        x, y, z = (np.zeros(n) for _ in range(3))
        self._check(self.lib.gomc_batch_get_coords(self.h, box, ptr(x), ptr(y), ptr(z)))
        return np.stack([x, y, z], axis=1)

    def layout(self, box: int) -> dict:
        """Device layout of a box: team size, k-vectors per segment, segment
        rounds, centre cull."""
        out = np.zeros(4, dtype=np.int32)
        self._check(self.lib.gomc_batch_box_layout(self.h, box, ptr(out)))
        return dict(team=int(out[0]), kseg=int(out[1]), rounds=int(out[2]), cull=bool(out[3]))

    def set_trace(self, box: int, cap: int):
        self._check(self.lib.gomc_batch_set_trace(self.h, box, cap))

    def trace(self):
        n = C.c_int64()
        self._check(self.lib.gomc_batch_get_trace(self.h, None, None, None, None, C.byref(n)))
        k = n.value
        acc = np.zeros(k, np.int32)
        dlj, dreal, drecip = np.zeros(k), np.zeros(k), np.zeros(k)
        self._check(self.lib.gomc_batch_get_trace(self.h, ptr(acc), ptr(dlj), ptr(dreal),
                                                  ptr(drecip), C.byref(n)))
        return dict(acc=acc, dlj=dlj, dreal=dreal, drecip=drecip)

    def last_kernel_ms(self) -> float:
        v = C.c_double()
        self._check(self.lib.gomc_batch_last_kernel_ms(self.h, C.byref(v)))
        return v.value
