"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from gomc_amd.systems import SystemDesc, System

LIB = Path(__file__).resolve().parent / "liboracle.so"
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_up = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p

_SIG = {
    "orc_create": (_vp, [C.POINTER(SystemDesc)]),
    "orc_destroy": (None, [_vp]),
    "orc_plant": (None, [C.c_int]),
    "orc_set_scales": (None, [_vp, _dp, _dp, C.c_double, C.c_double]),
    "orc_natoms": (C.c_int, [_vp]),
    "orc_nmols": (C.c_int, [_vp]),
    "orc_nkvec": (C.c_int, [_vp]),
    "orc_cutoffs": (None, [_vp, _dp]),
    "orc_system_total": (None, [_vp, _dp]),
    "orc_energies": (None, [_vp, _dp]),
    "orc_box_inter": (None, [_vp, _dp, _dp]),
    "orc_box_inter_brute": (None, [_vp, _dp, _dp]),
    "orc_mol_inter": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    "orc_mol_recip": (C.c_double, [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    "orc_swap_recip": (C.c_double, [_vp, C.c_int, C.c_int, _dp, _dp, _dp]),
    "orc_particle_inter": (None, [_vp, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip]),
    "orc_kvectors": (None, [_vp, _dp, _dp, _dp, _dp]),
    "orc_recip_sums": (None, [_vp, _dp, _dp]),
    "orc_self_corr_lrc": (None, [_vp, _dp]),
    "orc_tail_change": (C.c_double, [_vp, C.c_int, C.c_int]),
    "orc_get_coords": (None, [_vp, _dp, _dp, _dp]),
    "orc_get_com": (None, [_vp, _dp, _dp, _dp]),
    "orc_get_scales": (None, [_vp, _dp, _dp]),
    "orc_cell_csr": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "orc_cells_exhaustive": (C.c_int, [_vp]),
    "orc_run": (C.c_int64, [_vp, C.c_int64, C.c_int64, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp]),
    "orc_box_force": (None, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "orc_box_force_recip": (None, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "orc_torque": (None, [_vp, _dp, _dp, _dp]),
    "orc_virial": (None, [_vp, _dp]),
    "orc_exchange": (C.c_int, [C.POINTER(_vp), _ip, C.c_int, C.c_int64, C.c_uint32, C.c_int64, _ip]),
    "orc_mp_settings": (None, [_vp, _dp]),
    "orc_mol_boxes": (C.c_int, [_vp, _ip]),
    "orc_batch_run": (C.c_double, [C.c_int, C.POINTER(_vp), _lp, C.c_int]),
    "orc_mt_raw": (None, [C.c_uint32, C.c_int, _up]),
    "orc_mt_rand_int": (None, [C.c_uint32, C.c_int, C.c_uint32, _up]),
    "orc_philox": (None, [_u64p, _u64p, _u64p]),
    "orc_u01": (C.c_double, [C.c_uint64]),
    "orc_uneg11": (C.c_double, [C.c_uint64]),
    "orc_r123": (None, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, _dp]),
    "orc_rotation": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
    "orc_uniform_rotation": (None, [C.c_double, C.c_double, C.c_double, _dp]),
    "orc_pair_terms": (None, [C.POINTER(SystemDesc), C.c_int, C.c_int, C.c_double, C.c_double, _dp]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        for n, (r, a) in _SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as({"float64": _dp, "int32": _ip, "int64": _lp, "uint32": _up,
                             "uint64": _u64p}[a.dtype.name])


class Oracle:
    """One box in the CPU oracle (GOMC arithmetic, OpenMP threads)."""

    def __init__(self, system: System, threads: int = 1):
        self.L = lib()
        system.threads = threads
        self.system = system
        self._desc = system.desc()
        self.h = self.L.orc_create(C.byref(self._desc))
        self.natoms = self.L.orc_natoms(self.h)
        self.nmols = self.L.orc_nmols(self.h)
        self.nkvec = self.L.orc_nkvec(self.h)
        self.mol_start = system.mol_start()

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cutoffs(self):
        o = np.zeros(6)
        self.L.orc_cutoffs(self.h, _p(o))
        return dict(zip(("alpha", "rcut_box", "rcut_low", "recip_rcut", "volume", "beta"), o))

    def system_total(self):
        o = np.zeros(7)
        self.L.orc_system_total(self.h, _p(o))
        return o

    def energies(self):
        o = np.zeros(7)
        self.L.orc_energies(self.h, _p(o))
        return o

    def box_inter(self):
        a, b = C.c_double(), C.c_double()
        self.L.orc_box_inter(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def box_inter_brute(self):
        a, b = C.c_double(), C.c_double()
        self.L.orc_box_inter_brute(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    @staticmethod
    def _xyz3(xyz):
        xyz = np.asarray(xyz, dtype=np.float64)
        return tuple(np.ascontiguousarray(xyz[:, i]) for i in range(3))

    def mol_inter(self, m, new_xyz):
        x, y, z = self._xyz3(new_xyz)
        a, b = C.c_double(), C.c_double()
        ov = self.L.orc_mol_inter(self.h, m, _p(x), _p(y), _p(z), C.byref(a), C.byref(b))
        return a.value, b.value, ov

    def mol_recip(self, m, new_xyz, want_sums=False):
        x, y, z = self._xyz3(new_xyz)
        sr = np.zeros(self.nkvec) if want_sums else None
        si = np.zeros(self.nkvec) if want_sums else None
        d = self.L.orc_mol_recip(self.h, m, _p(x), _p(y), _p(z), _p(sr), _p(si))
        return (d, sr, si) if want_sums else d

    def swap_recip(self, kind, insert, xyz):
        x, y, z = self._xyz3(xyz)
        return self.L.orc_swap_recip(self.h, kind, int(insert), _p(x), _p(y), _p(z))

    def particle_inter(self, atom_type, charge, trials):
        x, y, z = self._xyz3(trials)
        n = len(x)
        en, real, ov = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
        self.L.orc_particle_inter(self.h, atom_type, charge, n, _p(x), _p(y), _p(z), _p(en),
                                  _p(real), _p(ov))
        return en, real, ov

    def kvectors(self):
        a = [np.zeros(self.nkvec) for _ in range(4)]
        self.L.orc_kvectors(self.h, *(_p(v) for v in a))
        return dict(zip(("kx", "ky", "kz", "prefact"), a))

    def recip_sums(self):
        a, b = np.zeros(self.nkvec), np.zeros(self.nkvec)
        self.L.orc_recip_sums(self.h, _p(a), _p(b))
        return a, b

    def self_corr_lrc(self):
        o = np.zeros(3)
        self.L.orc_self_corr_lrc(self.h, _p(o))
        return o

    def tail_change(self, kind, add):
        return self.L.orc_tail_change(self.h, kind, int(add))

    def coords(self):
        x, y, z = (np.zeros(self.natoms) for _ in range(3))
        self.L.orc_get_coords(self.h, _p(x), _p(y), _p(z))
        return np.stack([x, y, z], axis=1)

    def com(self):
        x, y, z = (np.zeros(self.nmols) for _ in range(3))
        self.L.orc_get_com(self.h, _p(x), _p(y), _p(z))
        return np.stack([x, y, z], axis=1)

    def scales(self):
        nk = len(self.system.kinds)
        d, r = np.zeros(nk), np.zeros(nk)
        self.L.orc_get_scales(self.h, _p(d), _p(r))
        return d, r

    def set_scales(self, disp, rot, mp_t_max, mp_r_max):
        """Fixed move scales per kind and MultiParticle step sizes."""
        d = np.ascontiguousarray(disp, dtype=np.float64)
        r = np.ascontiguousarray(rot, dtype=np.float64)
        self.L.orc_set_scales(self.h, _p(d), _p(r), float(mp_t_max), float(mp_r_max))

    def cell_csr(self):
        edge = np.zeros(3, np.int32)
        ncell = self.L.orc_cell_csr(self.h, None, None, None, _p(edge))
        cv = np.zeros(self.natoms, np.int32)
        cs = np.zeros(ncell + 1, np.int32)
        mp = np.zeros(self.natoms, np.int32)
        self.L.orc_cell_csr(self.h, _p(cv), _p(cs), _p(mp), _p(edge))
        return cv, cs, mp, tuple(int(e) for e in edge)

    def cells_exhaustive(self):
        return bool(self.L.orc_cells_exhaustive(self.h))

    def run(self, nsteps, start_step=0, mode=0, record=True):
        if not record:
            return self.L.orc_run(self.h, start_step, nsteps, mode, *([None] * 7)), None
        rec = dict(kind=np.zeros(nsteps, np.int32), mol=np.zeros(nsteps, np.int32),
                   acc=np.zeros(nsteps, np.int32), ovl=np.zeros(nsteps, np.int32),
                   dlj=np.zeros(nsteps), dreal=np.zeros(nsteps), drecip=np.zeros(nsteps))
        n = self.L.orc_run(self.h, start_step, nsteps, mode, _p(rec["kind"]), _p(rec["mol"]),
                           _p(rec["acc"]), _p(rec["ovl"]), _p(rec["dlj"]), _p(rec["dreal"]),
                           _p(rec["drecip"]))
        return n, rec

    def box_force(self):
        lj, real = C.c_double(), C.c_double()
        a = [np.zeros(self.natoms) for _ in range(3)] + [np.zeros(self.nmols) for _ in range(3)]
        self.L.orc_box_force(self.h, C.byref(lj), C.byref(real), *(_p(v) for v in a))
        return lj.value, real.value, np.stack(a[:3], 1), np.stack(a[3:], 1)

    def box_force_recip(self):
        a = [np.zeros(self.natoms) for _ in range(3)] + [np.zeros(self.nmols) for _ in range(3)]
        self.L.orc_box_force_recip(self.h, *(_p(v) for v in a))
        return np.stack(a[:3], 1), np.stack(a[3:], 1)

    def torque(self):
        t = [np.zeros(self.nmols) for _ in range(3)]
        self.L.orc_torque(self.h, *(_p(v) for v in t))
        return np.stack(t, 1)

    def volume(self):
        out = np.zeros(6)
        self.L.orc_cutoffs(self.h, _p(out))
        return out[4]

    def virial(self):
        """{inter, tail, real, recip, total, pressure (bar), total tensor diagonal}."""
        out = np.zeros(9)
        self.L.orc_virial(self.h, _p(out))
        return dict(inter=out[0], tail=out[1], real=out[2], recip=out[3], total=out[4],
                    pressure=out[5], tensor=out[6:9].copy())

    def mol_boxes(self):
        b = np.zeros(self.nmols, np.int32)
        n = self.L.orc_mol_boxes(self.h, _p(b))
        return n, b

    def mp_settings(self):
        out = np.zeros(2)
        self.L.orc_mp_settings(self.h, _p(out))
        return out


class planted:
    """Context manager that plants mistake `which` in the oracle (1: MP
    weight ratio with new and old forces swapped; 2: CBMC keeps the trial
    after the weighted pick; 3: N instead of N + 1 in the insertion
    coefficient) and removes it on exit."""

    def __init__(self, which: int):
        self.which = which

    def __enter__(self):
        lib().orc_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_plant(0)
        return False


def batch_run(oracles: list[Oracle], moves, threads: int) -> float:
    """Philox-stream moves on many boxes, one box per thread; seconds."""
    hs = (_vp * len(oracles))(*[o.h for o in oracles])
    mv = np.ascontiguousarray(moves, dtype=np.int64)
    return lib().orc_batch_run(len(oracles), hs, _p(mv), threads)


def mt_raw(seed, n):
    o = np.zeros(n, np.uint32)
    lib().orc_mt_raw(seed, n, _p(o))
    return o


def mt_rand_int(seed, n, bound):
    o = np.zeros(n, np.uint32)
    lib().orc_mt_rand_int(seed, n, bound, _p(o))
    return o


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint64)
    k = np.ascontiguousarray(key, np.uint64)
    o = np.zeros(4, np.uint64)
    lib().orc_philox(_p(c), _p(k), _p(o))
    return o


def u01(v):
    return lib().orc_u01(int(v))


def uneg11(v):
    return lib().orc_uneg11(int(v))


def r123(seed, step, counter, bound):
    o = np.zeros(8)
    lib().orc_r123(seed, step, counter, bound, _p(o))
    return o


def rotation(theta, axis):
    o = np.zeros(9)
    lib().orc_rotation(theta, *map(float, axis), _p(o))
    return o.reshape(3, 3)


def uniform_rotation(u1, u2, u3):
    o = np.zeros(9)
    lib().orc_uniform_rotation(u1, u2, u3, _p(o))
    return o.reshape(3, 3)


def pair_terms(system: System, t1, t2, dist_sq, qiqj):
    d = system.desc()
    o = np.zeros(5)
    lib().orc_pair_terms(C.byref(d), t1, t2, dist_sq, qiqj, _p(o))
    return o


def exchange(oracles: list, ladder: np.ndarray, attempt: int, seed: int, draws_before: int):
    """Replica exchange over oracle systems (see orc_exchange); returns
    (accepted per slot, draws made); `ladder` (int32) is updated in place."""
    L = lib()
    hs = (C.c_void_p * len(oracles))(*[o.h for o in oracles])
    acc = np.zeros(len(ladder), np.int32)
    d = L.orc_exchange(hs, _p(ladder), len(ladder), attempt, seed, draws_before, _p(acc))
    return acc, d
