"""Synthetic GOMC systems and the C layout that carries them.

`SystemDesc` is the ctypes mirror of `gomc_system_desc` (include/gomc_hip.h).
The generators build the workloads the benchmark and the tests run:

* rigid 3-site SPC/E-like water (O-H 1.0 A, H-O-H 109.47 deg, LJ on O only,
  sigma 3.166 A, eps/kB 78.2 K, q_O -0.8476 e, q_H +0.4238 e);
* united-atom methane (sigma 3.73 A, eps/kB 148 K) and rigid linear CO2
  (C-O 1.16 A; C: 2.80 A, 27 K, +0.70 e; O: 3.05 A, 79 K, -0.35 e);

placed by a seeded uniform draw with overlap rejection and uniform random
orientations, coordinates wrapped into [0, L).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SystemDesc(C.Structure):
    """ctypes mirror of gomc_system_desc (same layout as orc_system_desc)."""

    _fields_ = [
        ("box", C.c_double * 3),
        ("temperature", C.c_double),
        ("rcut", C.c_double),
        ("rcut_low", C.c_double),
        ("rcut_coulomb", C.c_double),
        ("tolerance", C.c_double),
        ("ewald", C.c_int32),
        ("electrostatic", C.c_int32),
        ("lrc", C.c_int32),
        ("threads", C.c_int32),
        ("n_atom_types", C.c_int32),
        ("n_kinds", C.c_int32),
        ("sigma", _dp),
        ("epsilon", _dp),
        ("n_rep", _dp),
        ("kind_natoms", _ip),
        ("kind_atom_type", _ip),
        ("kind_charge", _dp),
        ("n_mols", C.c_int32),
        ("_pad0", C.c_int32),
        ("mol_kind", _ip),
        ("x", _dp),
        ("y", _dp),
        ("z", _dp),
        ("seed", C.c_uint64),
        ("move_disp", C.c_double),
        ("move_rot", C.c_double),
        ("move_mp", C.c_double),
        ("adj_steps", C.c_int64),
        ("ensemble", C.c_int32),
        ("cbmc_first", C.c_int32),
        ("cbmc_nth", C.c_int32),
        ("cbmc_ang", C.c_int32),
        ("mol_box", _ip),
        ("kind_fugacity", _dp),
        ("move_swap", C.c_double),
        ("pressure", C.c_double),
        ("move_vol", C.c_double),
    ]


@dataclass
class MolKind:
    name: str
    atom_types: list[int]
    charges: list[float]
    geometry: np.ndarray  # (natoms, 3) body-frame coordinates, A


@dataclass
class System:
    """One box: force field, kinds, molecules and coordinates (numpy)."""

    box: tuple[float, float, float]
    temperature: float
    rcut: float
    rcut_low: float
    rcut_coulomb: float
    tolerance: float
    sigma: list[float]
    epsilon: list[float]
    n_rep: list[float]
    kinds: list[MolKind]
    mol_kind: np.ndarray
    xyz: np.ndarray  # (natoms, 3)
    seed: int = 12345
    move_disp: float = 0.5
    move_rot: float = 0.5
    move_mp: float = 0.0
    adj_steps: int = 1000
    ewald: bool = True
    electrostatic: bool = True
    lrc: bool = True
    threads: int = 1
    # GCMC: 1 selects the grand-canonical move set; mol_box 0 (box) / 1
    # (reservoir) per molecule; fugacity per kind in bar.
    ensemble: int = 0
    mol_box: np.ndarray | None = None
    fugacity: list[float] | None = None
    move_swap: float = 0.0
    cbmc_first: int = 16
    cbmc_nth: int = 8
    cbmc_ang: int = 100
    pressure: float = 1.0  # bar (NPT)
    move_vol: float = 0.0  # VolFreq (NPT)
    _keep: list = field(default_factory=list, repr=False)

    @property
    def natoms(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nmols(self) -> int:
        return int(self.mol_kind.shape[0])

    def mol_start(self) -> np.ndarray:
        lens = np.array([len(self.kinds[k].atom_types) for k in self.mol_kind], dtype=np.int64)
        return np.concatenate([[0], np.cumsum(lens)[:-1]])

    def desc(self) -> SystemDesc:
        """Build the C descriptor; the arrays it points to live on `self`."""
        d = SystemDesc()
        d.box[:] = self.box
        d.temperature = self.temperature
        d.rcut, d.rcut_low, d.rcut_coulomb, d.tolerance = (
            self.rcut, self.rcut_low, self.rcut_coulomb, self.tolerance)
        d.ewald, d.electrostatic, d.lrc, d.threads = (
            int(self.ewald), int(self.electrostatic), int(self.lrc), int(self.threads))
        sig = np.ascontiguousarray(self.sigma, dtype=np.float64)
        eps = np.ascontiguousarray(self.epsilon, dtype=np.float64)
        nrep = np.ascontiguousarray(self.n_rep, dtype=np.float64)
        kn = np.ascontiguousarray([len(k.atom_types) for k in self.kinds], dtype=np.int32)
        kt = np.ascontiguousarray([t for k in self.kinds for t in k.atom_types], dtype=np.int32)
        kq = np.ascontiguousarray([q for k in self.kinds for q in k.charges], dtype=np.float64)
        mk = np.ascontiguousarray(self.mol_kind, dtype=np.int32)
        x = np.ascontiguousarray(self.xyz[:, 0], dtype=np.float64)
        y = np.ascontiguousarray(self.xyz[:, 1], dtype=np.float64)
        z = np.ascontiguousarray(self.xyz[:, 2], dtype=np.float64)
        mb = np.ascontiguousarray(self.mol_box if self.mol_box is not None
                                  else np.zeros(len(mk)), dtype=np.int32)
        fug = np.ascontiguousarray(self.fugacity if self.fugacity is not None
                                   else np.zeros(len(self.kinds)), dtype=np.float64)
        self._keep = [sig, eps, nrep, kn, kt, kq, mk, x, y, z, mb, fug]
        d.n_atom_types = len(sig)
        d.n_kinds = len(self.kinds)
        d.sigma, d.epsilon, d.n_rep = (a.ctypes.data_as(_dp) for a in (sig, eps, nrep))
        d.kind_natoms = kn.ctypes.data_as(_ip)
        d.kind_atom_type = kt.ctypes.data_as(_ip)
        d.kind_charge = kq.ctypes.data_as(_dp)
        d.n_mols = len(mk)
        d.mol_kind = mk.ctypes.data_as(_ip)
        d.x, d.y, d.z = (a.ctypes.data_as(_dp) for a in (x, y, z))
        d.seed = int(self.seed)
        d.move_disp, d.move_rot, d.move_mp = self.move_disp, self.move_rot, self.move_mp
        d.adj_steps = int(self.adj_steps)
        d.ensemble = int(self.ensemble)
        d.cbmc_first, d.cbmc_nth, d.cbmc_ang = (int(self.cbmc_first), int(self.cbmc_nth),
                                                int(self.cbmc_ang))
        d.mol_box = mb.ctypes.data_as(_ip)
        d.kind_fugacity = fug.ctypes.data_as(_dp)
        d.move_swap = float(self.move_swap)
        d.pressure = float(self.pressure)
        d.move_vol = float(self.move_vol)
        return d


# ------------------------------------------------------------------ kinds --
def water_kind() -> MolKind:
    th = math.radians(109.47)
    geo = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [math.cos(th), math.sin(th), 0.0]])
    return MolKind("SPCE", [0, 1, 1], [-0.8476, 0.4238, 0.4238], geo)


WATER_TYPES = dict(sigma=[3.166, 0.0], epsilon=[78.2, 0.0], n_rep=[12.0, 12.0])


def methane_kind(type_index: int = 0) -> MolKind:
    return MolKind("CH4", [type_index], [0.0], np.zeros((1, 3)))


def co2_kind(c_type: int, o_type: int) -> MolKind:
    geo = np.array([[0.0, 0.0, 0.0], [1.16, 0.0, 0.0], [-1.16, 0.0, 0.0]])
    return MolKind("CO2", [c_type, o_type, o_type], [0.70, -0.35, -0.35], geo)


MIXTURE_TYPES = dict(sigma=[3.73, 2.80, 3.05], epsilon=[148.0, 27.0, 79.0], n_rep=[12.0] * 3)

WATER_MASS = 18.01528  # g/mol
AVOGADRO = 6.02214076e23


def water_number_density(rho_g_cm3: float = 0.997) -> float:
    """Molecules per A^3."""
    return rho_g_cm3 * AVOGADRO / WATER_MASS * 1e-24


# ------------------------------------------------------------- placement --
def place_molecules(kinds: list[MolKind], mol_kind: np.ndarray, box: float, seed: int,
                    min_heavy: float = 2.5, min_any: float = 1.5,
                    max_tries: int = 200000) -> np.ndarray:
    """Random sequential insertion (gomc_place_molecules): uniform centre,
    uniform orientation, rejected if a heavy-atom pair (first atom of each
    molecule) is closer than `min_heavy` or any other inter-molecular atom
    pair closer than `min_any`.  Returns wrapped coordinates (natoms, 3)."""
    from . import _lib
    from ._lib import ptr
    lib = _lib.load()
    kn = np.ascontiguousarray([len(k.atom_types) for k in kinds], dtype=np.int32)
    geo = np.ascontiguousarray(np.concatenate([k.geometry for k in kinds]), dtype=np.float64)
    mk = np.ascontiguousarray(mol_kind, dtype=np.int32)
    natoms = int(kn[mk].sum())
    x, y, z = (np.zeros(natoms) for _ in range(3))
    rc = lib.gomc_place_molecules(len(kinds), ptr(kn), ptr(geo), len(mk), ptr(mk), float(box),
                                  int(seed) & 0xFFFFFFFFFFFFFFFF, min_heavy, min_any, max_tries,
                                  ptr(x), ptr(y), ptr(z))
    if rc != 0:
        raise RuntimeError("could not place molecules without overlap; lower the density")
    return np.ascontiguousarray(np.stack([x, y, z], axis=1))


# ------------------------------------------------------------- workloads --
def water_box(nmol: int, box: float | None = None, temperature: float = 298.0,
              rcut: float = 10.0, tolerance: float = 1e-5, seed: int = 12345,
              place_seed: int | None = None, density: float = 0.997, **kw) -> System:
    if box is None:
        box = (nmol / water_number_density(density)) ** (1.0 / 3.0)
    rcut = min(rcut, box / 2.0)
    kinds = [water_kind()]
    mk = np.zeros(nmol, dtype=np.int32)
    xyz = place_molecules(kinds, mk, box, seed if place_seed is None else place_seed)
    return System(box=(box, box, box), temperature=temperature, rcut=rcut, rcut_low=1.0,
                  rcut_coulomb=rcut, tolerance=tolerance, kinds=kinds, mol_kind=mk, xyz=xyz,
                  seed=seed, **WATER_TYPES, **kw)


def config1(seed: int = 12345) -> System:
    """1000 waters, L = 31.07 A, 298 K, rc 10 A + LRC, tolerance 1e-5."""
    return water_box(1000, box=31.07, temperature=298.0, rcut=10.0, tolerance=1e-5, seed=seed)


def config2(seed: int = 12345) -> System:
    """10,000 waters, L = 66.9 A, 298 K, rc 12 A, tolerance 1e-6, MP 1:1."""
    return water_box(10000, box=66.9, temperature=298.0, rcut=12.0, tolerance=1e-6, seed=seed,
                     move_disp=0.25, move_rot=0.25, move_mp=0.5)


def config4(nboxes: int = 256, seed: int = 2024, nmin: int = 200, nmax: int = 4000) -> list[System]:
    """Independent NVT water boxes: N log-uniform in [nmin, nmax] at
    0.997 g/cm^3, T uniform in [280, 360] K, one seed per box.  Boxes with
    L < 20 A use rc = L/2 (GOMC needs L >= 2 rc)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nboxes)).astype(int)
    ts = rng.uniform(280.0, 360.0, size=nboxes)
    boxes = []
    for i, (n, t) in enumerate(zip(ns, ts)):
        boxes.append(water_box(int(n), temperature=float(t), rcut=10.0, tolerance=1e-5,
                               seed=int(seed * 1000 + i), place_seed=int(seed * 7919 + i)))
    return boxes


def mixture_box(n_ch4: int, n_co2: int, box: float, temperature: float = 300.0,
                rcut: float = 10.0, tolerance: float = 1e-5, seed: int = 12345, **kw) -> System:
    kinds = [methane_kind(0), co2_kind(1, 2)]
    mk = np.array([0] * n_ch4 + [1] * n_co2, dtype=np.int32)
    rng = np.random.default_rng(seed)
    rng.shuffle(mk)
    xyz = place_molecules(kinds, mk, box, seed, min_heavy=3.0, min_any=2.0)
    return System(box=(box, box, box), temperature=temperature, rcut=rcut, rcut_low=1.0,
                  rcut_coulomb=rcut, tolerance=tolerance, kinds=kinds, mol_kind=mk, xyz=xyz,
                  seed=seed, **MIXTURE_TYPES, **kw)


BAR_TO_K_PER_A3 = 0.007242971  # UnitConst.h:44


def gcmc_mixture(n_ch4: int, n_co2: int, box: float, temperature: float = 300.0,
                 fugacity_bar: tuple[float, float] = (10.0, 10.0), reservoir: int | None = None,
                 rcut: float = 10.0, tolerance: float = 1e-5, seed: int = 12345,
                 swap: float = 0.4, disp: float = 0.3, rot: float = 0.3, **kw) -> System:
    """Grand-canonical methane / CO2 box: `n_ch4` + `n_co2` molecules placed in
    the box, and a reservoir (GOMC's second GCMC box, no interactions) of
    `reservoir` molecules of each kind to insert from.  Fugacities in bar."""
    kinds = [methane_kind(0), co2_kind(1, 2)]
    rng = np.random.default_rng(seed)
    if reservoir is None:
        reservoir = max(n_ch4, n_co2, 16)
    mk_box = np.array([0] * n_ch4 + [1] * n_co2, dtype=np.int32)
    rng.shuffle(mk_box)
    mk_res = np.array([0] * reservoir + [1] * reservoir, dtype=np.int32)
    rng.shuffle(mk_res)
    parts, boxes = [], []
    if len(mk_box):
        parts.append(place_molecules(kinds, mk_box, box, seed, min_heavy=3.0, min_any=2.0))
        boxes.append(np.zeros(len(mk_box), np.int32))
    # reservoir molecules do not interact: any non-overlapping draw will do
    parts.append(place_molecules(kinds, mk_res, box, seed + 7, min_heavy=0.5, min_any=0.3))
    boxes.append(np.ones(len(mk_res), np.int32))
    xyz = np.concatenate(parts)
    mk = np.concatenate([mk_box, mk_res])
    rc = min(rcut, box / 2.0)
    return System(box=(box, box, box), temperature=temperature, rcut=rc, rcut_low=1.0,
                  rcut_coulomb=rc, tolerance=tolerance, kinds=kinds, mol_kind=mk, xyz=xyz,
                  seed=seed, ensemble=1, mol_box=np.concatenate(boxes),
                  fugacity=list(fugacity_bar), move_swap=swap, move_disp=disp, move_rot=rot,
                  **MIXTURE_TYPES, **kw)


def config3(seed: int = 12345) -> System:
    """GCMC methane + CO2, 40 A box, 300 K, at the densest state the box
    reaches: fugacities 960 / 400 bar hold about 980 molecules (283 CH4,
    693 CO2; device move loop, mean over 5e5 moves after 5e5).  The molecule
    count saturates there (the same ratio at 60 / 25 bar: 241; x8: 920;
    x32: 1023): 2000 molecules in (40 A)^3 would be 0.031 per A^3, about
    twice the number density of liquid methane or CO2 at 300 K, which no
    chemical potential reaches.  Seeded uniformly with 500 + 500 molecules;
    40 % transfers with 16 first-atom trials, 30 % translate, 30 % rotate,
    Ewald tolerance 1e-5."""
    return gcmc_mixture(500, 500, 40.0, temperature=300.0, fugacity_bar=(960.0, 400.0),
                        reservoir=1500, seed=seed, cbmc_first=16, cbmc_nth=8, cbmc_ang=100)

