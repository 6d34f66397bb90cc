"""GOMC's on-disk input formats: the control file (in.conf), PDB coordinates,
PSF structure and CHARMM / Mie parameter files, read into `systems.System`,
and the same formats written back out (restart-style PDB/PSF per box, a
Mie parameter file and a control file) so that a system round-trips.

What is read, and where the reference reads it:

* control file: whitespace-separated `Key value ...` lines, `#` comments
  (src/ConfigSetup.cpp:200-1420).  Keys used here: Parameters,
  ParaTypeCHARMM / ParaTypeEXOTIC / ParaTypeMie / ParaTypeMARTINI,
  Coordinates <box>, Structure <box>, Temperature, Rcut, RcutLow,
  RcutCoulomb <box>, LRC, Ewald, ElectroStatic, CachedFourier, Tolerance,
  Potential, Exclude, PRNG / Random_Seed, DisFreq, RotFreq,
  MultiParticleFreq, SwapFreq, VolFreq, AdjSteps, CBMC_First, CBMC_Nth,
  CBMC_Ang, Pressure, Fugacity, ChemPot, CellBasisVector1-3 <box>.  Other
  keys are kept in `GomcConfig.raw` and otherwise ignored (output control).
* PDB: fixed columns as src/PDBConst.h:116-169 lays them out (name 12-16,
  residue name 17-21, chain 21, residue number 22-26, x/y/z 30/38/46 width 8,
  occupancy 54-60, beta 60-66), CRYST1 a/b/c at 6/15/24 width 9
  (src/PDBSetup.cpp:78-135).  The first frame is read (up to END/ENDMDL).
* PSF: the !NATOM section `id segment resid resname name type charge mass`
  (src/MolSetup.cpp:1186-1213, whitespace separated) and the !NBOND pairs;
  molecules are the connected components of the bond graph and kinds are
  keyed by residue name (src/MolSetup.cpp:932-1060).
* parameters: NONBONDED (CHARMM: `type ignored eps Rmin/2`, eps in kcal/mol
  stored negative) and NONBONDED_MIE (`type eps[K] sigma n`); CHARMM values
  become eps[K] = -eps * 503.21959899 and sigma = Rmin/2 * 1.7817974362807,
  n = 12; eps = 0 resets n to 12 (src/FFSetup.cpp:205-279, the two constants
  at src/FFSetup.cpp:25-26).

ChemPot (K) is carried as the fugacity with the same transfer coefficient:
MoleculeTransfer::GetCoeff (src/moves/MoleculeTransfer.h:138-166) uses
beta * f for a fugacity f and exp(beta * mu) for a chemical potential, so
f = T exp(mu / T) K/A^3, divided by BAR_TO_K_MOLECULE_PER_A3 to give bar.

Limits (raised as errors, not approximated): orthogonal cells only,
`Potential VDW` only, no NBFIX pair overrides among the system's types,
molecules of at most the sizes the energy path takes, GEMC not built, no
nonzero frequency of a move that is not built (regrowth, intra-swap,
crankshaft, MEMC, targeted swap, NeMTMC, Brownian MultiParticle).  The move
frequencies the ensemble reads must sum to one within 0.001, as the
reference requires.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from .systems import MolKind, System

KCAL_PER_MOL_TO_K = 503.21959899  # src/FFSetup.cpp:25
RIJ_OVER_2_TO_SIG = 1.7817974362807  # src/FFSetup.cpp:26
BAR_TO_K_MOLECULE_PER_A3 = 0.007242971  # src/UnitConst.h:44

ENSEMBLES = {"NVT": 0, "GCMC": 1, "NPT": 2}


class GomcFormatError(ValueError):
    """A GOMC input file that cannot be read, or asks for what is not built."""


# ------------------------------------------------------------ control file --
def _bool(tok: str) -> bool:
    t = tok.lower()
    if t in ("true", "on", "yes", "1"):
        return True
    if t in ("false", "off", "no", "0"):
        return False
    raise GomcFormatError(f"not a boolean: {tok!r}")


@dataclass
class GomcConfig:
    """The settings of one GOMC control file (paths resolved against the
    file's directory)."""
    path: str
    parameters: list[str] = field(default_factory=list)
    charmm: bool = True
    coordinates: dict[int, str] = field(default_factory=dict)
    structure: dict[int, str] = field(default_factory=dict)
    cell: dict[int, np.ndarray] = field(default_factory=dict)  # box -> 3x3 rows
    temperature: float | None = None
    rcut: float | None = None
    rcut_low: float = 0.0
    rcut_coulomb: dict[int, float] = field(default_factory=dict)
    lrc: bool = True  # ConfigSetup.cpp:73, doTailCorr defaults on
    ewald: bool = False
    electrostatic: bool = False
    cached_fourier: bool = False
    tolerance: float | None = None
    potential: str = "VDW"
    exclude: str = "1-3"
    seed: int | None = None
    freq: dict[str, float] = field(default_factory=dict)
    adj_steps: int | None = None
    cbmc: dict[str, int] = field(default_factory=dict)
    pressure: float | None = None  # bar
    fugacity: dict[str, float] = field(default_factory=dict)  # bar
    chem_pot: dict[str, float] = field(default_factory=dict)  # K
    raw: list[list[str]] = field(default_factory=list)


_FREQ_KEYS = {"DisFreq": "disp", "RotFreq": "rot", "MultiParticleFreq": "mp",
              "SwapFreq": "swap", "VolFreq": "vol"}
# Move frequencies the reference reads (src/ConfigSetup.cpp:785-935) for
# moves this build does not run; a nonzero one is refused.  The MEMC keys
# share one frequency there (sys.moves.memc / intraMemc).
_UNBUILT_FREQ_KEYS = {
    "MultiParticleBrownianFreq": "mp_brownian", "IntraSwapFreq": "intra_swap",
    "RegrowthFreq": "regrowth", "CrankShaftFreq": "crankshaft",
    "IntraTargetedSwapFreq": "intra_targeted_swap",
    "IntraMEMC-1Freq": "intra_memc", "IntraMEMC-2Freq": "intra_memc",
    "IntraMEMC-3Freq": "intra_memc", "MEMC-1Freq": "memc", "MEMC-2Freq": "memc",
    "MEMC-3Freq": "memc", "MEMC-2-LiqFreq": "memc", "MEMC-3-LiqFreq": "memc",
    "TargetedSwapFreq": "targeted_swap", "NeMTMCFreq": "nemtmc"}
# The frequencies each ensemble's binary reads and sums to one
# (src/ConfigSetup.cpp:2036-2085): VolFreq is read only where the volume
# varies and SwapFreq, MEMC, targeted swap and NeMTMC only in GCMC/GEMC.
_ENSEMBLE_FREQS = {
    "NVT": ("disp", "rot", "mp", "mp_brownian", "intra_swap", "regrowth", "crankshaft",
            "intra_targeted_swap", "intra_memc"),
    "NPT": ("disp", "rot", "mp", "mp_brownian", "intra_swap", "regrowth", "crankshaft",
            "intra_targeted_swap", "intra_memc", "vol"),
    "GCMC": ("disp", "rot", "mp", "mp_brownian", "intra_swap", "regrowth", "crankshaft",
             "intra_targeted_swap", "intra_memc", "swap", "memc", "targeted_swap", "nemtmc"),
}
_CBMC_KEYS = {"CBMC_First": "first", "CBMC_Nth": "nth", "CBMC_Ang": "ang", "CBMC_Dih": "dih"}


_KNOWN_KEYS = ["Parameters", "ParaTypeCHARMM", "ParaTypeEXOTIC", "ParaTypeMie",
               "ParaTypeMARTINI", "Coordinates", "Structure", "CellBasisVector1",
               "CellBasisVector2", "CellBasisVector3", "Temperature", "Rcut", "RcutLow",
               "RcutCoulomb", "LRC", "Ewald", "ElectroStatic", "CachedFourier", "Tolerance",
               "Potential", "Exclude", "Random_Seed", "AdjSteps", "Pressure", "Fugacity",
               "ChemPot", *_FREQ_KEYS, *_UNBUILT_FREQ_KEYS, *_CBMC_KEYS]
_CANON = {k.upper(): k for k in _KNOWN_KEYS}


def read_conf(path: str) -> GomcConfig:
    """Parse a GOMC control file.  Keys are matched without regard to case,
    as the reference's CheckString does (src/ConfigSetup.cpp:194-210); a key
    GOMC requires that is missing is reported when a system is built from
    the file."""
    base = os.path.dirname(os.path.abspath(path))
    c = GomcConfig(path=os.path.abspath(path))
    here = lambda p: os.path.normpath(os.path.join(base, p))  # noqa: E731
    with open(path) as f:
        for line in f:
            tok = line.split("#", 1)[0].split()
            if not tok:
                continue
            k, v = _CANON.get(tok[0].upper(), tok[0]), tok[1:]
            c.raw.append(tok)
            try:
                if k == "Parameters":
                    c.parameters.append(here(v[0]))
                elif k == "ParaTypeCHARMM":
                    if _bool(v[0]):
                        c.charmm = True
                elif k in ("ParaTypeEXOTIC", "ParaTypeMie", "ParaTypeMARTINI"):
                    if _bool(v[0]):
                        c.charmm = False
                elif k == "Coordinates":
                    c.coordinates[int(v[0])] = here(v[1])
                elif k == "Structure":
                    c.structure[int(v[0])] = here(v[1])
                elif k in ("CellBasisVector1", "CellBasisVector2", "CellBasisVector3"):
                    b = int(v[0])
                    m = c.cell.setdefault(b, np.zeros((3, 3)))
                    m[int(k[-1]) - 1] = [float(x) for x in v[1:4]]
                elif k == "Temperature":
                    c.temperature = float(v[0])
                elif k == "Rcut":
                    c.rcut = float(v[0])
                elif k == "RcutLow":
                    c.rcut_low = max(float(v[0]), 0.0)
                elif k == "RcutCoulomb" and len(v) == 2:
                    c.rcut_coulomb[int(v[0])] = float(v[1])
                elif k == "LRC":
                    c.lrc = _bool(v[0])
                elif k == "Ewald":
                    c.ewald = _bool(v[0])
                elif k == "ElectroStatic":
                    c.electrostatic = _bool(v[0])
                elif k == "CachedFourier":
                    c.cached_fourier = _bool(v[0])
                elif k == "Tolerance":
                    c.tolerance = float(v[0])
                elif k == "Potential":
                    c.potential = v[0].upper()
                elif k == "Exclude":
                    c.exclude = v[0]
                elif k == "Random_Seed":
                    c.seed = int(v[0])
                elif k in _FREQ_KEYS:
                    c.freq[_FREQ_KEYS[k]] = float(v[0])
                elif k in _UNBUILT_FREQ_KEYS:
                    # the MEMC keys set their shared frequency only when > 0
                    if not (k.startswith(("MEMC", "IntraMEMC")) and float(v[0]) <= 0.0):
                        c.freq[_UNBUILT_FREQ_KEYS[k]] = float(v[0])
                elif k == "AdjSteps":
                    c.adj_steps = int(v[0])
                elif k in _CBMC_KEYS:
                    c.cbmc[_CBMC_KEYS[k]] = int(v[0])
                elif k == "Pressure":
                    c.pressure = float(v[0])
                elif k == "Fugacity":
                    c.fugacity[v[0]] = float(v[1])
                elif k == "ChemPot":
                    c.chem_pot[v[0]] = float(v[1])
            except (IndexError, ValueError) as e:
                raise GomcFormatError(f"{path}: cannot read line {line.strip()!r}: {e}") from e
    return c


# ------------------------------------------------------------------- PDB --
@dataclass
class PdbFrame:
    name: list[str]
    resname: list[str]
    resid: np.ndarray
    chain: list[str]
    xyz: np.ndarray  # (n, 3)
    occupancy: np.ndarray
    beta: np.ndarray
    cell: tuple[float, float, float] | None = None
    angles: tuple[float, float, float] | None = None


def _col(line: str, start: int, width: int) -> str:
    return line[start:start + width].strip()


def read_pdb(path: str) -> PdbFrame:
    """The first frame of a PDB file: ATOM/HETATM records and CRYST1."""
    name, res, resid, chain, xyz, occ, beta = [], [], [], [], [], [], []
    cell = angles = None
    with open(path) as f:
        for line in f:
            rec = line[:6].strip()
            if rec in ("END", "ENDMDL"):
                if xyz:
                    break
                continue
            if rec == "CRYST1":
                cell = tuple(float(_col(line, s, 9)) for s in (6, 15, 24))
                a = [_col(line, s, 7) for s in (33, 40, 47)]
                angles = tuple(float(x) for x in a) if all(a) else (90.0, 90.0, 90.0)
            elif rec in ("ATOM", "HETATM"):
                try:
                    xyz.append([float(_col(line, s, 8)) for s in (30, 38, 46)])
                except ValueError as e:
                    raise GomcFormatError(f"{path}: bad coordinates in {line.rstrip()!r}") from e
                name.append(_col(line, 12, 4))
                res.append(_col(line, 17, 4))
                chain.append(line[21:22])
                r = _col(line, 22, 4)
                resid.append(int(r) if r else 0)
                o, bt = _col(line, 54, 6), _col(line, 60, 6)
                occ.append(float(o) if o else 0.0)
                beta.append(float(bt) if bt else 0.0)
    return PdbFrame(name, res, np.array(resid, np.int64), chain,
                    np.array(xyz, np.float64).reshape(-1, 3), np.array(occ), np.array(beta),
                    cell, angles)


def write_pdb(path: str, xyz: np.ndarray, names: list[str], resnames: list[str],
              resids: np.ndarray, cell: tuple[float, float, float] | None = None,
              box_index: int = 0) -> None:
    """ATOM records in the reference's columns; the occupancy column holds the
    box index as GOMC's restart PDB does (src/PDBOutput.cpp)."""
    with open(path, "w") as f:
        if cell is not None:
            f.write("CRYST1%9.3f%9.3f%9.3f%7.2f%7.2f%7.2f P 1           1\n"
                    % (cell[0], cell[1], cell[2], 90.0, 90.0, 90.0))
        for i, (p, nm, rn, ri) in enumerate(zip(xyz, names, resnames, resids)):
            f.write("ATOM  %5d %-4s %-4s%1s%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n"
                    % ((i + 1) % 100000, nm[:4], rn[:4], "A", int(ri) % 10000,
                       p[0], p[1], p[2], float(box_index), 0.0))
        f.write("END\n")


# ------------------------------------------------------------------- PSF --
@dataclass
class PsfStructure:
    segment: list[str]
    resid: np.ndarray
    resname: list[str]
    name: list[str]
    type: list[str]
    charge: np.ndarray
    mass: np.ndarray
    bonds: np.ndarray  # (nbond, 2), 0-based
    # the other topology sections GOMC's kind map keeps (MolSetup.cpp:
    # 1214-1480), 0-based, in file order: !NTHETA, !NPHI, !NIMPHI, !NDON, !NACC
    angles: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int64))
    dihedrals: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int64))
    impropers: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int64))
    donors: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int64))
    acceptors: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int64))


def _psf_section(lines: list[str], tag: str, width: int, path: str) -> np.ndarray:
    """The index tuples of one PSF section ("!NTHETA" ...), 0-based; empty
    when the file has no such section."""
    i = 0
    while i < len(lines) and tag not in lines[i]:
        i += 1
    if i == len(lines):
        return np.zeros((0, width), np.int64)
    n = int(lines[i].split()[0])
    vals: list[int] = []
    i += 1
    while len(vals) < width * n and i < len(lines):
        t = lines[i].split()
        i += 1
        if t and t[0].startswith("!"):
            break
        vals.extend(int(x) - 1 for x in t)
    if len(vals) < width * n:
        raise GomcFormatError(f"{path}: found {len(vals) // width} of {n} entries of {tag}")
    return np.array(vals[:width * n], np.int64).reshape(-1, width)


def read_psf(path: str) -> PsfStructure:
    with open(path) as f:
        lines = f.readlines()
    i = 0
    while i < len(lines) and "!NATOM" not in lines[i]:
        i += 1
    if i == len(lines):
        raise GomcFormatError(f"{path}: no !NATOM section")
    natom = int(lines[i].split()[0])
    seg, resid, res, name, typ, q, m = [], [], [], [], [], [], []
    i += 1
    while len(seg) < natom:
        if i >= len(lines):
            raise GomcFormatError(f"{path}: found {len(seg)} of {natom} atoms")
        t = lines[i].split()
        i += 1
        if not t or t[0].startswith("!"):
            continue
        seg.append(t[1])
        resid.append(int(t[2]))
        res.append(t[3])
        name.append(t[4])
        typ.append(t[5])
        q.append(float(t[6]))
        m.append(float(t[7]))
    bonds: list[int] = []
    nb = 0
    while i < len(lines) and "!NBOND" not in lines[i]:
        i += 1
    if i < len(lines):
        nb = int(lines[i].split()[0])
        i += 1
        while len(bonds) < 2 * nb and i < len(lines):
            t = lines[i].split()
            i += 1
            if t and t[0].startswith("!"):
                break
            bonds.extend(int(x) - 1 for x in t)
        if len(bonds) < 2 * nb:
            raise GomcFormatError(f"{path}: found {len(bonds) // 2} of {nb} bonds")
    return PsfStructure(seg, np.array(resid, np.int64), res, name, typ, np.array(q),
                        np.array(m), np.array(bonds[:2 * nb] if bonds else [],
                                              np.int64).reshape(-1, 2),
                        _psf_section(lines, "!NTHETA", 3, path),
                        _psf_section(lines, "!NPHI", 4, path),
                        _psf_section(lines, "!NIMPHI", 4, path),
                        _psf_section(lines, "!NDON", 2, path),
                        _psf_section(lines, "!NACC", 2, path))


def write_psf(path: str, segment: list[str], resid, resname, name, typ, charge, mass,
              bonds: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("PSF\n\n       1 !NTITLE\n REMARKS written by gomc_amd\n\n")
        f.write("%8d !NATOM\n" % len(name))
        for i in range(len(name)):
            f.write("%8d %-4s %-4d %-4s %-4s %-4s %14.10f %13.4f           0\n"
                    % (i + 1, segment[i], int(resid[i]), resname[i], name[i], typ[i],
                       float(charge[i]), float(mass[i])))
        b = np.asarray(bonds, np.int64).reshape(-1, 2) + 1
        f.write("\n%8d !NBOND: bonds\n" % len(b))
        for j in range(0, len(b), 4):
            f.write("".join("%8d%8d" % (p[0], p[1]) for p in b[j:j + 4]) + "\n")
        f.write("\n%8d !NTHETA: angles\n\n%8d !NPHI: dihedrals\n\n" % (0, 0))


# ------------------------------------------------------------ parameters --
@dataclass
class NonBonded:
    epsilon: float  # K
    sigma: float  # A
    n: float


_SECTIONS = {"BONDS", "BOND", "ANGLES", "ANGLE", "THETA", "DIHEDRALS", "DIHEDRAL", "PHI",
             "IMPROPER", "IMPROPERS", "IMPHI", "NONBONDED", "NONBONDED_MIE", "NBFIX",
             "NBFIX_MIE", "CMAP", "HBOND", "END", "ATOMS", "RESI", "MASS"}


def read_parameters(paths: list[str] | str, charmm: bool) -> dict[str, NonBonded]:
    """Non-bonded parameters per atom type, in K and A.  A later file (or
    line) for the same type replaces an earlier one."""
    if isinstance(paths, str):
        paths = [paths]
    want = "NONBONDED" if charmm else "NONBONDED_MIE"
    out: dict[str, NonBonded] = {}
    for path in paths:
        section = None
        with open(path) as f:
            for line in f:
                s = line.split("!", 1)[0].strip()
                if not s or s.startswith("*"):
                    continue
                t = s.split()
                head = t[0].upper()
                if head in _SECTIONS:
                    if charmm and head.endswith("_MIE"):
                        raise GomcFormatError(f"{path}: {head} in a CHARMM-style file")
                    if not charmm and head in ("NONBONDED", "NBFIX"):
                        raise GomcFormatError(f"{path}: {head} in a Mie-style file")
                    section = head
                    continue
                if section != want:
                    continue
                try:
                    if charmm:
                        e = -float(t[2]) * KCAL_PER_MOL_TO_K
                        sig = float(t[3]) * RIJ_OVER_2_TO_SIG
                        n = 12.0
                    else:
                        e, sig, n = float(t[1]), float(t[2]), float(t[3])
                except (IndexError, ValueError):
                    # continuation of the NONBONDED header options (cutnb ...)
                    continue
                if abs(e) < 1e-20:
                    e, n = 0.0, 12.0
                if n - 6.0 < 1e-20:
                    raise GomcFormatError(f"{path}: Mie exponent of {t[0]} must be > 6")
                out[t[0]] = NonBonded(e, sig, n)
    return out


def read_nbfix_pairs(paths: list[str] | str, charmm: bool) -> list[tuple[str, str]]:
    """Atom-type pairs that an NBFIX (CHARMM) or NBFIX_MIE section overrides
    (src/FFSetup.cpp NBfix::Read).  The energy path blends every pair from
    the per-type parameters, so a system that uses such a pair is refused."""
    if isinstance(paths, str):
        paths = [paths]
    want = "NBFIX" if charmm else "NBFIX_MIE"
    pairs = []
    for path in paths:
        section = None
        with open(path) as f:
            for line in f:
                s = line.split("!", 1)[0].strip()
                if not s or s.startswith("*"):
                    continue
                t = s.split()
                if t[0].upper() in _SECTIONS:
                    section = t[0].upper()
                    continue
                if section == want and len(t) >= 4:
                    pairs.append((t[0], t[1]))
    return pairs


def write_parameters(path: str, params: dict[str, NonBonded]) -> None:
    with open(path, "w") as f:
        f.write("* written by gomc_amd\n\nNONBONDED_MIE\n!atom eps(K) sigma(A) n\n")
        for name, p in params.items():
            f.write(f"{name} {p.epsilon!r} {p.sigma!r} {p.n!r}\n")
        f.write("\nEND\n")


# ---------------------------------------------------------------- system --
def _molecules(natom: int, bonds: np.ndarray) -> list[list[int]]:
    """Connected components of the bond graph, each sorted, ordered by their
    first atom."""
    parent = list(range(natom))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for a, b in bonds:
        if not (0 <= a < natom and 0 <= b < natom):
            raise GomcFormatError(f"bond {a + 1}-{b + 1} names an atom outside 1..{natom}")
        ra, rb = find(int(a)), find(int(b))
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    groups: dict[int, list[int]] = {}
    for a in range(natom):
        groups.setdefault(find(a), []).append(a)
    return sorted(groups.values(), key=lambda g: g[0])


def _wrap(xyz: np.ndarray, L: np.ndarray) -> np.ndarray:
    w = xyz - L * np.floor(xyz / L)
    return np.where(w >= L, w - L, w)


def _read_box(conf: GomcConfig, b: int):
    if b not in conf.structure or b not in conf.coordinates:
        raise GomcFormatError(f"{conf.path}: no Structure/Coordinates for box {b}")
    psf = read_psf(conf.structure[b])
    pdb = read_pdb(conf.coordinates[b])
    if len(pdb.xyz) != len(psf.name):
        raise GomcFormatError(f"box {b}: {len(pdb.xyz)} PDB atoms but {len(psf.name)} PSF atoms")
    return psf, pdb


def _box_lengths(conf: GomcConfig, b: int, pdb: PdbFrame) -> tuple[float, float, float]:
    if b in conf.cell:
        m = conf.cell[b]
        if np.count_nonzero(m - np.diag(np.diag(m))):
            raise GomcFormatError(f"box {b}: only orthogonal cells are built")
        return tuple(float(x) for x in np.diag(m))
    if pdb.cell is not None:
        if any(abs(a - 90.0) > 1e-9 for a in pdb.angles):
            raise GomcFormatError(f"box {b}: only orthogonal cells are built")
        return pdb.cell
    raise GomcFormatError(f"box {b}: no CellBasisVector and no CRYST1 record")


def system_from_gomc(conf_path: str, ensemble: str | None = None, threads: int = 1) -> System:
    """Build a `systems.System` from a GOMC control file and the PDB, PSF and
    parameter files it names.

    `ensemble` is what GOMC fixes when the binary is built (GOMC_CPU_NVT,
    _GCMC, _NPT): "NVT", "GCMC" or "NPT".  Left as None it is GCMC when the
    file gives ChemPot or Fugacity, NPT when it gives Pressure and a nonzero
    VolFreq, else NVT.  NVT and NPT use box 0; GCMC uses box 0 as the
    simulation box and box 1 as the reservoir, as GOMC's GCMC does."""
    conf = read_conf(conf_path)
    if ensemble is None:
        if conf.chem_pot or conf.fugacity:
            ensemble = "GCMC"
        elif conf.pressure is not None and conf.freq.get("vol", 0.0) > 0.0:
            ensemble = "NPT"
        else:
            ensemble = "NVT"
    if ensemble not in ENSEMBLES:
        raise GomcFormatError(f"ensemble {ensemble!r} is not built (NVT, GCMC, NPT)")
    for key, val in (("Temperature", conf.temperature), ("Rcut", conf.rcut),
                     ("Parameters", conf.parameters), ("AdjSteps", conf.adj_steps)):
        if val is None or val == []:
            raise GomcFormatError(f"{conf.path}: {key} is not specified")
    if conf.potential != "VDW":
        raise GomcFormatError(f"Potential {conf.potential} is not built (VDW only)")
    if conf.ewald and conf.tolerance is None:
        raise GomcFormatError(f"{conf.path}: Tolerance is not specified")
    _check_freqs(conf, ensemble)
    params = read_parameters(conf.parameters, conf.charmm)

    boxes = [0, 1] if ensemble == "GCMC" else [0]
    psfs, pdbs = zip(*(_read_box(conf, b) for b in boxes))
    L = _box_lengths(conf, 0, pdbs[0])

    type_names: list[str] = []
    kinds: list[MolKind] = []
    kind_of: dict[str, int] = {}
    mol_kind, mol_box, xyz = [], [], []
    for bi, (psf, pdb) in enumerate(zip(psfs, pdbs)):
        Lb = np.array(_box_lengths(conf, boxes[bi], pdb))
        for atoms in _molecules(len(psf.name), psf.bonds):
            if atoms != list(range(atoms[0], atoms[0] + len(atoms))):
                raise GomcFormatError("a molecule's atoms must be consecutive in the PSF")
            rn = psf.resname[atoms[0]]
            types = []
            for a in atoms:
                t = psf.type[a]
                if t not in params:
                    raise GomcFormatError(f"atom type {t} has no non-bonded parameters")
                if t not in type_names:
                    type_names.append(t)
                types.append(type_names.index(t))
            q = [float(psf.charge[a]) for a in atoms]
            p = pdb.xyz[atoms]
            if rn not in kind_of:
                d = p - p[0]
                d -= Lb * np.round(d / Lb)
                kind_of[rn] = len(kinds)
                kinds.append(MolKind(rn, types, q, d))
            k = kinds[kind_of[rn]]
            if k.atom_types != types or not np.array_equal(k.charges, q):
                raise GomcFormatError(f"molecules of residue {rn} differ in atoms or charges")
            mol_kind.append(kind_of[rn])
            mol_box.append(bi)
            xyz.append(p)
    for a, b in read_nbfix_pairs(conf.parameters, conf.charmm):
        if a in type_names and b in type_names:
            raise GomcFormatError(f"NBFIX override of the {a}-{b} pair is not built")
    if max(len(k.atom_types) for k in kinds) > 3 and ensemble == "GCMC":
        raise GomcFormatError("transfers are built for molecules of at most 3 atoms")
    if conf.exclude == "1-2" and any(len(k.atom_types) >= 3 for k in kinds):
        raise GomcFormatError("Exclude 1-2 with 3-atom molecules is not built")

    fug = None
    if ensemble == "GCMC":
        fug = []
        for k in kinds:
            if k.name in conf.fugacity:
                fug.append(conf.fugacity[k.name])
            elif k.name in conf.chem_pot:
                T = conf.temperature
                fug.append(T * math.exp(conf.chem_pot[k.name] / T) / BAR_TO_K_MOLECULE_PER_A3)
            else:
                raise GomcFormatError(f"no ChemPot or Fugacity for residue {k.name}")

    # Every atom is wrapped into box 0's [0, L): box-0 atoms as the
    # reference's coordinate set-up wraps them; a reservoir molecule's
    # position carries only its shape (an insertion regrows it), and its
    # intramolecular vectors survive the wrap under the minimum image.
    xyz_all = _wrap(np.concatenate(xyz) if xyz else np.zeros((0, 3)), np.array(L))
    f = conf.freq
    # ConfigSetup::fillDefaults (src/ConfigSetup.cpp:1446-1449): Ewald turns
    # the electrostatics on whatever ElectroStatic says
    electro = conf.electrostatic or conf.ewald
    elec_ewald = conf.ewald
    rc = conf.rcut
    return System(
        box=L, temperature=conf.temperature, rcut=rc, rcut_low=conf.rcut_low,
        rcut_coulomb=conf.rcut_coulomb.get(0, rc),
        tolerance=conf.tolerance if conf.tolerance is not None else 1e-5,
        sigma=[params[t].sigma for t in type_names],
        epsilon=[params[t].epsilon for t in type_names],
        n_rep=[params[t].n for t in type_names],
        kinds=kinds, mol_kind=np.array(mol_kind, np.int32), xyz=xyz_all,
        seed=conf.seed if conf.seed is not None else 12345,
        move_disp=f.get("disp", 0.0), move_rot=f.get("rot", 0.0), move_mp=f.get("mp", 0.0),
        adj_steps=conf.adj_steps, ewald=elec_ewald, electrostatic=electro, lrc=conf.lrc,
        threads=threads, ensemble=ENSEMBLES[ensemble],
        mol_box=np.array(mol_box, np.int32) if ensemble == "GCMC" else None,
        fugacity=fug, move_swap=f.get("swap", 0.0) if ensemble == "GCMC" else 0.0,
        cbmc_first=conf.cbmc.get("first", 16), cbmc_nth=conf.cbmc.get("nth", 8),
        cbmc_ang=conf.cbmc.get("ang", 100),
        pressure=conf.pressure if conf.pressure is not None else 1.0,
        move_vol=f.get("vol", 0.0) if ensemble == "NPT" else 0.0)


def _check_freqs(conf: GomcConfig, ensemble: str) -> None:
    """The reference's move-frequency checks for `ensemble`
    (src/ConfigSetup.cpp:1975-2085): DisFreq must be given (and SwapFreq in
    GCMC, VolFreq in NPT), and the frequencies the ensemble reads must sum
    to one within 0.001.  A nonzero frequency of a move this build does not
    run is refused rather than dropped, which would change the move mix."""
    f = conf.freq
    need = {"NVT": ("disp",), "NPT": ("disp", "vol"), "GCMC": ("disp", "swap")}[ensemble]
    names = {v: k for k, v in {**_FREQ_KEYS, **_UNBUILT_FREQ_KEYS}.items()}
    for key in need:
        if key not in f:
            raise GomcFormatError(f"{conf.path}: {names[key]} is not specified")
    read = _ENSEMBLE_FREQS[ensemble]
    for key in set(_UNBUILT_FREQ_KEYS.values()) & set(read):
        if f.get(key, 0.0) != 0.0:
            raise GomcFormatError(f"{conf.path}: {names[key]} {f[key]} asks for a move that is "
                                  "not built (Translate, Rotate, MultiParticle, "
                                  "MoleculeTransfer and VolumeTransfer are)")
    total = sum(f.get(key, 0.0) for key in read)
    if abs(total - 1.0) > 0.001:
        raise GomcFormatError(f"{conf.path}: the sum of move frequencies is {total:g}, not one")


def gomc_box_records(s: System, type_names: list[str] | None = None
                     ) -> tuple[list[PsfStructure], list[PdbFrame]]:
    """Per box (box 1 holds a GCMC system's reservoir), the PSF structure and
    the PDB frame that `write_gomc` writes for `s`, as GOMC reads them back
    (coordinates rounded to the PDB's three decimals).  Atom masses are 1.0 (a
    System carries none; the energies do not use them) and each molecule's
    atoms are bonded as a chain so that it stays one connected component."""
    nt = len(s.sigma)
    names = type_names or [f"T{i}" for i in range(nt)]
    starts = s.mol_start()
    mb = s.mol_box if s.mol_box is not None else np.zeros(s.nmols, np.int32)
    nbox = 2 if s.ensemble == 1 else 1
    psfs, pdbs = [], []
    for b in range(nbox):
        seg, resid, res, nm, typ, q, mass, xyz, bonds = [], [], [], [], [], [], [], [], []
        r = 0
        for m in range(s.nmols):
            if mb[m] != b:
                continue
            k = s.kinds[s.mol_kind[m]]
            r += 1
            first = len(nm)
            for j, t in enumerate(k.atom_types):
                seg.append(k.name[:4])
                resid.append(r)
                res.append(k.name[:4])
                nm.append(f"{names[t][:3]}{j}")
                typ.append(names[t])
                q.append(k.charges[j])
                mass.append(1.0)
                xyz.append(s.xyz[starts[m] + j])
            bonds.extend([first + j, first + j + 1] for j in range(len(k.atom_types) - 1))
        xyz = np.array([[float("%.3f" % v) for v in p] for p in xyz]).reshape(-1, 3)
        n = len(nm)
        psfs.append(PsfStructure(seg, np.array(resid, np.int64), res, nm, typ, np.array(q),
                                 np.array(mass), np.array(bonds, np.int64).reshape(-1, 2)))
        pdbs.append(PdbFrame(nm, res, np.array(resid, np.int64), ["A"] * n, xyz,
                             np.full(n, float(b)), np.zeros(n), tuple(s.box),
                             (90.0, 90.0, 90.0)))
    return psfs, pdbs


def write_gomc(outdir: str, s: System, type_names: list[str] | None = None,
               prefix: str = "system") -> str:
    """Write `s` as GOMC input files (control file, PDB and PSF per box, a
    Mie-style parameter file) under `outdir`; returns the control file's path.
    The boxes are `gomc_box_records(s, type_names)`."""
    os.makedirs(outdir, exist_ok=True)
    nt = len(s.sigma)
    names = type_names or [f"T{i}" for i in range(nt)]
    write_parameters(os.path.join(outdir, f"{prefix}.par"),
                     {names[i]: NonBonded(float(s.epsilon[i]), float(s.sigma[i]), float(s.n_rep[i]))
                      for i in range(nt)})
    nbox = 2 if s.ensemble == 1 else 1
    L = s.box
    starts = s.mol_start()
    mb = s.mol_box if s.mol_box is not None else np.zeros(s.nmols, np.int32)
    psfs, _ = gomc_box_records(s, names)
    for b in range(nbox):
        ps = psfs[b]
        xyz = np.array([s.xyz[starts[m] + j] for m in range(s.nmols) if mb[m] == b
                        for j in range(len(s.kinds[s.mol_kind[m]].atom_types))]).reshape(-1, 3)
        write_pdb(os.path.join(outdir, f"{prefix}_BOX_{b}.pdb"), xyz, ps.name, ps.resname,
                  ps.resid, cell=L, box_index=b)
        write_psf(os.path.join(outdir, f"{prefix}_BOX_{b}.psf"), ps.segment, ps.resid,
                  ps.resname, ps.name, ps.type, ps.charge, ps.mass, ps.bonds)
    conf = os.path.join(outdir, "in.conf")
    tot = s.move_disp + s.move_rot + s.move_mp + s.move_swap + s.move_vol
    with open(conf, "w") as f:
        w = lambda *a: f.write(" ".join(str(x) for x in a) + "\n")  # noqa: E731
        w("ParaTypeCHARMM", "off")
        w("ParaTypeMie", "on")
        w("Parameters", f"{prefix}.par")
        for b in range(nbox):
            w("Coordinates", b, f"{prefix}_BOX_{b}.pdb")
            w("Structure", b, f"{prefix}_BOX_{b}.psf")
            for i in range(3):
                row = [0.0, 0.0, 0.0]
                row[i] = L[i]
                w(f"CellBasisVector{i + 1}", b, *(repr(x) for x in row))
        w("Temperature", repr(float(s.temperature)))
        w("Potential", "VDW")
        w("LRC", "true" if s.lrc else "false")
        w("Rcut", repr(float(s.rcut)))
        w("RcutLow", repr(float(s.rcut_low)))
        w("RcutCoulomb", 0, repr(float(s.rcut_coulomb)))
        w("Exclude", "1-3")
        w("Ewald", "true" if s.ewald else "false")
        w("ElectroStatic", "true" if s.electrostatic else "false")
        w("Tolerance", repr(float(s.tolerance)))
        w("PRNG", "INTSEED")
        w("Random_Seed", int(s.seed))
        w("AdjSteps", int(s.adj_steps))
        w("DisFreq", repr(s.move_disp / tot))
        w("RotFreq", repr(s.move_rot / tot))
        w("MultiParticleFreq", repr(s.move_mp / tot))
        if s.ensemble == 1:
            w("SwapFreq", repr(s.move_swap / tot))
            for k, fb in zip(s.kinds, s.fugacity):
                w("Fugacity", k.name[:4], repr(float(fb)))
        if s.ensemble == 2:
            w("VolFreq", repr(s.move_vol / tot))
            w("Pressure", repr(float(s.pressure)))
        w("CBMC_First", s.cbmc_first)
        w("CBMC_Nth", s.cbmc_nth)
        w("CBMC_Ang", s.cbmc_ang)
    return conf
