"""GOMC's checkpoint archive (`<name>_restart.chk`) read and written in
cereal's binary layout, and the restart files that go with it.

The reference writes its checkpoint with cereal 1.3.2's BinaryOutputArchive
(src/CheckpointOutput.cpp:56-90): the Checkpoint object (its serialize()
member order, src/Checkpoint.h:118-167) followed by the MoleculeLookup's
raw arrays.  cereal's binary layout, as this module writes and reads it:

* arithmetic values in native (little-endian) byte order, bool as one byte;
* std::string and every std::vector: a uint64 size, then the elements (a
  vector of arithmetic type as one block of bytes, std::vector<bool> one
  byte per element); a C array of arithmetic type as its bytes, no size;
* std::map: a uint64 size, then key and value of each entry in key order;
* a class whose serialize() takes a version: a uint32 version (0 for every
  GOMC class) the first time that class appears in the archive, none after.

The archive does not hold coordinates: GOMC restarts from it together with
the binary restart coordinates `<name>_BOX_<b>_restart.coor` (int32 atom
count, then x, y, z per atom as doubles; ExtendedSystemOutput.cpp:300-314)
and the cell file `<name>_BOX_<b>_restart.xsc`, whose atoms are listed in
MoleculeLookup order, box by box (ExtendedSystemOutput.cpp:255-287).
`write_restart` writes the three, `read_restart` reads them back and
`restore_box` continues a run from them.

A checkpoint is a dict with the keys of the reference's members (snake
case): gomc_version (5 bytes), step, true_step, save_array (625 words:
MT19937 state and `left`), seed_location, seed_left, seed_value, scale,
accept_percent, accepted, tries, temp_accepted, temp_tries ([box][move]
[kind]), mp_tries, mp_accepted, mp_interval_accepted, mp_interval_tries
([box][type]), mp_t_max, mp_r_max ([box]), is_single_move_accepted,
original_start, restarted_start, lookup (MoleculeLookup: counts, fixed /
swap / move kinds and the raw arrays), mol_setup (kind_map as a list of
[key, MolKind] in key order, mol_vars), pdb_atoms (pdb_setup::Atoms).
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

# move kinds of the reference's ensembles (src/MoveConst.h:21-52)
DISPLACE, ROTATE, MULTIPARTICLE = 0, 1, 2
MOVE_KINDS = {0: 9, 1: 13, 2: 10}       # NVT, GCMC, NPT
MOL_TRANSFER = 10                        # GCMC
VOL_TRANSFER = 9                         # NPT
BOX_TOTAL = {0: 1, 1: 2, 2: 1}           # src/EnsemblePreprocessor.h
MP_TYPES = 2                             # src/MoveSettings.h:21
MT_N = 624


class GomcCheckpointError(ValueError):
    """A checkpoint that cannot be read, or that does not fit the box."""


# --------------------------------------------------------------- writer --
class _Out:
    def __init__(self):
        self.parts: list[bytes] = []
        self.seen: set[str] = set()

    def version(self, cls: str):
        if cls not in self.seen:
            self.seen.add(cls)
            self.parts.append(struct.pack("<I", 0))

    def raw(self, fmt: str, *v):
        self.parts.append(struct.pack("<" + fmt, *v))

    def u32(self, v):
        self.raw("I", int(v))

    def i32(self, v):
        self.raw("i", int(v))

    def u64(self, v):
        self.raw("Q", int(v))

    def f64(self, v):
        self.raw("d", float(v))

    def boolean(self, v):
        self.raw("?", bool(v))

    def size(self, n):
        self.u64(n)

    def string(self, s: str):
        b = s.encode("latin-1")
        self.size(len(b))
        self.parts.append(b)

    def block(self, arr, dtype):
        a = np.ascontiguousarray(arr, dtype=dtype)
        self.parts.append(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())

    def vec(self, arr, dtype):
        a = np.asarray(arr, dtype=dtype).reshape(-1)
        self.size(a.size)
        self.block(a, dtype)

    def vec_bool(self, v):
        self.size(len(v))
        for x in v:
            self.boolean(x)

    def vec_str(self, v):
        self.size(len(v))
        for s in v:
            self.string(s)

    def nested(self, v, depth, dtype):
        self.size(len(v))
        for x in v:
            if depth == 1:
                self.vec(x, dtype)
            else:
                self.nested(x, depth - 1, dtype)

    def bytes(self) -> bytes:
        return b"".join(self.parts)


# --------------------------------------------------------------- reader --
class _In:
    def __init__(self, buf: bytes):
        self.b = memoryview(buf)
        self.p = 0
        self.seen: set[str] = set()

    def take(self, n):
        if n < 0 or self.p + n > len(self.b):
            raise GomcCheckpointError("checkpoint archive is truncated")
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def version(self, cls: str):
        if cls not in self.seen:
            self.seen.add(cls)
            v = self.u32()
            if v != 0:
                raise GomcCheckpointError(f"unknown class version {v} for {cls}")

    def raw(self, fmt: str):
        n = struct.calcsize("<" + fmt)
        return struct.unpack("<" + fmt, self.take(n))

    def u32(self):
        return self.raw("I")[0]

    def i32(self):
        return self.raw("i")[0]

    def u64(self):
        return self.raw("Q")[0]

    def f64(self):
        return self.raw("d")[0]

    def boolean(self):
        v = self.take(1)[0]
        if v > 1:
            raise GomcCheckpointError("malformed bool in checkpoint archive")
        return bool(v)

    def size(self, limit=1 << 31):
        n = self.u64()
        if n > limit or n > len(self.b) - self.p + 64:
            raise GomcCheckpointError("implausible container size in checkpoint archive")
        return n

    def string(self) -> str:
        n = self.size()
        return bytes(self.take(n)).decode("latin-1")

    def block(self, n, dtype):
        dt = np.dtype(dtype).newbyteorder("<")
        return np.frombuffer(self.take(n * dt.itemsize), dtype=dt).astype(np.dtype(dtype))

    def vec(self, dtype):
        return self.block(self.size(), dtype)

    def vec_bool(self):
        return [self.boolean() for _ in range(self.size())]

    def vec_str(self):
        return [self.string() for _ in range(self.size())]

    def nested(self, depth, dtype):
        n = self.size()
        return [self.vec(dtype) if depth == 1 else self.nested(depth - 1, dtype)
                for _ in range(n)]


# ------------------------------------------------- the archived classes --
_BOND_FIELDS = {"bonds": ("Bond", 2), "angles": ("Angle", 3), "dihedrals": ("Dihedral", 4),
                "impropers": ("Improper", 4), "donors": ("Bond", 2), "acceptors": ("Bond", 2)}


def _put_lookup(o: _Out, lk: dict):
    # MoleculeLookup::serialize (src/MoleculeLookup.h:194-204)
    o.version("MoleculeLookup")
    for key in ("mol_lookup_count", "atom_count", "box_and_kind_start_length",
                "box_and_kind_swappable_length", "num_kinds"):
        o.u32(lk[key])
    for key in ("fixed_molecule", "can_swap_kind", "can_move_kind"):
        o.vec(lk[key], np.uint32)


def _get_lookup(r: _In) -> dict:
    r.version("MoleculeLookup")
    lk = {key: r.u32() for key in ("mol_lookup_count", "atom_count",
                                   "box_and_kind_start_length",
                                   "box_and_kind_swappable_length", "num_kinds")}
    for key in ("fixed_molecule", "can_swap_kind", "can_move_kind"):
        lk[key] = r.vec(np.uint32)
    return lk


def _put_molkind(o: _Out, k: dict):
    # mol_setup::MolKind / Atom / Bond / Angle / Dihedral / Improper
    # serialize() (src/MolSetup.h:50-62,98-108,128-136,156-164,178-184,
    # 198-204,237-258)
    o.version("MolKind")
    o.size(len(k["atoms"]))
    for a in k["atoms"]:
        o.version("Atom")
        for key in ("name", "type", "residue", "segment"):
            o.string(a[key])
        o.f64(a["charge"])
        o.f64(a["mass"])
        o.u32(a["residue_id"])
        o.u32(a["kind"])
    for field, (cls, n) in _BOND_FIELDS.items():
        o.size(len(k[field]))
        for b in k[field]:
            o.version(cls)
            for v in b[:n]:
                o.u32(v)
            o.u32(b[n])  # kind
    o.u32(k["kind_index"])
    o.u32(k["first_atom_id"])
    o.u32(k["first_mol_id"])
    o.boolean(k["incomplete"])
    o.boolean(k["is_multi_residue"])
    o.vec(k["intra_molecule_res_ids"], np.uint32)
    o.string(k["molecule_name"])


def _get_molkind(r: _In) -> dict:
    r.version("MolKind")
    k = {"atoms": []}
    for _ in range(r.size()):
        r.version("Atom")
        a = {key: r.string() for key in ("name", "type", "residue", "segment")}
        a["charge"] = r.f64()
        a["mass"] = r.f64()
        a["residue_id"] = r.u32()
        a["kind"] = r.u32()
        k["atoms"].append(a)
    for field, (cls, n) in _BOND_FIELDS.items():
        k[field] = []
        for _ in range(r.size()):
            r.version(cls)
            k[field].append([r.u32() for _ in range(n + 1)])
    k["kind_index"] = r.u32()
    k["first_atom_id"] = r.u32()
    k["first_mol_id"] = r.u32()
    k["incomplete"] = r.boolean()
    k["is_multi_residue"] = r.boolean()
    k["intra_molecule_res_ids"] = r.vec(np.uint32)
    k["molecule_name"] = r.string()
    return k


_MOLVAR_UINTS = ("last_atom_index_in_box0", "number_mols_in_box0", "mol_kind_index",
                 "string_suffix_multi_residue", "string_suffix_non_multi_residue",
                 "molecule_iteration")
_MOLVAR_STRS = ("molecule_names", "molecule_kind_names", "molecule_segment_names",
                "unique_map_keys")


def _put_molsetup(o: _Out, ms: dict):
    # MolSetup::serialize (src/MolSetup.h:333-337): kindMap then molVars
    o.version("MolSetup")
    km = sorted(ms["kind_map"], key=lambda kv: kv[0].encode("latin-1"))
    o.size(len(km))
    for key, kind in km:
        o.string(key)
        _put_molkind(o, kind)
    mv = ms["mol_vars"]
    o.version("MoleculeVariables")
    o.vec(mv["start_idx_molecules"], np.uint32)
    o.vec(mv["molecule_kinds"], np.uint32)
    for key in _MOLVAR_STRS:
        o.vec_str(mv[key])
    for key in _MOLVAR_UINTS:
        o.u32(mv[key])


def _get_molsetup(r: _In) -> dict:
    r.version("MolSetup")
    km = []
    for _ in range(r.size()):
        key = r.string()
        km.append([key, _get_molkind(r)])
    r.version("MoleculeVariables")
    mv = {"start_idx_molecules": r.vec(np.uint32), "molecule_kinds": r.vec(np.uint32)}
    for key in _MOLVAR_STRS:
        mv[key] = r.vec_str()
    for key in _MOLVAR_UINTS:
        mv[key] = r.u32()
    return {"kind_map": km, "mol_vars": mv}


def _put_pdb(o: _Out, pa: dict, box_total: int):
    # pdb_setup::Atoms::serialize (src/PDBSetup.h:117-136)
    o.version("PdbAtoms")
    chain = pa["chain_letter"]
    o.size(len(chain))
    o.parts.append(chain.encode("latin-1"))
    for key in ("x", "y", "z", "beta", "occ"):
        o.vec(pa[key], np.float64)
    o.vec(pa["box"], np.uint32)
    o.vec_str(pa["res_names"])
    for key in ("restart", "first_res_in_file", "recalc_trajectory"):
        o.boolean(pa[key])
    o.u32(pa["curr_box"])
    o.u32(pa["count"])
    nab = np.asarray(pa["num_atoms_in_box"], np.uint32)
    if nab.size != box_total:
        raise GomcCheckpointError("num_atoms_in_box must hold one count per box")
    o.block(nab, np.uint32)


def _get_pdb(r: _In, box_total: int) -> dict:
    r.version("PdbAtoms")
    pa = {"chain_letter": bytes(r.take(r.size())).decode("latin-1")}
    for key in ("x", "y", "z", "beta", "occ"):
        pa[key] = r.vec(np.float64)
    pa["box"] = r.vec(np.uint32)
    pa["res_names"] = r.vec_str()
    for key in ("restart", "first_res_in_file", "recalc_trajectory"):
        pa[key] = r.boolean()
    pa["curr_box"] = r.u32()
    pa["count"] = r.u32()
    pa["num_atoms_in_box"] = r.block(box_total, np.uint32)
    return pa


_LOOKUP_ARRAYS = (("mol_lookup", np.uint32, "mol_lookup_count"),
                  ("box_and_kind_start", np.uint32, "box_and_kind_start_length"),
                  ("box_and_kind_swappable_counts", np.uint32, "box_and_kind_swappable_length"),
                  ("mol_index", np.int32, "atom_count"),
                  ("atom_index", np.int32, "atom_count"),
                  ("mol_kind", np.int32, "atom_count"),
                  ("atom_kind", np.int32, "atom_count"),
                  ("atom_charge", np.float64, "atom_count"))


def dumps(chk: dict, box_total: int) -> bytes:
    """The archive bytes CheckpointOutput::saveCheckpointFile writes for
    `chk` (src/CheckpointOutput.cpp:69-90)."""
    o = _Out()
    o.version("Checkpoint")
    ver = bytes(chk.get("gomc_version", b"2.80\0"))
    if len(ver) != 5:
        raise GomcCheckpointError("gomc_version is 5 bytes")
    o.parts.append(ver)
    o.u64(chk["step"])
    o.u64(chk["true_step"])
    sa = np.asarray(chk["save_array"], np.uint32)
    if sa.size != MT_N + 1:
        raise GomcCheckpointError("save_array holds 625 words")
    o.block(sa, np.uint32)
    o.u32(chk["seed_location"])
    o.u32(chk["seed_left"])
    o.u32(chk["seed_value"])
    o.nested(chk["scale"], 2, np.float64)
    o.nested(chk["accept_percent"], 2, np.float64)
    for key in ("accepted", "tries", "temp_accepted", "temp_tries"):
        o.nested(chk[key], 2, np.uint32)
    for key in ("mp_tries", "mp_accepted", "mp_interval_accepted", "mp_interval_tries"):
        o.nested(chk[key], 1, np.uint32)
    o.vec(chk["mp_t_max"], np.float64)
    o.vec(chk["mp_r_max"], np.float64)
    o.vec_bool(chk["is_single_move_accepted"])
    o.vec(chk["original_start"], np.uint32)
    o.vec(chk["restarted_start"], np.uint32)
    lk = chk["lookup"]
    _put_lookup(o, lk)
    _put_molsetup(o, chk["mol_setup"])
    _put_pdb(o, chk["pdb_atoms"], box_total)
    for key, dt, count in _LOOKUP_ARRAYS:
        a = np.asarray(lk[key], dt)
        if a.size != lk[count]:
            raise GomcCheckpointError(f"lookup array {key} must hold {count} entries")
        o.block(a, dt)
    return o.bytes()


def loads(buf: bytes, box_total: int) -> dict:
    """Inverse of dumps: the checkpoint dict of an archive's bytes
    (CheckpointSetup::loadCheckpointFile, src/CheckpointSetup.cpp:56-96)."""
    r = _In(buf)
    r.version("Checkpoint")
    chk = {"gomc_version": bytes(r.take(5)), "step": r.u64(), "true_step": r.u64(),
           "save_array": r.block(MT_N + 1, np.uint32), "seed_location": r.u32(),
           "seed_left": r.u32(), "seed_value": r.u32()}
    chk["scale"] = r.nested(2, np.float64)
    chk["accept_percent"] = r.nested(2, np.float64)
    for key in ("accepted", "tries", "temp_accepted", "temp_tries"):
        chk[key] = r.nested(2, np.uint32)
    for key in ("mp_tries", "mp_accepted", "mp_interval_accepted", "mp_interval_tries"):
        chk[key] = r.nested(1, np.uint32)
    chk["mp_t_max"] = r.vec(np.float64)
    chk["mp_r_max"] = r.vec(np.float64)
    chk["is_single_move_accepted"] = r.vec_bool()
    chk["original_start"] = r.vec(np.uint32)
    chk["restarted_start"] = r.vec(np.uint32)
    lk = _get_lookup(r)
    chk["lookup"] = lk
    chk["mol_setup"] = _get_molsetup(r)
    chk["pdb_atoms"] = _get_pdb(r, box_total)
    for key, dt, count in _LOOKUP_ARRAYS:
        lk[key] = r.block(lk[count], dt)
    if r.p != len(r.b):
        raise GomcCheckpointError("trailing bytes after the checkpoint archive")
    return chk


def write_checkpoint(path: str, chk: dict, box_total: int) -> None:
    with open(path, "wb") as f:
        f.write(dumps(chk, box_total))


def read_checkpoint(path: str, box_total: int) -> dict:
    with open(path, "rb") as f:
        return loads(f.read(), box_total)


# ------------------------------------------------------------- topology --
def alpha_suffix(n: int) -> str:
    """AlphaNum::uint2String (lib/AlphaNum.cpp:16-37): 0 -> A, 25 -> Z,
    26 -> AA, ..."""
    out = []
    v = int(n)
    while True:
        out.append(chr(ord("A") + v % 26))
        v = v // 26 - 1
        if v < 0:
            break
    return "".join(reversed(out))


def _components(natom: int, bonds: np.ndarray) -> list[list[int]]:
    """BondAdjacencyList's connected components (src/BondAdjacencyList.cpp:
    21-90): each sorted, then the components in order."""
    adj = [[] for _ in range(natom)]
    for a, b in np.asarray(bonds, np.int64).reshape(-1, 2):
        adj[a].append(int(b))
        adj[b].append(int(a))
    seen = [False] * natom
    comps = []
    for v in range(natom):
        if seen[v]:
            continue
        stack, comp = [v], []
        seen[v] = True
        while stack:
            u = stack.pop()
            comp.append(u)
            for w in adj[u]:
                if not seen[w]:
                    seen[w] = True
                    stack.append(w)
        comps.append(sorted(comp))
    return sorted(comps)


def _mol_bonds(bonds: np.ndarray, begin: int, n: int) -> list[list[int]]:
    """MolSetup::copyBondInfoIntoMapEntry (src/MolSetup.cpp:648-668): the
    adjacency lists are built by prepending each bond's two ends in file
    order; atoms are walked in order, each list from its head, a bond kept
    once (from its lower end), and the list reversed at the end."""
    head = {}
    for a, b in np.asarray(bonds, np.int64).reshape(-1, 2):
        head.setdefault(int(a), []).insert(0, int(b))
        head.setdefault(int(b), []).insert(0, int(a))
    out = []
    for i in range(begin, begin + n):
        for v in head.get(i, []):
            if i < v:
                out.append([i - begin, v - begin, 0])
    out.reverse()
    return out


def _atom_rec(psf, i: int, kind: int) -> dict:
    return {"name": psf.name[i], "type": psf.type[i], "residue": psf.resname[i],
            "segment": psf.segment[i], "charge": float(psf.charge[i]),
            "mass": float(psf.mass[i]), "residue_id": int(psf.resid[i]), "kind": int(kind)}


def _same_atoms(kind_atoms: list[dict], psf, comp: list[int]) -> bool:
    # mol_setup::Atom::operator== (src/MolSetup.h:87-93): type, charge,
    # residue and mass
    for a, i in zip(kind_atoms, comp):
        if not (a["type"] == psf.type[i] and a["charge"] == float(psf.charge[i])
                and a["residue"] == psf.resname[i] and a["mass"] == float(psf.mass[i])):
            return False
    return True


def _new_kind(psf, comp, type_index, kind_index, name) -> dict:
    return {"atoms": [_atom_rec(psf, i, type_index[psf.type[i]]) for i in comp],
            "bonds": _mol_bonds(psf.bonds, comp[0], len(comp)), "angles": [],
            "dihedrals": [], "impropers": [], "donors": [], "acceptors": [],
            "kind_index": kind_index, "first_atom_id": comp[0] + 1,
            "first_mol_id": int(psf.resid[comp[0]]), "incomplete": True,
            "is_multi_residue": False, "intra_molecule_res_ids": [], "molecule_name": name}


def _read_box_kinds(psf, box: int, mv: dict, prior: dict | None, prior_sizes: dict | None,
                    type_index: dict) -> tuple[dict, dict]:
    """createKindMap (src/MolSetup.cpp:356-645) for one box's PSF, molecules
    of one residue each, then the box's angle, dihedral, improper, donor and
    acceptor sections (ReadPSF, MolSetup.cpp:932-1182)."""
    kmap: dict[str, dict] = {}
    sizes: dict[int, list[str]] = {}
    offset = 0
    if box == 1 and mv["number_mols_in_box0"] != 0:
        offset = mv["last_atom_index_in_box0"] + 1
    comps = _components(len(psf.name), psf.bonds)
    for comp in comps:
        if any(psf.resid[i] != psf.resid[comp[0]] for i in comp):
            raise GomcCheckpointError("molecules of more than one residue are not built")
        seg = psf.segment[comp[0]]
        found = False
        if prior is not None:
            for key in prior_sizes.get(len(comp), []):
                if _same_atoms(prior[key]["atoms"], psf, comp):
                    mv["start_idx_molecules"].append(offset + comp[0])
                    mv["molecule_kinds"].append(prior[key]["kind_index"])
                    mv["molecule_names"].append(prior[key]["molecule_name"])
                    mv["molecule_segment_names"].append(seg)
                    if key not in kmap:
                        kmap[key] = _new_kind(psf, comp, type_index, prior[key]["kind_index"],
                                              prior[key]["molecule_name"])
                        sizes.setdefault(len(comp), []).append(key)
                    found = True
                    break
        if not found:
            match = None
            for key in sizes.get(len(comp), []):
                if _same_atoms(kmap[key]["atoms"], psf, comp):
                    match = key
                    break
            if match is not None:
                mv["start_idx_molecules"].append(offset + comp[0])
                mv["molecule_kinds"].append(kmap[match]["kind_index"])
                mv["molecule_names"].append(kmap[match]["molecule_name"])
                mv["molecule_segment_names"].append(seg)
            else:
                key = "MOL" + alpha_suffix(mv["string_suffix_non_multi_residue"])
                mv["string_suffix_non_multi_residue"] += 1
                name = psf.resname[comp[0]]
                kmap[key] = _new_kind(psf, comp, type_index, mv["mol_kind_index"], name)
                mv["start_idx_molecules"].append(offset + comp[0])
                mv["molecule_kinds"].append(mv["mol_kind_index"])
                mv["molecule_kind_names"].append(name)
                mv["unique_map_keys"].append(key)
                mv["molecule_names"].append(name)
                mv["molecule_segment_names"].append(seg)
                mv["mol_kind_index"] += 1
                sizes.setdefault(len(comp), []).append(key)
        mv["molecule_iteration"] += 1
    if box == 0:
        mv["number_mols_in_box0"] = len(comps)
        if comps:
            mv["last_atom_index_in_box0"] = comps[-1][-1]
    # the first molecule of each kind of this box takes the section entries
    # whose first atom lies in it (ReadPSFAngles and its siblings)
    first = sorted((k["first_atom_id"] - 1, key) for key, k in kmap.items())
    for field, width, dedupe in (("angles", 3, False), ("dihedrals", 4, True),
                                 ("impropers", 4, True), ("donors", 2, False),
                                 ("acceptors", 2, False)):
        for row in np.asarray(getattr(psf, field, np.zeros((0, width), np.int64))):
            for begin, key in first:
                n = len(kmap[key]["atoms"])
                if begin <= row[0] < begin + n:
                    entry = [int(v) - begin for v in row] + [0]
                    if not (dedupe and entry in kmap[key][field]):
                        kmap[key][field].append(entry)
                    break
    return kmap, sizes


def mol_setup_of(psfs: list, ff_types: list[str]) -> dict:
    """MolSetup as GOMC builds it from the PSF of each box
    (mol_setup::ReadCombinePSF, src/MolSetup.cpp:273-310), with each atom's
    kind its type's position in `ff_types` (the parameter file's order).  GOMC recomputes the kinds of
    atoms, bonds, angles and dihedrals from its parameter files when it loads
    a checkpoint (CheckpointSetup::SetMoleculeSetup -> MolSetup::AssignKinds,
    src/CheckpointSetup.cpp:173-176), so the bonded kinds are written as 0."""
    type_index = {t: i for i, t in enumerate(ff_types)}
    mv = {"start_idx_molecules": [], "molecule_kinds": [], "molecule_names": [],
          "molecule_kind_names": [], "molecule_segment_names": [], "unique_map_keys": [],
          "last_atom_index_in_box0": 0, "number_mols_in_box0": 0, "mol_kind_index": 0,
          "string_suffix_multi_residue": 0, "string_suffix_non_multi_residue": 0,
          "molecule_iteration": 0}
    kmap, sizes = _read_box_kinds(psfs[0], 0, mv, None, None, type_index)
    if len(psfs) > 1 and len(psfs[1].name) > 0:
        k1, _ = _read_box_kinds(psfs[1], 1, mv, kmap, sizes, type_index)
        for key, val in k1.items():
            kmap.setdefault(key, val)
    return {"kind_map": [[k, kmap[k]] for k in sorted(kmap, key=lambda x: x.encode("latin-1"))],
            "mol_vars": mv}


def pdb_atoms_of(pdbs: list, box_total: int) -> dict:
    """pdb_setup::Atoms as GOMC reads the input PDB of each box
    (src/PDBSetup.cpp:81-116): the box of an atom is the file it came from."""
    xyz = np.concatenate([p.xyz for p in pdbs]) if pdbs else np.zeros((0, 3))
    counts = [len(p.name) for p in pdbs] + [0] * (box_total - len(pdbs))
    return {"chain_letter": "".join((c or " ")[0] for p in pdbs for c in p.chain),
            "x": xyz[:, 0].copy(), "y": xyz[:, 1].copy(), "z": xyz[:, 2].copy(),
            "beta": np.concatenate([p.beta for p in pdbs]) if pdbs else np.zeros(0),
            "occ": np.concatenate([p.occupancy for p in pdbs]) if pdbs else np.zeros(0),
            "box": np.concatenate([np.full(len(p.name), b, np.uint32)
                                   for b, p in enumerate(pdbs)]) if pdbs else np.zeros(0),
            "res_names": [r for p in pdbs for r in p.resname], "restart": False,
            "first_res_in_file": False, "recalc_trajectory": False,
            "curr_box": box_total - 1, "count": int(sum(counts)),
            "num_atoms_in_box": np.array(counts[:box_total], np.uint32)}


# ----------------------------------------------- this library's box state --
# gomc_box_checkpoint's own format (gomc_amd/csrc/checkpoint.cpp, version 2):
# read here to take the Monte Carlo state out of a box, and rebuilt to put a
# GOMC checkpoint's state into one.  The field order is Serialize()'s.
_STATE_MAGIC = b"GOMCHCKP"


class _StateIn(_In):
    def vec(self, dtype):
        return self.block(self.u64(), dtype)


def _parse_state(blob: bytes) -> dict:
    r = _StateIn(blob)
    if bytes(r.take(8)) != _STATE_MAGIC or r.u32() != 2:
        raise GomcCheckpointError("not a version-2 box state")
    st = {"step": r.raw("q")[0], "natoms": r.i32(), "nmols": r.i32(), "nkinds": r.i32(),
          "ensemble": r.i32(), "mol_kind": r.vec(np.int32), "fingerprint": r.u64(),
          "L": [r.f64() for _ in range(3)], "half": [r.f64() for _ in range(3)],
          "volume": r.f64(), "vol_inv": r.f64()}
    for key in ("hx", "hy", "hz", "comx", "comy", "comz"):
        st[key] = r.vec(np.float64)
    st["mt_words"] = r.vec(np.uint32)
    st["mt_index"] = r.i32()
    for mv in range(2):
        st[f"scale{mv}"] = r.vec(np.float64)
        for key in ("tries", "accepted", "temp_tries", "temp_accepted"):
            st[f"{key}{mv}"] = r.vec(np.uint64)
    st["vol_scale"] = r.f64()
    for key in ("vol_tries", "vol_accepted", "vol_temp_tries", "vol_temp_accepted"):
        st[key] = r.u64()
    st["mp_t_max"] = r.f64()
    st["mp_r_max"] = r.f64()
    for i in range(2):
        for key in ("mp_tries", "mp_acc", "mp_itries", "mp_iacc"):
            st.setdefault(key, []).append(r.u64())
    st["mol_box"] = r.vec(np.int32)
    st["lookup"] = r.vec(np.uint32)
    st["bk_start"] = r.vec(np.uint32)
    st["swap_count"] = r.vec(np.uint32)
    st["kind_count"] = r.vec(np.int32)
    st["swap_tries"] = r.vec(np.uint64)
    st["swap_acc"] = r.vec(np.uint64)
    st["pot"] = [r.f64() for _ in range(7)]
    st["sum_r"] = r.vec(np.float64)
    st["sum_i"] = r.vec(np.float64)
    if r.p != len(r.b):
        raise GomcCheckpointError("trailing bytes after the box state")
    return st


def _build_state(st: dict) -> bytes:
    o = _Out()
    o.parts.append(_STATE_MAGIC)
    o.u32(2)
    o.raw("q", int(st["step"]))
    for key in ("natoms", "nmols", "nkinds", "ensemble"):
        o.i32(st[key])
    o.vec(st["mol_kind"], np.int32)
    o.u64(st["fingerprint"])
    for v in list(st["L"]) + list(st["half"]) + [st["volume"], st["vol_inv"]]:
        o.f64(v)
    for key in ("hx", "hy", "hz", "comx", "comy", "comz"):
        o.vec(st[key], np.float64)
    o.vec(st["mt_words"], np.uint32)
    o.i32(st["mt_index"])
    for mv in range(2):
        o.vec(st[f"scale{mv}"], np.float64)
        for key in ("tries", "accepted", "temp_tries", "temp_accepted"):
            o.vec(st[f"{key}{mv}"], np.uint64)
    o.f64(st["vol_scale"])
    for key in ("vol_tries", "vol_accepted", "vol_temp_tries", "vol_temp_accepted"):
        o.u64(st[key])
    o.f64(st["mp_t_max"])
    o.f64(st["mp_r_max"])
    for i in range(2):
        for key in ("mp_tries", "mp_acc", "mp_itries", "mp_iacc"):
            o.u64(st[key][i])
    o.vec(st["mol_box"], np.int32)
    o.vec(st["lookup"], np.uint32)
    o.vec(st["bk_start"], np.uint32)
    o.vec(st["swap_count"], np.uint32)
    o.vec(st["kind_count"], np.int32)
    o.vec(st["swap_tries"], np.uint64)
    o.vec(st["swap_acc"], np.uint64)
    for v in st["pot"]:
        o.f64(v)
    o.vec(st["sum_r"], np.float64)
    o.vec(st["sum_i"], np.float64)
    return o.bytes()


# ------------------------------------------------- box <-> GOMC archive --
class GomcTopology:
    """What GOMC's checkpoint keeps of a run's input files: the PSF structure
    and PDB frame of each box (box 1: a GCMC system's reservoir), and the
    atom type names in the order of the parameter file, whose positions GOMC
    takes as atom kinds (MolSetup::AssignKinds -> AssignAtomKinds,
    src/MolSetup.cpp:680-692)."""

    def __init__(self, psfs: list, pdbs: list, ff_types: list[str]):
        self.psfs, self.pdbs, self.ff_types = list(psfs), list(pdbs), list(ff_types)

    @classmethod
    def from_system(cls, s, type_names: list[str] | None = None) -> "GomcTopology":
        """The input files `gomc_io.write_gomc` writes for `s` (its initial
        state); their parameter file lists the types in the system's order."""
        from .gomc_io import gomc_box_records
        names = type_names or [f"T{i}" for i in range(len(s.sigma))]
        psfs, pdbs = gomc_box_records(s, names)
        return cls(psfs, pdbs, names)

    @classmethod
    def from_gomc(cls, conf_path: str, ensemble: str | None = None) -> "GomcTopology":
        """The PSF, PDB and parameter files a GOMC control file names."""
        from .gomc_io import read_conf, read_parameters, read_pdb, read_psf
        conf = read_conf(conf_path)
        if ensemble is None:
            ensemble = "GCMC" if (conf.chem_pot or conf.fugacity) else "NVT"
        boxes = [0, 1] if ensemble == "GCMC" else [0]
        psfs = [read_psf(conf.structure[b]) for b in boxes]
        pdbs = [read_pdb(conf.coordinates[b]) for b in boxes]
        return cls(psfs, pdbs, list(read_parameters(conf.parameters, conf.charmm)))

    def atom_kinds(self) -> np.ndarray:
        """The parameter-file index of every atom, box 0's PSF then box 1's."""
        idx = {t: i for i, t in enumerate(self.ff_types)}
        return np.array([idx[t] for p in self.psfs for t in p.type], np.int32)


def _restart_order(st: dict, nmols: int, starts: np.ndarray, lens: np.ndarray):
    """Atoms of the restart coordinate files: molecules in MoleculeLookup
    order, box 0 then box 1 (ExtendedSystemOutput::SetMolInBox,
    src/ExtendedSystemOutput.cpp:255-287)."""
    order = [int(m) for m in st["lookup"]] if len(st["lookup"]) else list(range(nmols))
    atoms = np.concatenate([np.arange(starts[m], starts[m] + lens[m]) for m in order]) \
        if order else np.zeros(0, np.int64)
    return order, atoms


def _check_topology(s, topo: GomcTopology):
    """GOMC's original atom order is box 0's PSF then box 1's: the system's
    molecules must come in that order, with the PSF's atom counts."""
    n_file = [len(p.name) for p in topo.psfs]
    lens = [len(s.kinds[k].atom_types) for k in s.mol_kind]
    mb = s.mol_box if s.mol_box is not None else np.zeros(s.nmols, np.int32)
    if sum(n_file) != sum(lens) or (len(mb) and np.any(np.diff(np.asarray(mb)) < 0)):
        raise GomcCheckpointError("the system's molecules are not in the input files' order "
                                  "(box 0's molecules first, then box 1's)")
    if sum(lens[m] for m in range(s.nmols) if mb[m] == 0) != n_file[0]:
        raise GomcCheckpointError("box 0's atom count differs from its PSF")


def checkpoint_of(box, topo: GomcTopology, step: int, true_step: int | None = None) -> dict:
    """The checkpoint GOMC writes after `step` moves of the run `box` holds
    (Checkpoint::Checkpoint, src/Checkpoint.cpp:9-31; CheckpointOutput::
    DoOutputRestart passes step + 1 of the last move, i.e. the moves done).

    Move settings go to [box][move][kind] (DISPLACE 0, ROTATE 1,
    MULTIPARTICLE 2; MOL_TRANSFER 10 under GCMC, VOL_TRANSFER 9 under NPT).
    The counts since the last adjustment of a transfer are not kept by this
    library and are written as zero (they steer nothing:
    MoveSettings::Adjust changes only the displacement, rotation and volume
    scales, src/MoveSettings.cpp:213-258)."""
    s = box.system
    ens = int(s.ensemble)
    nb, nmk, K = BOX_TOTAL[ens], MOVE_KINDS[ens], len(s.kinds)
    _check_topology(s, topo)
    st = _parse_state(box.checkpoint(step))
    starts = np.asarray(s.mol_start(), np.int64)
    lens = np.array([len(s.kinds[k].atom_types) for k in s.mol_kind], np.int64)
    # MT19937: state words, then `left` (MTRand::save, src/MersenneTwister.h:
    # 392-398); pNext - state = the words used since the last reload
    idx = int(st["mt_index"])
    save = np.concatenate([st["mt_words"], [MT_N - idx]]).astype(np.uint32)

    def zeros(dt):
        return [[np.zeros(K, dt) for _ in range(nmk)] for _ in range(nb)]
    scale, pct = zeros(np.float64), zeros(np.float64)
    cnt = {key: zeros(np.uint64) for key in ("accepted", "tries", "temp_accepted", "temp_tries")}
    for mv in (DISPLACE, ROTATE):
        scale[0][mv][:] = st[f"scale{mv}"]
        for key, src in (("tries", "tries"), ("accepted", "accepted"),
                         ("temp_tries", "temp_tries"), ("temp_accepted", "temp_accepted")):
            cnt[key][0][mv][:] = st[f"{src}{mv}"]
    # MultiParticle: one count per move for every kind (MoveSettings::Update,
    # src/MoveSettings.cpp:118-145); its interval counts are those since the
    # last adjustment
    mp_all = sum(st["mp_tries"]) + sum(st["mp_itries"])
    mp_acc = sum(st["mp_acc"]) + sum(st["mp_iacc"])
    cnt["tries"][0][MULTIPARTICLE][:] = mp_all
    cnt["accepted"][0][MULTIPARTICLE][:] = mp_acc
    cnt["temp_tries"][0][MULTIPARTICLE][:] = sum(st["mp_itries"])
    cnt["temp_accepted"][0][MULTIPARTICLE][:] = sum(st["mp_iacc"])
    if ens == 1:
        # reservoir box: GOMC's initial scales of its cell (MoveSettings::Init,
        # src/MoveSettings.cpp:42-57), never adjusted (no moves in box 1)
        scale[1][DISPLACE][:] = min(s.box) * 0.25
        scale[1][ROTATE][:] = math.pi / 4
        for b in range(2):
            cnt["tries"][b][MOL_TRANSFER][:] = st["swap_tries"][b * K:(b + 1) * K]
            cnt["accepted"][b][MOL_TRANSFER][:] = st["swap_acc"][b * K:(b + 1) * K]
    if ens == 2:
        scale[0][VOL_TRANSFER][:] = st["vol_scale"]
        for key, src in (("tries", "vol_tries"), ("accepted", "vol_accepted"),
                         ("temp_tries", "vol_temp_tries"),
                         ("temp_accepted", "vol_temp_accepted")):
            cnt[key][0][VOL_TRANSFER][:] = st[src]
    for b in range(nb):
        for m in range(nmk):
            t = cnt["tries"][b][m].astype(np.float64)
            pct[b][m] = np.where(t > 0, cnt["accepted"][b][m] / np.maximum(t, 1), 0.0)
    # lookup: GCMC keeps box 1 sections; NVT/NPT hold box 0 only
    nmol = s.nmols
    bk = np.asarray(st["bk_start"], np.uint32)[: nb * K + 1].copy()
    bk[-1] = nmol
    swapc = np.asarray(st["swap_count"], np.uint32)[: nb * K]
    order, atoms = _restart_order(st, nmol, starts, lens)
    restarted = np.concatenate([[0], np.cumsum(lens[order])]).astype(np.uint32)
    mol_of_atom = np.repeat(np.arange(nmol), lens)
    atom_in_mol = np.concatenate([np.arange(n) for n in lens]) if nmol else np.zeros(0)
    kind_of_atom = np.repeat(np.asarray(s.mol_kind), lens)
    type_of_atom = topo.atom_kinds()
    charge_of_atom = np.concatenate([s.kinds[k].charges for k in s.mol_kind])
    lookup = {"mol_lookup_count": nmol, "atom_count": int(lens.sum()),
              "box_and_kind_start_length": nb * K + 1, "box_and_kind_swappable_length": nb * K,
              "num_kinds": K, "fixed_molecule": np.zeros(nmol, np.uint32),
              "can_swap_kind": [k for k in range(K) if np.any(s.mol_kind == k)],
              "can_move_kind": [k for k in range(K) if np.any(s.mol_kind == k)],
              "mol_lookup": np.asarray(order, np.uint32), "box_and_kind_start": bk,
              "box_and_kind_swappable_counts": swapc, "mol_index": mol_of_atom,
              "atom_index": atom_in_mol, "mol_kind": kind_of_atom, "atom_kind": type_of_atom,
              "atom_charge": charge_of_atom}
    # canSwapKind / canMoveKind list kinds in the order molecules first show
    # them (MoleculeLookup::Init, src/MoleculeLookup.cpp:73-84)
    first_seen = list(dict.fromkeys(int(k) for k in s.mol_kind))
    lookup["can_swap_kind"] = first_seen
    lookup["can_move_kind"] = first_seen
    return {"gomc_version": b"2.80\0", "step": int(step),
            "true_step": int(step if true_step is None else true_step),
            "save_array": save, "seed_location": idx, "seed_left": MT_N - idx,
            "seed_value": int(s.seed) & 0xFFFFFFFF, "scale": scale, "accept_percent": pct,
            **cnt,
            "mp_tries": [[st["mp_tries"][0], st["mp_tries"][1]]] + [[0, 0]] * (nb - 1),
            "mp_accepted": [[st["mp_acc"][0], st["mp_acc"][1]]] + [[0, 0]] * (nb - 1),
            "mp_interval_accepted": [[st["mp_iacc"][0], st["mp_iacc"][1]]] + [[0, 0]] * (nb - 1),
            "mp_interval_tries": [[st["mp_itries"][0], st["mp_itries"][1]]] + [[0, 0]] * (nb - 1),
            "mp_t_max": [st["mp_t_max"]] * nb, "mp_r_max": [st["mp_r_max"]] * nb,
            "is_single_move_accepted": [True],
            "original_start": np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32),
            "restarted_start": restarted, "lookup": lookup,
            "mol_setup": mol_setup_of(topo.psfs, topo.ff_types),
            "pdb_atoms": pdb_atoms_of(topo.pdbs, nb)}


def _unwrap(x, ref, L):
    # BoxDimensions::UnwrapPBC, per axis (src/BoxDimensions.h:214-229)
    if abs(ref - x) > 0.5 * L:
        return x - L if ref < 0.5 * L else x + L
    return x


def write_restart(prefix: str, box, topo: GomcTopology, step: int,
                  true_step: int | None = None) -> list[str]:
    """The restart set GOMC writes at a checkpoint: `<prefix>_restart.chk`
    (CheckpointOutput), and per box `<prefix>_BOX_<b>_restart.coor` (atoms in
    MoleculeLookup order, unwrapped around their molecule's centre) and
    `<prefix>_BOX_<b>_restart.xsc` (ExtendedSystemOutput::DoOutputRestart,
    src/ExtendedSystemOutput.cpp:220-253).  Returns the paths."""
    s = box.system
    ens = int(s.ensemble)
    nb = BOX_TOTAL[ens]
    chk = checkpoint_of(box, topo, step, true_step)
    paths = [f"{prefix}_restart.chk"]
    write_checkpoint(paths[0], chk, nb)
    st = _parse_state(box.checkpoint(step))
    xyz = np.stack([st["hx"], st["hy"], st["hz"]], axis=1)
    com = np.stack([st["comx"], st["comy"], st["comz"]], axis=1)
    L = st["L"]
    starts = np.asarray(s.mol_start(), np.int64)
    lens = np.array([len(s.kinds[k].atom_types) for k in s.mol_kind], np.int64)
    mol_box = st["mol_box"] if len(st["mol_box"]) else np.zeros(s.nmols, np.int32)
    lookup = chk["lookup"]["mol_lookup"]
    for b in range(nb):
        rows = []
        for m in lookup:
            if mol_box[m] != b:
                continue
            for a in range(starts[m], starts[m] + lens[m]):
                rows.append([_unwrap(xyz[a, d], com[m, d], L[d]) for d in range(3)])
        arr = np.array(rows, np.float64).reshape(-1, 3)
        path = f"{prefix}_BOX_{b}_restart.coor"
        with open(path, "wb") as f:
            f.write(struct.pack("<i", len(arr)))
            f.write(arr.astype("<f8").tobytes())
        paths.append(path)
        xsc = f"{prefix}_BOX_{b}_restart.xsc"
        with open(xsc, "w") as f:
            f.write("#$LABELS step a_x a_y a_z b_x b_y b_z c_x c_y c_z o_x o_y o_z\n")
            vals = [L[0], 0.0, 0.0, 0.0, L[1], 0.0, 0.0, 0.0, L[2],
                    0.5 * L[0], 0.5 * L[1], 0.5 * L[2]]
            f.write(" ".join([str(int(step) if step else 0)] + ["%.12g" % v for v in vals]) + "\n")
        paths.append(xsc)
    return paths


def read_restart(prefix: str, ensemble: int) -> dict:
    """The restart set `write_restart` (or GOMC) wrote: the checkpoint, the
    coordinates of each box in file order and each box's edges."""
    nb = BOX_TOTAL[int(ensemble)]
    chk = read_checkpoint(f"{prefix}_restart.chk", nb)
    coords, cells = [], []
    for b in range(nb):
        with open(f"{prefix}_BOX_{b}_restart.coor", "rb") as f:
            buf = f.read()
        if len(buf) < 4:
            raise GomcCheckpointError(f"box {b}: restart coordinates are truncated")
        n = struct.unpack("<i", buf[:4])[0]
        if n < 0 or len(buf) != 4 + 24 * n:
            raise GomcCheckpointError(f"box {b}: restart coordinates hold {n} atoms "
                                      f"but {len(buf) - 4} bytes")
        coords.append(np.frombuffer(buf[4:], "<f8").reshape(n, 3).astype(np.float64))
        with open(f"{prefix}_BOX_{b}_restart.xsc") as f:
            rows = [ln.split() for ln in f if ln.strip() and not ln.startswith("#")]
        v = [float(x) for x in rows[-1][1:10]]
        if any(abs(v[i]) > 0 for i in (1, 2, 3, 5, 6, 7)):
            raise GomcCheckpointError(f"box {b}: only orthogonal cells are built")
        cells.append((v[0], v[4], v[8]))
    return {"checkpoint": chk, "coords": coords, "cells": cells}


def restore_box(system, restart: dict, device: int = 0, precision: int = 0):
    """A Box that continues the run a GOMC restart set describes; returns
    (box, step), step being the first move to run.

    As GOMC's restart (CheckpointSetup::SetCheckpointData, src/
    CheckpointSetup.cpp:101-189), the coordinates come from the restart
    files and the energies are computed anew from them; the MT19937 state,
    move settings and MoleculeLookup come from the archive, and the
    MultiParticle seed from its seedValue."""
    from dataclasses import replace

    from .api import Box
    chk = restart["checkpoint"]
    s = system
    ens = int(s.ensemble)
    nb, K, nmol = BOX_TOTAL[ens], len(s.kinds), s.nmols
    lk = chk["lookup"]
    lens = np.array([len(s.kinds[k].atom_types) for k in s.mol_kind], np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)])
    if (lk["mol_lookup_count"] != nmol or lk["num_kinds"] != K
            or lk["atom_count"] != int(lens.sum())
            or not np.array_equal(np.asarray(chk["original_start"], np.int64), starts)
            or not np.array_equal(np.asarray(lk["mol_kind"], np.int64),
                                  np.repeat(np.asarray(s.mol_kind), lens))):
        raise GomcCheckpointError("checkpoint belongs to another system")
    order = np.asarray(lk["mol_lookup"], np.int64)
    bk = np.asarray(lk["box_and_kind_start"], np.int64)
    if sorted(order.tolist()) != list(range(nmol)) or bk[0] != 0 or bk[-1] != nmol \
            or np.any(np.diff(bk) < 0):
        raise GomcCheckpointError("checkpoint molecule lookup is not a permutation")
    # box of each molecule from the lookup sections
    mol_box = np.zeros(nmol, np.int32)
    for sec in range(nb * K):
        b, k = divmod(sec, K)
        for m in order[bk[sec]:bk[sec + 1]]:
            if s.mol_kind[m] != k:
                raise GomcCheckpointError("lookup section holds a molecule of another kind")
            mol_box[m] = b
    rs = np.asarray(chk["restarted_start"], np.int64)
    if not np.array_equal(rs, np.concatenate([[0], np.cumsum(lens[order])])):
        raise GomcCheckpointError("restarted_start does not follow the lookup")
    file_xyz = np.concatenate(restart["coords"]) if restart["coords"] else np.zeros((0, 3))
    if len(file_xyz) != int(lens.sum()):
        raise GomcCheckpointError("restart coordinates do not hold every atom")
    # the cell: the system's, except under NPT, whose volume moves leave it
    # in the cell file (GOMC reads it there when restarting from an XSC file)
    L = np.array(restart["cells"][0] if ens == 2 else s.box, np.float64)
    xyz = np.zeros_like(file_xyz)
    for i, m in enumerate(order):
        xyz[starts[m]:starts[m + 1]] = file_xyz[rs[i]:rs[i + 1]]
    # back into [0, L) (BoxDimensions::WrapPBC); a coordinate that rounds to
    # L on the way back is taken to 0 as the box requires
    xyz = np.where(xyz >= L, xyz - L, np.where(xyz < 0.0, xyz + L, xyz))
    xyz = np.where(xyz >= L, xyz - L, xyz)
    sys2 = replace(s, box=tuple(float(v) for v in L), xyz=xyz,
                   mol_box=mol_box if ens == 1 else s.mol_box,
                   seed=int(chk["seed_value"]), _keep=[])
    box = Box(sys2, device=device)
    if precision:
        box.set_precision(precision)
    st = _parse_state(box.checkpoint(0))
    left = int(chk["save_array"][MT_N])
    if left != int(chk["seed_left"]) or MT_N - left != int(chk["seed_location"]):
        raise GomcCheckpointError("MT19937 read position and `left` disagree")
    st["mt_words"] = np.asarray(chk["save_array"][:MT_N], np.uint32)
    st["mt_index"] = MT_N - left
    for mv in (DISPLACE, ROTATE):
        st[f"scale{mv}"] = np.asarray(chk["scale"][0][mv], np.float64)
        for key in ("tries", "accepted", "temp_tries", "temp_accepted"):
            st[f"{key}{mv}"] = np.asarray(chk[key][0][mv], np.uint64)
    if ens == 2:
        st["vol_scale"] = float(chk["scale"][0][VOL_TRANSFER][0])
        for key, src in (("vol_tries", "tries"), ("vol_accepted", "accepted"),
                         ("vol_temp_tries", "temp_tries"),
                         ("vol_temp_accepted", "temp_accepted")):
            st[key] = int(chk[src][0][VOL_TRANSFER][0])
    st["mp_t_max"] = float(chk["mp_t_max"][0])
    st["mp_r_max"] = float(chk["mp_r_max"][0])
    st["mp_tries"] = [int(v) for v in chk["mp_tries"][0]]
    st["mp_acc"] = [int(v) for v in chk["mp_accepted"][0]]
    st["mp_itries"] = [int(v) for v in chk["mp_interval_tries"][0]]
    st["mp_iacc"] = [int(v) for v in chk["mp_interval_accepted"][0]]
    st["lookup"] = order.astype(np.uint32)
    if ens == 1:
        st["bk_start"] = bk.astype(np.uint32)
        st["swap_count"] = np.asarray(lk["box_and_kind_swappable_counts"], np.uint32)
        st["kind_count"] = np.array([np.sum((mol_box == 0) & (s.mol_kind == k))
                                     for k in range(K)], np.int32)
        st["swap_tries"] = np.concatenate([chk["tries"][b][MOL_TRANSFER] for b in range(2)]
                                          ).astype(np.uint64)
        st["swap_acc"] = np.concatenate([chk["accepted"][b][MOL_TRANSFER] for b in range(2)]
                                        ).astype(np.uint64)
    else:
        # NVT/NPT boxes keep two-box sections with an empty box 1
        st["bk_start"] = np.concatenate([bk, np.full(K, nmol)]).astype(np.uint32)
        st["swap_count"] = np.concatenate(
            [np.asarray(lk["box_and_kind_swappable_counts"], np.uint32),
             np.zeros(K, np.uint32)]) if len(st["swap_count"]) else st["swap_count"]
    box.restore(_build_state(st))
    return box, int(chk["step"])
