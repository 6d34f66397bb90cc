"""GOMC's checkpoint archive in cereal's binary layout (gomc_amd/gomc_chk.py).

* The writer produces, byte for byte, the archives that the reference's own
  cereal (lib/cereal) and MolSetup serializers (src/MolSetup.h) write for the
  same values (tests/golden/cereal, made by tests/golden/make_cereal_vectors.py
  from oracle/ref/ref_cereal_chk.cpp), for one box (NVT) and two (GCMC); the
  reader reads those archives back to the values.  The fixtures cover class
  versions written at a class's first appearance (Bond and Angle first inside
  one kind, Dihedral and Improper never), std::vector<bool>, C arrays,
  std::map key order, strings and the lookup's raw arrays after the object.
* Truncated and trailing-byte archives are refused.
* The kind map is GOMC's for the reference's example inputs: the TIP3
  reservoir (one kind MOLA named TIP3, bonds in copyBondInfoIntoMapEntry's
  order, the PSF's angle) and the AR_KR two-box GCMC input (box-1 molecules
  of kinds first seen in box 0 reuse their keys; molecule start indices
  continue after box 0's last atom).
"""
import json
import os

import numpy as np
import pytest

from gomc_amd import gomc_chk, gomc_io

HERE = os.path.dirname(__file__)
GOLD = os.path.join(HERE, "golden", "cereal")
ARKR = os.path.join(HERE, "golden", "gomc_inputs", "AR_KR", "Base", "in.conf")
TIP3 = os.path.join(HERE, "golden", "gomc_inputs", "TIP3")


def _fixture(name):
    with open(os.path.join(GOLD, f"{name}_restart.json")) as f:
        d = json.load(f)
    with open(os.path.join(GOLD, d["archive"]), "rb") as f:
        gold = f.read()
    chk = dict(d)
    chk["gomc_version"] = bytes(d["gomc_version"])
    return chk, gold, d["box_total"]


def _same(a, b, path="chk"):
    if isinstance(a, dict):
        keys = set(a) - {"box_total", "archive"}
        assert isinstance(b, dict) and keys == set(b) - {"box_total", "archive"}, \
            (path, keys ^ set(b))
        for k in keys:
            _same(a[k], b[k], f"{path}.{k}")
    elif isinstance(a, (list, tuple, np.ndarray)) and not isinstance(a, (str, bytes)):
        assert len(a) == len(b), (path, len(a), len(b))
        for i, (x, y) in enumerate(zip(a, b)):
            _same(x, y, f"{path}[{i}]")
    else:
        assert a == b or (isinstance(a, float) and float(a) == float(b)), (path, a, b)


@pytest.mark.parametrize("name", ["nvt", "gcmc"])
def test_writer_equals_cereal_bytes(name):
    chk, gold, nb = _fixture(name)
    assert gomc_chk.dumps(chk, nb) == gold


@pytest.mark.parametrize("name", ["nvt", "gcmc"])
def test_reader_reads_cereal_archive(name):
    chk, gold, nb = _fixture(name)
    back = gomc_chk.loads(gold, nb)
    _same(chk, back)
    assert gomc_chk.dumps(back, nb) == gold


def test_truncated_or_padded_archive_refused():
    _, gold, nb = _fixture("gcmc")
    with pytest.raises(gomc_chk.GomcCheckpointError):
        gomc_chk.loads(gold[:-1], nb)
    with pytest.raises(gomc_chk.GomcCheckpointError):
        gomc_chk.loads(gold + b"\0", nb)
    with pytest.raises(gomc_chk.GomcCheckpointError):
        gomc_chk.loads(gold, 1)  # a two-box archive read as one box


def test_alpha_suffix():
    # lib/AlphaNum.cpp: A..Z, then AA, AB, ...
    assert [gomc_chk.alpha_suffix(n) for n in (0, 1, 25, 26, 27, 51, 52, 701, 702)] == \
        ["A", "B", "Z", "AA", "AB", "AZ", "BA", "ZZ", "AAA"]


def test_kind_map_of_tip3_reservoir():
    psf = gomc_io.read_psf(os.path.join(TIP3, "water_reservior_1_molecule.psf"))
    ms = gomc_chk.mol_setup_of([psf], ["OT", "HT"])
    assert [k for k, _ in ms["kind_map"]] == ["MOLA"]
    k = ms["kind_map"][0][1]
    assert k["molecule_name"] == "TIP3" and k["first_atom_id"] == 1 and k["kind_index"] == 0
    assert [a["name"] for a in k["atoms"]] == ["OH2", "H1", "H2"]
    assert [a["kind"] for a in k["atoms"]] == [0, 1, 1]
    # PSF bonds 1-2, 1-3: prepended to the adjacency lists, walked from the
    # lower end, then reversed -> file order here
    assert k["bonds"] == [[0, 1, 0], [0, 2, 0]]
    assert k["angles"] == [[1, 0, 2, 0]]
    assert not k["is_multi_residue"] and k["incomplete"]
    mv = ms["mol_vars"]
    assert mv["start_idx_molecules"] == [0] and mv["unique_map_keys"] == ["MOLA"]
    assert mv["number_mols_in_box0"] == 1 and mv["last_atom_index_in_box0"] == 2
    assert mv["string_suffix_non_multi_residue"] == 1 and mv["molecule_iteration"] == 1


def test_kind_map_of_ar_kr_two_boxes():
    conf = gomc_io.read_conf(ARKR)
    psfs = [gomc_io.read_psf(conf.structure[b]) for b in (0, 1)]
    pdbs = [gomc_io.read_pdb(conf.coordinates[b]) for b in (0, 1)]
    s = gomc_io.system_from_gomc(ARKR)
    topo = gomc_chk.GomcTopology.from_gomc(ARKR)
    ms = gomc_chk.mol_setup_of(psfs, topo.ff_types)
    mv = ms["mol_vars"]
    # every molecule is listed once, in file order, box 1 after box 0
    n0, n1 = len(psfs[0].name), len(psfs[1].name)
    assert mv["molecule_iteration"] == s.nmols == n0 + n1  # one atom per molecule
    assert mv["start_idx_molecules"] == list(range(n0)) + [n0 + i for i in range(n1)]
    assert mv["number_mols_in_box0"] == n0 and mv["last_atom_index_in_box0"] == n0 - 1
    # kinds: one map entry per residue name, keys MOLA, MOLB, ... in order of
    # first appearance; molecule kinds agree with the system built from the
    # same files
    names = [k["molecule_name"] for _, k in ms["kind_map"]]
    assert sorted(names) == sorted({k.name for k in s.kinds})
    key_of = {k["molecule_name"]: key for key, k in ms["kind_map"]}
    kind_index = {k["molecule_name"]: k["kind_index"] for _, k in ms["kind_map"]}
    assert mv["molecule_kinds"] == [kind_index[s.kinds[k].name] for k in s.mol_kind]
    assert mv["unique_map_keys"] == [key_of[n] for n in mv["molecule_kind_names"]]
    assert mv["unique_map_keys"] == ["MO" + "L" + gomc_chk.alpha_suffix(i)
                                     for i in range(len(mv["unique_map_keys"]))]
    pa = gomc_chk.pdb_atoms_of(pdbs, 2)
    assert pa["count"] == n0 + n1 and list(pa["num_atoms_in_box"]) == [n0, n1]
    assert list(pa["box"]) == [0] * n0 + [1] * n1 and pa["curr_box"] == 1
    # the archive of this topology round-trips through the codec
    chk, _, _ = _fixture("gcmc")
    chk["mol_setup"], chk["pdb_atoms"] = ms, pa
    _same(chk, gomc_chk.loads(gomc_chk.dumps(chk, 2), 2))


def test_kind_map_of_written_system_matches_reading_its_files(tmp_path):
    """The topology of a synthetic system equals what reading the GOMC files
    `write_gomc` writes for it gives (GCMC: reservoir in box 1)."""
    from gomc_amd import systems
    s = systems.gcmc_mixture(5, 4, 24.0, reservoir=6, seed=3)
    conf = gomc_io.write_gomc(str(tmp_path), s)
    a = gomc_chk.GomcTopology.from_system(s)
    b = gomc_chk.GomcTopology.from_gomc(conf, "GCMC")
    assert a.ff_types == b.ff_types
    _same(gomc_chk.mol_setup_of(a.psfs, a.ff_types), gomc_chk.mol_setup_of(b.psfs, b.ff_types))
    assert np.array_equal(a.atom_kinds(), b.atom_kinds())
    _same(gomc_chk.pdb_atoms_of(a.pdbs, 2), gomc_chk.pdb_atoms_of(b.pdbs, 2))
