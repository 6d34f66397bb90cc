"""The C++ shim (gomc_amd/shim/gomc_shim.hpp): CalculateEnergy, Ewald and
CellList methods with the reference's names and return conventions over
the C ABI, compiled with the host compiler against libgomc_hip.so.

* CPU: a box cannot be made without a device, and the shim raises
  gomc_hip::Error with the library's code (GOMC_ERR_NODEV) rather than
  crashing.
* GPU: on a 200-water box the shim's SystemTotal, MoleculeInter,
  MolReciprocal (the change of the reciprocal energy, as
  Ewald::MolReciprocal returns it), ParticleInter (which adds into the
  caller's arrays), the CSR of GetCellListNeighbor and UpdateRecip /
  CommitMol equal the oracle's (energies 1e-11 / 1e-10 relative, the CSR
  element for element); a bad molecule index raises GOMC_ERR_ARG.
"""
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gomc_amd import systems
from oracle.pyoracle import Oracle

ROOT = Path(__file__).resolve().parent.parent


def _build(tmp_path):
    exe = tmp_path / "shim_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), str(ROOT / "tests" / "shim_check.cpp"),
                    f"-L{ROOT / 'gomc_amd'}", "-lgomc_hip", f"-Wl,-rpath,{ROOT / 'gomc_amd'}"],
                   check=True, capture_output=True)
    return exe


def _write(path, s, mol, trial, ptype, pq, ptrials):
    def vec(v):
        v = list(np.ravel(v))
        return f"{len(v)} " + " ".join(repr(float(a)) if isinstance(a, (float, np.floating))
                                       else str(int(a)) for a in v) + "\n"
    with open(path, "w") as f:
        f.write(vec(np.array([*s.box, s.temperature, s.rcut, s.rcut_low, s.rcut_coulomb,
                              s.tolerance], dtype=float)))
        f.write(vec([int(s.ewald), int(s.electrostatic), int(s.lrc)]))
        for a in (s.sigma, s.epsilon, s.n_rep):
            f.write(vec(np.asarray(a, dtype=float)))
        f.write(vec([len(k.atom_types) for k in s.kinds]))
        f.write(vec([t for k in s.kinds for t in k.atom_types]))
        f.write(vec(np.array([q for k in s.kinds for q in k.charges], dtype=float)))
        f.write(vec(s.mol_kind))
        for i in range(3):
            f.write(vec(np.asarray(s.xyz[:, i], dtype=float)))
        f.write(vec(np.concatenate([[float(mol)], np.asarray(trial, float).ravel()])))
        f.write(vec(np.concatenate([[float(ptype), float(pq)], np.asarray(ptrials, float).ravel()])))


def _case(tmp_path):
    s = systems.water_box(200, seed=13, rcut=8.0)
    mol = 7
    st = s.mol_start()[mol]
    trial = (s.xyz[st:st + 3] + np.array([0.31, -0.22, 0.17])) % np.array(s.box)
    rng = np.random.default_rng(2)
    ptrials = rng.uniform(0, s.box[0], (8, 3))
    inp = tmp_path / "box.txt"
    _write(inp, s, mol, trial, 0, -0.8476, ptrials)
    return s, mol, trial, ptrials, inp


def test_shim_raises_the_library_error_without_a_device(tmp_path):
    exe = _build(tmp_path)
    s, mol, trial, ptrials, inp = _case(tmp_path)
    out = json.loads(subprocess.run([str(exe), str(inp)], check=True, capture_output=True,
                                    text=True).stdout)
    if "error_code" not in out:
        pytest.skip("the library found a device: the GPU test covers the shim")
    assert out["error_code"] == -4 and "gomc_box_create" in out["error"]


@pytest.mark.gpu
def test_shim_methods_equal_the_oracle(tmp_path):
    exe = _build(tmp_path)
    s, mol, trial, ptrials, inp = _case(tmp_path)
    r = subprocess.run([str(exe), str(inp)], check=True, capture_output=True, text=True,
                       timeout=120)
    out = json.loads(r.stdout)
    assert "error_code" not in out, out
    o = Oracle(s, threads=4)
    tot = o.system_total()
    assert np.allclose(out["total"], tot, rtol=1e-11, atol=0)
    assert out["recip_setup"] == pytest.approx(tot[3], rel=1e-11)
    lj, coul, ov = o.mol_inter(mol, trial)
    assert out["mol_inter"][2] == ov
    assert abs(out["mol_inter"][0] - lj) <= 1e-10 * max(abs(lj), abs(tot[0]))
    assert abs(out["mol_inter"][1] - coul) <= 1e-10 * max(abs(coul), abs(tot[2]))
    d = o.mol_recip(mol, trial)
    assert abs(out["mol_recip"] - d) <= 1e-10 * max(abs(d), abs(tot[3]))
    # after the accept the reference energy moved by the delta, and the same
    # coordinates give no further change
    assert out["recip_after_accept"][0] == pytest.approx(tot[3] + d, rel=1e-11)
    assert abs(out["recip_after_accept"][1]) <= 1e-9 * abs(tot[3])
    # the rest runs after the commit: the oracle holds the committed molecule
    from dataclasses import replace
    xyz = s.xyz.copy()
    st = s.mol_start()[mol]
    xyz[st:st + 3] = trial
    oc = Oracle(replace(s, xyz=xyz, _keep=[]), threads=4)
    en, real, _ = oc.particle_inter(0, -0.8476, ptrials)
    assert np.allclose(np.array(out["particle_en_plus_1"]) - 1.0, en, rtol=1e-10, atol=1e-8)
    assert np.allclose(np.array(out["particle_real_plus_2"]) - 2.0, real, rtol=1e-10, atol=1e-8)
    cv, cs, _, _ = oc.cell_csr()
    assert np.array_equal(out["cell_vector"], cv) and np.array_equal(out["cell_start"], cs)
    assert out["bad_mol_error"] == -1
    # the committed molecule is part of the box: totals move by the trial's change
    assert np.allclose(out["total_after"], oc.system_total(), rtol=1e-11, atol=0)
