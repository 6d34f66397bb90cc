"""Systems read from GOMC input files run on the device with the oracle's
results: a TIP3 water lattice whose force field comes from the CHARMM
stream file of the reference's BPTI example (Ewald on), and the AR_KR
grand-canonical example with its reservoir box and chemical potentials
(transfer frequency raised so that transfers happen in a short run).

Bar (FP64): box totals within 1e-11 relative; move-kind, molecule, overlap
and accept sequences identical; per-move energy changes within 1e-10 of the
box totals; coordinates within 1e-9 A; running energies within 1e-10.
"""
import os
import shutil

import numpy as np
import pytest

from gomc_amd import gomc_io
from gomc_amd.api import Box
from oracle.pyoracle import Oracle
from test_gomc_io import ARKR, tip3_lattice

pytestmark = pytest.mark.gpu


def dE_close(a, b, total, rtol=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), abs(total))))


def same_run(s, n, device):
    b, o = Box(s), Oracle(s, threads=8)
    assert np.allclose(b.system_total(), o.system_total(), rtol=1e-11, atol=1e-9)
    e0 = o.energies()
    ng, rg = b.run(n, device=device)
    no, ro = o.run(n, mode=0)
    for key in ("kind", "mol", "ovl", "acc"):
        assert np.array_equal(rg[key], ro[key]), (key, np.nonzero(rg[key] != ro[key])[0][:5])
    assert ng == no
    assert dE_close(rg["dlj"], ro["dlj"], e0[0])
    assert dE_close(rg["dreal"], ro["dreal"], e0[2])
    ok = ro["ovl"] == 0
    assert dE_close(rg["drecip"][ok], ro["drecip"][ok], e0[3])
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)
    return ro


@pytest.mark.parametrize("device", [False, True])
def test_tip3_lattice_from_files(tmp_path, device):
    s = gomc_io.system_from_gomc(tip3_lattice(tmp_path), threads=8)
    ro = same_run(s, 2000, device)
    assert ro["acc"].sum() > 200


def ar_kr_gcmc(tmp_path, n_res=1000):
    """The AR_KR example with its reservoir (box 1, 65 A) refilled with
    `n_res` atoms, alternately AR and KR on a lattice: the example's three
    reservoir atoms run out after a few insertions, and GOMC stops a GCMC
    run whose reservoir is empty."""
    base = tmp_path / "Base"
    shutil.copytree(os.path.join(ARKR, "Base"), base)
    shutil.copy(os.path.join(ARKR, "Par_Vrabec01_Noble_Gases_Exotic.inp"), tmp_path)
    side = round(n_res ** (1 / 3))
    g = (np.arange(side) + 0.5) * (65.0 / side)
    xyz = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    res = ["AR" if i % 2 == 0 else "KR" for i in range(len(xyz))]
    ids = np.arange(1, len(xyz) + 1)
    gomc_io.write_pdb(str(base / "START_BOX_1.pdb"), xyz, res, res, ids, cell=(65.0,) * 3,
                      box_index=1)
    mass = [39.948 if r == "AR" else 83.798 for r in res]
    gomc_io.write_psf(str(base / "START_BOX_1.psf"), res, ids, res, res, res,
                      [0.0] * len(xyz), mass, np.zeros((0, 2), np.int64))
    conf = base / "in.conf"
    # more transfers than the example's 0.1 %, displacements lowered so that
    # the frequencies still sum to one
    txt = conf.read_text().replace("SwapFreq\t      0.001", "SwapFreq\t      0.3")
    conf.write_text(txt.replace("DisFreq               0.87", "DisFreq               0.571"))
    return str(conf)


@pytest.mark.parametrize("device", [False, True])
def test_ar_kr_gcmc_from_files(tmp_path, device):
    s = gomc_io.system_from_gomc(ar_kr_gcmc(tmp_path))
    assert s.move_swap == 0.3 and s.nmols == 3 + 1000 and s.ensemble == 1
    ro = same_run(s, 3000, device)
    k = ro["kind"]
    assert (k == 10).sum() > 500 and ro["acc"][k == 10].sum() > 5
