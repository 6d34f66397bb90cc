"""A run continued from GOMC's restart set (gomc_amd/gomc_chk.py): the
cereal-layout `_restart.chk`, the binary restart coordinates and the cell
files, written from a box and read back into a new one.

As GOMC's own restart (CheckpointSetup::SetCheckpointData), the restored box
takes the MT19937 state, move settings, MultiParticle settings and
MoleculeLookup from the archive and the coordinates from the restart files
(unwrapped around each molecule's centre, then wrapped back), and computes
its energies anew.  Bar: the fields read back equal the fields written; the
next moves have the same kind, molecule, overlap and accept sequences as the
uninterrupted run, per-move energy changes within 1e-10 of the larger of
|dE| and the total, running energies within 1e-10 relative, coordinates
within 1e-12 A -- for NVT water with MultiParticle moves (host driver) and
GCMC methane/CO2 (device move loop, reservoir in box 1).  A checkpoint of
another system is refused.
"""
import numpy as np
import pytest

from gomc_amd import gomc_chk, systems
from gomc_amd.api import Box

pytestmark = pytest.mark.gpu


def _same_fields(a, b, path="chk"):
    if isinstance(a, dict):
        assert set(a) == set(b), (path, set(a) ^ set(b))
        for k in a:
            _same_fields(a[k], b[k], f"{path}.{k}")
    elif isinstance(a, (list, tuple, np.ndarray)) and not isinstance(a, (str, bytes)):
        assert len(a) == len(b), (path, len(a), len(b))
        for i, (x, y) in enumerate(zip(a, b)):
            _same_fields(x, y, f"{path}[{i}]")
    else:
        assert a == b, (path, a, b)


def _continue(s, n1, n2, tmp_path, device=False):
    topo = gomc_chk.GomcTopology.from_system(s)
    a = Box(s)
    a.run(n1, device=device)
    prefix = str(tmp_path / "run")
    gomc_chk.write_restart(prefix, a, topo, n1)
    written = gomc_chk.checkpoint_of(a, topo, n1)
    rs = gomc_chk.read_restart(prefix, s.ensemble)
    _same_fields(written, rs["checkpoint"])
    b, step = gomc_chk.restore_box(s, rs)
    assert step == n1
    na, ra = a.run(n2, start_step=n1, device=device)
    nb, rb = b.run(n2, start_step=step, device=device)
    for key in ("kind", "mol", "ovl", "acc"):
        assert np.array_equal(ra[key], rb[key]), (key, np.nonzero(ra[key] != rb[key])[0][:5])
    tot = np.abs(a.energies()).max()
    for key in ("dlj", "dreal", "drecip"):
        assert np.all(np.abs(ra[key] - rb[key]) <= 1e-10 * np.maximum(np.abs(ra[key]), tot)), key
    assert na == nb
    assert np.allclose(a.energies(), b.energies(), rtol=1e-10, atol=1e-6)
    assert np.allclose(a.coords(), b.coords(), rtol=0, atol=1e-12)
    assert np.allclose(a.scales()[0], b.scales()[0], rtol=1e-15, atol=0)
    return a, b, ra


def test_nvt_water_with_multiparticle(tmp_path):
    s = systems.water_box(200, seed=21, adj_steps=100, move_disp=0.25, move_rot=0.25,
                          move_mp=0.5)
    a, b, ra = _continue(s, 400, 400, tmp_path)
    assert (ra["kind"] == 2).sum() > 50 and ra["acc"].sum() > 50
    assert np.array_equal(a.mp_settings(), b.mp_settings())


def test_gcmc_device_loop(tmp_path):
    s = systems.gcmc_mixture(60, 60, 25.0, fugacity_bar=(200.0, 60.0), reservoir=120, seed=3,
                             adj_steps=250)
    a, b, ra = _continue(s, 1500, 1500, tmp_path, device=True)
    assert ra["acc"][ra["kind"] == 10].sum() > 10
    assert np.array_equal(a.mol_boxes()[1], b.mol_boxes()[1])


def test_other_system_refused(tmp_path):
    s = systems.water_box(150, seed=5)
    a = Box(s)
    a.run(100)
    prefix = str(tmp_path / "w")
    gomc_chk.write_restart(prefix, a, gomc_chk.GomcTopology.from_system(s), 100)
    rs = gomc_chk.read_restart(prefix, 0)
    with pytest.raises(gomc_chk.GomcCheckpointError):
        gomc_chk.restore_box(systems.water_box(151, seed=5), rs)
