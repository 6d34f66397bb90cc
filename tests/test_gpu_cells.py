"""The trial cell table of the plain trial path.

In a box of more than 27 cells gomc_trial walks the cells within one cell
of each old and new position of the moving molecule, and
gomc_particle_inter the 27 cells around each CBMC trial position, as the
reference's MoleculeInter and ParticleInter walk CellList::EnumerateLocal
(src/CalculateEnergy.cpp:595,745).  The table they walk is edited by every
accepted displacement (CellList::RemoveMol/AddMol, src/CellList.h:28-29)
and rebuilt after anything else moves molecules or changes which molecules
live in the box.  The oracle scans every atom, so agreement with it is
agreement with the full scan.

* 2200 waters in a 40.4 A box (4 x 4 x 4 cells): the kind, molecule,
  overlap and accept sequences of 10^4 translate/rotate moves equal the
  oracle's; per-move energy changes within 1e-10 of the larger of |dE| and
  the box total; then the table holds every atom once, in the cell of its
  position (CellList::PositionToCell, src/CellList.h:88-101).
* GCMC methane / CO2 in a 44 A box: host-driver moves (transfers with CBMC
  growth on the table) alternated with device-loop runs, the whole sequence
  equal to the oracle's; the table then holds exactly the molecules in the
  box, each atom in the cell of its position.
* In a box of 27 cells there is no table: the trials scan every atom, and
  the read-back reports GOMC_ERR_STATE.
"""
import numpy as np
import pytest

from gomc_amd import systems
from gomc_amd.api import Box, GomcError
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def same_records(rg, ro, e0):
    for key in ("kind", "mol", "ovl", "acc"):
        assert np.array_equal(rg[key], ro[key]), (key, np.nonzero(rg[key] != ro[key])[0][:5])
    for key, e in (("dlj", e0[0]), ("dreal", e0[2])):
        a, b = rg[key], ro[key]
        assert np.all(np.abs(a - b) <= 1e-10 * np.maximum(np.abs(b), abs(e))), key
    ok = ro["ovl"] == 0
    a, b = rg["drecip"][ok], ro["drecip"][ok]
    assert np.all(np.abs(a - b) <= 1e-10 * np.maximum(np.abs(b), abs(e0[3]))), "drecip"


def cells_of(xyz, box, edge):
    """CellList::PositionToCell over an (n, 3) array of wrapped positions."""
    idx = []
    for d in range(3):
        i = (xyz[:, d] / (box[d] / edge[d])).astype(np.int64)
        i[i == edge[d]] -= 1
        idx.append(np.clip(i, 0, edge[d] - 1))
    return idx[0] * edge[1] * edge[2] + idx[1] * edge[2] + idx[2]


def check_table(b, live_atoms):
    cnt, slots = b.trial_cells()
    held = np.concatenate([slots[c, :cnt[c]] for c in range(len(cnt))])
    assert np.array_equal(np.sort(held), np.sort(live_atoms)), "each live atom once, no other"
    want = cells_of(b.coords()[live_atoms], b.system.box, b.edge)
    where = np.empty(b.natoms, dtype=np.int64)
    for c in range(len(cnt)):
        where[slots[c, :cnt[c]]] = c
    assert np.array_equal(where[live_atoms], want), "every atom in the cell of its position"


def test_nvt_water_64_cells_same_sequence_as_oracle():
    s = systems.water_box(2200, seed=31, adj_steps=500)
    b, o = Box(s), Oracle(s, threads=8)
    assert b.ncells == 64
    e0 = o.energies()
    n = 10000
    ng, rg = b.run(n)
    no, ro = o.run(n, mode=0)
    same_records(rg, ro, e0)
    assert ng == no and ng > 2000
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=0)
    assert np.allclose(b.coords(), o.coords(), rtol=0, atol=1e-9)
    check_table(b, np.arange(b.natoms))


def test_gcmc_64_cells_host_and_device_drivers_interleave():
    s = systems.gcmc_mixture(300, 300, 44.0, fugacity_bar=(60.0, 25.0), reservoir=600, seed=9,
                             adj_steps=200)
    b, o = Box(s), Oracle(s, threads=8)
    assert b.ncells == 64
    e0 = o.energies()
    parts = [(0, 1500, False), (1500, 1000, True), (2500, 1500, False), (4000, 500, True),
             (4500, 1500, False)]
    recs = []
    for start, n, dev in parts:
        _, r = b.run(n, start_step=start, device=dev)
        recs.append(r)
    rg = {k: np.concatenate([r[k] for r in recs]) for k in recs[0]}
    _, ro = o.run(6000, mode=0)
    assert ((ro["kind"] == 10) & (ro["acc"] == 1)).sum() > 20
    same_records(rg, ro, e0)
    n1, box1 = b.mol_boxes()
    n2, box2 = o.mol_boxes()
    assert n1 == n2 and np.array_equal(box1, box2)
    assert np.allclose(b.energies(), o.energies(), rtol=1e-10, atol=1e-6)
    ms = np.append(b.mol_start, b.natoms)
    live = np.concatenate([np.arange(ms[m], ms[m + 1]) for m in np.nonzero(box1 == 0)[0]])
    check_table(b, live)


def test_box_of_27_cells_has_no_table():
    b = Box(systems.water_box(300, seed=5))
    assert b.ncells == 27
    with pytest.raises(GomcError):
        b.trial_cells()
