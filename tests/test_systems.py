"""Synthetic inputs: placement honours its exclusion distances, is seeded,
and the descriptor layout matches the C struct."""
import ctypes

import numpy as np

from gomc_amd import systems


def _min_dists(s):
    L = s.box[0]
    xyz = s.xyz
    starts = s.mol_start()
    lens = [len(s.kinds[k].atom_types) for k in s.mol_kind]
    mol_of = np.repeat(np.arange(s.nmols), lens)
    heavy = np.zeros(s.natoms, bool)
    heavy[starts] = True
    d = xyz[:, None, :] - xyz[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))
    other = mol_of[:, None] != mol_of[None, :]
    hh = heavy[:, None] & heavy[None, :] & other
    return r[hh].min(), r[other].min()


def test_water_placement_exclusions_and_seed():
    s = systems.water_box(400, seed=5)
    assert s.xyz.min() >= 0.0 and s.xyz.max() < s.box[0]
    hh, anyp = _min_dists(s)
    assert hh >= 2.5 and anyp >= 1.5
    s2 = systems.water_box(400, seed=5)
    assert np.array_equal(s.xyz, s2.xyz)
    # rigid geometry preserved (O-H 1.0 A, H-O-H 109.47 deg)
    L = s.box[0]
    o, h1, h2 = s.xyz[0], s.xyz[1], s.xyz[2]
    v1 = h1 - o - L * np.round((h1 - o) / L)
    v2 = h2 - o - L * np.round((h2 - o) / L)
    assert abs(np.linalg.norm(v1) - 1.0) < 1e-12 and abs(np.linalg.norm(v2) - 1.0) < 1e-12
    ang = np.degrees(np.arccos(v1 @ v2))
    assert abs(ang - 109.47) < 1e-9


def test_config4_sizes():
    boxes = systems.config4(16, seed=3)
    n = np.array([b.nmols for b in boxes])
    assert n.min() >= 200 and n.max() <= 4000
    for b in boxes:
        assert b.rcut <= b.box[0] / 2 and 280 <= b.temperature <= 360


def test_desc_layout():
    # field offsets used by the C side (include/gomc_hip.h)
    assert systems.SystemDesc.seed.offset == 176
    assert systems.SystemDesc.ensemble.offset == 216
    assert systems.SystemDesc.mol_box.offset == 232
    assert systems.SystemDesc.move_swap.offset == 248
    assert systems.SystemDesc.pressure.offset == 256
    assert systems.SystemDesc.move_vol.offset == 264
    assert ctypes.sizeof(systems.SystemDesc) == 272
