"""Pins the oracle's generators.

* Philox4x64-10, u01, uneg11, Random123Wrapper: against vectors printed by
  the reference's own sources compiled here (tests/golden/ref_rng_kat.json,
  written by tests/golden/make_ref_vectors.py); its first three Philox
  cases are also Random123's published known-answer vectors.
* MT19937 / MTRand: against the same fixture, against numpy's legacy
  RandomState (init_genrand seeding), and against the C++ standard's known
  answer (10000th output of a default-seeded mt19937 is 4123659995).
* Rotation matrices: against the fixture and the axis-angle invariants.
"""
import json
from pathlib import Path

import numpy as np

from oracle import pyoracle as po

GOLD = json.loads((Path(__file__).parent / "golden" / "ref_rng_kat.json").read_text())


def test_philox_published_kat():
    pub = [
        ([0, 0, 0, 0], [0, 0], ["16554d9eca36314c", "db20fe9d672d0fdc", "d7e772cee186176b", "7e68b68aec7ba23b"]),
        ([2**64 - 1] * 4, [2**64 - 1] * 2,
         ["87b092c3013fe90b", "438c3c67be8d0224", "9cc7d7c69cd777b6", "a09caebf594f0ba0"]),
        ([0x243f6a8885a308d3, 0x13198a2e03707344, 0xa4093822299f31d0, 0x082efa98ec4e6c89],
         [0x452821e638d01377, 0xbe5466cf34e90c6c],
         ["a528f45403e61d95", "38c72dbd566e9788", "a5a1610e72fd18b5", "57bd43b5e52b7fe6"]),
    ]
    for ctr, key, out in pub:
        got = po.philox(ctr, key)
        assert [f"{int(v):016x}" for v in got] == out


def test_philox_and_uniform_against_reference_build():
    for case in GOLD["philox4x64_10"]:
        ctr = [int(v, 16) for v in case["ctr"]]
        key = [int(v, 16) for v in case["key"]]
        got = po.philox(ctr, key)
        assert [f"{int(v):016x}" for v in got] == case["out"]
        assert po.u01(got[0]) == case["u01"][0]
        assert po.u01(got[1]) == case["u01"][1]
        assert po.uneg11(got[2]) == case["uneg11"][0]
        assert po.uneg11(got[3]) == case["uneg11"][1]


def test_uniform_closed_forms():
    assert po.u01(0) == 2.0**-65
    assert po.u01(2**64 - 1) == 1.0
    assert po.uneg11(0) == 2.0**-64
    assert po.uneg11(2**63 - 1) == 1.0
    assert po.uneg11(2**63) == -1.0 + 2.0**-64


def test_random123_wrapper_against_reference_build():
    for c in GOLD["wrapper"]:
        o = po.r123(c["seed"], c["step"], c["counter"], 0.75)
        assert o[0] == c["rand"]
        assert list(o[1:4]) == c["sym"]
        assert list(o[4:7]) == c["sphere"]
        assert o[7] == c["symr"]


def test_mt19937_against_reference_build_and_numpy():
    for c in GOLD["mtrand"]:
        raw = po.mt_raw(c["seed"], 700)
        assert raw.tolist() == c["raw"]
        rs = np.random.RandomState(c["seed"])
        assert np.array_equal(rs.randint(0, 2**32, size=700, dtype=np.uint64).astype(np.uint32), raw)
        assert po.mt_rand_int(c["seed"], 64, 999).tolist() == c["randInt_999"]
        assert list(raw[:8] * (1.0 / 4294967295.0)) == c["rand"]
        assert list(raw[:8] * (1.0 / 4294967296.0)) == c["randExc"]


def test_mt19937_standard_known_answer():
    assert int(po.mt_raw(5489, 10000)[-1]) == 4123659995 == GOLD["mt_default_10000th"]


def test_rand_int_is_uniform_rejection():
    v = po.mt_rand_int(7, 200000, 999)
    assert v.max() <= 999 and v.min() >= 0
    counts = np.bincount(v, minlength=1000)
    assert abs(counts.mean() - 200.0) < 1e-9 and counts.std() < 3 * np.sqrt(200)


def test_rotations_against_reference_build_and_invariants():
    for c in GOLD["rotation"]:
        if "theta" in c:
            m = po.rotation(c["theta"], c["axis"])
            assert m.ravel().tolist() == c["m"]
            ax = np.array(c["axis"], dtype=float)
            ax = ax / np.linalg.norm(ax)
            assert np.allclose(m @ m.T, np.eye(3), atol=1e-14)
            assert np.allclose(m @ ax, ax, atol=1e-14)
            assert abs(np.trace(m) - (1 + 2 * np.cos(c["theta"]))) < 1e-14
        else:
            m = po.uniform_rotation(*c["uniform"])
            assert m.ravel().tolist() == c["m"]
            assert np.allclose(m @ m.T, np.eye(3), atol=1e-14)
            assert abs(np.linalg.det(m) - 1.0) < 1e-14
