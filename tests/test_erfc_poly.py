"""The batched kernel's FP64 erfc on [0, 4]: exp(-x^2) times a fitted
polynomial for erfcx (gomc_amd/csrc/erfcx_poly.hpp).  Pinned against
scipy.special.erfc on a dense grid (relative error below 5e-15), at
erfc(0) = 1 and against the series erfc(x) = 1 - 2x/sqrt(pi) + 2x^3/(3
sqrt(pi)) - ... near 0.  The kernel evaluates the same Horner form in
double precision."""
import math
import re
from pathlib import Path

import numpy as np
from scipy.special import erfc

HDR = Path(__file__).resolve().parents[1] / "gomc_amd" / "csrc" / "erfcx_poly.hpp"


def _coeffs():
    text = HDR.read_text()
    body = text[text.index("kErfcxPoly"):]
    vals = re.findall(r"[-+]?\d\.\d+e[-+]\d+", body)
    deg = int(re.search(r"kErfcxDeg = (\d+)", text).group(1))
    assert len(vals) == deg + 1
    return np.array([float(v) for v in vals])


def _consts():
    text = HDR.read_text()
    mid = float(re.search(r"kErfcxMid = ([0-9.e+-]+)", text).group(1))
    inv = float(re.search(r"kErfcxInvHalf = ([0-9.e+-]+)", text).group(1))
    use = float(re.search(r"kErfcxUse = ([0-9.e+-]+)", text).group(1))
    return mid, inv, use


def _erfc_poly(x):
    m = _coeffs()
    mid, inv, _ = _consts()
    t = (x - mid) * inv
    p = np.full_like(x, m[-1])
    for c in m[-2::-1]:
        p = p * t + c
    return np.exp(-(x * x)) * p


def test_erfc_poly_matches_scipy_on_range():
    assert _consts()[2] == 4.0
    x = np.linspace(0.0, 4.0, 400001)
    rel = np.abs(_erfc_poly(x) - erfc(x)) / erfc(x)
    assert rel.max() < 5e-15, rel.max()


def test_erfc_poly_closed_values():
    assert abs(_erfc_poly(np.array([0.0]))[0] - 1.0) < 2e-15
    x = np.array([1e-6, 1e-4])
    series = 1.0 - 2.0 * x / math.sqrt(math.pi) + 2.0 * x ** 3 / (3.0 * math.sqrt(math.pi))
    assert np.allclose(_erfc_poly(x), series, rtol=3e-15, atol=0)
