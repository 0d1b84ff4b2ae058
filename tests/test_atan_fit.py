"""The float32 arctangent of the Lambert map (atan_unit in csrc/lambert.h):
its coefficients are the ones fit_atan_unit.py derives, and their device
evaluation (Horner with fused multiply-adds, emulated exactly) stays within
the stated ulp bound of atan over [-1, 1]."""

import importlib.util
import os
import re
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ebsdtorch_amd", "csrc")
BOUND_ULP = 1.415  # the exhaustive maximum over (0, 1] (1.4144 ulp at t = 0.99157), as lambert.h states


def _fit():
    spec = importlib.util.spec_from_file_location("fit_atan_unit", os.path.join(CSRC, "fit_atan_unit.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


FIT = _fit()


def source_coefficients():
    txt = open(os.path.join(CSRC, "lambert.h")).read()
    body = txt[txt.index("float atan_unit(float t)"):]
    body = body[:body.index("return p * t;")]
    vals = re.findall(r"(-?\d\.\d+e[-+]\d+)f", body)
    return np.array([float(v) for v in vals], dtype=np.float32)


def test_source_holds_the_fitted_coefficients():
    src = source_coefficients()
    fit = FIT.coefficients()
    assert len(src) == FIT.DEGREE + 1
    # equal as float32 (the fit's float64 solve may differ by round-off in
    # the last float64 bits on another machine: allow one float32 step)
    steps = np.abs(src.view(np.int32).astype(np.int64) - fit.view(np.int32).astype(np.int64))
    assert steps.max() <= 1, (src, fit)


def test_fit_is_minimax():
    """The float64 Remez solution equioscillates: DEGREE + 2 alternating
    extrema of equal size, the characterisation of the minimax polynomial."""
    c = FIT.remez()
    grid = np.linspace(0, 1, 200_001)
    err = np.polynomial.polynomial.polyval(grid, c) - FIT.target(grid)
    peak = np.abs(err).max()
    assert peak < 4e-10
    sign = np.sign(err)
    runs = np.split(np.arange(len(grid)), np.flatnonzero(np.diff(sign) != 0) + 1)
    peaks = np.array([np.abs(err[r]).max() for r in runs])
    assert len(runs) == FIT.DEGREE + 2
    assert peaks.min() > 0.999 * peak


def test_fma_emulation_rounds_once():
    """fma32 against exact rational arithmetic, including sums that land on a
    float32 midpoint in float64."""
    rng = np.random.default_rng(0)
    a = rng.standard_normal(3000).astype(np.float32)
    b = rng.standard_normal(3000).astype(np.float32)
    c = (-(a.astype(np.float64) * b) * (1 + rng.standard_normal(3000) * 1e-7)).astype(np.float32)
    # constructed midpoints: a * b = 1 + 2^-24 (+/- 2^-48 tail from c)
    a2 = np.full(4, 1 + 2.0 ** -12, dtype=np.float32)
    b2 = np.full(4, 1 + 2.0 ** -12, dtype=np.float32)  # product 1 + 2^-11 + 2^-24
    c2 = np.array([0.0, -2.0 ** -11, -(2.0 ** -11) - 2.0 ** -40, -(2.0 ** -11) + 2.0 ** -40], dtype=np.float32)
    a, b, c = np.concatenate([a, a2]), np.concatenate([b, b2]), np.concatenate([c, c2])
    got = FIT.fma32(a, b, c)
    for i in range(len(a)):
        x = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        f = np.float32(float(x))
        cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
        best = min(cands, key=lambda v: (abs(Fraction(float(v)) - x), int(np.float32(v).view(np.uint32)) & 1))
        assert got[i] == best, (i, a[i], b[i], c[i])


def test_ulp_bound_over_unit_interval():
    coefs = source_coefficients()
    worst, _ = FIT.max_ulp_error(coefs, stride=211)
    assert worst <= BOUND_ULP
    # every float32 in the last 2^20 below 1, where the error peaks
    top = int(np.float32(1.0).view(np.uint32))
    t = np.arange(top - (1 << 20), top + 1, dtype=np.uint32).view(np.float32)
    assert FIT.ulp_error(coefs, t).max() <= BOUND_ULP
    # tiny arguments return t itself (P(0) = 1)
    tiny = np.array([1e-30, 1e-20, 1e-10, 2.0 ** -126, 2.0 ** -149], dtype=np.float32)
    assert np.array_equal(FIT.atan_unit(coefs, tiny), tiny)


def test_odd_symmetry():
    coefs = source_coefficients()
    t = np.linspace(0, 1, 100_001, dtype=np.float32)
    assert np.array_equal(FIT.atan_unit(coefs, -t), -FIT.atan_unit(coefs, t))


@pytest.mark.parametrize("mistake", ["dropped term", "sign", "transposed"])
def test_bound_catches_a_wrong_table(mistake):
    """A plausible slip in the table breaks the bound by far."""
    coefs = source_coefficients().copy()
    if mistake == "dropped term":
        coefs = coefs[1:]
    elif mistake == "sign":
        coefs[3] = -coefs[3]
    else:
        coefs[[4, 5]] = coefs[[5, 4]]
    t = np.linspace(0, 1, 10_001, dtype=np.float32)
    assert FIT.ulp_error(coefs, t).max() > 10 * BOUND_ULP
