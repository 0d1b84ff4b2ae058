"""Writes the coefficients of atan_unit in lambert.h and measures its error.

atan(t) for |t| <= 1 is evaluated as t * P(t^2) with P of degree 10. P is
the minimax (least maximum absolute error) polynomial for
f(s) = atan(sqrt(s)) / sqrt(s) on s in [0, 1], found by the Remez exchange
algorithm in float64 and then rounded to float32. The device code evaluates
P by Horner's rule with fused multiply-adds and multiplies by t once.

    python fit_atan_unit.py            # print the coefficient lines
    python fit_atan_unit.py --check    # also the max error over every float32 in [0, 1]

The error is measured in units in the last place of the float32 value of
atan(t), with each fused multiply-add rounded once to float32 as the device
does (emulated exactly in float64: the product of two float32 values is
exact there, and the one double-rounding case, a float64 sum that lands on a
float32 midpoint, is resolved with the sum's exact error term).
"""

from __future__ import annotations

import argparse

import numpy as np

DEGREE = 10


def target(s: np.ndarray) -> np.ndarray:
    """atan(sqrt(s)) / sqrt(s), with its limit 1 at s = 0."""
    s = np.asarray(s, dtype=np.float64)
    t = np.sqrt(s)
    out = np.ones_like(s)
    nz = t > 0
    out[nz] = np.arctan(t[nz]) / t[nz]
    return out


def remez(degree: int = DEGREE, grid_points: int = 200_001, iterations: int = 40) -> np.ndarray:
    """Minimax coefficients (constant term first) of the absolute error of a
    degree-`degree` polynomial against target() on [0, 1]."""
    n = degree + 2
    x = 0.5 - 0.5 * np.cos(np.pi * np.arange(n) / (n - 1))  # Chebyshev extrema
    grid = np.linspace(0.0, 1.0, grid_points)
    fg = target(grid)
    alternate = (-1.0) ** np.arange(n)
    for _ in range(iterations):
        system = np.hstack([np.vander(x, degree + 1, increasing=True), alternate[:, None]])
        sol = np.linalg.solve(system, target(x))
        coefs = sol[:-1]
        err = np.polynomial.polynomial.polyval(grid, coefs) - fg
        # one extremum per run of constant sign
        sign = np.sign(err)
        cuts = np.flatnonzero(np.diff(sign) != 0) + 1
        runs = np.split(np.arange(grid_points), cuts)
        ext = np.array([r[np.argmax(np.abs(err[r]))] for r in runs])
        while len(ext) > n:  # drop the smaller end
            ext = ext[1:] if abs(err[ext[0]]) < abs(err[ext[-1]]) else ext[:-1]
        if len(ext) < n:
            raise RuntimeError("the Remez exchange lost an alternation point")
        x = grid[ext]
    return coefs


def coefficients() -> np.ndarray:
    """float32 coefficients of P, highest power first (Horner order)."""
    return remez()[::-1].astype(np.float32)


def fma32(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """float32 a * b + c rounded once (round to nearest, ties to even)."""
    p = a.astype(np.float64) * b.astype(np.float64)  # exact
    c64 = c.astype(np.float64)
    s = p + c64
    bb = s - p
    e = (p - (s - bb)) + (c64 - bb)  # s + e == p + c exactly
    r = s.astype(np.float32)
    up = np.nextafter(r, np.float32(np.inf))
    dn = np.nextafter(r, np.float32(-np.inf))
    r64 = r.astype(np.float64)
    r = np.where((s == (r64 + up.astype(np.float64)) * 0.5) & (e > 0), up, r)
    r = np.where((s == (r64 + dn.astype(np.float64)) * 0.5) & (e < 0), dn, r)
    return r


def atan_unit(coefs: np.ndarray, t: np.ndarray) -> np.ndarray:
    """The device evaluation of atan_unit on float32 t."""
    t = np.asarray(t, dtype=np.float32)
    s = t * t
    p = np.full_like(t, coefs[0])
    for c in coefs[1:]:
        p = fma32(p, s, np.full_like(t, c))
    return p * t


def ulp_error(coefs: np.ndarray, t: np.ndarray) -> np.ndarray:
    """|atan_unit(t) - atan(t)| in float32 ulps of atan(t)."""
    got = atan_unit(coefs, t).astype(np.float64)
    ref = np.arctan(np.asarray(t, dtype=np.float64))
    return np.abs(got - ref) / np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)


def max_ulp_error(coefs: np.ndarray, stride: int = 1, chunk: int = 1 << 24) -> tuple[float, float]:
    """Max ulp error (and where) over the float32 values in (0, 1], every
    `stride`-th bit pattern."""
    top = int(np.float32(1.0).view(np.uint32))
    worst, at = 0.0, 0.0
    for lo in range(1, top + 1, chunk * stride):
        bits = np.arange(lo, min(lo + chunk * stride, top + 1), stride, dtype=np.uint32)
        t = bits.view(np.float32)
        e = ulp_error(coefs, t)
        i = int(np.argmax(e))
        if e[i] > worst:
            worst, at = float(e[i]), float(t[i])
    return worst, at


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--check", action="store_true", help="max ulp error over every float32 in (0, 1]")
    args = ap.parse_args()
    c = coefficients()
    print(f"  float p = {c[0]:.8e}f;")
    for v in c[1:]:
        print(f"  p = fmaf(p, s, {v:.8e}f);")
    if args.check:
        worst, at = max_ulp_error(c)
        print(f"max error {worst:.4f} ulp at t = {at!r}")


if __name__ == "__main__":
    main()
