"""Pins the float transcendentals of dspbr-pt_amd/csrc/pin_math.hpp.

The kernels and the oracle's float build share this one implementation of
sin, cos, exp, log, pow, atan2 and acos (GLSL leaves their precision open, so
the reference itself fixes no bits). Pinned here by what the mathematics
fixes, independently of the code under test:

* every constant of the header is re-derived in exact arithmetic
  (gen_pin_math.py: Machin's pi, series for ln 2 and atan(k/8), exact
  Taylor coefficients) and must match the header's block bit for bit;
* on dense and random argument sweeps (every exponent for sin/cos, so both
  range reductions run) each result must be the float64 NumPy value rounded
  to float32 — correct rounding for >= 99.999 % of arguments and never more
  than one ulp away. A dropped series term, a wrong sign, a wrong split of
  pi/2 or a wrong Payne-Hanek window breaks this by many ulps;
* the C99 Annex F special cases (signed zeros, infinities, NaN) equal
  NumPy's, which follows C99;
* psincos returns exactly psin and pcos.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_ref as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dspbr-pt_amd", "csrc")


def _ulp_dist(a, b):
    """Distance in ulps between float32 arrays (NaN pairs count 0)."""
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-(2 ** 31)) - ia, ia)
    ib = np.where(ib < 0, np.int64(-(2 ** 31)) - ib, ib)
    d = np.abs(ia - ib)
    both_nan = np.isnan(a) & np.isnan(b)
    one_nan = np.isnan(a) ^ np.isnan(b)
    d = np.where(both_nan, 0, d)
    return np.where(one_nan, 2 ** 40, d)


def sweep_args(seed=7):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32)
    bits = bits[~np.isnan(bits)]
    trig = np.concatenate([
        np.linspace(-40.0, 40.0, 1 << 20, dtype=np.float32),
        bits,
        # floats next to multiples of pi/2, where reduction cancels most
        np.nextafter(np.float32(np.arange(1, 4000) * (np.pi / 2)), np.float32(np.inf)),
        np.float32(np.arange(1, 4000) * (np.pi / 2)),
        np.array([2 ** 25, np.nextafter(np.float32(2 ** 25), np.float32(0)), 3.4e38, 1e30, 1e10], np.float32),
    ]).astype(np.float32)
    exp = np.concatenate([
        np.linspace(-104.5, 89.5, 1 << 20, dtype=np.float32),
        rng.uniform(-5, 5, 1 << 18).astype(np.float32),
    ])
    pos = np.abs(bits)
    pos = pos[np.isfinite(pos) & (pos > 0)]
    log = np.concatenate([pos, np.linspace(0.5, 2.0, 1 << 20, dtype=np.float32),
                          np.array([1.0, np.float32(2 ** -149), 3.4e38], np.float32)])
    acos = np.concatenate([np.linspace(-1.0, 1.0, 1 << 20, dtype=np.float32),
                           1.0 - rng.uniform(0, 1e-3, 1 << 16).astype(np.float32)])
    px = np.concatenate([rng.uniform(0, 10, 1 << 20), rng.uniform(0, 1, 1 << 18), np.abs(rng.standard_normal(1 << 18)) * 1e-3])
    py = np.concatenate([rng.uniform(-12, 12, 1 << 20), rng.uniform(0, 10, 1 << 18), rng.uniform(-3, 3, 1 << 18)])
    # integer exponents, including negative bases
    px = np.concatenate([px, rng.uniform(-4, 4, 1 << 16)]).astype(np.float32)
    py = np.concatenate([py, rng.integers(-9, 10, 1 << 16)]).astype(np.float32)
    ay = np.concatenate([bits[: 1 << 19], rng.standard_normal(1 << 19)]).astype(np.float32)
    ax = np.concatenate([bits[-(1 << 19):], rng.standard_normal(1 << 19)]).astype(np.float32)
    return {"trig": trig, "exp": exp, "log": log, "acos": acos, "pow": (px, py), "atan2": (ay, ax)}


SPECIAL = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 0.5, -0.5, 2.0, -2.0, 3.0, -3.0,
                    1e-45, -1e-45, 1e30, -1e30, 16777216.0, -16777217.0, 0.3, 2.5, -2.5], np.float32)


def _np_ref(name, x, y=None):
    with np.errstate(all="ignore"):
        xd = x.astype(np.float64)
        if name == "sin":
            r = np.sin(xd)
        elif name == "cos":
            r = np.cos(xd)
        elif name == "exp":
            r = np.exp(xd)
        elif name == "log":
            r = np.log(xd)
        elif name == "acos":
            r = np.arccos(xd)
        elif name == "pow":
            r = np.power(xd, y.astype(np.float64))
        elif name == "atan2":
            r = np.arctan2(xd, y.astype(np.float64))
        return r.astype(np.float32)


def _check(name, x, y=None):
    got = O.pin_math(name, x, y)
    ref = _np_ref(name, x, y)
    d = _ulp_dist(got, ref)
    exact = (d == 0).mean()
    return got, ref, d, exact


def test_constants_match_generator():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_pin_math.py")], capture_output=True, text=True,
                         check=True).stdout.strip()
    hdr = open(os.path.join(CSRC, "pin_math.hpp")).read()
    a = hdr.index("// ---- BEGIN GENERATED CONSTANTS")
    b = hdr.index("// ---- END GENERATED CONSTANTS ----") + len("// ---- END GENERATED CONSTANTS ----")
    assert hdr[a:b] == out


def test_generator_pi_and_ln2_against_known_values():
    sys.path.insert(0, CSRC)
    import gen_pin_math as G
    from fractions import Fraction
    pi = Fraction(G.pi_fixed(), 1 << G.PREC)
    # 50 decimal digits of pi and ln 2
    assert abs(pi - Fraction("3.14159265358979323846264338327950288419716939937510")) < Fraction(1, 10 ** 49)
    ln2 = Fraction(G.ln2_fixed(), 1 << G.PREC)
    assert abs(ln2 - Fraction("0.69314718055994530941723212145817656807550013436025")) < Fraction(1, 10 ** 49)
    # atan(1) = pi/4 through the Euler series
    assert abs(Fraction(G.atan_frac_fixed(8, 8), 1 << G.PREC) - pi / 4) < Fraction(1, 2 ** 600)


@pytest.mark.parametrize("name,key", [("sin", "trig"), ("cos", "trig"), ("exp", "exp"), ("log", "log"),
                                      ("acos", "acos"), ("pow", "pow"), ("atan2", "atan2")])
def test_correctly_rounded_on_sweeps(name, key):
    a = sweep_args()[key]
    x, y = (a if isinstance(a, tuple) else (a, None))
    got, ref, d, exact = _check(name, x, y)
    assert d.max() <= 1, (name, x[d.argmax()], None if y is None else y[d.argmax()], got[d.argmax()], ref[d.argmax()])
    assert exact >= 0.99999, (name, exact)


@pytest.mark.parametrize("name", ["sin", "cos", "exp", "log", "acos"])
def test_special_values_unary(name):
    got = O.pin_math(name, SPECIAL)
    ref = _np_ref(name, SPECIAL)
    assert (np.isnan(got) == np.isnan(ref)).all()
    fin = ~np.isnan(ref)
    # bit patterns, so signed zeros and infinities count
    assert (_ulp_dist(got[fin], ref[fin]) <= 1).all()
    zero = fin & (ref == 0)
    assert (np.signbit(got[zero]) == np.signbit(ref[zero])).all()


@pytest.mark.parametrize("name", ["pow", "atan2"])
def test_special_values_binary(name):
    x, y = np.meshgrid(SPECIAL, SPECIAL)
    x, y = x.ravel().copy(), y.ravel().copy()
    got = O.pin_math(name, x, y)
    ref = _np_ref(name, x, y)
    assert (np.isnan(got) == np.isnan(ref)).all(), [(a, b, g, r) for a, b, g, r in zip(x, y, got, ref)
                                                      if np.isnan(g) != np.isnan(r)][:5]
    fin = ~np.isnan(ref)
    d = _ulp_dist(got[fin], ref[fin])
    assert (d <= 1).all()
    zero = fin & (ref == 0)
    assert (np.signbit(got[zero]) == np.signbit(ref[zero])).all()


def test_sincos_equals_sin_and_cos():
    x = sweep_args()["trig"]
    s, c = O.pin_math("sincos_s", x), O.pin_math("sincos_c", x)
    assert np.array_equal(s.view(np.uint32), O.pin_math("sin", x).view(np.uint32))
    assert np.array_equal(c.view(np.uint32), O.pin_math("cos", x).view(np.uint32))
