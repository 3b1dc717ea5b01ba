"""The pinned transcendentals give the same bits on the GPU as in the
oracle's float build (pin_math.hpp compiled by hipcc for gfx950 and by gcc
for the host), over the argument sweeps of test_pin_math.py: every
exponent for sin/cos (both range reductions), exp/log over their whole
finite range, pow with negative bases and integer exponents, atan2 over
random bit patterns, and the C99 special values."""
import numpy as np
import pytest

import oracle_ref as O
from test_pin_math import SPECIAL, sweep_args

pytestmark = pytest.mark.gpu

FNS = [("sin", 0, "trig"), ("cos", 1, "trig"), ("exp", 2, "exp"), ("log", 3, "log"), ("pow", 4, "pow"),
       ("atan2", 5, "atan2"), ("acos", 6, "acos"), ("sincos_s", 7, "trig"), ("sincos_c", 8, "trig")]


def _same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint32) == b.view(np.uint32)) | nan


@pytest.mark.parametrize("name,fn,key", FNS)
def test_device_bits_equal_oracle(name, fn, key):
    import dspt._lib as L
    a = sweep_args()[key]
    if isinstance(a, tuple):
        x, y = a
        sx, sy = np.meshgrid(SPECIAL, SPECIAL)
        x = np.concatenate([x, sx.ravel()]).astype(np.float32)
        y = np.concatenate([y, sy.ravel()]).astype(np.float32)
    else:
        x, y = np.concatenate([a, SPECIAL]).astype(np.float32), None
    print(f"pin_math {name}: {x.size} arguments", flush=True)
    g = L.selftest_pin_math(fn, x, y)
    o = O.pin_math(name, x, y)
    same = _same_bits(g, o)
    bad = np.flatnonzero(~same)[:5]
    assert same.all(), [(x[i], None if y is None else y[i], g[i], o[i]) for i in bad]
