"""Formula parity of the engine's shading code against the oracle, on the CPU.

tests/devhost compiles the device headers (dev_shade.hpp, dev_bsdf.hpp) for
the host with the same flags as the oracle (-ffp-contract=off), so the two
run on the same libm. Material set-up, BSDF eval, pdf and sampling must then
agree bit for bit: a dropped term, a swapped operand or a different
association shows up as a mismatch here, independently of the GPU's
transcendental functions (which differ from glibc's by an ulp or two and
are the only source of the small per-sample drift the GPU parity tests allow).
"""
import ctypes
import os
import shutil
import subprocess
import zlib

import numpy as np
import pytest

import oracle_ref as O

HERE = os.path.dirname(os.path.abspath(__file__))
DEVHOST = os.path.join(HERE, "devhost")


@pytest.fixture(scope="module")
def devhost():
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available for the host build of the shading code")
    subprocess.run(["make", "-s", "-C", DEVHOST], check=True)
    return ctypes.CDLL(os.path.join(DEVHOST, "libdevhost.so"))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _unit(v):
    return v / np.linalg.norm(v)


def _geo(rng):
    n = _unit(rng.normal(size=3))
    t = _unit(np.cross(n, _unit(rng.normal(size=3))))
    ng = _unit(n + 0.2 * rng.normal(size=3))
    wi = _unit(rng.normal(size=3))
    g = np.zeros(16, np.float32)
    g[0:3], g[3:6], g[6:9], g[9], g[10:13] = n, ng, t, 1.0, wi
    return g


@pytest.mark.parametrize("profile", [0, 1])
@pytest.mark.parametrize("family", ["basic", "clearcoat", "sheen", "anisotropy", "specular", "transmission",
                                    "iridescence", "dispersion", "thin"])
def test_device_bsdf_matches_oracle_bitwise(devhost, family, profile):
    from dspt import scenes
    rng = np.random.default_rng(zlib.crc32(f"{family}/{profile}".encode()))
    n_eval = n_sample = 0
    for trial in range(60):
        if family == "basic":
            m = scenes.basic_material(rng)
        elif family == "thin":
            m = scenes.family_material(rng, "transmission")
            m.thinWalled = True
        else:
            m = scenes.family_material(rng, family)
        mat = np.ascontiguousarray(m.data, np.float32)
        geo = _geo(rng)
        for _ in range(16):
            wo = _unit(rng.normal(size=3)).astype(np.float32)
            ref = O.bsdf_eval(mat, geo, wo, profile=profile, use_double=False)
            dev = np.zeros(8)
            devhost.devhost_bsdf_eval(_p(mat), profile, _p(geo), _p(wo), _p(dev))
            np.testing.assert_array_equal(dev, ref, err_msg=f"eval {family} trial {trial}")
            n_eval += 1
        rnd = rng.uniform(0, 1, (32, 4)).astype(np.float32)
        ref = O.bsdf_sample(mat, geo, rnd, profile=profile, use_double=False)
        dev = np.zeros((32, 8))
        devhost.devhost_bsdf_sample(_p(mat), profile, _p(geo), _p(rnd), 32, _p(dev))
        np.testing.assert_array_equal(dev, ref, err_msg=f"sample {family} trial {trial}")
        n_sample += 32
    assert n_eval and n_sample
