import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dspbr-pt_amd")
for p in (PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libdspt.so")


@pytest.fixture(scope="session")
def sphere_field_small():
    """The sphere-field workload (full geometry) with its BVH, shared by tests."""
    from dspt import scenes, build_bvh
    wl = scenes.sphere_field()
    nodes, bidx = build_bvh(wl.scene.bvhPositionBuffer, wl.scene.bvhIndexBuffer)
    wl.scene.bvhNodes, wl.scene.bvhTriangleIndices = nodes, bidx
    return wl
