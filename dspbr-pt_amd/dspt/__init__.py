"""dspt — MI355X (gfx950) path-tracing engine behind the dspbr-pt renderer surface.

The compute path is libdspt.so (HIP kernels + C ABI, include/dspt.h); this
package is the host-side mirror of the reference's PathtracingRenderer /
PathtracingViewport and the synthetic workload generators.
"""
from ._lib import DsptError, Engine, build_bvh, lib, PROFILE_FULL, PROFILE_LEAN, SCHED_PASS, SCHED_STREAM, POLICY_RR, POLICY_FIFO, POLICY_SRPT, BVH_LBVH, BVH_PLOC, ORDER_NONE, ORDER_BUCKETS, ORDER_SHADE  # noqa: F401
from .renderer import PathtracingRenderer, PathtracingViewport, pack_frame_block  # noqa: F401
from .scene import Light, MaterialData, PathtracingSceneData, TexInfo, ViewportCamera  # noqa: F401
