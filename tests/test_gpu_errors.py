"""Failed calls free what they allocated.

The failure is forced from the host: dspt_set_memory_limit caps the device
memory a context may hold, so every allocating entry point fails part-way
(after some of its buffers exist). Each must return DSPT_ERR_MEMORY and leave
dspt_memory_usage where it was; after the limit is lifted the same calls
succeed. Admission also refuses a job larger than the device's free HBM
before allocating anything.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_MEMORY = -4


def test_failed_calls_leave_memory_unchanged(sphere_field_small):
    from dspt import DsptError, Engine, pack_frame_block
    from test_gpu_parity import _surface_rays, _upload
    wl = sphere_field_small
    s = wl.scene
    eng = Engine(0)
    tot = _upload(eng, wl)
    img = wl.ibl["image"]
    blk = pack_frame_block(wl.camera, 32, 32, max_bounces=2, ibl_size=(img["width"], img["height"]), ibl_total_sum=tot)
    eng.resize(32, 32)
    eng.set_frame(blk)
    eng.render(1, 1)  # the path pool exists
    eng.synchronize()
    m0 = eng.memory_usage()
    rays = _surface_rays(wl, 4096, 5)
    ref = eng.trace_closest(rays)
    # each call either fails with DSPT_ERR_MEMORY leaving the usage at m0, or
    # fits under the limit (its resource is then destroyed again); margins
    # from 4 KiB up make every call fail before its first allocation and
    # part-way through its allocations
    calls = {
        "upload_ibl": (lambda: eng.upload_ibl(img["data"], img["width"], img["height"]), None),
        "ibl_create": (lambda: eng.ibl_create(img["data"], img["width"], img["height"])[0], eng.ibl_destroy),
        "scene_create": (lambda: eng.scene_create(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes,
                                                  s.bvhTriangleIndices, s.getFlatMaterialBuffer()), eng.scene_destroy),
        "trace_closest": (lambda: eng.trace_closest(rays), None),
        "job_create": (lambda: eng.job_create(0, 0, blk, 256, 256, 4), eng.job_destroy),
    }
    margins = [4096 << k for k in range(0, 14)]  # 4 KiB .. 32 MiB, doubling
    for name, (call, undo) in calls.items():
        failed = []
        for margin in margins:
            eng.set_memory_limit(m0 + margin)
            print(f"{name}: limit usage + {margin} B", flush=True)
            try:
                res = call()
            except DsptError as e:
                assert e.code == ERR_MEMORY, (name, str(e))
                failed.append(margin)
            else:
                if undo:
                    undo(res)
            assert eng.memory_usage() == m0, (name, margin)
        assert len(failed) >= 2 and failed[0] == 4096 and max(failed) >= 256 << 10, (name, failed)
    eng.set_memory_limit(0)
    # the engine still works, and what succeeds is counted and released
    got = eng.trace_closest(rays)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    iid = eng.ibl_create(img["data"], img["width"], img["height"])[0]
    assert eng.memory_usage() > m0
    eng.ibl_destroy(iid)
    assert eng.memory_usage() == m0
    eng.close()


def test_admission_refuses_more_than_free_hbm(sphere_field_small):
    from dspt import DsptError, Engine, pack_frame_block
    from test_gpu_parity import _upload
    wl = sphere_field_small
    eng = Engine(0)
    tot = _upload(eng, wl)
    img = wl.ibl["image"]
    W = H = 16384
    blk = pack_frame_block(wl.camera, W, H, max_bounces=2, ibl_size=(img["width"], img["height"]), ibl_total_sum=tot)
    m0 = eng.memory_usage()
    # recorded samples of 100 frames of 2^28 pixels: 644 GB, above any free HBM
    with pytest.raises(DsptError) as e:
        eng.job_create(0, 0, blk, W, H, 100, record=True)
    assert e.value.code == ERR_MEMORY and "free" in str(e.value)
    assert eng.memory_usage() == m0
    eng.close()
