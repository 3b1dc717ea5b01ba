"""ctypes binding of the engine's C ABI (include/dspt.h).

The product path always goes through libdspt.so; there is no CPU fallback.
If the library is missing, loading raises immediately.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libdspt.so")

FRAME_FLOATS = 52
PROFILE_LEAN = 0
PROFILE_FULL = 1
CNT_NAMES = [
    "paths", "rays_primary", "rays_extension", "rays_shadow", "node_visits", "tri_tests",
    "shade_hits", "env_misses", "nee_samples",
]

# Every exported symbol of include/dspt.h with (restype, argtypes).
_c_int = ctypes.c_int
_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)

SIGNATURES = {
    "dspt_last_error": (ctypes.c_char_p, []),
    "dspt_version": (ctypes.c_char_p, []),
    "dspt_create": (_c_int, [_c_int, _c_int, ctypes.POINTER(_vp)]),
    "dspt_destroy": (_c_int, [_vp]),
    "dspt_upload_scene": (_c_int, [_vp, _fp, _c_i64, _u32p, _c_i64, _fp, _c_i64, _i32p, _c_i64, _fp, _c_i32]),
    "dspt_upload_ibl": (_c_int, [_vp, _fp, _c_i32, _c_i32, _dp]),
    "dspt_read_ibl_tables": (_c_int, [_vp, _fp, _fp, _fp, _fp]),
    "dspt_resize": (_c_int, [_vp, _c_i32, _c_i32]),
    "dspt_set_frame": (_c_int, [_vp, _fp]),
    "dspt_render": (_c_int, [_vp, _c_i32, _c_i32]),
    "dspt_render_region": (_c_int, [_vp, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32]),
    "dspt_clear_accum": (_c_int, [_vp]),
    "dspt_synchronize": (_c_int, [_vp]),
    "dspt_read_accum": (_c_int, [_vp, _fp]),
    "dspt_set_record_samples": (_c_int, [_vp, _c_int]),
    "dspt_read_samples": (_c_int, [_vp, _fp, _u32p, _i32p]),
    "dspt_get_counters": (_c_int, [_vp, _i64p]),
    "dspt_reset_counters": (_c_int, [_vp]),
    "dspt_trace_closest": (_c_int, [_vp, _fp, _c_i64, _i32p, _fp, _fp, _fp]),
    "dspt_trace_occluded": (_c_int, [_vp, _fp, _c_i64, _i32p]),
    "dspt_trace_closest_noalloc": (_c_int, [_vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _c_i32]),
    "dspt_build_bvh": (_c_int, [_fp, _c_i64, _u32p, _c_i64, _c_i32, _fp, _i64p, _i32p, _i64p]),
    "dspt_memory_usage": (_c_int, [_vp, _i64p]),
    "dspt_set_kernel_timing": (_c_int, [_vp, _c_int]),
    "dspt_set_bounce_graphs": (_c_int, [_vp, _c_int]),
    "dspt_set_ray_order": (_c_int, [_vp, _c_int]),
    "dspt_get_kernel_stats": (_c_int, [_vp, _dp]),
    "dspt_scene_create": (_c_int, [_vp, _fp, _c_i64, _u32p, _c_i64, _fp, _c_i64, _i32p, _c_i64, _fp, _c_i32, _i32p]),
    "dspt_scene_destroy": (_c_int, [_vp, _c_i32]),
    "dspt_scene_create_instanced": (_c_int, [_vp, _fp, _c_i64, _u32p, _c_i64, _i64p, _i64p, _c_i32, _fp, _i32p, _c_i32, _fp,
                                             _c_i32, _i32p]),
    "dspt_upload_scene_instanced": (_c_int, [_vp, _fp, _c_i64, _u32p, _c_i64, _i64p, _i64p, _c_i32, _fp, _i32p, _c_i32, _fp,
                                             _c_i32]),
    "dspt_trace_closest_instanced": (_c_int, [_vp, _fp, _c_i64, _i32p, _i32p, _fp, _fp, _fp]),
    "dspt_scene_create_gpu_bvh": (_c_int, [_vp, _fp, _c_i64, _u32p, _c_i64, _fp, _c_i32, _c_i32, _c_i32, _i32p, _dp]),
    "dspt_upload_scene_gpu_bvh": (_c_int, [_vp, _fp, _c_i64, _u32p, _c_i64, _fp, _c_i32, _c_i32, _c_i32, _dp]),
    "dspt_build_bvh_gpu": (_c_int, [_vp, _fp, _c_i64, _u32p, _c_i64, _c_i32, _c_i32, _fp, _i64p, _i32p, _dp]),
    "dspt_postprocess": (_c_int, [_vp, _c_i32, ctypes.c_float, _c_i32, _c_i32, _c_i32, _c_i32, _u8p, _fp]),
    "dspt_job_postprocess": (_c_int, [_vp, _c_i32, _c_i32, ctypes.c_float, _c_i32, _c_i32, _c_i32, _c_i32, _u8p, _fp]),
    "dspt_ibl_create": (_c_int, [_vp, _fp, _c_i32, _c_i32, _dp, _i32p]),
    "dspt_ibl_destroy": (_c_int, [_vp, _c_i32]),
    "dspt_job_create": (_c_int, [_vp, _c_i32, _c_i32, _fp, _c_i32, _c_i32, _c_i32, _c_i32, _i32p]),
    "dspt_job_destroy": (_c_int, [_vp, _c_i32]),
    "dspt_job_status": (_c_int, [_vp, _c_i32, _i32p, _i32p]),
    "dspt_job_read_accum": (_c_int, [_vp, _c_i32, _fp]),
    "dspt_job_read_samples": (_c_int, [_vp, _c_i32, _fp, _u32p, _i32p]),
    "dspt_jobs_step": (_c_int, [_vp, _i32p, _c_i32, _i32p]),
    "dspt_jobs_run": (_c_int, [_vp, _i32p, _c_i32]),
    "dspt_set_memory_budget": (_c_int, [_vp, _c_i64]),
    "dspt_set_memory_limit": (_c_int, [_vp, _c_i64]),
    "dspt_set_launch_config": (_c_int, [_vp, _c_i64, _c_i32, _c_i32]),
    "dspt_get_launch_config": (_c_int, [_vp, _i64p, _i32p, _i32p]),
    "dspt_set_schedule": (_c_int, [_vp, _c_i32, _c_i32, _c_i32, ctypes.c_float]),
    "dspt_get_schedule": (_c_int, [_vp, _i32p, _i32p, _i32p, _fp]),
    "dspt_set_policy": (_c_int, [_vp, _c_i32]),
    "dspt_selftest_pin_math": (_c_int, [_c_int, _c_i32, _fp, _fp, _fp, _c_i64]),
    "dspt_selftest_copy": (_c_int, [_c_int, _c_i64, _c_i32]),
    "dspt_scene_set_textures": (_c_int, [_vp, _c_i32, _fp, _c_i32, _i32p, _c_i32, ctypes.POINTER(ctypes.c_uint8), _c_i64]),
}
SCHED_STREAM = 0
SCHED_PASS = 1
POLICY_RR = 0
POLICY_FIFO = 1
POLICY_SRPT = 2
BVH_LBVH, BVH_PLOC = 0, 1
ORDER_NONE, ORDER_BUCKETS, ORDER_SHADE = 0, 1, 2
KERNEL_CLASSES = ["raygen", "trace_ext", "shade", "trace_shadow", "accumulate", "post", "order"]


class DsptError(RuntimeError):
    """Raised when an engine call returns a non-zero status (message from dspt_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"dspt error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libdspt.so (once; DSPT_LIBRARY may name another build of it, e.g.
    a tuning variant). Raises if the library has not been built."""
    global _lib
    if _lib is None:
        path = os.environ.get("DSPT_LIBRARY") or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(f"engine library not built: {path} (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != 0:
        msg = lib().dspt_last_error()
        raise DsptError(code, msg.decode() if msg else "")
    return code


def ptr(a, ctype=ctypes.c_float):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the engine must be C-contiguous"
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def build_bvh(positions, indices, max_leaf=4):
    """Engine-side builder of the TinyBVH bvh_gpu layout (nodes f32[16N], indices i32[M])."""
    pos = f32(positions).reshape(-1)
    ind = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    nn = ctypes.c_int64(0)
    nm = ctypes.c_int64(0)
    L = lib()
    check(L.dspt_build_bvh(ptr(pos), pos.size // 3, ptr(ind, ctypes.c_uint32), ind.size // 3, max_leaf, None,
                           ctypes.byref(nn), None, ctypes.byref(nm)))
    nodes = np.zeros(nn.value * 16, np.float32)
    bidx = np.zeros(nm.value, np.int32)
    check(L.dspt_build_bvh(ptr(pos), pos.size // 3, ptr(ind, ctypes.c_uint32), ind.size // 3, max_leaf, ptr(nodes),
                           ctypes.byref(nn), ptr(bidx, ctypes.c_int32), ctypes.byref(nm)))
    return nodes, bidx


def selftest_pin_math(fn, x, y=None, device=0):
    """Runs pinned transcendental `fn` (0 sin, 1 cos, 2 exp, 3 log, 4 pow,
    5 atan2, 6 acos, 7/8 sincos) on the GPU over float32 arrays."""
    x = f32(x).reshape(-1)
    yy = None if y is None else f32(y).reshape(-1)
    out = np.empty_like(x)
    check(lib().dspt_selftest_pin_math(device, fn, ptr(x), ptr(yy), ptr(out), x.size))
    return out


def selftest_copy(nbytes, reps=5, device=0):
    """reps launches of a plain device copy of nbytes (counter calibration)."""
    check(lib().dspt_selftest_copy(device, int(nbytes), int(reps)))


class Engine:
    """Thin RAII wrapper over one dspt_ctx (one device, one stream)."""

    def __init__(self, profile=PROFILE_LEAN, device=0):
        self._L = lib()
        h = ctypes.c_void_p()
        check(self._L.dspt_create(int(profile), int(device), ctypes.byref(h)))
        self.h = h
        self.W = self.H = 0
        self.ibl_size = (0, 0)

    def close(self):
        if getattr(self, "h", None):
            self._L.dspt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_scene(self, verts, tris, nodes, bvh_idx, mats):
        verts, nodes, mats = f32(verts), f32(nodes), f32(mats)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        bvh_idx = np.ascontiguousarray(bvh_idx, dtype=np.int32)
        check(self._L.dspt_upload_scene(self.h, ptr(verts), verts.size // 20, ptr(tris, ctypes.c_uint32), tris.size // 3,
                                        ptr(nodes), nodes.size // 16, ptr(bvh_idx, ctypes.c_int32), bvh_idx.size,
                                        ptr(mats), mats.size // 68))

    def upload_ibl(self, rgba, w, h):
        rgba = f32(rgba)
        assert rgba.size == 4 * w * h
        tot = ctypes.c_double(0)
        check(self._L.dspt_upload_ibl(self.h, ptr(rgba), int(w), int(h), ctypes.byref(tot)))
        self.ibl_size = (w, h)
        return tot.value

    def read_ibl_tables(self):
        w, h = self.ibl_size
        pdf = np.zeros(w * h, np.float32)
        cdf = np.zeros(w * h, np.float32)
        ypdf = np.zeros(h, np.float32)
        ycdf = np.zeros(h, np.float32)
        check(self._L.dspt_read_ibl_tables(self.h, ptr(pdf), ptr(cdf), ptr(ypdf), ptr(ycdf)))
        return pdf, cdf, ypdf, ycdf

    def resize(self, w, h):
        check(self._L.dspt_resize(self.h, int(w), int(h)))
        self.W, self.H = int(w), int(h)

    def set_frame(self, block):
        block = f32(block)
        assert block.size == FRAME_FLOATS
        check(self._L.dspt_set_frame(self.h, ptr(block)))

    def render(self, first_frame, num_frames):
        check(self._L.dspt_render(self.h, int(first_frame), int(num_frames)))

    def render_region(self, first_frame, num_frames, x0, y0, width, height):
        """Frames of the pixels [x0, x0+width) x [y0, y0+height) only (row 0 = bottom)."""
        check(self._L.dspt_render_region(self.h, int(first_frame), int(num_frames), int(x0), int(y0), int(width),
                                         int(height)))

    def clear_accum(self):
        check(self._L.dspt_clear_accum(self.h))

    def synchronize(self):
        check(self._L.dspt_synchronize(self.h))

    def read_accum(self):
        out = np.zeros(self.W * self.H * 4, np.float32)
        check(self._L.dspt_read_accum(self.h, ptr(out)))
        return out.reshape(self.H, self.W, 4)

    def set_record_samples(self, on):
        check(self._L.dspt_set_record_samples(self.h, 1 if on else 0))

    def read_samples(self, num_frames):
        n = num_frames * self.W * self.H
        rgba = np.zeros(n * 4, np.float32)
        sig = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.int32)
        check(self._L.dspt_read_samples(self.h, ptr(rgba), ptr(sig, ctypes.c_uint32), ptr(prim, ctypes.c_int32)))
        return rgba.reshape(num_frames, self.H, self.W, 4), sig.reshape(num_frames, self.H, self.W), prim.reshape(
            num_frames, self.H, self.W)

    def counters(self):
        out = np.zeros(16, np.int64)
        check(self._L.dspt_get_counters(self.h, ptr(out, ctypes.c_int64)))
        return {k: int(out[i]) for i, k in enumerate(CNT_NAMES)}

    def reset_counters(self):
        check(self._L.dspt_reset_counters(self.h))

    def trace_closest(self, rays):
        rays = f32(rays).reshape(-1, 8)
        n = rays.shape[0]
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        check(self._L.dspt_trace_closest(self.h, ptr(rays), n, ptr(tri, ctypes.c_int32), ptr(t), ptr(u), ptr(v)))
        return tri, t, u, v

    def trace_occluded(self, rays):
        rays = f32(rays).reshape(-1, 8)
        n = rays.shape[0]
        occ = np.zeros(n, np.int32)
        check(self._L.dspt_trace_occluded(self.h, ptr(rays), n, ptr(occ, ctypes.c_int32)))
        return occ

    def set_bounce_graphs(self, on):
        """Replay the per-chunk bounce loop as a captured HIP graph (default) or launch it eagerly."""
        check(self._L.dspt_set_bounce_graphs(self.h, 1 if on else 0))

    def set_ray_order(self, mode):
        """ORDER_NONE (compaction order), ORDER_BUCKETS (origin-cell + up/down buckets before each trace) or
        ORDER_SHADE (the default: BUCKETS, then shading grouped by the hit's material)."""
        check(self._L.dspt_set_ray_order(self.h, int(mode)))

    def set_kernel_timing(self, on):
        check(self._L.dspt_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_stats(self):
        out = np.zeros(8 * 4, np.float64)
        check(self._L.dspt_get_kernel_stats(self.h, ptr(out, ctypes.c_double)))
        out = out.reshape(8, 4)
        return {k: {"launches": int(out[i, 0]), "total_ms": float(out[i, 1]), "node_visits": int(out[i, 2]),
                    "tri_tests": int(out[i, 3])} for i, k in enumerate(KERNEL_CLASSES)}

    # ---- resident scenes / IBLs and jobs (batched + served modes)
    def scene_create(self, verts, tris, nodes, bvh_idx, mats):
        verts, nodes, mats = f32(verts), f32(nodes), f32(mats)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        bvh_idx = np.ascontiguousarray(bvh_idx, dtype=np.int32)
        sid = ctypes.c_int32(-1)
        check(self._L.dspt_scene_create(self.h, ptr(verts), verts.size // 20, ptr(tris, ctypes.c_uint32), tris.size // 3,
                                        ptr(nodes), nodes.size // 16, ptr(bvh_idx, ctypes.c_int32), bvh_idx.size,
                                        ptr(mats), mats.size // 68, ctypes.byref(sid)))
        return sid.value

    @staticmethod
    def _instanced_args(verts, tris, mesh_first, mesh_count, o2w, inst_mesh, mats):
        verts, mats = f32(verts), f32(mats)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        mf = np.ascontiguousarray(mesh_first, np.int64)
        mc = np.ascontiguousarray(mesh_count, np.int64)
        o2w = f32(o2w).reshape(-1)
        im = np.ascontiguousarray(inst_mesh, np.int32)
        if mf.size != mc.size or o2w.size != 12 * im.size:
            raise ValueError("mesh ranges or instance arrays have mismatched lengths")
        keep = (verts, tris, mf, mc, o2w, im, mats)
        args = (ptr(verts), verts.size // 20, ptr(tris, ctypes.c_uint32), tris.size // 3, ptr(mf, ctypes.c_int64),
                ptr(mc, ctypes.c_int64), mf.size, ptr(o2w), ptr(im, ctypes.c_int32), im.size, ptr(mats), mats.size // 68)
        return keep, args

    def scene_create_instanced(self, verts, tris, mesh_first, mesh_count, o2w, inst_mesh, mats):
        """Two-level scene: object-space meshes (triangle ranges of `tris`)
        placed by instances (o2w f32[I, 3, 4], inst_mesh i32[I])."""
        keep, args = self._instanced_args(verts, tris, mesh_first, mesh_count, o2w, inst_mesh, mats)
        sid = ctypes.c_int32(-1)
        check(self._L.dspt_scene_create_instanced(self.h, *args, ctypes.byref(sid)))
        del keep
        return sid.value

    def upload_scene_instanced(self, verts, tris, mesh_first, mesh_count, o2w, inst_mesh, mats):
        keep, args = self._instanced_args(verts, tris, mesh_first, mesh_count, o2w, inst_mesh, mats)
        check(self._L.dspt_upload_scene_instanced(self.h, *args))
        del keep

    def trace_closest_instanced(self, rays):
        rays = f32(rays).reshape(-1, 8)
        n = rays.shape[0]
        inst = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        check(self._L.dspt_trace_closest_instanced(self.h, ptr(rays), n, ptr(inst, ctypes.c_int32), ptr(tri, ctypes.c_int32),
                                                   ptr(t), ptr(u), ptr(v)))
        return inst, tri, t, u, v

    def _post(self, fn, args, size, tonemap_mode, exposure, gamma, fxaa, out_size, want_float):
        ow, oh = out_size or size
        u8 = np.zeros((oh, ow, 4), np.uint8)
        f = np.zeros((oh, ow, 4), np.float32) if want_float else None
        check(fn(self.h, *args, int(tonemap_mode), float(exposure), 1 if gamma else 0, 1 if fxaa else 0, int(ow), int(oh),
                 ptr(u8, ctypes.c_uint8), ptr(f)))
        return (u8, f) if want_float else u8

    def postprocess(self, tonemap_mode=0, exposure=1.0, gamma=True, fxaa=False, out_size=None, want_float=False):
        """Display chain (renderer.ts:665-700) on the single-image
        accumulation: RGBA8 canvas u8[oh, ow, 4] (row 0 = bottom), and the
        float display buffer when want_float."""
        return self._post(self._L.dspt_postprocess, (), (self.W, self.H), tonemap_mode, exposure, gamma, fxaa, out_size, want_float)

    def job_postprocess(self, jid, tonemap_mode=0, exposure=1.0, gamma=True, fxaa=False, out_size=None, want_float=False):
        w, h, _ = self._job_dims[jid]
        return self._post(self._L.dspt_job_postprocess, (int(jid),), (w, h), tonemap_mode, exposure, gamma, fxaa, out_size,
                          want_float)

    def scene_create_gpu_bvh(self, verts, tris, mats, max_leaf=4, method=1):
        """Resident scene whose BVH the GPU builds (method BVH_PLOC or
        BVH_LBVH); returns (scene id, build ms)."""
        verts, mats = f32(verts), f32(mats)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        sid, ms = ctypes.c_int32(-1), ctypes.c_double(0)
        check(self._L.dspt_scene_create_gpu_bvh(self.h, ptr(verts), verts.size // 20, ptr(tris, ctypes.c_uint32),
                                                tris.size // 3, ptr(mats), mats.size // 68, int(max_leaf), int(method),
                                                ctypes.byref(sid), ctypes.byref(ms)))
        return sid.value, ms.value

    def upload_scene_gpu_bvh(self, verts, tris, mats, max_leaf=4, method=1):
        verts, mats = f32(verts), f32(mats)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        ms = ctypes.c_double(0)
        check(self._L.dspt_upload_scene_gpu_bvh(self.h, ptr(verts), verts.size // 20, ptr(tris, ctypes.c_uint32),
                                                tris.size // 3, ptr(mats), mats.size // 68, int(max_leaf), int(method),
                                                ctypes.byref(ms)))
        return ms.value

    def build_bvh_gpu(self, verts, tris, max_leaf=4, method=1):
        """GPU BVH of a vertex stream (method BVH_PLOC or BVH_LBVH): (nodes f32[16 N], bvh indices i32[T], build ms)."""
        verts = f32(verts)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        T = tris.size // 3
        nodes = np.zeros(16 * max(1, 2 * T - 1), np.float32)
        bidx = np.zeros(max(1, T), np.int32)
        nn, ms = ctypes.c_int64(0), ctypes.c_double(0)
        check(self._L.dspt_build_bvh_gpu(self.h, ptr(verts), verts.size // 20, ptr(tris, ctypes.c_uint32), T, int(max_leaf),
                                         int(method), ptr(nodes), ctypes.byref(nn), ptr(bidx, ctypes.c_int32), ctypes.byref(ms)))
        return nodes[:16 * nn.value].copy(), bidx[:T], ms.value

    def scene_destroy(self, sid):
        check(self._L.dspt_scene_destroy(self.h, int(sid)))

    def scene_set_textures(self, scene_id, texinfos, array_dims, texels):
        """texinfos f32[8*X]; array_dims i32[A, 3] (w, h, layers); texels u8 (all arrays, RGBA8, row 0 first)."""
        ti = f32(texinfos)
        dims = np.ascontiguousarray(array_dims, np.int32).reshape(-1)
        tx = np.ascontiguousarray(texels, np.uint8).reshape(-1)
        check(self._L.dspt_scene_set_textures(self.h, int(scene_id), ptr(ti), ti.size // 8, ptr(dims, ctypes.c_int32),
                                              dims.size // 3, ptr(tx, ctypes.c_uint8), tx.size))

    def ibl_create(self, rgba, w, h):
        rgba = f32(rgba)
        tot = ctypes.c_double(0)
        iid = ctypes.c_int32(-1)
        check(self._L.dspt_ibl_create(self.h, ptr(rgba), int(w), int(h), ctypes.byref(tot), ctypes.byref(iid)))
        return iid.value, tot.value

    def ibl_destroy(self, iid):
        check(self._L.dspt_ibl_destroy(self.h, int(iid)))

    def job_create(self, scene, ibl, block, w, h, frames, record=False):
        block = f32(block)
        jid = ctypes.c_int32(-1)
        check(self._L.dspt_job_create(self.h, int(scene), int(ibl), ptr(block), int(w), int(h), int(frames),
                                      1 if record else 0, ctypes.byref(jid)))
        self._job_dims = getattr(self, "_job_dims", {})
        self._job_dims[jid.value] = (int(w), int(h), int(frames))
        return jid.value

    def job_destroy(self, jid):
        check(self._L.dspt_job_destroy(self.h, int(jid)))
        getattr(self, "_job_dims", {}).pop(jid, None)

    def job_status(self, jid):
        d = ctypes.c_int32(0)
        t = ctypes.c_int32(0)
        check(self._L.dspt_job_status(self.h, int(jid), ctypes.byref(d), ctypes.byref(t)))
        return d.value, t.value

    def job_read_accum(self, jid):
        w, h, _ = self._job_dims[jid]
        out = np.zeros(w * h * 4, np.float32)
        check(self._L.dspt_job_read_accum(self.h, int(jid), ptr(out)))
        return out.reshape(h, w, 4)

    def job_read_samples(self, jid):
        w, h, nf = self._job_dims[jid]
        n = nf * w * h
        rgba = np.zeros(n * 4, np.float32)
        sig = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.int32)
        check(self._L.dspt_job_read_samples(self.h, int(jid), ptr(rgba), ptr(sig, ctypes.c_uint32), ptr(prim, ctypes.c_int32)))
        return rgba.reshape(nf, h, w, 4), sig.reshape(nf, h, w), prim.reshape(nf, h, w)

    def jobs_step(self, job_ids=None):
        """One chunk of bounces over the listed jobs (all when None); returns
        how many of them are not finished (as of the previous chunk)."""
        active = ctypes.c_int32(0)
        if job_ids is None:
            check(self._L.dspt_jobs_step(self.h, None, 0, ctypes.byref(active)))
        else:
            ids = np.ascontiguousarray(job_ids, np.int32)
            check(self._L.dspt_jobs_step(self.h, ptr(ids, ctypes.c_int32), ids.size, ctypes.byref(active)))
        return active.value

    def jobs_run(self, job_ids=None):
        if job_ids is None:
            check(self._L.dspt_jobs_run(self.h, None, 0))
        else:
            ids = np.ascontiguousarray(job_ids, np.int32)
            check(self._L.dspt_jobs_run(self.h, ptr(ids, ctypes.c_int32), ids.size))

    def set_schedule(self, mode=SCHED_STREAM, pass_frames=0, chunk_iters=0, ring_factor=0.0):
        check(self._L.dspt_set_schedule(self.h, int(mode), int(pass_frames), int(chunk_iters), float(ring_factor)))

    def set_policy(self, policy):
        check(self._L.dspt_set_policy(self.h, int(policy)))

    def schedule(self):
        m, pf, ci = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        rf = ctypes.c_float(0)
        check(self._L.dspt_get_schedule(self.h, ctypes.byref(m), ctypes.byref(pf), ctypes.byref(ci), ctypes.byref(rf)))
        return {"mode": m.value, "pass_frames": pf.value, "chunk_iters": ci.value, "ring_factor": rf.value}

    def set_memory_limit(self, nbytes):
        check(self._L.dspt_set_memory_limit(self.h, int(nbytes)))

    def set_memory_budget(self, nbytes):
        check(self._L.dspt_set_memory_budget(self.h, int(nbytes)))

    def set_launch_config(self, max_paths=0, trace_blocks=0, shade_blocks=0):
        check(self._L.dspt_set_launch_config(self.h, int(max_paths), int(trace_blocks), int(shade_blocks)))

    def launch_config(self):
        mp = ctypes.c_int64(0)
        tb = ctypes.c_int32(0)
        sb = ctypes.c_int32(0)
        check(self._L.dspt_get_launch_config(self.h, ctypes.byref(mp), ctypes.byref(tb), ctypes.byref(sb)))
        return {"max_paths": mp.value, "trace_blocks": tb.value, "shade_blocks": sb.value}

    def memory_usage(self):
        b = ctypes.c_int64(0)
        check(self._L.dspt_memory_usage(self.h, ctypes.byref(b)))
        return b.value
