"""ctypes binding of the CPU oracle (oracle/liboracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
and only as the checker / CPU baseline.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_f = ctypes.POINTER(ctypes.c_float)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_d = ctypes.POINTER(ctypes.c_double)


class CScene(ctypes.Structure):
    _fields_ = [
        ("verts", _f), ("V", ctypes.c_int), ("tris", _u32), ("T", ctypes.c_int), ("nodes", _f), ("N", ctypes.c_int),
        ("bvh_idx", _i32), ("M", ctypes.c_int), ("mats", _f), ("K", ctypes.c_int), ("env", _f), ("ew", ctypes.c_int),
        ("eh", ctypes.c_int), ("pdf", _f), ("cdf", _f), ("ypdf", _f), ("ycdf", _f), ("profile", ctypes.c_int),
        ("texinfo", _f), ("num_tex", ctypes.c_int), ("texdims", _i32), ("num_texarr", ctypes.c_int),
        ("texels", ctypes.POINTER(ctypes.c_uint8)),
    ]


def build_oracle():
    import subprocess
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        srcs = [os.path.join(ORACLE_DIR, f) for f in os.listdir(ORACLE_DIR) if f.endswith((".hpp", ".cpp"))]
        srcs.append(os.path.join(ROOT, "dspbr-pt_amd", "csrc", "pin_math.hpp"))
        if not os.path.exists(ORACLE_LIB) or max(os.path.getmtime(f) for f in srcs) > os.path.getmtime(ORACLE_LIB):
            build_oracle()
        L = ctypes.CDLL(ORACLE_LIB)
        L.oracle_ibl_cdf.restype = ctypes.c_double
        L.oracle_pin_math.argtypes = [ctypes.c_int, _f, _f, _f, ctypes.c_int64]
        L.oracle_camera_ray.argtypes = [_f, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        ctypes.c_float, _f]
        L.oracle_ggx_albedo_fits.argtypes = [ctypes.c_double, ctypes.c_double, _d]
        L.oracle_ibl_cdf.argtypes = [_f, ctypes.c_int, ctypes.c_int, _f, _f, _f, _f]
        L.oracle_rng_seq.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32, _u32, _f, _u32]
        L.oracle_rng_step.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _u32]
        L.oracle_trace.argtypes = [ctypes.POINTER(CScene), _f, ctypes.c_int, _i32, _f, _f, _f, _i64, ctypes.c_int]
        L.oracle_render.argtypes = [ctypes.POINTER(CScene), _f, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, _f, _u32, _i32, _i64,
                                    ctypes.c_int, ctypes.c_int]
        L.oracle_bsdf_eval.argtypes = [_f, ctypes.c_int, _f, _f, ctypes.c_int, _d]
        L.oracle_bsdf_sample.argtypes = [_f, ctypes.c_int, _f, _f, ctypes.c_int, ctypes.c_int, _d]
        L.oracle_ibl_sample.argtypes = [ctypes.POINTER(CScene), _f, _f, ctypes.c_int, _f]
        L.oracle_ibl_lookup.argtypes = [ctypes.POINTER(CScene), _f, _f, ctypes.c_int, _f]
        L.oracle_tex_value.argtypes = [ctypes.POINTER(CScene), ctypes.c_float, _f, ctypes.c_int, _f]
        L.oracle_postprocess.argtypes = [_f, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, ctypes.POINTER(ctypes.c_uint8)]
        L.oracle_thin_film.argtypes = [_f, _f, ctypes.c_int, ctypes.c_int, _d]
        L.oracle_dispersion_iors.argtypes = [_f, ctypes.c_int, ctypes.c_int, _d]
        assert L.oracle_scene_struct_size() == ctypes.sizeof(CScene)
        assert L.oracle_frame_struct_size() == 52 * 4
        _lib = L
    return _lib


def _p(a, t=ctypes.c_float):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


class OracleScene:
    """Holds the numpy buffers alive for the C struct."""

    def __init__(self, verts, tris, nodes, bvh_idx, mats, ibl=None, profile=0):
        self.verts = np.ascontiguousarray(verts, np.float32).reshape(-1)
        self.tris = np.ascontiguousarray(tris, np.uint32).reshape(-1)
        self.nodes = np.ascontiguousarray(nodes, np.float32).reshape(-1)
        self.bvh_idx = np.ascontiguousarray(bvh_idx, np.int32).reshape(-1)
        self.mats = np.ascontiguousarray(mats, np.float32).reshape(-1)
        self.c = CScene()
        self.c.verts, self.c.V = _p(self.verts), self.verts.size // 20
        self.c.tris, self.c.T = _p(self.tris, ctypes.c_uint32), self.tris.size // 3
        self.c.nodes, self.c.N = _p(self.nodes), self.nodes.size // 16
        self.c.bvh_idx, self.c.M = _p(self.bvh_idx, ctypes.c_int32), self.bvh_idx.size
        self.c.mats, self.c.K = _p(self.mats), self.mats.size // 68
        self.c.profile = profile
        self.total = None
        if ibl is not None:
            self.set_ibl(ibl)

    def set_textures(self, texinfos, array_dims, texels):
        """Material textures (same arguments as the engine's scene_set_textures)."""
        self.texinfo = np.ascontiguousarray(texinfos, np.float32).reshape(-1)
        self.texdims = np.ascontiguousarray(array_dims, np.int32).reshape(-1)
        self.texels = np.ascontiguousarray(texels, np.uint8).reshape(-1)
        self.c.texinfo, self.c.num_tex = _p(self.texinfo), self.texinfo.size // 8
        self.c.texdims, self.c.num_texarr = _p(self.texdims, ctypes.c_int32), self.texdims.size // 3
        self.c.texels = _p(self.texels, ctypes.c_uint8)

    def tex_value(self, tex_id, uv):
        """get_texture_value for probes: f32[n, 4] at uv f32[n, 2]."""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 4), np.float32)
        lib().oracle_tex_value(ctypes.byref(self.c), float(tex_id), _p(uv), uv.shape[0], _p(out))
        return out

    def set_ibl(self, ibl):
        img = ibl["image"]
        w, h = int(img["width"]), int(img["height"])
        self.env = np.ascontiguousarray(img["data"], np.float32).reshape(-1)
        self.pdf, self.cdf, self.ypdf, self.ycdf, self.total = ibl_cdf(self.env, w, h)
        self.c.env, self.c.ew, self.c.eh = _p(self.env), w, h
        self.c.pdf, self.c.cdf, self.c.ypdf, self.c.ycdf = _p(self.pdf), _p(self.cdf), _p(self.ypdf), _p(self.ycdf)

    def trace(self, rays, threads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        cnt = np.zeros(2, np.int64)
        lib().oracle_trace(ctypes.byref(self.c), _p(rays), n, _p(tri, ctypes.c_int32), _p(t), _p(u), _p(v),
                           _p(cnt, ctypes.c_int64), threads)
        return tri, t, u, v, cnt

    def render(self, block, W, H, first_frame, nframes, rect=None, accum=None, record=True, threads=8, use_double=False):
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, W, H)
        rw, rh = x1 - x0, y1 - y0
        block = np.ascontiguousarray(block, np.float32)
        if accum is None:
            accum = np.zeros(W * H * 4, np.float32)
        samples = np.zeros(nframes * rw * rh * 4, np.float32) if record else None
        sigs = np.zeros(nframes * rw * rh, np.uint32) if record else None
        prim = np.zeros(nframes * rw * rh, np.int32) if record else None
        cnt = np.zeros(3, np.int64)
        lib().oracle_render(ctypes.byref(self.c), _p(block), W, H, x0, y0, x1, y1, first_frame, nframes, _p(accum),
                            _p(samples), _p(sigs, ctypes.c_uint32), _p(prim, ctypes.c_int32), _p(cnt, ctypes.c_int64),
                            threads, 1 if use_double else 0)
        out = {"accum": accum.reshape(H, W, 4), "counters": cnt}
        if record:
            out["samples"] = samples.reshape(nframes, rh, rw, 4)
            out["sigs"] = sigs.reshape(nframes, rh, rw)
            out["prim"] = prim.reshape(nframes, rh, rw)
        return out

    def ibl_sample(self, block, r01):
        r01 = np.ascontiguousarray(r01, np.float32).reshape(-1, 2)
        out = np.zeros((r01.shape[0], 4), np.float32)
        lib().oracle_ibl_sample(ctypes.byref(self.c), _p(np.ascontiguousarray(block, np.float32)), _p(r01),
                                r01.shape[0], _p(out))
        return out

    def ibl_lookup(self, block, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros((dirs.shape[0], 4), np.float32)
        lib().oracle_ibl_lookup(ctypes.byref(self.c), _p(np.ascontiguousarray(block, np.float32)), _p(dirs),
                                dirs.shape[0], _p(out))
        return out


def ibl_cdf(rgba, w, h):
    rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1)
    pdf = np.zeros(w * h, np.float32)
    cdf = np.zeros(w * h, np.float32)
    ypdf = np.zeros(h, np.float32)
    ycdf = np.zeros(h, np.float32)
    tot = lib().oracle_ibl_cdf(_p(rgba), w, h, _p(pdf), _p(cdf), _p(ypdf), _p(ycdf))
    return pdf, cdf, ypdf, ycdf, tot


def rng_seq(seed, px, py, n):
    """-> initial state (x, y), states after each step [n, 2], u32 outputs, floats."""
    states = np.zeros(2 * n, np.uint32)
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    st = np.zeros(2, np.uint32)
    lib().oracle_rng_seq(seed, px, py, n, _p(states, ctypes.c_uint32), _p(u, ctypes.c_uint32), _p(f),
                         _p(st, ctypes.c_uint32))
    return st, states.reshape(n, 2), u, f


def rng_step(x, y, n):
    states = np.zeros(2 * n, np.uint32)
    lib().oracle_rng_step(int(x), int(y), n, _p(states, ctypes.c_uint32))
    return states.reshape(n, 2)


def thin_film(rows, base3=None, use_double=True):
    """rows: (n, 5) {outsideIor, filmIor, baseF0 (< 0: per channel from base3),
    thicknessNm, cosTheta1} -> (n, 3) thin-film reflectance (KF:789-915)."""
    rows = np.ascontiguousarray(rows, np.float32).reshape(-1, 5)
    b3 = None if base3 is None else np.ascontiguousarray(base3, np.float32).reshape(-1, 3)
    out = np.zeros((rows.shape[0], 3), np.float64)
    rc = lib().oracle_thin_film(_p(rows), _p(b3), rows.shape[0], 1 if use_double else 0, _p(out, ctypes.c_double))
    assert rc == 0
    return out


def dispersion_iors(ior_disp, use_double=True):
    """(n, 2) {ior, dispersion} -> (n, 3) channel IORs (KF:1289-1298)."""
    a = np.ascontiguousarray(ior_disp, np.float32).reshape(-1, 2)
    out = np.zeros((a.shape[0], 3), np.float64)
    lib().oracle_dispersion_iors(_p(a), a.shape[0], 1 if use_double else 0, _p(out, ctypes.c_double))
    return out


def bsdf_eval(mat, geo, wo, profile=0, use_double=True):
    out = np.zeros(8, np.float64)
    lib().oracle_bsdf_eval(_p(np.ascontiguousarray(mat, np.float32)), profile, _p(np.ascontiguousarray(geo, np.float32)),
                           _p(np.ascontiguousarray(wo, np.float32)), 1 if use_double else 0, _p(out, ctypes.c_double))
    return out


def bsdf_sample(mat, geo, rnd, profile=0, use_double=True):
    rnd = np.ascontiguousarray(rnd, np.float32).reshape(-1, 4)
    out = np.zeros((rnd.shape[0], 8), np.float64)
    lib().oracle_bsdf_sample(_p(np.ascontiguousarray(mat, np.float32)), profile, _p(np.ascontiguousarray(geo, np.float32)),
                             _p(rnd), rnd.shape[0], 1 if use_double else 0, _p(out, ctypes.c_double))
    return out


def postprocess(accum, mode, exposure=1.0, gamma=True, fxaa=False, out_size=None):
    """renderer.ts:665-700 on an RGBA32F accumulation f32[H, W, 4] (row 0 =
    bottom): (display buffer f32[oh, ow, 4], 8-bit canvas u8[oh, ow, 4])."""
    a = np.ascontiguousarray(accum, np.float32)
    H, W = a.shape[0], a.shape[1]
    ow, oh = out_size or (W, H)
    f = np.zeros((oh, ow, 4), np.float32)
    u8 = np.zeros((oh, ow, 4), np.uint8)
    lib().oracle_postprocess(_p(a), W, H, int(mode), float(exposure), int(bool(gamma)), int(bool(fxaa)), ow, oh, _p(f),
                             _p(u8, ctypes.c_uint8))
    return f, u8


PIN_FNS = {"sin": 0, "cos": 1, "exp": 2, "log": 3, "pow": 4, "atan2": 5, "acos": 6, "sincos_s": 7, "sincos_c": 8}


def pin_math(name, x, y=None):
    """The oracle's pinned float transcendental `name` over float32 arrays
    (atan2 takes (y-coordinate, x-coordinate) as (x, y))."""
    x = np.ascontiguousarray(x, np.float32)
    yy = None if y is None else np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    r = lib().oracle_pin_math(PIN_FNS[name], x.ctypes.data_as(_f), None if yy is None else yy.ctypes.data_as(_f),
                              out.ctypes.data_as(_f), x.size)
    assert r == 0
    return out


def camera_ray(block, W, H, px, py, r0=0.0, r1=0.0):
    """(origin, direction) of the oracle's camera ray for pixel (px, py) with jitter (r0, r1)."""
    out = np.zeros(6, np.float32)
    b = np.ascontiguousarray(block, np.float32)
    lib().oracle_camera_ray(b.ctypes.data_as(_f), W, H, px, py, r0, r1, out.ctypes.data_as(_f))
    return out[:3], out[3:]


def ggx_albedo_fits(alpha, cos_t):
    """(approxAvgAlbedoGGX(alpha), approxDirAlbedoGGX(alpha, cos_t)) of the oracle, in double."""
    out = np.zeros(2, np.float64)
    lib().oracle_ggx_albedo_fits(float(alpha), float(cos_t), out.ctypes.data_as(_d))
    return out[0], out[1]
