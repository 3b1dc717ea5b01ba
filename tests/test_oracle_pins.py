"""Pins of the oracle functions that test_oracle.py leaves to invariants:
sheen, dielectric transmission (thick and thin), the dielectric specular F0
with specular tint, diffuse transmission, the energy-compensation lobe, the
camera ray and the point light. Each is checked against what the physics or
the published method fixes, computed here independently of the oracle's
code, so that a dropped term, a wrong exponent or a swapped operand fails:

  sheen (KF:1798-1889)          the Charlie distribution of Estevez & Kulla 2017,
                                D = (2 + 1/a) sin^(1/a)(th_h) / (2 pi), integrates
                                D cos(th_h) to 1 over the hemisphere (quadrature here);
                                the oracle's lobe equals D V / (4 nv nl) nl with the
                                paper's shadowing V = 1 / (1 + L(nv) + L(nl)),
                                Lambda from the paper's fitted curve and its table;
                                reciprocity f(v, l) = f(l, v)
  dielectric (KF:1132-1286,     Snell's law for the refracted samples at the roughness
   2183-2328, 2881-3145)        floor (air -> material, as the reference evaluates
                                every interface); the share of reflected samples is
                                Schlick's F(theta) with F0 = ((n-1)/(n+1))^2, within
                                the known gap to the exact unpolarised Fresnel
                                reflectance; thin-walled refraction goes straight
                                through; smooth reflection keeps the energy (weight 1)
  specular F0 (KF:357-422)      normal-incidence directional albedo of a smooth
                                dielectric = ((1-n)/(1+n))^2 * specular * specularColor
  diffuse transmission          back side: c * t / pi * |cos| (Lambert through the
   (KF:470-473)                 surface); front side keeps albedo (1 - t) / pi * cos
  energy compensation           the extra lobe of a metal is Kulla & Conty's 2017
   (KF:5-91, base specular)     f_ms = (1-E(mu_o))(1-E(mu_i)) / (pi (1-E_avg)) *
                                F_avg^2 E_avg / (1 - F_avg (1 - E_avg)), with F_avg the
                                cosine-weighted hemispherical mean of Schlick's F
                                (quadrature here) and E, E_avg the reference's curves;
                                its hemispherical integral has the closed form that
                                follows from those curves
  camera (renderer.frag:        the centre pixel without jitter looks at the target;
   129-142)                     a pixel's ray makes the angle that the film geometry
                                (tan(fov/2) at unit distance, aspect W/H) gives; the
                                jitter moves the ray by r/W of the film's half-width
  point light (lighting.glsl    a Lambert plane seen from the light's position:
   2-22)                        L = (1/0.9) rho/pi E cos / d^2 at every pixel
"""
import math

import numpy as np
import pytest

import oracle_ref as O


def _mat(**kw):
    from dspt import MaterialData
    return MaterialData(**kw).data


def _geo(n=(0, 0, 1), wi=(0.3, -0.2, 0.9)):
    n = np.asarray(n, np.float64)
    n /= np.linalg.norm(n)
    wi = np.asarray(wi, np.float64)
    wi /= np.linalg.norm(wi)
    t = np.cross(n, [-n[2], n[0], -n[1]])
    t /= np.linalg.norm(t)
    return np.concatenate([n, n, t, [1.0], wi]).astype(np.float32), wi.astype(np.float32).astype(np.float64)


def _f32(v):
    return np.asarray(v, np.float32).astype(np.float64)


# ------------------------------------------------------------------ sheen
# Estevez & Kulla 2017, "Production Friendly Microfacet Sheen BRDF", table of
# the fitted curve L(x) = a / (1 + b x^c) + d x + e at r = 0 and r = 1, mixed by
# (1 - r)^2 where r is the sheen roughness alpha.
EK = {"a": (25.3245, 21.5473), "b": (3.32435, 3.82987), "c": (0.16801, 0.19823), "d": (-1.27393, -1.97760),
      "e": (-4.85967, -4.32054)}


def _ek_L(x, alpha):
    t = (1.0 - alpha) ** 2
    p = {k: v[1] * (1 - t) + v[0] * t for k, v in EK.items()}
    return p["a"] / (1.0 + p["b"] * abs(x) ** p["c"]) + p["d"] * x + p["e"]


def _ek_lambda(mu, alpha):
    if abs(mu) < 0.5:
        return math.exp(_ek_L(mu, alpha))
    return math.exp(2.0 * _ek_L(0.5, alpha) - _ek_L(1.0 - mu, alpha))


def _charlie_D(alpha, nh):
    s2 = max(1.0 - nh * nh, 0.001)  # the reference floors sin^2 at 1e-3
    inv = 1.0 / max(alpha, 1e-5)
    return (2.0 + inv) * s2 ** (0.5 * inv) / (2.0 * math.pi)


@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.6, 1.0])
def test_charlie_distribution_is_normalised(alpha):
    th = np.linspace(0, math.pi / 2, 200001)
    f = np.array([_charlie_D(alpha, math.cos(x)) for x in th[::50]])
    t = th[::50]
    integral = np.trapezoid(f * np.cos(t) * np.sin(t) * 2 * math.pi, t)
    assert abs(integral - 1.0) < 2e-3, integral


def test_sheen_lobe_is_estevez_kulla():
    rng = np.random.default_rng(3)
    for r in (0.25, 0.5, 0.8, 1.0):
        sc = np.array([0.9, 0.6, 0.3])
        mat = _mat(baseColorFactor=[0, 0, 0], specularFactor=0.0, sheenColorFactor=sc, sheenRoughnessFactor=r)
        alpha = float(np.float32(r) * np.float32(r))
        for wi in ((0.3, -0.2, 0.9), (0.9, 0.1, 0.3), (0.0, 0.0, 1.0)):
            geo, v = _geo(wi=wi)
            for _ in range(12):
                l = rng.normal(size=3)
                l[2] = abs(l[2]) + 0.05
                l = _f32(l / np.linalg.norm(l))
                e = O.bsdf_eval(mat, geo, l.astype(np.float32))
                nv, nl = v[2], l[2]
                h = (v + l) / np.linalg.norm(v + l)
                V = 1.0 / (1.0 + _ek_lambda(nv, alpha) + _ek_lambda(nl, alpha))
                ref = _f32(sc) * (V * _charlie_D(alpha, h[2]) / max(4 * nv * nl, 1e-5)) * nl
                np.testing.assert_allclose(e[:3], ref, rtol=2e-4)
                # reciprocity of the BRDF: f(v, l) = e / nl
                geo2, _ = _geo(wi=l)
                e2 = O.bsdf_eval(mat, geo2, v.astype(np.float32))
                np.testing.assert_allclose(e[:3] / nl, e2[:3] / nv, rtol=1e-4)


# ------------------------------------------------------------------ dielectric transmission
def _fresnel_exact(cos_i, n):
    s2 = (1 - cos_i * cos_i) / (n * n)
    cos_t = np.sqrt(1 - s2)
    rs = (cos_i - n * cos_t) / (cos_i + n * cos_t)
    rp = (n * cos_i - cos_t) / (n * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp)


def _schlick(cos_i, n):
    f0 = ((n - 1) / (n + 1)) ** 2
    return f0 + (1 - f0) * (1 - cos_i) ** 5


@pytest.mark.parametrize("ior", [1.33, 1.5, 2.0])
def test_smooth_dielectric_snell_and_fresnel_split(ior):
    mat = _mat(baseColorFactor=[1, 1, 1], transmissionFactor=1.0, roughnessFactor=0.0, ior=ior, thinWalled=0)
    N = 2000
    for deg in (0.0, 25.0, 50.0, 70.0):
        th = math.radians(deg)
        geo, v = _geo(wi=(math.sin(th), 0.0, math.cos(th)))
        rng = np.random.default_rng(int(deg) + 7)
        rnd = np.stack([rng.uniform(size=N), rng.uniform(size=N), rng.uniform(size=N),
                        (np.arange(N) + 0.5) / N], -1).astype(np.float32)
        s = O.bsdf_sample(mat, geo, rnd)
        d = s[:, :3]
        refl = d[:, 2] > 0
        cos_i = v[2]
        # share of reflected samples: Schlick, within its gap to exact Fresnel
        assert abs(refl.mean() - _schlick(cos_i, ior)) < 2.0 / N, (deg, refl.mean(), _schlick(cos_i, ior))
        assert abs(refl.mean() - _fresnel_exact(cos_i, ior)) < 0.02
        # Snell: sin(t) = sin(i) / n for the refracted directions (roughness floor 0.002)
        sin_t = np.sqrt(np.maximum(0, 1 - d[~refl, 2] ** 2))
        assert np.abs(sin_t - math.sin(th) / ior).max() < 3e-3
        # the refracted ray stays in the plane of incidence, on the far side of the normal
        assert np.abs(d[~refl, 1]).max() < 3e-3 and (d[~refl, 0] <= 3e-3).all()
        # reflection: the mirror direction, energy kept (weight 1 for a white interface)
        mirror = np.array([-v[0], -v[1], v[2]])
        assert np.abs(d[refl] - mirror).max() < 3e-3
        np.testing.assert_allclose(s[refl, 4:7], 1.0, rtol=2e-2)


def test_thin_walled_transmission_goes_straight_through():
    mat = _mat(baseColorFactor=[1, 1, 1], transmissionFactor=1.0, roughnessFactor=0.0, ior=1.5, thinWalled=1)
    N = 1000
    for deg in (10.0, 45.0, 65.0):
        th = math.radians(deg)
        geo, v = _geo(wi=(math.sin(th), 0.0, math.cos(th)))
        rng = np.random.default_rng(5)
        rnd = np.stack([rng.uniform(size=N), rng.uniform(size=N), rng.uniform(size=N),
                        (np.arange(N) + 0.5) / N], -1).astype(np.float32)
        s = O.bsdf_sample(mat, geo, rnd)
        d = s[:, :3]
        through = d[:, 2] < 0
        assert through.mean() > 0.5
        assert np.abs(d[through] + v).max() < 3e-3


@pytest.mark.parametrize("ior,spec,tint", [(1.5, 1.0, (1, 1, 1)), (1.8, 0.6, (1.0, 0.5, 0.25)), (1.3, 0.3, (0.2, 0.9, 1.0))])
def test_dielectric_f0_with_specular_tint(ior, spec, tint):
    mat = _mat(baseColorFactor=[0, 0, 0], metallicFactor=0.0, roughnessFactor=0.0, ior=ior, specularFactor=spec,
               specularColorFactor=tint)
    geo, v = _geo(wi=(0.0, 0.0, 1.0))
    rng = np.random.default_rng(9)
    s = O.bsdf_sample(mat, geo, rng.uniform(size=(400, 4)).astype(np.float32))
    up = s[:, 2] > 0
    assert up.mean() > 0.99
    albedo = s[up, 4:7].mean(0)
    f0 = ((1 - _f32(ior)) / (1 + _f32(ior))) ** 2 * _f32(spec) * _f32(tint)
    np.testing.assert_allclose(albedo, f0, rtol=5e-3)


def test_diffuse_transmission_is_back_side_lambert():
    rng = np.random.default_rng(4)
    for _ in range(4):
        b, c, t = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3), float(rng.uniform(0.1, 1))
        mat = _mat(baseColorFactor=b, specularFactor=0.0, diffuseTransmissionFactor=t, diffuseTransmissionColorFactor=c)
        geo, v = _geo(wi=(0.2, 0.1, 0.9))
        for _ in range(10):
            l = rng.normal(size=3)
            l /= np.linalg.norm(l)
            l32 = _f32(l)
            e = O.bsdf_eval(mat, geo, l32.astype(np.float32))
            if l32[2] < -1e-3:
                np.testing.assert_allclose(e[:3], _f32(c) * _f32(t) / math.pi * abs(l32[2]), rtol=1e-5)
            elif l32[2] > 1e-3:
                np.testing.assert_allclose(e[:3], _f32(b) * (1 - _f32(t)) / math.pi * l32[2], rtol=1e-5)


# ------------------------------------------------------------------ energy compensation
def _ref_E(alpha, mu):
    """The reference's analytic curves (material_kernel.glsl:5-91)."""
    return min(max(1 - math.sqrt(alpha) * (0.5 + 0.5 * (1 - min(max(mu, 0), 1)) ** 5), 0), 1)


def _ref_Eavg(alpha):
    r = math.sqrt(alpha)
    return min(max(1 - 0.58 * r + 0.08 * r * r, 0), 1)


def _ggx_ss(f0, alpha, v, l):
    h = (v + l) / np.linalg.norm(v + l)
    nv, nl, nh, vh = v[2], l[2], h[2], v @ h
    D = alpha ** 2 / (math.pi * (nh * nh * (alpha ** 2 - 1) + 1) ** 2)

    def lam(c):
        return 0.5 * (-1 + math.sqrt(1 + alpha ** 2 * (1 - c * c) / (c * c)))
    G = 1.0 / (1 + lam(nv) + lam(nl))
    F = f0 + (1 - f0) * (1 - vh) ** 5
    return F * D * G / (4 * nv * nl)


def _schlick_avg(f0):
    mu = np.linspace(0, 1, 200001)
    return np.trapezoid(2 * mu * (f0 + (1 - f0) * (1 - mu) ** 5), mu)


@pytest.mark.parametrize("f0", [1.0, 0.5])
def test_metal_energy_compensation_is_kulla_conty(f0):
    rng = np.random.default_rng(6)
    favg = _schlick_avg(f0)
    for r in (0.3, 0.6, 0.9):
        alpha = float(np.float32(r) * np.float32(r))
        mat = _mat(baseColorFactor=[f0, f0, f0], metallicFactor=1.0, roughnessFactor=r)
        ea = _ref_Eavg(alpha)
        for wi in ((0.3, 0.0, 0.95), (0.8, 0.2, 0.4)):
            geo, v = _geo(wi=wi)
            for _ in range(10):
                l = rng.normal(size=3)
                l[2] = abs(l[2]) + 0.1
                l = _f32(l / np.linalg.norm(l))
                e = O.bsdf_eval(mat, geo, l.astype(np.float32))[0]
                ms = (1 - _ref_E(alpha, l[2])) * (1 - _ref_E(alpha, v[2])) / (math.pi * (1 - ea))
                fms = favg * favg * ea / (1 - favg * (1 - ea))
                ref = (_ggx_ss(_f32(f0), alpha, v, l) + ms * fms) * l[2]
                np.testing.assert_allclose(e, ref, rtol=5e-5)


def test_energy_compensation_integral_closed_form():
    # integrating the compensation lobe over l gives, from the reference's curves,
    # (1 - E(mu_v)) * (1 - 2 int E(mu) mu dmu) / (1 - E_avg) * F_ms; with
    # E(mu) = 1 - sqrt(a) (0.5 + 0.5 (1-mu)^5), 2 int (1-mu)^5 mu = 1/21
    for r in (0.3, 0.7):
        alpha = float(np.float32(r) * np.float32(r))
        ea = _ref_Eavg(alpha)
        ebar = 1 - math.sqrt(alpha) * (0.5 + 0.5 / 21)
        mat = _mat(baseColorFactor=[1, 1, 1], metallicFactor=1.0, roughnessFactor=r)
        geo, v = _geo(wi=(0.5, 0.0, 0.8660254))
        # the oracle lobe minus the textbook single-scattering term, integrated over the hemisphere
        nth, nph = 400, 64
        th = (np.arange(nth) + 0.5) / nth * (math.pi / 2)
        ph = (np.arange(nph) + 0.5) / nph * 2 * math.pi
        tot = 0.0
        for t in th:
            for p in ph:
                l = np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])
                l32 = _f32(l)
                e = O.bsdf_eval(mat, geo, l32.astype(np.float32))[0]
                tot += (e - _ggx_ss(1.0, alpha, v, l32) * l32[2]) * math.sin(t)
        tot *= (math.pi / 2 / nth) * (2 * math.pi / nph)
        expect = (1 - _ref_E(alpha, v[2])) * (1 - ebar) / (1 - ea)
        assert abs(tot - expect) < 2e-3 * max(expect, 1e-3) + 1e-4, (r, tot, expect)


# ------------------------------------------------------------------ camera
def test_camera_rays_follow_the_film_geometry():
    from dspt import ViewportCamera, pack_frame_block
    W, H, fov = 101, 61, 50.0
    pos, tgt = np.array([1.0, 2.0, 7.0]), np.array([-0.5, 0.5, 0.0])
    cam = ViewportCamera(position=tuple(pos), target=tuple(tgt), fov=fov, near=0.05)
    blk = pack_frame_block(cam, W, H)
    fwd = (tgt - pos) / np.linalg.norm(tgt - pos)
    o, d = O.camera_ray(blk, W, H, W // 2, H // 2)
    np.testing.assert_allclose(o, pos, rtol=1e-6)
    np.testing.assert_allclose(d, fwd, atol=2e-6)
    M = np.asarray(cam.matrixWorld, np.float64).reshape(4, 4).T
    right, up = M[:3, 0], M[:3, 1]
    th = math.tan(math.radians(fov) / 2)
    for px, py, r0, r1 in ((0, 0, 0.0, 0.0), (W - 1, H - 1, 0.0, 0.0), (17, 44, 0.3, 0.9), (70, 3, 1.0, 0.5)):
        _, d = O.camera_ray(blk, W, H, px, py, r0, r1)
        x = ((2 * (px + 0.5) / W - 1) * (W / H) + r0 / W) * th
        y = ((2 * (py + 0.5) / H - 1) + r1 / H) * th
        want = fwd + x * right + y * up
        want /= np.linalg.norm(want)
        np.testing.assert_allclose(d, want, atol=3e-6)
    # the top row's rays make (almost) half the vertical field of view with the axis
    _, d = O.camera_ray(blk, W, H, W // 2, H - 1, 0.0, 0.0)
    ang = math.degrees(math.acos(float(np.clip(d @ fwd, -1, 1))))
    assert abs(ang - math.degrees(math.atan((1 - 1 / H) * th))) < 1e-3


# ------------------------------------------------------------------ point light
def test_point_light_inverse_square_on_lambert_plane():
    from dspt import MaterialData, PathtracingSceneData, ViewportCamera, build_bvh, pack_frame_block, scenes
    from dspt.scene import Light
    rho, E, h = 0.7, 25.0, 2.0
    s = PathtracingSceneData()
    s.addMaterial(MaterialData(baseColorFactor=[rho, rho, rho], specularFactor=0.0, roughnessFactor=1.0))
    mb = scenes.MeshBuilder()
    P = np.array([[-1e3, 0, -1e3], [1e3, 0, -1e3], [1e3, 0, 1e3], [-1e3, 0, 1e3]], np.float64)
    mb.add(P, np.tile([0, 1, 0], (4, 1)), None, [[0, 2, 1], [0, 3, 2]], 0)
    mb.finish(s)
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    sc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
    lp = np.array([0.0, h, 0.0])
    cam = ViewportCamera(position=tuple(lp), target=(0.0, 0.0, 0.0), up=(0, 0, -1), fov=70.0)
    W, H = 24, 18
    blk = pack_frame_block(cam, W, H, max_bounces=1, use_ibl=False, show_background=False, clamp_threshold=0.0,
                           light=Light(position=tuple(lp), emission=(E, E, E)))
    out = sc.render(blk, W, H, 1, 1)
    L = out["samples"][0, ..., 0].astype(np.float64)
    assert (out["prim"] >= 0).all()
    for py in range(H):
        for px in range(W):
            r = O.rng_seq(1 * 1, px, py, 2)[3]  # the jitter: first two draws of seed frame * maxBounces
            _, d = O.camera_ray(blk, W, H, px, py, float(r[0]), float(r[1]))
            d = d.astype(np.float64)
            dist = h / -d[1]  # the ray from the light's position meets y = 0 at this distance
            cos = -d[1]
            want = (1 / 0.9) * rho / math.pi * E * cos / (dist * dist)
            assert abs(L[py, px] - want) < 2e-3 * want, (px, py, L[py, px], want)
