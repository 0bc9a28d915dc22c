"""Analytic spheres: the mitsuba `sphere` shape, sphere area emitters and the
sphere texcoord (resource/shape.cpp:113-132, render/geometry.h:94-100,
render/emitter/sphere.h, optix/util.h:38-43 and :117-128).

CPU tests pin the oracle's sphere functions against float64 closed forms and
the irradiance of a spherical light, and check the product's loader and host
traversal; the tests marked `gpu` compare the HIP path with the oracle at
equal seeds (same tolerances as the triangle scenes).
"""
import os

import numpy as np
import pytest

import oracle
import pupiloptixlab_amd as pt

REF_SCENES = "/root/reference/data/static"


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def ulp_err(got, want):
    return np.abs(np.float64(got) - want) / np.spacing(np.float32(max(abs(want), 1e-30)))


# ---------------------------------------------------------------- unit functions
def test_atan2_and_acos_accuracy():
    rng = np.random.default_rng(1)
    ys, xs = rng.normal(size=4000).astype(np.float32), rng.normal(size=4000).astype(np.float32)
    ys[:8] = [0, 0, 1, -1, 0, 1e-30, 3, -3]
    xs[:8] = [1, -1, 0, 0, 2, 1, 3, 3]
    worst = 0.0
    for y, x in zip(ys, xs):
        want = np.arctan2(np.float64(y), np.float64(x))
        got = oracle.atan2(y, x)
        assert abs(got - want) <= 4e-7 * max(1.0, abs(want)), (y, x, got, want)
        if abs(want) > 1e-3:
            worst = max(worst, ulp_err(got, want))
    assert worst <= 4.0
    assert oracle.atan2(0.0, 0.0) == 0.0
    for z in np.concatenate([np.linspace(-1, 1, 2001), [-1.0, 1.0, 0.0, 1 - 2 ** -24]]).astype(np.float32):
        want = np.arccos(np.float64(z))
        assert abs(oracle.acos(z) - want) <= 5e-7, z
    assert oracle.acos(1.0) == 0.0 and oracle.acos(2.0) == 0.0


def test_sphere_texcoord_known_values():
    cases = {(1, 0, 0): (0.0, 0.5), (0, 1, 0): (0.25, 0.5), (-1, 0, 0): (0.5, 0.5), (0, -1, 0): (0.75, 0.5),
             (0, 0, 1): (0.0, 0.0), (0, 0, -1): (0.0, 1.0)}
    for d, uv in cases.items():
        assert np.allclose(oracle.sphere_texcoord(np.array(d, np.float32)), uv, atol=2e-7), d
    # phi in [0, 2pi): u in [0, 1); v = theta / pi
    rng = np.random.default_rng(2)
    v = rng.normal(size=(500, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for p in v.astype(np.float32):
        u, w = oracle.sphere_texcoord(p)
        phi = np.arctan2(float(p[1]), float(p[0])) % (2 * np.pi)
        assert 0.0 <= u < 1.0 and abs(u - phi / (2 * np.pi)) < 1e-6
        assert abs(w - np.arccos(np.clip(float(p[2]), -1, 1)) / np.pi) < 2e-6


def roots64(o, d, c, r):
    o, d, c = (np.asarray(x, np.float64) for x in (o, d, c))
    f = o - c
    a, b, cc = d @ d, 2 * f @ d, f @ f - r * r
    disc = b * b - 4 * a * cc
    if disc < 0:
        return None
    s = np.sqrt(disc)
    return sorted([(-b - s) / (2 * a), (-b + s) / (2 * a)])


def test_ray_sphere_matches_float64_roots():
    rng = np.random.default_rng(3)
    hits = 0
    for _ in range(3000):
        c = rng.uniform(-5, 5, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.01, 2.0))
        o = (c + rng.normal(size=3) * rng.uniform(0.0, 20.0)).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        if rng.uniform() < 0.5:
            d = (c - o + rng.normal(size=3) * r).astype(np.float32)  # aim near the sphere
        d /= np.float32(np.linalg.norm(d))
        ref = roots64(o, d, c, float(r))
        hit, t = oracle.intersect_sphere(o, d, c, r, 1e-3, 1e16)
        want = None if ref is None else next((x for x in ref if x >= 1e-3), None)
        if want is None:
            if hit:  # only tolerated at grazing incidence
                f = np.float64(o) - c
                dist = np.linalg.norm(f - (f @ d) * np.float64(d))
                assert abs(dist - r) <= 1e-4 * max(1.0, np.linalg.norm(f))
            continue
        if not hit:
            f = np.float64(o) - c
            dist = np.linalg.norm(f - (f @ d) * np.float64(d))
            assert abs(dist - r) <= 1e-4 * max(1.0, np.linalg.norm(f)) or abs(want - 1e-3) < 1e-5
            continue
        hits += 1
        scale = max(1.0, np.linalg.norm(np.float64(o) - c) + r)
        assert abs(t - want) <= 2e-6 * scale, (t, want)
    assert hits > 500


def test_ray_sphere_surface_origin_and_range():
    c = np.array([0, 0, 0], np.float32)
    # Leaving the surface outward: no hit; inward: the far side at 2 r cos.
    o = np.array([0, 0, 1], np.float32)
    assert not oracle.intersect_sphere(o, np.array([0, 0, 1], np.float32), c, 1.0, 1e-3, 1e16)[0]
    hit, t = oracle.intersect_sphere(o, np.array([0, 0, -1], np.float32), c, 1.0, 1e-3, 1e16)
    assert hit and t == pytest.approx(2.0, abs=1e-6)
    d = np.array([0, 0.6, -0.8], np.float32)
    hit, t = oracle.intersect_sphere(o, d, c, 1.0, 1e-3, 1e16)
    assert hit and t == pytest.approx(1.6, abs=1e-6)
    # From inside the nearest root is behind: the far one is taken.
    hit, t = oracle.intersect_sphere(c, np.array([1, 0, 0], np.float32), c, 2.0, 1e-3, 1e16)
    assert hit and t == pytest.approx(2.0, abs=1e-6)
    # tmax is inclusive like the triangle test, shorter ranges miss.
    hit, t = oracle.intersect_sphere(np.array([0, 0, 5], np.float32), np.array([0, 0, -1], np.float32), c, 1.0,
                                     0.0, 4.0)
    assert hit and t == 4.0
    assert not oracle.intersect_sphere(np.array([0, 0, 5], np.float32), np.array([0, 0, -1], np.float32), c, 1.0,
                                       0.0, 3.999)[0]


# ---------------------------------------------------------------- loader
SPHERES_XML = """<scene version="3.0.0">
  <integrator type="path"><integer name="max_depth" value="{depth}"/></integrator>
  <sensor type="perspective">
    <float name="fov" value="{fov}"/>
    <transform name="to_world"><lookat origin="{cam}" target="{target}" up="0, 1, 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="{w}"/><integer name="height" value="{h}"/></film>
  </sensor>
  <bsdf type="diffuse" id="floor"><rgb name="reflectance" value="{albedo}"/></bsdf>
  <shape type="rectangle">
    <transform name="to_world"><scale x="50" y="50"/><rotate x="1" angle="-90"/></transform>
    <ref id="floor"/>
  </shape>
  <shape type="sphere">
    <point name="center" x="0" y="{height}" z="0"/><float name="radius" value="{radius}"/>
    <bsdf type="diffuse"><rgb name="reflectance" value="0"/></bsdf>
    <emitter type="area"><rgb name="radiance" value="{radiance}"/></emitter>
  </shape>
  {extra}
</scene>"""


def sphere_light_scene(tmp_path, w=2, h=2, fov=0.1, depth=2, albedo=0.5, height=1.0, radius=0.25, radiance=10.0,
                       cam="0, 1, 3", target="0, 0, 0", extra=""):
    text = SPHERES_XML.format(w=w, h=h, fov=fov, depth=depth, albedo=albedo, height=height, radius=radius,
                              radiance=radiance, cam=cam, target=target, extra=extra)
    return pt.HostScene.load_xml(write(tmp_path, "spheres.xml", text))


def test_xml_sphere_placement_and_flags(tmp_path):
    extra = """<shape type="sphere">
      <float name="radius" value="2"/><point name="center" x="1" y="2" z="3"/>
      <transform name="to_world"><scale value="0.5"/><translate value="10, 0, 0"/></transform>
      <boolean name="flip_normals" value="true"/>
      <bsdf type="diffuse"/>
    </shape>"""
    s = sphere_light_scene(tmp_path, extra=extra)
    d = s.desc
    assert d.num_spheres == 2 and d.num_triangles == 2
    m = s.spheres()
    assert np.allclose(m[0], [[0.25, 0, 0, 0], [0, 0.25, 0, 1], [0, 0, 0.25, 0]])
    # transform * T(center) * S(radius): scale 0.5 then +10 in x
    assert np.allclose(m[1], [[1, 0, 0, 10.5], [0, 1, 0, 1], [0, 0, 1, 1.5]])
    shapes = s.shapes()
    sph = [sh for sh in shapes if sh.flags & pt.SHAPE_SPHERE]
    assert [(sh.first_tri, sh.num_tris) for sh in sph] == [(0, 1), (1, 1)]
    # flip_normals is shared by every sphere: the last parsed value wins
    assert all(sh.flags & pt.SHAPE_FLIP_NORMALS for sh in sph)
    assert sph[0].is_emitter == 1 and sph[1].is_emitter == 0


def test_sphere_emitter_table(tmp_path):
    extra = """<shape type="sphere"><float name="radius" value="0.5"/><point name="center" x="3" y="1" z="0"/>
      <bsdf type="diffuse"/><emitter type="area"><rgb name="radiance" value="1, 4, 2"/></emitter></shape>"""
    s = sphere_light_scene(tmp_path, extra=extra)
    sel, area, env = oracle.emitter_table(s.desc)
    assert len(sel) == 2 and env == 0.0
    assert area == pytest.approx([4 * np.pi * 0.25 ** 2, 4 * np.pi * 0.25], rel=1e-6)
    w = np.array([10.0 * area[0], 4.0 * area[1]])  # max-channel radiance x area
    assert sel == pytest.approx(w / w.sum(), rel=1e-5)


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="reference scenes not present")
def test_reference_scenes_load():
    counts = {}
    for name in sorted(os.listdir(REF_SCENES)):
        if name.endswith(".xml"):
            s = pt.HostScene.load_xml(os.path.join(REF_SCENES, name))
            counts[name] = s.desc.num_spheres
    assert len(counts) == 6 and sum(counts.values()) >= 20


# ---------------------------------------------------------------- traversal
def random_rays(n, rng, lo, hi):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.full((n, 1), 1e-3), d, np.full((n, 1), 1e16)], axis=1).astype(np.float32)


def mixed_scene(tmp_path):
    extra = "".join(
        f"""<shape type="sphere"><float name="radius" value="{0.05 + 0.03 * (i % 5)}"/>
          <point name="center" x="{-1 + 0.25 * (i % 9)}" y="{0.1 + 0.2 * (i // 9)}" z="{0.3 * ((i * 7) % 5) - 0.6}"/>
          <bsdf type="diffuse"/></shape>""" for i in range(40))
    extra += """<shape type="cube"><transform name="to_world"><scale value="0.2"/><translate value="0.3, 0.2, 0.1"/>
        </transform><bsdf type="diffuse"/></shape>"""
    return sphere_light_scene(tmp_path, extra=extra)


def test_host_bvh_with_spheres_matches_brute_force(tmp_path):
    s = mixed_scene(tmp_path)
    d = s.desc
    rng = np.random.default_rng(4)
    rays = random_rays(20000, rng, -1.2, 1.2)
    t_o, p_o, b1_o, b2_o = oracle.trace(d, rays, brute=True)
    t_h, p_h, b1_h, b2_h, _, _ = pt.host_trace_rays(d, rays)
    sph = (p_o != pt.MISS) & ((p_o & pt.PRIM_SPHERE) != 0)
    assert sph.sum() > 2000
    assert np.array_equal(p_o, p_h)
    assert np.array_equal(t_o.view(np.uint32), t_h.view(np.uint32))
    # shadow queries agree on occlusion
    a_o = oracle.trace(d, rays, any_hit=True, brute=True)[1] != pt.MISS
    a_h = pt.host_trace_rays(d, rays, any_hit=True)[1] != pt.MISS
    assert np.array_equal(a_o, a_h)


# ---------------------------------------------------------------- integrator
def test_sphere_light_irradiance_closed_form(tmp_path):
    """A diffuse floor under a spherical light: the point below the center
    receives E = pi L (r/d)^2, so its radiance is albedo * L * (r/d)^2, with
    direct lighting only (max depth 2), in both MIS modes."""
    albedo, L, r, h = 0.5, 10.0, 0.25, 1.0
    s = sphere_light_scene(tmp_path, albedo=albedo, radiance=L, radius=r, height=h)
    want = albedo * L * (r / h) ** 2
    for mode in (pt.MIS_REFERENCE, pt.MIS_CONSISTENT):
        out = oracle.render(s.desc, spp=16384, max_depth=2, mis_mode=mode)
        got = float(out["accum"][..., :3].mean())
        assert got == pytest.approx(want, rel=0.02), (mode, got, want)
    # a camera ray that sees the sphere returns its radiance (primary emission)
    s2 = sphere_light_scene(tmp_path, cam="0, 1, 3", target="0, 1, 0", depth=1)
    out = oracle.render(s2.desc, spp=2, max_depth=1)
    assert np.allclose(out["accum"][..., :3], L)
    assert np.all(out["prim"] == pt.PRIM_SPHERE)


# ---------------------------------------------------------------- GPU parity
REL_TOL, PRIM_TOL = 1e-3, 1e-4


def gpu_vs_oracle(tracer, scene, spp, **kw):
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=spp, **kw)
    acc, prim = tracer.read(pt.BUF_ACCUM), tracer.read(pt.BUF_PRIM_ID)
    ref = oracle.render(scene.desc, spp=spp, **{k: v for k, v in kw.items() if k in ("max_depth", "mis_mode")})
    g, r = acc[..., :3].astype(np.float64), ref["accum"][..., :3].astype(np.float64)
    err = np.abs(g - r) / np.maximum(np.abs(r), 1e-6 / REL_TOL)
    print(f"bit-identical pixels {float(np.mean(np.all(g == r, axis=-1))):.4f}")
    return float(np.mean(prim != ref["prim"])), float(np.mean(err.max(axis=-1) > REL_TOL)), acc, ref


def materials_scene(tmp_path):
    """Nine spheres with the seven materials over a checkerboard floor, lit by
    a rectangle light and a sphere light, seen at a slant."""
    mats = ['<bsdf type="diffuse"><rgb name="reflectance" value="0.7, 0.3, 0.2"/></bsdf>',
            '<bsdf type="conductor"><string name="material" value="Au"/></bsdf>',
            '<bsdf type="roughconductor"><string name="material" value="Cu"/><float name="alpha" value="0.2"/></bsdf>',
            '<bsdf type="dielectric"><float name="int_ior" value="1.5"/></bsdf>',
            '<bsdf type="roughdielectric"><float name="alpha" value="0.3"/></bsdf>',
            '<bsdf type="plastic"><rgb name="diffuse_reflectance" value="0.2, 0.4, 0.8"/></bsdf>',
            '<bsdf type="roughplastic"><float name="alpha" value="0.15"/></bsdf>',
            '<bsdf type="twosided"><bsdf type="diffuse"><texture type="checkerboard" name="reflectance">'
            '<rgb name="color0" value="0.9"/><rgb name="color1" value="0.1"/></texture></bsdf></bsdf>',
            '<bsdf type="roughconductor"><float name="alpha" value="0.05"/></bsdf>']
    extra = "".join(f"""<shape type="sphere"><transform name="to_world"><scale value="0.45"/>
        <translate value="{-1.1 + 1.1 * (i % 3)}, 0.45, {-1.1 + 1.1 * (i // 3)}"/></transform>{m}</shape>"""
                    for i, m in enumerate(mats))
    extra += """<shape type="rectangle"><transform name="to_world"><scale value="0.5"/><rotate x="1" angle="90"/>
        <translate value="0, 3, 0"/></transform><bsdf type="diffuse"><rgb name="reflectance" value="0"/></bsdf>
        <emitter type="area"><rgb name="radiance" value="8"/></emitter></shape>"""
    return sphere_light_scene(tmp_path, w=96, h=64, fov=45, depth=8, height=2.2, radius=0.2, radiance=20,
                              cam="0, 2.5, 4.5", target="0, 0.3, 0", extra=extra)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [pt.MIS_REFERENCE, pt.MIS_CONSISTENT])
def test_gpu_sphere_materials_parity(tracer, tmp_path, mode):
    print(f"case: nine spheres 96x64 8spp depth 8 mis={mode}", flush=True)
    s = materials_scene(tmp_path)
    prim_mis, outliers, acc, _ = gpu_vs_oracle(tracer, s, 8, mis_mode=mode)
    print(f"prim mismatch {prim_mis:.2e} radiance outliers {outliers:.2e}")
    assert prim_mis <= PRIM_TOL and outliers == 0.0
    assert np.isfinite(acc).all() and acc[..., :3].mean() > 0.01


@pytest.mark.gpu
def test_gpu_sphere_light_closed_form(tracer, tmp_path):
    print("case: sphere light over a floor, 2x2, 16384 spp, depth 2", flush=True)
    s = sphere_light_scene(tmp_path)
    prim_mis, outliers, acc, _ = gpu_vs_oracle(tracer, s, 4096, max_depth=2)
    assert prim_mis == 0.0 and outliers == 0.0
    tracer.render(spp=12288, max_depth=2)
    got = float(tracer.read(pt.BUF_ACCUM)[..., :3].mean())
    assert got == pytest.approx(0.5 * 10.0 * 0.25 ** 2, rel=0.02)


@pytest.mark.gpu
def test_gpu_trace_rays_spheres(tracer, tmp_path):
    print("case: pt_trace_rays on 40 spheres + cube, 20000 rays", flush=True)
    s = mixed_scene(tmp_path)
    tracer.set_scene(s)
    tracer.build_bvh()
    rays = random_rays(20000, np.random.default_rng(5), -1.2, 1.2)
    t_g, p_g, _, _ = tracer.trace_rays(rays)
    t_o, p_o, _, _ = oracle.trace(s.desc, rays, brute=True)
    assert np.array_equal(p_g, p_o)
    assert np.array_equal(t_g.view(np.uint32), t_o.view(np.uint32))
    occ_g = tracer.trace_rays(rays, any_hit=True)[1] != pt.MISS
    assert np.array_equal(occ_g, oracle.trace(s.desc, rays, any_hit=True, brute=True)[1] != pt.MISS)


@pytest.mark.gpu
def test_gpu_batch_with_spheres(tracer, tmp_path):
    print("case: batch of two sphere scenes vs pt_render", flush=True)
    scenes = [(materials_scene(tmp_path), 3), (mixed_scene(tmp_path), 2)]
    ref = []
    for sc, spp in scenes:
        tracer.set_scene(sc)
        tracer.build_bvh()
        tracer.render(spp=spp, mis_mode=pt.MIS_REFERENCE, aov=False)
        ref.append(tracer.read(pt.BUF_ACCUM))
    tracer.batch_clear()
    for sc, spp in scenes:
        tracer.batch_add_view(sc, spp)
    tracer.batch_render(mis_mode=pt.MIS_REFERENCE)
    for v, r in enumerate(ref):
        assert np.array_equal(tracer.batch_read(v), r), f"view {v}"
    tracer.batch_clear()
