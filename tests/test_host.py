"""Host-side checks (no GPU): C ABI exports and error paths, scene loading,
procedural scenes, the compressed wide BVH against brute force, EXR output."""
import ctypes as C
import math
import os
import re
import struct

import numpy as np
import pytest

import oracle
import pupiloptixlab_amd as pt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- C ABI
def declared_functions():
    text = open(os.path.join(ROOT, "include", "pupil_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(pt_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = pt.lib()
    names = declared_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pt._SIGNATURES), "ctypes signatures out of sync with the header"


def test_struct_layouts_match_header():
    # offsets the C side relies on (checked against the header's field order)
    assert C.sizeof(pt.TextureDesc) == 64
    assert C.sizeof(pt.MaterialDesc) == 20 + 4 * 64
    assert pt.SceneDesc.width.offset == 208
    assert pt.SceneDesc.num_spheres.offset == 220
    assert pt.SceneDesc.spheres.offset == 224
    assert pt.SceneDesc.bitmaps.offset == 240 and pt.SceneDesc.env_to_world.offset == 256
    assert C.sizeof(pt.SceneDesc) == 296
    assert C.sizeof(pt.SphereDesc) == 48 and C.sizeof(pt.BitmapDesc) == 24
    assert C.sizeof(pt.BatchParams) == 40 and pt.BatchParams.guide_train_spp.offset == 32
    assert pt.BatchParams.train_paths_per_launch.offset == 36
    assert pt.GuideInfo.grids.offset == 180 and C.sizeof(pt.GuideInfo) == 184


def test_create_without_device_fails_cleanly():
    # In a container without a GPU pt_create must return an error, not abort.
    h = C.c_void_p()
    rc = pt.lib().pt_create(0, C.byref(h))
    if rc == pt.PT_OK:
        pt.lib().pt_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc in (-2, -1)
    assert pt.lib().pt_last_error(None)


def test_null_arguments_are_rejected():
    lib = pt.lib()
    assert lib.pt_render(None, None) == -1
    assert lib.pt_set_scene(None, None) == -1
    assert lib.pt_read_buffer(None, 0, None, 0) == -1
    assert lib.pt_scene_get_desc(None, None) == -1
    assert lib.pt_write_exr(None, None, 0, 0) == -1


# ---------------------------------------------------------------- procedural scenes
def test_terrain_triangle_count_and_materials():
    s = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=0)
    d = s.desc
    assert d.num_triangles == 2 * 1023 * 1023 + 256 * 1280 + 8 == 2420746
    assert (d.width, d.height, d.max_depth) == (1920, 1080, 16)
    types = {m.type for m in s.materials()}
    assert types == set(range(1, 8))
    emitters = [sh for sh in s.shapes() if sh.is_emitter]
    assert len(emitters) == 4 and all(5 <= sh.radiance.color0[0] <= 20 for sh in emitters)
    assert d.env_type == pt.ENV_CONSTANT and abs(d.env_radiance[0] - 0.1) < 1e-7


def test_generators_are_deterministic_per_seed():
    a = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=4, grid=64, spheres=16)
    b = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=4, grid=64, spheres=16)
    c = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=5, grid=64, spheres=16)
    assert np.array_equal(a.positions(), b.positions())
    assert not np.array_equal(a.positions(), c.positions())


def test_cornell_scene_layout():
    s = pt.HostScene.generate(pt.SCENE_CORNELL)
    d = s.desc
    assert (d.width, d.height, d.max_depth) == (512, 512, 8)
    assert d.num_triangles == 5 * 2 + 2 * 12 + 2
    light = [sh for sh in s.shapes() if sh.is_emitter]
    assert len(light) == 1 and tuple(light[0].radiance.color0) == (17.0, 12.0, 4.0)
    p = s.positions()
    assert p[:, 1].min() == pytest.approx(0.0, abs=1e-6) and p[:, 1].max() == pytest.approx(2.0, abs=1e-6)


def test_door_room_size():
    s = pt.HostScene.generate(pt.SCENE_DOOR_ROOM)
    d = s.desc
    assert 400_000 < d.num_triangles < 700_000
    light = [sh for sh in s.shapes() if sh.is_emitter]
    assert len(light) == 1 and light[0].radiance.color0[0] == 200.0


def test_door_room_variants():
    """Seed 0 is the canonical room (door slit centered at z = 3); other seeds
    move the door, its closet and light along the x = 4 wall and the camera."""
    base = pt.HostScene.generate(pt.SCENE_DOOR_ROOM, width=32, height=18)
    p0 = base.positions()
    light = lambda s: s.positions()[-15:]  # the light (last shape): 1 m x 0.5 m in 0.25 m cells, 5 x 3 vertices
    assert np.allclose(light(base)[:, 2].min(), 2.75) and np.allclose(light(base)[:, 2].max(), 3.25)
    zs = set()
    for seed in (1, 2, 3):
        v = pt.HostScene.generate(pt.SCENE_DOOR_ROOM, seed=seed, width=32, height=18)
        q = v.positions()
        assert q.min(0).tolist() == p0.min(0).tolist() and q.max(0).tolist() == p0.max(0).tolist()
        zl = light(v)[:, 2]
        assert zl.max() - zl.min() == pytest.approx(0.5, abs=1e-5) and 1.25 <= zl.min() and zl.max() <= 4.75
        zs.add(round(float(zl.min()), 4))
        assert list(v.desc.camera.camera_to_world) != list(base.desc.camera.camera_to_world)
    assert len(zs) == 3
    again = pt.HostScene.generate(pt.SCENE_DOOR_ROOM, seed=2, width=32, height=18)
    assert np.array_equal(again.positions(), pt.HostScene.generate(pt.SCENE_DOOR_ROOM, seed=2, width=32,
                                                                   height=18).positions())


# ---------------------------------------------------------------- wide BVH (host traversal) vs brute force
def random_rays(n, rng, lo=-1.2, hi=1.2):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.full((n, 1), 1e-3, np.float32), d, np.full((n, 1), 1e16, np.float32)], axis=1)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_wide_bvh_matches_brute_force(seed):
    s = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=seed, width=8, height=8, grid=48, spheres=24)
    rng = np.random.default_rng(seed)
    rays = random_rays(20000, rng)
    t_b, p_b, b1_b, b2_b = oracle.trace(s.desc, rays, brute=True)
    t_w, p_w, b1_w, b2_w, nodes, tris = pt.host_trace_rays(s.desc, rays)
    assert np.array_equal(p_b, p_w)
    assert np.array_equal(t_b.view(np.uint32), t_w.view(np.uint32))
    assert np.array_equal(b1_b.view(np.uint32), b1_w.view(np.uint32))
    # any-hit with finite tmax
    rays[:, 7] = rng.uniform(0.05, 1.5, len(rays))
    a_b = oracle.trace(s.desc, rays, any_hit=True, brute=True)[1] != pt.MISS
    a_w = pt.host_trace_rays(s.desc, rays, any_hit=True)[1] != pt.MISS
    assert np.array_equal(a_b, a_w)
    assert nodes / len(rays) < 60  # the tree prunes


def test_wide_bvh_degenerate_and_duplicate_triangles():
    """Coplanar stacks, duplicates and zero-area triangles: the (t, prim) rule
    picks the lowest primitive id on ties, independent of tree shape."""
    s = pt.HostScene.generate(pt.SCENE_CORNELL, width=8, height=8)
    d = s.desc
    pos = np.array(s.positions(), np.float32)
    idx = np.array(s.indices(), np.uint32)
    # duplicate every triangle and add degenerate ones
    pos2 = np.concatenate([pos, pos[:3] * 0 + pos[0]], axis=0)
    deg = np.array([[len(pos), len(pos) + 1, len(pos) + 2]], np.uint32)
    idx2 = np.concatenate([idx, idx, deg], axis=0)
    shapes = (pt.ShapeDesc * 1)()
    shapes[0].material, shapes[0].first_tri, shapes[0].num_tris = 0, 0, len(idx2)
    nd = pt.SceneDesc()
    C.memmove(C.byref(nd), C.byref(d), C.sizeof(pt.SceneDesc))
    nd.num_vertices, nd.num_triangles, nd.num_shapes = len(pos2), len(idx2), 1
    pos2 = np.ascontiguousarray(pos2)
    idx2 = np.ascontiguousarray(idx2)
    nd.positions = pos2.ctypes.data_as(C.POINTER(C.c_float))
    nd.indices = idx2.ctypes.data_as(C.POINTER(C.c_uint32))
    nd.normals = None
    nd.texcoords = None
    nd.shapes = shapes
    rays = random_rays(10000, np.random.default_rng(3), -0.9, 0.9)
    rays[:, 1] += 1.0
    _, p_b, _, _ = oracle.trace(nd, rays, brute=True)
    _, p_w, _, _, _, _ = pt.host_trace_rays(nd, rays)
    assert np.array_equal(p_b, p_w)
    hit = p_w != pt.MISS
    assert hit.mean() > 0.5 and np.all(p_w[hit] < len(idx))  # duplicates lose the tie


# ---------------------------------------------------------------- XML loader (resource/scene.cpp, shape.cpp)
def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


XML_BASIC = """<scene version="3.0.0">
  <default name="resx" value="64"/>
  <default name="res" value="7"/>
  <default name="max_depth" value="5"/>
  <integrator type="path"><integer name="max_depth" value="$max_depth"/></integrator>
  <sensor type="perspective">
    <float name="fov" value="60"/>
    <string name="fov_axis" value="y"/>
    <transform name="to_world"><matrix value="-1 0 0 0 0 1 0 1 0 0 -1 6 0 0 0 1"/></transform>
    <film type="hdrfilm"><integer name="width" value="$resx"/><integer name="height" value="32"/></film>
    <sampler type="independent"><integer name="sample_count" value="4"/></sampler>
  </sensor>
  <bsdf type="twosided" id="white"><bsdf type="diffuse"><rgb name="reflectance" value="0.5, 0.25, 0.125"/></bsdf></bsdf>
  <shape type="rectangle">
    <transform name="to_world">
      <translate value="1, 2, 3"/>
      <rotate y="1" angle="90"/>
      <scale value="2"/>
    </transform>
    <ref id="white"/>
  </shape>
  <shape type="cube">
    <bsdf type="roughplastic"><float name="alpha" value="0.2"/><rgb name="diffuse_reflectance" value="0.1"/></bsdf>
  </shape>
  <shape type="rectangle">
    <boolean name="flip_normals" value="true"/>
    <bsdf type="conductor"><string name="material" value="Au"/></bsdf>
    <emitter type="area"><rgb name="radiance" value="3, 4, 5"/></emitter>
  </shape>
  <emitter type="constant"><rgb name="radiance" value="0.2"/></emitter>
</scene>
"""


def test_xml_scene_semantics(tmp_path):
    s = pt.HostScene.load_xml(write(tmp_path, "basic.xml", XML_BASIC))
    d = s.desc
    assert (d.width, d.height, d.max_depth) == (64, 32, 5)
    assert d.num_triangles == 2 + 12 + 2
    # transform order is scale -> rotate -> translate regardless of element order
    p = s.positions()[:4]
    # rectangle corner (-1,-1,0) -> scale 2 -> (-2,-2,0) -> rot y 90 -> (0,-2,2) -> +(1,2,3)
    assert np.allclose(p[0], [1.0, 0.0, 5.0], atol=1e-5)
    m = s.materials()
    assert m[0].type == pt.MAT_DIFFUSE and m[0].twosided == 1
    assert tuple(round(c, 6) for c in m[0].tex[0].color0) == (0.5, 0.25, 0.125)
    assert m[1].type == pt.MAT_ROUGH_PLASTIC and abs(m[1].tex[2].color0[0] - 0.2) < 1e-7
    assert abs(m[1].int_ior - 1.49) < 1e-7 and abs(m[1].ext_ior - 1.000277) < 1e-7  # defaults
    assert m[2].type == pt.MAT_CONDUCTOR and abs(m[2].tex[1].color0[0] - 0.14282) < 1e-6
    shapes = s.shapes()
    assert shapes[2].is_emitter and tuple(shapes[2].radiance.color0) == (3.0, 4.0, 5.0)
    # the rectangle flip flag is shared by all rectangles: the last one wins
    assert shapes[0].flags & pt.SHAPE_FLIP_NORMALS and shapes[2].flags & pt.SHAPE_FLIP_NORMALS
    assert not shapes[1].flags & pt.SHAPE_FLIP_NORMALS
    assert d.env_type == pt.ENV_CONSTANT and abs(d.env_radiance[1] - 0.2) < 1e-7
    # camera: mitsuba matrix with columns 0 and 2 negated; fov along y
    c2w = np.array(d.camera.camera_to_world).reshape(4, 4)
    assert np.allclose(c2w[:3, :3], np.diag([1, 1, 1]), atol=1e-6) and np.allclose(c2w[:3, 3], [0, 1, 6])
    s2c = np.array(d.camera.sample_to_camera, np.float64).reshape(4, 4)
    v = s2c @ np.array([1.0, 1.0, 0.0, 1.0])
    v = v[:3] / v[3]
    assert v[1] / -v[2] == pytest.approx(math.tan(math.radians(30)), rel=1e-5)
    assert v[0] / v[1] == pytest.approx(2.0, rel=1e-5)


CONDUCTOR_XML = """<scene version="3.0.0">
  <sensor type="perspective"><float name="fov" value="40"/>
    <film type="hdrfilm"><integer name="width" value="8"/><integer name="height" value="8"/></film></sensor>
  {shapes}
</scene>
"""


def test_xml_named_conductors(tmp_path):
    """render/material/ior.h:59-174: every named metal loads its eta / k (Al,
    Cr, Fe and Ni among them); an unknown name falls back to the mirror (eta 0,
    k 1) as material.cpp:72-80 does, and the loader records a warning."""
    named = {"Al": ((1.65394, 0.87850, 0.52012), (9.20430, 6.25621, 4.82675)),
             "Cr": ((4.36041, 2.91052, 1.65119), (5.19538, 4.22239, 3.74700)),
             "Au": ((0.14282, 0.37414, 1.43944), (3.97472, 2.38066, 1.59981)),
             "none": ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))}
    shapes = "".join(f'<shape type="rectangle"><bsdf type="{"roughconductor" if i % 2 else "conductor"}">'
                     f'<string name="material" value="{n}"/></bsdf></shape>'
                     for i, n in enumerate(list(named) + ["Unobtainium"]))
    s = pt.HostScene.load_xml(write(tmp_path, "metals.xml", CONDUCTOR_XML.format(shapes=shapes)))
    m = s.materials()
    for i, (name, (eta, k)) in enumerate(named.items()):
        assert tuple(m[i].tex[1].color0) == pytest.approx(eta, abs=1e-6), name
        assert tuple(m[i].tex[2].color0) == pytest.approx(k, abs=1e-6), name
    assert tuple(m[4].tex[1].color0) == (0.0, 0.0, 0.0) and tuple(m[4].tex[2].color0) == (1.0, 1.0, 1.0)
    w = s.warnings()
    assert len(w) == 1 and "Unobtainium" in w[0]


def test_xml_conductor_table_complete():
    """All 73 conductor entries of ior.h load by name (none of them falls back
    to the mirror, which only "none" is)."""
    names = ["a-C", "Ag", "Al", "Al2O3", "AlAs", "AlAs_palik", "AlSb", "AlSb_palik", "Au", "Be", "Be_palik", "Cr",
             "CsI", "CsI_palik", "Cu", "Cu_palik", "Cu2O", "Cu2O_palik", "CuO", "CuO_palik", "d-C", "d-C_palik",
             "Fe", "GaN", "Ge", "Hg", "Hg_palik", "HgTe", "HgTe_palik", "Ir", "Ir_palik", "K", "K_palik", "Li",
             "Li_palik", "MgO", "MgO_palik", "Mo", "Mo_palik", "Na", "Na_palik", "Nb", "Nb_palik", "Ni",
             "Ni_palik", "Rh", "Rh_palik", "Se", "Se_palik", "SiC", "SiC_palik", "SnTe", "SnTe_palik", "Ta",
             "Ta_palik", "Te", "Te_palik", "ThF4", "ThF4_palik", "TiC", "TiC_palik", "TiN", "TiN_palik", "TiO2",
             "TiO2_palik", "VC", "VC_palik", "V", "V_palik", "VN", "VN_palik", "W", "none"]
    import tempfile
    import pathlib
    d = pathlib.Path(tempfile.mkdtemp())
    shapes = "".join(f'<shape type="rectangle"><bsdf type="conductor"><string name="material" value="{n}"/>'
                      f'</bsdf></shape>' for n in names)
    s = pt.HostScene.load_xml(write(d, "all.xml", CONDUCTOR_XML.format(shapes=shapes)))
    assert s.warnings() == []
    m = s.materials()
    assert len(m) == len(names)
    mirror = [n for n, mm in zip(names, m) if tuple(mm.tex[1].color0) == (0.0, 0.0, 0.0)]
    assert mirror == ["none"]


def test_xml_fov_axis_x_and_lookat(tmp_path):
    xml = """<scene version="3.0.0">
      <sensor type="perspective"><float name="fov" value="90"/>
        <transform name="to_world"><lookat origin="0, 0, 5" target="0, 0, 0" up="0, 1, 0"/></transform>
        <film type="hdrfilm"><integer name="width" value="200"/><integer name="height" value="100"/></film>
      </sensor>
      <shape type="cube"/>
    </scene>"""
    s = pt.HostScene.load_xml(write(tmp_path, "la.xml", xml))
    d = s.desc
    s2c = np.array(d.camera.sample_to_camera, np.float64).reshape(4, 4)
    v = s2c @ np.array([1.0, 1.0, 0.0, 1.0])
    v = v[:3] / v[3]
    # fov 90 on x -> tan(half fov_y) = tan(45 deg) * 100/200
    assert v[1] / -v[2] == pytest.approx(0.5, rel=1e-5)
    c2w = np.array(d.camera.camera_to_world).reshape(4, 4)
    # lookat is flipped twice (loader and sensor): camera looks down -z toward the origin
    assert np.allclose(c2w[:3, 2], [0, 0, 1], atol=1e-6) and np.allclose(c2w[:3, 3], [0, 0, 5])
    assert s.materials()[0].type == pt.MAT_DIFFUSE  # no bsdf: diffuse 0.5 default


def test_xml_obj_shape(tmp_path):
    (tmp_path / "quad.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                                       "vn 0 0 1\nf 1/1/1 2/2/1 3/3/1 4/4/1\n")
    xml = """<scene version="3.0.0"><shape type="obj"><string name="filename" value="quad.obj"/></shape></scene>"""
    s = pt.HostScene.load_xml(write(tmp_path, "o.xml", xml))
    assert s.desc.num_triangles == 2
    sh = s.shapes()[0]
    assert sh.flags & pt.SHAPE_FLIP_TEXCOORDS and sh.flags & pt.SHAPE_HAS_NORMALS


def tri_area(p):
    return 0.5 * abs((p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1]))


@pytest.mark.parametrize("rot", [0, 1, 2, 3])
def test_obj_concave_quad_fans_from_the_concave_corner(tmp_path, rot):
    """A dart-shaped quad (corner 2 concave) written with its corners rotated:
    the two triangles must tile the quad (areas sum to its area, 0.5), which
    only the fan from the concave corner does.  A convex quad keeps the fan
    from its first corner (0,1,2), (0,2,3)."""
    dart = [(0, 0), (2, 1), (1, 1), (2, 2)]  # reflex angle at (1, 1)
    order = [(k + rot) % 4 for k in range(4)]
    lines = [f"v {dart[k][0]} {dart[k][1]} 0" for k in order] + ["f 1 2 3 4"]
    (tmp_path / "dart.obj").write_text("\n".join(lines) + "\n")
    xml = """<scene version="3.0.0"><shape type="obj"><string name="filename" value="dart.obj"/></shape></scene>"""
    s = pt.HostScene.load_xml(write(tmp_path, "d.xml", xml))
    p = s.positions().reshape(-1, 3)
    idx = np.ctypeslib.as_array(s.desc.indices, shape=(s.desc.num_triangles * 3,)).reshape(-1, 3)
    areas = [tri_area(p[t]) for t in idx]
    assert len(areas) == 2 and sum(areas) == pytest.approx(0.5, abs=1e-6)
    # the concave corner (1, 1) is the shared fan apex
    apex = p[idx[0, 0]]
    assert tuple(apex[:2]) == (1.0, 1.0) and tuple(p[idx[1, 0]][:2]) == (1.0, 1.0)
    (tmp_path / "sq.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    s = pt.HostScene.load_xml(write(tmp_path, "q.xml", xml.replace("dart.obj", "sq.obj")))
    q = s.positions().reshape(-1, 3)[:, :2]
    assert q.tolist() == [[0, 0], [1, 0], [1, 1], [0, 0], [1, 1], [0, 1]]


def _inside(poly, pt2):
    """Even-odd rule for a 2D point against a polygon (independent of the loader)."""
    x, y = pt2
    c = False
    for k in range(len(poly)):
        (x0, y0), (x1, y1) = poly[k], poly[(k + 1) % len(poly)]
        if (y0 > y) != (y1 > y) and x < x0 + (y - y0) * (x1 - x0) / (y1 - y0):
            c = not c
    return c


def _signed_area(poly):
    return 0.5 * sum(poly[k][0] * poly[(k + 1) % len(poly)][1] - poly[(k + 1) % len(poly)][0] * poly[k][1]
                     for k in range(len(poly)))


CONCAVE_POLYGONS = {
    "L": [(0, 0), (3, 0), (3, 1), (1, 1), (1, 3), (0, 3)],
    "comb": [(0, 0), (5, 0), (5, 3), (4, 3), (4, 1), (3, 1), (3, 3), (2, 3), (2, 1), (1, 1), (1, 3), (0, 3)],
    "star": [(np.cos(a) * r, np.sin(a) * r) for a, r in
             zip(np.linspace(0, 2 * np.pi, 10, endpoint=False), [2.0, 0.8] * 5)],
}


@pytest.mark.parametrize("name", sorted(CONCAVE_POLYGONS))
@pytest.mark.parametrize("rot", [0, 1, 3])
@pytest.mark.parametrize("plane,flip", [("xy", False), ("xy", True), ("yz", False), ("zx", True)])
def test_obj_concave_polygon_is_ear_clipped(tmp_path, name, rot, plane, flip):
    """Faces with five or more corners are ear-clipped on their own plane:
    n - 2 triangles that keep the polygon's winding, tile it (areas sum to
    the shoelace area) and lie inside it (every centroid passes the even-odd
    test), whatever corner the face starts at, its winding and its plane."""
    poly = CONCAVE_POLYGONS[name]
    if flip:
        poly = poly[::-1]
    poly = poly[rot:] + poly[:rot]
    axes = {"xy": (0, 1, 2), "yz": (1, 2, 0), "zx": (2, 0, 1)}[plane]
    lines = []
    for u, v in poly:
        xyz = [0.0, 0.0, 0.0]
        xyz[axes[0]], xyz[axes[1]], xyz[axes[2]] = float(u), float(v), 0.25
        lines.append("v %.9g %.9g %.9g" % tuple(xyz))
    lines.append("f " + " ".join(str(k + 1) for k in range(len(poly))))
    (tmp_path / "p.obj").write_text("\n".join(lines) + "\n")
    xml = """<scene version="3.0.0"><shape type="obj"><string name="filename" value="p.obj"/></shape></scene>"""
    s = pt.HostScene.load_xml(write(tmp_path, "p.xml", xml))
    p = s.positions().reshape(-1, 3)[:, [axes[0], axes[1]]].astype(np.float64)
    idx = np.ctypeslib.as_array(s.desc.indices, shape=(s.desc.num_triangles * 3,)).reshape(-1, 3)
    assert len(idx) == len(poly) - 2
    area = _signed_area(poly)
    tris = [p[t] for t in idx]
    signed = [_signed_area([tuple(q) for q in t]) for t in tris]
    assert all(a * area > 0 for a in signed)  # winding kept, no degenerate triangle
    assert sum(abs(a) for a in signed) == pytest.approx(abs(area), rel=1e-6)
    assert all(_inside(poly, t.mean(axis=0)) for t in tris)


def test_obj_convex_polygon_fans_from_first_corner(tmp_path):
    """A convex pentagon keeps the fan (0, 1, 2), (0, 2, 3), (0, 3, 4)."""
    pent = [(np.cos(a), np.sin(a)) for a in np.linspace(0, 2 * np.pi, 5, endpoint=False)]
    lines = ["v %.9g %.9g 0" % q for q in pent] + ["f 1 2 3 4 5"]
    (tmp_path / "pent.obj").write_text("\n".join(lines) + "\n")
    xml = """<scene version="3.0.0"><shape type="obj"><string name="filename" value="pent.obj"/></shape></scene>"""
    s = pt.HostScene.load_xml(write(tmp_path, "pent.xml", xml))
    p = s.positions().reshape(-1, 3)[:, :2]
    want = np.array([pent[k] for t in ((0, 1, 2), (0, 2, 3), (0, 3, 4)) for k in t], np.float32)
    assert np.array_equal(p, want)


@pytest.mark.parametrize("text,code", [
    ("<scene><shape type='cube'>", pt.STATUS and -6),
    ("<notscene/>", -6),
    ("<scene><emitter type='envmap'><string name='filename' value='sky.exr'/></emitter></scene>", -6),
    ("<scene><shape type='obj'><string name='filename' value='missing.obj'/></shape></scene>", -5),
])
def test_xml_errors(tmp_path, text, code):
    path = write(tmp_path, "bad.xml", text)
    with pytest.raises(pt.PupilError) as e:
        pt.HostScene.load_xml(path)
    assert e.value.code in (code, -5, -6)


def test_xml_missing_file():
    with pytest.raises(pt.PupilError) as e:
        pt.HostScene.load_xml("/nonexistent/scene.xml")
    assert e.value.code == -5


# ---------------------------------------------------------------- EXR (util/texture.cpp:23-85)
def read_exr(path):
    data = open(path, "rb").read()
    assert data[:4] == bytes([0x76, 0x2F, 0x31, 0x01])
    i = 8
    attrs = {}
    while data[i] != 0:
        j = data.index(b"\0", i)
        name = data[i:j].decode()
        k = data.index(b"\0", j + 1)
        size = struct.unpack("<i", data[k + 1:k + 5])[0]
        attrs[name] = data[k + 5:k + 5 + size]
        i = k + 5 + size
    i += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack(f"<{h}Q", data[i:i + 8 * h])
    img = np.zeros((h, w, 3), np.float32)
    for y, off in enumerate(offs):
        yy, size = struct.unpack("<2i", data[off:off + 8])
        row = np.frombuffer(data[off + 8:off + 8 + size], np.float32).reshape(3, w)  # B, G, R
        img[yy] = row[::-1].T
    return img


def test_exr_round_trip_and_row_order(tmp_path):
    h, w = 5, 7
    rgba = np.random.default_rng(0).random((h, w, 4)).astype(np.float32)
    path = str(tmp_path / "x.exr")
    pt.write_exr(path, rgba)
    img = read_exr(path)
    # buffer row 0 is the image bottom: the file's first scanline is the last row
    assert np.array_equal(img, rgba[::-1, :, :3])
