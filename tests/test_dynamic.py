"""Dynamic scene updates: RenderObject::UpdateTransform with the world's
RenderInstanceTransform / RenderInstanceUpdate handling (world/render_object.
cpp:72-75, world/world.cpp:45-59), the emitter reset (world/emitter.cpp:
169-262) and the IAS update (world/ias_manager.cpp:116-151, 187-205).

The BVH is refit on the GPU (topology kept).  Traversal results do not depend
on the tree, so a render after pt_set_shape_transform must equal, bit for
bit, a render of a freshly built scene whose geometry was moved the same way
(here with the same float32 operations in numpy).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import pupiloptixlab_amd as pt

pytestmark = pytest.mark.gpu


def rot(axis, deg):
    a = np.radians(deg)
    c, s = np.cos(a), np.sin(a)
    x, y, z = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def affine(r3, t, scale=1.0):
    m = np.zeros((3, 4), np.float32)
    m[:, :3] = (np.asarray(r3) * scale).astype(np.float32)
    m[:, 3] = np.asarray(t, np.float32)
    return m


def inv_transpose_rows(m):
    """(A^-1)^T of the 3x3 part by cofactors in float64, as float32 rows."""
    a = m[:, :3].astype(np.float64)
    (a00, a01, a02), (a10, a11, a12), (a20, a21, a22) = a
    c00, c01, c02 = a11 * a22 - a12 * a21, a02 * a21 - a01 * a22, a01 * a12 - a02 * a11
    c10, c11, c12 = a12 * a20 - a10 * a22, a00 * a22 - a02 * a20, a02 * a10 - a00 * a12
    c20, c21, c22 = a10 * a21 - a11 * a20, a01 * a20 - a00 * a21, a00 * a11 - a01 * a10
    det = a00 * c00 + a01 * c10 + a02 * c20
    it = np.array([[c00, c10, c20], [c01, c11, c21], [c02, c12, c22]]) / det
    return it.astype(np.float32), det


def apply_rows(m, v, translate=True):
    """Row-by-row float32 evaluation m[r,0]*x + m[r,1]*y + m[r,2]*z (+ m[r,3])."""
    out = np.empty_like(v)
    for r in range(3):
        acc = m[r, 0] * v[:, 0] + m[r, 1] * v[:, 1] + m[r, 2] * v[:, 2]
        out[:, r] = acc + m[r, 3] if translate else acc
    return out


class MovedScene:
    """A copy of a scene description with some shapes moved, in float32."""

    def __init__(self, scene):
        d = scene.desc
        self.base = d
        self.pos = np.array(np.ctypeslib.as_array(d.positions, shape=(d.num_vertices, 3)), np.float32)
        self.nrm = (np.array(np.ctypeslib.as_array(d.normals, shape=(d.num_vertices, 3)), np.float32)
                    if d.normals else None)
        self.idx = np.array(np.ctypeslib.as_array(d.indices, shape=(d.num_triangles, 3)), np.uint32)
        self.pos0, self.nrm0 = self.pos.copy(), None if self.nrm is None else self.nrm.copy()
        self.shapes = (pt.ShapeDesc * d.num_shapes)()
        C.memmove(self.shapes, d.shapes, C.sizeof(pt.ShapeDesc) * d.num_shapes)
        self.flags0 = [self.shapes[i].flags for i in range(d.num_shapes)]
        self.sph = None
        if d.num_spheres:
            self.sph = (pt.SphereDesc * d.num_spheres)()
            C.memmove(self.sph, d.spheres, C.sizeof(pt.SphereDesc) * d.num_spheres)
            self.sph0 = [np.array(self.sph[k].to_world, np.float32).reshape(3, 4) for k in range(d.num_spheres)]

    def move(self, shape, m):
        sh = self.shapes[shape]
        if sh.flags & pt.SHAPE_SPHERE:
            for k in range(sh.first_tri, sh.first_tri + sh.num_tris):
                a = self.sph0[k]
                o = np.empty((3, 4), np.float32)
                for r in range(3):
                    for c in range(3):
                        o[r, c] = m[r, 0] * a[0, c] + m[r, 1] * a[1, c] + m[r, 2] * a[2, c]
                    o[r, 3] = m[r, 0] * a[0, 3] + m[r, 1] * a[1, 3] + m[r, 2] * a[2, 3] + m[r, 3]
                self.sph[k].to_world[:] = o.reshape(-1).tolist()
            return
        verts = np.unique(self.idx[sh.first_tri:sh.first_tri + sh.num_tris])
        self.pos[verts] = apply_rows(m, self.pos0[verts])
        it, det = inv_transpose_rows(m)
        if self.nrm is not None:
            self.nrm[verts] = apply_rows(np.concatenate([it, np.zeros((3, 1), np.float32)], 1), self.nrm0[verts], False)
        self.shapes[shape].flags = self.flags0[shape] ^ (pt.SHAPE_NEG_FACE_NORMAL if det < 0 else 0)

    def desc(self):
        d = pt.SceneDesc()
        C.memmove(C.byref(d), C.byref(self.base), C.sizeof(pt.SceneDesc))
        self._keep = (np.ascontiguousarray(self.pos), None if self.nrm is None else np.ascontiguousarray(self.nrm))
        d.positions = self._keep[0].ctypes.data_as(C.POINTER(C.c_float))
        if self.nrm is not None:
            d.normals = self._keep[1].ctypes.data_as(C.POINTER(C.c_float))
        d.shapes = self.shapes
        if self.sph is not None:
            d.spheres = self.sph
        return d


def render(tracer, spp=4, **kw):
    tracer.render(spp=spp, mis_mode=pt.MIS_CONSISTENT, **kw)
    return tracer.read(pt.BUF_ACCUM)


def test_update_matches_fresh_build_cornell(tracer):
    print("case: Cornell box, move a box and the light, refit vs fresh build", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=96, height=96)
    tracer.set_scene(scene)
    tracer.build_bvh()
    before = render(tracer)
    m_box = affine(rot([0, 1, 0], 25) @ np.diag([1.0, 1.0, -1.0]), [0.3, 0.05, -0.2], 0.8)  # mirrored
    m_light = affine(rot([1, 0, 0], 180), [0.0, 1.9, 0.0])  # turned upside down
    tracer.set_shape_transform(5, m_box)
    tracer.set_shape_transform(7, m_light)
    after = render(tracer)
    st = tracer.stats()
    assert st["refits"] == 1
    moved = MovedScene(scene)
    moved.move(5, m_box)
    moved.move(7, m_light)
    fresh = moved.desc()
    tracer.set_scene(fresh)
    tracer.build_bvh()
    ref = render(tracer)
    assert not np.array_equal(before, after)
    assert np.array_equal(after, ref)
    # the refit tree answers rays exactly like brute force on the moved scene
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.set_shape_transform(5, m_box)
    tracer.set_shape_transform(7, m_light)
    rng = np.random.default_rng(1)
    n = 20000
    o = rng.uniform(-1, 1, (n, 3)).astype(np.float32) + np.float32([0, 1, 0])
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 1e-3), d, np.full((n, 1), 1e16)], 1).astype(np.float32)
    t_g, p_g, _, _ = tracer.trace_rays(rays)
    t_o, p_o, _, _ = oracle.trace(fresh, rays, brute=True)
    assert np.array_equal(p_g, p_o) and np.array_equal(t_g.view(np.uint32), t_o.view(np.uint32))
    # moving back restores the original image
    eye = affine(np.eye(3), [0, 0, 0])
    tracer.set_shape_transform(5, eye)
    tracer.set_shape_transform(7, eye)
    assert np.array_equal(render(tracer), before)


def test_update_terrain_with_normals_and_spheres(tracer, tmp_path):
    print("case: terrain patch + icosphere moved, sphere emitter moved, refit vs fresh", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=4, width=128, height=72, grid=96, spheres=12)
    tracer.set_scene(scene)
    tracer.build_bvh()
    m1 = affine(rot([0.3, 1, 0.2], 40), [0.1, 0.2, -0.1], 1.3)
    m2 = affine(rot([1, 0, 0], -15), [0.0, 0.05, 0.0])
    tracer.set_shape_transform(3, m1)
    tracer.set_shape_transform(17, m2)
    after = render(tracer, spp=3)
    moved = MovedScene(scene)
    moved.move(3, m1)
    moved.move(17, m2)
    tracer.set_scene(moved.desc())
    tracer.build_bvh()
    assert np.array_equal(after, render(tracer, spp=3))


SPHERE_XML = """<scene version="3.0.0"><integrator type="path"><integer name="max_depth" value="4"/></integrator>
  <sensor type="perspective"><float name="fov" value="45"/>
    <transform name="to_world"><lookat origin="0, 2, 5" target="0, 0.5, 0" up="0, 1, 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="80"/><integer name="height" value="60"/></film></sensor>
  <shape type="rectangle"><transform name="to_world"><scale x="20" y="20"/><rotate x="1" angle="-90"/></transform>
    <bsdf type="diffuse"/></shape>
  <shape type="sphere"><point name="center" x="-0.8" y="0.5" z="0"/><float name="radius" value="0.5"/>
    <bsdf type="roughplastic"/></shape>
  <shape type="sphere"><point name="center" x="0.8" y="1.5" z="0"/><float name="radius" value="0.25"/>
    <bsdf type="diffuse"><rgb name="reflectance" value="0"/></bsdf>
    <emitter type="area"><rgb name="radiance" value="20, 15, 10"/></emitter></shape>
</scene>"""


def test_update_spheres_and_errors(tracer, tmp_path):
    print("case: sphere placement updates and rejected updates", flush=True)
    p = tmp_path / "s.xml"
    p.write_text(SPHERE_XML)
    scene = pt.HostScene.load_xml(str(p))
    tracer.set_scene(scene)
    tracer.build_bvh()
    base = render(tracer)
    m_light = affine(rot([0, 0, 1], 30), [-0.5, 0.3, 0.4], 1.5)
    m_ball = affine(rot([0, 1, 0], 90), [0.2, 0.0, -0.3])
    tracer.set_shape_transform(2, m_light)
    tracer.set_shape_transform(1, m_ball)
    after = render(tracer)
    moved = MovedScene(scene)
    moved.move(2, m_light)
    moved.move(1, m_ball)
    tracer.set_scene(moved.desc())
    tracer.build_bvh()
    assert np.array_equal(after, render(tracer))
    # rejected updates leave the scene as it was
    tracer.set_scene(scene)
    tracer.build_bvh()
    squash = affine(np.diag([1.0, 0.3, 1.0]), [0, 0, 0])
    with pytest.raises(pt.PupilError) as e:
        tracer.set_shape_transform(1, squash)
    assert e.value.code == -8
    with pytest.raises(pt.PupilError) as e:
        tracer.set_shape_transform(0, affine(np.zeros((3, 3)), [0, 0, 0]))
    assert e.value.code == -1
    with pytest.raises(pt.PupilError):
        tracer.set_shape_transform(99, squash)
    assert np.array_equal(render(tracer), base)


def test_refit_timing_terrain(tracer):
    print("case: refit of the 2.4M-triangle terrain after moving one icosphere", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=0, width=64, height=36)
    tracer.set_scene(scene)
    tracer.build_bvh()
    before = tracer.stats()
    tracer.set_shape_transform(20, affine(np.eye(3), [0.01, 0.02, 0.0]))
    tracer.refit_bvh()
    st = tracer.stats()
    print(f"refit ms {st['ms_refit'] - before['ms_refit']:.2f}")
    assert st["refits"] == before["refits"] + 1
    assert st["ms_refit"] - before["ms_refit"] < 500.0
    img = render(tracer, spp=1)
    assert np.isfinite(img).all()
