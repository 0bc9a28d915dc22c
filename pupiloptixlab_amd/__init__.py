"""MI355X-native wavefront MIS path tracer (PupilOptixLab's path-tracing pass).

Python mirror of the pass interface over the C ABI in ``include/pupil_hip.h``.
The compute path is the gfx950 kernels in ``libpupil_hip.so`` (built in-tree by
``__graft_entry__.build()`` or ``make -C pupiloptixlab_amd``); importing this
package fails loudly when that library is missing.

``PathTracer`` stands in for ``Pupil::pt::PTPass`` (example/path_tracer/
pt_pass.h:21-41): ``set_scene`` is ``SetScene`` (pt_pass.cpp:127-240),
``render`` runs ``OnRun`` once per frame (pt_pass.cpp:39-57) and ``read``
fetches the named buffers the pass allocates ("final result", "pt accum
buffer", "albedo", "normal", "test").
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpupil_hip.so")

PT_OK = 0
STATUS = {
    0: "PT_OK", -1: "PT_ERR_INVALID_ARG", -2: "PT_ERR_HIP", -3: "PT_ERR_NO_SCENE", -4: "PT_ERR_CAPACITY",
    -5: "PT_ERR_IO", -6: "PT_ERR_PARSE", -7: "PT_ERR_OOM", -8: "PT_ERR_UNSUPPORTED", -9: "PT_ERR_NO_BVH",
    -10: "PT_ERR_NO_GUIDE",
}

MAT_DIFFUSE, MAT_DIELECTRIC, MAT_ROUGH_DIELECTRIC, MAT_CONDUCTOR, MAT_ROUGH_CONDUCTOR, MAT_PLASTIC, \
    MAT_ROUGH_PLASTIC = range(1, 8)
TEX_RGB, TEX_CHECKER, TEX_BITMAP = 0, 1, 2
SHAPE_FLIP_NORMALS, SHAPE_FLIP_TEXCOORDS, SHAPE_HAS_NORMALS, SHAPE_HAS_TEXCOORDS, SHAPE_NEG_FACE_NORMAL, \
    SHAPE_SPHERE = 1, 2, 4, 8, 16, 32
PRIM_SPHERE = 1 << 30  # primitive ids of spheres: PRIM_SPHERE | sphere index
ENV_NONE, ENV_CONSTANT, ENV_MAP = 0, 1, 2
FILTER_POINT, FILTER_LINEAR = 0, 1
WRAP_REPEAT, WRAP_CLAMP, WRAP_MIRROR, WRAP_BORDER = 0, 1, 2, 3
MIS_REFERENCE, MIS_CONSISTENT = 0, 1
RENDER_AOV, RENDER_FP64_ACCUM, RENDER_TIMING, RENDER_GUIDE, RENDER_GUIDE_TRAIN = 1, 2, 4, 8, 16
RENDER_ONE_CHAIN = 32
DENOISE_ALBEDO, DENOISE_NORMAL = 1, 2
GUIDE_OUTPUTS, GUIDE_LEVELS = 41, 16
GUIDE_NO_RADIANCE_L2, GUIDE_NO_GUIDE_CE, GUIDE_NO_MIXTURE_CE = 1, 2, 4
BATCH_GRAPH, BATCH_AUTOTUNE, BATCH_GUIDE_GATE = 1, 2, 4
BUF_FRAME, BUF_ACCUM, BUF_ALBEDO, BUF_NORMAL, BUF_TEST, BUF_PRIM_ID, BUF_ACCUM64 = range(7)
SCENE_CORNELL, SCENE_TERRAIN, SCENE_DOOR_ROOM, SCENE_FURNACE, SCENE_CORNELL_VARIANT = 1, 2, 3, 4, 5
MISS = 0xFFFFFFFF


class TextureDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("bitmap", C.c_uint32), ("color0", C.c_float * 3),
                ("color1", C.c_float * 3), ("uv_row0", C.c_float * 4), ("uv_row1", C.c_float * 4)]


class MaterialDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("twosided", C.c_uint32), ("nonlinear", C.c_uint32),
                ("int_ior", C.c_float), ("ext_ior", C.c_float), ("tex", TextureDesc * 4)]


class ShapeDesc(C.Structure):
    _fields_ = [("material", C.c_uint32), ("flags", C.c_uint32), ("first_tri", C.c_uint32),
                ("num_tris", C.c_uint32), ("is_emitter", C.c_int32), ("radiance", TextureDesc)]


class CameraDesc(C.Structure):
    _fields_ = [("sample_to_camera", C.c_float * 16), ("camera_to_world", C.c_float * 16)]


class DenoiseParams(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("iterations", C.c_uint32), ("sigma_luminance", C.c_float),
                ("sigma_normal", C.c_float), ("sigma_albedo", C.c_float)]


class SphereDesc(C.Structure):
    _fields_ = [("to_world", C.c_float * 12)]


class BitmapDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("filter", C.c_uint32), ("wrap", C.c_uint32),
                ("rgba", C.POINTER(C.c_float))]


class SceneDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("num_shapes", C.c_uint32),
                ("num_materials", C.c_uint32), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("texcoords", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("shapes", C.POINTER(ShapeDesc)),
                ("materials", C.POINTER(MaterialDesc)), ("env_type", C.c_uint32),
                ("env_radiance", C.c_float * 3), ("camera", CameraDesc), ("width", C.c_uint32),
                ("height", C.c_uint32), ("max_depth", C.c_uint32), ("num_spheres", C.c_uint32),
                ("spheres", C.POINTER(SphereDesc)), ("num_bitmaps", C.c_uint32), ("bitmaps", C.POINTER(BitmapDesc)),
                ("env_bitmap", C.c_uint32), ("env_scale", C.c_float), ("env_to_world", C.c_float * 9)]


class BvhOpts(C.Structure):
    _fields_ = [("num_bins", C.c_uint32), ("threads", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("mis_mode", C.c_uint32), ("accumulate", C.c_uint32), ("flags", C.c_uint32),
                ("spp_per_launch", C.c_uint32), ("block_size", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64), ("shadow_node_visits", C.c_uint64),
                ("shadow_tri_tests", C.c_uint64), ("ms_total", C.c_double),
                ("ms_raygen", C.c_double), ("ms_trace", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_accumulate", C.c_double), ("launches_trace", C.c_uint64),
                ("launches_shade", C.c_uint64), ("launches_shadow", C.c_uint64), ("bvh_nodes", C.c_uint32),
                ("bvh_leaf_tris", C.c_uint32), ("bvh_depth", C.c_uint32), ("bvh_build_ms", C.c_double),
                ("ms_guide", C.c_double), ("guide_steps", C.c_uint64), ("guide_samples", C.c_uint64),
                ("ms_guide_infer", C.c_double), ("guide_queries", C.c_uint64), ("ms_refit", C.c_double),
                ("refits", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class GuideConfig(C.Structure):
    _fields_ = [("log2_table", C.c_uint32), ("n_min", C.c_uint32), ("n_max", C.c_uint32), ("seed", C.c_uint32),
                ("learning_rate", C.c_float), ("batch_size", C.c_uint32), ("steps_per_frame", C.c_uint32),
                ("train_stride", C.c_uint32), ("loss_flags", C.c_uint32), ("fixed_selection", C.c_float),
                ("l2_weight", C.c_float), ("guided_depth", C.c_uint32)]


class GuideInfo(C.Structure):
    _fields_ = [("enabled", C.c_uint32), ("log2_table", C.c_uint32), ("steps", C.c_uint32),
                ("skipped", C.c_uint32), ("loss_scale", C.c_float), ("last_loss", C.c_float),
                ("last_samples", C.c_uint32), ("resolution", C.c_uint32 * 16), ("dense", C.c_uint32 * 16),
                ("bounds_lo", C.c_float * 3), ("bounds_inv", C.c_float * 3), ("grids", C.c_uint32)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if isinstance(v, C.Array) else v
        return d


class BatchParams(C.Structure):
    _fields_ = [("rr_depth", C.c_uint32), ("mis_mode", C.c_uint32), ("render_flags", C.c_uint32),
                ("batch_flags", C.c_uint32), ("paths_per_launch", C.c_uint32), ("shade_block", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("guide_train_spp", C.c_uint32), ("train_paths_per_launch", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("paths_per_launch", C.c_uint32), ("shade_block", C.c_uint32), ("shade_blocks_per_cu", C.c_uint32),
                ("trace_blocks_per_cu", C.c_uint32), ("candidates", C.c_uint32), ("probe_paths", C.c_uint32),
                ("best_ns_per_path", C.c_double), ("default_ns_per_path", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [("views", C.c_uint32), ("launches", C.c_uint32), ("paths", C.c_uint64),
                ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64), ("triangles", C.c_uint64),
                ("ms_total", C.c_double), ("build_ms", C.c_double), ("guide_gated_off", C.c_uint32),
                ("pad", C.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchGateInfo(C.Structure):
    _fields_ = [("guided_mean", C.c_double), ("guided_median", C.c_double), ("plain_mean", C.c_double),
                ("plain_median", C.c_double), ("guided_frames", C.c_uint32), ("plain_frames", C.c_uint32),
                ("guide_off", C.c_uint32), ("pad", C.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class GenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32), ("grid", C.c_uint32),
                ("spheres", C.c_uint32), ("max_depth", C.c_uint32), ("albedo", C.c_float), ("radiance", C.c_float),
                ("triangles", C.c_uint32)]


# Every symbol the C header declares, with its ctypes signature.
_SIGNATURES = {
    "pt_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pt_destroy": (None, [C.c_void_p]),
    "pt_set_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc)]),
    "pt_build_bvh": (C.c_int, [C.c_void_p, C.POINTER(BvhOpts)]),
    "pt_set_camera": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "pt_reset_accumulation": (C.c_int, [C.c_void_p]),
    "pt_set_shape_transform": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "pt_denoise": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pt_denoise_image": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "pt_refit_bvh": (C.c_int, [C.c_void_p]),
    "pt_render": (C.c_int, [C.c_void_p, C.POINTER(RenderParams)]),
    "pt_render_noalloc": (C.c_int, [C.c_void_p, C.POINTER(RenderParams)]),
    "pt_read_buffer": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "pt_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "pt_set_counting": (C.c_int, [C.c_void_p, C.c_int]),
    "pt_last_error": (C.c_char_p, [C.c_void_p]),
    "pt_trace_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
    "pt_bandwidth_probe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "pt_scene_load_xml": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    "pt_scene_get_warnings": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "pt_scene_generate": (C.c_int, [C.c_int, C.POINTER(GenParams), C.POINTER(C.c_void_p)]),
    "pt_scene_get_desc": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc)]),
    "pt_scene_destroy": (None, [C.c_void_p]),
    "pt_write_exr": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "pt_load_image": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_void_p, C.c_size_t,
                                C.c_char_p, C.c_size_t]),
    "pt_host_trace_rays": (C.c_int, [C.POINTER(SceneDesc), C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "pt_guide_enable": (C.c_int, [C.c_void_p, C.POINTER(GuideConfig)]),
    "pt_guide_disable": (C.c_int, [C.c_void_p]),
    "pt_guide_get_info": (C.c_int, [C.c_void_p, C.POINTER(GuideInfo)]),
    "pt_guide_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "pt_guide_enable_batch": (C.c_int, [C.c_void_p, C.POINTER(GuideConfig)]),
    "pt_set_random_seed": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pt_guide_query_grid": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]),
    "pt_guide_grid_bounds": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "pt_guide_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.POINTER(C.c_float)]),
    "pt_guide_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_float)]),
    "pt_guide_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pt_guide_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pt_guide_read_samples": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "pt_batch_clear": (C.c_int, [C.c_void_p]),
    "pt_batch_add_view": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.c_uint32, C.POINTER(C.c_uint32)]),
    "pt_batch_render": (C.c_int, [C.c_void_p, C.POINTER(BatchParams)]),
    "pt_batch_read": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_size_t]),
    "pt_batch_reset": (C.c_int, [C.c_void_p]),
    "pt_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(BatchParams), C.POINTER(BatchTuning)]),
    "pt_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    "pt_batch_get_gate": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(BatchGateInfo)]),
}

_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Loads libpupil_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C pupiloptixlab_amd` (there is no CPU fallback)")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class PupilError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {STATUS.get(code, code)}")
        self.code = code


def _check(code: int, what: str, ctx=None):
    if code != PT_OK:
        msg = lib().pt_last_error(ctx)
        raise PupilError(code, f"{what} ({msg.decode() if msg else ''})")


class HostScene:
    """A world-space scene on the host (resource::Scene + World packing)."""

    def __init__(self, handle):
        self._h = handle
        self.desc = SceneDesc()
        _check(lib().pt_scene_get_desc(self._h, C.byref(self.desc)), "pt_scene_get_desc")

    @classmethod
    def generate(cls, kind: int, seed: int = 0, width: int = 0, height: int = 0, grid: int = 0,
                 spheres: int = 0, max_depth: int = 0, albedo: float = 0.0, radiance: float = 0.0,
                 triangles: int = 0) -> "HostScene":
        gp = GenParams(seed, width, height, grid, spheres, max_depth, albedo, radiance, triangles)
        h = C.c_void_p()
        _check(lib().pt_scene_generate(kind, C.byref(gp), C.byref(h)), "pt_scene_generate")
        return cls(h)

    @classmethod
    def load_xml(cls, path: str) -> "HostScene":
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        code = lib().pt_scene_load_xml(path.encode(), C.byref(h), err, 512)
        if code != PT_OK:
            raise PupilError(code, f"pt_scene_load_xml({path}): {err.value.decode()}")
        return cls(h)

    def warnings(self) -> list:
        """Loader warnings (unknown conductor names, BSDF and shape types)."""
        n = lib().pt_scene_get_warnings(self._h, None, 0)
        if n <= 0:
            return []
        buf = C.create_string_buffer(256 * n + 1)
        lib().pt_scene_get_warnings(self._h, buf, len(buf))
        return [w for w in buf.value.decode().split("\n") if w]

    def close(self):
        if self._h:
            lib().pt_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # numpy views (valid while the scene lives)
    def positions(self):
        d = self.desc
        return np.ctypeslib.as_array(d.positions, shape=(d.num_vertices, 3))

    def indices(self):
        d = self.desc
        return np.ctypeslib.as_array(d.indices, shape=(d.num_triangles, 3))

    def shapes(self):
        d = self.desc
        return [d.shapes[i] for i in range(d.num_shapes)]

    def bitmaps(self):
        """(width, height, filter, wrap, rgba array) of each bitmap."""
        d = self.desc
        out = []
        for i in range(d.num_bitmaps):
            b = d.bitmaps[i]
            out.append((b.width, b.height, b.filter, b.wrap,
                        np.ctypeslib.as_array(b.rgba, shape=(b.height, b.width, 4))))
        return out

    def spheres(self):
        """Object -> world 3x4 matrices of the analytic spheres (N x 3 x 4)."""
        d = self.desc
        if d.num_spheres == 0:
            return np.zeros((0, 3, 4), dtype=np.float32)
        return np.ctypeslib.as_array(C.cast(d.spheres, C.POINTER(C.c_float)), shape=(d.num_spheres, 3, 4))

    def materials(self):
        d = self.desc
        return [d.materials[i] for i in range(d.num_materials)]

    @property
    def width(self):
        return self.desc.width

    @property
    def height(self):
        return self.desc.height


def host_trace_rays(desc: SceneDesc, rays: np.ndarray, any_hit: bool = False):
    """Traces rays (N x 8: origin, tmin, dir, tmax) on the CPU through the same
    compressed wide BVH the GPU uses.  Returns (t, prim, b1, b2, node_visits,
    tri_tests)."""
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    hits = np.zeros((n, 4), dtype=np.uint32)
    nv, tv = C.c_uint64(), C.c_uint64()
    _check(lib().pt_host_trace_rays(C.byref(desc), rays.ctypes.data, n, int(any_hit), hits.ctypes.data,
                                    C.byref(nv), C.byref(tv)), "pt_host_trace_rays")
    return (hits[:, 0].view(np.float32), hits[:, 1].copy(), hits[:, 2].view(np.float32),
            hits[:, 3].view(np.float32), nv.value, tv.value)


def load_image(path: str) -> np.ndarray:
    """util::BitmapTexture::Load: EXR / Radiance HDR / PNG as (h, w, 4) float32,
    row 0 = top row of the file."""
    w, h = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    code = lib().pt_load_image(path.encode(), C.byref(w), C.byref(h), None, 0, err, 512)
    if code != PT_OK:
        raise PupilError(code, f"pt_load_image({path}): {err.value.decode()}")
    out = np.zeros((h.value, w.value, 4), np.float32)
    _check(lib().pt_load_image(path.encode(), C.byref(w), C.byref(h), out.ctypes.data, out.size, err, 512),
           "pt_load_image")
    return out


def write_exr(path: str, rgba: np.ndarray):
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    _check(lib().pt_write_exr(path.encode(), rgba.ctypes.data, w, h), "pt_write_exr")


class PathTracer:
    """Headless stand-in for the reference path-tracing pass on one GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().pt_create(device, C.byref(h)), "pt_create")
        self._ctx = h
        self.width = 0
        self.height = 0

    def close(self):
        if self._ctx:
            lib().pt_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scene(self, scene) -> None:
        desc = scene.desc if isinstance(scene, HostScene) else scene
        _check(lib().pt_set_scene(self._ctx, C.byref(desc)), "pt_set_scene", self._ctx)
        self.width, self.height = desc.width, desc.height

    def build_bvh(self, num_bins: int = 0, threads: int = 0) -> None:
        _check(lib().pt_build_bvh(self._ctx, C.byref(BvhOpts(num_bins, threads))), "pt_build_bvh", self._ctx)

    def set_camera(self, sample_to_camera, camera_to_world) -> None:
        a = (C.c_float * 16)(*[float(x) for x in np.asarray(sample_to_camera).ravel()])
        b = (C.c_float * 16)(*[float(x) for x in np.asarray(camera_to_world).ravel()])
        _check(lib().pt_set_camera(self._ctx, a, b), "pt_set_camera", self._ctx)

    def reset(self) -> None:
        _check(lib().pt_reset_accumulation(self._ctx), "pt_reset_accumulation", self._ctx)

    def set_random_seed(self, frame: int) -> None:
        """Frame index of the next frame (samples of frame f are drawn from key f)."""
        _check(lib().pt_set_random_seed(self._ctx, frame), "pt_set_random_seed", self._ctx)

    def set_shape_transform(self, shape: int, transform) -> None:
        """RenderObject::UpdateTransform: the shape's pt_set_scene geometry under
        `transform` (3x4, or 4x4 whose last row is ignored)."""
        m = np.ascontiguousarray(np.asarray(transform, np.float32).reshape(-1)[:12])
        _check(lib().pt_set_shape_transform(self._ctx, shape, m.ctypes.data), "pt_set_shape_transform", self._ctx)

    def refit_bvh(self) -> None:
        _check(lib().pt_refit_bvh(self._ctx), "pt_refit_bvh", self._ctx)

    # ---- denoising (DenoisePass) ---------------------------------------------------
    def denoise(self, albedo: bool = True, normal: bool = True, iterations: int = 0, sigma_luminance: float = 0.0,
                sigma_normal: float = 0.0, sigma_albedo: float = 0.0) -> np.ndarray:
        """Denoises the accumulation buffer into the frame buffer and returns it."""
        p = DenoiseParams((DENOISE_ALBEDO if albedo else 0) | (DENOISE_NORMAL if normal else 0), iterations,
                          sigma_luminance, sigma_normal, sigma_albedo)
        _check(lib().pt_denoise(self._ctx, C.byref(p)), "pt_denoise", self._ctx)
        return self.read(BUF_FRAME)

    def denoise_image(self, rgba, albedo=None, normal=None, iterations: int = 0, sigma_luminance: float = 0.0,
                      sigma_normal: float = 0.0, sigma_albedo: float = 0.0) -> np.ndarray:
        rgba = np.ascontiguousarray(rgba, np.float32)
        h, w = rgba.shape[:2]
        a = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
        n = None if normal is None else np.ascontiguousarray(normal, np.float32)
        flags = (DENOISE_ALBEDO if a is not None else 0) | (DENOISE_NORMAL if n is not None else 0)
        p = DenoiseParams(flags, iterations, sigma_luminance, sigma_normal, sigma_albedo)
        out = np.zeros((h, w, 4), np.float32)
        _check(lib().pt_denoise_image(self._ctx, w, h, rgba.ctypes.data, None if a is None else a.ctypes.data,
                                      None if n is None else n.ctypes.data, out.ctypes.data, C.byref(p)),
               "pt_denoise_image", self._ctx)
        return out

    def render(self, spp: int = 1, max_depth: int = 0, rr_depth: int = 2, mis_mode: int = MIS_REFERENCE,
               accumulate: bool = True, aov: bool = True, fp64: bool = False, timing: bool = False,
               spp_per_launch: int = 0, guide: bool = False, guide_train: bool = False,
               one_chain: bool = False) -> None:
        """Renders `spp` frames.  guide: sample with the learned proposal;
        guide_train: train it on this render's paths (both need
        guide_enable() and mis_mode=MIS_CONSISTENT); one_chain: run each
        launch as one chain of kernels (same image)."""
        flags = (RENDER_AOV if aov else 0) | (RENDER_FP64_ACCUM if fp64 else 0) | (RENDER_TIMING if timing else 0)
        flags |= (RENDER_GUIDE if guide else 0) | (RENDER_GUIDE_TRAIN if guide_train else 0)
        flags |= RENDER_ONE_CHAIN if one_chain else 0
        p = RenderParams(spp, max_depth, rr_depth, mis_mode, int(accumulate), flags, spp_per_launch, 0)
        _check(lib().pt_render(self._ctx, C.byref(p)), "pt_render", self._ctx)

    def read(self, which: int) -> np.ndarray:
        w, h = self.width, self.height
        shapes = {BUF_FRAME: ((h, w, 4), np.float32), BUF_ACCUM: ((h, w, 4), np.float32),
                  BUF_ALBEDO: ((h, w, 3), np.float32), BUF_NORMAL: ((h, w, 3), np.float32),
                  BUF_TEST: ((h, w), np.float32), BUF_PRIM_ID: ((h, w), np.uint32),
                  BUF_ACCUM64: ((h, w, 3), np.float64)}
        shape, dt = shapes[which]
        out = np.empty(shape, dtype=dt)
        _check(lib().pt_read_buffer(self._ctx, which, out.ctypes.data, out.nbytes), "pt_read_buffer", self._ctx)
        return out

    def stats(self) -> dict:
        s = Stats()
        _check(lib().pt_get_stats(self._ctx, C.byref(s)), "pt_get_stats", self._ctx)
        return s.as_dict()

    def set_counting(self, enable: bool) -> None:
        _check(lib().pt_set_counting(self._ctx, int(enable)), "pt_set_counting", self._ctx)

    def bandwidth_probe(self, nbytes: int = 1 << 30, iters: int = 10) -> float:
        """Achievable device copy bandwidth in GB/s (read + write)."""
        g = C.c_double()
        _check(lib().pt_bandwidth_probe(self._ctx, nbytes, iters, C.byref(g)), "pt_bandwidth_probe", self._ctx)
        return g.value

    def trace_rays(self, rays: np.ndarray, any_hit: bool = False):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), dtype=np.uint32)
        _check(lib().pt_trace_rays(self._ctx, rays.ctypes.data, n, int(any_hit), hits.ctypes.data),
               "pt_trace_rays", self._ctx)
        return hits[:, 0].view(np.float32), hits[:, 1].copy(), hits[:, 2].view(np.float32), hits[:, 3].view(np.float32)

    # ---- learned guiding ---------------------------------------------------------
    def guide_enable(self, log2_table: int = 19, n_min: int = 16, n_max: int = 2048, seed: int = 0,
                     learning_rate: float = 1e-2, batch_size: int = 16384, steps_per_frame: int = 8,
                     train_stride: int = 1, loss_flags: int = 0, fixed_selection: float = 0.0,
                     l2_weight: float = 0.0, guided_depth: int = 0) -> None:
        cfg = GuideConfig(log2_table, n_min, n_max, seed, learning_rate, batch_size, steps_per_frame, train_stride,
                          loss_flags, fixed_selection, l2_weight, guided_depth)
        _check(lib().pt_guide_enable(self._ctx, C.byref(cfg)), "pt_guide_enable", self._ctx)
        self._guide_entries = 1 << log2_table
        self._guide_grids = 1

    def guide_enable_batch(self, log2_table: int = 19, n_min: int = 16, n_max: int = 2048, seed: int = 0,
                           learning_rate: float = 1e-2, batch_size: int = 16384, steps_per_frame: int = 8,
                           train_stride: int = 1, loss_flags: int = 0, fixed_selection: float = 0.0,
                           l2_weight: float = 0.0, guided_depth: int = 0) -> None:
        """Guide for the batch: one hash grid per batch view, one shared MLP."""
        cfg = GuideConfig(log2_table, n_min, n_max, seed, learning_rate, batch_size, steps_per_frame, train_stride,
                          loss_flags, fixed_selection, l2_weight, guided_depth)
        _check(lib().pt_guide_enable_batch(self._ctx, C.byref(cfg)), "pt_guide_enable_batch", self._ctx)
        self._guide_entries = 1 << log2_table
        self._guide_grids = self.guide_info()["grids"]

    def guide_grid_bounds(self, grid: int):
        """(lo, inv) of hash grid `grid`: u = (p - lo) * inv."""
        lo = (C.c_float * 3)()
        inv = (C.c_float * 3)()
        _check(lib().pt_guide_grid_bounds(self._ctx, grid, lo, inv), "pt_guide_grid_bounds", self._ctx)
        return np.array(list(lo), np.float32), np.array(list(inv), np.float32)

    def guide_disable(self) -> None:
        _check(lib().pt_guide_disable(self._ctx), "pt_guide_disable", self._ctx)

    def guide_info(self) -> dict:
        info = GuideInfo()
        _check(lib().pt_guide_get_info(self._ctx, C.byref(info)), "pt_guide_get_info", self._ctx)
        return info.as_dict()

    def guide_query(self, positions: np.ndarray, grid=None) -> np.ndarray:
        """Raw network outputs (N x 41) at world-space positions (N x 3), in
        hash grid `grid` (None: pt_guide_query, grid 0)."""
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((pos.shape[0], GUIDE_OUTPUTS), dtype=np.float32)
        if grid is None:
            _check(lib().pt_guide_query(self._ctx, pos.ctypes.data, pos.shape[0], out.ctypes.data),
                   "pt_guide_query", self._ctx)
        else:
            _check(lib().pt_guide_query_grid(self._ctx, grid, pos.ctypes.data, pos.shape[0], out.ctypes.data),
                   "pt_guide_query_grid", self._ctx)
        return out

    @staticmethod
    def guide_samples(pos, wdir, target, fs_over_p, p_bsdf, valid=None, grid=None) -> np.ndarray:
        """Packs training samples into the N x 12 float layout of the C ABI."""
        n = len(target)
        s = np.zeros((n, 12), dtype=np.float32)
        s[:, 0:3] = pos
        if grid is not None:
            s[:, 3] = grid
        s[:, 4:7] = wdir
        s[:, 8] = target
        s[:, 9] = fs_over_p
        s[:, 10] = p_bsdf
        s[:, 11] = 1.0 if valid is None else valid
        return s

    def guide_train(self, samples: np.ndarray, fp_norm: float = 1.0) -> float:
        s = np.ascontiguousarray(samples, dtype=np.float32)
        loss = C.c_float()
        _check(lib().pt_guide_train(self._ctx, s.ctypes.data, s.shape[0], fp_norm, C.byref(loss)), "pt_guide_train",
               self._ctx)
        return loss.value

    def guide_gradients(self, samples: np.ndarray, fp_norm: float = 1.0):
        """(table_grad [16, T, 2], weight_grad [4, 64, 64], mean loss) without a step."""
        s = np.ascontiguousarray(samples, dtype=np.float32)
        tg = np.zeros(self._guide_table_shape(), dtype=np.float32)
        wg = np.zeros((4, 64, 64), dtype=np.float32)
        loss = C.c_float()
        _check(lib().pt_guide_gradients(self._ctx, s.ctypes.data, s.shape[0], fp_norm, tg.ctypes.data,
                                        wg.ctypes.data, C.byref(loss)), "pt_guide_gradients", self._ctx)
        return tg, wg, loss.value

    def _guide_table_shape(self):
        g = getattr(self, "_guide_grids", 1)
        return (GUIDE_LEVELS, self._guide_entries, 2) if g == 1 else (g, GUIDE_LEVELS, self._guide_entries, 2)

    def guide_read_samples(self, max_samples: int = 1 << 20) -> np.ndarray:
        """Training samples collected by the last training launch (N x 12)."""
        out = np.zeros((max_samples, 12), dtype=np.float32)
        n = C.c_uint32()
        _check(lib().pt_guide_read_samples(self._ctx, out.ctypes.data, max_samples, C.byref(n)),
               "pt_guide_read_samples", self._ctx)
        return out[:n.value]

    def guide_get_params(self):
        tb = np.zeros(self._guide_table_shape(), dtype=np.float32)
        w = np.zeros((4, 64, 64), dtype=np.float32)
        _check(lib().pt_guide_get_params(self._ctx, tb.ctypes.data, w.ctypes.data), "pt_guide_get_params",
               self._ctx)
        return tb, w

    def guide_set_params(self, table=None, weights=None) -> None:
        tb = None if table is None else np.ascontiguousarray(table, dtype=np.float32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        _check(lib().pt_guide_set_params(self._ctx, None if tb is None else tb.ctypes.data,
                                         None if w is None else w.ctypes.data), "pt_guide_set_params", self._ctx)

    # ---- batched views -------------------------------------------------------------
    def batch_clear(self) -> None:
        _check(lib().pt_batch_clear(self._ctx), "pt_batch_clear", self._ctx)
        self._batch_films = []

    def batch_add_view(self, scene, spp: int) -> int:
        desc = scene.desc if isinstance(scene, HostScene) else scene
        idx = C.c_uint32()
        _check(lib().pt_batch_add_view(self._ctx, C.byref(desc), spp, C.byref(idx)), "pt_batch_add_view",
               self._ctx)
        if not hasattr(self, "_batch_films"):
            self._batch_films = []
        self._batch_films.append((desc.width, desc.height))
        return idx.value

    @staticmethod
    def _batch_params(rr_depth, mis_mode, fp64, timing, graph, autotune, paths_per_launch, shade_block,
                      shade_blocks_per_cu, trace_blocks_per_cu, guide=False, guide_train=False, guide_train_spp=0,
                      guide_gate=False, train_paths_per_launch=0):
        rf = (RENDER_FP64_ACCUM if fp64 else 0) | (RENDER_TIMING if timing else 0)
        rf |= (RENDER_GUIDE if guide else 0) | (RENDER_GUIDE_TRAIN if guide_train else 0)
        bf = (BATCH_GRAPH if graph else 0) | (BATCH_AUTOTUNE if autotune else 0) | (BATCH_GUIDE_GATE if guide_gate else 0)
        return BatchParams(rr_depth, mis_mode, rf, bf, paths_per_launch, shade_block, shade_blocks_per_cu,
                           trace_blocks_per_cu, guide_train_spp, train_paths_per_launch)

    def batch_render(self, rr_depth: int = 2, mis_mode: int = MIS_REFERENCE, fp64: bool = False,
                     timing: bool = False, graph: bool = False, autotune: bool = False, paths_per_launch: int = 0,
                     shade_block: int = 0, shade_blocks_per_cu: int = 0, trace_blocks_per_cu: int = 0,
                     guide: bool = False, guide_train: bool = False, guide_train_spp: int = 0,
                     guide_gate: bool = False, train_paths_per_launch: int = 0) -> None:
        """Renders every view's frames.  guide / guide_train (after
        guide_enable_batch, MIS_CONSISTENT): sample the batch guide / train it
        on the first guide_train_spp frames of each view (0 = all).
        guide_gate (with guide_train): switch the guide off for views whose
        guided training frames show no lower error than their plain ones.
        train_paths_per_launch: width of a training launch (0 = 8M paths)."""
        p = self._batch_params(rr_depth, mis_mode, fp64, timing, graph, autotune, paths_per_launch, shade_block,
                               shade_blocks_per_cu, trace_blocks_per_cu, guide, guide_train, guide_train_spp,
                               guide_gate, train_paths_per_launch)
        _check(lib().pt_batch_render(self._ctx, C.byref(p)), "pt_batch_render", self._ctx)

    def batch_autotune(self, rr_depth: int = 2, mis_mode: int = MIS_REFERENCE, graph: bool = False,
                       guide: bool = False) -> dict:
        p = self._batch_params(rr_depth, mis_mode, False, False, graph, False, 0, 0, 0, 0, guide)
        t = BatchTuning()
        _check(lib().pt_batch_autotune(self._ctx, C.byref(p), C.byref(t)), "pt_batch_autotune", self._ctx)
        return t.as_dict()

    def batch_reset(self) -> None:
        _check(lib().pt_batch_reset(self._ctx), "pt_batch_reset", self._ctx)

    def batch_read(self, view: int, which: int = BUF_ACCUM) -> np.ndarray:
        w, h = self._batch_films[view]
        if which == BUF_ACCUM64:
            out = np.empty((h, w, 3), dtype=np.float64)
        else:
            out = np.empty((h, w, 4), dtype=np.float32)
        _check(lib().pt_batch_read(self._ctx, view, which, out.ctypes.data, out.nbytes), "pt_batch_read", self._ctx)
        return out

    def batch_stats(self) -> dict:
        s = BatchStats()
        _check(lib().pt_batch_get_stats(self._ctx, C.byref(s)), "pt_batch_get_stats", self._ctx)
        return s.as_dict()

    def batch_gate(self, view: int) -> dict:
        """What the last gated guided render measured for `view`: mean and
        median per-frame relMSE of its guided and BSDF-only evaluation frames,
        their counts, and whether its guide was switched off."""
        g = BatchGateInfo()
        _check(lib().pt_batch_get_gate(self._ctx, view, C.byref(g)), "pt_batch_get_gate", self._ctx)
        return g.as_dict()
