"""ctypes bindings of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU baseline as the
checker of the GPU path tracer.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("mis_mode", C.c_uint32), ("accumulate", C.c_uint32), ("row_begin", C.c_uint32),
                ("row_end", C.c_uint32), ("threads", C.c_uint32)]


class OrCounts(C.Structure):
    _fields_ = [("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64), ("node_visits", C.c_uint64),
                ("tri_tests", C.c_uint64), ("build_ms", C.c_double), ("render_ms", C.c_double),
                ("threads", C.c_uint32)]


_lib = None
_fp = C.POINTER(C.c_float)


def build():
    subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_render.restype = C.c_int
        L.oracle_render.argtypes = [C.c_void_p, C.POINTER(OrParams), C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(OrCounts)]
        L.oracle_trace.restype = C.c_int
        L.oracle_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p]
        L.oracle_rng.restype = None
        L.oracle_rng.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
        L.oracle_sincos.restype = None
        L.oracle_sincos.argtypes = [C.c_float, _fp, _fp]
        L.oracle_camera_ray.restype = None
        L.oracle_camera_ray.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_void_p]
        L.oracle_bsdf_eval.restype = None
        L.oracle_bsdf_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_bsdf_sample.restype = None
        L.oracle_bsdf_sample.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p]
        L.oracle_fresnel_dielectric.restype = C.c_float
        L.oracle_fresnel_dielectric.argtypes = [C.c_float, C.c_float, _fp]
        L.oracle_fresnel_conductor.restype = C.c_float
        L.oracle_fresnel_conductor.argtypes = [C.c_float, C.c_float, C.c_float]
        L.oracle_ggx_d.restype = C.c_float
        L.oracle_ggx_d.argtypes = [C.c_void_p, C.c_float]
        L.oracle_ggx_pdf.restype = C.c_float
        L.oracle_ggx_pdf.argtypes = [C.c_void_p, C.c_void_p, C.c_float]
        L.oracle_ggx_sample.restype = None
        L.oracle_ggx_sample.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_void_p]
        L.oracle_intersect.restype = C.c_int
        L.oracle_intersect.argtypes = [C.c_void_p] * 5 + [C.c_void_p]
        L.oracle_bitmap_sample.restype = None
        L.oracle_bitmap_sample.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
        L.oracle_atan2.restype = C.c_float
        L.oracle_atan2.argtypes = [C.c_float, C.c_float]
        L.oracle_acos.restype = C.c_float
        L.oracle_acos.argtypes = [C.c_float]
        L.oracle_sphere_texcoord.restype = None
        L.oracle_sphere_texcoord.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_intersect_sphere.restype = C.c_int
        L.oracle_intersect_sphere.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float,
                                              _fp]
        L.oracle_emitter_table.restype = C.c_int
        L.oracle_emitter_table.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                           C.POINTER(C.c_uint32), _fp]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def render(desc, spp, max_depth=0, rr_depth=2, mis_mode=0, accumulate=True, rows=None, threads=0):
    """Renders from a reset accumulator; desc is a pt_scene_desc-layout struct.
    Returns dict of accum (h,w,4), albedo, normal, test, prim, sum64, counts."""
    w, h = desc.width, desc.height
    r0, r1 = (0, h) if rows is None else rows
    acc = np.zeros((h, w, 4), np.float32)
    alb = np.zeros((h, w, 3), np.float32)
    nrm = np.zeros((h, w, 3), np.float32)
    tst = np.zeros((h, w), np.float32)
    prim = np.full((h, w), 0xFFFFFFFF, np.uint32)
    s64 = np.zeros((h, w, 3), np.float64)
    cnt = OrCounts()
    p = OrParams(spp, max_depth, rr_depth, mis_mode, int(accumulate), r0, r1, threads)
    rc = lib().oracle_render(C.byref(desc), C.byref(p), acc.ctypes.data, alb.ctypes.data, nrm.ctypes.data,
                             tst.ctypes.data, prim.ctypes.data, s64.ctypes.data, C.byref(cnt))
    assert rc == 0
    return {"accum": acc, "albedo": alb, "normal": nrm, "test": tst, "prim": prim, "sum64": s64,
            "counts": {k: getattr(cnt, k) for k, _ in cnt._fields_}}


def trace(desc, rays, any_hit=False, brute=False):
    rays = _f32(rays)
    n = rays.shape[0]
    hits = np.zeros((n, 4), np.uint32)
    assert lib().oracle_trace(C.byref(desc), rays.ctypes.data, n, int(any_hit), int(brute), hits.ctypes.data) == 0
    return hits[:, 0].view(np.float32), hits[:, 1].copy(), hits[:, 2].view(np.float32), hits[:, 3].view(np.float32)


def rng(pixel, sample, n):
    out = np.zeros(n, np.float32)
    seeds = np.zeros(n, np.uint32)
    lib().oracle_rng(pixel, sample, n, out.ctypes.data, seeds.ctypes.data)
    return out, seeds


def sincos(x):
    s, c = C.c_float(), C.c_float()
    lib().oracle_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def camera_ray(desc, x, y, jx, jy):
    out = np.zeros(6, np.float32)
    lib().oracle_camera_ray(C.byref(desc), x, y, jx, jy, out.ctypes.data)
    return out[:3], out[3:]


def bsdf_eval(mat, wo, wi):
    out = np.zeros(4, np.float32)
    wo, wi = _f32(wo), _f32(wi)
    lib().oracle_bsdf_eval(C.byref(mat), wo.ctypes.data, wi.ctypes.data, out.ctypes.data)
    return out[:3].copy(), float(out[3])


def bsdf_sample(mat, wo, seed):
    out = np.zeros(8, np.float32)
    wo = _f32(wo)
    s = C.c_uint32(seed)
    lib().oracle_bsdf_sample(C.byref(mat), wo.ctypes.data, C.byref(s), out.ctypes.data)
    return {"wi": out[:3].copy(), "f": out[3:6].copy(), "pdf": float(out[6]), "lobe": int(out[7]), "seed": s.value}


def fresnel_dielectric(eta, cos_i):
    ct = C.c_float()
    f = lib().oracle_fresnel_dielectric(eta, cos_i, C.byref(ct))
    return f, ct.value


def fresnel_conductor(eta, k, cos_i):
    return lib().oracle_fresnel_conductor(eta, k, cos_i)


def ggx_d(wh, alpha):
    return lib().oracle_ggx_d(_f32(wh).ctypes.data, alpha)


def ggx_pdf(wo, wh, alpha):
    return lib().oracle_ggx_pdf(_f32(wo).ctypes.data, _f32(wh).ctypes.data, alpha)


def ggx_sample(wo, alpha, u1, u2):
    out = np.zeros(3, np.float32)
    lib().oracle_ggx_sample(_f32(wo).ctypes.data, alpha, u1, u2, out.ctypes.data)
    return out


def intersect(o, d, v0, v1, v2):
    out = np.zeros(3, np.float32)
    hit = lib().oracle_intersect(_f32(o).ctypes.data, _f32(d).ctypes.data, _f32(v0).ctypes.data,
                                 _f32(v1).ctypes.data, _f32(v2).ctypes.data, out.ctypes.data)
    return bool(hit), out


class OrBitmap(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("filter", C.c_uint32), ("wrap", C.c_uint32),
                ("rgba", C.POINTER(C.c_float))]


def bitmap_sample(rgba, u, v, filter=1, wrap=0):
    """Lookup in an (h, w, 4) float image at texture coordinates (u, v)."""
    img = np.ascontiguousarray(rgba, np.float32)
    b = OrBitmap(img.shape[1], img.shape[0], filter, wrap, img.ctypes.data_as(C.POINTER(C.c_float)))
    out = np.zeros(3, np.float32)
    lib().oracle_bitmap_sample(C.byref(b), float(u), float(v), out.ctypes.data)
    return out


def atan2(y, x):
    return lib().oracle_atan2(float(y), float(x))


def acos(z):
    return lib().oracle_acos(float(z))


def sphere_texcoord(d):
    out = np.zeros(2, np.float32)
    lib().oracle_sphere_texcoord(_f32(d).ctypes.data, out.ctypes.data)
    return out


def intersect_sphere(o, d, center, radius, tmin=0.0, tmax=np.inf):
    t = C.c_float()
    hit = lib().oracle_intersect_sphere(_f32(o).ctypes.data, _f32(d).ctypes.data, _f32(center).ctypes.data,
                                        float(radius), float(tmin), float(tmax), C.byref(t))
    return bool(hit), t.value


def emitter_table(desc, capacity=1 << 20):
    sel = np.zeros(capacity, np.float32)
    area = np.zeros(capacity, np.float32)
    cnt = C.c_uint32()
    env = C.c_float()
    lib().oracle_emitter_table(C.byref(desc), sel.ctypes.data, area.ctypes.data, capacity, C.byref(cnt), C.byref(env))
    n = cnt.value
    return sel[:n].copy(), area[:n].copy(), env.value
