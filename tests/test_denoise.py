"""Denoising: the DenoisePass of the reference (system/denoise_pass.cpp:31-93,
optix/denoiser.h:9-38) with the closed OptiX model replaced by an
edge-avoiding a-trous wavelet filter (pupiloptixlab_amd/csrc/pt_denoise.hip).

CPU tests pin the float64 restatement (oracle/denoise.py) by closed forms;
GPU tests compare the kernels with it and measure what the filter gains on a
rendered image against a converged render.
"""
import numpy as np
import pytest

import oracle.denoise as dn
import pupiloptixlab_amd as pt


def scene_images(h=40, w=56, seed=0, noise=0.3):
    """Two materials (albedo step), two normal regions, noisy irradiance."""
    rng = np.random.default_rng(seed)
    alb = np.zeros((h, w, 3))
    alb[:, : w // 2] = [0.8, 0.3, 0.2]
    alb[:, w // 2:] = [0.2, 0.5, 0.9]
    nrm = np.zeros((h, w, 3))
    nrm[: h // 2] = [0, 1, 0]
    nrm[h // 2:] = [0, 0, 1]
    irr = np.where(np.arange(h)[:, None, None] < h // 2, 2.0, 0.5) * np.ones((h, w, 3))
    clean = alb * irr
    noisy = clean * (1 + noise * rng.standard_normal((h, w, 1)))
    rgba = np.concatenate([noisy, np.ones((h, w, 1))], -1)
    return rgba, alb, nrm, clean


def test_constant_image_is_a_fixed_point():
    img = np.ones((20, 30, 4)) * [0.3, 0.6, 0.9, 1]
    out = dn.denoise(img)
    assert np.allclose(out[..., :3], img[..., :3], rtol=1e-12)
    alb = np.full((20, 30, 3), 0.5)
    nrm = np.zeros((20, 30, 3))
    nrm[..., 2] = 1
    out = dn.denoise(img, alb, nrm)
    assert np.allclose(out[..., :3], img[..., :3], rtol=1e-12)


def test_albedo_steps_and_normal_edges_do_not_bleed():
    rgba, alb, nrm, clean = scene_images(noise=0.0)
    out = dn.denoise(rgba, alb, nrm)
    assert np.allclose(out[..., :3], clean, rtol=1e-9)  # demodulated irradiance is flat per normal region
    # without guides the albedo step is blurred
    blurred = dn.denoise(rgba, sigma_luminance=1e6)
    assert np.abs(blurred[..., :3] - clean).max() > 0.05


def test_noise_shrinks_in_flat_regions():
    rgba, alb, nrm, clean = scene_images(noise=0.3, seed=1)
    out = dn.denoise(rgba, alb, nrm)
    err_in = np.mean((rgba[..., :3] - clean) ** 2)
    err_out = np.mean((out[..., :3] - clean) ** 2)
    assert err_out < err_in / 10


def test_zero_normals_blend_only_with_each_other():
    h, w = 16, 16
    rgba = np.ones((h, w, 4))
    rgba[:, : w // 2, :3] = 5.0  # "sky" (no surface) on the left
    nrm = np.zeros((h, w, 3))
    nrm[:, w // 2:] = [0, 1, 0]
    out = dn.denoise(rgba, None, nrm, sigma_luminance=1e6)
    assert np.allclose(out[:, : w // 2, :3], 5.0) and np.allclose(out[:, w // 2:, :3], 1.0)


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("guides", ["none", "albedo", "normal", "both"])
def test_gpu_filter_matches_restatement(tracer, guides):
    print(f"case: denoise 56x40 image, guides={guides}", flush=True)
    rgba, alb, nrm, _ = scene_images(seed=2)
    rgba[5:9, 10:14, :3] *= 6.0  # a bright patch
    a = alb if guides in ("albedo", "both") else None
    n = nrm if guides in ("normal", "both") else None
    got = tracer.denoise_image(rgba, a, n)
    want = dn.denoise(rgba.astype(np.float32), None if a is None else a.astype(np.float32),
                      None if n is None else n.astype(np.float32))
    assert np.allclose(got, want, rtol=2e-4, atol=1e-5)


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


@pytest.mark.gpu
def test_gpu_denoise_rendered_image(tracer):
    print("case: Cornell 256x256, 4 spp denoised vs 1024 spp", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=256, height=256)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.set_random_seed(1 << 24)  # reference samples independent of frames 0, 1, ...
    tracer.render(spp=1024, mis_mode=pt.MIS_CONSISTENT, aov=False)
    ref = tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    tracer.reset()
    tracer.render(spp=4, mis_mode=pt.MIS_CONSISTENT)
    noisy = tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    den = tracer.denoise()[..., :3].astype(np.float64)
    assert np.array_equal(tracer.read(pt.BUF_ACCUM)[..., :3], noisy.astype(np.float32))  # input untouched
    r_noisy, r_den = relmse(noisy, ref), relmse(den, ref)
    print(f"relMSE noisy {r_noisy:.4f} denoised {r_den:.4f} gain {r_noisy / r_den:.1f}x")
    assert r_den < r_noisy / 4
    # the unguided filter also helps, less
    plain = tracer.denoise(albedo=False, normal=False)[..., :3].astype(np.float64)
    assert relmse(plain, ref) < r_noisy
