#!/usr/bin/env python3
"""Throughput benchmark of the plain wavefront MIS path tracer on one MI355X.

Workload: the procedural terrain scene at 1920x1080 -- a 1024x1024-vertex fBm
heightfield (2,093,058 triangles), 256 subdivision-3 icospheres (327,680
triangles), four area-light quads, materials drawn from all seven BSDF types,
constant environment 0.1 -- max depth 16, seed 0.  One step is one frame
(one sample per pixel); the default run traces 64 frames (the 64-spp render).

Prints one JSON line (rank 0).  `value` is Msamples/s with the scene and BVH
resident on the GPU before the timed region starts.  `relmse_guided_vs_plain`
is the second half of the metric: relMSE of the plain tracer divided by that
of the guided tracer at equal wall-clock on the door-room scene (1080p, depth
16, against a 4096-spp FP64 plain reference; see guided_experiment).  `batched` is the
multi-view batch (64 Cornell variants), `guided_batched` guided plus batched (2 door-room + 2
terrain variants at 1080p with per-view hash grids and a shared MLP; `--guided-batch-full` runs
the 16 + 16 views).  The roofline block uses
the kernel that takes longest alone; its algorithmic bytes come from a counting
pass over the same frames (identical rays: the tracer is deterministic) and its
time from HIP events recorded during the timed region on the streams of the
launch's concurrent chains (its active time: the union of its dispatches'
intervals); `roofline.single_chain` repeats the frames with every launch as one
chain and prices the kernel alone.  The CPU baseline is the oracle (scalar C++,
multithreaded) on a bounded band of rows of the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Reference renders draw frames from this index on, so they share no samples
# with the renders they are compared with (frames 0, 1, ...).
REF_SEED = 1 << 24
# Algorithmic bytes per unit: a ray record (origin, direction = 28 B) and a
# hit record (16 B) per traced ray, an 80-B compressed node per node visit, a
# 36-B triangle per triangle test, a 4-B result per shadow ray, and the ray
# state a shading step reads and writes per path vertex.
RAY_BYTES, HIT_BYTES, NODE_BYTES, TRI_BYTES = 28, 16, 80, 36
SHADOW_RESULT_BYTES = 4
SHADE_BYTES_PER_VERTEX = 280


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_sample(scene_desc, target_seconds, threads):
    """The oracle on a band of rows through the middle of the terrain frame,
    sized (rows x spp) to about target_seconds of work on `threads` threads."""
    import oracle

    w, h = scene_desc.width, scene_desc.height
    rows = 8 if threads == 1 else 32
    y0 = h // 2 - rows // 2
    r = oracle.render(scene_desc, spp=1, rows=(y0, y0 + rows), threads=threads)
    per_row_ms = max(r["counts"]["render_ms"], 1e-3) / rows
    budget_rows = target_seconds * 1000.0 / per_row_ms
    spp = int(max(1, min(64, budget_rows // h)))
    rows = int(max(4, min(h, budget_rows / spp)))
    y0 = max(0, h // 2 - rows // 2)
    r = oracle.render(scene_desc, spp=spp, rows=(y0, y0 + rows), threads=threads)
    c = r["counts"]
    samples = rows * w * spp
    secs = c["render_ms"] / 1000.0
    return {"msamples_per_s": samples / secs / 1e6, "mrays_per_s": (c["rays_extension"] + c["rays_shadow"]) / secs / 1e6,
            "threads": int(c["threads"]), "rows": (y0, y0 + rows - 1), "spp": spp, "samples": samples,
            "seconds": secs}


def cpu_baseline(scene_desc, target_seconds, threads):
    """The oracle (scalar C++, the same BVH traversal rules and integrator) on
    `threads` threads and on one thread, on bounded samples of the same frame."""
    many = cpu_sample(scene_desc, target_seconds * 0.6, threads)
    one = cpu_sample(scene_desc, target_seconds * 0.4, 1)
    visible = os.cpu_count() or 1
    return {
        "value": many["msamples_per_s"],
        "unit": "Msamples/s",
        "cores": many["threads"],
        "kind": "port",
        "sample": (f"oracle (scalar C++, {many['threads']} threads on {visible} cores visible, {cpu_model()}) on "
                   f"the same terrain frame, rows {many['rows'][0]}-"
                   f"{many['rows'][1]} of 1080, {many['spp']} spp, depth 16: {many['samples']} samples in "
                   f"{many['seconds']:.2f} s, {many['mrays_per_s']:.3f} Mrays/s; 1 thread: rows {one['rows'][0]}-"
                   f"{one['rows'][1]}, {one['spp']} spp, {one['msamples_per_s']:.4f} Msamples/s, "
                   f"{one['mrays_per_s']:.3f} Mrays/s; BVH build excluded"),
        "single_core_msamples_per_s": round(one["msamples_per_s"], 5),
        "cores_visible": visible,
        "cpu_model": cpu_model(),
    }


COPY_PROBE_BYTES = 1 << 30
# Multiply-adds of one network evaluation as executed on MFMA: 32x64 + 3 x 64x64.
GUIDE_FLOPS_PER_QUERY = 2 * (32 * 64 + 3 * 64 * 64)
MFMA_FP16_PEAK_TFLOPS = 2500.0


def relmse(img, ref):
    import numpy as np

    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def batched_experiment(pt, views, seed):
    """Config 3: `views` independent views in one wavefront loop -- seeded
    Cornell-box variants (triangles log-uniform in [50K, 2M], random camera,
    film from {256^2, 512^2, 768x432, 1024^2, 1920x1080}, spp from {4, 16, 64},
    depth 8).  The launch shape comes from the autotuner; one launch's kernels
    are replayed from a hipGraph.  Reports whole-batch Msamples/s with and
    without the graph, and the tuner's gain over the default shape."""
    import numpy as np

    rng = np.random.default_rng(seed)
    t = pt.PathTracer(0)
    t0 = time.perf_counter()
    tris = 0
    samples = 0
    for v in range(views):
        scene = pt.HostScene.generate(pt.SCENE_CORNELL_VARIANT, seed=seed * 1000 + v)
        spp = int(rng.choice([4, 16, 64]))
        t.batch_add_view(scene, spp)
        tris += scene.desc.num_triangles
        samples += scene.width * scene.height * spp
        scene.close()
    t_build = time.perf_counter() - t0
    log(f"batched: {views} views, {tris} tris, {samples} samples per render, scenes + BVHs in {t_build:.1f}s")
    tuning = t.batch_autotune(graph=True)
    log(f"batched: tuning {tuning}")
    t.batch_render(graph=True)  # warm-up
    t.batch_reset()
    st0 = t.batch_stats()
    t0 = time.perf_counter()
    t.batch_render(graph=True)
    t_graph = time.perf_counter() - t0
    st1 = t.batch_stats()
    rays = (st1["rays_extension"] - st0["rays_extension"]) + (st1["rays_shadow"] - st0["rays_shadow"])
    t.batch_reset()
    t0 = time.perf_counter()
    t.batch_render(graph=False)
    t_direct = time.perf_counter() - t0
    st = t.batch_stats()
    finite = all(bool(np.all(np.isfinite(t.batch_read(v)))) for v in range(views))
    t.close()
    out = {
        "workload": (f"{views} seeded Cornell-box variants (50K-2M tris each, random camera, film from "
                     "{256^2, 512^2, 768x432, 1024^2, 1920x1080}, spp from {4, 16, 64}, depth 8) in one wavefront "
                     "loop"),
        "views": views,
        "triangles": tris,
        "samples": samples,
        "build_seconds": round(t_build, 2),
        "msamples_per_s_graph": round(samples / t_graph / 1e6, 2),
        "msamples_per_s_direct": round(samples / t_direct / 1e6, 2),
        "seconds_graph": round(t_graph, 3),
        "seconds_direct": round(t_direct, 3),
        "launches_per_render": st["launches"] // 3 if st["launches"] else None,
        # Rays per sample explain the gap to the single terrain view: closed
        # Cornell boxes keep every path bouncing to depth 8 (extension and
        # shadow rays at each vertex); terrain paths escape to the sky early.
        "rays_per_sample": round(rays / samples, 3),
        "mrays_per_s_graph": round(rays / t_graph / 1e6, 1),
        "tuning": tuning,
        "all_finite": finite,
    }
    log(f"batched: {json.dumps(out)}")
    return out


# Hash-table entries per level (log2) of the guided render judged at equal
# time: a 4 MiB table (2^16 entries x 16 levels) makes inference faster, which
# at equal time outweighs the larger table's accuracy on this workload (door
# room 1080p, 256 training frames, four chains per launch: ratio 3.74-3.83 at
# 2^16, 3.51-3.80 at 2^17, 3.64-3.66 at 2^18, 3.02-3.65 at 2^19, the
# library's default).  The guided reference keeps the default.
GUIDED_LOG2_TABLE = 16


def guided_experiment(pt, ref_spp, plain_spp, train_spp, ref_g_spp=1024):
    """relMSE at equal wall-clock, guided vs plain, on the door room (room lit
    through a 5 cm door gap, ~500K triangles, 1920x1080, max depth 16).

    Reference: ref_spp plain frames accumulated in FP64 (the headline ratio);
    both renders are also judged against a guided FP64 reference of ref_g_spp
    frames (vs_guided_reference).  Plain: plain_spp frames, timed.  Guided: trained online on its own paths during the first
    train_spp frames, then rendering continues, in calls sized to the time left,
    until the plain run's time is used up; the guided time includes training.  Both use
    balance-heuristic MIS with the emitter selection probability on both sides."""
    import numpy as np

    scene = pt.HostScene.generate(pt.SCENE_DOOR_ROOM, seed=0)
    d = scene.desc
    t = pt.PathTracer(0)
    t.set_scene(scene)
    t.build_bvh()
    mis = pt.MIS_CONSISTENT
    t0 = time.perf_counter()
    t.set_random_seed(REF_SEED)  # reference samples independent of the tested frames 0, 1, ...
    t.render(spp=ref_spp, mis_mode=mis, fp64=True, aov=False)
    ref = t.read(pt.BUF_ACCUM64) / ref_spp
    t_ref = time.perf_counter() - t0
    log(f"guided experiment: reference {ref_spp} spp in {t_ref:.1f}s")
    # A second, guided reference (its own network trained on frames 2^23 on,
    # rendering frames 2^24 on): unbiased like the plain one, with about a
    # tenth of a plain 1024-frame image's error.  Against the plain reference
    # its own noise (~0.03 relMSE at 4096 spp) is most of what the guided
    # render's error reads.
    t.reset()
    t.guide_enable(seed=1000)
    t.set_random_seed(REF_SEED >> 1)
    t.render(spp=128, mis_mode=mis, aov=False, guide=True, guide_train=True)
    t.reset()
    t.set_random_seed(REF_SEED)
    t.render(spp=ref_g_spp, mis_mode=mis, fp64=True, aov=False, guide=True)
    ref_g = t.read(pt.BUF_ACCUM64) / ref_g_spp
    t.guide_disable()
    log(f"guided experiment: guided reference {ref_g_spp} spp")
    t.reset()
    t.render(spp=4, mis_mode=mis, aov=False)  # warm-up
    t.reset()
    t0 = time.perf_counter()
    t.render(spp=plain_spp, mis_mode=mis, aov=False)
    t_plain = time.perf_counter() - t0
    img_plain = t.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    e_plain = relmse(img_plain, ref)
    e_plain_g = relmse(img_plain, ref_g)
    t.reset()
    t.guide_enable(log2_table=GUIDED_LOG2_TABLE)
    s0 = t.stats()
    t0 = time.perf_counter()
    t.render(spp=train_spp, mis_mode=mis, aov=False, guide=True, guide_train=True, timing=True)
    t_train = time.perf_counter() - t0
    # Then as many guided frames as fit in the rest of the plain render's
    # time, in calls sized from the measured frame time (at most 64 frames, a
    # launch of the library's default width; the plain render's calls use the
    # same), so the guided side stops within a frame of the plain time.
    n = train_spp
    per_frame = None
    while True:
        remaining = t_plain - (time.perf_counter() - t0)
        k = 8 if per_frame is None else min(64, int(remaining / per_frame))
        if remaining <= 0.0 or k < 1:
            break
        t1 = time.perf_counter()
        t.render(spp=k, mis_mode=mis, aov=False, guide=True, timing=True)
        per_frame = (time.perf_counter() - t1) / k
        n += k
    t_guided = time.perf_counter() - t0
    img = t.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    e_guided = relmse(img, ref)
    e_guided_g = relmse(img, ref_g)
    s1 = t.stats()
    info = t.guide_info()
    q = s1["guide_queries"] - s0["guide_queries"]
    ms_infer = s1["ms_guide_infer"] - s0["ms_guide_infer"]
    infer_tflops = q * GUIDE_FLOPS_PER_QUERY / (ms_infer / 1000.0) / 1e12 if ms_infer > 0 else None
    t.close()
    out = {
        "scene": (f"door room: 4x6x3 m room, walls ~500K tris diffuse 0.6, emitter (radiance 200) behind a 5 cm "
                  f"door gap; {d.width}x{d.height}, max depth {d.max_depth}, {d.num_triangles} tris"),
        "reference_spp": ref_spp,
        "plain_spp": plain_spp,
        "plain_seconds": round(t_plain, 3),
        "relmse_plain": round(e_plain, 5),
        "guided_spp": n,
        "guided_train_spp": train_spp,
        "guide_log2_table": GUIDED_LOG2_TABLE,
        "guided_train_seconds": round(t_train, 3),
        "guided_seconds": round(t_guided, 3),
        "relmse_guided": round(e_guided, 5),
        "ratio": round(e_plain / e_guided, 3),
        "mean_rel_diff": round(float(img.mean() / ref.mean() - 1.0), 5),
        "vs_guided_reference": {"reference": f"{ref_g_spp} guided FP64 frames after 128 training frames",
                                "relmse_plain": round(e_plain_g, 5), "relmse_guided": round(e_guided_g, 5),
                                "ratio": round(e_plain_g / e_guided_g, 3)},
        "guide_steps": info["steps"],
        "guide_skipped_steps": info["skipped"],
        "guide_queries": q,
        "ms_guide_infer": round(ms_infer, 2),
        "ms_guide_train": round(s1["ms_guide"] - s0["ms_guide"], 2),
        "infer_TFLOPs": round(infer_tflops, 2) if infer_tflops else None,
        "infer_mfma_frac": round(infer_tflops / MFMA_FP16_PEAK_TFLOPS, 4) if infer_tflops else None,
    }
    log(f"guided experiment: {json.dumps(out)}")
    return out


def batched_guided_experiment(pt, n_door, n_terrain, width, height, door_ref_spp, terrain_ref_spp, train_spp,
                              terrain_grids, seed, gate=True):
    """Guided plus batched: n_door seeded door-room variants (door and camera
    moved) and n_terrain seeded terrain variants (heightfield grid drawn from
    terrain_grids, 256 icospheres, all seven BSDFs) at width x height, spp per
    view from {16, 64, 256} (each kind cycles through them), in one wavefront
    loop.  The guide has one hash grid per view and one shared MLP, trained
    online on the first train_spp frames of every view (round-robin over the
    views), with optimizer steps per frame scaled to the film size (8 per
    1920x1080 frame, at least 1).

    Plain: the batch with BSDF sampling (hipGraph replay, default launch
    width), timed; rendered again m = round(T_guided / T_plain) times in all so
    that both sides use about the same wall-clock.  Guided: one call, training
    included (training launches of ~2M paths, the rest at the default width).
    relMSE per view against FP64 references (door_ref_spp guided frames after
    128 training frames / terrain_ref_spp plain frames, single-view renders,
    frames disjoint from the tested ones)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    views = []
    for i in range(n_door):
        views.append(("door", pt.HostScene.generate(pt.SCENE_DOOR_ROOM, seed=1 + i, width=width, height=height)))
    for i in range(n_terrain):
        grid = int(rng.choice(terrain_grids))
        views.append(("terrain", pt.HostScene.generate(pt.SCENE_TERRAIN, seed=100 + i, width=width, height=height,
                                                       grid=grid)))
        log(f"guided batch: terrain variant {i} generated ({views[-1][1].desc.num_triangles} tris)")
    spps = [(256, 64, 16)[i % 3] for i in range(n_door)] + [(64, 256, 16)[i % 3] for i in range(n_terrain)]
    tris = [v[1].desc.num_triangles for v in views]
    t = pt.PathTracer(0)
    mis = pt.MIS_CONSISTENT
    refs = []
    t_ref0 = time.perf_counter()
    for i, (kind, scene) in enumerate(views):
        t.set_scene(scene)
        t.build_bvh()
        n = door_ref_spp if kind == "door" else terrain_ref_spp
        if kind == "door":
            # A guided reference: unbiased like a plain one, but its own error
            # at 1024 frames is about a tenth of a plain 1024-frame image's,
            # which would otherwise dominate both sides' relMSE.  Its network
            # is trained on frames 2^23 on and it renders frames 2^24 on, so
            # it shares no samples with the renders it judges.
            t.guide_enable(seed=1000 + i)
            t.set_random_seed(REF_SEED >> 1)
            t.render(spp=128, mis_mode=mis, aov=False, guide=True, guide_train=True)
            t.reset()
            t.set_random_seed(REF_SEED)
            t.render(spp=n, mis_mode=mis, fp64=True, aov=False, guide=True)
            t.guide_disable()
        else:
            t.set_random_seed(REF_SEED)
            t.render(spp=n, mis_mode=mis, fp64=True, aov=False)
        refs.append(t.read(pt.BUF_ACCUM64) / n)
        log(f"guided batch: reference {len(refs)}/{len(views)} ({kind}, {n} spp) at {time.perf_counter() - t_ref0:.1f}s")
    log(f"guided batch: {len(views)} views, {sum(tris)} tris, references in {time.perf_counter() - t_ref0:.1f}s")
    for (kind, scene), spp in zip(views, spps):
        t.batch_add_view(scene, spp)
    t_build = time.perf_counter() - t0
    log(f"guided batch: views added ({t_build:.1f}s in all)")
    samples = sum(width * height * spp for spp in spps)

    def errors():
        return [relmse(t.batch_read(k)[..., :3].astype(np.float64), refs[k]) for k in range(len(views))]

    t.batch_render(mis_mode=mis, graph=True)  # warm-up
    t.batch_reset()
    t0 = time.perf_counter()
    t.batch_render(mis_mode=mis, graph=True)
    t_plain1 = time.perf_counter() - t0
    e_plain1 = errors()
    log(f"guided batch: plain {t_plain1:.2f}s")
    t.batch_reset()
    spf = max(1, int(round(8 * width * height / (1920 * 1080))))
    t.guide_enable_batch(seed=seed, steps_per_frame=spf)
    t0 = time.perf_counter()
    t.batch_render(mis_mode=mis, graph=True, guide=True, guide_train=True, guide_train_spp=train_spp,
                   guide_gate=gate)
    t_guided = time.perf_counter() - t0
    e_guided = errors()
    info = t.guide_info()
    gated_off = t.batch_stats()["guide_gated_off"] if gate else 0
    gates = [t.batch_gate(k) for k in range(len(views))] if gate else [None] * len(views)
    log(f"guided batch: guided {t_guided:.2f}s")
    m = max(1, int(round(t_guided / t_plain1)))
    t.batch_reset()
    t0 = time.perf_counter()
    for r in range(m):
        t.batch_render(mis_mode=mis, graph=True)
        log(f"guided batch: plain equal-time render {r + 1}/{m}")
    t_plain = time.perf_counter() - t0
    e_plain = errors()
    t.close()
    per_view = [{"scene": kind, "triangles": tr, "spp": spp, "relmse_plain_equal_time": round(ep, 5),
                 "relmse_plain_same_spp": round(e1, 5), "relmse_guided": round(eg, 5),
                 **({"gate": {"guide_off": g["guide_off"],
                              "frame_relmse_guided": [round(g["guided_mean"], 4), round(g["guided_median"], 4)],
                              "frame_relmse_plain": [round(g["plain_mean"], 4), round(g["plain_median"], 4)],
                              "frames": [g["guided_frames"], g["plain_frames"]]}} if g else {})}
                for (kind, _), tr, spp, ep, e1, eg, g in zip(views, tris, spps, e_plain, e_plain1, e_guided, gates)]
    mean = lambda xs: float(np.mean(xs)) if xs else None
    # The plain side ran m whole batches (m T_plain, not exactly T_guided):
    # scaled to T_guided with relMSE proportional to 1 / time.
    scale = m * t_plain1 / t_guided
    ep_scaled = [e * scale for e in e_plain]
    kinds = {}
    for kind in ("door", "terrain"):
        idx = [k for k, v in enumerate(views) if v[0] == kind]
        if idx:
            ep = mean([e_plain[k] for k in idx])
            eg = mean([e_guided[k] for k in idx])
            kinds[kind] = {"relmse_plain_equal_time": round(ep, 5), "relmse_guided": round(eg, 5),
                           "ratio": round(ep / eg, 3),
                           "ratio_time_scaled": round(mean([ep_scaled[k] for k in idx]) / eg, 3),
                           "min_view_ratio_time_scaled": round(min(ep_scaled[k] / e_guided[k] for k in idx), 3)}
    ep, eg = mean(e_plain), mean(e_guided)
    # Per-view equal-time ratios; their geometric mean weighs every view alike
    # (the ratio of mean relMSEs is dominated by the noisiest, lowest-spp views).
    geo = float(np.exp(np.mean([np.log(a / b) for a, b in zip(e_plain, e_guided)])))
    out = {
        "workload": (f"{n_door} door-room variants + {n_terrain} terrain variants (grids {list(terrain_grids)}), "
                     f"{width}x{height}, spp {spps}, depth 16, one wavefront loop; guide: one hash grid per view, "
                     f"shared MLP, trained on the first {train_spp} frames of each view, {spf} steps per frame"
                     f"{', per-view gate' if gate else ''}; references: door views 1024 guided FP64 frames, "
                     f"terrain views {terrain_ref_spp} plain FP64 frames"),
        "views": len(views),
        "triangles": sum(tris),
        "samples": samples,
        "build_seconds": round(t_build, 2),
        "plain_seconds_same_spp": round(t_plain1, 3),
        "plain_msamples_per_s": round(samples / t_plain1 / 1e6, 2),
        "guided_seconds": round(t_guided, 3),
        "guided_msamples_per_s": round(samples / t_guided / 1e6, 2),
        "plain_renders_equal_time": m,
        "plain_seconds_equal_time": round(t_plain, 3),
        "relmse_plain_equal_time": round(ep, 5),
        "relmse_guided": round(eg, 5),
        "ratio": round(ep / eg, 3),
        "ratio_time_scaled": round(mean(ep_scaled) / eg, 3),
        "ratio_geomean_per_view": round(geo, 3),
        "by_scene": kinds,
        "guide_gate": gate,
        "views_gated_off": gated_off,
        "guide_steps": info["steps"],
        "guide_skipped_steps": info["skipped"],
        "per_view": per_view,
    }
    log(f"guided batch: {json.dumps({k: v for k, v in out.items() if k != 'per_view'})}")
    return out


def summarize_pmc(fetch_dir, write_dir, frames_per_launch, out_path, kernels, chains=1):
    """Per-dispatch HBM bytes of each kernel from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by the factor that
    maps the counter onto the known bytes of the copy-probe kernel.  The runs
    render frames_per_launch frames per launch (the bench's launch shape) in
    `chains` concurrent chains, so one bounce (a logical launch, the unit of
    the bench's per-kernel figures) is `chains` dispatches of each kernel."""
    import csv
    import glob

    def per_kernel(d, counter):
        vals = {}
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") != counter:
                        continue
                    vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return vals

    out = {"frames_per_launch": int(frames_per_launch), "chains": int(chains), "kernels": {}}
    for counter, d, key in (("FETCH_SIZE", fetch_dir, "fetch"), ("WRITE_SIZE", write_dir, "write")):
        vals = per_kernel(d, counter)
        copy = [v for k, vs in vals.items() if "k_copy" in k for v in vs]
        if not copy:
            raise SystemExit(f"missing {counter} samples for k_copy in {d}")
        factor = (COPY_PROBE_BYTES / 1024.0) / (sum(copy) / len(copy))
        out[f"{key}_copy_factor"] = factor
        for name, prefix in kernels.items():
            kern = [v for k, vs in vals.items() if prefix in k for v in vs]
            if not kern:
                raise SystemExit(f"missing {counter} samples for {prefix} in {d}")
            e = out["kernels"].setdefault(name, {"kernel": prefix})
            e[f"{key}_kib_per_dispatch_raw"] = sum(kern) / len(kern)
            e[f"{key}_bytes_per_dispatch"] = e[f"{key}_kib_per_dispatch_raw"] * 1024.0 * factor
            e[f"{key}_dispatches"] = len(kern)
    for e in out["kernels"].values():
        e["hbm_bytes_per_dispatch"] = int(e["fetch_bytes_per_dispatch"] + e["write_bytes_per_dispatch"])
        e["hbm_bytes_per_launch"] = e["hbm_bytes_per_dispatch"] * int(chains)
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_terrain_1080p.json")
TRAFFIC_KERNELS = {"trace": "k_trace<false, false>", "shade": "k_shade<false, ", "shadow": "k_shadow<false, false>"}


def measured_traffic(frames_per_launch):
    """HBM bytes per logical launch (one bounce, all chains) of each kernel
    from the committed counter summary, for launches of frames_per_launch
    frames: measured at that shape, or scaled linearly in frames from the
    measured shape (marked so)."""
    if not os.path.exists(TRAFFIC_FILE):
        return {}, None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    scale = frames_per_launch / t["frames_per_launch"]
    note = None if scale == 1 else f"scaled from {t['frames_per_launch']}-frame launches x {scale:.3f}"
    return {k: v.get("hbm_bytes_per_launch", v["hbm_bytes_per_dispatch"]) * scale
            for k, v in t["kernels"].items()}, note


def run_guided_batch(pt, full, train_spp=32):
    if full:
        return batched_guided_experiment(pt, 16, 16, 1920, 1080, 1024, 256, train_spp, [512, 1024, 1448], 0)
    return batched_guided_experiment(pt, 2, 2, 1920, 1080, 1024, 256, train_spp, [512], 0)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-pmc":
        # bench.py --summarize-pmc FETCH_DIR WRITE_DIR FRAMES_PER_LAUNCH OUT_JSON [CHAINS]
        summarize_pmc(sys.argv[2], sys.argv[3], int(sys.argv[4]), sys.argv[5], TRAFFIC_KERNELS,
                      int(sys.argv[6]) if len(sys.argv) > 6 else 1)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64, help="frames (samples per pixel) timed")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--spp-per-launch", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single-chain", action="store_true",
                    help="skip the one-chain pass that times the dominant kernel alone (counter runs use this)")
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--spheres", type=int, default=256)
    ap.add_argument("--no-guided", action="store_true", help="skip the guided-vs-plain relMSE experiment")
    ap.add_argument("--guided-ref-spp", type=int, default=4096)
    ap.add_argument("--guided-plain-spp", type=int, default=1024)
    ap.add_argument("--guided-train-spp", type=int, default=256)
    ap.add_argument("--guided-only", action="store_true",
                    help="run only the guided experiment and print its JSON (profiling runs)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched multi-view experiment")
    ap.add_argument("--batched-views", type=int, default=64)
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched experiment and print its JSON (profiling runs)")
    ap.add_argument("--no-guided-batch", action="store_true", help="skip the guided + batched experiment")
    ap.add_argument("--guided-batch-only", action="store_true",
                    help="run only the guided + batched experiment and print its JSON")
    ap.add_argument("--guided-batch-train-spp", type=int, default=32,
                    help="guided + batched: frames of each view that train the guide")
    ap.add_argument("--guided-batch-full", action="store_true",
                    help="guided + batched at full scale: 16 + 16 views at 1920x1080, terrain grids up to "
                         "1448^2 (4.2M tris)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (a multi-GPU run would split "
              "the frames of every pixel across GPUs and sum the per-GPU means).", file=sys.stderr)
        sys.exit(2)

    import pupiloptixlab_amd as pt

    if args.batched_only:
        print(json.dumps(batched_experiment(pt, args.batched_views, 0)), flush=True)
        return
    if args.guided_only:
        print(json.dumps(guided_experiment(pt, args.guided_ref_spp, args.guided_plain_spp, args.guided_train_spp)),
              flush=True)
        return
    if args.guided_batch_only:
        print(json.dumps(run_guided_batch(pt, args.guided_batch_full, args.guided_batch_train_spp)), flush=True)
        return

    t0 = time.perf_counter()
    scene = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=0, grid=args.grid, spheres=args.spheres)
    d = scene.desc
    log(f"scene: {d.num_triangles} tris, {d.width}x{d.height}, depth {d.max_depth}, "
        f"gen {time.perf_counter() - t0:.2f}s")
    tracer = pt.PathTracer(0)
    tracer.set_scene(scene)
    t1 = time.perf_counter()
    tracer.build_bvh()
    st = tracer.stats()
    log(f"bvh: {st['bvh_nodes']} nodes, depth {st['bvh_depth']}, build {time.perf_counter() - t1:.2f}s")
    copy_gbps = tracer.bandwidth_probe(COPY_PROBE_BYTES, 5)
    log(f"device copy bandwidth: {copy_gbps:.0f} GB/s (read + write)")

    npix = d.width * d.height
    # Every render writes the albedo / normal / test / primitive-id buffers of
    # its last frame, as each launch of the reference does (main.cu:97-104).
    if args.warmup:
        tracer.render(spp=args.warmup, aov=True, spp_per_launch=args.spp_per_launch)
    # Counting pass over exactly the frames of the timed run (deterministic rays).
    tracer.reset()
    tracer.set_counting(True)
    before = tracer.stats()
    tracer.render(spp=args.steps, aov=True, spp_per_launch=args.spp_per_launch)
    after = tracer.stats()
    counted = {k: after[k] - before[k] for k in ("rays_extension", "rays_shadow", "node_visits", "tri_tests",
                                                  "shadow_node_visits", "shadow_tri_tests")}
    tracer.set_counting(False)
    log(f"counted: {counted['rays_extension']} ext rays, {counted['rays_shadow']} shadow rays, "
        f"{counted['node_visits']} nodes, {counted['tri_tests']} tris")

    # Timed region: the pass is blocking, so the call brackets all GPU work.
    tracer.reset()
    base = tracer.stats()
    t_start = time.perf_counter()
    tracer.render(spp=args.steps, aov=True, timing=True, spp_per_launch=args.spp_per_launch)
    elapsed = time.perf_counter() - t_start
    s = tracer.stats()
    ms = {k: s[k] - base[k] for k in ("ms_raygen", "ms_trace", "ms_shade", "ms_shadow", "ms_accumulate")}
    launches = {k: s[k] - base[k] for k in ("launches_trace", "launches_shade", "launches_shadow")}
    rays_ext = s["rays_extension"] - base["rays_extension"]
    rays_shadow = s["rays_shadow"] - base["rays_shadow"]
    assert rays_ext == counted["rays_extension"], "timed and counting passes diverged"

    # The same frames once more (untimed for `value`) with each launch as one
    # chain of kernels: the dominant kernel's event time when it runs alone.
    alone, elapsed1 = None, None
    if not args.no_single_chain:
        tracer.reset()
        base1 = tracer.stats()
        t1 = time.perf_counter()
        tracer.render(spp=args.steps, aov=True, timing=True, spp_per_launch=args.spp_per_launch, one_chain=True)
        elapsed1 = time.perf_counter() - t1
        s1 = tracer.stats()
        alone = {k: (s1["ms_" + k] - base1["ms_" + k], s1["launches_" + k] - base1["launches_" + k])
                 for k in ("trace", "shade", "shadow")}

    samples = args.steps * npix
    value = samples / elapsed / 1e6
    kernels = {
        "trace": (ms["ms_trace"], launches["launches_trace"],
                  rays_ext * (RAY_BYTES + HIT_BYTES) + counted["node_visits"] * NODE_BYTES
                  + counted["tri_tests"] * TRI_BYTES),
        "shadow": (ms["ms_shadow"], launches["launches_shadow"],
                   rays_shadow * (RAY_BYTES + SHADOW_RESULT_BYTES) + counted["shadow_node_visits"] * NODE_BYTES
                   + counted["shadow_tri_tests"] * TRI_BYTES),
        "shade": (ms["ms_shade"], launches["launches_shade"], rays_ext * SHADE_BYTES_PER_VERTEX),
    }
    # A launch here is one bounce of the pass (all concurrent chains), and a
    # kernel's time is its active time: the union of its dispatches' intervals
    # (they share the GPU with the other chains' kernels meanwhile).  The
    # one-chain pass gives each kernel's time alone; the dominant kernel is
    # the one that takes longest alone.
    dominant = max(kernels, key=lambda k: (alone or {k: kernels[k]})[k][0])
    dms, dlaunch, dbytes = kernels[dominant]
    achieved = dbytes / (dms / 1000.0) / 1e9 if dms > 0 else 0.0
    single_chain = None
    if alone:
        ams, alaunch = alone[dominant]
        achieved_alone = dbytes / (ams / 1000.0) / 1e9 if ams > 0 else 0.0
        single_chain = {"avg_launch_ms": round(ams / max(1, alaunch), 4), "achieved": round(achieved_alone, 1),
                        "frac": round(achieved_alone / HBM_PEAK_GBS, 4),
                        "msamples_per_s": round(samples / elapsed1 / 1e6, 3),
                        "pass": "the same frames again with every launch as one chain (not part of value)"}
    # Frames per launch of the timed render (pt_render's default width: up to
    # 64 frames and 128M paths per launch).
    frames_per_launch = args.spp_per_launch or max(1, min(64, (128 << 20) // npix))
    frames_per_launch = min(frames_per_launch, args.steps)
    traffic_by_kernel, traffic_note = measured_traffic(frames_per_launch)
    per_kernel = {}
    for k, (kms, kl, kb) in kernels.items():
        e = {"ms_total": round(kms, 3), "launches": kl, "ms_per_launch": round(kms / max(1, kl), 4),
             "algorithmic_GBps": round(kb / (kms / 1000.0) / 1e9, 1) if kms > 0 else None,
             "algorithmic_bytes_per_launch": int(kb / max(1, kl))}
        if k in traffic_by_kernel and kms > 0:
            e["hbm_bytes_per_launch"] = int(traffic_by_kernel[k])
            e["hbm_GBps"] = round(traffic_by_kernel[k] / (kms / max(1, kl) / 1000.0) / 1e9, 1)
        per_kernel[k] = e
    traffic = per_kernel[dominant].get("hbm_bytes_per_launch")
    hbm_gbps = per_kernel[dominant].get("hbm_GBps")
    # What the counters say bounds the kernel: HBM bandwidth if its measured
    # traffic runs near the device's copy bandwidth, else latency and issue.
    if hbm_gbps is None:
        limiter = "unmeasured (no counter summary)"
    elif hbm_gbps >= 0.6 * copy_gbps:
        limiter = f"hbm bandwidth (counter traffic {hbm_gbps:.0f} GB/s)"
    else:
        limiter = (f"latency and instruction issue: counter traffic {hbm_gbps:.0f} GB/s is "
                   f"{hbm_gbps / HBM_PEAK_GBS:.1%} of HBM peak; node fetches are served by L2 / Infinity Cache")
    log(f"timed: {elapsed:.3f}s, {value:.2f} Msamples/s, kernels {json.dumps(per_kernel)}")

    cpu = None
    if not args.no_cpu_baseline:
        # Every core the process sees, at most 16 (a one-GPU job's share of a
        # host), plus a one-thread figure.
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(d, args.cpu_seconds, threads)
        log(f"cpu baseline: {cpu}")

    tracer.close()
    guided = None
    if not args.no_guided:
        guided = guided_experiment(pt, args.guided_ref_spp, args.guided_plain_spp, args.guided_train_spp)
    batched = None
    if not args.no_batched:
        batched = batched_experiment(pt, args.batched_views, 0)
    guided_batch = None
    if not args.no_guided_batch:
        guided_batch = run_guided_batch(pt, args.guided_batch_full, args.guided_batch_train_spp)

    result = {
        "metric": "Msamples/s at 1080p on 1 GPU; relMSE at equal time, guided vs plain PT",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("terrain: 1024x1024-vertex fBm heightfield + 256 icospheres + 4 area lights, 7 BSDF types, "
                         "constant env 0.1; 1920x1080, max depth 16, seed 0; plain PT with MIS"),
            "triangles": d.num_triangles,
            "width": d.width,
            "height": d.height,
            "spp": args.steps,
            "max_depth": d.max_depth,
            "parallelism": "single GPU",
        },
        "mrays_per_s": round((rays_ext + rays_shadow) / elapsed / 1e6, 2),
        "kernels": per_kernel,
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_shape": traffic_note or f"measured at this launch shape ({frames_per_launch} frames per launch)",
            "traffic_GBps": hbm_gbps,
            "limiter": limiter,
            "single_chain": single_chain,
            "algorithmic_bytes_per_launch": int(dbytes / max(1, dlaunch)),
            "avg_launch_ms": round(dms / max(1, dlaunch), 4),
            "launch": ("one bounce of the pass (all concurrent chains); avg_launch_ms is the kernel's active time "
                       "(union of its dispatch intervals) per bounce"),
            "frames_per_launch": frames_per_launch,
            "measured_copy_GBps": round(copy_gbps, 1),
        },
        "aov_written": True,
        "cpu_baseline": cpu,
        "relmse_guided_vs_plain": guided["ratio"] if guided else None,
        "guided": guided,
        "batched": batched,
        "guided_batched": guided_batch,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
