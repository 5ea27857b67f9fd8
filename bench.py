"""Benchmark: 1920x1080 many-lights ReSTIR on one MI355X.

Workload (one step = one frame of the per-pixel hot path over the whole image):
  scene   512x512-vertex value-noise height field (522,242 triangles) + 500 instances of a
          320-triangle UV sphere (682,242 instanced triangles), 4096 lights (2048 point, 2048 rect),
          intensities 10^U[0,4], blackbody colours U[2000,10000] K;
  path    RaycastPass G-buffer ray -> jittered camera ray -> 4 hero wavelengths -> RIS with
          M = 32 candidates (16 light + 16 BSDF), skipVisibility (one shadow ray for the winner),
          temporal reservoir reuse (the reference's merge), then spatial reuse of 3 neighbours
          within 30 px (an extension beyond the reference), max depth 1 -> spectral->RGB -> accumulation;
          the reference's own pipeline (temporal reuse only) is timed beside it;
  camera  the 64-frame linear dolly of the workload, 1 spp per frame: every camera move clears the
          reservoirs and the accumulation and restarts the frame count (renderer.cpp:401-406), so
          each frame starts from a cleared view at frame index 0; the static-camera variant
          (progressive frames with temporal reuse) is reported beside it.

Prints one JSON line. `value` = ReSTIR frames/s at 1080p with scene, reservoirs and accumulation
resident in HBM. The CPU baseline is the oracle (a faithful C++ restatement of the reference's CPU
path) timed on a bounded sample of the same frames on this host. Further entries: the box at
256x256 (16 spp, PathTracer and RIS), the depth-6 frame, the 16-view batch, the large scene and the
serving stream.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes: what each kernel must move at least once per unit of work.
NODE_BYTES = 32    # one BVH node record (6 fp32 bounds + 2 int)
TRI_BYTES = 36     # three fp32x3 vertices read by the ray-triangle test
GBUF_BYTES = 64    # per pixel G-buffer record, written by the G-buffer kernel, read once after
RES_BYTES = 160    # per pixel reservoir (80 B) read + written
ACCUM_BYTES = 48   # per pixel accumulation rd+wr 32 + colour write 16
RGB2_BYTES = 32    # deferred path: visible + occluded colour written by shading (one read back: 16)
FLAG_BYTES = 4     # deferred path: per pixel visibility flag
QUERY_BYTES = 32   # deferred path: one shadow query (origin, tmax, direction, pixel)
AUX_BYTES = 32     # spatial reuse: per pixel {ray direction, wavelength u}, {throughput}
KERNEL_NAMES = {"gbuffer": "gbuffer_kernel", "integrate": "shade_direct_kernel", "shadow": "shadow_persistent_kernel",
                "accumulate": "accumulate_deferred_kernel", "firefly": "firefly_accumulate_kernel",
                "spatial": "spatial_kernel"}


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _section(name, fn):
    """Runs one measurement after the headline; a failure is reported in the line as {"error": ...}
    (and logged) rather than losing the headline's numbers."""
    _log(name)
    try:
        out = fn()
    except Exception as e:  # noqa: BLE001 -- any failure of an optional measurement is reported
        _log(f"{name} failed: {type(e).__name__}: {e}")
        return {"error": f"{type(e).__name__}: {e}"}
    _log(f"{name} {out}")
    return out


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_many_lights_pmc.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.strip().split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def host_cpu():
    """The host's logical CPU count and model name (from /proc/cpuinfo)."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_threads():
    """One worker per hardware thread of this GPU's share of the host (OMP_NUM_THREADS where the
    job scheduler sets it, else every logical CPU)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    return oracle_py


def cpu_baseline(scene, pipe, rows=540, timed=2):
    """The oracle (a faithful C++20 restatement of the reference's CPU path: 16x16 tiles, per-pixel
    mt19937_64, a UniformLightSampler with its set of every light built per path) on dolly frames:
    each frame from a cleared state at frame index 0, as the reference renders a moving camera. One
    warm-up frame, then the median of `timed` frames over the central `rows` pixel rows, scaled to
    the whole frame. The hoisted-light-set variant (one light set per frame, identical output) is
    timed the same way beside it."""
    oracle_py = _oracle()
    threads = cpu_threads()
    rows = min(rows, scene.height)
    y0 = (scene.height - rows) // 2
    rect = (0, y0, scene.width, y0 + rows)
    out = {}
    for hoist in (False, True):
        o = oracle_py.Oracle(scene, hoist_light_set=hoist, threads=threads)
        secs = []
        for k in range(1 + timed):
            o.reset()
            o.render(scene.frame(0, camera_index=k % len(scene.cameras)), pipe, rect=rect, debug=False)
            if k:
                secs.append(o.last_seconds * scene.height / rows)
        o.close()
        out[hoist] = secs
    med = float(np.median(out[False]))
    return {"value": round(1.0 / med, 5), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"median of {timed} dolly frames (after one warm-up), each {scene.width}x{rows} pixels (the "
                      f"central rows) from a cleared state, scaled by {scene.height / rows:g} to the full frame; "
                      "faithful per-path light-sampler construction",
            "frame_seconds": [round(v, 3) for v in out[False]],
            "hoisted_light_set_frames_per_s": round(1.0 / float(np.median(out[True])), 5),
            "host": host_cpu(), "compiler": "g++ -O2 -std=c++20 -ffp-contract=off (oracle/Makefile)"}


def box_measurement(gpu, abi, ctx, cpu=True, cpu_frames=5):
    """Config 0: the box of 12 triangles with two icospheres (4840 triangles) and one rect light at
    256x256, 16 spp (frames 0 .. 15), depth 4: PathTracer (MIS) and RIS with M = 16 (8 light + 8
    BSDF candidates, visibility tested, without and with the reservoir). GPU: frames/s (= spp/s) over
    the 16 frames after an untimed pass, frame by frame, and the 16-spp image through
    hdr_render_frames; CPU: the oracle on whole frames, median of `cpu_frames`
    frames after one warm-up."""
    from hdrestir_amd import scenes
    sc = scenes.box_spheres(256, 256, frames=16)
    pipes = {"pt_d4": abi.pipeline("mis", max_depth=4),
             "ris16_d4": abi.pipeline("ris", candidate_count=16, use_reservoir=False, max_depth=4),
             "restir16_d4": abi.pipeline("ris", candidate_count=16, use_reservoir=True, max_depth=4)}
    gs = gpu.GpuScene(ctx, sc)
    r = gpu.Renderer(ctx, gs, sc.width, sc.height)
    out = {"workload": "box_spheres_256_16spp_depth4", "triangles": sc.triangle_counts()[0], "spp": 16}
    for name, pipe in pipes.items():
        for _ in range(2):  # an untimed pass, then the timed one
            r.reset()
            r.synchronize()
            t0 = time.perf_counter()
            for f in range(16):
                r.render(sc.frame(f), pipe, read_color=False)
            r.synchronize()
            ms = (time.perf_counter() - t0) * 1000.0 / 16
        out[name] = {"gpu_ms_per_frame": round(ms, 4), "gpu_frames_per_s": round(1000.0 / ms, 2),
                     "gpu_image_ms_16spp": round(16 * ms, 3)}
        # the same 16 frames of the static view through hdr_render_frames: frames that keep no state
        # from frame to frame share launches (ReSTIR's reservoirs do, so it renders frame by frame)
        for _ in range(2):
            r.reset()
            r.synchronize()
            t0 = time.perf_counter()
            r.render_frames(sc.frame(0), pipe, 16)
            r.synchronize()
            ims = (time.perf_counter() - t0) * 1000.0
        out[name]["gpu_image_ms_16spp_render_frames"] = round(ims, 3)
    r.close()
    gs.close()
    if cpu:
        oracle_py = _oracle()
        threads = cpu_threads()
        for name, pipe in pipes.items():
            o = oracle_py.Oracle(sc, hoist_light_set=False, threads=threads)
            secs = []
            for f in range(1 + cpu_frames):
                o.render(sc.frame(f), pipe, debug=False)
                if f:
                    secs.append(o.last_seconds)
            o.close()
            med = float(np.median(secs))
            out[name]["cpu_frames_per_s"] = round(1.0 / med, 3)
            out[name]["cpu_cores"] = threads
        out["cpu_sample"] = f"whole 256x256 frames 1 .. {cpu_frames} after frame 0, median; the oracle, faithful"
    return out


def serving_requests(rng, n, many_view, box_view, only_restir_1080p=False, small_only=False):
    """The serving stream's request mix: resolution uniform over {256^2, 512^2, 1280x720,
    1920x1080}, spp = round(2^U[0,6]), PathTracer (MIS, depth 4) or ReSTIR (M = 32, 1 temporal +
    3 spatial) with equal probability, scene uniform over the resident ones: the 256² box,
    the many-lights scene seen from its camera and from a random point of a 40 m sphere (the
    16-view workload's cameras)."""
    from hdrestir_amd import scenes
    sizes = [(256, 256), (512, 512)] if small_only else [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    reqs = []
    for _ in range(n):
        if only_restir_1080p:
            w, h, integ, scene = 1920, 1080, "restir", 1
        else:
            w, h = sizes[int(rng.integers(len(sizes)))]
            integ = "pt" if rng.uniform() < 0.5 else "restir"
            scene = int(rng.integers(3))
        spp = int(round(2.0 ** rng.uniform(0.0, 6.0)))
        if scene == 0:
            view = box_view
        elif scene == 1:
            view = many_view
        else:
            view = scenes.look_at_origin_from_sphere(rng, 40.0)
        reqs.append(dict(w=w, h=h, spp=spp, integ=integ, scene=scene, view=view))
    return reqs


SERVER_MAX_ACTIVE = 64  # requests on the device at once when batched


def run_server(gpu, abi, ctx, gscenes, reqs, batched, arrivals=None, images=False):
    """Feeds `reqs` to a server (all at t = 0, or at the given arrival times in seconds) and steps
    it until every request has finished. Returns (makespan s, latencies ms, images by index)."""
    from hdrestir_amd import scenes
    pipes = {"pt": abi.pipeline("mis", max_depth=4),
             "restir": abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=1,
                                    spatial_count=3, spatial_radius=30.0)}
    srv = gpu.Server(ctx, batched=batched, max_active=SERVER_MAX_ACTIVE, hbm_budget_bytes=48 << 30)
    ids, lat, imgs = {}, [], {}
    t0 = time.perf_counter()
    nxt = 0
    order = list(range(len(reqs)))
    while len(lat) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(order) and (arrivals is None or arrivals[nxt] <= now):
            q = reqs[order[nxt]]
            proj = scenes.perspective(50.0, q["w"] / q["h"], 0.05, 1000.0)
            fr = abi.frame(q["view"], proj, 0)
            ids[srv.submit(gscenes[q["scene"]], q["w"], q["h"], q["spp"], fr, pipes[q["integ"]])] = order[nxt]
            nxt += 1
        st = srv.stats()
        if st["queued"] == 0 and st["active"] == 0 and arrivals is not None and nxt < len(order):
            time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.01)))
            continue
        srv.step()
        done_t = time.perf_counter() - t0
        for res, img in srv.results(images=images):
            k = ids[res["id"]]
            arrival = 0.0 if arrivals is None else arrivals[k]
            lat.append((done_t - arrival) * 1000.0)
            if images:
                imgs[k] = img
    makespan = time.perf_counter() - t0
    srv.close()
    return makespan, np.array(lat), imgs


def serving_measurement(gpu, abi, ctx, sc_many, g_many, n_requests, seed=1):
    """Config-2 serving: batched (continuous batching) vs the plain per-request path."""
    from hdrestir_amd import scenes
    box = scenes.box_spheres(256, 256, frames=1)
    g_box = gpu.GpuScene(ctx, box)
    gscenes = {0: g_box, 1: g_many, 2: g_many}
    many_view, box_view = sc_many.cameras[0][0], box.cameras[0][0]
    out = {"requests": n_requests, "scenes": "256x256 box, 1080p many-lights from its camera, many-lights from a "
           "40 m sphere (the 16-view workload's cameras); the last two share one resident upload",
           "mix": "resolution U{256^2, 512^2, 1280x720, 1920x1080}, spp round(2^U[0,6]), PathTracer depth 4 or "
                  "ReSTIR M=32 1 temporal + 3 spatial (50/50)"}
    # same images alone and in the mix: small requests, images compared exactly
    rng = np.random.default_rng(seed + 100)
    check = [q for q in serving_requests(rng, 40, many_view, box_view) if q["w"] <= 512][:10]
    _, _, ia = run_server(gpu, abi, ctx, gscenes, check, batched=True, images=True)
    _, _, ib = run_server(gpu, abi, ctx, gscenes, check, batched=False, images=True)
    out["identical_images_batched_vs_plain"] = all(np.array_equal(ia[k], ib[k]) for k in range(len(check)))
    # identical images have identical MSE against any reference (the 4096-spp references of
    # tests/test_gpu_convergence.py included): the ratio the serving target bounds by 1.02
    out["mse_ratio_batched_vs_plain"] = 1.0 if out["identical_images_batched_vs_plain"] else None
    # saturated throughput on the same request list
    for name, only, small in (("mixed", False, False), ("restir_1080p", True, False), ("small", False, True)):
        rng = np.random.default_rng(seed)
        count = max(8, n_requests // 4) if only else n_requests
        reqs = serving_requests(rng, count, many_view, box_view, only, small)
        mb, lb, _ = run_server(gpu, abi, ctx, gscenes, reqs, batched=True)
        mp, lp, _ = run_server(gpu, abi, ctx, gscenes, reqs, batched=False)
        out[name] = {"requests": len(reqs), "batched_requests_per_s": round(len(reqs) / mb, 3),
                     "plain_requests_per_s": round(len(reqs) / mp, 3), "speedup": round(mp / mb, 3),
                     "batched_p50_ms": round(float(np.percentile(lb, 50)), 2),
                     "batched_p99_ms": round(float(np.percentile(lb, 99)), 2),
                     "plain_p50_ms": round(float(np.percentile(lp, 50)), 2),
                     "plain_p99_ms": round(float(np.percentile(lp, 99)), 2)}
        _log(f"serving {name}: {out[name]}")
    # Poisson arrivals: the workload's rates and one below the measured capacity (the all-at-once
    # batched rate of the mix), where p50 / p99 are steady-state latencies rather than queue growth
    capacity = out["mixed"]["batched_requests_per_s"]
    out["capacity_requests_per_s"] = capacity
    out["poisson"] = {}
    below = max(1, int(capacity * 0.4))
    for lam in (below, 50, 200, 800):
        rng = np.random.default_rng(seed + lam)
        reqs = serving_requests(rng, n_requests, many_view, box_view)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(reqs)))
        mb, lb, _ = run_server(gpu, abi, ctx, gscenes, reqs, batched=True, arrivals=arrivals)
        out["poisson"][str(lam)] = {"requests": len(reqs), "requests_per_s": round(len(reqs) / mb, 3),
                                    "p50_ms": round(float(np.percentile(lb, 50)), 2),
                                    "p99_ms": round(float(np.percentile(lb, 99)), 2),
                                    "below_capacity": bool(lam < capacity)}
        _log(f"serving poisson {lam}/s: {out['poisson'][str(lam)]}")
    g_box.close()
    return out


def deep_path_measurement(gpu, abi, ctx, sc, gs, steps, depth=6, m=32, skip_visibility=True):
    """The many-lights frame at path depth `depth` with ReSTIR at every vertex: M = 32 with one shadow
    ray per vertex (skipVisibility), or M candidates whose NEE half each traces its shadow ray
    (without skipVisibility). The wavefront path. Frames/s and rays/s of one 1080p view, static
    camera."""
    pipe = abi.pipeline("ris", candidate_count=m, use_reservoir=True, skip_visibility=skip_visibility,
                        max_depth=depth)
    r = gpu.Renderer(ctx, gs, sc.width, sc.height)
    r.set_counting(True)
    r.render(sc.frame(0, camera_index=0), pipe, read_color=False)
    r.synchronize()
    c = r.counters()
    r.set_counting(False)
    r.reset()
    for i in range(2):
        r.render(sc.frame(i, camera_index=0), pipe, read_color=False)
    r.synchronize()
    r.stage_times()
    t0 = time.perf_counter()
    for i in range(steps):
        r.render(sc.frame(2 + i, camera_index=0), pipe, read_color=False)
    r.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    rays = c["primary_rays"] + c["bounce_rays"] + c["shadow_rays"]
    r.close()
    name = f"many_lights_1080p_restir_depth{depth}" if skip_visibility else \
        f"many_lights_1080p_restir{m}_depth{depth}_visibility"
    return {"workload": name, "max_depth": depth, "candidate_count": m, "skip_visibility": skip_visibility,
            "frames": steps, "ms_per_frame": round(ms, 3), "frames_per_s": round(1000.0 / ms, 3),
            "rays_per_frame": {"primary": c["primary_rays"], "bounce": c["bounce_rays"], "shadow": c["shadow_rays"]},
            "mrays_per_s": round(rays / ms / 1e3, 2)}


def large_scene_measurement(gpu, abi, ctx, frames=3):
    """The large-scene workload at 1080p: 2,093,058-triangle displaced grid + 4000 instances of a
    256-triangle sphere (3.1 M instanced triangles), 65,536 point / rect lights over 5 decades, a
    2048x1024 environment map with log-normal texels (dome CDFs built on the device), ReSTIR M = 32
    with one shadow ray at every vertex of depth-8 paths, the reference's temporal reuse plus 5
    spatial neighbours within 30 px at the primary vertex. With an environment the generated BSDF
    candidates are traced (the wavefront's traced shading launches). The same frames without
    spatial reuse are timed beside it."""
    from hdrestir_amd import scenes
    t0 = time.perf_counter()
    sc = scenes.large_scene(frames=1)
    built = time.perf_counter() - t0
    t0 = time.perf_counter()
    gs = gpu.GpuScene(ctx, sc)
    uploaded = time.perf_counter() - t0
    view, proj = sc.cameras[0]
    pipe = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=8,
                        spatial_count=5, spatial_radius=30.0)
    pipe_t = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=8)
    r = gpu.Renderer(ctx, gs, sc.width, sc.height)

    def timed(p):
        r.reset()
        r.render(abi.frame(view, proj, 0), p, read_color=False)
        r.synchronize()
        r.stage_times()
        t0 = time.perf_counter()
        for f in range(1, 1 + frames):
            r.render(abi.frame(view, proj, f), p, read_color=False)
        r.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / frames, r.stage_times()
    ms, st = timed(pipe)
    ms_t, _ = timed(pipe_t)
    out = {"workload": "large_scene_1080p_restir_depth8_spatial5", "triangles_unique": sc.triangle_counts()[0],
           "triangles_instanced": sc.triangle_counts()[1], "lights": len(sc.lights), "dome": [2048, 1024],
           "max_depth": 8, "candidate_count": 32, "spatial_reuse": {"neighbours": 5, "radius_px": 30},
           "scene_generate_s": round(built, 2), "scene_upload_s": round(uploaded, 2),
           "scene_device_bytes": gs.device_bytes(), "frames": frames, "ms_per_frame": round(ms, 2),
           "frames_per_s": round(1000.0 / ms, 3),
           "spatial_stage_ms_per_frame": round(st["spatial"][0] / max(1, st["spatial"][1]), 3),
           "temporal_only_ms_per_frame": round(ms_t, 2)}
    r.close()
    gs.close()
    return out


def multiview_measurement(gpu, abi, ctx, gs, steps, seed=3):
    """16 views of the many-lights scene in one batch: sizes uniform in [320, 1920] x [240, 1080],
    cameras on a 40 m sphere around the origin, ReSTIR M = 32 at every vertex of depth-6 paths.
    The launch shape is autotuned, steps replay captured graphs; the same frames rendered view by
    view (hdr_render_frame) are timed beside it."""
    from hdrestir_amd import scenes
    rng = np.random.default_rng(seed)
    pipe = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=6)
    views, cams = [], []
    for _ in range(16):
        w, h = int(rng.integers(320, 1921)), int(rng.integers(240, 1081))
        cams.append((scenes.look_at_origin_from_sphere(rng, 40.0), scenes.perspective(50.0, w / h, 0.05, 1000.0)))
        views.append(gpu.Renderer(ctx, gs, w, h))
    pixels = sum(v.W * v.H for v in views)

    def frames(f):
        return [abi.frame(v, p, f) for v, p in cams]
    batch = gpu.Batch(ctx)
    chosen, tuned_ms = batch.autotune(views, frames(0), [pipe] * 16, reps=2)
    for v in views:
        v.reset()
    batch.render(views, frames(0), [pipe] * 16)
    batch.synchronize()
    t0 = time.perf_counter()
    for f in range(1, 1 + steps):
        batch.render(views, frames(f), [pipe] * 16)
    batch.synchronize()
    ms_batched = (time.perf_counter() - t0) * 1000.0 / steps
    for v in views:
        v.reset()
    for v, fr in zip(views, frames(0)):
        v.render(fr, pipe, read_color=False)
    for v in views:
        v.synchronize()
    t0 = time.perf_counter()
    for f in range(1, 1 + steps):
        for v, fr in zip(views, frames(f)):
            v.render(fr, pipe, read_color=False)
    for v in views:
        v.synchronize()
    ms_plain = (time.perf_counter() - t0) * 1000.0 / steps
    batch.close()
    for v in views:
        v.close()
    return {"workload": "many_lights_16_views_restir_depth6", "views": 16, "pixels": pixels,
            "sizes": [[v.W, v.H] for v in views], "steps": steps, "autotuned_config": chosen,
            "autotune_step_ms": round(tuned_ms, 3), "batched_ms_per_step": round(ms_batched, 3),
            "batched_mpixels_per_s": round(pixels / ms_batched / 1e3, 2),
            "plain_ms_per_step": round(ms_plain, 3), "speedup": round(ms_plain / ms_batched, 3)}


def post_process_measurement(r, frame, pipe, reps=8):
    """hdr_view_post_process (DenoiserPass's chromaticity blur, PostProcessPass's lens flare and
    chromatic aberration) on the view's 1080p colour: each pass alone and all three, wall time per
    call from launch to the end of its kernels (a fresh frame before each call, outside the timing)."""
    out = {}
    cases = (("chromaticity_blur", dict(chromaticity_blur=True)), ("lens_flare", dict(lens_flare=True)),
             ("chromatic_aberration", dict(chromatic_aberration=4.0)),
             ("all", dict(chromaticity_blur=True, lens_flare=True, chromatic_aberration=4.0)))
    for name, kw in cases:
        ts = []
        for k in range(reps + 1):
            r.reset()
            r.render(frame, pipe, read_color=False)
            r.synchronize()
            t0 = time.perf_counter()
            r.post_process(**kw)
            r.synchronize()
            if k:
                ts.append((time.perf_counter() - t0) * 1e3)
        out[name + "_ms"] = round(float(np.median(ts)), 4)
    return out


def dolly_reprojection_measurement(gpu, abi, ctx, sc, gs, r, pipe, warmup, steps, ref_frames=256):
    """Config 1's dolly with reprojected temporal reuse (an extension: hdr_pipeline_desc::
    temporal_reprojection). Each move restarts the accumulation and the frame index (as on the
    headline) but keeps the reservoirs, and the previous frame's reservoir where each pixel's hit was
    seen joins the spatial pass. Frames/s over the headline's frames, and the mean squared error of
    the 1-spp frames at dolly frames 16, 32, 48 and 63 (rendered in order from frame 0) against
    converged images of those cameras -- RIS M = 32 without a reservoir (unbiased), `ref_frames`
    frames -- next to the headline's frames, whose moves clear the reservoirs."""
    import ctypes as C
    rp = abi.PipelineDesc()
    C.memmove(C.byref(rp), C.byref(pipe), C.sizeof(rp))
    rp.temporal_reprojection = 1
    n = len(sc.cameras)

    def dolly(i):
        return sc.frame(0, camera_index=i % n)

    r.reset()
    for i in range(warmup):
        r.reset_accumulation()
        r.render(dolly(i), rp, read_color=False)
    r.synchronize()
    r.stage_times()
    t0 = time.perf_counter()
    for i in range(steps):
        r.reset_accumulation()
        r.render(dolly(warmup + i), rp, read_color=False)
    r.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    stages = {k: round(v[0] / v[1], 4) for k, v in r.stage_times().items() if v[1] > 0}

    checks = (16, 32, 48, n - 1)
    images = {"reset_each_move": {}, "reprojection": {}}
    for name, p in (("reset_each_move", pipe), ("reprojection", rp)):
        r.reset()
        for i in range(n):
            if name == "reset_each_move":
                r.reset()
            else:
                r.reset_accumulation()
            r.render(dolly(i), p, read_color=False)
            if i in checks:
                images[name][i] = r.read_color()[..., :3].astype(np.float64).copy()
    ref_pipe = abi.pipeline("ris", candidate_count=32, use_reservoir=False, skip_visibility=True, max_depth=1)
    v = gpu.Renderer(ctx, gs, sc.width, sc.height)
    mse = {"reset_each_move": [], "reprojection": []}
    for i in checks:
        v.reset()
        for f in range(ref_frames):
            v.render(sc.frame(f, camera_index=i), ref_pipe, read_color=False)
        ref = v.read_color()[..., :3].astype(np.float64)
        for name in mse:
            mse[name].append(float(np.mean((images[name][i] - ref) ** 2)))
    v.close()
    ratio = [b / a for a, b in zip(mse["reset_each_move"], mse["reprojection"])]
    return {"frames_per_s": round(1000.0 / ms, 3), "ms_per_step": round(ms, 4), "stage_ms": stages,
            "mse_frames": list(checks), "mse_reset_each_move": [round(x, 6) for x in mse["reset_each_move"]],
            "mse_reprojection": [round(x, 6) for x in mse["reprojection"]],
            "mse_ratio_reprojection_to_reset": [round(x, 3) for x in ratio],
            "reference": f"RIS M=32 without a reservoir, {ref_frames} frames per camera"}


def run_frames(r, desc, pipe, warmup, steps, reset_each):
    """`warmup` untimed frames desc(0 ..), then `steps` timed frames; with reset_each every frame starts
    from a cleared view (a camera move). Returns (ms per frame by the host clock around the timed
    frames, synchronised on both sides; {stage: (summed device ms, launches)} from HIP events on the
    streams the kernels run on)."""
    if not reset_each:
        r.reset()
    for i in range(warmup):
        if reset_each:
            r.reset()
        r.render(desc(i), pipe, read_color=False)
    r.synchronize()
    r.stage_times()  # clears the per-kernel sums
    t0 = time.perf_counter()
    for i in range(steps):
        if reset_each:
            r.reset()
        r.render(desc(warmup + i), pipe, read_color=False)
    r.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    return ms, {k: v for k, v in r.stage_times().items() if v[1] > 0}


def count_frames(r, desc, pipe, warmup, steps, reset_each):
    """The frame sequence of run_frames with the work counters on (untimed, deterministic, so the
    timed pass repeats exactly this work); counters averaged over the `steps` frames."""
    r.set_counting(True)
    r.reset()
    for i in range(warmup):
        if reset_each:
            r.reset()
        r.render(desc(i), pipe, read_color=False)
    counts = None
    for i in range(steps):
        if reset_each:
            r.reset()
        r.render(desc(warmup + i), pipe, read_color=False)
        r.synchronize()
        c = r.counters()
        counts = c if counts is None else {k: counts[k] + c[k] for k in c}
    r.set_counting(False)
    return {k: v / steps for k, v in counts.items()}


# stages on the view's main stream: the frame's critical path (the G-buffer kernel of the next frame
# runs on the second stream beside them)
CRITICAL_STAGES = ("integrate", "spatial", "shadow", "accumulate", "firefly")


def kernel_table(stages, counts, P):
    """Per-kernel average time, algorithmic bytes and achieved GB/s, and the roofline entry of the
    critical-path kernel with the largest summed time."""
    pn, pt = counts["primary_node_visits"], counts["primary_triangle_tests"]
    sn, st = counts["node_visits"] - pn, counts["triangle_tests"] - pt
    q = counts["shadow_rays"]
    deferred = "shadow" in stages
    if deferred and "spatial" in stages:
        # shading: G-buffer read, reservoir read, post-temporal reservoir and aux record written;
        # spatial: every pixel's post-temporal reservoir, G-buffer and aux record read once (the
        # neighbours' re-reads are cache traffic), final reservoir, colour pair and flag written
        alg = {"gbuffer": NODE_BYTES * pn + TRI_BYTES * pt + GBUF_BYTES * P,
               "integrate": (GBUF_BYTES + RES_BYTES + AUX_BYTES) * P,
               "spatial": (RES_BYTES // 2 + GBUF_BYTES + AUX_BYTES + RES_BYTES // 2 + RGB2_BYTES + FLAG_BYTES) * P
                          + QUERY_BYTES * q,
               "shadow": QUERY_BYTES * q + NODE_BYTES * sn + TRI_BYTES * st,
               "accumulate": (RGB2_BYTES // 2 + FLAG_BYTES + ACCUM_BYTES) * P}
    elif deferred:
        alg = {"gbuffer": NODE_BYTES * pn + TRI_BYTES * pt + GBUF_BYTES * P,
               "integrate": (GBUF_BYTES + RES_BYTES + RGB2_BYTES + FLAG_BYTES) * P + QUERY_BYTES * q,
               "shadow": QUERY_BYTES * q + NODE_BYTES * sn + TRI_BYTES * st,
               "accumulate": (RGB2_BYTES // 2 + FLAG_BYTES + ACCUM_BYTES) * P}
    else:
        alg = {"gbuffer": NODE_BYTES * pn + TRI_BYTES * pt + GBUF_BYTES * P,
               "integrate": NODE_BYTES * sn + TRI_BYTES * st + (GBUF_BYTES + RES_BYTES + ACCUM_BYTES) * P}
    kernels = {}
    for k, (ms, n) in stages.items():
        avg = ms / n
        name = KERNEL_NAMES[k] if (deferred or k != "integrate") else "integrate_kernel"
        e = {"avg_ms": round(avg, 4), "launches": n}
        if k in alg:
            e["algorithmic_bytes_per_launch"] = round(alg[k])
            e["achieved_GBs"] = round(alg[k] / (avg * 1e-3) / 1e9, 2)
            e["hbm_traffic_bytes_per_launch"] = read_traffic(name)
        kernels[name] = e
    crit = [k for k in stages if k in CRITICAL_STAGES and k in alg]
    dom_stage = max(crit, key=lambda k: stages[k][0])
    dom = KERNEL_NAMES[dom_stage] if (deferred or dom_stage != "integrate") else "integrate_kernel"
    dom_avg = stages[dom_stage][0] / stages[dom_stage][1]
    achieved = alg[dom_stage] / (dom_avg * 1e-3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(dom),
            "traffic_source": f"profiles/{dom}_many_lights_pmc.txt (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of "
                              "this workload, scaled by a plain copy)",
            "avg_ms": round(dom_avg, 4), "algorithmic_bytes_per_launch": round(alg[dom_stage]),
            "selection": "the critical-path kernel (view's main stream) with the largest summed time"}
    return kernels, roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=270,
                    help="pixel rows of each frame the CPU baseline renders (centred)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES bytes 4 times (the reference "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    ap.add_argument("--extended", action="store_true",
                    help="also run the request-serving, post-process and reprojected-dolly measurements "
                         "(off by default: two runs that included them ended in a GPU memory fault whose "
                         "cause is not shown)")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-serving measurement")
    ap.add_argument("--no-deep", action="store_true",
                    help="skip the depth-6 single-view, 16-view batch and large-scene measurements")
    ap.add_argument("--deep-only", action="store_true",
                    help="only the depth-6 single-view measurement, --steps frames (for profilers)")
    ap.add_argument("--large-only", action="store_true",
                    help="only the large-scene measurement (for profilers)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the counting pass and the timed frames of the headline pipeline (for profilers)")
    ap.add_argument("--serving-requests", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the path runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by pixels and by requests across GPUs)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration, "launches": 4}), flush=True)
        return

    from hdrestir_amd import abi, gpu, scenes
    if args.large_only:
        ctx = gpu.Context(0)
        print(json.dumps(large_scene_measurement(gpu, abi, ctx)), flush=True)
        ctx.close()
        return
    _log("building scene")
    sc = scenes.many_lights(frames=64, dolly=True)  # the 1080p many-lights scene's 64-frame linear dolly
    if args.deep_only:
        ctx = gpu.Context(0)
        gs = gpu.GpuScene(ctx, sc)
        print(json.dumps(deep_path_measurement(gpu, abi, ctx, sc, gs, args.steps)), flush=True)
        gs.close()
        ctx.close()
        return
    # 1 temporal (the reference's merge) + 3 spatial neighbours within 30 px (an extension)
    pipe = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=1,
                        spatial_count=3, spatial_radius=30.0)
    pipe_temporal = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=1)
    ctx = gpu.Context(0)
    gs = gpu.GpuScene(ctx, sc)
    r = gpu.Renderer(ctx, gs, sc.width, sc.height)
    P = sc.width * sc.height

    # the headline workload: a 64-frame linear dolly at 1 spp per frame. Each camera move clears the
    # reservoirs and the accumulation and restarts the frame count (renderer.cpp:401-406), so every
    # frame is rendered from a cleared view at frame index 0.
    def dolly(i):
        return sc.frame(0, camera_index=i % len(sc.cameras))

    def static(i):  # the progressive variant: one camera, temporal reuse across frames
        return sc.frame(i, camera_index=0)

    counts = count_frames(r, dolly, pipe, args.warmup, args.steps, reset_each=True)
    _log(f"counters per dolly frame {counts}")
    ms_per_step, stages = run_frames(r, dolly, pipe, args.warmup, args.steps, reset_each=True)
    fps = 1000.0 / ms_per_step
    img = r.read_color()
    kernels, roof = kernel_table(stages, counts, P)

    # the same frames with the reference's own pipeline (temporal reuse only)
    temporal_fps = None
    if not args.headline_only:
        ms_t, _ = run_frames(r, dolly, pipe_temporal, args.warmup, args.steps, reset_each=True)
        temporal_fps = 1000.0 / ms_t
    # static camera: progressive accumulation, temporal reuse from the previous frames
    static_line = None
    if not args.headline_only:
        counts_s = count_frames(r, static, pipe, args.warmup, args.steps, reset_each=False)
        ms_s, stages_s = run_frames(r, static, pipe, args.warmup, args.steps, reset_each=False)
        kernels_s, roof_s = kernel_table(stages_s, counts_s, P)
        rays_s = sum(counts_s[k] for k in ("primary_rays", "bounce_rays", "shadow_rays", "candidate_rays"))
        static_line = {"camera": "static (progressive, temporal reuse across frames)",
                       "frames_per_s": round(1000.0 / ms_s, 3), "ms_per_step": round(ms_s, 4),
                       "mrays_per_s": round(rays_s / ms_s / 1e3, 2), "shadow_rays_per_frame": counts_s["shadow_rays"],
                       "kernels": kernels_s, "roofline": roof_s}
        _log(f"static camera {static_line['frames_per_s']} frames/s")

    rays = counts["primary_rays"] + counts["bounce_rays"] + counts["shadow_rays"] + counts["candidate_rays"]
    ref_rays = rays + counts["candidate_rays_skipped"]
    line = {
        "metric": "ReSTIR frames/s at 1080p (many-lights, M=32, 1 temporal + 3 spatial within 30 px, 64-frame "
                  "linear dolly); Mrays/s alongside",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": "many_lights_1080p_restir", "width": sc.width, "height": sc.height,
                   "triangles_unique": sc.triangle_counts()[0], "triangles_instanced": sc.triangle_counts()[1],
                   "lights": len(sc.lights), "candidate_count": 32, "skip_visibility": True, "max_depth": 1,
                   "temporal_reuse": True, "spatial_reuse": 3, "spatial_radius_px": 30, "camera": "linear_dolly",
                   "dolly_frames": len(sc.cameras), "spp_per_frame": 1,
                   "camera_move": "clears reservoirs and accumulation, frame index restarts at 0 (renderer.cpp:401-406)",
                   "parallelism": "single-gpu"},
        "mrays_per_s": round(rays * fps / 1e6, 2),
        "mrays_per_s_reference_equivalent": round(ref_rays * fps / 1e6, 2),
        "rays_per_frame": {"primary": counts["primary_rays"], "bounce": counts["bounce_rays"],
                           "shadow": counts["shadow_rays"], "bsdf_candidate_traced": counts["candidate_rays"],
                           "skipped_cannot_contribute": counts["candidate_rays_skipped"]},
        "temporal_only_frames_per_s": None if temporal_fps is None else round(temporal_fps, 3),
        "kernels": kernels,
        "roofline": roof,
        "static_camera": static_line,
        "image_mean_rgb": [round(float(v), 3) for v in img[..., :3].reshape(-1, 3).mean(0)],
    }
    if not args.headline_only and not args.no_deep:
        line["deep_paths"] = _section(
            "deep paths", lambda: deep_path_measurement(gpu, abi, ctx, sc, gs, max(4, args.steps // 4)))
        line["visibility_paths"] = _section(
            "visibility paths", lambda: deep_path_measurement(gpu, abi, ctx, sc, gs, max(4, args.steps // 8), depth=4,
                                                              m=16, skip_visibility=False))
        line["multiview"] = _section("multi-view batch", lambda: multiview_measurement(gpu, abi, ctx, gs, 4))
        line["large_scene"] = _section("large scene", lambda: large_scene_measurement(gpu, abi, ctx))
    if not args.headline_only:
        line["box_256"] = _section("256x256 box, 16 spp",
                                   lambda: box_measurement(gpu, abi, ctx, cpu=not args.no_cpu_baseline))
    if args.extended and not args.no_serving and not args.headline_only:
        line["serving"] = _section("serving",
                                   lambda: serving_measurement(gpu, abi, ctx, sc, gs, args.serving_requests))
    # the headline view's extensions last (the view stays open until here)
    if args.extended and not args.headline_only:
        line["post_process"] = _section("post process", lambda: post_process_measurement(r, dolly(0), pipe))
        line["dolly_reprojection"] = _section(
            "dolly with reprojected temporal reuse",
            lambda: dolly_reprojection_measurement(gpu, abi, ctx, sc, gs, r, pipe, args.warmup, args.steps))
    try:
        r.close()
        gs.close()
        ctx.close()
    except Exception as e:  # noqa: BLE001 -- a device error after the measurements
        _log(f"closing the GPU context failed: {type(e).__name__}: {e}")
    if not args.no_cpu_baseline:
        _log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(sc, pipe, rows=args.cpu_rows)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
