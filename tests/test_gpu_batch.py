"""Batches of views (hdr_render_batch): one frame for many views in merged launches gives every
view exactly the frame hdr_render_frame gives it -- across scenes, sizes, kernel paths (deferred
shading, spatial reuse, integration megakernel, inlined path tracer, wavefront) and frames of
temporal reuse."""
import numpy as np
import pytest

import parity

from hdrestir_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _cases():
    box = scenes.box_spheres(64, 48, frames=5)
    box_small = scenes.box_spheres(40, 24, frames=5)
    feat = scenes.features(48, 40, frames=5, dome=False, emissive=False, opaque=True)
    feat_dome = scenes.features(32, 32, frames=5)
    feat_dome_opaque = scenes.features(32, 32, frames=5, dome=True, emissive=False, mesh_light=False, opaque=True)
    return [
        (box, abi.pipeline("ris", candidate_count=16, use_reservoir=True, skip_visibility=True, max_depth=1)),
        (feat, abi.pipeline("ris", candidate_count=8, use_reservoir=True, skip_visibility=True, max_depth=1,
                            spatial_count=3, spatial_radius=10.0)),
        (box_small, abi.pipeline("mis", max_depth=4)),
        (feat_dome, abi.pipeline("ris", candidate_count=8, use_reservoir=True, max_depth=3)),
        (box, abi.pipeline("ris", candidate_count=4, use_reservoir=False, skip_visibility=True, max_depth=1)),
        # the wavefront path, two depths in one group
        (box_small, abi.pipeline("ris", candidate_count=16, use_reservoir=True, skip_visibility=True, max_depth=5)),
        (feat, abi.pipeline("ris", candidate_count=8, use_reservoir=False, skip_visibility=True, max_depth=3)),
        # spatial reuse on the wavefront path, with and without traced candidates
        (box, abi.pipeline("ris", candidate_count=8, use_reservoir=True, skip_visibility=True, max_depth=4,
                           spatial_count=3, spatial_radius=12.0)),
        (feat_dome_opaque, abi.pipeline("ris", candidate_count=8, use_reservoir=True, skip_visibility=True,
                                        max_depth=3, spatial_count=2, spatial_radius=8.0)),
    ]


def test_batch_matches_single_view_frames(gpu_ctx):
    from hdrestir_amd import gpu
    cases = _cases()
    gscenes = {}
    for sc, _ in cases:
        if id(sc) not in gscenes:
            gscenes[id(sc)] = gpu.GpuScene(gpu_ctx, sc)
    single = [gpu.Renderer(gpu_ctx, gscenes[id(sc)], sc.width, sc.height) for sc, _ in cases]
    batched = [gpu.Renderer(gpu_ctx, gscenes[id(sc)], sc.width, sc.height) for sc, _ in cases]
    batch = gpu.Batch(gpu_ctx)
    for f in range(5):
        for r, (sc, pipe) in zip(single, cases):
            r.render(sc.frame(f), pipe, read_color=False)
        batch.render(batched, [sc.frame(f) for sc, _ in cases], [pipe for _, pipe in cases])
    batch.synchronize()
    assert batch.last_timing() > 0.0
    for k, (rs, rb) in enumerate(zip(single, batched)):
        a = rs.read_color().copy()
        b = rb.read_color().copy()
        assert np.array_equal(a, b), k
        if cases[k][1].use_reservoir:
            assert rs.reservoirs().tobytes() == rb.reservoirs().tobytes(), k
    # a batch of one equals the single-view path too, and later single frames follow the batch
    for r, (sc, pipe) in zip(batched, cases):
        r.render(sc.frame(5 % 5), pipe, read_color=False)
    for r, (sc, pipe) in zip(single, cases):
        r.render(sc.frame(5 % 5), pipe, read_color=False)
    for rs, rb in zip(single, batched):
        assert np.array_equal(rs.read_color(), rb.read_color())
    batch.close()
    for r in single + batched:
        r.close()
    for g in gscenes.values():
        g.close()


def test_batch_refuses_duplicates_and_counting(gpu_ctx):
    from hdrestir_amd import gpu
    sc = scenes.box_spheres(16, 16, frames=1)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, 16, 16)
    batch = gpu.Batch(gpu_ctx)
    pipe = abi.pipeline("mis", max_depth=2)
    with pytest.raises(gpu.HdrError):
        batch.render([r, r], [sc.frame(0), sc.frame(0)], [pipe, pipe])
    r.set_counting(True)
    with pytest.raises(gpu.HdrError):
        batch.render([r], [sc.frame(0)], [pipe])
    batch.close()
    r.close()
    gs.close()


def test_graphs_launch_shapes_and_autotune_keep_results(gpu_ctx):
    """Captured-graph replay, every launch-shape variant (queue waves, traversal and shading
    occupancy) and the autotuner's choice render exactly the frames of plain launches."""
    from hdrestir_amd import gpu
    cases = _cases()
    gscenes = {}
    for sc, _ in cases:
        if id(sc) not in gscenes:
            gscenes[id(sc)] = gpu.GpuScene(gpu_ctx, sc)

    def views():
        return [gpu.Renderer(gpu_ctx, gscenes[id(sc)], sc.width, sc.height) for sc, _ in cases]

    pipes = [pipe for _, pipe in cases]
    tune_views = views()
    tuner = gpu.Batch(gpu_ctx)
    chosen, ms = tuner.autotune(tune_views, [sc.frame(0) for sc, _ in cases], pipes, reps=1)
    print("autotuned", chosen, ms)
    assert ms > 0.0 and chosen["use_graphs"] == 1
    configs = [abi.launch_config(use_graphs=0),
               abi.launch_config(**chosen),
               abi.launch_config(queue_waves=64, trace_occupancy=8, shade_occupancy=3, use_graphs=1),
               abi.launch_config(queue_waves=65536, trace_occupancy=4, use_graphs=1),
               abi.launch_config(trace_occupancy=1, shade_occupancy=1, use_graphs=1)]
    results = []
    for cfg in configs:
        vs = views()
        b = gpu.Batch(gpu_ctx)
        b.set_launch_config(cfg)
        for f in range(6):  # more steps than argument slots: graphs are replayed
            b.render(vs, [sc.frame(f % 5) for sc, _ in cases], pipes)
        b.synchronize()
        results.append([(v.read_color().copy(), v.reservoirs().tobytes()) for v in vs])
        b.close()
        for v in vs:
            v.close()
    for k in range(1, len(results)):
        for j, ((ca, ra), (cb, rb)) in enumerate(zip(results[0], results[k])):
            assert np.array_equal(ca, cb), (k, j)
            assert ra == rb, (k, j)
    # single views take the launch shape too
    sc, pipe = cases[5]
    imgs = []
    for cfg in (abi.launch_config(), abi.launch_config(queue_waves=128, trace_occupancy=6, shade_occupancy=3)):
        r = gpu.Renderer(gpu_ctx, gscenes[id(sc)], sc.width, sc.height)
        r.set_launch_config(cfg)
        for f in range(3):
            r.render(sc.frame(f), pipe, read_color=False)
        imgs.append(r.read_color().copy())
        r.close()
    assert np.array_equal(imgs[0], imgs[1])
    tuner.close()
    for v in tune_views:
        v.close()
    for g in gscenes.values():
        g.close()


def test_sixteen_views_of_the_many_lights_scene_batched_with_graphs_and_autotune(gpu_ctx, oracle_mod):
    """The multi-view workload: 16 viewports of the many-lights scene (sizes uniform in
    [320, 1920] x [240, 1080], cameras on a 40 m sphere), ReSTIR M=32 at every vertex of depth-6
    paths, steps replayed from captured graphs with the autotuned launch shape: every view's image and
    reservoirs equal those of the same frames through hdr_render_frame, and one view matches the
    oracle on a crop."""
    from hdrestir_amd import gpu
    sc = scenes.many_lights(frames=1, dolly=False)
    gs = gpu.GpuScene(gpu_ctx, sc)
    rng = np.random.default_rng(3)
    pipe = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=6)
    sizes, cams = [], []
    for _ in range(16):
        w, h = int(rng.integers(320, 1921)), int(rng.integers(240, 1081))
        sizes.append((w, h))
        cams.append((scenes.look_at_origin_from_sphere(rng, 40.0), scenes.perspective(50.0, w / h, 0.05, 1000.0)))
    views = [gpu.Renderer(gpu_ctx, gs, w, h) for w, h in sizes]
    singles = [gpu.Renderer(gpu_ctx, gs, w, h) for w, h in sizes]

    def frames(f):
        return [abi.frame(v, p, f) for v, p in cams]
    batch = gpu.Batch(gpu_ctx)
    chosen, ms = batch.autotune(views, frames(0), [pipe] * 16, reps=1)
    print("autotuned", chosen, f"{ms:.2f} ms per step")
    assert chosen["use_graphs"] == 1
    for v in views:
        v.reset()
    for f in range(2):
        batch.render(views, frames(f), [pipe] * 16)
        for s, fr in zip(singles, frames(f)):
            s.render(fr, pipe, read_color=False)
    batch.synchronize()
    for k, (v, s) in enumerate(zip(views, singles)):
        assert np.array_equal(v.read_color(), s.read_color()), k
        assert np.array_equal(v.reservoirs(), s.reservoirs()), k
    # view 0 against the oracle on a crop (the view's own size and camera)
    w, h = sizes[0]
    vsc = scenes.many_lights(w, h, frames=1, dolly=False)
    o = oracle_mod.Oracle(vsc, threads=8)
    x0, y0 = w // 2 - 24, h // 2 - 16
    rect = (x0, y0, x0 + 48, y0 + 32)
    probe = gpu.Renderer(gpu_ctx, gs, w, h)
    for f in range(2):
        fr = frames(f)[0]
        probe.render(fr, pipe, debug=True)
        o.render(fr, pipe, rect=rect)
        sel = np.zeros((h, w), bool)
        sel[rect[1]:rect[3], rect[0]:rect[2]] = True
        sel = sel.reshape(-1)
        same = (probe.debug["draws"] == o.debug["draws"]) & (probe.debug["light_hash"] == o.debug["light_hash"]) & \
               (probe.debug["select_hash"] == o.debug["select_hash"])
        parity.check(f"multi-view view 0 frame {f}", same[sel], probe.debug["sample_rgb"][sel],
                     o.debug["sample_rgb"][sel])
    assert np.array_equal(probe.read_color(), views[0].read_color())
    batch.close()
    for v in views + singles + [probe]:
        v.close()
    gs.close()
