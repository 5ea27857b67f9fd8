"""Serving (hdr_server_*): continuous batching and the plain per-request path give every request
the image it gets rendered alone; admission keeps the views within the device-memory budget."""
import numpy as np
import pytest

from hdrestir_amd import abi, scenes

pytestmark = pytest.mark.gpu

RESTIR = abi.pipeline("ris", candidate_count=16, use_reservoir=True, skip_visibility=True, max_depth=1,
                      spatial_count=3, spatial_radius=8.0)
PT = abi.pipeline("mis", max_depth=4)


def _requests(box, feat):
    # (scene, width, height, spp, pipe): uneven sizes and sample counts, both integrators
    return [(box, 48, 32, 3, RESTIR), (feat, 40, 40, 5, PT), (box, 64, 48, 1, PT), (feat, 32, 24, 4, RESTIR),
            (box, 24, 24, 6, RESTIR), (feat, 56, 40, 2, PT), (box, 40, 24, 5, PT), (feat, 48, 32, 3, RESTIR)]


def _alone(gpu, ctx, gs, sc, w, h, spp, pipe):
    r = gpu.Renderer(ctx, gs, w, h)
    for f in range(spp):
        img = r.render(_frame(sc, w, h, f), pipe)
    out = img.copy()
    r.close()
    return out


def _frame(sc, w, h, f):
    view, _ = sc.cameras[0]
    return abi.frame(view, scenes.perspective(50.0, w / h, 0.05, 100.0), f)


@pytest.mark.parametrize("batched", [True, False])
def test_server_results_match_requests_rendered_alone(gpu_ctx, batched):
    from hdrestir_amd import gpu
    box = scenes.box_spheres(64, 64, frames=1)
    feat = scenes.features(64, 64, frames=1, dome=False, emissive=False, opaque=True)
    gbox, gfeat = gpu.GpuScene(gpu_ctx, box), gpu.GpuScene(gpu_ctx, feat)
    gsc = {id(box): gbox, id(feat): gfeat}
    srv = gpu.Server(gpu_ctx, batched=batched, max_active=5)
    reqs = _requests(box, feat)
    ids = {}
    for k, (sc, w, h, spp, pipe) in enumerate(reqs):
        ids[srv.submit(gsc[id(sc)], w, h, spp, _frame(sc, w, h, 0), pipe)] = k
    got = {}
    steps = 0
    while len(got) < len(reqs):
        srv.step()
        steps += 1
        assert steps < 200
        for res, img in srv.results(images=True):
            got[ids[res["id"]]] = (res, img)
    for k, (sc, w, h, spp, pipe) in enumerate(reqs):
        res, img = got[k]
        assert res["latency_ms"] >= res["queued_ms"] >= 0.0
        ref = _alone(gpu, gpu_ctx, gsc[id(sc)], sc, w, h, spp, pipe)
        assert np.array_equal(img.reshape(-1, 4), ref.reshape(-1, 4)), k
    st = srv.stats()
    assert st["queued"] == 0 and st["active"] == 0
    srv.close()
    gbox.close()
    gfeat.close()


def test_server_admission_respects_memory_budget(gpu_ctx):
    from hdrestir_amd import gpu
    box = scenes.box_spheres(64, 64, frames=1)
    gbox = gpu.GpuScene(gpu_ctx, box)
    budget = 6 << 20  # room for about two of these views at once
    srv = gpu.Server(gpu_ctx, batched=True, max_active=16, hbm_budget_bytes=budget)
    for k in range(8):
        srv.submit(gbox, 64, 48, 2 + k % 3, _frame(box, 64, 48, 0), RESTIR)
    done = 0
    waited = False
    while done < 8:
        srv.step()
        st = srv.stats()
        assert st["hbm_in_use"] <= budget
        waited |= st["queued"] > 0 and st["active"] > 0
        done += len(srv.results())
    assert waited  # requests had to wait for memory, and all of them finished
    srv.close()
    gbox.close()


def test_server_rejects_an_unsupported_request_and_serves_the_rest(gpu_ctx):
    """A request whose pipeline its scene cannot run (spatial reuse on a scene with translucent
    materials and a mesh light) is refused at submission, as hdr_render_frame would refuse it; the
    valid request next to it is served as usual."""
    from hdrestir_amd import gpu
    box = scenes.box_spheres(64, 64, frames=1)
    feat = scenes.features(32, 32, frames=1)
    gbox, gfeat = gpu.GpuScene(gpu_ctx, box), gpu.GpuScene(gpu_ctx, feat)
    srv = gpu.Server(gpu_ctx, batched=True, max_active=4)
    with pytest.raises(gpu.HdrError, match="UNSUPPORTED"):
        srv.submit(gfeat, 32, 32, 2, _frame(feat, 32, 32, 0), RESTIR)
    rid = srv.submit(gbox, 48, 32, 3, _frame(box, 48, 32, 0), RESTIR)
    got = []
    for _ in range(50):
        srv.step()
        got += srv.results(images=True)
        if got:
            break
    assert len(got) == 1 and got[0][0]["id"] == rid and got[0][0]["status"] == 0
    ref = _alone(gpu, gpu_ctx, gbox, box, 48, 32, 3, RESTIR)
    assert np.array_equal(got[0][1], ref)
    srv.close()
    gbox.close()
    gfeat.close()


def test_plan_bytes_bound_the_views_device_memory(gpu_ctx):
    """hdr_view_plan_bytes (what admission charges) covers what a view holds after rendering frames
    of the pipeline, within the argument ring and status word it does not report, on an opaque scene
    and on a translucent one (RNG spill slots: 2 x 2.5 KB per pixel)."""
    from hdrestir_amd import gpu
    cases = [(scenes.box_spheres(48, 40, frames=1), (PT, RESTIR, abi.pipeline("ris", candidate_count=32, max_depth=6,
                                                                             skip_visibility=True))),
             (scenes.features(48, 40, frames=1), (PT, abi.pipeline("ris", candidate_count=8, max_depth=5)))]
    plans = {}
    for sc, pipes in cases:
        gs = gpu.GpuScene(gpu_ctx, sc)
        for k, pipe in enumerate(pipes):
            plan = gs.plan_bytes(48, 40, pipe)
            r = gpu.Renderer(gpu_ctx, gs, 48, 40)
            for f in range(2):
                r.render(_frame(sc, 48, 40, f), pipe, read_color=False)
            r.synchronize()
            held = r.device_bytes()
            print(sc.name, k, "plan", plan, "held", held)
            assert held <= plan <= held + (64 << 10), (plan, held)
            plans[(sc.name, k)] = plan
            r.close()
        gs.close()
    assert plans[("features", 0)] > plans[("box_spheres", 0)] + 48 * 40 * 4096  # spill slots counted


def test_server_on_its_own_thread(gpu_ctx):
    """With own_thread the server admits and renders on its own host thread: the caller only
    submits and waits for results, and every image equals the request rendered alone."""
    from hdrestir_amd import gpu
    box = scenes.box_spheres(64, 64, frames=1)
    feat = scenes.features(64, 64, frames=1, dome=False, emissive=False, opaque=True)
    gbox, gfeat = gpu.GpuScene(gpu_ctx, box), gpu.GpuScene(gpu_ctx, feat)
    gsc = {id(box): gbox, id(feat): gfeat}
    srv = gpu.Server(gpu_ctx, batched=True, max_active=5, own_thread=True)
    with pytest.raises(gpu.HdrError, match="UNSUPPORTED"):
        srv.step()
    reqs = _requests(box, feat)
    ids = {srv.submit(gsc[id(sc)], w, h, spp, _frame(sc, w, h, 0), pipe): k
           for k, (sc, w, h, spp, pipe) in enumerate(reqs)}
    got = {}
    for _ in range(400):
        for res, img in srv.results(images=True, wait_ms=50):
            got[ids[res["id"]]] = img
        if len(got) == len(reqs):
            break
    assert len(got) == len(reqs)
    for k, (sc, w, h, spp, pipe) in enumerate(reqs):
        ref = _alone(gpu, gpu_ctx, gsc[id(sc)], sc, w, h, spp, pipe)
        assert np.array_equal(got[k], ref), k
    srv.close()
    gbox.close()
    gfeat.close()


def test_served_mix_over_three_resident_scenes(gpu_ctx):
    """The serving workload's mix on its three resident scenes -- the box (256^2 .. 1280x720), the
    many-lights scene from its camera and from a point of a 40 m sphere -- with 1080p ReSTIR
    (M=32, 3 spatial neighbours within 30 px) and 1080p PathTracer (depth 4) requests among smaller
    ones, served with continuous batching: every image equals the request rendered alone."""
    from hdrestir_amd import gpu
    box = scenes.box_spheres(256, 256, frames=1)
    many = scenes.many_lights(frames=1, dolly=False)
    gbox, gmany = gpu.GpuScene(gpu_ctx, box), gpu.GpuScene(gpu_ctx, many)
    rng = np.random.default_rng(5)
    sphere_view = scenes.look_at_origin_from_sphere(rng, 40.0)
    restir = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=1,
                          spatial_count=3, spatial_radius=30.0)
    pt = abi.pipeline("mis", max_depth=4)
    box_view, many_view = box.cameras[0][0], many.cameras[0][0]
    # (scene handle, view matrix, width, height, spp, pipeline)
    reqs = [(gmany, many_view, 1920, 1080, 3, restir), (gmany, sphere_view, 1920, 1080, 2, pt),
            (gbox, box_view, 256, 256, 5, pt), (gbox, box_view, 512, 512, 4, restir),
            (gmany, many_view, 1280, 720, 2, pt), (gmany, sphere_view, 512, 512, 3, restir),
            (gbox, box_view, 1280, 720, 1, restir), (gmany, many_view, 256, 256, 6, restir)]

    def frame(view, w, h, f):
        return abi.frame(view, scenes.perspective(50.0, w / h, 0.05, 1000.0), f)
    srv = gpu.Server(gpu_ctx, batched=True, max_active=64, hbm_budget_bytes=64 << 30)
    ids = {srv.submit(g, w, h, spp, frame(v, w, h, 0), p): k for k, (g, v, w, h, spp, p) in enumerate(reqs)}
    got = {}
    for _ in range(200):
        srv.step()
        for res, img in srv.results(images=True):
            assert res["status"] == 0
            got[ids[res["id"]]] = img
        if len(got) == len(reqs):
            break
    assert len(got) == len(reqs)
    for k, (g, v, w, h, spp, p) in enumerate(reqs):
        r = gpu.Renderer(gpu_ctx, g, w, h)
        for f in range(spp):
            r.render(frame(v, w, h, f), p, read_color=False)
        assert np.array_equal(got[k], r.read_color()), k
        r.close()
    srv.close()
    gbox.close()
    gmany.close()
