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
