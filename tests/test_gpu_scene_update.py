"""Scene updates (hdr_scene_update): instance transforms, instance lists, materials and lights change
between frames while the meshes' BLAS stay resident. The reference rebuilds the instance list, the
TLAS and the lights in Scene::BuildRenderState (scene.cpp:118-271) before every frame, so an updated
scene must render exactly as the same scene uploaded from scratch -- bit for bit -- and match the
oracle built from the edited descriptor at the usual bars (streams 100 %, colours 1e-3 of the
pixel's largest component)."""
import numpy as np
import pytest

import parity

from hdrestir_amd import abi, scenes

pytestmark = pytest.mark.gpu

PT = abi.pipeline("mis", max_depth=4)
RIS = abi.pipeline("ris", candidate_count=16, use_reservoir=True, max_depth=3)
DEFERRED = abi.pipeline("ris", candidate_count=16, use_reservoir=True, skip_visibility=True, max_depth=1,
                        spatial_count=3, spatial_radius=8.0)


def _stream_match(r, o):
    return (r.debug["draws"] == o.debug["draws"]) & (r.debug["light_hash"] == o.debug["light_hash"]) & \
           (r.debug["select_hash"] == o.debug["select_hash"]) & \
           (r.debug["n_light_picks"] == o.debug["n_light_picks"])


def _edit_box(sc):
    """Moves a sphere, recolours the other's material, adds a third sphere sharing the first
    sphere's mesh, retargets and dims the light."""
    sc.instances[1]["transform"] = scenes._translate_scale((0.35, -0.7, -0.3), 0.3)
    sc.materials[sc.instances[2]["material"]]["diffuse_color"] = (0.1, 0.8, 0.2)
    sc.materials[sc.instances[2]["material"]]["roughness"] = 0.2
    sc.add_instance(sc.instances[1]["mesh"], sc.add_material(diffuse_color=(0.9, 0.3, 0.1), metallic=1.0,
                                                             roughness=0.3),
                    scenes._translate_scale((-0.3, -0.5, -0.6), 0.15))
    sc.lights[0]["intensity"] = 12.0
    sc.lights[0]["color"] = np.asarray(scenes.blackbody_rgb(3500.0), np.float32)
    sc.lights[0]["transform"] = scenes._rect_light_down((0.2, 0.97, -0.3), 0.5)
    sc.changed()


def _render(gpu, ctx, gs, sc, pipe, frames, debug=False):
    r = gpu.Renderer(ctx, gs, sc.width, sc.height)
    for f in range(frames):
        img = r.render(sc.frame(f), pipe, debug=debug)
    return r, img.copy()


@pytest.mark.parametrize("pname", ["pt", "ris", "deferred"])
def test_updated_scene_renders_as_a_fresh_upload(gpu_ctx, pname):
    """Frames before the update prime the view (reservoirs, accumulation, the G-buffer of the
    unchanged camera); after the update and a view reset every frame equals the edited scene
    uploaded from scratch, so nothing of the old instances, TLAS, materials, lights or G-buffer
    survives."""
    from hdrestir_amd import gpu
    pipe = {"pt": PT, "ris": RIS, "deferred": DEFERRED}[pname]
    sc = scenes.box_spheres(64, 48, frames=4)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r, before = _render(gpu, gpu_ctx, gs, sc, pipe, 2)
    _edit_box(sc)
    gs.update(sc)
    r.reset()
    fresh = gpu.GpuScene(gpu_ctx, sc)
    rf = gpu.Renderer(gpu_ctx, fresh, sc.width, sc.height)
    for f in range(3):
        a = r.render(sc.frame(f), pipe).copy()
        b = rf.render(sc.frame(f), pipe).copy()
        assert np.array_equal(a, b), (pname, f)
    assert not np.array_equal(a, before)
    r.close()
    rf.close()
    fresh.close()
    gs.close()


@pytest.mark.parametrize("pname", ["pt", "ris"])
def test_updated_scene_matches_the_oracle(gpu_ctx, oracle_mod, pname):
    from hdrestir_amd import gpu
    pipe = {"pt": PT, "ris": RIS}[pname]
    sc = scenes.box_spheres(64, 64, frames=3)
    gs = gpu.GpuScene(gpu_ctx, sc)
    _edit_box(sc)
    gs.update(sc)
    r = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
    o = oracle_mod.Oracle(sc, threads=8)
    for f in range(3):
        img = r.render(sc.frame(f), pipe, debug=True).copy()
        ref = o.render(sc.frame(f), pipe).copy()
        assert np.array_equal(r.debug["prim_instance"], o.debug["prim_instance"])
        assert (r.debug["prim_instance"] == 3).any()  # the added sphere is in view
        parity.check(f"updated box {pname} frame {f}", _stream_match(r, o), r.debug["sample_rgb"],
                     o.debug["sample_rgb"], img, ref)
    r.close()
    gs.close()


def test_update_growing_the_instance_list_rebuilds_the_tlas(gpu_ctx, oracle_mod):
    """Three times the instances (the TLAS outgrows the node room reserved at upload, the state
    records their allocation), then back to fewer than uploaded: closest hits equal a fresh upload's
    and the oracle's on camera rays and random interior rays, after each update."""
    from hdrestir_amd import gpu
    sc = scenes.many_lights(96, 64, n_lights=16, n_spheres=40, grid=64, frames=1)
    gs = gpu.GpuScene(gpu_ctx, sc)
    bytes0 = gs.device_bytes()
    rng = np.random.default_rng(11)
    sph = sc.instances[1]["mesh"]
    for _ in range(80):
        x, z = rng.uniform(-40, 40, 2)
        sc.add_instance(sph, sc.instances[1 + int(rng.integers(0, 40))]["material"],
                        scenes._translate_scale((float(x), float(rng.uniform(0, 8)), float(z)),
                                                float(rng.uniform(0.3, 1.5))))
    sc.changed()
    co, cd = oracle_mod.camera_rays(sc.frame(0), sc.width, sc.height, centre=True)
    o2 = rng.uniform((-45, -2, -45), (45, 10, 45), (4000, 3)).astype(np.float32)
    d2 = rng.normal(size=(4000, 3))
    d2 = (d2 / np.linalg.norm(d2, axis=1, keepdims=True)).astype(np.float32)
    org = np.concatenate([co, o2]).astype(np.float32)
    dirs = np.concatenate([cd, d2]).astype(np.float32)
    for step in range(2):
        gs.update(sc)
        fresh = gpu.GpuScene(gpu_ctx, sc)
        gi, gt, gtt = gs.intersect(org, dirs)
        fi, ft, ftt = fresh.intersect(org, dirs)
        assert np.array_equal(gi, fi) and np.array_equal(gt, ft) and np.array_equal(gtt, ftt), step
        oi, ot, ott = oracle_mod.Oracle(sc, threads=8).intersect(org, dirs)
        assert np.array_equal(gi, oi), step
        hit = oi >= 0
        assert (gt[hit] == ot[hit]).mean() >= 0.9999
        assert (np.abs(gtt[hit] - ott[hit]) <= 1e-5 * np.abs(ott[hit])).all()
        if step == 0:
            assert (gi >= 41).sum() > 100  # rays hit the added spheres
            assert gs.device_bytes() > bytes0
        fresh.close()
        # second round: fewer instances than at upload (the rooms are reused)
        del sc.instances[20:]
        sc.changed()
    gs.close()


def test_dome_kept_or_rebuilt_by_texels(gpu_ctx):
    """A light update that leaves the dome's texels alone keeps its device CDFs; new texels
    rebuild them. Either way the CDFs and the image equal a fresh upload's."""
    from hdrestir_amd import gpu
    sc = scenes.features(40, 32, frames=2)
    dome = sc.environment
    dh, dw = sc.lights[dome]["dome"].shape[:2]
    gs = gpu.GpuScene(gpu_ctx, sc)
    pipe = abi.pipeline("mis", max_depth=3)
    for edit in ("intensity", "texels"):
        if edit == "intensity":
            sc.lights[0]["intensity"] = 30.0
            sc.lights[dome]["intensity"] = 0.5
        else:
            tex = sc.lights[dome]["dome"].copy()
            tex[: dh // 2] *= 4.0
            sc.lights[dome]["dome"] = tex
        sc.changed()
        gs.update(sc)
        fresh = gpu.GpuScene(gpu_ctx, sc)
        for a, b in zip(gs.dome_cdf(dome, dw, dh), fresh.dome_cdf(dome, dw, dh)):
            assert np.array_equal(np.asarray(a), np.asarray(b)), edit
        ra, ia = _render(gpu, gpu_ctx, gs, sc, pipe, 2)
        rb, ib = _render(gpu, gpu_ctx, fresh, sc, pipe, 2)
        assert np.array_equal(ia, ib), edit
        ra.close()
        rb.close()
        fresh.close()
    gs.close()


def test_invalid_update_leaves_the_scene_unchanged(gpu_ctx):
    """An update whose mesh count differs from the upload's, or whose instance names a missing
    material, is refused with HDR_ERR_INVALID_ARGUMENT and the scene renders as before."""
    from hdrestir_amd import gpu
    sc = scenes.box_spheres(48, 32, frames=2)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r0, before = _render(gpu, gpu_ctx, gs, sc, PT, 2)
    r0.close()
    bad = scenes.box_spheres(48, 32, frames=2)
    bad.add_mesh(*scenes.box_mesh(0.5))
    bad.changed()
    with pytest.raises(gpu.HdrError, match="INVALID_ARGUMENT"):
        gs.update(bad)
    bad2 = scenes.box_spheres(48, 32, frames=2)
    bad2.instances[1]["material"] = 99
    bad2.changed()
    with pytest.raises(gpu.HdrError, match="INVALID_ARGUMENT"):
        gs.update(bad2)
    r1, after = _render(gpu, gpu_ctx, gs, sc, PT, 2)
    assert np.array_equal(before, after)
    r1.close()
    gs.close()
