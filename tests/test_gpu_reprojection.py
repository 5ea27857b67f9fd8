"""Reprojected temporal reuse (an extension: hdr_pipeline_desc::temporal_reprojection) on the
device against the oracle's restatement of the same extension.

A camera path of six frames -- A, A, B, C, C, D -- where each move clears only the accumulation
(hdr_view_reset_accumulation; the frame index restarts at 0, as after any move), so frames 2, 3 and
5 reproject the previous frame's reservoirs into
the spatial pass and frames 1 and 4 merge the same pixel as the reference does. On the deferred path
(depth 1) and the wavefront (depth 3, with and without traced candidates): integer streams identical
on every pixel (the selection hash carries the chosen domain, so the reprojection map and every
resampling decision are compared), colours and reservoirs within the parity bars, and the
reprojected domain found for most pixels. Frames of a static camera render exactly as without the
flag; unsupported uses are refused."""
import numpy as np
import pytest

import parity

from hdrestir_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _translate(view, t):
    T = np.eye(4)
    T[3, :3] = -np.asarray(t, np.float64)
    return view @ T


# the camera of each frame, moved in its own axes (A, A, B, C, C, D)
STEPS = [(0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.01, 0.004, -0.02), (0.018, 0.006, -0.05), (0.018, 0.006, -0.05),
         (0.03, -0.004, -0.06)]

CASES = {
    # deferred path (RIS at depth 1, skipVisibility, nothing BSDF-sampled reaches an emitter)
    "box_deferred_m16_s3": (lambda: scenes.box_spheres(48, 40, frames=1), 16, 1, 3, 10.0),
    "features_deferred_m8_s2": (lambda: scenes.features(48, 40, dome=False, emissive=False, opaque=True), 8, 1, 2,
                                6.0),
    # wavefront path, lean shading launches
    "box_wave_d3_m8_s3": (lambda: scenes.box_spheres(48, 40, frames=1), 8, 3, 3, 10.0),
    # wavefront path with traced candidates (environment)
    "features_dome_wave_d2_m8_s2": (lambda: scenes.features(48, 40, dome=True, emissive=False, mesh_light=False,
                                                            opaque=True), 8, 2, 2, 6.0),
}


def _match(r, o):
    return (r.debug["draws"] == o.debug["draws"]) & (r.debug["light_hash"] == o.debug["light_hash"]) & \
           (r.debug["select_hash"] == o.debug["select_hash"])


@pytest.mark.parametrize("case", list(CASES))
def test_reprojection_parity_on_a_camera_path(gpu_ctx, oracle_mod, case):
    from hdrestir_amd import gpu
    make, m, depth, k, radius = CASES[case]
    sc = make()
    pipe = abi.pipeline("ris", candidate_count=m, use_reservoir=True, skip_visibility=True, max_depth=depth,
                        spatial_count=k, spatial_radius=radius, reprojection=True)
    view0, proj = sc.cameras[0]
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
    o = oracle_mod.Oracle(sc, threads=8)
    r.stage_times()
    index = -1
    for f, step in enumerate(STEPS):
        moved = f > 0 and step != STEPS[f - 1]
        if moved:  # a camera move: the accumulation and the frame count restart, the reservoirs stay
            r.reset_accumulation()
            o.reset_accumulation()
            index = -1
        index += 1
        fr = abi.frame(_translate(view0, step), proj, index)
        img = r.render(fr, pipe, debug=True).copy()
        ref = o.render(fr, pipe).copy()
        assert np.array_equal(r.debug["prim_instance"], o.debug["prim_instance"])
        used = (o.reprojection() >= 0).mean()
        print(case, "frame", f, "moved" if moved else "static", "reprojected pixels", round(float(used), 3))
        if moved:
            assert used > 0.5, used
        else:
            assert used == 0.0
        same = _match(r, o)
        parity.check(f"{case} frame {f}", same, r.debug["sample_rgb"], o.debug["sample_rgb"], img, ref)
        gr, orr = r.reservoirs(), o.reservoirs()
        assert np.array_equal(gr["flags_light"], orr["flags_light"])
        assert np.array_equal(gr["count"], orr["count"])
        for key in ("dir", "target", "w", "wsum"):
            x, y = gr[key], orr[key]
            assert parity.rel_err(x.reshape(len(x), -1), y.reshape(len(y), -1)).max() < parity.MAX_REL, key
        assert r.sample_count() == index + 1  # frames accumulated since the last move
    assert r.stage_times()["spatial"][1] == len(STEPS)
    r.close()
    gs.close()


def test_static_camera_frames_are_unchanged_by_the_flag(gpu_ctx):
    from hdrestir_amd import gpu
    sc = scenes.box_spheres(64, 48, frames=1)
    base = dict(candidate_count=16, use_reservoir=True, skip_visibility=True, spatial_count=3, spatial_radius=10.0)
    gs = gpu.GpuScene(gpu_ctx, sc)
    for depth in (1, 3):
        a = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
        b = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
        for f in range(4):
            fr = sc.frame(f, camera_index=0)
            a.render(fr, abi.pipeline("ris", max_depth=depth, **base), read_color=False)
            b.render(fr, abi.pipeline("ris", max_depth=depth, reprojection=True, **base), read_color=False)
        assert np.array_equal(a.read_color(), b.read_color())
        assert np.array_equal(a.reservoirs(), b.reservoirs())
        a.close()
        b.close()
    gs.close()


def test_reprojection_refusals(gpu_ctx):
    from hdrestir_amd import gpu
    sc = scenes.box_spheres(16, 16, frames=1)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, 16, 16)
    base = dict(candidate_count=8, use_reservoir=True, skip_visibility=True, max_depth=1, reprojection=True)
    with pytest.raises(gpu.HdrError, match="UNSUPPORTED"):
        r.render(sc.frame(0), abi.pipeline("ris", **base))  # no spatial reuse
    with pytest.raises(gpu.HdrError, match="INVALID"):
        r.render(sc.frame(0), abi.pipeline("ris", spatial_count=8, spatial_radius=4.0, **base))
    batch = gpu.Batch(gpu_ctx)
    with pytest.raises(gpu.HdrError, match="UNSUPPORTED"):
        batch.render([r], [sc.frame(0)], [abi.pipeline("ris", spatial_count=2, spatial_radius=4.0, **base)])
    batch.close()
    r.close()
    gs.close()


def test_reset_accumulation_keeps_the_reservoirs(gpu_ctx):
    """hdr_view_reset_accumulation clears the accumulation and the sample count and keeps the
    reservoirs (the next frame at frame index 0 shows its own sample alone); hdr_view_reset clears
    both."""
    from hdrestir_amd import gpu
    sc = scenes.box_spheres(64, 48, frames=1)
    pipe = abi.pipeline("ris", candidate_count=8, use_reservoir=True, skip_visibility=True, max_depth=1,
                        spatial_count=2, spatial_radius=6.0)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
    for f in range(3):
        r.render(sc.frame(f, camera_index=0), pipe, read_color=False)
    before = r.reservoirs().copy()
    assert r.sample_count() == 3
    r.reset_accumulation()
    assert r.sample_count() == 0
    assert np.array_equal(r.reservoirs(), before)  # the reservoirs stay
    img = r.render(sc.frame(0, camera_index=0), pipe, debug=True).copy()
    assert r.sample_count() == 1
    # the accumulation restarted: at frame index 0 the colour is this frame's sample alone
    np.testing.assert_array_equal(img.reshape(-1, 4)[:, :3], r.debug["sample_rgb"])
    r.reset()
    assert r.sample_count() == 0
    assert (r.reservoirs()["flags_light"] == 0).all()  # hdr_view_reset clears them
    r.close()
    gs.close()
