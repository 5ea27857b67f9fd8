"""Resolve, upscale and the progressive render loop (hdr_view_resolve, hdr_rendered_size,
hdr_view_sample_count, hdr_render_progressive).

Bars: the upscale is bit-exact against UpscalePass's nearest-texel rule restated in numpy
(upscale_pass.cpp:17-31: source x = min(sw - 1, x * sw // dw)); the format conversions are bit-exact
against numpy's float32 -> float16 (round to nearest even) and the clamp-and-truncate integer rules
of HdRestirRenderBuffer::_WriteOutput (hd_restir_render_buffer.cpp:46-74); the progressive loop's
image equals the frames 0 .. target-1 rendered one by one."""
import numpy as np
import pytest

from hdrestir_amd import abi, scenes

pytestmark = pytest.mark.gpu

PT = abi.pipeline("mis", max_depth=3)


def _nearest(src, dw, dh):
    sh, sw = src.shape[:2]
    sx = np.minimum(sw - 1, (np.arange(dw, dtype=np.int64) * sw) // dw)
    sy = np.minimum(sh - 1, (np.arange(dh, dtype=np.int64) * sh) // dh)
    return src[sy][:, sx]


def _frame(sc, W, H, f):
    view, _ = sc.cameras[0]
    return abi.frame(view, scenes.perspective(60.0, W / H, 0.01, 100.0), f)


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4, 6])
def test_upscaled_outputs_follow_the_nearest_texel_rule(gpu_ctx, level):
    """A 67 x 45 window at each resolution level (6 clamps to 4): the view renders at
    ceil(window / 2^level); colour and the three AOVs resolve to the window by the nearest texel."""
    from hdrestir_amd import gpu
    W, H = 67, 45
    rw, rh = gpu.rendered_size(W, H, level)
    div = 1 << min(level, 4)
    assert (rw, rh) == (max(1, -(-W // div)), max(1, -(-H // div)))
    sc = scenes.box_spheres(rw, rh, frames=1)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, rw, rh)
    assert not r.resolve(abi.HDR_AOV_DEPTH, W, H, components=1).any()  # enabled, zero until a frame
    for f in range(2):
        r.render(_frame(sc, W, H, f), PT, read_color=False)
    color = r.read_color().reshape(rh, rw, 4)
    got = r.resolve(abi.HDR_AOV_COLOR, W, H)
    assert got.shape == (H, W, 4)
    assert np.array_equal(got, _nearest(color, W, H))
    if level == 0:
        assert np.array_equal(got, color)
    d, a, n = r.aovs()
    for aov, src in ((abi.HDR_AOV_DEPTH, d.reshape(rh, rw, 1)), (abi.HDR_AOV_ALBEDO, a.reshape(rh, rw, 3)),
                     (abi.HDR_AOV_NORMAL, n.reshape(rh, rw, 3))):
        c = src.shape[2]
        assert np.array_equal(r.resolve(aov, W, H, components=c), _nearest(src, W, H)), aov
    assert (d != 0).any() and (a != 0).any()
    r.close()
    gs.close()


def test_render_buffer_formats(gpu_ctx):
    """Colour and depth into every render-buffer format and component count: Float16 rounds to
    nearest even, UNorm8 / SNorm8 clamp then truncate, Int32 truncates, missing components are 0,
    extra source components are dropped."""
    from hdrestir_amd import gpu
    W, H = 40, 28
    sc = scenes.box_spheres(W, H, frames=1)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, W, H)
    r.aovs()
    for f in range(2):
        r.render(_frame(sc, W, H, f), PT, read_color=False)
    color = r.read_color().reshape(H, W, 4).copy()
    depth = r.aovs()[0].reshape(H, W, 1)
    c32 = color.astype(np.float32)
    assert np.array_equal(r.resolve(abi.HDR_AOV_COLOR, W, H, abi.HDR_FORMAT_FLOAT16, 4).view(np.uint16),
                          c32.astype(np.float16).view(np.uint16))
    assert np.array_equal(r.resolve(abi.HDR_AOV_COLOR, W, H, abi.HDR_FORMAT_UNORM8, 4),
                          np.clip(c32 * np.float32(255.0), 0.0, 255.0).astype(np.uint8))
    nrm = r.aovs()[2].reshape(H, W, 3)  # both signs (colours are >= 0)
    assert (nrm < 0).any()
    assert np.array_equal(r.resolve(abi.HDR_AOV_NORMAL, W, H, abi.HDR_FORMAT_SNORM8, 3),
                          np.clip(nrm * np.float32(127.0), -128.0, 127.0).astype(np.int8))
    assert np.array_equal(r.resolve(abi.HDR_AOV_DEPTH, W, H, abi.HDR_FORMAT_INT32, 1), np.trunc(depth).astype(np.int32))
    three = r.resolve(abi.HDR_AOV_COLOR, W, H, abi.HDR_FORMAT_FLOAT32, 3)
    assert np.array_equal(three, color[..., :3])
    four = r.resolve(abi.HDR_AOV_DEPTH, W, H, abi.HDR_FORMAT_FLOAT32, 4)
    assert np.array_equal(four[..., :1], depth) and not four[..., 1:].any()
    with pytest.raises(gpu.HdrError, match="INVALID_ARGUMENT"):
        r.resolve(abi.HDR_AOV_COLOR, W, H, 7, 4)
    r.close()
    gs.close()


def test_progressive_loop_converges_and_cancels(gpu_ctx):
    """target 5, at most 2 frames per call: 2, 4, 5 samples, then no more frames; the image equals
    frames 0 .. 4 rendered one by one. A set cancel flag stops the loop with HDR_ERR_CANCELLED before
    any frame; a reset starts over from sample 0."""
    from hdrestir_amd import gpu
    W, H = 48, 32
    sc = scenes.box_spheres(W, H, frames=1)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, W, H)
    fr = _frame(sc, W, H, 0)
    seen = []
    for _ in range(4):
        conv = r.render_progressive(fr, PT, 5, 2)
        seen.append((conv, r.sample_count()))
    assert seen == [(False, 2), (False, 4), (True, 5), (True, 5)]
    img = r.read_color().copy()
    ref = gpu.Renderer(gpu_ctx, gs, W, H)
    for f in range(5):
        want = ref.render(_frame(sc, W, H, f), PT).copy()
    assert np.array_equal(img, want)
    r.reset()
    assert r.sample_count() == 0
    gpu_ctx.set_cancel(1)
    try:
        with pytest.raises(gpu.HdrError, match="CANCELLED"):
            r.render_progressive(fr, PT, 5, 2)
        assert r.sample_count() == 0
    finally:
        gpu_ctx.set_cancel(0)
    assert r.render_progressive(fr, PT, 1, 10) and r.sample_count() == 1
    r.close()
    ref.close()
    gs.close()


@pytest.mark.parametrize("case", ["pt", "pt_firefly", "restir_deferred", "wavefront_d3"])
def test_variance_metrics_match_the_oracle(gpu_ctx, oracle_mod, case):
    """AccumulationPass's Variance metrics on every accumulation path (the megakernel's or lean path
    tracer's, the firefly kernel's, the deferred path's and the wavefront's accumulation): per-pixel
    luminance sums within the colour bar (1e-3 of the oracle's value, floor 1e-3) on every pixel,
    the statistics within 1e-3; no statistics after one frame; a reset clears the sums; metrics off
    refuses the read."""
    from hdrestir_amd import gpu
    import parity
    W, H = 48, 32
    sc = scenes.box_spheres(W, H, frames=6)
    pipe = {"pt": abi.pipeline("mis", max_depth=3),
            "pt_firefly": abi.pipeline("mis", max_depth=3, firefly=True),
            "restir_deferred": abi.pipeline("ris", candidate_count=16, use_reservoir=True, skip_visibility=True,
                                            max_depth=1, spatial_count=3, spatial_radius=8.0),
            "wavefront_d3": abi.pipeline("ris", candidate_count=16, use_reservoir=True, skip_visibility=True,
                                         max_depth=3)}[case]
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, W, H)
    r.set_metrics(True)
    o = oracle_mod.Oracle(sc, threads=8)
    o.set_firefly_in_place(False)
    for f in range(5):
        r.render(sc.frame(f), pipe, read_color=False)
        o.render(sc.frame(f), pipe)
        if f == 0:
            assert not r.variance()[0]["valid"] and not o.variance()[0]
    st, ls, lq = r.variance()
    ok, ost, ols, olq = o.variance()
    assert ok and st["valid"] and st["sample_count"] == 5
    rel = np.abs(ls - ols) / np.maximum(np.abs(ols), parity.REL_FLOOR)
    print(case, "lum sum max rel", rel.max(), st, ost)
    assert rel.max() < parity.MAX_REL
    relq = np.abs(lq - olq) / np.maximum(np.abs(olq), parity.REL_FLOOR)
    assert relq.max() < 2 * parity.MAX_REL
    for k in ("mean_estimator_variance", "max_estimator_variance", "mean_luminance"):
        assert abs(st[k] - ost[k]) <= 1e-3 * abs(ost[k]) + 1e-9, (k, st[k], ost[k])
    assert st["mean_estimator_variance"] > 0.0
    r.reset()
    st0, ls0, _ = r.variance()
    assert st0["sample_count"] == 0 and not st0["valid"] and not ls0.any()
    r.set_metrics(False)
    with pytest.raises(gpu.HdrError, match="INVALID_ARGUMENT"):
        r.variance()
    r.close()
    gs.close()
