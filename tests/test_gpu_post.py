"""The passes after accumulation on the device (hdr_view_post_process): DenoiserPass's chromaticity
blur and PostProcessPass's lens flare and chromatic aberration, against the oracle's restatement of
the reference's CPU loops (oracle/post_process.cpp).

Bar: bit-exact. Every sum is formed in the CPU loop's order (the flare's gather walks each target's
sources in the scatter's raster order) with the same float operations, so the images are equal to
the last bit, on rendered frames of odd size with pixels bright enough to flare."""
import numpy as np
import pytest

from hdrestir_amd import abi, scenes

pytestmark = pytest.mark.gpu

PT = abi.pipeline("mis", max_depth=3)
W, H = 97, 61


def _bright_box():
    sc = scenes.box_spheres(W, H, frames=4)
    for l in sc.lights:
        l["exposure"] = 2.0  # x4: the light and its neighbourhood pass the flare's threshold of 2
    sc.changed()
    return sc


def _render(gpu_ctx, sc, frames):
    from hdrestir_amd import gpu
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, W, H)
    for f in range(frames):
        r.render(sc.frame(f), PT, read_color=False)
    return gs, r


CONFIGS = {
    "chroma": dict(chromaticity_blur=True),
    "flare": dict(lens_flare=True),
    "aberration": dict(chromatic_aberration=5.5),
    "all": dict(chromaticity_blur=True, lens_flare=True, chromatic_aberration=3.25),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_post_process_matches_the_oracle(gpu_ctx, oracle_mod, name):
    kw = CONFIGS[name]
    sc = _bright_box()
    gs, r = _render(gpu_ctx, sc, 3)
    before = r.read_color().reshape(H, W, 4).copy()
    lum = 0.2126 * before[..., 0] + 0.7152 * before[..., 1] + 0.0722 * before[..., 2]
    assert (lum > 2.0).sum() >= 20 and (lum <= 2.0).sum() >= 20
    exp = oracle_mod.post_process(before, chroma_blur=kw.get("chromaticity_blur", False),
                                  lens_flare=kw.get("lens_flare", False),
                                  chromatic_aberration=kw.get("chromatic_aberration", 0.0))
    r.post_process(**kw)
    got = r.read_color().reshape(H, W, 4)
    assert not np.array_equal(got, before)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), np.abs(got - exp).max()
    r.close()
    gs.close()


def test_next_frame_writes_the_plain_mean_again(gpu_ctx):
    """The accumulation is untouched: the frame after a post-process equals the same frame of a
    view that was never post-processed."""
    sc = _bright_box()
    gs, r = _render(gpu_ctx, sc, 2)
    _, r2 = _render(gpu_ctx, sc, 2)
    r.post_process(chromaticity_blur=True, lens_flare=True, chromatic_aberration=2.0)
    r.render(sc.frame(2), PT, read_color=False)
    r2.render(sc.frame(2), PT, read_color=False)
    assert np.array_equal(r.read_color(), r2.read_color())
    r.close()
    r2.close()
    gs.close()


def test_post_process_arguments(gpu_ctx):
    from hdrestir_amd import gpu
    sc = _bright_box()
    gs, r = _render(gpu_ctx, sc, 1)
    before = r.read_color().copy()
    r.post_process()  # every pass off: the reference's passes return early
    assert np.array_equal(r.read_color(), before)
    with pytest.raises(gpu.HdrError, match="UNSUPPORTED"):
        r.post_process(oidn=True)
    for bad in (float("nan"), float("inf"), 2.0e6):
        with pytest.raises(gpu.HdrError, match="INVALID"):
            r.post_process(chromatic_aberration=bad)
    assert np.array_equal(r.read_color(), before)
    r.close()
    gs.close()
