"""Convergence against 4096-spp references: the per-pixel MSE of each estimator at a stated spp, and
the mean radiance of the unbiased estimators inside a 99 % interval (batch means over 32 batches).

The references are the CPU oracle's 4096-spp PathTracer (MIS) images, committed under tests/golden/
(written by tests/golden/make_convergence_refs.py): the whole 256 x 256 box image, and a 64 x 64
crop of the many-lights view. The GPU renders the same 4096 frames (same seeds) first, and its image
must equal the oracle's to rounding (the whole image for the box, the crop for the many-lights
view); the box bars then use the oracle's image, the many-lights bars the GPU image that equals
the oracle's on the crop. The CI checks use the GPU run's batch means.

Workloads: the box of 12 triangles with two icospheres and one rect light, 256 x 256, depth 4
(PathTracer, RIS M = 16 without and with the reservoir, ReSTIR with one shadow ray with and without
3 spatial neighbours); the many-lights scene (height field, 500
spheres, 4096 point / rect lights) from its camera at 480 x 270, direct lighting (depth 1: RIS
M = 32, ReSTIR M = 32 with one shadow ray, with and without 3 spatial neighbours).

Relative MSE = mean over pixels and RGB of (image - reference)^2 / mean of reference^2. The
thresholds are the stated bars; the measured values are printed.
"""
import os

import numpy as np
import pytest

from hdrestir_amd import abi, scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

REF_SPP = 4096
BATCHES = 32


def _render_mean(r, sc_frame, pipe, first, n, batch=None):
    """Renders frames first .. first+n-1 into a freshly reset view. Returns the mean image of those
    frames (float64, RGB) and, with `batch`, the per-batch mean radiance (n / batch batches). The
    view's colour is its accumulated sum divided by frame_index + 1 (the renderer's progressive
    normalisation), so the sum is recovered by multiplying back."""
    r.reset()
    means = []
    prev_sum = np.zeros(3)
    for i in range(n):
        r.render(sc_frame(first + i), pipe, read_color=False)
        if batch and (i + 1) % batch == 0:
            img = r.read_color()[..., :3].astype(np.float64)
            tot = img.mean(axis=(0, 1)) * (first + i + 1)  # sum of the per-frame image means so far
            means.append((tot - prev_sum) / batch)
            prev_sum = tot
    img = r.read_color()[..., :3].astype(np.float64) * (first + n) / n
    return img, np.array(means)


def _rel_mse(img, ref):
    return float(((img - ref) ** 2).mean() / (ref ** 2).mean())


def _ci_check(batch_means, ref_batch_means, name):
    """The difference of the two mean radiances inside its 99 % interval: standard errors from the
    batch means of each run (32 batches each), t quantile 2.744 (31 degrees of freedom)."""
    m, rm = batch_means.mean(0), ref_batch_means.mean(0)
    se = np.sqrt(batch_means.var(0, ddof=1) / len(batch_means) + ref_batch_means.var(0, ddof=1) / len(ref_batch_means))
    z = np.abs(m - rm) / np.maximum(se, 1e-12)
    print(f"{name}: mean {m} reference {rm} se {se} |z| {z}")
    assert (np.abs(m - rm) <= 2.744 * se).all(), (m, rm, se)


@pytest.fixture(scope="module")
def box(gpu_ctx):
    from hdrestir_amd import gpu
    sc = scenes.box_spheres(256, 256, frames=1)
    view, proj = sc.cameras[0]
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
    fr = lambda f: abi.frame(view, proj, f)  # noqa: E731
    gref, ref_bm = _render_mean(r, fr, abi.pipeline("mis", max_depth=4), 0, REF_SPP, batch=REF_SPP // BATCHES)
    ref = np.load(os.path.join(GOLDEN, "box_pt_d4_4096spp.npz"))["mean"].astype(np.float64)
    e = _rel_mse(gref, ref)  # float16 storage of the oracle's image: ~1e-7
    print(f"box: GPU 4096-spp image vs the oracle's: relative MSE {e:.3g}")
    assert e < 1e-6, e
    assert np.allclose(gref.mean((0, 1)), ref.mean((0, 1)), rtol=1e-4)
    yield r, fr, ref, ref_bm
    r.close()
    gs.close()


@pytest.fixture(scope="module")
def many(gpu_ctx):
    from hdrestir_amd import gpu
    sc = scenes.many_lights(480, 270, frames=1, dolly=False)
    view, proj = sc.cameras[0]
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
    fr = lambda f: abi.frame(view, proj, f)  # noqa: E731
    ref, ref_bm = _render_mean(r, fr, abi.pipeline("mis", max_depth=1), 0, REF_SPP, batch=REF_SPP // BATCHES)
    g = np.load(os.path.join(GOLDEN, "many_lights_pt_d1_crop_4096spp.npz"))
    x0, y0, x1, y1 = (int(v) for v in g["rect"])
    crop, oref = ref[y0:y1, x0:x1], g["mean"].astype(np.float64)
    e = _rel_mse(crop, oref)
    print(f"many-lights: GPU 4096-spp crop vs the oracle's: relative MSE {e:.3g}")
    assert e < 1e-6, e
    assert np.allclose(crop.mean((0, 1)), oref.mean((0, 1)), rtol=1e-4)
    yield r, fr, ref, ref_bm
    r.close()
    gs.close()


# (pipeline, spp for the MSE bar, relative-MSE bar, unbiased -> CI test)
BOX_CASES = {
    "pt_d4": (abi.pipeline("mis", max_depth=4), 64, 0.05, True),
    "ris16_d4": (abi.pipeline("ris", candidate_count=16, use_reservoir=False, max_depth=4), 64, 0.03, True),
    "restir16_d4": (abi.pipeline("ris", candidate_count=16, use_reservoir=True, max_depth=4), 64, 0.03, False),
    # spatial reuse (an extension) at depth 4: the primary vertex's direct light resampled with 3
    # neighbours, the later vertices as without it
    "restir16_d4_skipvis_spatial3": (abi.pipeline("ris", candidate_count=16, use_reservoir=True,
                                                  skip_visibility=True, max_depth=4, spatial_count=3,
                                                  spatial_radius=30.0), 64, 0.06, False),
    "restir16_d4_skipvis": (abi.pipeline("ris", candidate_count=16, use_reservoir=True, skip_visibility=True,
                                         max_depth=4), 64, 0.03, False),
}


@pytest.mark.parametrize("name", list(BOX_CASES))
def test_box_convergence(box, name):
    r, fr, ref, ref_bm = box
    pipe, spp, bar, unbiased = BOX_CASES[name]
    first = 100000  # seeds disjoint from the reference's frames 0 .. 4095
    img, _ = _render_mean(r, fr, pipe, first, spp)
    e = _rel_mse(img, ref)
    print(f"box {name}: relative MSE at {spp} spp = {e:.5f}")
    assert e < bar, e
    if unbiased:
        _, bm = _render_mean(r, fr, pipe, first + spp, BATCHES * 8, batch=8)
        _ci_check(bm, ref_bm, f"box {name}")


MANY_CASES = {
    "pt_d1": (abi.pipeline("mis", max_depth=1), 64, 1.0, True),
    "ris32_d1": (abi.pipeline("ris", candidate_count=32, use_reservoir=False, max_depth=1), 64, 0.11, True),
    "restir32_d1_skipvis": (abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True,
                                         max_depth=1), 64, 1.0, False),
    "restir32_d1_skipvis_spatial3": (abi.pipeline("ris", candidate_count=32, use_reservoir=True,
                                                  skip_visibility=True, max_depth=1, spatial_count=3,
                                                  spatial_radius=30.0), 64, 0.75, False),
}


@pytest.mark.parametrize("name", list(MANY_CASES))
def test_many_lights_convergence(many, name):
    r, fr, ref, ref_bm = many
    pipe, spp, bar, unbiased = MANY_CASES[name]
    first = 100000
    img, _ = _render_mean(r, fr, pipe, first, spp)
    e = _rel_mse(img, ref)
    print(f"many-lights {name}: relative MSE at {spp} spp = {e:.5f}")
    assert e < bar, e
    if unbiased:
        _, bm = _render_mean(r, fr, pipe, first + spp, BATCHES * 4, batch=4)
        _ci_check(bm, ref_bm, f"many-lights {name}")
