"""Parity bars shared by the GPU tests (imported by them; not a test module).

Streams: every compared pixel's integer streams (draw count, FNV hash of every light-selection
index, FNV hash of every RIS / reservoir selection) equal the oracle's, unless a test states a lower
bar with its reason. Colours: on stream-identical pixels the largest difference is below 1e-3 of
the pixel's largest oracle component (floor 1e-3); the number of components above 1e-5 is printed.
"""
import numpy as np

REL_FLOOR = 1e-3
MAX_REL = 1e-3


def rel_err(a, b):
    """|a - b| relative to max(|b|, 1e-3); for (n, 3) colours relative to the pixel's largest
    component (the RGB of a spectral sample is a difference of CIE terms, so a small channel of a
    saturated colour carries the absolute rounding of the large ones)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.abs(b)
    if b.ndim == 2 and b.shape[1] == 3:
        scale = np.repeat(scale.max(axis=1, keepdims=True), 3, axis=1)
    return np.abs(a - b) / np.maximum(scale, REL_FLOOR)


def check(label, same, a, b, img=None, ref=None, stream_bar=1.0, max_rel=MAX_REL, outlier_pixels=0):
    """same: per-pixel stream match; a, b: GPU / oracle per-frame samples (n, 3) of all pixels;
    img, ref: accumulated (H, W, 4) images, compared on stream-identical pixels. outlier_pixels:
    stream-identical pixels a test allows above max_rel, with its stated reason."""
    same = np.asarray(same).reshape(-1)
    n_bad = int((~same).sum())
    rel = rel_err(np.asarray(a)[same], np.asarray(b)[same])
    worst = float(rel.max()) if rel.size else 0.0
    print(f"{label}: stream parity {same.mean():.6f} ({n_bad} pixels differ), colour max rel {worst:.3g}, "
          f"components > 1e-5: {int((rel > 1e-5).sum())} of {rel.size}", flush=True)
    if stream_bar >= 1.0:
        assert n_bad == 0, f"{label}: {n_bad} pixels with different streams"
    else:
        assert same.mean() >= stream_bar, (label, same.mean())
    n_out = int((rel.max(axis=1) >= max_rel).sum()) if rel.ndim == 2 and rel.size else int((rel >= max_rel).sum())
    assert n_out <= outlier_pixels, (label, worst, n_out)
    if img is not None:
        px = same.reshape(img.shape[0], img.shape[1])
        ra = rel_err(img[px][:, :3], ref[px][:, :3])
        wi = float(ra.max()) if ra.size else 0.0
        assert wi < max_rel, (label, "accumulated", wi)
