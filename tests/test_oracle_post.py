"""Pins the oracle's passes after accumulation (oracle/post_process.cpp) against what the
mathematics fixes. CPU only:

  * chromaticity blur (denoiser.cpp:54-94): in the lifting YCoCg basis Y = (r + 2g + b) / 4,
    Co = (r - b) / 2, Cg = (-r + 2g - b) / 4 with inverse r = Y + Co - Cg, g = Y + Cg, b = Y - Co - Cg,
    the output keeps each pixel's own luma and takes the mean chroma of its 3x3 neighbourhood with
    edge pixels repeated (float64 brute force over the neighbourhood), clamped at 0; a constant image
    is returned unchanged;
  * lens flare (post_process.cpp:23-109): a lone pixel of colour (c, c, c), c > 2, adds
    0.1 (c - 2) / (|d| + 1) at offsets d = -B, -B + 2, ..., B (d != 0, B = max(1, min(W, H) / 20))
    along its row and column and half of that along both diagonals, nothing elsewhere; the flare is
    additive over separated sources; an image whose Rec.709 luminance stays at or under 2 is
    unchanged, and a pure green pixel scatters (every channel's excess over 2) just above 2 / 0.7152
    and not just below;
  * chromatic aberration (post_process.cpp:111-142): red comes from a pixel closer to the image
    centre along the pixel's radius, blue from one further out, both ca r / r_max away (up to the
    truncation to whole pixels); green, alpha and the centre pixel are kept, and ca = 0 is the
    identity.
Planted in the oracle, each of these fails a test here: Cg with +0.25 r, the neighbours outside the
image skipped instead of clamped, the luma instead of the chroma averaged; the Rec.601 luminance weights, offsets in
steps of 1, the diagonals at full weight, no diagonals; red and blue shifted the other way, the
shift without the r / r_max factor.
"""
import numpy as np
import pytest


def _post(oracle_mod, img, **kw):
    return oracle_mod.post_process(img, **kw)


def _ycocg(rgb):
    r, g, b = (rgb[..., k].astype(np.float64) for k in range(3))
    return (r + 2 * g + b) / 4, (r - b) / 2, (-r + 2 * g - b) / 4


def test_chroma_blur_keeps_luma_and_averages_chroma(oracle_mod):
    rng = np.random.default_rng(1)
    h, w = 23, 31
    img = rng.uniform(0.5, 2.0, (h, w, 4)).astype(np.float32)
    out = _post(oracle_mod, img, chroma_blur=True)
    y, co, cg = _ycocg(img)
    pad = lambda a: np.pad(a, 1, mode="edge")
    co_m = sum(pad(co)[1 + dy:1 + dy + h, 1 + dx:1 + dx + w] for dy in (-1, 0, 1) for dx in (-1, 0, 1)) / 9
    cg_m = sum(pad(cg)[1 + dy:1 + dy + h, 1 + dx:1 + dx + w] for dy in (-1, 0, 1) for dx in (-1, 0, 1)) / 9
    exp = np.stack([y + co_m - cg_m, y + cg_m, y - co_m - cg_m], -1)
    assert (exp > 0).all()  # no clamping on this image
    np.testing.assert_allclose(out[..., :3], exp, rtol=2e-6, atol=2e-6)
    np.testing.assert_array_equal(out[..., 3], img[..., 3])
    yo, _, _ = _ycocg(out)
    np.testing.assert_allclose(yo, y, rtol=2e-6, atol=2e-6)


def test_chroma_blur_of_a_constant_image_is_the_image(oracle_mod):
    img = np.empty((9, 12, 4), np.float32)
    img[...] = (0.3, 1.7, 0.9, 0.5)
    out = _post(oracle_mod, img, chroma_blur=True)
    np.testing.assert_allclose(out, img, rtol=1e-6, atol=1e-7)


def test_chroma_blur_clamps_at_zero(oracle_mod):
    img = np.zeros((8, 8, 4), np.float32)
    img[4, 4, :3] = (0.0, 0.0, 9.0)  # a blue spike on black
    out = _post(oracle_mod, img, chroma_blur=True)
    assert (out[..., :3] >= 0).all()
    # a black neighbour: luma 0, mean chroma Co = -4.5 / 9, Cg = -2.25 / 9, so r = g = -0.25 -> 0, b = 0.75
    np.testing.assert_allclose(out[4, 3, :3], (0.0, 0.0, 0.75), rtol=1e-6)
    assert out[4, 3, 0] == 0.0 and out[4, 3, 1] == 0.0


def _flare_closed_form(h, w, x0, y0, c):
    B = max(1, min(w, h) // 20)
    add = np.zeros((h, w), np.float64)
    for d in range(-B, B + 1, 2):
        if d == 0:
            continue
        wgt = 0.1 * (c - 2.0) / (abs(d) + 1.0)
        for (x, y, f) in ((x0 + d, y0, 1.0), (x0, y0 + d, 1.0), (x0 + d, y0 + d, 0.5), (x0 + d, y0 - d, 0.5)):
            if 0 <= x < w and 0 <= y < h:
                add[y, x] += wgt * f
    return add


@pytest.mark.parametrize("h,w", [(100, 120), (61, 47)])
def test_lens_flare_of_one_bright_pixel(oracle_mod, h, w):
    x0, y0, c = w // 2 - 3, h // 2 + 2, 7.0
    img = np.zeros((h, w, 4), np.float32)
    img[..., 3] = 1.0
    img[y0, x0, :3] = c
    out = _post(oracle_mod, img, lens_flare=True)
    exp = _flare_closed_form(h, w, x0, y0, c)
    exp[y0, x0] += c
    for k in range(3):
        np.testing.assert_allclose(out[..., k], exp, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(out[..., 3], img[..., 3])


def test_lens_flare_is_additive_and_thresholded(oracle_mod):
    h, w = 80, 80
    img = np.full((h, w, 4), 1.5, np.float32)  # luminance 1.5 everywhere: under the threshold
    out = _post(oracle_mod, img, lens_flare=True)
    np.testing.assert_array_equal(out, img)
    img[10, 12, :3] = 5.0
    img[50, 60, :3] = 3.0
    out = _post(oracle_mod, img, lens_flare=True)
    exp = _flare_closed_form(h, w, 12, 10, 5.0) + _flare_closed_form(h, w, 60, 50, 3.0)
    np.testing.assert_allclose(out[..., 1] - img[..., 1], exp, rtol=1e-5, atol=1e-6)


def test_lens_flare_threshold_uses_rec709_luminance(oracle_mod):
    h, w = 60, 60
    img = np.zeros((h, w, 4), np.float32)
    img[20, 20, 1] = 2.9  # luminance 0.7152 * 2.9 = 2.07: scatters (-2, 0.9, -2)
    img[40, 40, 1] = 2.7  # 1.93: does not
    out = _post(oracle_mod, img, lens_flare=True)
    exp = _flare_closed_form(h, w, 20, 20, 2.9)
    np.testing.assert_allclose(out[..., 1] - img[..., 1], exp, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(out[..., 0], exp * (-2.0 / 0.9), rtol=1e-5, atol=1e-6)


def test_chromatic_aberration_shifts_red_in_and_blue_out(oracle_mod):
    h, w, ca = 64, 96, 6.0
    ys, xs = np.mgrid[0:h, 0:w]
    img = np.zeros((h, w, 4), np.float32)
    code = (xs + 1000 * ys).astype(np.float32)
    img[..., 0] = code
    img[..., 1] = code + 0.5
    img[..., 2] = code
    img[..., 3] = 0.25
    out = _post(oracle_mod, img, chromatic_aberration=ca)
    np.testing.assert_array_equal(out[..., 1], img[..., 1])
    np.testing.assert_array_equal(out[..., 3], img[..., 3])
    cx, cy = w * 0.5, h * 0.5
    rmax = 0.5 * np.hypot(w, h)
    rx, ry = out[..., 0].astype(int) % 1000, out[..., 0].astype(int) // 1000
    bx, by = out[..., 2].astype(int) % 1000, out[..., 2].astype(int) // 1000
    dx, dy = xs - cx, ys - cy
    r = np.hypot(dx, dy)
    inner = (xs > 8) & (xs < w - 9) & (ys > 8) & (ys < h - 9) & (r > 2)  # no clamping at the borders
    shift = ca * r / rmax
    for sx, sy, sign in ((rx, ry, -1), (bx, by, 1)):
        mx, my = sx - xs, sy - ys
        # whole-pixel truncation of (x + sign * dir * shift): within one pixel per axis of the exact point
        assert (np.abs(mx - sign * dx / np.maximum(r, 1e-9) * shift)[inner] < 1.0 + 1e-4).all()
        assert (np.abs(my - sign * dy / np.maximum(r, 1e-9) * shift)[inner] < 1.0 + 1e-4).all()
        # and on the right side of the radius where the shift reaches a whole pixel
        far = inner & (shift > 2.5)
        assert ((mx * dx + my * dy)[far] * sign > 0).all()
    c0 = (h // 2, w // 2)
    assert out[c0][0] == img[c0][0] and out[c0][2] == img[c0][2]
    np.testing.assert_array_equal(_post(oracle_mod, img, chromatic_aberration=0.0), img)
