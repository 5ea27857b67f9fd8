"""Writes the CPU oracle's 4096-spp reference images used by tests/test_gpu_convergence.py.

- box_pt_d4_4096spp.npz: the box of 12 triangles with two icospheres and one rect light, 256 x 256,
  PathTracer (MIS) at depth 4, mean of frames 0 .. 4095 (float16 RGB; the mean radiance of every
  pixel, rounded to half precision).
- many_lights_pt_d1_crop_4096spp.npz: the many-lights scene (height field, 500 spheres, 4096
  lights) from its camera at 480 x 270, PathTracer at depth 1, the 64 x 64 crop at x 208 .. 271,
  y 103 .. 166, mean of frames 0 .. 4095 (float32 RGB).

Both are the average of the oracle's frames on the reference's own seeds (frameIndex 0 .. 4095), so
the GPU's 4096-frame image of the same frames must equal them up to float rounding. Run from the
repository root: python tests/golden/make_convergence_refs.py [--threads N]. About 10 minutes on 8
threads.
"""
import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from hdrestir_amd import abi, scenes  # noqa: E402
import oracle_py  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
SPP = 4096
CROP = (208, 103, 272, 167)  # x0, y0, x1, y1 of the many-lights crop


def mean_image(sc, pipe, rect, threads):
    view, proj = sc.cameras[0]
    o = oracle_py.Oracle(sc, threads=threads)
    x0, y0, x1, y1 = rect
    acc = np.zeros((y1 - y0, x1 - x0, 3), np.float64)
    for f in range(SPP):
        img = o.render(abi.frame(view, proj, f), pipe, rect=rect, debug=False)
        # the oracle's colour is the running mean sum / (f + 1); the frame itself is the difference
        # of consecutive sums, so keep the last mean only
        if f == SPP - 1:
            acc[:] = img[y0:y1, x0:x1, :3]
        if f % 512 == 0:
            print(f"  frame {f}", flush=True)
    o.close()
    return acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    sc = scenes.box_spheres(256, 256, frames=1)
    img = mean_image(sc, abi.pipeline("mis", max_depth=4), (0, 0, 256, 256), a.threads)
    np.savez_compressed(os.path.join(HERE, "box_pt_d4_4096spp.npz"), mean=img.astype(np.float16))
    print("box done", img.mean(axis=(0, 1)))
    sc = scenes.many_lights(480, 270, frames=1, dolly=False)
    img = mean_image(sc, abi.pipeline("mis", max_depth=1), CROP, a.threads)
    np.savez_compressed(os.path.join(HERE, "many_lights_pt_d1_crop_4096spp.npz"), mean=img.astype(np.float32),
                        rect=np.array(CROP, np.int32))
    print("many-lights crop done", img.mean(axis=(0, 1)))


if __name__ == "__main__":
    main()
