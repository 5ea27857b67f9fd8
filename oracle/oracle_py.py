"""ctypes front end of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
CPU-baseline leg, to check results against. The product library never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from hdrestir_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [C.POINTER(abi.SceneDesc), C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.oracle_destroy.argtypes = [P]
        L.oracle_reset.argtypes = [P]
        L.oracle_reset_accumulation.argtypes = [P]
        L.oracle_read_reprojection.argtypes = [P, P]
        L.oracle_set_firefly_in_place.argtypes = [P, C.c_int32]
        L.oracle_render_frame.restype = C.c_double
        L.oracle_render_frame.argtypes = [P, C.POINTER(abi.FrameDesc), C.POINTER(abi.PipelineDesc), C.c_int32,
                                          C.c_int32, C.c_int32, C.c_int32, P, P]
        L.oracle_read_reservoirs.argtypes = [P, P]
        L.oracle_read_variance.restype = C.c_int32
        L.oracle_read_variance.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P, P]
        L.oracle_read_aovs.argtypes = [P, P, P, P]
        L.oracle_read_counters.argtypes = [P, C.POINTER(abi.FrameCounters)]
        L.oracle_intersect.argtypes = [P, P, P, C.c_int32, P, P, P]
        L.oracle_dome_cdf.argtypes = [P, C.c_int32, P, P, P]
        L.oracle_dome_cdf.restype = C.c_int32
        L.oracle_rng_raw.argtypes = [P, C.c_int32, C.c_int32, P]
        L.oracle_rng_float.argtypes = [P, C.c_int32, C.c_int32, P]
        L.oracle_camera_rays.argtypes = [C.POINTER(abi.FrameDesc), C.c_int32, C.c_int32, C.c_int32, P, P]
        L.oracle_post_process.restype = C.c_int32
        L.oracle_post_process.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One CPU renderer instance over a scene (== one RenderPipeline with its persistent store)."""

    def __init__(self, scene, hoist_light_set=True, threads=0):
        self.scene = scene
        self.W, self.H = scene.width, scene.height
        self.h = lib().oracle_create(C.byref(scene.desc), self.W, self.H, int(hoist_light_set), threads)
        self.color = np.zeros((self.H, self.W, 4), np.float32)
        self.debug = np.zeros(self.H * self.W, abi.debug_dtype())
        self.last_seconds = 0.0

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def reset(self):
        lib().oracle_reset(self.h)

    def reset_accumulation(self):
        lib().oracle_reset_accumulation(self.h)

    def reprojection(self):
        """Per pixel (H x W), the previous frame's pixel index the last frame reprojected it to (-1: none)."""
        out = np.empty(self.H * self.W, np.int32)
        lib().oracle_read_reprojection(self.h, _p(out))
        return out.reshape(self.H, self.W)

    def set_firefly_in_place(self, v):
        lib().oracle_set_firefly_in_place(self.h, int(v))

    def render(self, frame, pipe, rect=None, debug=True):
        x0, y0, x1, y1 = rect if rect else (0, 0, self.W, self.H)
        self.last_seconds = lib().oracle_render_frame(self.h, C.byref(frame), C.byref(pipe), x0, y0, x1, y1,
                                                      _p(self.color), _p(self.debug) if debug else None)
        if self.last_seconds < 0:
            raise ValueError("the oracle does not define this pipeline (spatial reuse needs ReSTIR: RIS with the "
                             "reservoir; temporal reprojection needs spatial reuse with at most 7 neighbours)")
        return self.color

    def reservoirs(self):
        out = np.zeros(self.H * self.W, abi.reservoir_dtype())
        lib().oracle_read_reservoirs(self.h, _p(out))
        return out

    def variance(self, rect=None):
        """AccumulationPass's Variance metrics over `rect` (default: the whole image): (valid,
        {mean_estimator_variance, max_estimator_variance, mean_luminance}, per-pixel luminance sums,
        per-pixel sums of squares)."""
        x0, y0, x1, y1 = rect if rect else (0, 0, self.W, self.H)
        out = np.zeros(3, np.float64)
        ls = np.zeros(self.H * self.W, np.float64)
        lq = np.zeros(self.H * self.W, np.float64)
        ok = lib().oracle_read_variance(self.h, x0, y0, x1, y1, _p(out), _p(ls), _p(lq))
        stats = dict(mean_estimator_variance=out[0], max_estimator_variance=out[1], mean_luminance=out[2])
        return bool(ok), stats, ls, lq

    def aovs(self):
        d = np.zeros(self.H * self.W, np.float32)
        a = np.zeros((self.H * self.W, 3), np.float32)
        n = np.zeros((self.H * self.W, 3), np.float32)
        lib().oracle_read_aovs(self.h, _p(d), _p(a), _p(n))
        return d, a, n

    def counters(self):
        c = abi.FrameCounters()
        lib().oracle_read_counters(self.h, C.byref(c))
        return c.as_dict()

    def dome_cdf(self, light, w, h):
        """(row_cdf, col_cdf (h, w+1), total luminance) of dome light `light`."""
        row = np.zeros(h + 1, np.float32)
        col = np.zeros((h, w + 1), np.float32)
        tot = C.c_float()
        if lib().oracle_dome_cdf(self.h, light, _p(row), _p(col), C.byref(tot)) != 0:
            raise ValueError("not a dome light with texels")
        return row, col, float(tot.value)

    def intersect(self, origins, dirs):
        origins = np.ascontiguousarray(origins, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = origins.shape[0]
        inst = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        lib().oracle_intersect(self.h, _p(origins), _p(dirs), n, _p(inst), _p(tri), _p(t))
        return inst, tri, t


def rng_raw(seeds, n_draws):
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.zeros((len(seeds), n_draws), np.uint64)
    lib().oracle_rng_raw(_p(seeds), len(seeds), n_draws, _p(out))
    return out


def rng_float(seeds, n_draws):
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.zeros((len(seeds), n_draws), np.float32)
    lib().oracle_rng_float(_p(seeds), len(seeds), n_draws, _p(out))
    return out


def camera_rays(frame, width, height, centre):
    o = np.zeros((height * width, 3), np.float32)
    d = np.zeros((height * width, 3), np.float32)
    lib().oracle_camera_rays(C.byref(frame), width, height, int(centre), _p(o), _p(d))
    return o, d


def post_process(rgba, chroma_blur=False, lens_flare=False, chromatic_aberration=0.0):
    """DenoiserPass's prefilter then PostProcessPass on an H x W x 4 float32 image; returns a new image."""
    out = np.ascontiguousarray(rgba, dtype=np.float32).copy()
    h, w = out.shape[:2]
    if lib().oracle_post_process(_p(out), w, h, int(bool(chroma_blur)), int(bool(lens_flare)),
                                 float(chromatic_aberration)) != 0:
        raise ValueError("empty image")
    return out
