"""Python front end of libhdrgpu.so (the MI355X hot path) over its C-ABI.

Mirrors the reference's pass interface for the per-pixel path: a `Renderer` owns one view
(RaycastPass -> {RIS,}PathTracePass -> AccumulationPass with device-resident reservoirs and
accumulation) over a resident scene. There is no CPU fallback: if the shared library is
missing or no HIP device is present, construction raises.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhdrgpu.so")
CSRC = os.path.join(_HERE, "csrc")
_lib = None

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# the rendering kernels are split over several translation units so they compile in parallel
SOURCES = ["k_wshade.hip", "k_wshade_traced.hip", "k_mega.hip", "k_ptlean.hip", "k_wave0.hip", "k_trace.hip",
           "k_deferred.hip", "k_post.hip", "hdrgpu.hip",
           "host_scene.cpp", "server.cpp"]
OBJ_DIR = os.path.join(_HERE, "_obj")
ADAPTER_SMOKE = os.path.join(_HERE, "host", "build", "adapter_smoke")


def build(verbose=False, jobs=None):
    """Compiles libhdrgpu.so for gfx950 in-tree (hipcc cross-compiles without a GPU): every source
    to an object in parallel, then one link."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or max(1, min(len(SOURCES), os.cpu_count() or 1))
    pending = list(SOURCES)
    running, objs = [], []
    while pending or running:
        while pending and len(running) < jobs:
            s = pending.pop(0)
            obj = os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o")
            objs.append(obj)
            cmd = ["/opt/rocm/bin/hipcc"] + HIPCC_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, s)]
            running.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        s, p = running.pop(0)
        _, err = p.communicate()
        if p.returncode != 0:
            for _, q in running:
                q.kill()
                q.wait()
            raise RuntimeError(f"hipcc failed on {s}:\n" + err[-4000:])
    tmp = LIB_PATH + ".tmp"
    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-shared", "-o", tmp] + objs,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB_PATH)
    # the C++20 pass adapter's smoke program, linked against the library just built
    host = os.path.join(_HERE, "host")
    os.makedirs(os.path.join(host, "build"), exist_ok=True)
    r = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-o", ADAPTER_SMOKE,
                        os.path.join(host, "adapter_smoke.cpp"), "-L" + _HERE, "-lhdrgpu",
                        "-Wl,-rpath,$ORIGIN/../.."], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed on the pass adapter:\n" + r.stderr[-4000:])
    if verbose:
        print("built", LIB_PATH, "and", ADAPTER_SMOKE)
    return LIB_PATH


class HdrError(RuntimeError):
    pass


def lib():
    """Loads libhdrgpu.so and declares the C-ABI signatures (no GPU is touched here)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HdrError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        I = C.c_int32
        L.hdr_last_error.restype = C.c_char_p
        L.hdr_abi_version.restype = C.c_uint32
        sigs = {
            "hdr_context_create": [I, C.POINTER(P)],
            "hdr_context_destroy": [P],
            "hdr_context_set_cancel": [P, I],
            "hdr_scene_upload": [P, C.POINTER(abi.SceneDesc), C.POINTER(P)],
            "hdr_scene_destroy": [P],
            "hdr_scene_update": [P, C.POINTER(abi.SceneDesc)],
            "hdr_scene_device_bytes": [P, C.POINTER(C.c_uint64)],
            "hdr_scene_read_dome": [P, I, P, P, C.POINTER(C.c_float)],
            "hdr_view_create": [P, I, I, C.POINTER(P)],
            "hdr_view_destroy": [P],
            "hdr_view_reset": [P],
            "hdr_view_reset_accumulation": [P],
            "hdr_view_device_bytes": [P, C.POINTER(C.c_uint64)],
            "hdr_render_frame": [P, P, C.POINTER(abi.FrameDesc), C.POINTER(abi.PipelineDesc), P, P],
            "hdr_view_read_color": [P, P],
            "hdr_view_read_reservoirs": [P, P],
            "hdr_view_read_aovs": [P, P, P, P],
            "hdr_view_color_device_ptr": [P, C.POINTER(P)],
            "hdr_rendered_size": [I, I, I, C.POINTER(I), C.POINTER(I)],
            "hdr_view_set_metrics": [P, I],
            "hdr_view_read_variance": [P, C.POINTER(abi.VarianceStats), P, P],
            "hdr_view_sample_count": [P, C.POINTER(I)],
            "hdr_view_resolve": [P, I, I, I, I, I, P],
            "hdr_view_post_process": [P, C.POINTER(abi.PostDesc)],
            "hdr_render_frames": [P, P, C.POINTER(abi.FrameDesc), C.POINTER(abi.PipelineDesc), I],
            "hdr_render_progressive": [P, P, C.POINTER(abi.FrameDesc), C.POINTER(abi.PipelineDesc), I, I,
                                       C.POINTER(I)],
            "hdr_view_synchronize": [P],
            "hdr_view_last_timing": [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)],
            "hdr_view_read_stage_times": [P, P, P],
            "hdr_batch_create": [P, C.POINTER(P)],
            "hdr_batch_destroy": [P],
            "hdr_render_batch": [P, I, P, P, P, P],
            "hdr_batch_synchronize": [P],
            "hdr_batch_last_timing": [P, C.POINTER(C.c_float)],
            "hdr_view_set_launch_config": [P, C.POINTER(abi.LaunchConfig)],
            "hdr_batch_set_launch_config": [P, C.POINTER(abi.LaunchConfig)],
            "hdr_batch_get_launch_config": [P, C.POINTER(abi.LaunchConfig)],
            "hdr_batch_autotune": [P, I, P, P, P, P, I, C.POINTER(abi.LaunchConfig), C.POINTER(C.c_float)],
            "hdr_server_create": [P, C.POINTER(abi.ServerDesc), C.POINTER(P)],
            "hdr_server_destroy": [P],
            "hdr_server_submit": [P, C.POINTER(abi.RequestDesc), C.POINTER(C.c_uint64)],
            "hdr_server_step": [P, C.POINTER(C.c_int32)],
            "hdr_server_peek_result": [P, C.POINTER(abi.RequestResult), C.POINTER(C.c_int32)],
            "hdr_server_take_result": [P, C.c_uint64, P],
            "hdr_server_wait_result": [P, I, C.POINTER(abi.RequestResult), C.POINTER(C.c_int32)],
            "hdr_view_plan_bytes": [P, I, I, C.POINTER(abi.PipelineDesc), C.POINTER(C.c_uint64)],
            "hdr_batch_poll": [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
            "hdr_server_get_stats": [P, C.POINTER(abi.ServerStats)],
            "hdr_scene_intersect": [P, P, P, I, P, P, P],
            "hdr_rng_raw": [P, P, I, I, P],
            "hdr_rng_float": [P, P, I, I, P],
            "hdr_rng_raw_noalloc": [P, P, I, I, P, P],
            "hdr_view_set_counting": [P, I],
            "hdr_view_read_counters": [P, C.POINTER(abi.FrameCounters)],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def check(status):
    if status != abi.HDR_OK:
        msg = lib().hdr_last_error().decode(errors="replace")
        raise HdrError(f"{abi.STATUS_NAMES.get(status, status)}: {msg}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().hdr_context_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().hdr_context_destroy(self.h)
            self.h = None

    def set_cancel(self, flag):
        check(lib().hdr_context_set_cancel(self.h, int(flag)))

    def rng_raw(self, seeds, n_draws):
        seeds = np.ascontiguousarray(seeds, np.uint32)
        out = np.zeros((len(seeds), n_draws), np.uint64)
        check(lib().hdr_rng_raw(self.h, _p(seeds), len(seeds), n_draws, _p(out)))
        return out

    def rng_raw_noalloc(self, seeds_ptr, n, n_draws, out_ptr, stream=0):
        """hdr_rng_raw_noalloc on device pointers (queued on `stream`; the caller synchronises)."""
        check(lib().hdr_rng_raw_noalloc(self.h, C.c_void_p(seeds_ptr), int(n), int(n_draws), C.c_void_p(out_ptr),
                                        C.c_void_p(stream)))

    def rng_float(self, seeds, n_draws):
        seeds = np.ascontiguousarray(seeds, np.uint32)
        out = np.zeros((len(seeds), n_draws), np.float32)
        check(lib().hdr_rng_float(self.h, _p(seeds), len(seeds), n_draws, _p(out)))
        return out


def rendered_size(window_width, window_height, level):
    """hdr_rendered_size: the (width, height) a window renders at for a resolution level."""
    w, h = C.c_int32(), C.c_int32()
    check(lib().hdr_rendered_size(window_width, window_height, level, C.byref(w), C.byref(h)))
    return int(w.value), int(h.value)


class GpuScene:
    """A scene resident in HBM (BVHs, light and material tables)."""

    def __init__(self, ctx, scene):
        h = C.c_void_p()
        check(lib().hdr_scene_upload(ctx.h, C.byref(scene.desc), C.byref(h)))
        self.h, self.ctx, self.scene = h, ctx, scene

    def device_bytes(self):
        v = C.c_uint64()
        check(lib().hdr_scene_device_bytes(self.h, C.byref(v)))
        return int(v.value)

    def update(self, scene):
        """hdr_scene_update: the instances, materials and lights of `scene` (the uploaded scene,
        edited, with the same meshes) replace the resident ones; the BLAS stay."""
        check(lib().hdr_scene_update(self.h, C.byref(scene.desc)))
        self.scene = scene

    def dome_cdf(self, light, w, h):
        """(row_cdf, col_cdf (h, w+1), total luminance) of dome light `light`, as built on the device."""
        row = np.zeros(h + 1, np.float32)
        col = np.zeros((h, w + 1), np.float32)
        tot = C.c_float()
        check(lib().hdr_scene_read_dome(self.h, light, _p(row), _p(col), C.byref(tot)))
        return row, col, float(tot.value)

    def intersect(self, origins, dirs):
        origins = np.ascontiguousarray(origins, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = origins.shape[0]
        inst = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        check(lib().hdr_scene_intersect(self.h, _p(origins), _p(dirs), n, _p(inst), _p(tri), _p(t)))
        return inst, tri, t

    def plan_bytes(self, width, height, pipe):
        """Device bytes a view of width x height holds to render frames of `pipe` on this scene
        (hdr_view_plan_bytes; validates the pipeline as a render would)."""
        b = C.c_uint64()
        check(lib().hdr_view_plan_bytes(self.h, int(width), int(height), C.byref(pipe), C.byref(b)))
        return int(b.value)

    def close(self):
        if self.h:
            lib().hdr_scene_destroy(self.h)
            self.h = None


class Renderer:
    """One render view (== one RenderPipeline instance with its persistent store) on the GPU."""

    def __init__(self, ctx, gscene, width, height):
        h = C.c_void_p()
        check(lib().hdr_view_create(ctx.h, width, height, C.byref(h)))
        self.h, self.ctx, self.gscene = h, ctx, gscene
        self.W, self.H = width, height
        self.color = np.zeros((height, width, 4), np.float32)
        self.debug = np.zeros(height * width, abi.debug_dtype())

    def close(self):
        if self.h:
            lib().hdr_view_destroy(self.h)
            self.h = None

    def reset(self):
        check(lib().hdr_view_reset(self.h))

    def reset_accumulation(self):
        """hdr_view_reset_accumulation: a camera move under reprojected temporal reuse."""
        check(lib().hdr_view_reset_accumulation(self.h))

    def render(self, frame, pipe, read_color=True, debug=False):
        check(lib().hdr_render_frame(self.h, self.gscene.h, C.byref(frame), C.byref(pipe),
                                     _p(self.color) if read_color else None,
                                     _p(self.debug) if debug else None))
        return self.color

    def render_frames(self, frame, pipe, count):
        """hdr_render_frames: frames frame.frame_index .. + count - 1, as many render() calls would."""
        check(lib().hdr_render_frames(self.h, self.gscene.h, C.byref(frame), C.byref(pipe), int(count)))

    def synchronize(self):
        check(lib().hdr_view_synchronize(self.h))

    def last_timing(self):
        """(G-buffer kernel ms, integration kernel ms, whole frame ms) of the last frame."""
        g, a, b = C.c_float(), C.c_float(), C.c_float()
        check(lib().hdr_view_last_timing(self.h, C.byref(g), C.byref(a), C.byref(b)))
        return float(g.value), float(a.value), float(b.value)

    STAGES = ("gbuffer", "integrate", "shadow", "accumulate", "firefly", "spatial")

    def stage_times(self):
        """{stage: (summed device ms, launches)} over the frames since the previous call (HIP events
        on the view's stream around each kernel); clears the sums."""
        ms = np.zeros(len(self.STAGES), np.float64)
        n = np.zeros(len(self.STAGES), np.int32)
        check(lib().hdr_view_read_stage_times(self.h, _p(ms), _p(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.STAGES)}

    def read_color(self):
        check(lib().hdr_view_read_color(self.h, _p(self.color)))
        return self.color

    def reservoirs(self):
        out = np.zeros(self.H * self.W, abi.reservoir_dtype())
        check(lib().hdr_view_read_reservoirs(self.h, _p(out)))
        return out

    def aovs(self):
        d = np.zeros(self.H * self.W, np.float32)
        a = np.zeros((self.H * self.W, 3), np.float32)
        n = np.zeros((self.H * self.W, 3), np.float32)
        check(lib().hdr_view_read_aovs(self.h, _p(d), _p(a), _p(n)))
        return d, a, n

    _FORMAT_DTYPES = {abi.HDR_FORMAT_FLOAT32: np.float32, abi.HDR_FORMAT_FLOAT16: np.float16,
                      abi.HDR_FORMAT_UNORM8: np.uint8, abi.HDR_FORMAT_SNORM8: np.int8,
                      abi.HDR_FORMAT_INT32: np.int32}

    def resolve(self, aov, window_width, window_height, fmt=abi.HDR_FORMAT_FLOAT32, components=4):
        """hdr_view_resolve: output `aov` upscaled to the window and converted to `components` of
        `fmt`; an array (window_height, window_width, components)."""
        out = np.zeros((window_height, window_width, components), self._FORMAT_DTYPES.get(fmt, np.float32))
        check(lib().hdr_view_resolve(self.h, aov, window_width, window_height, fmt, components, _p(out)))
        return out

    def post_process(self, chromaticity_blur=False, lens_flare=False, chromatic_aberration=0.0, oidn=False):
        """hdr_view_post_process: DenoiserPass's chromaticity blur, then PostProcessPass's lens flare
        and chromatic aberration, on the view's Color in place (read it with read_color)."""
        d = abi.post_desc(chromaticity_blur, lens_flare, chromatic_aberration, oidn)
        check(lib().hdr_view_post_process(self.h, C.byref(d)))

    def set_metrics(self, on):
        """hdr_view_set_metrics: per-pixel luminance sums of the accumulated samples (Variance)."""
        check(lib().hdr_view_set_metrics(self.h, int(on)))

    def variance(self):
        """hdr_view_read_variance: ({sample_count, valid, mean_estimator_variance,
        max_estimator_variance, mean_luminance}, per-pixel luminance sums, per-pixel sums of squares)."""
        st = abi.VarianceStats()
        st.struct_size = C.sizeof(abi.VarianceStats)
        ls = np.zeros(self.H * self.W, np.float64)
        lq = np.zeros(self.H * self.W, np.float64)
        check(lib().hdr_view_read_variance(self.h, C.byref(st), _p(ls), _p(lq)))
        stats = {k: getattr(st, k) for k, _ in abi.VarianceStats._fields_ if k not in ("struct_size", "_pad")}
        return stats, ls, lq

    def sample_count(self):
        v = C.c_int32()
        check(lib().hdr_view_sample_count(self.h, C.byref(v)))
        return int(v.value)

    def render_progressive(self, frame, pipe, target_samples, max_frames):
        """hdr_render_progressive: renders `frame` with frame_index = the sample count until the view
        holds target_samples samples (at most max_frames frames); returns whether it converged."""
        conv = C.c_int32()
        check(lib().hdr_render_progressive(self.h, self.gscene.h, C.byref(frame), C.byref(pipe), target_samples,
                                           max_frames, C.byref(conv)))
        return bool(conv.value)

    def set_counting(self, on):
        check(lib().hdr_view_set_counting(self.h, int(on)))

    def set_launch_config(self, cfg):
        check(lib().hdr_view_set_launch_config(self.h, C.byref(cfg)))

    def counters(self):
        c = abi.FrameCounters()
        check(lib().hdr_view_read_counters(self.h, C.byref(c)))
        return c.as_dict()

    def device_bytes(self):
        v = C.c_uint64()
        check(lib().hdr_view_device_bytes(self.h, C.byref(v)))
        return int(v.value)


class Batch:
    """One frame for each of many views in merged launches (hdr_render_batch)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().hdr_batch_create(ctx.h, C.byref(self.h)))

    @staticmethod
    def _arrays(renderers, frames, pipes):
        n = len(renderers)
        return (n, (C.c_void_p * n)(*[r.h for r in renderers]), (C.c_void_p * n)(*[r.gscene.h for r in renderers]),
                (abi.FrameDesc * n)(*frames), (abi.PipelineDesc * n)(*pipes))

    def render(self, renderers, frames, pipes):
        check(lib().hdr_render_batch(self.h, *self._arrays(renderers, frames, pipes)))

    def set_launch_config(self, cfg):
        check(lib().hdr_batch_set_launch_config(self.h, C.byref(cfg)))

    def launch_config(self):
        c = abi.LaunchConfig()
        check(lib().hdr_batch_get_launch_config(self.h, C.byref(c)))
        return c.as_dict()

    def autotune(self, renderers, frames, pipes, reps=2):
        """Times candidate launch shapes on these views and keeps the fastest; returns
        (chosen config dict, its step ms). The views' frames advance (reset them afterwards)."""
        c = abi.LaunchConfig()
        ms = C.c_float()
        check(lib().hdr_batch_autotune(self.h, *self._arrays(renderers, frames, pipes), reps, C.byref(c),
                                       C.byref(ms)))
        return c.as_dict(), float(ms.value)

    def synchronize(self):
        check(lib().hdr_batch_synchronize(self.h))

    def last_timing(self):
        ms = C.c_float()
        check(lib().hdr_batch_last_timing(self.h, C.byref(ms)))
        return float(ms.value)

    def close(self):
        if self.h:
            lib().hdr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Render-request server (hdr_server_*): continuous batching (batched=True) or the plain
    per-request path, with admission by device-memory footprint."""

    def __init__(self, ctx, batched=True, max_active=64, hbm_budget_bytes=64 << 30, own_thread=False):
        self.ctx = ctx
        d = abi.ServerDesc()
        d.struct_size = C.sizeof(abi.ServerDesc)
        d.max_active = max_active
        d.hbm_budget_bytes = hbm_budget_bytes
        d.batched = int(batched)
        d.own_thread = int(own_thread)
        self.h = C.c_void_p()
        check(lib().hdr_server_create(ctx.h, C.byref(d), C.byref(self.h)))

    def submit(self, gscene, width, height, spp, frame, pipe):
        r = abi.RequestDesc()
        r.struct_size = C.sizeof(abi.RequestDesc)
        r.width, r.height, r.spp = width, height, spp
        r.scene = gscene.h
        r.frame = frame
        r.pipe = pipe
        rid = C.c_uint64()
        check(lib().hdr_server_submit(self.h, C.byref(r), C.byref(rid)))
        return int(rid.value)

    def step(self):
        n = C.c_int32()
        check(lib().hdr_server_step(self.h, C.byref(n)))
        return int(n.value)

    def results(self, images=False, wait_ms=None):
        """Takes every finished request: list of (result dict, image or None). wait_ms: with the
        server's own thread, wait up to that long for the first result."""
        out = []
        while True:
            res = abi.RequestResult()
            avail = C.c_int32()
            if wait_ms is not None and not out:
                check(lib().hdr_server_wait_result(self.h, int(wait_ms), C.byref(res), C.byref(avail)))
            else:
                check(lib().hdr_server_peek_result(self.h, C.byref(res), C.byref(avail)))
            if not avail.value:
                return out
            img = None
            if images and res.status == abi.HDR_OK:
                img = np.zeros((res.height, res.width, 4), np.float32)
            check(lib().hdr_server_take_result(self.h, C.c_uint64(res.id), _p(img) if img is not None else None))
            out.append(({"id": int(res.id), "width": res.width, "height": res.height, "spp": res.spp,
                         "status": int(res.status), "queued_ms": res.queued_ms, "latency_ms": res.latency_ms},
                        img))

    def stats(self):
        st = abi.ServerStats()
        check(lib().hdr_server_get_stats(self.h, C.byref(st)))
        return {k: int(getattr(st, k)) for k, _ in st._fields_}

    def close(self):
        if self.h:
            lib().hdr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
