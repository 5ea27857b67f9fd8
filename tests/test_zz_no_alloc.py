"""hdr_rng_raw_noalloc (the allocation-free variant of hdr_rng_raw, spill slots in the library's
static device scratch) against hdr_rng_raw: bit-exact on two calls in a row, the second smaller,
both past output 311 (the scratch in use) and past 624 on the first; plus the capacity check."""
import ctypes as C

import numpy as np
import pytest

from hdrestir_amd import gpu

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libhdrgpu.so is linked against (already loaded by the context)."""
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


def _noalloc(ctx, seeds, n_draws):
    """The caller's device buffers (allocated here, by the test), the entry point on the null stream."""
    hip = _hip()
    d_seeds, d_out = C.c_void_p(), C.c_void_p()
    out = np.zeros((len(seeds), n_draws), np.uint64)
    assert hip.hipMalloc(C.byref(d_seeds), seeds.nbytes) == 0
    assert hip.hipMalloc(C.byref(d_out), out.nbytes) == 0
    try:
        assert hip.hipMemcpy(d_seeds, seeds.ctypes.data, seeds.nbytes, 1) == 0  # host -> device
        ctx.rng_raw_noalloc(d_seeds.value, len(seeds), n_draws, d_out.value, 0)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data, d_out, out.nbytes, 2) == 0  # device -> host
    finally:
        hip.hipFree(d_seeds)
        hip.hipFree(d_out)
    return out


def test_noalloc_matches_rng_raw_on_two_calls(gpu_ctx):
    rs = np.random.default_rng(11)
    cases = [(5000, 700), (1200, 400)]  # second call smaller; both use the scratch
    for n, nd in cases:
        seeds = rs.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        print(f"noalloc case n={n} n_draws={nd}", flush=True)
        got = _noalloc(gpu_ctx, seeds, nd)
        ref = gpu_ctx.rng_raw(seeds, nd)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), int((got != ref).sum())


def test_noalloc_without_spill_and_capacity_error(gpu_ctx):
    rs = np.random.default_rng(12)
    seeds = rs.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32)
    got = _noalloc(gpu_ctx, seeds, 64)  # below output 312: no scratch, any n
    assert np.array_equal(got, gpu_ctx.rng_raw(seeds, 64))
    with pytest.raises(gpu.HdrError, match="CAPACITY"):
        gpu_ctx.rng_raw_noalloc(1, 13444, 400, 1, 0)  # rejected on the host before any launch
