"""The view's device status word: a frame whose pixels draw past the random numbers their streams
were set up for is reported as HDR_ERR_CAPACITY at the view's next synchronisation, instead of
returning silently wrong images. The condition is forced from the host with the library's test hook
HDRGPU_TEST_WITHHOLD_SPILL=1 (the spill slots a frame needs are withheld; the kernels' draws past
the missing slot return 0 and touch no memory)."""
import numpy as np
import pytest

from hdrestir_amd import abi, scenes

pytestmark = pytest.mark.gpu


def test_rng_overrun_is_reported(gpu_ctx, monkeypatch):
    from hdrestir_amd import gpu
    sc = scenes.box_spheres(64, 64, frames=1)
    # the wavefront path at depth 8 with M = 32: deep streams reach the third twist (output 624 on)
    pipe = abi.pipeline("ris", candidate_count=32, use_reservoir=True, skip_visibility=True, max_depth=8)
    gs = gpu.GpuScene(gpu_ctx, sc)
    r = gpu.Renderer(gpu_ctx, gs, sc.width, sc.height)
    try:
        r.render(sc.frame(0), pipe, debug=True)
        print("pixels past output 624:", int((r.debug["draws"] > 624).sum()))
        assert (r.debug["draws"] > 624).any()
        monkeypatch.setenv("HDRGPU_TEST_WITHHOLD_SPILL", "1")
        r.reset()
        with pytest.raises(gpu.HdrError, match="CAPACITY"):
            r.render(sc.frame(0), pipe)
        with pytest.raises(gpu.HdrError, match="CAPACITY"):  # sticky until the view is reset
            r.synchronize()
        monkeypatch.delenv("HDRGPU_TEST_WITHHOLD_SPILL")
        r.reset()
        img = r.render(sc.frame(0), pipe)
        assert np.isfinite(img).all()
    finally:
        r.close()
        gs.close()
