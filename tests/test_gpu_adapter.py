"""The C++20 pass adapter (hdrestir_amd/host/gpu_pass.h) on the GPU: RIS and PathTracer frames of a
lit box agree in mean radiance within 10 %, ClearPersistentBuffers works, a cancelled context
refuses the next frame with HDR_ERR_CANCELLED."""
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_cpp_adapter_smoke():
    from hdrestir_amd import gpu
    r = subprocess.run([gpu.ADAPTER_SMOKE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "adapter smoke ok" in r.stdout
