"""The C-ABI library loads and exports every entry point include/hdrgpu.h declares, and the
ctypes mirror of its structs matches the C layout. CPU only (no compute calls)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from hdrestir_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hdrgpu.h")


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:hdr_status|const char \*|uint32_t)\s*(hdr_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "hdr_render_frame" in names and "hdr_scene_upload" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from hdrestir_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libhdrgpu.so not built")
    lib = C.CDLL(gpu.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.hdr_abi_version.restype = C.c_uint32
    assert lib.hdr_abi_version() == 3


STRUCTS = {
    "hdr_mesh_desc": abi.MeshDesc, "hdr_material_desc": abi.MaterialDesc, "hdr_instance_desc": abi.InstanceDesc,
    "hdr_light_desc": abi.LightDesc, "hdr_scene_desc": abi.SceneDesc, "hdr_camera_params": abi.CameraParams,
    "hdr_frame_desc": abi.FrameDesc, "hdr_pipeline_desc": abi.PipelineDesc, "hdr_reservoir": abi.Reservoir,
    "hdr_pixel_debug": abi.PixelDebug, "hdr_frame_counters": abi.FrameCounters,
    "hdr_server_desc": abi.ServerDesc, "hdr_request_desc": abi.RequestDesc, "hdr_request_result": abi.RequestResult,
    "hdr_server_stats": abi.ServerStats, "hdr_launch_config": abi.LaunchConfig,
    "hdr_variance_stats": abi.VarianceStats, "hdr_post_desc": abi.PostDesc,
}


def test_ctypes_mirror_matches_c_layout():
    prog = "#include <stdio.h>\n#include \"hdrgpu.h\"\nint main(void){\n"
    for name in STRUCTS:
        prog += f'printf("{name} %zu\\n", sizeof({name}));\n'
    prog += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "sz.c")
        exe = os.path.join(d, "sz")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    sizes = dict(line.split() for line in out.strip().splitlines())
    for name, cls in STRUCTS.items():
        assert int(sizes[name]) == C.sizeof(cls), (name, sizes[name], C.sizeof(cls))
    assert C.sizeof(abi.Reservoir) == 80


def test_reservoir_and_debug_dtypes_match_structs():
    assert abi.reservoir_dtype().itemsize == C.sizeof(abi.Reservoir)
    assert abi.debug_dtype().itemsize == C.sizeof(abi.PixelDebug)


def test_product_path_does_not_reference_the_oracle():
    """libhdrgpu.so and the package must never load the CPU checker."""
    pkg = os.path.join(ROOT, "hdrestir_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower().replace("oracles", ""), f


def test_cpp_adapter_reports_missing_device():
    """The C++20 pass adapter turns a failing C-ABI call into hdrgpu::Error (no GPU here)."""
    from hdrestir_amd import gpu
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present; the GPU suite runs the adapter for real")
    if not os.path.exists(gpu.ADAPTER_SMOKE):
        pytest.skip("adapter not built")
    r = subprocess.run([gpu.ADAPTER_SMOKE, "--no-device"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "no device" in r.stdout


def test_rendered_size_is_ceil_of_window_over_two_to_the_level():
    """hdr_rendered_size (host-only): CameraFrame::RenderedWidth/Height, ceil(window / 2^level)
    with the level clamped to 0 .. 4 and at least 1 pixel; non-positive windows are refused."""
    from hdrestir_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libhdrgpu.so not built")
    for w, h in ((1920, 1080), (67, 45), (1, 1), (3, 1000)):
        for level in range(-2, 8):
            div = 2 ** min(max(level, 0), 4)
            assert gpu.rendered_size(w, h, level) == (max(1, -(-w // div)), max(1, -(-h // div))), (w, h, level)
    assert gpu.rendered_size(1920, 1080, 2) == (480, 270) and gpu.rendered_size(1921, 1081, 1) == (961, 541)
    with pytest.raises(gpu.HdrError, match="INVALID_ARGUMENT"):
        gpu.rendered_size(0, 10, 1)
