"""ctypes mirror of include/hdrgpu.h (POD descriptors shared by the library and its tests).

The layouts here must stay byte-identical to the C header; tests/test_abi.py checks the
sizes against the values the compiled library reports.
"""
import ctypes as C

HDR_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "DEVICE", 4: "UNSUPPORTED",
                5: "CANCELLED", 6: "CAPACITY"}

HDR_LIGHT_DISTANT, HDR_LIGHT_DOME, HDR_LIGHT_RECT, HDR_LIGHT_POINT = 0, 1, 2, 3
HDR_DIRECT_MIS, HDR_DIRECT_RIS = 0, 1
HDR_AOV_COLOR, HDR_AOV_DEPTH, HDR_AOV_ALBEDO, HDR_AOV_NORMAL = 0, 1, 2, 3
HDR_FORMAT_FLOAT32, HDR_FORMAT_FLOAT16, HDR_FORMAT_UNORM8, HDR_FORMAT_SNORM8, HDR_FORMAT_INT32 = 0, 1, 2, 3, 4

HDR_RES_HAS_SAMPLE = 0x01
HDR_RES_HAS_CANDIDATE = 0x02
HDR_RES_VIS_TESTED = 0x04
HDR_RES_PDF_SOLID_ANGLE = 0x08
HDR_RES_LIGHT_DELTA = 0x10

F3 = C.c_float * 3
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("points", FP), ("num_points", C.c_int32),
                ("indices", IP), ("num_triangles", C.c_int32),
                ("normals", FP), ("num_normals", C.c_int32),
                ("uvs", FP), ("num_uvs", C.c_int32),
                ("colors", FP), ("num_colors", C.c_int32)]


class MaterialDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("diffuse_color", F3), ("metallic", C.c_float),
                ("roughness", C.c_float), ("specular_color", F3), ("specular", C.c_float), ("opacity", C.c_float),
                ("ior", C.c_float), ("transmission", C.c_float), ("transmission_color", F3),
                ("transmission_depth", C.c_float), ("emission_color", F3), ("emission", C.c_float),
                ("coat", C.c_float), ("coat_color", F3), ("coat_roughness", C.c_float),
                ("coat_ior", C.c_float), ("sheen", C.c_float), ("sheen_color", F3),
                ("sheen_roughness", C.c_float), ("subsurface", C.c_float), ("subsurface_color", F3),
                ("thin_walled", C.c_int32)]


class InstanceDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("transform", C.c_float * 16), ("mesh", C.c_int32),
                ("material", C.c_int32), ("light", C.c_int32), ("_pad", C.c_int32)]


class LightDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("type", C.c_int32), ("color", F3), ("intensity", C.c_float),
                ("exposure", C.c_float), ("transform", C.c_double * 16), ("width", C.c_float), ("height", C.c_float),
                ("shaping_cone_angle", C.c_float), ("shaping_cone_softness", C.c_float),
                ("dome_pixels", FP), ("dome_width", C.c_int32), ("dome_height", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("meshes", C.POINTER(MeshDesc)), ("num_meshes", C.c_int32),
                ("materials", C.POINTER(MaterialDesc)), ("num_materials", C.c_int32),
                ("instances", C.POINTER(InstanceDesc)), ("num_instances", C.c_int32),
                ("lights", C.POINTER(LightDesc)), ("num_lights", C.c_int32),
                ("environment", C.c_int32)]


class CameraParams(C.Structure):
    _fields_ = [("enable_dof", C.c_int32), ("enable_exposure", C.c_int32), ("focal_length", C.c_float),
                ("f_stop", C.c_float), ("focus_distance", C.c_float), ("bokeh_blades", C.c_int32),
                ("lens_distortion", C.c_float), ("iso", C.c_float), ("shutter_speed", C.c_float)]


class FrameDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("view", C.c_double * 16), ("proj", C.c_double * 16),
                ("frame_index", C.c_int32), ("has_camera_params", C.c_int32), ("camera", CameraParams)]


class PipelineDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("direct_light", C.c_int32), ("candidate_count", C.c_int32),
                ("use_reservoir", C.c_int32), ("skip_visibility", C.c_int32), ("max_depth", C.c_int32),
                ("max_reflection_bounces", C.c_int32), ("max_refraction_bounces", C.c_int32),
                ("enable_subsurface", C.c_int32), ("render_ibl_background", C.c_int32),
                ("roulette_aggressiveness", C.c_float), ("enable_firefly_filter", C.c_int32),
                ("spatial_count", C.c_int32), ("spatial_radius", C.c_float),
                ("temporal_reprojection", C.c_int32)]


class Reservoir(C.Structure):
    _fields_ = [("dir", F3), ("dist", C.c_float), ("color", F3), ("pdf", C.c_float),
                ("light_normal", F3), ("target", C.c_float), ("throughput", C.c_float * 4),
                ("wsum", C.c_float), ("count", C.c_int32), ("w", C.c_float), ("flags_light", C.c_uint32)]


class PixelDebug(C.Structure):
    _fields_ = [("draws", C.c_uint32), ("light_hash", C.c_uint32), ("select_hash", C.c_uint32),
                ("n_light_picks", C.c_int32), ("prim_instance", C.c_int32), ("prim_triangle", C.c_int32),
                ("prim_t", C.c_float), ("sample_rgb", F3)]


class FrameCounters(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("bounce_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("candidate_rays", C.c_uint64), ("candidate_rays_skipped", C.c_uint64),
                ("node_visits", C.c_uint64), ("triangle_tests", C.c_uint64),
                ("instance_visits", C.c_uint64), ("rng_draws", C.c_uint64),
                ("primary_node_visits", C.c_uint64), ("primary_triangle_tests", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def pipeline(direct="ris", candidate_count=16, use_reservoir=True, skip_visibility=False, max_depth=4,
             max_refl=8, max_refr=8, enable_subsurface=True, render_ibl_background=True,
             roulette=1.0, firefly=False, spatial_count=0, spatial_radius=30.0, reprojection=False):
    """PathTracePassSettings + RISPipelineSettings as an hdr_pipeline_desc."""
    p = PipelineDesc()
    p.struct_size = C.sizeof(PipelineDesc)
    p.direct_light = HDR_DIRECT_RIS if direct == "ris" else HDR_DIRECT_MIS
    p.candidate_count = candidate_count
    p.use_reservoir = int(use_reservoir)
    p.skip_visibility = int(skip_visibility)
    p.max_depth = max_depth
    p.max_reflection_bounces = max_refl
    p.max_refraction_bounces = max_refr
    p.enable_subsurface = int(enable_subsurface)
    p.render_ibl_background = int(render_ibl_background)
    p.roulette_aggressiveness = roulette
    p.enable_firefly_filter = int(firefly)
    p.spatial_count = spatial_count
    p.spatial_radius = spatial_radius
    p.temporal_reprojection = int(reprojection)
    return p


def frame(view, proj, frame_index):
    f = FrameDesc()
    f.struct_size = C.sizeof(FrameDesc)
    for i in range(16):
        f.view[i] = float(view.flat[i])
        f.proj[i] = float(proj.flat[i])
    f.frame_index = frame_index
    f.has_camera_params = 0
    return f


def debug_dtype():
    import numpy as np
    return np.dtype([("draws", "<u4"), ("light_hash", "<u4"), ("select_hash", "<u4"),
                     ("n_light_picks", "<i4"), ("prim_instance", "<i4"), ("prim_triangle", "<i4"),
                     ("prim_t", "<f4"), ("sample_rgb", "<f4", 3)])


def reservoir_dtype():
    import numpy as np
    return np.dtype([("dir", "<f4", 3), ("dist", "<f4"), ("color", "<f4", 3), ("pdf", "<f4"),
                     ("light_normal", "<f4", 3), ("target", "<f4"), ("throughput", "<f4", 4),
                     ("wsum", "<f4"), ("count", "<i4"), ("w", "<f4"), ("flags_light", "<u4")])



class PostDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("chromaticity_blur", C.c_int32), ("oidn", C.c_int32),
                ("lens_flare", C.c_int32), ("chromatic_aberration", C.c_float)]


def post_desc(chromaticity_blur=False, lens_flare=False, chromatic_aberration=0.0, oidn=False):
    """Denoiser::Config (its chromaticity blur; OIDN is not part of the library) and PostProcess::Config."""
    d = PostDesc()
    d.struct_size = C.sizeof(PostDesc)
    d.chromaticity_blur, d.oidn, d.lens_flare = int(bool(chromaticity_blur)), int(bool(oidn)), int(bool(lens_flare))
    d.chromatic_aberration = float(chromatic_aberration)
    return d


class ServerDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_active", C.c_int32), ("hbm_budget_bytes", C.c_uint64),
                ("batched", C.c_int32), ("own_thread", C.c_int32)]


class LaunchConfig(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("queue_waves", C.c_int32), ("trace_occupancy", C.c_int32),
                ("shade_occupancy", C.c_int32), ("use_graphs", C.c_int32), ("lds_stack", C.c_int32),
                ("_reserved", C.c_int32 * 2)]

    def as_dict(self):
        return {"queue_waves": self.queue_waves, "trace_occupancy": self.trace_occupancy,
                "shade_occupancy": self.shade_occupancy, "use_graphs": self.use_graphs,
                "lds_stack": self.lds_stack}


def launch_config(queue_waves=0, trace_occupancy=0, shade_occupancy=0, use_graphs=1, lds_stack=0):
    c = LaunchConfig()
    c.struct_size = C.sizeof(LaunchConfig)
    c.queue_waves, c.trace_occupancy, c.shade_occupancy, c.use_graphs = (queue_waves, trace_occupancy,
                                                                         shade_occupancy, use_graphs)
    c.lds_stack = lds_stack
    return c


class RequestDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("scene", C.c_void_p), ("frame", FrameDesc), ("pipe", PipelineDesc)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("status", C.c_int32), ("queued_ms", C.c_double), ("latency_ms", C.c_double)]


class VarianceStats(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("sample_count", C.c_int32), ("valid", C.c_int32), ("_pad", C.c_int32),
                ("mean_estimator_variance", C.c_double), ("max_estimator_variance", C.c_double),
                ("mean_luminance", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int32), ("active", C.c_int32), ("finished", C.c_int32),
                ("pooled_views", C.c_int32), ("hbm_in_use", C.c_uint64), ("steps", C.c_uint64)]
