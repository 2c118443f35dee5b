"""restir_amd: the ReSTIR GI/DI hot path on MI355X (gfx950), behind a C ABI.

This package is the Python side of the C ABI declared in include/restir_amd.h
and include/restir_amd_scene.h: ctypes mirrors of the ABI structs, a loader
for the in-tree librestir_amd.so (it fails loudly when the library is
missing; there is no fallback path), the host-side scene helpers, and
`Settings`/`make_frame`, which pack the per-frame uniforms the way the
reference renderer does (src/renderer.cpp:497-531, 614-621, 632-640,
693-731; defaults from src/settings.h:183-227).
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librestir_amd.so")

RGI_OK = 0
RGI_E_INVALID, RGI_E_DEVICE, RGI_E_NOMEM = -1, -2, -3
RGI_RAY_SCREEN, RGI_RAY_HYBRID, RGI_RAY_BVH = 0, 1, 2
RGI_GRID_ATOMIC, RGI_GRID_DETERMINISTIC = 0, 1
RGI_BATCH_PASSES, RGI_BATCH_GRAPH = 0, 1
RGI_SCENE_CORNELL, RGI_SCENE_ATRIUM, RGI_SCENE_BOXFIELD = 0, 1, 2
RGI_OPT_GI_TEMPORAL_WAVES, RGI_OPT_GI_SPATIAL_WAVES, RGI_OPT_TRACE_PERSISTENT = 1, 2, 3
RGI_OPT_TILE_ORDER = 4
RGI_OPT_TRIANGLE_TEST = 5
RGI_OPT_TILE_SHAPE = 6
RGI_TUNE_TRIALS = 8
RGI_TUNE_TEMPORAL_WAVES4, RGI_TUNE_SPATIAL_WAVES5, RGI_TUNE_TILE_ORDER, RGI_TUNE_TILE_SHAPE = 1, 2, 4, 8


def tuning_of(config):
    """The context options of an autotuner configuration bit set."""
    return {
        RGI_OPT_GI_TEMPORAL_WAVES: 4 if config & RGI_TUNE_TEMPORAL_WAVES4 else 0,
        RGI_OPT_GI_SPATIAL_WAVES: 5 if config & RGI_TUNE_SPATIAL_WAVES5 else 0,
        RGI_OPT_TILE_ORDER: 2 if config & RGI_TUNE_TILE_ORDER else 0,
        RGI_OPT_TILE_SHAPE: 1 if config & RGI_TUNE_TILE_SHAPE else 0,
    }

F, I32, U32, P = ctypes.c_float, ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p


class Frame(ctypes.Structure):
    """rgi_frame (include/restir_amd.h)."""
    _fields_ = [
        ("inv_view_proj", F * 16), ("view_proj", F * 16), ("prev_view_proj", F * 16),
        ("camera_position", F * 4), ("prev_camera_position", F * 4),
        ("light_position", F * 4), ("light_color", F * 4),
        ("frame_index", U32), ("history_valid", I32), ("origin_bottom_left", I32),
        ("multiple_bounces", I32), ("homogeneous_depth", I32),
        ("gi_width", I32), ("gi_height", I32), ("width", I32), ("height", I32),
        ("scene_scale", F), ("boiling_filter", F), ("jacobian_clamp", F), ("outlier_clamp", F),
        ("lighting_surface", F), ("lighting_emission", F),
        ("initial_candidates", I32), ("temporal_taps", I32), ("spatial_taps", I32), ("trace_steps", I32),
        ("grid_min_cell", F), ("grid_pixel_size", F), ("grid_cell_pixels", F), ("world_space_reuse", I32),
        ("grid_table_size", U32), ("ray_mode", I32),
        ("sky_zenith", F * 4), ("sky_horizon", F * 4), ("sky_ground", F * 4),
        ("shadow_view_proj", F * 96), ("shadow_info", F * 4), ("shadows_enabled", I32),
        ("shadow_width", I32), ("shadow_height", I32),
        ("di_light_count", I32), ("di_candidates", I32), ("di_spatial_taps", I32), ("di_history_valid", I32),
        ("reserved", I32 * 3),
    ]


class Scene(ctypes.Structure):
    _fields_ = [("bvh", P), ("node_count", U32), ("triangle_offset", U32), ("triangle_count", U32),
                ("light_count", U32), ("lights", P)]


class GBuffer(ctypes.Structure):
    _fields_ = [("width", I32), ("height", I32), ("depth", P), ("albedo", P), ("normal", P),
                ("shading_normal", P), ("emission", P)]


class GiReservoirs(ctypes.Structure):
    _fields_ = [("sample", P), ("receiver", P), ("normal", P)]


class GiInputs(ctypes.Structure):
    _fields_ = [("history_sample", P), ("history_receiver", P), ("history_normal", P), ("history_color", P),
                ("history_moments", P), ("surface_light", P), ("shadow_map", P), ("grid", P)]


class DiReservoirs(ctypes.Structure):
    _fields_ = [("sample", P), ("receiver", P), ("normal", P)]


class CSettings(ctypes.Structure):
    """rgi_settings (include/restir_amd_renderer.h)."""
    _fields_ = [
        ("initial_candidates", I32), ("temporal_taps", I32), ("spatial_taps", I32), ("trace_steps", I32),
        ("di_candidates", I32), ("di_spatial_taps", I32), ("ray_mode", I32), ("world_space_reuse", I32),
        ("multiple_bounces", I32), ("clamp_jacobian", I32), ("clamp_outliers", I32), ("boiling_filter", I32),
        ("restir_di", I32), ("emissive_lighting", I32), ("sky", I32), ("shadows", I32), ("grid_mode", I32),
        ("grid_table_size", U32), ("sky_zenith", F * 3), ("sky_horizon", F * 3), ("sky_ground", F * 3),
        ("sky_intensity", F), ("light_position", F * 3), ("light_color", F * 3), ("origin_bottom_left", I32),
        ("homogeneous_depth", I32), ("shadow_face_size", I32), ("denoiser", I32), ("atrous_iterations", I32),
        ("composite", I32), ("specular_gi", I32), ("gi_divisor", I32), ("reserved", I32 * 2),
    ]


class View(ctypes.Structure):
    """rgi_view: eye, view-projection and its inverse."""
    _fields_ = [("eye", F * 4), ("view_proj", F * 16), ("inv_view_proj", F * 16)]


class RendererOutputs(ctypes.Structure):
    _fields_ = [("gi_sample", P), ("gi_receiver", P), ("gi_normal", P), ("color", P), ("moments", P), ("grid", P),
                ("di_diffuse", P), ("di_specular", P), ("filtered_history", P), ("denoised", P), ("hdr", P),
                ("specular_gi", P)]


class Shadow(ctypes.Structure):
    """rgi_shadow: the point light's six face matrices, their inverses, u_shadowInfo."""
    _fields_ = [("view_proj", F * 96), ("inv_view_proj", F * 96), ("info", F * 4), ("light_position", F * 4),
                ("near_plane", F), ("face_size", I32), ("homogeneous_depth", I32), ("reserved", I32)]


class CompositeInputs(ctypes.Structure):
    """rgi_composite_inputs: the composite pass's planes and u_compositeOptions."""
    _fields_ = [("gi", P), ("di_diffuse", P), ("di_specular", P), ("specular", P), ("shadow_map", P),
                ("resolved_diffuse", P), ("options", F * 4), ("specular_weight", F), ("reserved", F * 3)]


class Mesh(ctypes.Structure):
    _fields_ = [("triangle_count", U32), ("positions", P), ("albedo", P), ("emission", P), ("material", P),
                ("material_count", U32), ("materials", P)]


class NrcConfig(ctypes.Structure):
    """rgi_nrc_config (include/restir_amd_nrc.h)."""
    _fields_ = [("log2_table", U32), ("base_resolution", U32), ("max_resolution", U32), ("seed", U32),
                ("learning_rate", F), ("beta1", F), ("beta2", F), ("epsilon", F),
                ("aabb_min", F * 3), ("aabb_max", F * 3), ("ema_decay", F), ("reserved", I32 * 3)]


class NrcSettings(ctypes.Structure):
    """rgi_nrc_settings (include/restir_amd_nrc.h): the cache in the frame driver."""
    _fields_ = [("cache", NrcConfig), ("batch", U32), ("warmup", U32), ("precision", I32), ("resolve", I32),
                ("records", I32), ("reserved", I32 * 3)]


RGI_NRC_BF16, RGI_NRC_FP16, RGI_NRC_FP8 = 0, 1, 2
RGI_NRC_RECORDS_PIXELS, RGI_NRC_RECORDS_VERTICES, RGI_NRC_RECORDS_MIXED = 0, 1, 2
RGI_NRC_MLP_PARAMS = 5 * 64 * 64 + 3 * 64
RGI_NRC_LEVELS = 16


ABI_STRUCTS = [Frame, Scene, GBuffer, GiReservoirs, GiInputs, DiReservoirs, Shadow, CompositeInputs, NrcConfig,
               NrcSettings]


class RestirError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded librestir_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RestirError("librestir_amd.so is not built (%s): run python -m restir_amd.build" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.rgi_last_error.restype = ctypes.c_char_p
    L.rgi_abi_sizeof.restype = U32
    L.rgi_abi_sizeof.argtypes = [ctypes.c_int]
    L.rgi_create.argtypes = [ctypes.POINTER(P)]
    L.rgi_destroy.argtypes = [P]
    L.rgi_destroy.restype = None
    L.rgi_grid_clear.argtypes = [P, U32, P]
    L.rgi_grid_insert.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(GBuffer), P, ctypes.c_int, P]
    L.rgi_grid_insert_noalloc.argtypes = [ctypes.POINTER(Frame), ctypes.POINTER(GBuffer), P, P]
    L.rgi_gi_temporal.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(Scene), ctypes.POINTER(GBuffer),
                                  ctypes.POINTER(GiInputs), ctypes.POINTER(GiReservoirs), P]
    L.rgi_gi_spatial.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(Scene), ctypes.POINTER(GBuffer),
                                 ctypes.POINTER(GiInputs), ctypes.POINTER(GiReservoirs), P, P, P]
    L.rgi_di_temporal.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(Scene), ctypes.POINTER(GBuffer),
                                  ctypes.POINTER(DiReservoirs), ctypes.POINTER(DiReservoirs), P]
    L.rgi_di_spatial.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(Scene), ctypes.POINTER(GBuffer),
                                 ctypes.POINTER(DiReservoirs), P, P, P]
    L.rgi_trace.argtypes = [P, ctypes.POINTER(Scene), P, U32, P, P, ctypes.c_int, P]
    L.rgi_gbuffer_primary.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(Scene), P, U32, P, P, P, P, P, P]
    L.rgi_specular_gi.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(Scene), ctypes.POINTER(GBuffer),
                                  ctypes.POINTER(GiInputs), P, P, P, P]
    L.rgi_composite.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(GBuffer), ctypes.POINTER(CompositeInputs), P, P]
    L.rgi_svgf_variance.argtypes = [P, ctypes.POINTER(Frame), P, P, P, P, P, P]
    L.rgi_svgf_atrous.argtypes = [P, ctypes.POINTER(Frame), P, P, P, I32, P, P]
    L.rgi_shadow_setup.argtypes = [P, F, I32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Shadow)]
    L.rgi_shadow_apply.argtypes = [ctypes.POINTER(Shadow), ctypes.c_int, ctypes.POINTER(Frame)]
    L.rgi_shadow_atlas.argtypes = [P, ctypes.POINTER(Shadow), ctypes.POINTER(Scene), P, P]
    L.rgi_scene_generate.restype = ctypes.c_int64
    L.rgi_scene_generate.argtypes = [ctypes.c_int, U32, U32, ctypes.POINTER(Mesh), U32]
    L.rgi_mesh_bounds.restype = None
    L.rgi_mesh_bounds.argtypes = [ctypes.POINTER(Mesh), P, P, P, P]
    L.rgi_bvh_capacity.restype = ctypes.c_uint64
    L.rgi_bvh_capacity.argtypes = [U32]
    L.rgi_bvh_build.restype = ctypes.c_int64
    L.rgi_bvh_build.argtypes = [ctypes.POINTER(Mesh), P, P]
    L.rgi_lights_build.restype = ctypes.c_int64
    L.rgi_lights_build.argtypes = [P, U32, U32, I32, F, U32, P, P, F, P, U32]
    L.rgi_camera_matrices.restype = None
    L.rgi_camera_matrices.argtypes = [P, P, F, F, F, F, ctypes.c_int, P, P, P, P]
    L.rgi_settings_default.restype = None
    L.rgi_settings_default.argtypes = [ctypes.POINTER(CSettings)]
    L.rgi_settings_sizeof.restype = U32
    L.rgi_frame_pack.argtypes = [ctypes.POINTER(CSettings), ctypes.POINTER(Scene), F, U32, ctypes.POINTER(View),
                                 ctypes.POINTER(View), I32, I32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Frame)]
    L.rgi_renderer_create.argtypes = [ctypes.POINTER(P), I32, I32, ctypes.POINTER(CSettings), ctypes.POINTER(Scene), F]
    L.rgi_renderer_destroy.restype = None
    L.rgi_renderer_destroy.argtypes = [P]
    L.rgi_renderer_render.argtypes = [P, ctypes.POINTER(View), ctypes.POINTER(GBuffer), P]
    L.rgi_renderer_render_graph.argtypes = [P, ctypes.POINTER(View), ctypes.POINTER(GBuffer), P]
    L.rgi_batch_create.argtypes = [ctypes.POINTER(P), P, U32]
    L.rgi_batch_render.argtypes = [P, P, P, P]
    L.rgi_batch_set_mode.argtypes = [P, ctypes.c_int]
    L.rgi_batch_destroy.argtypes = [P]
    L.rgi_batch_destroy.restype = None
    L.rgi_renderer_set_reference_mode.argtypes = [P, ctypes.c_int]
    L.rgi_renderer_invalidate_history.restype = None
    L.rgi_renderer_invalidate_history.argtypes = [P]
    L.rgi_renderer_outputs_get.argtypes = [P, ctypes.POINTER(RendererOutputs)]
    L.rgi_renderer_last_frame.argtypes = [P, ctypes.POINTER(Frame)]
    L.rgi_context_set_option.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.rgi_context_get_option.argtypes = [P, ctypes.c_int]
    L.rgi_renderer_set_option.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.rgi_renderer_autotune.argtypes = [P, P, P, U32, U32, P, P]
    L.rgi_reference_accumulate.argtypes = [ctypes.POINTER(GBuffer), P, P, P, P, P]
    L.rgi_reference_error.argtypes = [P, ctypes.POINTER(GBuffer), P, P, P, P, U32, P, P]
    L.rgi_reference_accumulate_composite.argtypes = [ctypes.POINTER(GBuffer), P, P, P]
    L.rgi_reference_error_composite.argtypes = [P, ctypes.POINTER(GBuffer), P, P, U32, P, P]
    L.rgi_nrc_config_default.restype = None
    L.rgi_nrc_config_default.argtypes = [ctypes.POINTER(NrcConfig)]
    L.rgi_nrc_level_resolution.argtypes = [ctypes.POINTER(NrcConfig), U32, ctypes.POINTER(U32)]
    L.rgi_nrc_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(NrcConfig)]
    L.rgi_nrc_destroy.restype = None
    L.rgi_nrc_destroy.argtypes = [P]
    L.rgi_nrc_train.argtypes = [P, P, P, U32, P, P]
    L.rgi_nrc_train_batch.argtypes = [P, P, P, P, P, U32, P]
    L.rgi_nrc_infer.argtypes = [P, P, U32, ctypes.c_int, P, P]
    L.rgi_nrc_gradients.argtypes = [P, P, P, U32, P, P, P, P]
    L.rgi_nrc_get_params.argtypes = [P, P, P, P]
    L.rgi_nrc_get_average.argtypes = [P, P, P, P]
    L.rgi_nrc_set_params.argtypes = [P, P, P, P]
    L.rgi_nrc_steps.argtypes = [P, ctypes.POINTER(U32)]
    L.rgi_nrc_query_gbuffer.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(GBuffer), P, ctypes.c_int, P, P]
    L.rgi_nrc_gather_gbuffer.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(GBuffer), P, P, U32, U32, P, P, P]
    L.rgi_nrc_gather_vertices.argtypes = [P, ctypes.POINTER(Frame), ctypes.POINTER(GBuffer), ctypes.POINTER(GiReservoirs),
                                          P, P, U32, U32, P, P, P]
    L.rgi_nrc_settings_default.restype = None
    L.rgi_nrc_settings_default.argtypes = [ctypes.POINTER(NrcSettings)]
    L.rgi_renderer_set_nrc.argtypes = [P, ctypes.POINTER(NrcSettings)]
    L.rgi_renderer_nrc_state.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P),
                                         ctypes.POINTER(U32)]
    if L.rgi_settings_sizeof() != ctypes.sizeof(CSettings):
        raise RestirError("ABI mismatch: rgi_settings")
    for i, cls in enumerate(ABI_STRUCTS):
        size = L.rgi_abi_sizeof(i)
        if size != ctypes.sizeof(cls):
            raise RestirError("ABI mismatch: %s is %d bytes in C, %d in Python" % (cls.__name__, size, ctypes.sizeof(cls)))
    _lib = L
    return L


def shadow_setup(light_position, scene_radius, face_size=512, homogeneous_depth=False, origin_bottom_left=False):
    """rgi_shadow for a light (renderer.cpp:527-564)."""
    sh = Shadow()
    p = np.asarray(light_position, np.float32)[:3].copy()
    check(lib().rgi_shadow_setup(_ptr(p), float(scene_radius), int(face_size), int(homogeneous_depth),
                                 int(origin_bottom_left), ctypes.byref(sh)))
    return sh


def check(status):
    if status != RGI_OK:
        raise RestirError("restir_amd error %d: %s" % (status, lib().rgi_last_error().decode()))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------ scene

class HostScene:
    """A procedural scene on the host: mesh, BVH (bvh.h layout), light list."""

    def __init__(self, kind=RGI_SCENE_CORNELL, seed=1, size=0, extra_lights=0, extra_intensity=0.3, light_seed=1,
                 mesh=None):
        L = lib()
        if mesh is None:
            count = L.rgi_scene_generate(kind, seed, size, None, 0)
            if count < 0:
                raise RestirError("unknown scene kind %d" % kind)
            self.positions = np.zeros((count, 9), np.float32)
            self.albedo = np.zeros((count, 3), np.float32)
            self.emission = np.zeros((count, 3), np.float32)
            self.material = np.zeros(count, np.uint32)
            self.materials = np.zeros((16, 4), np.float32)
            mesh = Mesh(0, self.positions.ctypes.data, self.albedo.ctypes.data, self.emission.ctypes.data,
                        self.material.ctypes.data, 0, self.materials.ctypes.data)
            got = L.rgi_scene_generate(kind, seed, size, ctypes.byref(mesh), count)
            assert got == count
            self.material_count = min(mesh.material_count, 16)
        else:
            positions, albedo, emission, materials = mesh
            count = len(positions)
            self.positions = np.ascontiguousarray(positions, np.float32).reshape(count, 9)
            self.albedo = np.ascontiguousarray(albedo, np.float32).reshape(count, 3)
            self.emission = np.ascontiguousarray(emission, np.float32).reshape(count, 3)
            self.material = np.zeros(count, np.uint32)
            self.materials = np.zeros((16, 4), np.float32)
            self.materials[0, :len(materials)] = materials
            self.material_count = 1
            mesh = Mesh(count, self.positions.ctypes.data, self.albedo.ctypes.data, self.emission.ctypes.data,
                        self.material.ctypes.data, 1, self.materials.ctypes.data)
        self.triangle_count = int(count)
        self.bounds_min = np.zeros(3, np.float32)
        self.bounds_max = np.zeros(3, np.float32)
        self.center = np.zeros(3, np.float32)
        radius = ctypes.c_float()
        L.rgi_mesh_bounds(ctypes.byref(mesh), _ptr(self.bounds_min), _ptr(self.bounds_max), _ptr(self.center),
                          ctypes.byref(radius))
        self.radius = float(radius.value)
        cap = L.rgi_bvh_capacity(count)
        buf = np.zeros(cap, np.float32)
        depth = ctypes.c_uint32()
        nodes = L.rgi_bvh_build(ctypes.byref(mesh), _ptr(buf), ctypes.byref(depth))
        if nodes < 0:
            raise RestirError("BVH build failed")
        self.node_count = int(nodes)
        self.triangle_offset = 2 * self.node_count
        self.bvh_depth = int(depth.value)
        self.bvh = buf[: (2 * self.node_count + 4 * count) * 4].copy()
        # buildLightList's scale of an intensity of 1: the squared scene scale (renderer.cpp:497-503).
        self.scene_scale = float(np.float32(self.radius) / np.float32(1.75))
        nl = L.rgi_lights_build(_ptr(self.bvh), self.node_count, count, extra_lights, extra_intensity, light_seed,
                                _ptr(self.bounds_min), _ptr(self.bounds_max), self._light_scale(), None, 0)
        self.lights = np.zeros((max(nl, 1), 16), np.float32)
        L.rgi_lights_build(_ptr(self.bvh), self.node_count, count, extra_lights, extra_intensity, light_seed,
                           _ptr(self.bounds_min), _ptr(self.bounds_max), self._light_scale(),
                           _ptr(self.lights), max(nl, 1))
        self.light_count = int(nl)
        self.lights = self.lights[: max(nl, 1)]

    @classmethod
    def from_mesh(cls, positions, albedo, emission, material=(1.0, 0.0, 0.0, 0.0), extra_lights=0,
                  extra_intensity=0.3, light_seed=1):
        """A scene of given triangles (n x 9 vertex coordinates), per-triangle
        albedo and emission (n x 3), and one material (roughness, metallic)
        for all of them, through the same BVH and light-list builders."""
        return cls(extra_lights=extra_lights, extra_intensity=extra_intensity, light_seed=light_seed,
                   mesh=(positions, albedo, emission, np.asarray(material, np.float32)))

    def _light_scale(self):
        s = np.float32(self.scene_scale)
        return float(s * s)

    def rgi_scene(self, bvh_ptr, lights_ptr):
        return Scene(bvh_ptr, self.node_count, self.triangle_offset, self.triangle_count, self.light_count, lights_ptr)


def camera_matrices(eye, target, fovy=48.0, aspect=1.0, near=0.05, far=10.0, homogeneous_depth=False):
    L = lib()
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    view, proj, vp, inv = (np.zeros(16, np.float32) for _ in range(4))
    L.rgi_camera_matrices(_ptr(e), _ptr(t), fovy, aspect, near, far, int(homogeneous_depth), _ptr(view), _ptr(proj),
                          _ptr(vp), _ptr(inv))
    return view, proj, vp, inv


def swing_yaw(frame, period=16, yaw=0.64, amplitude=0.32):
    """Yaw of the benchmark's camera path: from the reference's starting yaw
    (main.cpp:61) towards 0 and back over `period` frames, so the Cornell
    box's open front stays in view (at most 7 degrees per frame)."""
    return yaw - amplitude * (1.0 - math.cos(2.0 * math.pi * (frame % period) / period))


def orbit_camera(scene, frame=0, width=320, height=240, yaw=0.64, pitch=0.255, distance=3.17, yaw_step=0.0,
                 homogeneous_depth=False, yaw_of=None):
    """The reference's starting framing (main.cpp:60-62, camera.cpp:112-122): `distance`
    scene radii from the centre at (yaw, pitch), fovy 48; yaw advances by yaw_step per
    frame, or follows yaw_of(frame) when given."""
    y = yaw_of(frame) if yaw_of is not None else yaw + yaw_step * frame
    r = distance * scene.radius
    c = scene.center
    eye = np.array([c[0] + r * math.cos(pitch) * math.sin(y), c[1] + r * math.sin(pitch),
                    c[2] + r * math.cos(pitch) * math.cos(y)], np.float32)
    far = max(float(np.linalg.norm(eye - c)) + 2.0 * scene.radius, 2.5 * scene.radius)
    view, proj, vp, inv = camera_matrices(eye, c, 48.0, width / height, 0.05 * scene.radius, far, homogeneous_depth)
    return eye, vp, inv


# --------------------------------------------------------------- settings

class Settings:
    """The hot path's part of the reference's Settings (settings.h:183-227)."""

    def __init__(self, **kw):
        self.initial_candidates = 2
        self.temporal_taps = 2
        self.spatial_taps = 4
        self.trace_steps = 32
        self.di_candidates = 8
        self.di_spatial_taps = 3
        self.ray_mode = RGI_RAY_HYBRID
        self.world_space_reuse = True
        self.multiple_bounces = True
        self.clamp_jacobian = True
        self.clamp_outliers = True
        self.boiling_filter = True
        self.restir_di = True          # DirectLighting::ReStirDi (else GiOnly)
        self.emissive_lighting = True
        self.sky = False               # off for the Cornell box (main.cpp:80)
        self.sky_zenith = (0.30, 0.50, 0.95)
        self.sky_horizon = (0.85, 0.88, 0.92)
        self.sky_ground = (0.32, 0.29, 0.26)
        self.sky_intensity = 1.0
        self.light_color = (0.0, 0.0, 0.0)  # main point light off
        self.light_position = (0.0, 1.7, 0.0)
        self.shadows = False
        self.shadow_face_size = 1      # index into 256, 512, 1024, 2048
        self.denoiser = True           # SVGF variance + a-trous (settings.h:190)
        self.atrous_iterations = 5     # kMaxAtrousIterations
        self.composite = True          # fs_composite.sc into an HDR image
        self.specular_gi = True        # fs_specular.sc (settings.h:196)
        self.grid_mode = RGI_GRID_ATOMIC
        self.grid_table_size = 65536
        # GI at the window size over this divisor (kGiResolutions,
        # settings.h:18-23). The reference's default is 2 (Half) and so is
        # rgi_settings_default's; the benchmark configurations give the GI
        # size as the G-buffer size, so the default here is 1.
        self.gi_divisor = 1
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)

    def reference_mode(self):
        """Effective settings while the reference accumulates (renderer.cpp:437-452)."""
        s = Settings(**self.__dict__)
        s.temporal_taps = 0
        s.spatial_taps = 0
        s.clamp_jacobian = False
        s.clamp_outliers = False
        s.boiling_filter = False
        s.di_spatial_taps = 0
        return s


def c_settings(settings):
    """rgi_settings from a Settings."""
    c = CSettings()
    lib().rgi_settings_default(ctypes.byref(c))
    s = settings
    for name in ("initial_candidates", "temporal_taps", "spatial_taps", "trace_steps", "di_candidates",
                 "di_spatial_taps", "ray_mode"):
        setattr(c, name, int(getattr(s, name)))
    for name in ("world_space_reuse", "multiple_bounces", "clamp_jacobian", "clamp_outliers", "boiling_filter",
                 "restir_di", "emissive_lighting", "sky", "shadows"):
        setattr(c, name, int(bool(getattr(s, name))))
    c.grid_mode = int(s.grid_mode)
    c.grid_table_size = int(s.grid_table_size)
    c.sky_zenith[:] = list(s.sky_zenith)
    c.sky_horizon[:] = list(s.sky_horizon)
    c.sky_ground[:] = list(s.sky_ground)
    c.sky_intensity = s.sky_intensity
    c.light_position[:] = list(s.light_position)
    c.light_color[:] = list(s.light_color)
    c.shadow_face_size = int(s.shadow_face_size)
    c.denoiser = int(bool(s.denoiser))
    c.atrous_iterations = int(s.atrous_iterations)
    c.composite = int(bool(s.composite))
    c.specular_gi = int(bool(s.specular_gi))
    c.gi_divisor = int(s.gi_divisor)
    return c


def make_view(eye, vp, inv_vp):
    v = View()
    v.eye[:] = [float(eye[0]), float(eye[1]), float(eye[2]), 1.0]
    v.view_proj[:] = [float(x) for x in vp]
    v.inv_view_proj[:] = [float(x) for x in inv_vp]
    return v


def make_frame(settings, scene, frame_index, eye, vp, inv_vp, prev_eye, prev_vp, width, height, history_valid,
               di_history_valid=False):
    """Packs rgi_frame as Renderer::render packs its uniforms (rgi_frame_pack)."""
    f = Frame()
    view = make_view(eye, vp, inv_vp)
    prev = make_view(prev_eye, prev_vp, inv_vp)
    sc = Scene(None, scene.node_count, scene.triangle_offset, scene.triangle_count, scene.light_count, None)
    check(lib().rgi_frame_pack(ctypes.byref(c_settings(settings)), ctypes.byref(sc), scene.radius, frame_index,
                               ctypes.byref(view), ctypes.byref(prev), width, height, int(history_valid),
                               int(di_history_valid), ctypes.byref(f)))
    return f


def frame_bytes(frame):
    return bytes(frame)
