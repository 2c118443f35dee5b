"""Shared set-up of the tests: seeded scenes and cameras, and one frame of the
hot path run either by the oracle (host arrays) or by the HIP library (device
tensors), with the pass order and ping-pong of Renderer::render
(src/renderer.cpp:626-833): DI temporal -> DI spatial -> grid clear -> grid
insert -> GI temporal -> GI spatial.
"""
import ctypes
import time

import numpy as np

import restir_amd as ra

import orc

GI_PLANES_F32 = ("sample", "receiver", "normal")


def cornell(extra_lights=4, seed=1):
    return ra.HostScene(ra.RGI_SCENE_CORNELL, seed=seed, extra_lights=extra_lights, light_seed=seed)


def frame_for(hs, settings, width, height, frame_index, history_valid, di_history_valid=False, yaw_step=0.0,
              jitter_seed=None, distance=3.17):
    eye, vp, inv = ra.orbit_camera(hs, frame_index, width, height, yaw_step=yaw_step, distance=distance)
    if frame_index > 0:
        peye, pvp, _ = ra.orbit_camera(hs, frame_index - 1, width, height, yaw_step=yaw_step, distance=distance)
    else:
        peye, pvp = eye, vp
    return ra.make_frame(settings, hs, frame_index, eye, vp, inv, peye, pvp, width, height, history_valid,
                         di_history_valid)


def gi_size(width, height, divisor):
    """The GI size of a window (renderer.cpp:519-520)."""
    return max(1, width // divisor), max(1, height // divisor)


def empty_history(width, height, table_size, divisor=1):
    n = width * height
    gw, gh = gi_size(width, height, divisor)
    ng = gw * gh
    return {
        "gi_sample": np.zeros((ng, 4), np.float32),
        "gi_receiver": np.zeros((ng, 4), np.float32),
        "gi_normal": np.zeros((ng, 4), np.float32),
        "color": np.zeros((ng, 4), np.uint16),
        "moments": np.zeros((ng, 4), np.uint16),
        "filtered": np.zeros((ng, 4), np.uint16),
        "specular_gi": np.zeros((ng, 4), np.uint16),
        "hdr": np.zeros((n, 4), np.uint16),
        "grid": np.zeros(34 * table_size, np.uint32),
        "di_sample": np.zeros((n, 4), np.float32),
        "di_receiver": np.zeros((n, 4), np.float32),
        "di_normal": np.zeros((n, 4), np.uint16),
    }


# ------------------------------------------------------------------ oracle

def oracle_gbuffer(hs, frame):
    L = orc.lib()
    n = frame.width * frame.height
    g = {
        "depth": np.zeros(n, np.float32),
        "albedo": np.zeros(n, np.uint32),
        "normal": np.zeros(n, np.uint32),
        "shading": np.zeros(n, np.uint32),
        "emission": np.zeros((n, 4), np.uint16),
    }
    scene = orc.host_scene(hs)
    L.orc_gbuffer_primary(ctypes.byref(frame), ctypes.byref(scene), orc.ptr(hs.materials), hs.material_count,
                          orc.ptr(g["depth"]), orc.ptr(g["albedo"]), orc.ptr(g["normal"]), orc.ptr(g["shading"]),
                          orc.ptr(g["emission"]))
    return g


def host_gbuffer(g, width, height):
    return ra.GBuffer(width, height, g["depth"].ctypes.data, g["albedo"].ctypes.data, g["normal"].ctypes.data,
                      g["shading"].ctypes.data, g["emission"].ctypes.data)


def frame_shadow(hs, frame):
    """The rgi_shadow a frame's shadow uniforms came from (rgi_frame_pack)."""
    return ra.shadow_setup(list(frame.light_position)[:3], hs.radius, frame.shadow_width // 3,
                           frame.homogeneous_depth, frame.origin_bottom_left)


def oracle_shadow_atlas(hs, frame):
    sh = frame_shadow(hs, frame)
    atlas = np.zeros(6 * sh.face_size * sh.face_size, np.float32)
    orc.lib().orc_shadow_atlas(ctypes.byref(sh), ctypes.byref(orc.host_scene(hs)), orc.ptr(atlas))
    return atlas


def oracle_denoise(frame, out, atrous_iterations):
    """The denoiser of Renderer::render (renderer.cpp:857-907) on the oracle:
    returns (filtered history, denoised)."""
    L = orc.lib()
    n = frame.gi_width * frame.gi_height
    targets = [np.zeros((n, 4), np.uint16), np.zeros((n, 4), np.uint16)]
    filtered = None
    rec, nor = out["gi_receiver"], out["gi_normal"]
    L.orc_svgf_variance(ctypes.byref(frame), orc.ptr(out["color"]), orc.ptr(out["moments"]), orc.ptr(rec),
                        orc.ptr(nor), orc.ptr(targets[0]))
    denoised = out["color"]
    for it in range(atrous_iterations):
        src = targets[0] if it == 0 else filtered if it == 1 else targets[(it - 1) & 1]
        dst = np.zeros((n, 4), np.uint16) if it == 0 else targets[it & 1]
        L.orc_svgf_atrous(ctypes.byref(frame), orc.ptr(src), orc.ptr(rec), orc.ptr(nor), 1 << it, orc.ptr(dst))
        if it == 0:
            filtered = dst
        denoised = dst.copy()
    return filtered, denoised


def oracle_composite(frame, g, out, atlas, di_on):
    """The composite of Renderer::render's final image (renderer.cpp:967-990)."""
    n = g["depth"].shape[0]
    hdr = np.zeros((n, 4), np.uint16)
    spec = out.get("specular_gi")
    ci = ra.CompositeInputs(out["denoised"].ctypes.data, out["diffuse"].ctypes.data if di_on else None,
                            out["specular"].ctypes.data if di_on else None,
                            spec.ctypes.data if spec is not None else None,
                            atlas.ctypes.data if atlas is not None else None)
    ci.options[:] = [1.0, 1.0, 1.0, 1.0]
    ci.specular_weight = 1.0 if spec is not None else 0.0
    orc.lib().orc_composite(ctypes.byref(frame), ctypes.byref(host_gbuffer(g, frame.width, frame.height)),
                            ctypes.byref(ci), orc.ptr(hdr))
    return hdr


class _Timed:
    """The oracle library with each call's wall time added to `times[name]`."""

    def __init__(self, lib, times):
        self.lib, self.times = lib, times

    def __getattr__(self, name):
        fn = getattr(self.lib, name)
        if self.times is None:
            return fn

        def call(*a):
            t0 = time.perf_counter()
            r = fn(*a)
            self.times[name] = self.times.get(name, 0.0) + time.perf_counter() - t0
            return r
        return call


def oracle_frame(hs, frame, g, hist, insert=True, denoiser=False, atrous_iterations=5, composite=False,
                 specular=False, times=None):
    """One frame on the oracle; returns this frame's outputs (host arrays).
    With `times` (a dict), each oracle pass's wall time in seconds is added
    under its function name."""
    L = _Timed(orc.lib(), times)
    w, h, T = frame.width, frame.height, frame.grid_table_size
    n = w * h                              # window-size planes: G-buffer, DI, composite
    ng = frame.gi_width * frame.gi_height  # GI-size planes
    scene = orc.host_scene(hs)
    gb = host_gbuffer(g, w, h)
    out = {}
    atlas = oracle_shadow_atlas(hs, frame) if frame.shadows_enabled else None
    shadow_ptr = atlas.ctypes.data if atlas is not None else None
    if frame.di_light_count > 0:
        prev = ra.DiReservoirs(hist["di_sample"].ctypes.data, hist["di_receiver"].ctypes.data,
                               hist["di_normal"].ctypes.data)
        out["di_sample"] = np.zeros((n, 4), np.float32)
        out["di_receiver"] = np.zeros((n, 4), np.float32)
        out["di_normal"] = np.zeros((n, 4), np.uint16)
        dres = ra.DiReservoirs(out["di_sample"].ctypes.data, out["di_receiver"].ctypes.data, out["di_normal"].ctypes.data)
        L.orc_di_temporal(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(prev), ctypes.byref(dres))
        out["diffuse"] = np.zeros((n, 4), np.uint16)
        out["specular"] = np.zeros((n, 4), np.uint16)
        L.orc_di_spatial(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(dres),
                         orc.ptr(out["diffuse"]), orc.ptr(out["specular"]))
        surface_light = out["diffuse"]
    elif frame.lighting_surface == 1.0 and frame.lighting_emission == 0.0:
        # ReSTIR DI selected without lights: a black surface light (renderer.cpp:690-695).
        surface_light = np.zeros_like(g["emission"])
    else:
        surface_light = g["emission"]
    grid = np.zeros(34 * T, np.uint32)
    L.orc_grid_clear(orc.ptr(grid), T)
    if insert:
        L.orc_grid_insert(ctypes.byref(frame), ctypes.byref(gb), orc.ptr(grid))
    out["grid"] = grid
    # With the denoiser, the GI history colour is the filtered history (renderer.cpp:774-776).
    history_color = hist["filtered"] if denoiser else hist["color"]
    gin = ra.GiInputs(hist["gi_sample"].ctypes.data, hist["gi_receiver"].ctypes.data, hist["gi_normal"].ctypes.data,
                      history_color.ctypes.data, hist["moments"].ctypes.data, surface_light.ctypes.data, shadow_ptr,
                      hist["grid"].ctypes.data)
    for k in GI_PLANES_F32:
        out["gi_" + k] = np.zeros((ng, 4), np.float32)
    gres = ra.GiReservoirs(out["gi_sample"].ctypes.data, out["gi_receiver"].ctypes.data, out["gi_normal"].ctypes.data)
    L.orc_gi_temporal(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(gin), ctypes.byref(gres))
    sin = ra.GiInputs(None, hist["gi_receiver"].ctypes.data, hist["gi_normal"].ctypes.data, history_color.ctypes.data,
                      hist["moments"].ctypes.data, surface_light.ctypes.data, shadow_ptr, grid.ctypes.data)
    out["color"] = np.zeros((ng, 4), np.uint16)
    out["moments"] = np.zeros((ng, 4), np.uint16)
    L.orc_gi_spatial(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(sin), ctypes.byref(gres),
                     orc.ptr(out["color"]), orc.ptr(out["moments"]))
    out["surface_light"] = surface_light
    if atlas is not None:
        out["shadow_atlas"] = atlas
    if specular:
        # Indirect specular (renderer.cpp:836-855), with last frame's composite.
        spin = ra.GiInputs(None, hist["gi_receiver"].ctypes.data, hist["gi_normal"].ctypes.data,
                           history_color.ctypes.data, hist["moments"].ctypes.data, surface_light.ctypes.data,
                           shadow_ptr, grid.ctypes.data)
        out["specular_gi"] = np.zeros((ng, 4), np.uint16)
        prev_hdr = hist["hdr"] if composite else None
        L.orc_specular_gi(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(spin),
                          prev_hdr.ctypes.data if prev_hdr is not None else None, hist["specular_gi"].ctypes.data,
                          orc.ptr(out["specular_gi"]))
    out["denoised"] = out["color"]
    if denoiser:
        filtered, denoised = oracle_denoise(frame, out, atrous_iterations)
        if filtered is not None:
            out["filtered"] = filtered
        out["denoised"] = denoised
    if composite:
        out["hdr"] = oracle_composite(frame, g, out, atlas, frame.di_light_count > 0)
    return out


def next_history(hist, out):
    """The ping-pong of renderer.cpp:683, 1119-1124 (denoiser off)."""
    h = dict(hist)
    for k in ("gi_sample", "gi_receiver", "gi_normal", "color", "moments", "grid"):
        h[k] = out[k]
    for k in ("filtered", "specular_gi", "hdr"):
        if k in out:
            h[k] = out[k]
    if "di_sample" in out:
        for k in ("di_sample", "di_receiver", "di_normal"):
            h[k] = out[k]
    return h


def oracle_sequence(hs, settings, width, height, frames, yaw_step=0.0, distance=3.17):
    """Runs `frames` frames on the oracle; returns (frames, gbuffers, histories, outputs)."""
    return list(oracle_frames(hs, settings, width, height, frames, yaw_step=yaw_step, distance=distance))


def camera(hs, i, width, height, yaw_step=0.0, distance=3.17, yaw_of=None):
    """Frame i's camera of a path: a constant yaw step, or yaw_of(i)."""
    return ra.orbit_camera(hs, i, width, height, yaw_step=yaw_step, distance=distance, yaw_of=yaw_of)


def oracle_frames(hs, settings, width, height, frames, yaw_step=0.0, distance=3.17, yaw_of=None):
    """oracle_sequence as a generator (each frame's arrays can be dropped
    once compared): yields (frame, gbuffer, history, outputs)."""
    hist = empty_history(width, height, settings.grid_table_size, settings.gi_divisor)
    di_on = settings.restir_di and hs.light_count > 0
    for i in range(frames):
        eye, vp, inv = camera(hs, i, width, height, yaw_step, distance, yaw_of)
        peye, pvp, _ = camera(hs, i - 1, width, height, yaw_step, distance, yaw_of) if i > 0 else (eye, vp, inv)
        f = ra.make_frame(settings, hs, i, eye, vp, inv, peye, pvp, width, height, i > 0, i > 0 and di_on)
        g = oracle_gbuffer(hs, f)
        out = oracle_frame(hs, f, g, hist, denoiser=settings.denoiser, atrous_iterations=settings.atrous_iterations,
                           composite=settings.composite, specular=settings.specular_gi)
        yield f, g, hist, out
        hist = next_history(hist, out)


# ---------------------------------------------------------------- device

class Device:
    """The HIP library over torch device tensors (plumbing only)."""

    def __init__(self):
        import torch
        self.torch = torch
        self.L = ra.lib()
        ctx = ctypes.c_void_p()
        ra.check(self.L.rgi_create(ctypes.byref(ctx)))
        self.ctx = ctx

    def close(self):
        if self.ctx:
            self.L.rgi_destroy(self.ctx)
            self.ctx = None

    @property
    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def put(self, a):
        t = self.torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).cuda()
        return t

    def empty(self, nbytes):
        return self.torch.zeros(nbytes, dtype=self.torch.uint8, device="cuda")

    def get(self, t, like):
        self.torch.cuda.synchronize()
        return t.cpu().numpy().view(like.dtype).reshape(like.shape)

    def scene(self, hs):
        self.bvh_t = self.put(hs.bvh)
        self.lights_t = self.put(hs.lights)
        return hs.rgi_scene(self.bvh_t.data_ptr(), self.lights_t.data_ptr())

    def gbuffer(self, g, width, height):
        d = {k: self.put(v) for k, v in g.items()}
        gb = ra.GBuffer(width, height, d["depth"].data_ptr(), d["albedo"].data_ptr(), d["normal"].data_ptr(),
                        d["shading"].data_ptr(), d["emission"].data_ptr())
        return gb, d


def grid_view(grid, table):
    """What readers of a grid can see (hashgrid.sh:155-177): checksums, counts
    and each slot's first min(count, 32) entries; entries past the count are
    never read, and a cleared grid keeps whatever they held."""
    T = table
    cs, cnt = grid[:T].copy(), grid[T:2 * T].copy()
    entries = grid[2 * T:].reshape(T, 32).copy()
    listed = np.minimum(cnt, 32)
    entries[np.arange(32)[None, :] >= listed[:, None]] = 0
    return cs, cnt, entries


def replicate(plane, width, height, k):
    """A width x height plane (rows of texels, any trailing shape) with every
    texel repeated k x k times: the same image at k times the resolution."""
    a = plane.reshape((height, width) + plane.shape[1:])
    a = np.repeat(np.repeat(a, k, axis=0), k, axis=1)
    return np.ascontiguousarray(a.reshape((height * k * width * k,) + plane.shape[1:]))


def gi_passes_oracle(hs, frame, g, hist, surface_light, previous_hdr):
    """Grid insert, GI temporal, GI spatial, indirect specular and the
    denoiser's first two steps on the oracle for one frame; returns the
    GI-size outputs."""
    L = orc.lib()
    w, h, T = frame.width, frame.height, frame.grid_table_size
    ng = frame.gi_width * frame.gi_height
    scene = orc.host_scene(hs)
    gb = host_gbuffer(g, w, h)
    out = {"grid": np.zeros(34 * T, np.uint32)}
    L.orc_grid_clear(orc.ptr(out["grid"]), T)
    L.orc_grid_insert(ctypes.byref(frame), ctypes.byref(gb), orc.ptr(out["grid"]))
    gin = ra.GiInputs(hist["gi_sample"].ctypes.data, hist["gi_receiver"].ctypes.data, hist["gi_normal"].ctypes.data,
                      hist["filtered"].ctypes.data, hist["moments"].ctypes.data, surface_light.ctypes.data, None,
                      hist["grid"].ctypes.data)
    for k in GI_PLANES_F32:
        out["gi_" + k] = np.zeros((ng, 4), np.float32)
    gres = ra.GiReservoirs(out["gi_sample"].ctypes.data, out["gi_receiver"].ctypes.data, out["gi_normal"].ctypes.data)
    L.orc_gi_temporal(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(gin), ctypes.byref(gres))
    sin = ra.GiInputs(None, hist["gi_receiver"].ctypes.data, hist["gi_normal"].ctypes.data,
                      hist["filtered"].ctypes.data, hist["moments"].ctypes.data, surface_light.ctypes.data, None,
                      out["grid"].ctypes.data)
    out["color"] = np.zeros((ng, 4), np.uint16)
    out["moments"] = np.zeros((ng, 4), np.uint16)
    L.orc_gi_spatial(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(sin), ctypes.byref(gres),
                     orc.ptr(out["color"]), orc.ptr(out["moments"]))
    out["specular_gi"] = np.zeros((ng, 4), np.uint16)
    L.orc_specular_gi(ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb), ctypes.byref(sin),
                      previous_hdr.ctypes.data, hist["specular_gi"].ctypes.data, orc.ptr(out["specular_gi"]))
    out["filtered"], out["denoised"] = oracle_denoise(frame, out, 2)
    return out
