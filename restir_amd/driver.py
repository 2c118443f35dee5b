"""Frame driver over the C ABI: device buffers with the reference's ping-pong
(renderer.cpp:626-833, 1119-1124) and the pass order DI temporal -> DI
spatial -> grid clear -> grid insert -> GI temporal -> GI spatial.

Device memory comes from torch (plumbing only); every pass is a
librestir_amd.so kernel launched on torch's current stream.
"""
import ctypes

import numpy as np

import restir_amd as ra


class GpuScene:
    def __init__(self, hs):
        import torch
        self.hs = hs
        self.bvh = torch.from_numpy(hs.bvh.view(np.uint8)).cuda()
        self.lights = torch.from_numpy(hs.lights.view(np.uint8)).cuda()
        self.materials = torch.from_numpy(hs.materials.view(np.uint8)).cuda()
        self.scene = hs.rgi_scene(self.bvh.data_ptr(), self.lights.data_ptr())


class GBufferSet:
    """G-buffer planes on the device, produced by rgi_gbuffer_primary."""

    def __init__(self, ctx, gscene, frame):
        import torch
        n = frame.width * frame.height
        z = lambda b: torch.empty(n * b, dtype=torch.uint8, device="cuda")
        self.depth, self.albedo, self.normal, self.shading, self.emission = z(4), z(4), z(4), z(4), z(8)
        L = ra.lib()
        ra.check(L.rgi_gbuffer_primary(ctx, ctypes.byref(frame), ctypes.byref(gscene.scene),
                                       gscene.materials.data_ptr(), gscene.hs.material_count, self.depth.data_ptr(),
                                       self.albedo.data_ptr(), self.normal.data_ptr(), self.shading.data_ptr(),
                                       self.emission.data_ptr(), stream()))
        self.gb = ra.GBuffer(frame.width, frame.height, self.depth.data_ptr(), self.albedo.data_ptr(),
                             self.normal.data_ptr(), self.shading.data_ptr(), self.emission.data_ptr())


def stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class FrameState:
    """One parity of the per-frame buffers."""

    def __init__(self, n, table):
        import torch
        z = lambda b: torch.zeros(b, dtype=torch.uint8, device="cuda")
        self.gi = [z(16 * n) for _ in range(3)]       # sample, receiver, normal
        self.color, self.moments = z(8 * n), z(8 * n)
        self.grid = z(4 * 34 * table)
        self.di = [z(16 * n), z(16 * n), z(8 * n)]   # sample, receiver, normal(half)

    def gi_res(self):
        return ra.GiReservoirs(*(t.data_ptr() for t in self.gi))

    def di_res(self):
        return ra.DiReservoirs(*(t.data_ptr() for t in self.di))


class Driver:
    """Runs frames of the hot path; `pass_hook(name)` brackets each pass."""

    def __init__(self, width, height, table, grid_mode=ra.RGI_GRID_ATOMIC):
        import torch
        self.L = ra.lib()
        ctx = ctypes.c_void_p()
        ra.check(self.L.rgi_create(ctypes.byref(ctx)))
        self.ctx = ctx
        n = width * height
        self.state = [FrameState(n, table), FrameState(n, table)]
        self.read = 0
        self.diffuse = torch.zeros(8 * n, dtype=torch.uint8, device="cuda")
        self.specular = torch.zeros(8 * n, dtype=torch.uint8, device="cuda")
        self.table = table
        self.grid_mode = grid_mode

    def close(self):
        if self.ctx:
            self.L.rgi_destroy(self.ctx)
            self.ctx = None

    def render(self, frame, gscene, gbuf, hook=None):
        L, s = self.L, stream()
        r, w = self.state[self.read], self.state[1 - self.read]
        hook = hook or (lambda name, start: None)
        f = ctypes.byref(frame)
        sc = ctypes.byref(gscene.scene)
        gb = ctypes.byref(gbuf.gb)
        if frame.di_light_count > 0:
            hook("di_temporal", True)
            ra.check(L.rgi_di_temporal(self.ctx, f, sc, gb, ctypes.byref(r.di_res()), ctypes.byref(w.di_res()), s))
            hook("di_temporal", False)
            hook("di_spatial", True)
            ra.check(L.rgi_di_spatial(self.ctx, f, sc, gb, ctypes.byref(w.di_res()), self.diffuse.data_ptr(),
                                      self.specular.data_ptr(), s))
            hook("di_spatial", False)
            surface_light = self.diffuse.data_ptr()
        else:
            surface_light = gbuf.emission.data_ptr()
        hook("grid_clear", True)
        ra.check(L.rgi_grid_clear(w.grid.data_ptr(), self.table, s))
        hook("grid_clear", False)
        hook("grid_insert", True)
        ra.check(L.rgi_grid_insert(self.ctx, f, gb, w.grid.data_ptr(), self.grid_mode, s))
        hook("grid_insert", False)
        gin = ra.GiInputs(r.gi[0].data_ptr(), r.gi[1].data_ptr(), r.gi[2].data_ptr(), r.color.data_ptr(),
                          r.moments.data_ptr(), surface_light, None, r.grid.data_ptr())
        hook("gi_temporal", True)
        ra.check(L.rgi_gi_temporal(self.ctx, f, sc, gb, ctypes.byref(gin), ctypes.byref(w.gi_res()), s))
        hook("gi_temporal", False)
        sin = ra.GiInputs(None, r.gi[1].data_ptr(), r.gi[2].data_ptr(), r.color.data_ptr(), r.moments.data_ptr(),
                          surface_light, None, w.grid.data_ptr())
        hook("gi_spatial", True)
        ra.check(L.rgi_gi_spatial(self.ctx, f, sc, gb, ctypes.byref(sin), ctypes.byref(w.gi_res()),
                                  w.color.data_ptr(), w.moments.data_ptr(), s))
        hook("gi_spatial", False)
        self.read = 1 - self.read
        return w
