"""GPU: every benchmark configuration at its own size, against the oracle.

- C2: the Cornell box at 1920 x 1080 with a 2^20-cell grid, the benchmark's
  16-frame camera path through the frame driver (DI, grid, GI, indirect
  specular, denoiser, composite), every plane of every frame bit-exact in
  the deterministic insert mode; the atomic insert's invariants at that size;
  the traversal microbenchmark's 2^24 rays, hit ids and distances bit-exact.
- C3: the ~256k-triangle atrium at 1080p with a 2^22-cell grid and DI over
  its 128 lights, eager and graph-captured frames bit-exact against the
  oracle; its 8 views as one batch, bit-exact against 8 renderers alone
  (and two of them against the oracle).
- C4: the neural radiance cache at its configured size (16 levels x 2^19
  entries) on 65,536 training records gathered from the Cornell box and the
  atrium at 1080p: loss and gradients of every layer and of the table
  against the oracle.
- C5: the 16-scene uneven batch with a cache per renderer, bit-exact against
  the same renderers alone without a cache over its first frames (the
  cache's warm-up, during which the frame is the plain path), and every
  scene alone against the oracle.
- DI over more than 256 lights (the light table read from global memory
  instead of LDS), through the C ABI against the oracle.
"""
import ctypes

import numpy as np
import pytest

import restir_amd as ra
from restir_amd import nrc as rnrc
from restir_amd.workloads import UNEVEN_SCENES, surface_rays

import common
import orc

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
ATRIUM_DISTANCE = 0.7


@pytest.fixture(scope="module")
def torch():
    t = pytest.importorskip("torch")
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def say(*a):
    print(*a, flush=True)


def upload(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).cuda()


def download(torch, ptr, like):
    hip = ra.lib()
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    out = np.empty_like(like)
    assert hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), out.nbytes, 2) == 0
    return out


def swing(k):
    return ra.swing_yaw(k, 16)


def atrium():
    return ra.HostScene(ra.RGI_SCENE_ATRIUM, seed=1, size=262144, extra_lights=64, light_seed=1)


def atrium_settings(**kw):
    """The benchmark's atrium workload: 2^22 cells, DI over every light."""
    return ra.Settings(grid_table_size=1 << 22, grid_mode=ra.RGI_GRID_DETERMINISTIC, denoiser=False,
                       composite=False, specular_gi=False, **kw)


ALL_VARIANTS = ra.tuning_of(ra.RGI_TUNE_TEMPORAL_WAVES4 | ra.RGI_TUNE_SPATIAL_WAVES5 | ra.RGI_TUNE_TILE_ORDER |
                            ra.RGI_TUNE_TILE_SHAPE)


class Renderer:
    """One rgi_renderer over device scene buffers."""

    def __init__(self, torch, hs, settings, w, h, variants=None, nrc=None, gscene=None):
        self.torch, self.hs, self.s, self.w, self.h = torch, hs, settings, w, h
        self.L = ra.lib()
        if gscene is None:
            self.bvh, self.lights = upload(torch, hs.bvh), upload(torch, hs.lights)
            self.scene = hs.rgi_scene(self.bvh.data_ptr(), self.lights.data_ptr())
        else:
            self.scene = gscene
        self.r = ctypes.c_void_p()
        ra.check(self.L.rgi_renderer_create(ctypes.byref(self.r), w, h, ctypes.byref(ra.c_settings(settings)),
                                            ctypes.byref(self.scene), hs.radius))
        for opt, value in (variants or {}).items():
            ra.check(self.L.rgi_renderer_set_option(self.r, opt, value))
        if nrc is not None:
            ra.check(self.L.rgi_renderer_set_nrc(self.r, ctypes.byref(nrc)))

    def close(self):
        if self.r:
            self.L.rgi_renderer_destroy(self.r)
            self.r = None

    def render(self, view, gb, stream, graph=False):
        fn = self.L.rgi_renderer_render_graph if graph else self.L.rgi_renderer_render
        ra.check(fn(self.r, ctypes.byref(view), ctypes.byref(gb), stream))

    def outputs(self):
        out = ra.RendererOutputs()
        ra.check(self.L.rgi_renderer_outputs_get(self.r, ctypes.byref(out)))
        return out

    def checks(self, di_on):
        out = self.outputs()
        c = [("gi_sample", out.gi_sample), ("gi_receiver", out.gi_receiver), ("gi_normal", out.gi_normal),
             ("color", out.color), ("moments", out.moments), ("grid", out.grid)]
        if di_on:
            c += [("diffuse", out.di_diffuse), ("specular", out.di_specular)]
        if self.s.denoiser:
            c += [("filtered", out.filtered_history), ("denoised", out.denoised)]
        if self.s.composite:
            c += [("hdr", out.hdr)]
        if self.s.specular_gi:
            c += [("specular_gi", out.specular_gi)]
        return c


def compare(torch, what, checks, want, table):
    for name, ptr in checks:
        got = download(torch, ptr, want[name])
        if name == "grid":
            for part, a, b in zip(("checksums", "counts", "entries"), common.grid_view(got, table),
                                  common.grid_view(want[name], table)):
                assert np.array_equal(a, b), "%s: grid %s differ" % (what, part)
        else:
            eq = got == want[name]
            assert eq.all(), "%s: %s differs in %d of %d values" % (what, name, int((~eq).sum()), eq.size)


def device_gbuffer(torch, g, w, h):
    planes = {k: upload(torch, v) for k, v in g.items()}
    gb = ra.GBuffer(w, h, planes["depth"].data_ptr(), planes["albedo"].data_ptr(), planes["normal"].data_ptr(),
                    planes["shading"].data_ptr(), planes["emission"].data_ptr())
    return gb, planes


def run_against_oracle(torch, hs, s, w, h, frames, distance, variants=None, graph=False, yaw_of=swing, tag=""):
    """The frame driver against the oracle frame by frame (each frame's
    arrays dropped once compared)."""
    rend = Renderer(torch, hs, s, w, h, variants)
    # A captured frame needs a stream other than the null stream.
    side = torch.cuda.Stream() if graph else None
    stream = ctypes.c_void_p((side or torch.cuda.current_stream()).cuda_stream)
    try:
        for i, (f, g, hist, want) in enumerate(common.oracle_frames(hs, s, w, h, frames, distance=distance,
                                                                     yaw_of=yaw_of)):
            say("  %s frame %d" % (tag, i))
            eye, vp, inv = common.camera(hs, i, w, h, distance=distance, yaw_of=yaw_of)
            gb, keep = device_gbuffer(torch, g, w, h)
            torch.cuda.synchronize()
            rend.render(ra.make_view(eye, vp, inv), gb, stream, graph)
            used = ra.Frame()
            ra.check(rend.L.rgi_renderer_last_frame(rend.r, ctypes.byref(used)))
            assert bytes(used) == bytes(f), "frame %d: packed uniforms differ" % i
            compare(torch, "%s frame %d" % (tag, i), rend.checks(f.di_light_count > 0), want, s.grid_table_size)
    finally:
        rend.close()


# ----------------------------------------------------------------- C2

def test_c2_cornell_1080p_16_frames(torch):
    say("case C2 cornell 1920x1080 T=2^20, 16 frames of the benchmark path, deterministic insert")
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(grid_table_size=1 << 20, grid_mode=ra.RGI_GRID_DETERMINISTIC)
    run_against_oracle(torch, hs, s, W, H, 16, 3.17, tag="C2")


def test_c2_atomic_insert_1080p_invariants(torch):
    """The atomic insert at 1080p and 2^20 cells: every listed entry is a
    pixel with that cell's checksum, lists hold min(count, 32) distinct
    pixels, and with no pixel dropped at this load the multiset of
    (checksum, count) equals the deterministic (oracle) insert's."""
    say("case C2 atomic insert 1920x1080 T=2^20")
    hs = common.cornell(extra_lights=4)
    T = 1 << 20
    s = ra.Settings(grid_table_size=T)
    f = common.frame_for(hs, s, W, H, 5, True, yaw_step=0.0)
    g = common.oracle_gbuffer(hs, f)
    want = np.zeros(34 * T, np.uint32)
    hg = common.host_gbuffer(g, W, H)
    orc.lib().orc_grid_insert(ctypes.byref(f), ctypes.byref(hg), orc.ptr(want))
    slot0, cs = np.zeros(W * H, np.uint32), np.zeros(W * H, np.uint32)
    orc.lib().orc_grid_keys(ctypes.byref(f), ctypes.byref(hg), orc.ptr(slot0), orc.ptr(cs), None)
    d = common.Device()
    try:
        gb, keep = d.gbuffer(g, W, H)
        grid = d.empty(34 * T * 4)
        ra.check(d.L.rgi_grid_clear(grid.data_ptr(), T, d.stream))
        ra.check(d.L.rgi_grid_insert(d.ctx, ctypes.byref(f), ctypes.byref(gb), grid.data_ptr(), ra.RGI_GRID_ATOMIC,
                                     d.stream))
        got = d.get(grid, want)
    finally:
        d.close()
    cnt, wcnt = got[T:2 * T], want[T:2 * T]
    occupied = np.nonzero(got[:T])[0]
    assert cnt.sum() == (cs != 0).sum() == wcnt.sum(), "no pixel is dropped at this load"
    entries = got[2 * T:].reshape(T, 32)
    k = np.minimum(cnt[occupied], 32)
    for s_, kk in zip(occupied, k):
        e = entries[s_, :kk]
        assert len(np.unique(e)) == kk
        assert np.all(cs[e] == got[s_])
    assert sorted(zip(got[:T][occupied], cnt[occupied])) == \
        sorted(zip(want[:T][want[:T] != 0], wcnt[want[:T] != 0]))
    say("  cells %d, surface pixels %d" % (len(occupied), int((cs != 0).sum())))


def test_c2_traversal_2pow24_rays(torch):
    """The traversal microbenchmark's 2^24 cosine rays from points uniform
    over the Cornell box's area, closest hit, in both launch forms (one ray
    per lane, persistent waves): hit triangle ids and distances bit-exact
    against the oracle."""
    say("case C2 traversal 2^24 rays")
    hs = common.cornell(extra_lights=4)
    rays = surface_rays(hs, 1 << 24, seed=7)
    n = rays.shape[0]
    want_t, want_d = np.zeros(n, np.int32), np.zeros(n, np.float32)
    orc.lib().orc_trace(ctypes.byref(orc.host_scene(hs)), orc.ptr(rays), n, orc.ptr(want_t), orc.ptr(want_d), 0)
    d = common.Device()
    try:
        scene = d.scene(hs)
        rt = d.put(rays)
        got = []
        for persistent in (0, 1):
            ra.check(d.L.rgi_context_set_option(d.ctx, ra.RGI_OPT_TRACE_PERSISTENT, persistent))
            ht, hd = d.empty(4 * n), d.empty(4 * n)
            ra.check(d.L.rgi_trace(d.ctx, ctypes.byref(scene), rt.data_ptr(), n, ht.data_ptr(), hd.data_ptr(), 0,
                                   d.stream))
            got.append((d.get(ht, want_t), d.get(hd, want_d)))
            del ht, hd
    finally:
        d.close()
    for persistent, (got_t, got_d) in enumerate(got):
        assert np.array_equal(got_t, want_t), "persistent=%d: %d hit ids differ" % (persistent,
                                                                                    int((got_t != want_t).sum()))
        assert np.array_equal(got_d, want_d), "persistent=%d" % persistent
    say("  hit rate %.4f" % (want_t >= 0).mean())


# ----------------------------------------------------------------- C3

@pytest.mark.parametrize("graph", [False, True])
def test_c3_atrium_1080p(torch, graph):
    say("case C3 atrium 1920x1080 T=2^22 graph=%d" % graph)
    hs = atrium()
    assert hs.light_count == 128 and hs.triangle_count > 250000
    run_against_oracle(torch, hs, atrium_settings(), W, H, 3, ATRIUM_DISTANCE, ALL_VARIANTS if graph else None,
                       graph=graph, tag="C3 graph" if graph else "C3")


def test_c3_atrium_8_views_batch(torch):
    """8 views of the atrium at 1080p (yaw offsets of 45 degrees) as one
    batch: every renderer's planes bit-exact against the same renderer
    alone, frame by frame; views 0 and 5 also against the oracle."""
    say("case C3 atrium 8 views batch 1920x1080")
    hs = atrium()
    s = atrium_settings()
    bvh, lights = upload(torch, hs.bvh), upload(torch, hs.lights)
    scene = hs.rgi_scene(bvh.data_ptr(), lights.data_ptr())
    L = ra.lib()
    frames = 3
    offsets = [v * np.pi / 4 for v in range(8)]
    paths = [(lambda k, o=o: swing(k) + o) for o in offsets]
    batched = [Renderer(torch, hs, s, W, H, gscene=scene) for _ in range(8)]
    alone = [Renderer(torch, hs, s, W, H, gscene=scene) for _ in range(8)]
    oracle = {v: common.oracle_frames(hs, s, W, H, frames, distance=ATRIUM_DISTANCE, yaw_of=paths[v]) for v in (0, 5)}
    batch = ctypes.c_void_p()
    arr = (ctypes.c_void_p * 8)(*[r.r.value for r in batched])
    ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, 8))
    side = torch.cuda.Stream()
    bs = ctypes.c_void_p(side.cuda_stream)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    gbs = {}
    try:
        for i in range(frames):
            say("  frame %d" % i)
            views, gbuf = [], []
            for v in range(8):
                eye, vp, inv = common.camera(hs, i, W, H, distance=ATRIUM_DISTANCE, yaw_of=paths[v])
                f = ra.make_frame(s, hs, i, eye, vp, inv, eye, vp, W, H, False)
                g = common.oracle_gbuffer(hs, f)
                gb, keep = device_gbuffer(torch, g, W, H)
                gbs[v] = keep
                views.append(ra.make_view(eye, vp, inv))
                gbuf.append(gb)
            torch.cuda.synchronize()
            va = (ra.View * 8)(*views)
            ga = (ra.GBuffer * 8)(*gbuf)
            with torch.cuda.stream(side):
                ra.check(L.rgi_batch_render(batch, va, ga, bs))
            side.synchronize()
            for v in range(8):
                alone[v].render(views[v], gbuf[v], stream)
            torch.cuda.synchronize()
            for v in range(8):
                for (name, pa), (_, pb) in zip(batched[v].checks(True), alone[v].checks(True)):
                    n_bytes = {"grid": 34 * s.grid_table_size * 4}.get(name, W * H * (16 if name.startswith("gi_") else 8))
                    a = download(torch, pa, np.zeros(n_bytes, np.uint8))
                    b = download(torch, pb, np.zeros(n_bytes, np.uint8))
                    if name == "grid":
                        a, b = a.view(np.uint32), b.view(np.uint32)
                        for part, x, y in zip(("checksums", "counts", "entries"),
                                              common.grid_view(a, s.grid_table_size),
                                              common.grid_view(b, s.grid_table_size)):
                            assert np.array_equal(x, y), "view %d frame %d: grid %s differ" % (v, i, part)
                    else:
                        assert np.array_equal(a, b), "view %d frame %d: %s differs from the renderer alone" % (
                            v, i, name)
            for v, gen in oracle.items():
                f, g, hist, want = next(gen)
                compare(torch, "batch view %d frame %d" % (v, i), batched[v].checks(True), want, s.grid_table_size)
    finally:
        L.rgi_batch_destroy(batch)
        for r in batched + alone:
            r.close()


# ----------------------------------------------------------------- C4

@pytest.mark.parametrize("scene", ["cornell", "atrium"])
def test_c4_cache_gradients_at_config_size(torch, scene):
    """The cache at its configured size (16 levels x 2 features x 2^19
    entries, 64-wide MLP) after a few frames of training in the frame
    driver: the loss and the gradients of one 65,536-record batch gathered
    from a 1080p frame, every MLP layer and the table, against the oracle's
    FP64 gradients with the GPU's rounding points."""
    say("case C4 cache gradients %s 2^19 x 65536" % scene)
    hs = atrium() if scene == "atrium" else common.cornell(extra_lights=4)
    dist = ATRIUM_DISTANCE if scene == "atrium" else 3.17
    s = atrium_settings() if scene == "atrium" else ra.Settings(grid_table_size=1 << 20)
    batch = 65536
    ns = rnrc.renderer_settings(hs, batch=batch, warmup=2)
    assert ns.cache.log2_table == 19
    rend = Renderer(torch, hs, s, W, H, nrc=ns)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = ra.lib()
    try:
        keep = []
        for i in range(4):
            eye, vp, inv = common.camera(hs, i, W, H, distance=dist, yaw_of=swing)
            f = ra.make_frame(s, hs, i, eye, vp, inv, eye, vp, W, H, False)
            gb, planes = device_gbuffer(torch, common.oracle_gbuffer(hs, f), W, H)
            keep.append(planes)
            rend.render(ra.make_view(eye, vp, inv), gb, stream)
        cache, _, _, trained = rnrc.renderer_state(rend.r)
        assert trained == 4
        used = ra.Frame()
        ra.check(L.rgi_renderer_last_frame(rend.r, ctypes.byref(used)))
        out = rend.outputs()
        rec = torch.zeros(batch * 12, dtype=torch.float32, device="cuda")
        tgt = torch.zeros(batch * 4, dtype=torch.float32, device="cuda")
        ra.check(L.rgi_nrc_gather_gbuffer(ctypes.c_void_p(cache), ctypes.byref(used), ctypes.byref(gb),
                                          ctypes.c_void_p(out.di_diffuse), ctypes.c_void_p(out.filtered_history),
                                          batch, 99, rec.data_ptr(), tgt.data_ptr(), stream))
        mlp = torch.empty(ra.RGI_NRC_MLP_PARAMS, dtype=torch.float32, device="cuda")
        cfg = ns.cache
        tab = torch.empty(ra.RGI_NRC_LEVELS * (1 << cfg.log2_table) * 2, dtype=torch.float32, device="cuda")
        ra.check(L.rgi_nrc_get_params(ctypes.c_void_p(cache), mlp.data_ptr(), tab.data_ptr(), stream))
        gm = torch.empty_like(mlp)
        gt = torch.empty_like(tab)
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        ra.check(L.rgi_nrc_gradients(ctypes.c_void_p(cache), rec.data_ptr(), tgt.data_ptr(), batch, gm.data_ptr(),
                                     gt.data_ptr(), loss.data_ptr(), stream))
        torch.cuda.synchronize()
        records = rec.cpu().numpy().reshape(batch, 12)
        targets = tgt.cpu().numpy().reshape(batch, 4)
        model = orc.NrcModel(rnrc.level_resolutions(cfg), cfg.log2_table, list(cfg.aabb_min), list(cfg.aabb_max),
                             mlp.cpu().numpy(), tab.cpu().numpy())
        wl, wm, wt = model.gradients(records, targets, emulate=1)
        gl, gm, gt = float(loss.cpu()[0]), gm.cpu().numpy().astype(np.float64), gt.cpu().numpy().astype(np.float64)
    finally:
        rend.close()
    say("  loss gpu %.6g oracle %.6g; surface records %d" % (gl, wl, int((targets[:, 3] > 0).sum())))
    assert (targets[:, 3] > 0).all()
    assert abs(gl - wl) < 1e-3 * wl
    for k in range(6):
        sl = slice(k * 4096, k * 4096 + 4096) if k < 5 else slice(20480, 20672)
        e = np.linalg.norm(gm[sl] - wm[sl]) / np.linalg.norm(wm[sl])
        assert e < 2e-2, (k, e)
    assert np.linalg.norm(gt - wt) < 5e-2 * np.linalg.norm(wt)
    assert np.array_equal(gt != 0.0, wt != 0.0)


# ----------------------------------------------------------------- C5

def uneven_settings(w, h):
    """The benchmark's uneven-batch renderer settings (deterministic insert)."""
    return ra.Settings(grid_table_size=1 << 20 if w * h <= (1 << 21) else 1 << 21,
                       grid_mode=ra.RGI_GRID_DETERMINISTIC, denoiser=False, composite=False, specular_gi=False)


def uneven_scene(kind, size, seed):
    return ra.HostScene(kind, seed=seed, size=size, extra_lights=4 if kind != ra.RGI_SCENE_ATRIUM else 64,
                        light_seed=1)


def test_c5_uneven_batch_with_cache(torch):
    """The 16 scenes of 36 to 1M triangles at 640x360 to 2560x1440, each
    renderer with its own cache, as one batch (first frame eager, then
    captured graphs): bit-exact against the same renderers alone without a
    cache over the cache's warm-up (4 trained frames), while the caches train
    every frame."""
    say("case C5 uneven batch of 16 with caches")
    L = ra.lib()
    frames = 3
    batched, alone, scenes = [], [], []
    for name, kind, size, seed, w, h, dist in UNEVEN_SCENES:
        hs = uneven_scene(kind, size, seed)
        bvh, lights = upload(torch, hs.bvh), upload(torch, hs.lights)
        scene = hs.rgi_scene(bvh.data_ptr(), lights.data_ptr())
        scenes.append((name, hs, (bvh, lights, upload(torch, hs.materials)), w, h, dist))
        s = uneven_settings(w, h)
        batched.append(Renderer(torch, hs, s, w, h, gscene=scene, nrc=rnrc.renderer_settings(hs)))
        alone.append(Renderer(torch, hs, s, w, h, gscene=scene))
    batch = ctypes.c_void_p()
    arr = (ctypes.c_void_p * 16)(*[r.r.value for r in batched])
    ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, 16))
    side = torch.cuda.Stream()
    bs = ctypes.c_void_p(side.cuda_stream)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ctx = ctypes.c_void_p()
    ra.check(L.rgi_create(ctypes.byref(ctx)))
    try:
        for i in range(frames):
            say("  frame %d" % i)
            views, gbuf, keep = [], [], []
            for (name, hs, bufs, w, h, dist), r in zip(scenes, batched):
                eye, vp, inv = common.camera(hs, i, w, h, distance=dist, yaw_of=swing)
                f = ra.make_frame(r.s, hs, i, eye, vp, inv, eye, vp, w, h, False)
                n = w * h
                t = {k: torch.empty(n * b, dtype=torch.uint8, device="cuda")
                     for k, b in (("depth", 4), ("albedo", 4), ("normal", 4), ("shading", 4), ("emission", 8))}
                ra.check(L.rgi_gbuffer_primary(ctx, ctypes.byref(f), ctypes.byref(r.scene),
                                               bufs[2].data_ptr(), hs.material_count,
                                               *(t[k].data_ptr() for k in ("depth", "albedo", "normal", "shading",
                                                                          "emission")), stream))
                keep.append(t)
                views.append(ra.make_view(eye, vp, inv))
                gbuf.append(ra.GBuffer(w, h, *(t[k].data_ptr() for k in ("depth", "albedo", "normal", "shading",
                                                                         "emission"))))
            torch.cuda.synchronize()
            with torch.cuda.stream(side):
                ra.check(L.rgi_batch_render(batch, (ra.View * 16)(*views), (ra.GBuffer * 16)(*gbuf), bs))
            side.synchronize()
            for k in range(16):
                alone[k].render(views[k], gbuf[k], stream)
            torch.cuda.synchronize()
            for k, (name, hs, bufs, w, h, dist) in enumerate(scenes):
                T = batched[k].s.grid_table_size
                for (pname, pa), (_, pb) in zip(batched[k].checks(True), alone[k].checks(True)):
                    nb = 34 * T * 4 if pname == "grid" else w * h * (16 if pname.startswith("gi_") else 8)
                    a = download(torch, pa, np.zeros(nb, np.uint8))
                    b = download(torch, pb, np.zeros(nb, np.uint8))
                    if pname == "grid":
                        for part, x, y in zip(("checksums", "counts", "entries"), common.grid_view(a.view(np.uint32), T),
                                              common.grid_view(b.view(np.uint32), T)):
                            assert np.array_equal(x, y), "%s frame %d: grid %s differ" % (name, i, part)
                    else:
                        assert np.array_equal(a, b), "%s frame %d: %s differs from the renderer alone" % (
                            name, i, pname)
        for r in batched:
            assert rnrc.renderer_state(r.r)[3] == frames, "every cache trained once per frame"
    finally:
        L.rgi_batch_destroy(batch)
        L.rgi_destroy(ctx)
        for r in batched + alone:
            r.close()


@pytest.mark.parametrize("case", UNEVEN_SCENES, ids=[c[0] for c in UNEVEN_SCENES])
def test_c5_scene_matches_oracle(torch, case):
    name, kind, size, seed, w, h, dist = case
    say("case C5 scene %s %dx%d" % (name, w, h))
    hs = uneven_scene(kind, size, seed)
    run_against_oracle(torch, hs, uneven_settings(w, h), w, h, 2, dist, tag=name)


# ------------------------------------------------------------ DI lights

@pytest.mark.parametrize("order,guide", [(0, True), ("wide", True), (0, False)])
def test_di_over_many_lights(torch, order, guide):
    """DI temporal and spatial through the C ABI on a box field lit by more
    than 256 lights (above the LDS light table's capacity: the light list
    is read from global memory), two frames, against the oracle; with a
    context the light choice starts from the CDF's guide table, without one
    it searches the whole CDF."""
    hs = ra.HostScene(ra.RGI_SCENE_BOXFIELD, seed=1, size=840, extra_lights=4, light_seed=1)
    say("case DI over %d lights order=%s guide=%s" % (hs.light_count, order, guide))
    assert hs.light_count > 256
    w, h = 640, 360
    s = ra.Settings()
    d = common.Device()
    opts = {ra.RGI_OPT_TILE_SHAPE: 1} if order == "wide" else {}
    try:
        for opt, value in opts.items():
            ra.check(d.L.rgi_context_set_option(d.ctx, opt, value))
        scene = d.scene(hs)
        n = w * h
        prev = {k: np.zeros((n, 4), t) for k, t in (("sample", np.float32), ("receiver", np.float32),
                                                     ("normal", np.uint16))}
        for i in range(2):
            f = common.frame_for(hs, s, w, h, i, i > 0, i > 0, yaw_step=2 * np.pi / 64, distance=0.5)
            g = common.oracle_gbuffer(hs, f)
            want = {k: np.zeros_like(v) for k, v in prev.items()}
            hg = common.host_gbuffer(g, w, h)
            hsc = orc.host_scene(hs)
            pr = ra.DiReservoirs(prev["sample"].ctypes.data, prev["receiver"].ctypes.data, prev["normal"].ctypes.data)
            wr = ra.DiReservoirs(want["sample"].ctypes.data, want["receiver"].ctypes.data, want["normal"].ctypes.data)
            orc.lib().orc_di_temporal(ctypes.byref(f), ctypes.byref(hsc), ctypes.byref(hg), ctypes.byref(pr),
                                      ctypes.byref(wr))
            wd, ws = np.zeros((n, 4), np.uint16), np.zeros((n, 4), np.uint16)
            orc.lib().orc_di_spatial(ctypes.byref(f), ctypes.byref(hsc), ctypes.byref(hg), ctypes.byref(wr),
                                     orc.ptr(wd), orc.ptr(ws))
            gb, keep = d.gbuffer(g, w, h)
            dp = {k: d.put(v) for k, v in prev.items()}
            do = {k: d.empty(v.nbytes) for k, v in want.items()}
            dd, dsp = d.empty(n * 8), d.empty(n * 8)
            ra.check(d.L.rgi_di_temporal(d.ctx if guide else None, ctypes.byref(f), ctypes.byref(scene), ctypes.byref(gb),
                                         ctypes.byref(ra.DiReservoirs(*(dp[k].data_ptr() for k in
                                                                        ("sample", "receiver", "normal")))),
                                         ctypes.byref(ra.DiReservoirs(*(do[k].data_ptr() for k in
                                                                        ("sample", "receiver", "normal")))),
                                         d.stream))
            ra.check(d.L.rgi_di_spatial(d.ctx, ctypes.byref(f), ctypes.byref(scene), ctypes.byref(gb),
                                        ctypes.byref(ra.DiReservoirs(*(do[k].data_ptr() for k in
                                                                       ("sample", "receiver", "normal")))),
                                        dd.data_ptr(), dsp.data_ptr(), d.stream))
            for k in want:
                got = d.get(do[k], want[k])
                assert np.array_equal(got, want[k]), "frame %d: DI %s differs" % (i, k)
            assert np.array_equal(d.get(dd, wd), wd), "frame %d: DI diffuse differs" % i
            assert np.array_equal(d.get(dsp, ws), ws), "frame %d: DI specular differs" % i
            lit = (wd.view(np.float16)[:, :3] != 0).any(1).mean()
            say("  frame %d lit %.3f" % (i, lit))
            assert lit > 0.1
            prev = want
    finally:
        d.close()
