"""GPU: the C++ frame driver (rgi_renderer_*) against a frame sequence run
by the oracle with the same cameras, and the FP64 reference accumulation and
error metric against the oracle. Bit-exact."""
import ctypes

import numpy as np
import pytest

import restir_amd as ra

import common
import orc

pytestmark = pytest.mark.gpu

W, H = 160, 120


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
    """Copy `like.nbytes` from a device pointer the library owns."""
    hip = ra.lib()  # hipMemcpy resolves through the library's HIP runtime
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    out = np.empty_like(like)
    assert hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), out.nbytes, 2) == 0
    return out


# Every launch variant away from its default (GI spatial's default is the
# 5-wave form, so the bit set leaves it at the compiler's budget).
ALL_VARIANTS = ra.tuning_of(ra.RGI_TUNE_TEMPORAL_WAVES4 | ra.RGI_TUNE_TILE_ORDER)


ATRIUM_DISTANCE = 0.7  # camera inside the courtyard: about 70% of pixels hit geometry


def atrium():
    """The ~256k-triangle atrium with 64 point lights besides its emitters."""
    return ra.HostScene(ra.RGI_SCENE_ATRIUM, seed=1, size=262144, extra_lights=64, light_seed=1)


@pytest.mark.parametrize("frames,kw,variants,scene", [
    (4, {}, {}, "cornell"), (3, {"restir_di": False, "sky": True}, {}, "cornell"),
    (3, {"ray_mode": ra.RGI_RAY_BVH}, {}, "cornell"), (4, {}, ALL_VARIANTS, "cornell"),
    (3, {"shadows": True, "light_color": (1.0, 0.9, 0.8), "light_position": (0.1, 1.2, -0.2),
         "shadow_face_size": 0}, {}, "cornell"),
    (3, {}, ALL_VARIANTS, "atrium"),
    (4, {"gi_divisor": 2}, {}, "cornell"), (3, {"gi_divisor": 4}, ALL_VARIANTS, "cornell"),
    (3, {"gi_divisor": 2}, ALL_VARIANTS, "atrium")])
def test_renderer_matches_oracle_sequence(torch, frames, kw, variants, scene):
    run_renderer_case(torch, frames, kw, variants, scene, W, H)


@pytest.mark.parametrize("width,height,divisor", [(97, 61, 1), (7, 5, 1), (97, 61, 2), (7, 5, 4)])
def test_renderer_matches_oracle_at_ragged_sizes(torch, width, height, divisor):
    """Sizes that are no multiple of the 16 x 16 blocks, the 4 x 8 groups or
    the 8 x 8 insert tiles: every pass's edge handling, all launch variants;
    at divisors 2 and 4 also GI sizes that are floors (48 x 30, 1 x 1)."""
    run_renderer_case(torch, 3, {"gi_divisor": divisor}, ALL_VARIANTS, "cornell", width, height)


def run_renderer_case(torch, frames, kw, variants, scene, W, H):
    say("case renderer scene=%s frames=%d %s variants=%s size=%dx%d" % (scene, frames, kw, variants, W, H))
    hs = atrium() if scene == "atrium" else common.cornell(extra_lights=4)
    distance = ATRIUM_DISTANCE if scene == "atrium" else 3.17
    s = ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC, **kw)
    yaw = 2 * np.pi / 64
    seq = common.oracle_sequence(hs, s, W, H, frames, yaw_step=yaw, distance=distance)
    L = ra.lib()
    bvh, lights = upload(torch, hs.bvh), upload(torch, hs.lights)
    scene = hs.rgi_scene(bvh.data_ptr(), lights.data_ptr())
    r = ctypes.c_void_p()
    ra.check(L.rgi_renderer_create(ctypes.byref(r), W, H, ctypes.byref(ra.c_settings(s)), ctypes.byref(scene),
                                   hs.radius))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    try:
        for opt, value in variants.items():
            ra.check(L.rgi_renderer_set_option(r, opt, value))
        for i, (f, g, hist, want) in enumerate(seq):
            eye, vp, inv = ra.orbit_camera(hs, i, W, H, yaw_step=yaw, distance=distance)
            view = ra.make_view(eye, vp, inv)
            planes = {k: upload(torch, v) for k, v in g.items()}
            gb = ra.GBuffer(W, H, planes["depth"].data_ptr(), planes["albedo"].data_ptr(), planes["normal"].data_ptr(),
                            planes["shading"].data_ptr(), planes["emission"].data_ptr())
            ra.check(L.rgi_renderer_render(r, ctypes.byref(view), ctypes.byref(gb), stream))
            used = ra.Frame()
            ra.check(L.rgi_renderer_last_frame(r, ctypes.byref(used)))
            assert bytes(used) == bytes(f), "frame %d: packed uniforms differ" % i
            out = ra.RendererOutputs()
            ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(out)))
            checks = [("gi_sample", out.gi_sample), ("gi_receiver", out.gi_receiver), ("gi_normal", out.gi_normal),
                      ("color", out.color), ("moments", out.moments), ("grid", out.grid)]
            if f.di_light_count > 0:
                checks += [("diffuse", out.di_diffuse), ("specular", out.di_specular)]
            if s.denoiser:
                checks += [("filtered", out.filtered_history), ("denoised", out.denoised)]
            if s.composite:
                checks += [("hdr", out.hdr)]
            if s.specular_gi:
                checks += [("specular_gi", out.specular_gi)]
            for name, ptr in checks:
                got = download(torch, ptr, want[name])
                if name == "grid":
                    for part, a, b in zip(("checksums", "counts", "entries"), common.grid_view(got, s.grid_table_size),
                                          common.grid_view(want[name], s.grid_table_size)):
                        assert np.array_equal(a, b), "frame %d: grid %s differ" % (i, part)
                else:
                    assert np.array_equal(got, want[name]), "frame %d: %s differs" % (i, name)
    finally:
        L.rgi_renderer_destroy(r)


def test_autotune_keeps_images_bit_exact(torch):
    """rgi_renderer_autotune times every trial (RGI_TUNE_TRIALS of them, each
    > 0 ms), leaves its chosen configuration set on the renderer, and the
    frames rendered afterwards are bit-identical to a default renderer's."""
    say("case autotune")
    L = ra.lib()
    W2, H2 = 160, 96
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC)
    bvh, lights = upload(torch, hs.bvh), upload(torch, hs.lights)
    scene = hs.rgi_scene(bvh.data_ptr(), lights.data_ptr())
    yaw = 2 * np.pi / 64
    views, gbs, keep = [], [], []
    for i in range(4):
        eye, vp, inv = ra.orbit_camera(hs, i, W2, H2, yaw_step=yaw)
        f = common.frame_for(hs, s, W2, H2, i, i > 0, yaw_step=yaw)
        planes = {k: upload(torch, v) for k, v in common.oracle_gbuffer(hs, f).items()}
        keep.append(planes)
        gbs.append(ra.GBuffer(W2, H2, planes["depth"].data_ptr(), planes["albedo"].data_ptr(),
                              planes["normal"].data_ptr(), planes["shading"].data_ptr(), planes["emission"].data_ptr()))
        views.append(ra.make_view(eye, vp, inv))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rs = []
    for _ in range(2):
        r = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(r), W2, H2, ctypes.byref(ra.c_settings(s)), ctypes.byref(scene),
                                       hs.radius))
        rs.append(r)
    try:
        ms = (ctypes.c_float * ra.RGI_TUNE_TRIALS)()
        frames = 2
        chosen = L.rgi_renderer_autotune(rs[0], (ra.View * 4)(*views), (ra.GBuffer * 4)(*gbs), 4, frames, ms, stream)
        assert chosen >= 0, chosen
        assert all(v > 0.0 for v in ms), list(ms)
        # The default renderer replays the frames the trials rendered (one
        # untimed and `frames` timed per trial, cycling the views): the two
        # renderers then hold the same history only if every trial's launch
        # variants rendered the same images.
        for k in range(ra.RGI_TUNE_TRIALS * (1 + frames)):
            ra.check(L.rgi_renderer_render(rs[1], ctypes.byref(views[k % 4]), ctypes.byref(gbs[k % 4]), stream))
        for i in range(4):
            for r in rs:
                ra.check(L.rgi_renderer_render(r, ctypes.byref(views[i]), ctypes.byref(gbs[i]), stream))
            outs = []
            for r in rs:
                o = ra.RendererOutputs()
                ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(o)))
                outs.append(o)
            for name in ("gi_sample", "color", "hdr"):
                like = np.zeros((W2 * H2, 4), np.float32 if name == "gi_sample" else np.uint16)
                a = download(torch, getattr(outs[0], name), like)
                b = download(torch, getattr(outs[1], name), like)
                assert np.array_equal(a, b), "frame %d: %s differs after tuning (config %d)" % (i, name, chosen)
        say("  autotune chose %d, trials %s" % (chosen, [round(v, 4) for v in ms]))
    finally:
        for r in rs:
            L.rgi_renderer_destroy(r)


def test_reference_accumulate_and_error(torch):
    say("case reference accumulate/error")
    hs = common.cornell(extra_lights=4)
    s = ra.Settings()
    seq = common.oracle_sequence(hs, s, W, H, 3, yaw_step=0.0)
    L = ra.lib()
    n = W * H
    want_acc = np.zeros(3 * n, np.float64)
    acc = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    keep = []
    for f, g, hist, out in seq:
        gh = common.host_gbuffer(g, W, H)
        orc.lib().orc_reference_accumulate(ctypes.byref(gh), orc.ptr(out["color"]), orc.ptr(out["diffuse"]),
                                           orc.ptr(out["specular"]), orc.ptr(want_acc))
        planes = {k: upload(torch, v) for k, v in g.items()}
        c, d, sp = upload(torch, out["color"]), upload(torch, out["diffuse"]), upload(torch, out["specular"])
        keep += [planes, c, d, sp]
        gb = ra.GBuffer(W, H, planes["depth"].data_ptr(), planes["albedo"].data_ptr(), planes["normal"].data_ptr(),
                        planes["shading"].data_ptr(), planes["emission"].data_ptr())
        ra.check(L.rgi_reference_accumulate(ctypes.byref(gb), c.data_ptr(), d.data_ptr(), sp.data_ptr(),
                                            acc.data_ptr(), stream))
    got_acc = acc.cpu().numpy()
    assert np.array_equal(got_acc, want_acc)
    ctx = ctypes.c_void_p()
    ra.check(L.rgi_create(ctypes.byref(ctx)))
    res = torch.zeros(5, dtype=torch.float64, device="cuda")
    ra.check(L.rgi_reference_error(ctx, ctypes.byref(gb), c.data_ptr(), d.data_ptr(), sp.data_ptr(), acc.data_ptr(), 3,
                                   res.data_ptr(), stream))
    got = res.cpu().numpy()
    L.rgi_destroy(ctx)
    want = np.zeros(5, np.float64)
    f, g, hist, out = seq[-1]
    gh = common.host_gbuffer(g, W, H)
    orc.lib().orc_reference_error(ctypes.byref(gh), orc.ptr(out["color"]), orc.ptr(out["diffuse"]),
                                  orc.ptr(out["specular"]), orc.ptr(want_acc), 3, orc.ptr(want))
    assert np.array_equal(got, want), (got, want)

    # The same over the composite (the image the reference measures).
    want_acc = np.zeros(3 * n, np.float64)
    acc.zero_()
    for f, g, hist, out in seq:
        gh = common.host_gbuffer(g, W, H)
        orc.lib().orc_reference_accumulate_composite(ctypes.byref(gh), orc.ptr(out["hdr"]), orc.ptr(want_acc))
        planes = {k: upload(torch, v) for k, v in g.items()}
        hdr = upload(torch, out["hdr"])
        keep += [planes, hdr]
        gb = ra.GBuffer(W, H, planes["depth"].data_ptr(), planes["albedo"].data_ptr(), planes["normal"].data_ptr(),
                        planes["shading"].data_ptr(), planes["emission"].data_ptr())
        ra.check(L.rgi_reference_accumulate_composite(ctypes.byref(gb), hdr.data_ptr(), acc.data_ptr(), stream))
    assert np.array_equal(acc.cpu().numpy(), want_acc)
    ctx = ctypes.c_void_p()
    ra.check(L.rgi_create(ctypes.byref(ctx)))
    ra.check(L.rgi_reference_error_composite(ctx, ctypes.byref(gb), hdr.data_ptr(), acc.data_ptr(), 3, res.data_ptr(),
                                             stream))
    got = res.cpu().numpy()
    L.rgi_destroy(ctx)
    orc.lib().orc_reference_error_composite(ctypes.byref(gh), orc.ptr(out["hdr"]), orc.ptr(want_acc), 3,
                                            orc.ptr(want))
    assert np.array_equal(got, want), (got, want)


def test_render_graph_matches_eager(torch):
    """rgi_renderer_render_graph (the frame as one updated-and-replayed HIP
    graph) against rgi_renderer_render: every output plane bit-identical on
    every frame of a camera sequence, with eager work and synchronisations
    between the replays (the first graph frame renders eagerly, the rest are
    captured)."""
    say("case renderer graph vs eager")
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC, shadows=True, light_color=(1.0, 0.9, 0.8),
                    light_position=(0.1, 1.2, -0.2), shadow_face_size=0)
    L = ra.lib()
    bvh, lights = upload(torch, hs.bvh), upload(torch, hs.lights)
    scene = hs.rgi_scene(bvh.data_ptr(), lights.data_ptr())
    rs = []
    for _ in range(2):
        r = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(r), W, H, ctypes.byref(ra.c_settings(s)), ctypes.byref(scene),
                                       hs.radius))
        rs.append(r)
    side = torch.cuda.Stream()
    eager_stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    graph_stream = ctypes.c_void_p(side.cuda_stream)
    yaw = 2 * np.pi / 64
    keep = []
    try:
        for i in range(5):
            eye, vp, inv = ra.orbit_camera(hs, i, W, H, yaw_step=yaw)
            f = common.frame_for(hs, s, W, H, i, i > 0, yaw_step=yaw)
            g = common.oracle_gbuffer(hs, f)
            planes = {k: upload(torch, v) for k, v in g.items()}
            keep.append(planes)
            gb = ra.GBuffer(W, H, planes["depth"].data_ptr(), planes["albedo"].data_ptr(), planes["normal"].data_ptr(),
                            planes["shading"].data_ptr(), planes["emission"].data_ptr())
            view = ra.make_view(eye, vp, inv)
            torch.cuda.synchronize()
            ra.check(L.rgi_renderer_render(rs[0], ctypes.byref(view), ctypes.byref(gb), eager_stream))
            ra.check(L.rgi_renderer_render_graph(rs[1], ctypes.byref(view), ctypes.byref(gb), graph_stream))
            torch.cuda.synchronize()
            scratch = torch.ones(1 << 20, device="cuda") * 2.0  # eager work between replays
            assert float(scratch.sum().item()) == float(1 << 21)
            outs = [ra.RendererOutputs(), ra.RendererOutputs()]
            for r, o in zip(rs, outs):
                ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(o)))
            n = W * H
            for name, like in (("gi_sample", np.zeros((n, 4), np.float32)), ("gi_receiver", np.zeros((n, 4), np.float32)),
                               ("color", np.zeros((n, 4), np.uint16)), ("moments", np.zeros((n, 4), np.uint16)),
                               ("di_diffuse", np.zeros((n, 4), np.uint16)), ("denoised", np.zeros((n, 4), np.uint16)),
                               ("hdr", np.zeros((n, 4), np.uint16)), ("specular_gi", np.zeros((n, 4), np.uint16))):
                a = download(torch, getattr(outs[0], name), like)
                b = download(torch, getattr(outs[1], name), like)
                assert np.array_equal(a, b), "frame %d: %s differs between graph and eager" % (i, name)
    finally:
        for r in rs:
            L.rgi_renderer_destroy(r)


@pytest.mark.parametrize("mode", [ra.RGI_BATCH_PASSES, ra.RGI_BATCH_GRAPH])
def test_batch_matches_single_renderers(torch, mode):
    """rgi_batch_render (renderers of different scenes and sizes: each pass
    over every view in one launch, or one HIP graph with each renderer on its
    own stream) against each renderer rendered alone: bit-identical outputs
    on every frame, also after a batch frame that failed part-way."""
    say("case batch vs single renderers, mode %d" % mode)
    L = ra.lib()
    cases = [(common.cornell(extra_lights=4), 96, 64),
             (ra.HostScene(ra.RGI_SCENE_BOXFIELD, seed=3, size=60, extra_lights=4), 80, 48),
             (atrium(), 64, 64)]
    s = ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC)
    keep, batch_r, single_r, scenes = [], [], [], []
    side = torch.cuda.Stream()
    bs = ctypes.c_void_p(side.cuda_stream)
    es = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    batch = ctypes.c_void_p()
    try:
        for hs, w, h in cases:
            bvh, lights = upload(torch, hs.bvh), upload(torch, hs.lights)
            keep += [bvh, lights]
            scene = hs.rgi_scene(bvh.data_ptr(), lights.data_ptr())
            scenes.append(scene)
            for lst in (batch_r, single_r):
                r = ctypes.c_void_p()
                ra.check(L.rgi_renderer_create(ctypes.byref(r), w, h, ctypes.byref(ra.c_settings(s)),
                                               ctypes.byref(scene), hs.radius))
                lst.append(r)
        arr = (ctypes.c_void_p * len(batch_r))(*[r.value for r in batch_r])
        ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, len(batch_r)))
        ra.check(L.rgi_batch_set_mode(batch, mode))
        yaw = 2 * np.pi / 64
        for i in range(4):
            views, gbs = [], []
            for hs, w, h in cases:
                eye, vp, inv = ra.orbit_camera(hs, i, w, h, yaw_step=yaw)
                f = common.frame_for(hs, s, w, h, i, i > 0, yaw_step=yaw)
                g = common.oracle_gbuffer(hs, f)
                planes = {k: upload(torch, v) for k, v in g.items()}
                keep.append(planes)
                gbs.append(ra.GBuffer(w, h, planes["depth"].data_ptr(), planes["albedo"].data_ptr(),
                                      planes["normal"].data_ptr(), planes["shading"].data_ptr(),
                                      planes["emission"].data_ptr()))
                views.append(ra.make_view(eye, vp, inv))
            torch.cuda.synchronize()
            for r, v, gb in zip(single_r, views, gbs):
                ra.check(L.rgi_renderer_render(r, ctypes.byref(v), ctypes.byref(gb), es))
            va = (ra.View * len(views))(*views)
            ga = (ra.GBuffer * len(gbs))(*gbs)
            if i == 2:
                # A batch frame that fails part-way (the last renderer's
                # G-buffer has the wrong size: found after the others were
                # captured, or planned) leaves every renderer as it was: the
                # retry below must still match the renderers alone.
                bad = list(gbs)
                bad[-1] = ra.GBuffer(cases[-1][1] - 1, cases[-1][2], *(getattr(gbs[-1], k) for k in (
                    "depth", "albedo", "normal", "shading_normal", "emission")))
                assert L.rgi_batch_render(batch, va, (ra.GBuffer * len(bad))(*bad), bs) == ra.RGI_E_INVALID
                torch.cuda.synchronize()
            ra.check(L.rgi_batch_render(batch, va, ga, bs))
            torch.cuda.synchronize()
            for k, ((hs, w, h), rb, rsg) in enumerate(zip(cases, batch_r, single_r)):
                ob, osg = ra.RendererOutputs(), ra.RendererOutputs()
                ra.check(L.rgi_renderer_outputs_get(rb, ctypes.byref(ob)))
                ra.check(L.rgi_renderer_outputs_get(rsg, ctypes.byref(osg)))
                for name in ("gi_sample", "color", "hdr", "specular_gi", "denoised"):
                    like = np.zeros((w * h, 4), np.float32 if name == "gi_sample" else np.uint16)
                    a = download(torch, getattr(ob, name), like)
                    b = download(torch, getattr(osg, name), like)
                    assert np.array_equal(a, b), "frame %d scene %d: %s differs" % (i, k, name)
    finally:
        if batch:
            L.rgi_batch_destroy(batch)
        for r in batch_r + single_r:
            L.rgi_renderer_destroy(r)


def _batch_views(torch, cases, settings, i, yaw, keep):
    views, gbs = [], []
    for (hs, w, h, div), s in zip(cases, settings):
        eye, vp, inv = ra.orbit_camera(hs, i, w, h, yaw_step=yaw)
        f = common.frame_for(hs, s, w, h, i, i > 0, yaw_step=yaw)
        g = common.oracle_gbuffer(hs, f)
        planes = {k: upload(torch, v) for k, v in g.items()}
        keep.append(planes)
        gbs.append(ra.GBuffer(w, h, planes["depth"].data_ptr(), planes["albedo"].data_ptr(),
                              planes["normal"].data_ptr(), planes["shading"].data_ptr(), planes["emission"].data_ptr()))
        views.append(ra.make_view(eye, vp, inv))
    return views, gbs


def _batch_and_single(torch, L, cases, settings, keep):
    batch_r, single_r = [], []
    for (hs, w, h, div), s in zip(cases, settings):
        bvh, lights = upload(torch, hs.bvh), upload(torch, hs.lights)
        keep += [bvh, lights]
        scene = hs.rgi_scene(bvh.data_ptr(), lights.data_ptr())
        for lst in (batch_r, single_r):
            r = ctypes.c_void_p()
            ra.check(L.rgi_renderer_create(ctypes.byref(r), w, h, ctypes.byref(ra.c_settings(s)),
                                           ctypes.byref(scene), hs.radius))
            lst.append(r)
    return batch_r, single_r


def test_batch_mixes_gi_sizes(torch):
    """A pass-major batch whose views have GI at the window size and at the
    reference's half and quarter size (both forms of every batched GI kernel
    in one frame, the box field's DI through the light guide), against each
    renderer alone: bit-identical outputs on every frame."""
    say("case batch with mixed GI sizes")
    L = ra.lib()
    cases = [(common.cornell(extra_lights=4), 96, 64, 1),
             (ra.HostScene(ra.RGI_SCENE_BOXFIELD, seed=3, size=60, extra_lights=300), 80, 48, 2),
             (atrium(), 128, 96, 4),
             (common.cornell(extra_lights=4), 64, 64, 2)]
    keep = []
    settings = [ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC, gi_divisor=div) for _, _, _, div in cases]
    es = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    side = torch.cuda.Stream()
    bs = ctypes.c_void_p(side.cuda_stream)
    batch = ctypes.c_void_p()
    batch_r, single_r = [], []
    try:
        batch_r, single_r = _batch_and_single(torch, L, cases, settings, keep)
        arr = (ctypes.c_void_p * len(batch_r))(*[r.value for r in batch_r])
        ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, len(batch_r)))
        yaw = 2 * np.pi / 64
        for i in range(3):
            views, gbs = _batch_views(torch, cases, settings, i, yaw, keep)
            torch.cuda.synchronize()
            for r, v, gb in zip(single_r, views, gbs):
                ra.check(L.rgi_renderer_render(r, ctypes.byref(v), ctypes.byref(gb), es))
            ra.check(L.rgi_batch_render(batch, (ra.View * len(views))(*views), (ra.GBuffer * len(gbs))(*gbs), bs))
            torch.cuda.synchronize()
            for k, ((hs, w, h, div), rb, rsg) in enumerate(zip(cases, batch_r, single_r)):
                ob, osg = ra.RendererOutputs(), ra.RendererOutputs()
                ra.check(L.rgi_renderer_outputs_get(rb, ctypes.byref(ob)))
                ra.check(L.rgi_renderer_outputs_get(rsg, ctypes.byref(osg)))
                gw, gh = max(w // div, 1), max(h // div, 1)
                for name, n, dt in (("gi_sample", gw * gh, np.float32), ("color", gw * gh, np.uint16),
                                    ("di_diffuse", w * h, np.uint16), ("hdr", w * h, np.uint16),
                                    ("denoised", gw * gh, np.uint16)):
                    like = np.zeros((n, 4), dt)
                    a = download(torch, getattr(ob, name), like)
                    b = download(torch, getattr(osg, name), like)
                    assert np.array_equal(a, b), "frame %d view %d (divisor %d): %s differs" % (i, k, div, name)
    finally:
        if batch:
            L.rgi_batch_destroy(batch)
        for r in batch_r + single_r:
            L.rgi_renderer_destroy(r)


def test_batch_atomic_grids(torch):
    """The batched atomic insert (one launch for every view's grid) against
    each renderer's own atomic insert on the same G-buffers, first frame:
    the same cells with the same checksums and counts, and each cell of at
    most 32 pixels lists the same pixels (the order of the atomics places
    them in any order, and decides which pixels a fuller cell keeps)."""
    say("case batch atomic grids")
    L = ra.lib()
    cases = [(common.cornell(extra_lights=4), 96, 64, 1), (atrium(), 128, 96, 2), (common.cornell(), 64, 48, 4)]
    settings = [ra.Settings(grid_mode=ra.RGI_GRID_ATOMIC, gi_divisor=div) for _, _, _, div in cases]
    keep = []
    es = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    batch = ctypes.c_void_p()
    batch_r, single_r = [], []
    try:
        batch_r, single_r = _batch_and_single(torch, L, cases, settings, keep)
        arr = (ctypes.c_void_p * len(batch_r))(*[r.value for r in batch_r])
        ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, len(batch_r)))
        views, gbs = _batch_views(torch, cases, settings, 0, 0.0, keep)
        torch.cuda.synchronize()
        for r, v, gb in zip(single_r, views, gbs):
            ra.check(L.rgi_renderer_render(r, ctypes.byref(v), ctypes.byref(gb), es))
        ra.check(L.rgi_batch_render(batch, (ra.View * len(views))(*views), (ra.GBuffer * len(gbs))(*gbs), es))
        torch.cuda.synchronize()
        T = settings[0].grid_table_size
        for k, (rb, rsg) in enumerate(zip(batch_r, single_r)):
            cells = []
            for r in (rb, rsg):
                o = ra.RendererOutputs()
                ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(o)))
                cs, cnt, entries = common.grid_view(download(torch, o.grid, np.zeros(34 * T, np.uint32)), T)
                occ = np.nonzero(cs)[0]
                cells.append({int(cs[s_]): (int(cnt[s_]), tuple(sorted(entries[s_, :cnt[s_]])) if cnt[s_] <= 32 else None)
                              for s_ in occ})
            assert cells[0] == cells[1], "view %d: the batched atomic insert's cells differ" % k
            assert len(cells[0]) > 0
    finally:
        if batch:
            L.rgi_batch_destroy(batch)
        for r in batch_r + single_r:
            L.rgi_renderer_destroy(r)
