"""The neural radiance cache's HIP kernels against the oracle (oracle/nrc.cpp).

Inference in BF16 is compared with the oracle's emulated mode (the same
rounding points; the kernels differ only in the order of FP32 accumulation
inside the matrix cores), FP16 and FP8 with its exact mode. The gradients of
one batch are compared layer by layer and over the table; the Adam step is
checked bit for bit on the MLP weights against the oracle's FP32 Adam applied
to the GPU's own gradients. Training is shown to reduce the loss and the
error on held-out points.
"""
import ctypes

import numpy as np
import pytest

import orc
import restir_amd as ra
from restir_amd import nrc as rnrc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def setup_cache(seed, log2_table=14, table_scale=0.5):
    rng = np.random.default_rng(seed)
    nc = rnrc.NeuralCache(log2_table=log2_table, seed=seed, aabb_min=(-1, -1, -1), aabb_max=(1, 1, 1))
    mlp = np.empty(ra.RGI_NRC_MLP_PARAMS, np.float32)
    mlp[:20480] = rng.uniform(-1, 1, 20480) * np.sqrt(6.0 / 128.0)
    mlp[20480:] = rng.uniform(-1, 1, 192) * np.sqrt(6.0 / 67.0)
    table = rng.uniform(-table_scale, table_scale, nc.table_floats).astype(np.float32)
    nc.set_params(torch.from_numpy(mlp).cuda(), torch.from_numpy(table).cuda())
    res = rnrc.level_resolutions(nc.cfg)
    model = orc.NrcModel(res, log2_table, [-1, -1, -1], [1, 1, 1], mlp, table)
    return nc, model, rng


def records(rng, n):
    rec = np.zeros((n, 12), np.float32)
    rec[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    nrm = rng.normal(size=(n, 3))
    rec[:, 4:7] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    rec[:, 8:11] = rng.uniform(0, 1, (n, 3))
    return rec


def targets(rng, n):
    t = np.zeros((n, 4), np.float32)
    t[:, 0:3] = rng.uniform(0, 2, (n, 3))
    t[:, 3] = (rng.uniform(size=n) < 0.9).astype(np.float32)
    return t


def rel(a, b, floor):
    return np.abs(a - b) / (np.abs(b) + floor)


@pytest.mark.parametrize("count", [4096, 1000, 17])
def test_nrc_infer_bf16_matches_oracle(count):
    nc, model, rng = setup_cache(11)
    rec = records(rng, count)
    print("nrc infer bf16 count", count, flush=True)
    out = nc.infer(torch.from_numpy(rec).cuda(), ra.RGI_NRC_BF16).cpu().numpy()
    want = model.forward(rec, emulate=1)
    assert np.all(out[:, 3] == 0.0)
    r = rel(out[:, :3].astype(np.float64), want, 0.05)
    assert np.median(r) < 2e-3, np.median(r)
    assert np.quantile(r, 0.99) < 3e-2, np.quantile(r, 0.99)


def test_nrc_infer_fp16_fp8_close_to_exact():
    nc, model, rng = setup_cache(12)
    rec = records(rng, 8192)
    exact = model.forward(rec, emulate=0)
    scale = np.sqrt(np.mean(exact ** 2))
    # FP8 activation scales come from a training batch: take one step first,
    # then compare against the function inference evaluates, that of the
    # parameters' moving average.
    tg = targets(rng, 8192)
    nc.train(torch.from_numpy(rec).cuda(), torch.from_numpy(tg).cuda())
    mlp, table = nc.average()
    model.mlp[:] = mlp.cpu().numpy()
    model.table[:] = table.cpu().numpy()
    exact = model.forward(rec, emulate=0)
    err = {}
    for name, prec in (("bf16", ra.RGI_NRC_BF16), ("fp16", ra.RGI_NRC_FP16), ("fp8", ra.RGI_NRC_FP8)):
        print("nrc infer", name, flush=True)
        out = nc.infer(torch.from_numpy(rec).cuda(), prec).cpu().numpy()[:, :3].astype(np.float64)
        err[name] = np.sqrt(np.mean((out - exact) ** 2)) / scale
    print("relative RMS error against the exact function:", err)
    assert err["fp16"] < 2e-3
    assert err["bf16"] < 2e-2
    assert err["fp8"] < 0.15
    assert err["fp16"] < err["bf16"] < err["fp8"]


def test_nrc_gradients_match_oracle():
    nc, model, rng = setup_cache(13)
    n = 2048
    rec, tg = records(rng, n), targets(rng, n)
    print("nrc gradients", n, flush=True)
    loss, mg, tgrad = nc.gradients(torch.from_numpy(rec).cuda(), torch.from_numpy(tg).cuda())
    loss, mg, tgrad = float(loss.cpu()[0]), mg.cpu().numpy().astype(np.float64), tgrad.cpu().numpy().astype(np.float64)
    wl, wm, wt = model.gradients(rec, tg, emulate=1)
    assert abs(loss - wl) < 1e-3 * wl, (loss, wl)
    for k in range(6):
        sl = slice(k * 4096, k * 4096 + 4096) if k < 5 else slice(20480, 20672)
        e = np.linalg.norm(mg[sl] - wm[sl]) / np.linalg.norm(wm[sl])
        assert e < 2e-2, (k, e)
    assert np.linalg.norm(tgrad - wt) < 5e-2 * np.linalg.norm(wt)
    # The same entries are touched.
    assert np.array_equal(tgrad != 0.0, wt != 0.0)


def test_nrc_adam_step_matches_oracle():
    nc, model, rng = setup_cache(14)
    n = 3000
    rec, tg = torch.from_numpy(records(rng, n)).cuda(), torch.from_numpy(targets(rng, n)).cuda()
    _, mg, tgrad = nc.gradients(rec, tg)
    mlp0, table0 = nc.params()
    print("nrc train step", flush=True)
    nc.train(rec, tg)
    assert nc.steps() == 1
    mlp1, table1 = nc.params()
    cfg = nc.cfg
    p = mlp0.cpu().numpy().copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    g = mg.cpu().numpy()
    c1, c2 = np.float32(1 - cfg.beta1), np.float32(1 - cfg.beta2)
    orc.lib().orc_nrc_adam(orc.ptr(p), orc.ptr(m), orc.ptr(v), orc.ptr(g), p.size, cfg.learning_rate, cfg.beta1,
                           cfg.beta2, cfg.epsilon, float(np.float32(1 - np.float64(cfg.beta1))),
                           float(np.float32(1 - np.float64(cfg.beta2))), 0)
    assert np.array_equal(p.view(np.uint32), mlp1.cpu().numpy().view(np.uint32))
    # Table: untouched entries keep their value bit for bit; touched ones move
    # by the learning rate against the gradient's sign.
    t0, t1, gt = table0.cpu().numpy(), table1.cpu().numpy(), tgrad.cpu().numpy()
    untouched = gt == 0.0
    assert np.array_equal(t0[untouched].view(np.uint32), t1[untouched].view(np.uint32))
    big = np.abs(gt) > 1e-5
    step = (t1[big].astype(np.float64) - t0[big]) / cfg.learning_rate
    assert np.allclose(step, -np.sign(gt[big]), atol=2e-3)


def smooth_radiance(rec):
    p, n = rec[:, 0:3], rec[:, 4:7]
    r = 0.6 + 0.4 * np.sin(2.0 * p[:, 0]) * np.cos(1.5 * p[:, 1]) + 0.2 * n[:, 1]
    g = 0.5 + 0.3 * np.cos(3.0 * p[:, 2]) + 0.2 * p[:, 0] * p[:, 1]
    b = 0.4 + 0.3 * rec[:, 8] + 0.2 * np.sin(4.0 * p[:, 1] + p[:, 2])
    return np.stack([r, g, b], axis=1).astype(np.float32)


def test_nrc_moving_average_follows_its_closed_form():
    """After set_params the moving average is the parameters; each training
    step then moves it to d e + (1 - d) w with d = min(ema_decay, (1 + t) /
    (10 + t)) at step t, both products rounded in FP32 and then summed, bit
    for bit; and inference reads it: BF16 and FP16 predictions equal those of
    a cache without the average given the averaged parameters."""
    nc, model, rng = setup_cache(31)
    decay = np.float32(nc.cfg.ema_decay)
    assert decay > 0
    pm, pt = nc.params()
    em, et = nc.average()
    assert torch.equal(pm.view(torch.int32), em.view(torch.int32))
    assert torch.equal(pt.view(torch.int32), et.view(torch.int32))
    prev_m, prev_t = em.cpu().numpy(), et.cpu().numpy()
    rec = torch.from_numpy(records(rng, 4096)).cuda()
    tgt = torch.from_numpy(targets(rng, 4096)).cuda()
    for t in (1, 2):
        nc.train(rec, tgt)
        wm, wt = (x.cpu().numpy() for x in nc.params())
        em, et = (x.cpu().numpy() for x in nc.average())
        d = np.minimum(decay, np.float32(1.0 + t) / np.float32(10.0 + t))
        om = np.float32(1.0) - d
        want_m = d * prev_m + om * wm
        want_t = d * prev_t + om * wt
        assert want_m.dtype == np.float32 and want_t.dtype == np.float32
        assert np.array_equal(em.view(np.uint32), want_m.view(np.uint32)), t
        assert np.array_equal(et.view(np.uint32), want_t.view(np.uint32)), t
        assert not np.array_equal(em, wm)  # the average lags the parameters
        prev_m, prev_t = em, et
    plain = rnrc.NeuralCache(log2_table=nc.cfg.log2_table, seed=31, aabb_min=(-1, -1, -1), aabb_max=(1, 1, 1),
                             ema_decay=0.0)
    plain.set_params(torch.from_numpy(prev_m).cuda(), torch.from_numpy(prev_t).cuda())
    q = torch.from_numpy(records(rng, 1000)).cuda()
    for prec in (ra.RGI_NRC_BF16, ra.RGI_NRC_FP16):
        a, b = nc.infer(q, prec), plain.infer(q, prec)
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), prec
    plain.close()
    nc.close()


def test_nrc_training_learns_a_smooth_field():
    rng = np.random.default_rng(15)
    nc = rnrc.NeuralCache(log2_table=16, seed=3, learning_rate=0.01, aabb_min=(-1, -1, -1), aabb_max=(1, 1, 1))
    held = records(rng, 8192)
    held_t = smooth_radiance(held)

    def held_error():
        out = nc.infer(torch.from_numpy(held).cuda()).cpu().numpy()[:, :3]
        return float(np.mean((out - held_t) ** 2) / np.mean(held_t ** 2))

    before = held_error()
    losses = []
    for step in range(300):
        rec = records(rng, 16384)
        t = np.concatenate([smooth_radiance(rec), np.ones((16384, 1), np.float32)], axis=1)
        losses.append(float(nc.train(torch.from_numpy(rec).cuda(), torch.from_numpy(t).cuda()).cpu()[0]))
    after = held_error()
    print("held-out relative MSE before %.4f after %.5f; loss %.4f -> %.5f" % (before, after, losses[0], losses[-1]))
    assert np.mean(losses[-20:]) < 0.05 * losses[0]
    assert after < 0.02 * before


# ------------------------------------------------------------ in the renderer

def _half_plane(ptr, n):
    """An RGBA16F device plane the library owns, as float32 (n, 4)."""
    out = torch.empty(n * 4, dtype=torch.float16, device="cuda")
    hip = ra.lib()
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    assert hip.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), n * 8, 3) == 0
    return out.float().cpu().numpy().reshape(n, 4)


def _setup_scene(W, H, settings):
    from restir_amd import driver
    hs = ra.HostScene(ra.RGI_SCENE_CORNELL, seed=1, extra_lights=4, light_seed=1)
    gscene = driver.GpuScene(hs)
    ctx = ctypes.c_void_p()
    ra.check(ra.lib().rgi_create(ctypes.byref(ctx)))
    eye, vp, inv = ra.orbit_camera(hs, 0, W, H, yaw_of=ra.swing_yaw)
    f = ra.make_frame(settings, hs, 0, eye, vp, inv, eye, vp, W, H, False)
    gbuf = driver.GBufferSet(ctx, gscene, f)
    return hs, gscene, ctx, gbuf, ra.make_view(eye, vp, inv)


def _renderer(hs, gscene, settings, W, H, nrc=None):
    r = ctypes.c_void_p()
    ra.check(ra.lib().rgi_renderer_create(ctypes.byref(r), W, H, ctypes.byref(ra.c_settings(settings)),
                                          ctypes.byref(gscene.scene), hs.radius))
    if nrc is not None:
        ra.check(ra.lib().rgi_renderer_set_nrc(r, ctypes.byref(nrc)))
    return r


def test_renderer_cache_falls_back_to_plain_path():
    """Until the cache has trained `warmup` frames, and while the reference
    accumulates, a renderer with a cache renders the plain path bit for bit."""
    from restir_amd import driver
    W, H = 320, 240
    st = ra.Settings(grid_table_size=1 << 16, grid_mode=ra.RGI_GRID_DETERMINISTIC)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    s = driver.stream()
    plain = _renderer(hs, gscene, st, W, H)
    cached = _renderer(hs, gscene, st, W, H, rnrc.renderer_settings(hs, batch=4096, warmup=3))
    out_p, out_c = ra.RendererOutputs(), ra.RendererOutputs()
    try:
        for frame in range(3):
            print("fallback frame", frame, flush=True)
            for r in (plain, cached):
                ra.check(L.rgi_renderer_render(r, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
            ra.check(L.rgi_renderer_outputs_get(plain, ctypes.byref(out_p)))
            ra.check(L.rgi_renderer_outputs_get(cached, ctypes.byref(out_c)))
            for plane in ("color", "hdr"):
                a = _half_plane(getattr(out_p, plane), W * H)
                b = _half_plane(getattr(out_c, plane), W * H)
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (frame, plane)
        assert rnrc.renderer_state(cached)[3] == 3
        # Reference accumulation never uses or trains the cache.
        for r in (plain, cached):
            ra.check(L.rgi_renderer_set_reference_mode(r, 1))
            ra.check(L.rgi_renderer_render(r, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        ra.check(L.rgi_renderer_outputs_get(plain, ctypes.byref(out_p)))
        ra.check(L.rgi_renderer_outputs_get(cached, ctypes.byref(out_c)))
        a, b = _half_plane(out_p.color, W * H), _half_plane(out_c.color, W * H)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert rnrc.renderer_state(cached)[3] == 3
    finally:
        L.rgi_renderer_destroy(plain)
        L.rgi_renderer_destroy(cached)
        L.rgi_destroy(ctx)


def test_renderer_cache_beats_the_surface_light_it_replaces():
    """What the trained cache gains: the GI passes read, at an on-screen
    sample point, the DI light plus last frame's indirect light (the
    reference), or the cache's radiance. Against the mean of that quantity
    over 48 frames of the plain renderer, the trained cache's radiance has a
    several times lower error than one frame's plain value."""
    from restir_amd import driver
    W, H = 320, 240
    n = W * H
    st = ra.Settings(grid_table_size=1 << 16)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    s = driver.stream()
    depth = torch.zeros(n, dtype=torch.float32, device="cuda")
    depth.copy_(gbuf.depth.view(torch.float32))
    surface = depth.cpu().numpy() < 0.99999
    alb = gbuf.albedo.view(torch.int32).cpu().numpy().view(np.uint32)
    nrm = gbuf.normal.view(torch.int32).cpu().numpy().view(np.uint32)
    albedo = np.stack([(alb >> (8 * c)) & 255 for c in range(3)], axis=1) / 255.0
    albedo = albedo * (1.0 - ((nrm >> 24) & 255) / 255.0)[:, None]
    plain = _renderer(hs, gscene, st, W, H)
    cached = _renderer(hs, gscene, st, W, H, rnrc.renderer_settings(hs, batch=16384, warmup=4))
    out = ra.RendererOutputs()
    try:
        frames = []
        for frame in range(64):
            ra.check(L.rgi_renderer_render(plain, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
            if frame >= 16:
                ra.check(L.rgi_renderer_outputs_get(plain, ctypes.byref(out)))
                di = _half_plane(out.di_diffuse, n)[:, :3]
                hist = _half_plane(out.filtered_history, n)[:, :3]
                frames.append(di + albedo * hist)
        mean = np.mean(frames, axis=0)
        one = frames[-1]
        for frame in range(64):
            ra.check(L.rgi_renderer_render(cached, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        _, rad_ptr, _, trained = rnrc.renderer_state(cached)
        assert trained == 64
        rad = _half_plane(rad_ptr, n)
        assert np.all(np.isfinite(rad)) and np.all(rad[:, :3] >= 0)
        assert np.all(rad[~surface] == 0) and np.all(rad[surface, 3] == 1)
        err_plain = np.mean((one[surface] - mean[surface]) ** 2)
        err_cache = np.mean((rad[surface, :3] - mean[surface]) ** 2)
        print("surface radiance MSE against the 48-frame mean: one plain frame %.3e, cache %.3e (ratio %.1f)"
              % (err_plain, err_cache, err_plain / err_cache), flush=True)
        assert err_cache < 0.5 * err_plain
    finally:
        L.rgi_renderer_destroy(plain)
        L.rgi_renderer_destroy(cached)
        L.rgi_destroy(ctx)


def test_renderer_resolve_lowers_composite_error():
    """What the resolved cache gains on the frame's image: against an FP64
    accumulation of 1024 reference frames at a static camera, the composite
    of a renderer whose cache resolves each pixel's diffuse light has a
    several times lower MSE than the plain renderer's after 96 frames (one
    frame's DI and GI noise against the cache's average over every frame it
    has learnt from). Before the cache is warm the image is the plain
    path's, bit for bit."""
    from restir_amd import driver
    W, H = 320, 240
    n = W * H
    ref_frames, frames = 1024, 96
    st = ra.Settings(grid_table_size=1 << 16, grid_mode=ra.RGI_GRID_DETERMINISTIC)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    s = driver.stream()
    acc = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    res = torch.zeros(5, dtype=torch.float64, device="cuda")
    out_p, out_c = ra.RendererOutputs(), ra.RendererOutputs()
    ref = _renderer(hs, gscene, st, W, H)
    plain = _renderer(hs, gscene, st, W, H)
    cached = _renderer(hs, gscene, st, W, H, rnrc.renderer_settings(hs, batch=16384, warmup=4, resolve=1))
    try:
        ra.check(L.rgi_renderer_set_reference_mode(ref, 1))
        for _ in range(ref_frames):
            ra.check(L.rgi_renderer_render(ref, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
            ra.check(L.rgi_renderer_outputs_get(ref, ctypes.byref(out_p)))
            ra.check(L.rgi_reference_accumulate_composite(ctypes.byref(gbuf.gb), out_p.hdr, acc.data_ptr(), s))
        for frame in range(frames):
            for r in (plain, cached):
                ra.check(L.rgi_renderer_render(r, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
            if frame < 4:
                ra.check(L.rgi_renderer_outputs_get(plain, ctypes.byref(out_p)))
                ra.check(L.rgi_renderer_outputs_get(cached, ctypes.byref(out_c)))
                a, b = _half_plane(out_p.hdr, n), _half_plane(out_c.hdr, n)
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), frame
        mse = {}
        for name, r in (("plain", plain), ("resolved", cached)):
            ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(out_p)))
            hdr = _half_plane(out_p.hdr, n)
            assert np.all(np.isfinite(hdr))
            ra.check(L.rgi_reference_error_composite(ctx, ctypes.byref(gbuf.gb), out_p.hdr, acc.data_ptr(),
                                                     ref_frames, res.data_ptr(), s))
            r5 = res.cpu().numpy()
            mse[name] = float(r5[1])
            print("%s: composite MSE %.4e, mean luminance %.5f (reference %.5f)" % (name, r5[1], r5[2], r5[3]),
                  flush=True)
            if name == "resolved":
                assert abs(r5[2] - r5[3]) < 0.1 * r5[3]
        print("plain / resolved MSE: %.1f" % (mse["plain"] / mse["resolved"]), flush=True)
        assert mse["resolved"] < 0.5 * mse["plain"]
    finally:
        for r in (ref, plain, cached):
            L.rgi_renderer_destroy(r)
        L.rgi_destroy(ctx)


def test_renderer_cache_in_graph_frames():
    """Graph frames with a resolving cache: the training step forks onto the
    cache's stream and joins inside each captured frame, so every frame
    trains once, the state read afterwards is the last frame's, and the
    resolved image is finite and close to the eager renderer's mean (the
    table gradient's FP32 atomics make the two caches differ in their last
    bits, so the check is statistical)."""
    from restir_amd import driver
    W, H = 160, 120
    n = W * H
    st = ra.Settings(grid_table_size=1 << 16, grid_mode=ra.RGI_GRID_DETERMINISTIC)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    s = driver.stream()
    settings = rnrc.renderer_settings(hs, batch=4096, warmup=2, resolve=1)
    side = torch.cuda.Stream()  # capture needs a stream other than the null stream
    gs = ctypes.c_void_p(side.cuda_stream)
    torch.cuda.synchronize()
    graph = _renderer(hs, gscene, st, W, H, settings)
    eager = _renderer(hs, gscene, st, W, H, settings)
    out = ra.RendererOutputs()
    try:
        frames = 12
        for frame in range(frames):
            ra.check(L.rgi_renderer_render_graph(graph, ctypes.byref(view), ctypes.byref(gbuf.gb), gs))
            ra.check(L.rgi_renderer_render(eager, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        torch.cuda.synchronize()
        means = []
        for r in (graph, eager):
            cache, rad_ptr, loss_ptr, trained = rnrc.renderer_state(r)
            assert cache and rad_ptr and trained == frames
            loss = torch.empty(1, dtype=torch.float32, device="cuda")
            L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            assert L.hipMemcpy(loss.data_ptr(), loss_ptr, 4, 3) == 0
            assert np.isfinite(float(loss.item())) and float(loss.item()) > 0
            ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(out)))
            hdr = _half_plane(out.hdr, n)
            assert np.all(np.isfinite(hdr))
            means.append(hdr[:, :3].mean())
        print("graph / eager resolved mean: %.5f / %.5f" % tuple(means), flush=True)
        assert abs(means[0] - means[1]) < 0.05 * means[1]
    finally:
        L.rgi_renderer_destroy(graph)
        L.rgi_renderer_destroy(eager)
        L.rgi_destroy(ctx)


def test_gbuffer_query_matches_records_query():
    """The per-pixel query reads a pixel's inputs exactly as the training
    gather records them: gathering one record per surface pixel (batch = the
    surface count puts record i on the i-th surface pixel in row-major order)
    and querying those records gives the G-buffer query's radiance at those
    pixels, bit for bit after its clamp and FP16 store, and 0 on the sky; the
    records query is itself checked against the oracle above."""
    from restir_amd import driver
    W, H = 320, 240
    n = W * H
    st = ra.Settings(grid_table_size=1 << 16)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    s = driver.stream()
    eye, vp, inv = ra.orbit_camera(hs, 0, W, H, yaw_of=ra.swing_yaw)
    frame = ra.make_frame(st, hs, 0, eye, vp, inv, eye, vp, W, H, False)
    cfg = rnrc.renderer_settings(hs).cache
    nc = rnrc.NeuralCache(cfg)
    rng = np.random.default_rng(21)
    mlp = np.empty(ra.RGI_NRC_MLP_PARAMS, np.float32)
    mlp[:20480] = rng.uniform(-1, 1, 20480) * np.sqrt(6.0 / 128.0)
    mlp[20480:] = rng.uniform(-1, 1, 192) * np.sqrt(6.0 / 67.0)
    table = rng.uniform(-0.5, 0.5, nc.table_floats).astype(np.float32)
    nc.set_params(torch.from_numpy(mlp).cuda(), torch.from_numpy(table).cuda())
    surface = gbuf.depth.view(torch.float32) < 0.99999
    ns = int(surface.sum())
    assert n // 10 < ns < n  # the box covers about a fifth of this view
    rec = torch.empty((ns, 12), dtype=torch.float32, device="cuda")
    tgt = torch.empty((ns, 4), dtype=torch.float32, device="cuda")
    ra.check(L.rgi_nrc_gather_gbuffer(nc.handle, ctypes.byref(frame), ctypes.byref(gbuf.gb), None, None, ns, 5,
                                      rec.data_ptr(), tgt.data_ptr(), s))
    assert torch.all(tgt[:, 3] == 1)
    for prec in (ra.RGI_NRC_BF16, ra.RGI_NRC_FP16, ra.RGI_NRC_FP8):
        print("gbuffer query precision", prec, flush=True)
        rad = torch.empty(n * 4, dtype=torch.float16, device="cuda")
        ra.check(L.rgi_nrc_query_gbuffer(nc.handle, ctypes.byref(frame), ctypes.byref(gbuf.gb), None, prec,
                                         rad.data_ptr(), s))
        y = nc.infer(rec, prec)
        want = torch.clamp(y[:, :3], min=0.0).half()
        got = rad.view(n, 4)
        assert torch.equal(got[surface][:, :3].view(torch.int16), want.view(torch.int16)), prec
        assert torch.all(got[surface][:, 3] == 1)
        assert torch.all(got[~surface] == 0)
    nc.close()
    L.rgi_destroy(ctx)


def test_gather_stratifies_over_surface_pixels():
    """A smaller batch: every record is a surface pixel's record (one of the
    full gather's), records follow the surface pixels' row-major order, and
    the stratification leaves no gap wider than two strata; a G-buffer with
    no surface pixel gives weight-0 records."""
    from restir_amd import driver
    W, H = 320, 240
    st = ra.Settings(grid_table_size=1 << 16)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    s = driver.stream()
    eye, vp, inv = ra.orbit_camera(hs, 0, W, H, yaw_of=ra.swing_yaw)
    frame = ra.make_frame(st, hs, 0, eye, vp, inv, eye, vp, W, H, False)
    nc = rnrc.NeuralCache(rnrc.renderer_settings(hs).cache)
    ns = int((gbuf.depth.view(torch.float32) < 0.99999).sum())
    full = torch.empty((ns, 12), dtype=torch.float32, device="cuda")
    full_t = torch.empty((ns, 4), dtype=torch.float32, device="cuda")
    ra.check(L.rgi_nrc_gather_gbuffer(nc.handle, ctypes.byref(frame), ctypes.byref(gbuf.gb), None, None, ns, 1,
                                      full.data_ptr(), full_t.data_ptr(), s))
    batch = 1000
    rec = torch.empty((batch, 12), dtype=torch.float32, device="cuda")
    tgt = torch.empty((batch, 4), dtype=torch.float32, device="cuda")
    ra.check(L.rgi_nrc_gather_gbuffer(nc.handle, ctypes.byref(frame), ctypes.byref(gbuf.gb), None, None, batch, 9,
                                      rec.data_ptr(), tgt.data_ptr(), s))
    assert torch.all(tgt[:, 3] == 1)
    f, r = full.cpu().numpy(), rec.cpu().numpy()
    index = {row.tobytes(): i for i, row in enumerate(f)}
    assert len(index) == ns  # surface pixels have distinct records
    picked = np.array([index[row.tobytes()] for row in r])
    assert np.all(np.diff(picked) >= 0)
    stratum = ns / batch
    assert picked[0] < stratum and picked[-1] >= ns - stratum - 1
    assert np.max(np.diff(picked)) < 2 * stratum + 1
    # No surface at all: a G-buffer whose depth is the far plane everywhere.
    sky = torch.ones_like(gbuf.depth.view(torch.float32))
    gb = ra.GBuffer(W, H, sky.data_ptr(), gbuf.albedo.data_ptr(), gbuf.normal.data_ptr(), gbuf.shading.data_ptr(),
                    gbuf.emission.data_ptr())
    ra.check(L.rgi_nrc_gather_gbuffer(nc.handle, ctypes.byref(frame), ctypes.byref(gb), None, None, batch, 9,
                                      rec.data_ptr(), tgt.data_ptr(), s))
    torch.cuda.synchronize()
    assert torch.all(tgt[:, 3] == 0)
    nc.close()
    L.rgi_destroy(ctx)


def test_gather_at_path_vertices():
    """Records at ReSTIR GI path vertices: with one record per surface pixel,
    a pixel whose reservoir holds the surface point of another pixel (its
    reconstructed position, as the pixel gather records it) gives that
    pixel's record, target included; a pixel whose reservoir is empty (M = 0),
    holds a sky sample or a point that the depth buffer does not show keeps
    its own record."""
    from restir_amd import driver
    W, H = 320, 240
    n = W * H
    st = ra.Settings(grid_table_size=1 << 16)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    s = driver.stream()
    eye, vp, inv = ra.orbit_camera(hs, 0, W, H, yaw_of=ra.swing_yaw)
    frame = ra.make_frame(st, hs, 0, eye, vp, inv, eye, vp, W, H, False)
    nc = rnrc.NeuralCache(rnrc.renderer_settings(hs).cache)
    depth = gbuf.depth.view(torch.float32)
    surface = (depth < 0.99999).cpu().numpy()
    ns = int(surface.sum())
    # Targets from a colour plane that differs per pixel, so that a record's
    # target names its pixel as well as its inputs do.
    color = torch.from_numpy(np.random.default_rng(3).uniform(0, 1, (n, 4)).astype(np.float16)).cuda()
    full = torch.empty((ns, 12), dtype=torch.float32, device="cuda")
    full_t = torch.empty((ns, 4), dtype=torch.float32, device="cuda")
    ra.check(L.rgi_nrc_gather_gbuffer(nc.handle, ctypes.byref(frame), ctypes.byref(gbuf.gb), None, color.data_ptr(), ns,
                                      1, full.data_ptr(), full_t.data_ptr(), s))
    torch.cuda.synchronize()
    f = full.cpu().numpy()
    pix = np.flatnonzero(surface)                  # record i <-> surface pixel pix[i]
    rng = np.random.default_rng(7)
    perm = rng.permutation(ns)                     # pixel pix[i]'s sample: pixel pix[perm[i]]'s point
    kind = rng.integers(0, 4, ns)                  # 0, 1: moved; 2: empty reservoir; 3: sky sample
    sample = np.zeros((n, 4), np.float32)
    normal = np.zeros((n, 4), np.float32)
    sample[pix, :3] = f[perm, :3]
    sample[pix, 3] = np.where(kind == 2, 0.0, 3.0)
    normal[pix, 3] = np.where(kind == 3, 30000000.0, 5.0)
    # A few moved samples pushed 50 scene units further along the eye's ray
    # through them: they project onto the same pixel, whose depth
    # reconstructs a point 50 units away, so the depth buffer does not show
    # them.
    off = (kind == 1) & (rng.uniform(size=ns) < 0.25)
    ray = f[perm[off], :3] - np.asarray(eye, np.float32)[None, :]
    ray /= np.linalg.norm(ray, axis=1, keepdims=True)
    sample[pix[off], :3] += np.float32(50.0 * frame.scene_scale) * ray
    ts, tn = torch.from_numpy(sample).cuda(), torch.from_numpy(normal).cuda()
    res = ra.GiReservoirs(ts.data_ptr(), ts.data_ptr(), tn.data_ptr())
    rec = torch.empty((ns, 12), dtype=torch.float32, device="cuda")
    tgt = torch.empty((ns, 4), dtype=torch.float32, device="cuda")
    ra.check(L.rgi_nrc_gather_vertices(nc.handle, ctypes.byref(frame), ctypes.byref(gbuf.gb), ctypes.byref(res), None,
                                       color.data_ptr(), ns, 1, rec.data_ptr(), tgt.data_ptr(), s))
    torch.cuda.synchronize()
    r, t, ft = rec.cpu().numpy(), tgt.cpu().numpy(), full_t.cpu().numpy()
    moved = (kind <= 1) & ~off
    assert moved.sum() > ns // 3
    assert np.array_equal(r[moved], f[perm[moved]]) and np.array_equal(t[moved], ft[perm[moved]])
    kept = ~moved
    assert np.array_equal(r[kept], f[kept]) and np.array_equal(t[kept], ft[kept])
    nc.close()
    L.rgi_destroy(ctx)


def test_train_batch_matches_single_steps():
    """rgi_nrc_train_batch (one Adam step of several caches of different
    table sizes and record counts in one chain of launches) against
    rgi_nrc_train on each cache from the same state: after the first step
    the MLP parameters and their moving average are bit-identical (the MLP
    gradient is a fixed-order reduction) and the table within its last bits
    (its gradient accumulates with FP32 atomics in no fixed order); after
    three steps both stay within a small tolerance, and every step count
    advanced."""
    L = ra.lib()
    cases = [(21, 14, 4096), (22, 15, 1000), (23, 13, 17)]
    batch_side, single_side, data = [], [], []
    for seed, log2_table, n in cases:
        a, _, rng = setup_cache(seed, log2_table)
        b, _, _ = setup_cache(seed, log2_table)
        batch_side.append(a)
        single_side.append(b)
        data.append([(torch.from_numpy(records(rng, n)).cuda(), torch.from_numpy(targets(rng, n)).cuda())
                     for _ in range(3)])
    P = ctypes.c_void_p
    losses = [torch.zeros(1, dtype=torch.float32, device="cuda") for _ in cases]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for step in range(3):
        print("nrc train batch, step", step, flush=True)
        caches = (P * 3)(*[c.handle.value for c in batch_side])
        recs = (P * 3)(*[d[step][0].data_ptr() for d in data])
        tgts = (P * 3)(*[d[step][1].data_ptr() for d in data])
        counts = (ra.U32 * 3)(*[n for _, _, n in cases])
        lp = (P * 3)(*[l.data_ptr() for l in losses])
        ra.check(L.rgi_nrc_train_batch(caches, recs, tgts, counts, lp, 3, s))
        for c, d in zip(single_side, data):
            c.train(d[step][0], d[step][1])
        torch.cuda.synchronize()
        for k, (a, b) in enumerate(zip(batch_side, single_side)):
            (ma, ta), (mb, tb) = a.params(), b.params()
            (ea, fa), (eb, fb) = a.average(), b.average()
            ma, ta, mb, tb = (x.cpu().numpy() for x in (ma, ta, mb, tb))
            ea, eb = ea.cpu().numpy(), eb.cpu().numpy()
            # Entries whose summed gradient lands on zero in one order and
            # not in the other take or skip a whole Adam step: a few per
            # hundred thousand at most.
            d = np.abs(ta - tb)
            if step == 0:
                assert np.array_equal(ma, mb), "cache %d: MLP differs after one step" % k
                assert np.array_equal(ea, eb), "cache %d: MLP average differs after one step" % k
                assert np.count_nonzero(d > 1e-6) <= max(4, d.size // 100000), (k, np.count_nonzero(d > 1e-6))
            else:
                assert np.max(np.abs(ma - mb)) < 1e-3 * np.max(np.abs(mb)), (k, step)
                assert np.count_nonzero(d > 1e-4) <= max(8, d.size // 10000), (k, step, np.count_nonzero(d > 1e-4))
            assert float(losses[k].cpu()[0]) > 0.0
            sa, sb = ra.U32(), ra.U32()
            ra.check(L.rgi_nrc_steps(a.handle, ctypes.byref(sa)))
            ra.check(L.rgi_nrc_steps(b.handle, ctypes.byref(sb)))
            assert sa.value == sb.value == step + 1
    for c in batch_side + single_side:
        c.close()


def test_renderer_caches_in_pass_major_batch():
    """Three renderers with resolving caches as one pass-major batch (the
    caches' queries and one chain of training steps on the batch's side
    stream, left running across calls) against the same renderers rendered
    alone: every batch frame trains each cache once, and the resolved images
    are finite and close to the eager renderers' mean (the table gradient's
    FP32 atomics make the caches differ in their last bits, so the check is
    statistical)."""
    W, H = 160, 120
    n = W * H
    st = ra.Settings(grid_table_size=1 << 16, grid_mode=ra.RGI_GRID_DETERMINISTIC)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, st)
    L = ra.lib()
    from restir_amd import driver
    s = driver.stream()
    settings = rnrc.renderer_settings(hs, batch=4096, warmup=2, resolve=1)
    batched = [_renderer(hs, gscene, st, W, H, settings) for _ in range(3)]
    alone = [_renderer(hs, gscene, st, W, H, settings) for _ in range(3)]
    batch = ctypes.c_void_p()
    out = ra.RendererOutputs()
    try:
        arr = (ctypes.c_void_p * 3)(*[r.value for r in batched])
        ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, 3))
        ra.check(L.rgi_batch_set_mode(batch, ra.RGI_BATCH_PASSES))
        frames = 12
        views = (ra.View * 3)(view, view, view)
        gbufs = (ra.GBuffer * 3)(gbuf.gb, gbuf.gb, gbuf.gb)
        for frame in range(frames):
            print("pass-major batch with caches, frame", frame, flush=True)
            ra.check(L.rgi_batch_render(batch, views, gbufs, s))
            for r in alone:
                ra.check(L.rgi_renderer_render(r, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
            if frame == 0:
                # The first step trains equal caches on the frame's gathered
                # records: the batch's one gather chain and the renderers'
                # own gathers give the same records exactly when the MLPs
                # after that step are bit-identical (its gradient is a
                # fixed-order reduction).
                torch.cuda.synchronize()
                for rb, rs in zip(batched, alone):
                    mb = torch.empty(ra.RGI_NRC_MLP_PARAMS, dtype=torch.float32, device="cuda")
                    ms = torch.empty(ra.RGI_NRC_MLP_PARAMS, dtype=torch.float32, device="cuda")
                    cb, cs = rnrc.renderer_state(rb)[0], rnrc.renderer_state(rs)[0]
                    ra.check(L.rgi_nrc_get_params(ctypes.c_void_p(cb), mb.data_ptr(), None, s))
                    ra.check(L.rgi_nrc_get_params(ctypes.c_void_p(cs), ms.data_ptr(), None, s))
                    torch.cuda.synchronize()
                    assert torch.equal(mb, ms), "the batch's first training step differs from the renderer's own"
        torch.cuda.synchronize()
        means = []
        for r in batched + alone:
            cache, rad_ptr, loss_ptr, trained = rnrc.renderer_state(r)
            assert cache and rad_ptr and trained == frames
            loss = torch.empty(1, dtype=torch.float32, device="cuda")
            L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            assert L.hipMemcpy(loss.data_ptr(), loss_ptr, 4, 3) == 0
            assert np.isfinite(float(loss.item())) and float(loss.item()) > 0
            steps = ra.U32()
            ra.check(L.rgi_nrc_steps(ctypes.c_void_p(cache), ctypes.byref(steps)))
            assert steps.value == frames
            ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(out)))
            hdr = _half_plane(out.hdr, n)
            assert np.all(np.isfinite(hdr))
            means.append(hdr[:, :3].mean())
        print("batch / alone resolved means:", ["%.5f" % m for m in means], flush=True)
        ref = float(np.mean(means[3:]))
        for m in means:
            assert abs(m - ref) < 0.05 * ref
    finally:
        if batch:
            L.rgi_batch_destroy(batch)
        for r in batched + alone:
            L.rgi_renderer_destroy(r)
        L.rgi_destroy(ctx)


def test_batch_cache_queries_match_single_renderers():
    """A pass-major batch queries all its caches in one launch per precision
    (the query plane, and the resolved plane where the query adds emission):
    with every cache given the same parameters after the warm-up frames (so
    that the batch's and the single renderers' caches are equal), the first
    frame that queries them is bit-identical between the batch and the
    renderers alone: query planes, GI reservoirs and colour, composite."""
    W, H = 160, 120
    n = W * H
    base = ra.Settings(grid_table_size=1 << 16, grid_mode=ra.RGI_GRID_DETERMINISTIC)
    hs, gscene, ctx, gbuf, view = _setup_scene(W, H, base)
    L = ra.lib()
    from restir_amd import driver
    s = driver.stream()
    variants = [(base, rnrc.renderer_settings(hs, batch=4096, warmup=3, resolve=1)),
                (ra.Settings(grid_table_size=1 << 16, grid_mode=ra.RGI_GRID_DETERMINISTIC, restir_di=False),
                 rnrc.renderer_settings(hs, batch=4096, warmup=3, resolve=1)),
                (base, rnrc.renderer_settings(hs, batch=2048, warmup=3, precision=ra.RGI_NRC_FP16))]
    batched = [_renderer(hs, gscene, st, W, H, ns) for st, ns in variants]
    alone = [_renderer(hs, gscene, st, W, H, ns) for st, ns in variants]
    batch = ctypes.c_void_p()
    try:
        arr = (ctypes.c_void_p * 3)(*[r.value for r in batched])
        ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, 3))
        views = (ra.View * 3)(view, view, view)
        gbufs = (ra.GBuffer * 3)(gbuf.gb, gbuf.gb, gbuf.gb)

        def frame():
            ra.check(L.rgi_batch_render(batch, views, gbufs, s))
            for r in alone:
                ra.check(L.rgi_renderer_render(r, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
            torch.cuda.synchronize()

        for _ in range(3):
            frame()
        probe = rnrc.NeuralCache(variants[0][1].cache)
        rng = np.random.default_rng(9)
        mlp = torch.from_numpy((rng.uniform(-1, 1, ra.RGI_NRC_MLP_PARAMS) * 0.15).astype(np.float32)).cuda()
        table = torch.from_numpy((rng.uniform(-1, 1, probe.table_floats) * 0.05).astype(np.float32)).cuda()
        probe.close()
        for r in batched + alone:
            cache = rnrc.renderer_state(r)[0]
            ra.check(L.rgi_nrc_set_params(ctypes.c_void_p(cache), mlp.data_ptr(), table.data_ptr(), s))
        torch.cuda.synchronize()
        print("batch cache queries: the querying frame", flush=True)
        frame()
        for k, (rb, ra_) in enumerate(zip(batched, alone)):
            planes = []
            for r in (rb, ra_):
                cache, rad, loss, trained = rnrc.renderer_state(r)
                assert trained == 4
                o = ra.RendererOutputs()
                ra.check(L.rgi_renderer_outputs_get(r, ctypes.byref(o)))
                planes.append({"query": _half_plane(rad, n), "color": _half_plane(o.color, n),
                               "hdr": _half_plane(o.hdr, n)})
            for name in ("query", "color", "hdr"):
                a, b = planes[0][name], planes[1][name]
                assert np.array_equal(a, b), "renderer %d: %s differs between the batch and the renderer alone" % (
                    k, name)
            assert np.any(planes[0]["query"][:, :3] != 0.0)
    finally:
        if batch:
            L.rgi_batch_destroy(batch)
        for r in batched + alone:
            L.rgi_renderer_destroy(r)
        L.rgi_destroy(ctx)
