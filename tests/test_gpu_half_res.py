"""GPU: the reference's default half-resolution GI (kGiResolutions,
src/settings.h:18-23, 218; renderer.cpp:519-520, 769-798, 1329-1331)
through the C ABI. The G-buffer is the window's size; grid insert, GI
temporal, GI spatial, indirect specular and the denoiser run over GI pixels
with their planes at the GI size; the composite upsamples the GI planes to
the window. Every output is compared bit for bit with the oracle on the same
inputs, and against the identity the sampling rules fix (a G-buffer whose
texels are repeated 2 x 2 gives the divisor-1 result)."""
import ctypes

import numpy as np
import pytest

import restir_amd as ra

import common
import orc

pytestmark = pytest.mark.gpu

W, H = 320, 240


@pytest.fixture(scope="module")
def dev():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = common.Device()
    yield d
    d.close()


def say(*a):
    print(*a, flush=True)


def assert_same(name, got, want):
    eq = (got == want) | (np.isnan(got) & np.isnan(want)) if got.dtype.kind == "f" else got == want
    bad = int((~eq).sum())
    if bad:
        idx = np.argwhere(~eq)[:5]
        detail = "; ".join("%s got %r want %r" % (tuple(i), got[tuple(i)], want[tuple(i)]) for i in idx)
        raise AssertionError("%s: %d of %d values differ: %s" % (name, bad, eq.size, detail))


def gi_passes_device(dev, hs, frame, g, hist, surface_light, previous_hdr, options=None):
    """common.gi_passes_oracle through the C ABI."""
    L = dev.L
    T = frame.grid_table_size
    ng = frame.gi_width * frame.gi_height
    before = {opt: L.rgi_context_get_option(dev.ctx, opt) for opt in (options or {})}
    for opt, value in (options or {}).items():
        ra.check(L.rgi_context_set_option(dev.ctx, opt, value))
    try:
        scene = dev.scene(hs)
        gb, keep = dev.gbuffer(g, frame.width, frame.height)
        d = {k: dev.put(hist[k]) for k in ("gi_sample", "gi_receiver", "gi_normal", "filtered", "moments", "grid",
                                            "specular_gi")}
        sl, hdr = dev.put(surface_light), dev.put(previous_hdr)
        grid = dev.empty(34 * T * 4)
        ra.check(L.rgi_grid_clear(grid.data_ptr(), T, dev.stream))
        ra.check(L.rgi_grid_insert(dev.ctx, ctypes.byref(frame), ctypes.byref(gb), grid.data_ptr(),
                                   ra.RGI_GRID_DETERMINISTIC, dev.stream))
        gin = ra.GiInputs(d["gi_sample"].data_ptr(), d["gi_receiver"].data_ptr(), d["gi_normal"].data_ptr(),
                          d["filtered"].data_ptr(), d["moments"].data_ptr(), sl.data_ptr(), None, d["grid"].data_ptr())
        res_t = [dev.empty(ng * 16) for _ in range(3)]
        res = ra.GiReservoirs(*(t.data_ptr() for t in res_t))
        ra.check(L.rgi_gi_temporal(dev.ctx, ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb),
                                   ctypes.byref(gin), ctypes.byref(res), dev.stream))
        sin = ra.GiInputs(None, d["gi_receiver"].data_ptr(), d["gi_normal"].data_ptr(), d["filtered"].data_ptr(),
                          d["moments"].data_ptr(), sl.data_ptr(), None, grid.data_ptr())
        color, moments = dev.empty(ng * 8), dev.empty(ng * 8)
        ra.check(L.rgi_gi_spatial(dev.ctx, ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb),
                                  ctypes.byref(sin), ctypes.byref(res), color.data_ptr(), moments.data_ptr(),
                                  dev.stream))
        spec = dev.empty(ng * 8)
        ra.check(L.rgi_specular_gi(dev.ctx, ctypes.byref(frame), ctypes.byref(scene), ctypes.byref(gb),
                                   ctypes.byref(sin), hdr.data_ptr(), d["specular_gi"].data_ptr(), spec.data_ptr(),
                                   dev.stream))
        var, filt, den = dev.empty(ng * 8), dev.empty(ng * 8), dev.empty(ng * 8)
        ra.check(L.rgi_svgf_variance(dev.ctx, ctypes.byref(frame), color.data_ptr(), moments.data_ptr(),
                                     res_t[1].data_ptr(), res_t[2].data_ptr(), var.data_ptr(), dev.stream))
        ra.check(L.rgi_svgf_atrous(dev.ctx, ctypes.byref(frame), var.data_ptr(), res_t[1].data_ptr(),
                                   res_t[2].data_ptr(), 1, filt.data_ptr(), dev.stream))
        ra.check(L.rgi_svgf_atrous(dev.ctx, ctypes.byref(frame), filt.data_ptr(), res_t[1].data_ptr(),
                                   res_t[2].data_ptr(), 2, den.data_ptr(), dev.stream))
        f4 = np.zeros((ng, 4), np.float32)
        h4 = np.zeros((ng, 4), np.uint16)
        return {
            "grid": dev.get(grid, np.zeros(34 * T, np.uint32)),
            "gi_sample": dev.get(res_t[0], f4), "gi_receiver": dev.get(res_t[1], f4),
            "gi_normal": dev.get(res_t[2], f4), "color": dev.get(color, h4), "moments": dev.get(moments, h4),
            "specular_gi": dev.get(spec, h4), "filtered": dev.get(filt, h4), "denoised": dev.get(den, h4),
        }
    finally:
        for opt, value in before.items():
            ra.check(L.rgi_context_set_option(dev.ctx, opt, value))


KEYS = ("grid", "gi_sample", "gi_receiver", "gi_normal", "color", "moments", "specular_gi", "filtered", "denoised")


@pytest.fixture(scope="module")
def half_seq():
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC, gi_divisor=2)
    seq = common.oracle_sequence(hs, s, W, H, 3, yaw_step=2 * np.pi / 64)
    return hs, s, seq


@pytest.mark.parametrize("variant", ["default", "xcd_blocks_waves4", "wide_tiles"])
def test_half_res_gi_passes_match_oracle(dev, half_seq, variant):
    hs, s, seq = half_seq
    f, g, hist, out = seq[2]
    assert (f.width, f.height, f.gi_width, f.gi_height) == (W, H, W // 2, H // 2)
    say("case half-res GI passes %s" % variant)
    opts = {"default": {},
            "xcd_blocks_waves4": {ra.RGI_OPT_TILE_ORDER: 2, ra.RGI_OPT_GI_TEMPORAL_WAVES: 4,
                                  ra.RGI_OPT_GI_SPATIAL_WAVES: 5},
            "wide_tiles": {ra.RGI_OPT_TILE_SHAPE: 1}}[variant]
    want = common.gi_passes_oracle(hs, f, g, hist, out["surface_light"], hist["hdr"])
    # The helper's frame is the sequence's own (the same passes, same inputs).
    for k in ("grid", "gi_sample", "gi_receiver", "gi_normal", "color", "moments"):
        assert np.array_equal(want[k], out[k]), k
    got = gi_passes_device(dev, hs, f, g, hist, out["surface_light"], hist["hdr"], opts)
    for k in KEYS:
        if k == "grid":
            for part, a, b in zip(("checksums", "counts", "entries"), common.grid_view(got[k], f.grid_table_size),
                                  common.grid_view(want[k], f.grid_table_size)):
                assert_same("grid." + part, a, b)
        else:
            assert_same(k, got[k], want[k])
    say("  GI texels lit %.3f" % (want["color"].view(np.float16)[:, :3] != 0).any(1).mean())


def test_half_res_over_repeated_gbuffer_equals_full_res(dev):
    """The GPU at divisor 2 over a 2 x 2-repeated G-buffer gives the oracle's
    divisor-1 result over the original, bit for bit."""
    w, h = 160, 120
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC)
    seq = common.oracle_sequence(hs, s, w, h, 2, yaw_step=2 * np.pi / 64)
    f, g, hist, out = seq[1]
    say("case half-res repeated G-buffer %dx%d" % (w, h))
    want = common.gi_passes_oracle(hs, f, g, hist, out["surface_light"], hist["hdr"])
    f2 = ra.Frame.from_buffer_copy(f)
    f2.width, f2.height = 2 * w, 2 * h
    g2 = {k: common.replicate(v, w, h, 2) for k, v in g.items()}
    got = gi_passes_device(dev, hs, f2, g2, hist, common.replicate(out["surface_light"], w, h, 2),
                           common.replicate(hist["hdr"], w, h, 2))
    for k in KEYS:
        if k == "grid":
            for part, a, b in zip(("checksums", "counts", "entries"), common.grid_view(got[k], f.grid_table_size),
                                  common.grid_view(want[k], f.grid_table_size)):
                assert_same("grid." + part, a, b)
        else:
            assert_same(k, got[k], want[k])


@pytest.mark.parametrize("divisor", [2, 4])
def test_half_res_composite_matches_oracle(dev, divisor):
    """rgi_composite with GI and indirect specular planes at the GI size,
    upsampled over same-surface texels (fs_composite.sc:48-80) to the window."""
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(gi_divisor=divisor, shadows=True, light_color=(1.0, 0.9, 0.8), light_position=(0.1, 1.2, -0.2),
                    shadow_face_size=0)
    seq = common.oracle_sequence(hs, s, W, H, 2, yaw_step=2 * np.pi / 64)
    f, g, hist, out = seq[-1]
    say("case half-res composite divisor=%d" % divisor)
    assert out["hdr"].shape[0] == W * H and out["denoised"].shape[0] == f.gi_width * f.gi_height
    gb, keep = dev.gbuffer(g, W, H)
    d = {k: dev.put(out[k]) for k in ("denoised", "diffuse", "specular", "specular_gi", "shadow_atlas")}
    ci = ra.CompositeInputs(d["denoised"].data_ptr(), d["diffuse"].data_ptr(), d["specular"].data_ptr(),
                            d["specular_gi"].data_ptr(), d["shadow_atlas"].data_ptr())
    ci.options[:] = [1.0, 1.0, 1.0, 1.0]
    ci.specular_weight = 1.0
    o = dev.empty(W * H * 8)
    ra.check(dev.L.rgi_composite(dev.ctx, ctypes.byref(f), ctypes.byref(gb), ctypes.byref(ci), o.data_ptr(),
                                 dev.stream))
    assert_same("hdr", dev.get(o, out["hdr"]), out["hdr"])
