"""The oracle at the reference's default half-resolution GI (kGiResolutions,
src/settings.h:18-23, 218; renderer.cpp:519-520, 769-798): GI passes over a
window-size G-buffer, their own planes at the GI size.

Pinned by an identity the sampling rules fix exactly: a G-buffer whose
texels are each repeated 2 x 2 is the same image at twice the resolution.
Point sampling reads texel floor(uv * 2W) of it, whose source texel is
floor(floor(uv * 2W) / 2) = floor(uv * W) (uv * 2W is exact in binary
floating point), so GI at divisor 2 over the repeated G-buffer must equal
GI at divisor 1 over the original, bit for bit. A pass that indexed a
G-buffer plane by GI pixel, or a GI plane at the G-buffer's size, fails it.
"""
import ctypes

import numpy as np
import pytest

import restir_amd as ra

import common
import orc

W, H = 64, 48


@pytest.fixture(scope="module")
def full_res():
    """Two frames at divisor 1 (W x H window): the second frame's inputs and
    outputs, with the first frame's outputs as history."""
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(grid_mode=ra.RGI_GRID_DETERMINISTIC)
    seq = common.oracle_sequence(hs, s, W, H, 2, yaw_step=2 * np.pi / 64)
    return hs, s, seq


def test_frame_pack_half_res_uniforms(full_res):
    """rgi_frame_pack at a 2W x 2H window and divisor 2 gives the W x H
    frame's uniforms with only the window size (u_screenSize) changed: the
    GI size, its texel size and u_gridInfo.y follow the GI height
    (renderer.cpp:708-720)."""
    hs, s, seq = full_res
    f = seq[1][0]
    s2 = ra.Settings(**dict(s.__dict__, gi_divisor=2))
    f2 = common.frame_for(hs, s2, 2 * W, 2 * H, 1, True, True, yaw_step=2 * np.pi / 64)
    assert (f2.width, f2.height, f2.gi_width, f2.gi_height) == (2 * W, 2 * H, W, H)
    g = ra.Frame.from_buffer_copy(f2)
    g.width, g.height = W, H
    assert bytes(g) == bytes(f)
    # Floors and the lower bound of 1 (renderer.cpp:519-520).
    for (ww, hh, d), want in [((97, 61, 2), (48, 30)), ((7, 5, 4), (1, 1)), ((1920, 1080, 4), (480, 270))]:
        fr = common.frame_for(hs, ra.Settings(gi_divisor=d), ww, hh, 0, False)
        assert (fr.gi_width, fr.gi_height) == want


def test_gi_passes_over_repeated_gbuffer_equal_full_res(full_res):
    hs, s, seq = full_res
    f, g, hist, want = seq[1]
    assert hist["gi_sample"].any() and f.history_valid
    surface_light = want["surface_light"]
    hdr = hist["hdr"]
    full = common.gi_passes_oracle(hs, f, g, hist, surface_light, hdr)
    f2 = ra.Frame.from_buffer_copy(f)
    f2.width, f2.height = 2 * W, 2 * H
    g2 = {k: common.replicate(v, W, H, 2) for k, v in g.items()}
    hist2 = dict(hist)
    half = common.gi_passes_oracle(hs, f2, g2, hist2, common.replicate(surface_light, W, H, 2),
                                   common.replicate(hdr, W, H, 2))
    # The divisor-1 run is itself the renderer's frame (a check of the helper).
    for k in ("gi_sample", "gi_receiver", "gi_normal", "color", "moments"):
        assert np.array_equal(full[k], want[k]), k
    for k in ("grid", "gi_sample", "gi_receiver", "gi_normal", "color", "moments", "specular_gi", "filtered",
              "denoised"):
        assert np.array_equal(full[k], half[k]), k
    assert (full["color"].view(np.float16)[:, :3] != 0).any(1).mean() > 0.1


def test_gi_passes_over_unrepeated_gbuffer_differ(full_res):
    """The identity is not vacuous: a 2W x 2H G-buffer rendered at that
    resolution (not repeated) gives different GI."""
    hs, s, seq = full_res
    f, g, hist, want = seq[1]
    s2 = ra.Settings(**dict(s.__dict__, gi_divisor=2))
    f2 = common.frame_for(hs, s2, 2 * W, 2 * H, 1, True, True, yaw_step=2 * np.pi / 64)
    g2 = common.oracle_gbuffer(hs, f2)
    sl2 = common.replicate(want["surface_light"], W, H, 2)
    half = common.gi_passes_oracle(hs, f2, g2, hist, sl2, common.replicate(hist["hdr"], W, H, 2))
    full = common.gi_passes_oracle(hs, f, g, hist, want["surface_light"], hist["hdr"])
    assert not np.array_equal(full["gi_receiver"], half["gi_receiver"])


def test_gi_entry_points_reject_a_gbuffer_smaller_than_the_gi_size():
    """Host-side checks only (no GPU call is made): every entry point over
    GI pixels rejects a G-buffer smaller than the frame's GI size, and one
    with no GI size; the specular pass included."""
    lib = ra.lib()
    hs = common.cornell(extra_lights=0)
    f = common.frame_for(hs, ra.Settings(), 32, 24, 0, False)
    dummy = np.zeros(64, np.float32)
    p = dummy.ctypes.data
    g = ra.GBuffer(16, 12, p, p, p, p, p)
    scene = hs.rgi_scene(None, None)
    scene.node_count = 0
    gin = ra.GiInputs(p, p, p, p, p, p, None, p)
    res = ra.GiReservoirs(p, p, p)
    calls = {
        "grid_insert": lambda: lib.rgi_grid_insert(None, ctypes.byref(f), ctypes.byref(g), p, 0, None),
        "gi_temporal": lambda: lib.rgi_gi_temporal(None, ctypes.byref(f), ctypes.byref(scene), ctypes.byref(g),
                                                   ctypes.byref(gin), ctypes.byref(res), None),
        "gi_spatial": lambda: lib.rgi_gi_spatial(None, ctypes.byref(f), ctypes.byref(scene), ctypes.byref(g),
                                                 ctypes.byref(gin), ctypes.byref(res), p, p, None),
        "specular_gi": lambda: lib.rgi_specular_gi(None, ctypes.byref(f), ctypes.byref(scene), ctypes.byref(g),
                                                   ctypes.byref(gin), p, p, dummy[32:].ctypes.data, None),
    }
    for name, call in calls.items():
        assert call() == -1, name
        assert b"smaller than the frame's GI size" in lib.rgi_last_error(), name
    f.gi_width = 0
    assert lib.rgi_svgf_variance(None, ctypes.byref(f), p, p, p, p, p, None) == -1
    assert b"GI size" in lib.rgi_last_error()
    ci = ra.CompositeInputs(p, None, None, None, None)
    g.width, g.height = 32, 24
    assert lib.rgi_composite(None, ctypes.byref(f), ctypes.byref(g), ctypes.byref(ci), p, None) == -1
    assert b"GI size" in lib.rgi_last_error()
