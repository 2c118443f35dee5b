"""Benchmark of the ReSTIR GI/DI hot path on one MI355X.

Workload (the 1080p Cornell configuration): the procedural open-front
Cornell box (36 triangles, emissive ceiling quad of radiance 17) seen at
1920 x 1080 along a 16-frame camera path that swings from the reference's
starting yaw (0.64 rad) to 0 and back, keeping the open front in view (at
most 7 degrees per frame), ReSTIR DI over
the 2 emissive triangles plus 4 seeded point lights, a 2^20-cell hash grid,
default settings (2 candidates, 2 temporal taps, 4 spatial taps, 32 march
steps, hybrid rays, clamps and boiling filter on). A step is one GI frame:
DI temporal, DI spatial, grid clear, grid insert, GI temporal, GI spatial,
on G-buffers already resident in HBM. The separate traversal benchmark traces
2^24 cosine rays from uniform surface points (closest hit, tmax 4 x scene
scale) and reports Mrays/s.

Beside the metric it reports the neural radiance cache: the metric's frame
with the cache on, its kernels alone (2M per-pixel queries in BF16, FP16 and
FP8; one training step of the default 16384 surface records) with the MLP's matrix-core
throughput, and the GI and composite MSE of the plain path and the cache
against the same 4096-frame reference.

Prints one JSON line. --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import restir_amd as ra  # noqa: E402
from restir_amd import driver  # noqa: E402
from restir_amd.workloads import UNEVEN_SCENES, surface_rays  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per pixel of each pass (compulsory traffic with the
# reference's formats: each pass's reads and writes of its input and output
# planes, reservoirs and grid entries, once per pixel).
BYTES_PER_PIXEL = {
    "grid_insert": 20.0,
    "gi_temporal": 220.0,
    "gi_spatial": 508.0,
    "di_temporal": 96.0,
    "di_spatial": 240.0,
}


def measured_traffic(pass_name, options, width, height):
    """HBM bytes per launch of the kernel variant the autotuner chose, from
    the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload
    (calibrated against a copy of known size), or None when none matches."""
    path = os.path.join(ROOT, "profiles", "cornell1080p_kernels.json")
    if (width, height) != (1920, 1080) or not os.path.exists(path):
        return None
    with open(path) as fh:
        summary = json.load(fh)
    waves = {"gi_temporal": options.get(ra.RGI_OPT_GI_TEMPORAL_WAVES, 0),
             "gi_spatial": options.get(ra.RGI_OPT_GI_SPATIAL_WAVES, 0)}.get(pass_name, 0)
    suffix = "_w%d" % waves if waves else ""
    # GI temporal is several kernels (surface-tile listing, candidate march,
    # BVH walks, resampling): its traffic is their sum.
    k = summary.get("passes", {}).get(pass_name + suffix)
    if k is not None:
        kernel = " + ".join(k["kernels"])
    else:
        kernel = {"grid_insert": "k_grid_insert_atomic"}.get(pass_name, "k_" + pass_name) + suffix
        k = summary["kernels"].get(kernel)
    if k is None or "fetch_bytes" not in k or "write_bytes" not in k:
        return None
    return {"bytes": round(k["fetch_bytes"] + k["write_bytes"]), "fetch_bytes": round(k["fetch_bytes"]),
            "write_bytes": round(k["write_bytes"]), "kernel": kernel, "source": "profiles/cornell1080p_kernels.txt"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hs, settings, width, height, table, reps=5):
    """The oracle on the host cores: the third frame of the sequence (reuse
    and history active), all passes, `reps` times at up to 16 threads
    (median per pass and per frame), then once on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import common
    import orc
    threads = min(os.cpu_count() or 1, 16)
    orc.lib().orc_set_threads(threads)
    s = ra.Settings(**settings.__dict__)
    s.grid_table_size = table
    hist = common.empty_history(width, height, table)
    frames = []
    for i in range(3):
        eye, vp, inv = ra.orbit_camera(hs, i, width, height, yaw_of=ra.swing_yaw)
        peye, pvp, _ = ra.orbit_camera(hs, max(i - 1, 0), width, height, yaw_of=ra.swing_yaw)
        f = ra.make_frame(s, hs, i, eye, vp, inv, peye, pvp, width, height, i > 0, i > 0)
        g = common.oracle_gbuffer(hs, f)
        frames.append((f, g))
        if i < 2:
            hist = common.next_history(hist, common.oracle_frame(hs, f, g, hist))
    f, g = frames[2]
    totals, per_pass = [], {}
    for _ in range(reps):
        times = {}
        t0 = time.perf_counter()
        common.oracle_frame(hs, f, g, hist, times=times)
        totals.append(time.perf_counter() - t0)
        for k, v in times.items():
            per_pass.setdefault(k.replace("orc_", ""), []).append(v)
    orc.lib().orc_set_threads(1)
    t0 = time.perf_counter()
    common.oracle_frame(hs, f, g, hist)
    single = time.perf_counter() - t0
    orc.lib().orc_set_threads(threads)
    return {"value": float(np.median(totals)) * 1e3, "unit": "ms per GI frame", "cores": threads, "kind": "port",
            "sample": "oracle, one steady-state %dx%d frame (frame 3 of the path), all passes, median of %d"
                      % (width, height, reps),
            "pass_ms_median": {k: round(float(np.median(v)) * 1e3, 2) for k, v in sorted(per_pass.items())},
            "single_thread_ms": round(single * 1e3, 1), "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count()}


def measure_mse(L, hs, gscene, gbuf, view, W, H, T, mode, ref_frames, frames, ctx, s):
    """MSE against a `ref_frames`-frame reference at one static camera
    (frame 0 of the path), on the composite HDR image that the reference
    accumulates and measures (fs_reference.sc, cs_error.sc; no TAA here):
    the reference renders with every biased shortcut and the denoiser off
    (renderer.cpp:437-452) and accumulates in FP64; the path under test then
    renders `frames` frames from an empty history with the reference's
    default settings (denoiser on), and its last composite is measured.
    The same without the denoiser is reported beside it."""
    import torch
    n = W * H
    acc = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    acc_gi = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    out = ra.RendererOutputs()
    res = torch.zeros(5, dtype=torch.float64, device="cuda")
    results = {}
    for name, denoise in (("display", True), ("no_denoiser", False)):
        settings = ra.Settings(grid_table_size=T, grid_mode=mode, denoiser=denoise)
        rend = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                       ctypes.byref(gscene.scene), hs.radius))
        if name == "display":
            ra.check(L.rgi_renderer_set_reference_mode(rend, 1))
            for _ in range(ref_frames):
                ra.check(L.rgi_renderer_render(rend, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
                ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(out)))
                ra.check(L.rgi_reference_accumulate_composite(ctypes.byref(gbuf.gb), out.hdr, acc.data_ptr(), s))
                ra.check(L.rgi_reference_accumulate(ctypes.byref(gbuf.gb), out.color, None, None, acc_gi.data_ptr(), s))
            ra.check(L.rgi_renderer_set_reference_mode(rend, 0))
            L.rgi_renderer_invalidate_history(rend)
        for _ in range(frames):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(out)))
        ra.check(L.rgi_reference_error_composite(ctx, ctypes.byref(gbuf.gb), out.hdr, acc.data_ptr(), ref_frames,
                                                 res.data_ptr(), s))
        r = res.cpu().numpy()
        results[name] = {"relative_mse": float(r[0]), "mse_rgb": float(r[1]), "mean_luminance": float(r[2]),
                         "reference_mean_luminance": float(r[3]), "pixels": int(r[4])}
        L.rgi_renderer_destroy(rend)
    line = dict(results["display"])
    line.update({"reference_frames": ref_frames, "frames": frames,
                 "method": "ReSTIR GI + DI, indirect specular, SVGF denoiser (reference defaults), composite; no TAA",
                 "image": "composite HDR (fs_composite.sc), static camera, geometry pixels",
                 "no_denoiser": results["no_denoiser"]})
    return line, acc, acc_gi


NRC_FLOP_PER_QUERY = 2 * (5 * 64 * 64 + 3 * 64)  # the MLP's multiply-adds per query (logical shapes)


# The plain path given more samples, to at least the resolved cache's frame
# time: twice the GI candidates, GI spatial taps and DI spatial taps, three
# times the DI candidates.
PLAIN_HEAVY = {"initial_candidates": 4, "di_candidates": 24, "di_spatial_taps": 6, "spatial_taps": 8}


def nrc_quality(L, hs, gscene, gbuf, view, W, H, T, mode, acc, acc_gi, ref_frames, frames, ctx, s):
    """The neural radiance cache against the plain path, each rendered
    `frames` frames from an empty history at the static camera with the
    reference's default settings, measured against the same `ref_frames`
    reference accumulations: the GI estimate alone (the GI frame's output,
    colour x albedo) and the composite, each the mean of the MSE of the last
    16 frames (one frame's MSE varies by about 15% from run to run). Variants:
    plain; the cache standing in for path continuation (its default records,
    65,536 records, records at pixels, mixed records); the cache with the
    composite resolved from it (pixel and path-vertex records); the plain path
    with PLAIN_HEAVY's extra samples. Each variant's full frame (denoiser
    and composite included) is timed over the 32 frames before the measured
    ones."""
    import torch
    from restir_amd import nrc as rnrc
    res = torch.zeros(5, dtype=torch.float64, device="cuda")
    out = ra.RendererOutputs()
    result = {"frames": frames, "reference_frames": ref_frames, "plain_heavy_settings": PLAIN_HEAVY}
    # By default the cache trains on 16,384 records per frame at ReSTIR GI path
    # vertices (at surface pixels when it resolves the composite); beside it
    # the same at the fourth configuration's 65,536 records, with records at
    # surface pixels, every other one at a path vertex, and the resolved
    # composite from path-vertex records.
    variants = {"cache": {}, "resolved": {"resolve": 1}, "cache_65536": {"batch": 65536},
                "cache_pixel_records": {"records": ra.RGI_NRC_RECORDS_PIXELS},
                "cache_mixed_records": {"records": ra.RGI_NRC_RECORDS_MIXED},
                "resolved_vertex_records": {"resolve": 1, "records": ra.RGI_NRC_RECORDS_VERTICES}}
    measured = 16
    for name in ("plain", "cache", "resolved", "plain_heavy", "cache_65536", "cache_pixel_records",
                 "cache_mixed_records", "resolved_vertex_records"):
        settings = ra.Settings(grid_table_size=T, grid_mode=mode, **(PLAIN_HEAVY if name == "plain_heavy" else {}))
        rend = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                       ctypes.byref(gscene.scene), hs.radius))
        if name in variants:
            ra.check(L.rgi_renderer_set_nrc(rend, ctypes.byref(rnrc.renderer_settings(hs, **variants[name]))))
        # frames - 48 frames to settle, 32 timed, then 16 measured.
        timed = 32
        for k in range(frames - measured - timed):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(timed):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / timed
        gi_sum, comp_sum = np.zeros(4), np.zeros(4)
        for k in range(measured):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
            ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(out)))
            ra.check(L.rgi_reference_error(ctx, ctypes.byref(gbuf.gb), out.color, None, None, acc_gi.data_ptr(),
                                           ref_frames, res.data_ptr(), s))
            gi_sum += res.cpu().numpy()[:4]
            ra.check(L.rgi_reference_error_composite(ctx, ctypes.byref(gbuf.gb), out.hdr, acc.data_ptr(), ref_frames,
                                                     res.data_ptr(), s))
            comp_sum += res.cpu().numpy()[:4]
        line = {"ms_per_frame": round(ms, 4)}
        r = gi_sum / measured
        line["gi"] = {"mse_rgb": float(r[1]), "relative_mse": float(r[0]), "mean_luminance": float(r[2]),
                      "reference_mean_luminance": float(r[3])}
        r = comp_sum / measured
        line["composite"] = {"mse_rgb": float(r[1]), "relative_mse": float(r[0]), "mean_luminance": float(r[2]),
                             "reference_mean_luminance": float(r[3])}
        result[name] = line
        L.rgi_renderer_destroy(rend)
    comp = lambda k: result[k]["composite"]["mse_rgb"]
    result["gi_mse_plain_over_cache"] = round(result["plain"]["gi"]["mse_rgb"] / result["cache"]["gi"]["mse_rgb"], 4)
    for k in ("cache_65536", "cache_pixel_records", "cache_mixed_records"):
        result["gi_mse_plain_over_" + k] = round(result["plain"]["gi"]["mse_rgb"] / result[k]["gi"]["mse_rgb"], 4)
    result["composite_mse_plain_over_resolved"] = round(comp("plain") / comp("resolved"), 3)
    result["composite_mse_plain_over_resolved_vertex_records"] = round(comp("plain") / comp("resolved_vertex_records"), 3)
    result["composite_mse_plain_heavy_over_resolved"] = round(comp("plain_heavy") / comp("resolved"), 3)
    result["plain_heavy_frame_at_least_as_long"] = result["plain_heavy"]["ms_per_frame"] >= result["resolved"]["ms_per_frame"]
    return result


def nrc_motion_quality(L, hs, gscene, views, gbufs, T, mode, ctx, s, ref_frames=2048, loops=8, measured=(0, 4, 8, 12)):
    """The resolved cache against the plain path with the camera moving: the
    metric's swing path, `loops` times round, with the composite measured at
    the `measured` views of the last loop against `ref_frames` reference
    accumulations of each of those views (static camera, FP64)."""
    import torch
    from restir_amd import nrc as rnrc
    W, H = gbufs[0].gb.width, gbufs[0].gb.height
    n = W * H
    settings = ra.Settings(grid_table_size=T, grid_mode=mode)
    out = ra.RendererOutputs()
    res = torch.zeros(5, dtype=torch.float64, device="cuda")
    accs = {}
    for v in measured:
        acc = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
        rend = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                       ctypes.byref(gscene.scene), hs.radius))
        ra.check(L.rgi_renderer_set_reference_mode(rend, 1))
        for _ in range(ref_frames):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(views[v]), ctypes.byref(gbufs[v].gb), s))
            ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(out)))
            ra.check(L.rgi_reference_accumulate_composite(ctypes.byref(gbufs[v].gb), out.hdr, acc.data_ptr(), s))
        L.rgi_renderer_destroy(rend)
        accs[v] = acc
    path = len(views)
    result = {"reference_frames": ref_frames, "path_views": path, "loops": loops, "measured_views": list(measured)}
    for name in ("plain", "resolved"):
        rend = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                       ctypes.byref(gscene.scene), hs.radius))
        if name == "resolved":
            ra.check(L.rgi_renderer_set_nrc(rend, ctypes.byref(rnrc.renderer_settings(hs, resolve=1))))
        mses = []
        for k in range(loops * path):
            v = k % path
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(views[v]), ctypes.byref(gbufs[v].gb), s))
            if k >= (loops - 1) * path and v in accs:
                ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(out)))
                ra.check(L.rgi_reference_error_composite(ctx, ctypes.byref(gbufs[v].gb), out.hdr, accs[v].data_ptr(),
                                                         ref_frames, res.data_ptr(), s))
                mses.append(float(res.cpu().numpy()[1]))
        result[name] = {"composite_mse_rgb": [round(m, 9) for m in mses], "mean": sum(mses) / len(mses)}
        L.rgi_renderer_destroy(rend)
    result["composite_mse_plain_over_resolved"] = round(result["plain"]["mean"] / result["resolved"]["mean"], 3)
    del accs
    return result


def resolve_quality(L, hs, gscene, gbuf, view, T, mode, ctx, s, ref_frames=1024, frames=256):
    """The resolved cache against the plain path on another scene: the
    composite of each after `frames` frames at one static camera (default
    settings, denoiser on), against an FP64 accumulation of `ref_frames`
    reference frames of that camera."""
    import torch
    from restir_amd import nrc as rnrc
    W, H = gbuf.gb.width, gbuf.gb.height
    settings = ra.Settings(grid_table_size=T, grid_mode=mode)
    acc = torch.zeros(3 * W * H, dtype=torch.float64, device="cuda")
    res = torch.zeros(5, dtype=torch.float64, device="cuda")
    out = ra.RendererOutputs()

    def renderer(cache):
        rend = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                       ctypes.byref(gscene.scene), hs.radius))
        if cache:
            ra.check(L.rgi_renderer_set_nrc(rend, ctypes.byref(rnrc.renderer_settings(hs, resolve=1))))
        return rend

    rend = renderer(False)
    ra.check(L.rgi_renderer_set_reference_mode(rend, 1))
    for _ in range(ref_frames):
        ra.check(L.rgi_renderer_render(rend, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(out)))
        ra.check(L.rgi_reference_accumulate_composite(ctypes.byref(gbuf.gb), out.hdr, acc.data_ptr(), s))
    L.rgi_renderer_destroy(rend)
    result = {"reference_frames": ref_frames, "frames": frames, "camera": "static, view 0 of the swing path"}
    for name in ("plain", "resolved"):
        rend = renderer(name == "resolved")
        for _ in range(frames):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(view), ctypes.byref(gbuf.gb), s))
        ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(out)))
        ra.check(L.rgi_reference_error_composite(ctx, ctypes.byref(gbuf.gb), out.hdr, acc.data_ptr(), ref_frames,
                                                 res.data_ptr(), s))
        r = res.cpu().numpy()
        result[name] = {"composite_mse_rgb": float(r[1]), "relative_mse": float(r[0]), "mean_luminance": float(r[2]),
                        "reference_mean_luminance": float(r[3])}
        L.rgi_renderer_destroy(rend)
    result["composite_mse_plain_over_resolved"] = round(result["plain"]["composite_mse_rgb"] /
                                                        result["resolved"]["composite_mse_rgb"], 3)
    del acc
    return result


def cache_frames(hs, W, H, T, frames_in_path, steps):
    """The metric's swing path and settings with the neural radiance cache on:
    16 warm-up frames, then `steps` timed frames."""
    import torch
    from restir_amd import nrc as rnrc
    L = ra.lib()
    settings = ra.Settings(grid_table_size=T, denoiser=False, composite=False, specular_gi=False)
    gscene = driver.GpuScene(hs)
    ctx = ctypes.c_void_p()
    ra.check(L.rgi_create(ctypes.byref(ctx)))
    views, gbufs = [], []
    for i in range(frames_in_path):
        eye, vp, inv = ra.orbit_camera(hs, i, W, H, yaw_of=lambda k: ra.swing_yaw(k, frames_in_path))
        f = ra.make_frame(settings, hs, i, eye, vp, inv, eye, vp, W, H, False)
        gbufs.append(driver.GBufferSet(ctx, gscene, f))
        views.append(ra.make_view(eye, vp, inv))
    s = driver.stream()
    rend = ctypes.c_void_p()
    ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                   ctypes.byref(gscene.scene), hs.radius))
    ra.check(L.rgi_renderer_set_nrc(rend, ctypes.byref(rnrc.renderer_settings(hs))))

    def render(k):
        ra.check(L.rgi_renderer_render(rend, ctypes.byref(views[k % frames_in_path]),
                                       ctypes.byref(gbufs[k % frames_in_path].gb), s))

    for k in range(16):
        render(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        render(k)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    L.rgi_renderer_destroy(rend)
    L.rgi_destroy(ctx)
    return {"workload": "cornell_1080p_swing%d with the neural radiance cache" % frames_in_path,
            "ms_per_frame_with_cache": round(ms, 5), "steps": steps}


def nrc_timing(L, hs, settings, views, gbufs, gscene, W, H, frames_in_path, steps, plain_ms, s):
    """Cost of the cache: the metric's frame with the cache on (same path,
    same settings; the default 16,384 and 65,536 training records per frame),
    and its kernels alone with HIP events: 2M queries per precision (the 1080p
    frame's per-pixel query count) and one training step of 16,384 and of
    65,536 records, with the MLP's matrix-core throughput."""
    import torch
    from restir_amd import nrc as rnrc

    def cache_frame_ms(nrc_settings):
        rend = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                       ctypes.byref(gscene.scene), hs.radius))
        ra.check(L.rgi_renderer_set_nrc(rend, ctypes.byref(nrc_settings)))
        for k in range(16):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(views[k % frames_in_path]),
                                           ctypes.byref(gbufs[k % frames_in_path].gb), s))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            ra.check(L.rgi_renderer_render(rend, ctypes.byref(views[k % frames_in_path]),
                                           ctypes.byref(gbufs[k % frames_in_path].gb), s))
        torch.cuda.synchronize()
        L.rgi_renderer_destroy(rend)
        return (time.perf_counter() - t0) * 1e3 / steps

    frame_ms = cache_frame_ms(rnrc.renderer_settings(hs))
    # The cache as the fourth configuration sets it: 65,536 training records
    # per frame (ReSTIR GI path vertices, as by default).
    frame_ms_65536 = cache_frame_ms(rnrc.renderer_settings(hs, batch=65536))

    # The kernels alone, on the frame's own data: the per-pixel query of the
    # first camera's G-buffer and one training step on records gathered from
    # it (DI diffuse and GI colour planes of a rendered frame).
    st = rnrc.renderer_settings(hs)
    cache = rnrc.NeuralCache(st.cache)
    plain = ctypes.c_void_p()
    ra.check(L.rgi_renderer_create(ctypes.byref(plain), W, H, ctypes.byref(ra.c_settings(settings)),
                                   ctypes.byref(gscene.scene), hs.radius))
    for k in range(4):
        ra.check(L.rgi_renderer_render(plain, ctypes.byref(views[0]), ctypes.byref(gbufs[0].gb), s))
    outs = ra.RendererOutputs()
    ra.check(L.rgi_renderer_outputs_get(plain, ctypes.byref(outs)))
    frame = ra.Frame()
    ra.check(L.rgi_renderer_last_frame(plain, ctypes.byref(frame)))
    batch = st.batch
    train_rec = torch.empty((batch, 12), dtype=torch.float32, device="cuda")
    train_tgt = torch.empty((batch, 4), dtype=torch.float32, device="cuda")
    radiance = torch.empty(W * H * 4, dtype=torch.float16, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def gather():
        ra.check(L.rgi_nrc_gather_gbuffer(cache.handle, ctypes.byref(frame), ctypes.byref(gbufs[0].gb), outs.di_diffuse,
                                          outs.color, batch, 7, train_rec.data_ptr(), train_tgt.data_ptr(), s))

    gather_ms = timed(gather)
    train_ms = timed(lambda: cache.train(train_rec, train_tgt))
    big_rec = torch.empty((65536, 12), dtype=torch.float32, device="cuda")
    big_tgt = torch.empty((65536, 4), dtype=torch.float32, device="cuda")
    gi_res = ra.GiReservoirs(outs.gi_sample, outs.gi_receiver, outs.gi_normal)
    ra.check(L.rgi_nrc_gather_vertices(cache.handle, ctypes.byref(frame), ctypes.byref(gbufs[0].gb), ctypes.byref(gi_res),
                                       outs.di_diffuse, outs.color, 65536, 7, big_rec.data_ptr(), big_tgt.data_ptr(), s))
    train_65536_ms = timed(lambda: cache.train(big_rec, big_tgt))
    queries = W * H
    # Waves whose 16 pixels are all sky skip the encoding and the MLP, so the
    # useful work is the surface pixels' queries.
    surface_queries = int((gbufs[0].depth.view(torch.float32) < 0.99999).sum().item())
    infer_ms = {}
    for name, prec in (("bf16", ra.RGI_NRC_BF16), ("fp16", ra.RGI_NRC_FP16), ("fp8", ra.RGI_NRC_FP8)):
        infer_ms[name] = timed(lambda: ra.check(L.rgi_nrc_query_gbuffer(cache.handle, ctypes.byref(frame),
                                                                        ctypes.byref(gbufs[0].gb), None, prec,
                                                                        radiance.data_ptr(), s)))
    L.rgi_renderer_destroy(plain)
    cache.close()
    tflops = surface_queries * NRC_FLOP_PER_QUERY / (infer_ms["bf16"] * 1e-3) / 1e12
    return {"frame_ms_with_cache": round(frame_ms, 5), "frame_ms_plain": round(plain_ms, 5),
            "frame_ms_with_cache_65536": round(frame_ms_65536, 5),
            "gather_ms": round(gather_ms, 5), "train_ms": round(train_ms, 5), "train_batch": batch,
            "train_ms_65536": round(train_65536_ms, 5),
            "query_ms": {k: round(v, 5) for k, v in infer_ms.items()}, "queries": queries,
            "surface_queries": surface_queries,
            "mfma": {"kernel": "k_nrc_infer_gbuffer (bf16)", "achieved": round(tflops, 3), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 6), "flop_per_query": NRC_FLOP_PER_QUERY,
                     "bound": "hash-table gathers (L1 address processing), not the matrix cores"}}


def trace_benchmark(L, ctx, scene, rays, s):
    """rgi_trace over `rays` (closest hit): both launch forms timed with HIP
    events (5 repetitions each), the faster kept on the context, then the
    watertight triangle test on the same rays: its time and how often its hit
    differs from the reference's test. Options are left at their defaults,
    except the launch form."""
    import torch
    count = rays.shape[0]
    rays_t = torch.from_numpy(rays.view(np.uint8).reshape(-1)).cuda()
    hit = torch.empty(count * 4, dtype=torch.uint8, device="cuda")
    dist = torch.empty(count * 4, dtype=torch.uint8, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def trace():
        ra.check(L.rgi_trace(ctx, ctypes.byref(scene), rays_t.data_ptr(), count, hit.data_ptr(), dist.data_ptr(), 0, s))

    def timed(reps=5):
        trace()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            trace()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    variants = {}
    for persistent in (0, 1):
        ra.check(L.rgi_context_set_option(ctx, ra.RGI_OPT_TRACE_PERSISTENT, persistent))
        variants[persistent] = timed()
    best = min(variants, key=variants.get)
    ra.check(L.rgi_context_set_option(ctx, ra.RGI_OPT_TRACE_PERSISTENT, best))
    trace()
    ms = variants[best]
    hit_rate = float((hit.view(torch.int32) >= 0).float().mean().item())
    mt_hit = hit.view(torch.int32).clone()
    ra.check(L.rgi_context_set_option(ctx, ra.RGI_OPT_TRIANGLE_TEST, 1))
    wt_ms = timed()
    wt_hit = hit.view(torch.int32)
    watertight = {"ms": round(wt_ms, 4), "mrays_per_s": round(count / (wt_ms * 1e-3) / 1e6, 2),
                  "hit_rate": round(float((wt_hit >= 0).float().mean().item()), 5),
                  "differs_from_moeller_trumbore": round(float((wt_hit != mt_hit).float().mean().item()), 7)}
    ra.check(L.rgi_context_set_option(ctx, ra.RGI_OPT_TRIANGLE_TEST, 0))
    del rays_t, hit, dist, mt_hit
    return {"rays": count, "persistent": best, "ms_by_variant": {str(k): round(v, 4) for k, v in variants.items()},
            "ms": ms, "mrays_per_s": count / (ms * 1e-3) / 1e6, "hit_rate": hit_rate, "watertight": watertight}


def run_workload(hs, settings, W, H, frames_in_path, steps, warmup, tune_frames, distance=3.17):
    """One workload through the C++ frame driver: the swing camera path's
    G-buffers produced once and resident in HBM, the autotuner, `steps` timed
    frames (wall clock between device synchronizations), then per-pass HIP
    event timing of up to 32 frames through the per-pass entry points."""
    import torch
    T = settings.grid_table_size
    gscene = driver.GpuScene(hs)
    drv = driver.Driver(W, H, T, settings.grid_mode)
    L = ra.lib()
    s = driver.stream()
    cams = [ra.orbit_camera(hs, i, W, H, yaw_of=lambda k: ra.swing_yaw(k, frames_in_path), distance=distance)
            for i in range(frames_in_path)]
    views = [ra.make_view(*c) for c in cams]
    gbufs = []
    for i in range(frames_in_path):
        eye, vp, inv = cams[i]
        f = ra.make_frame(settings, hs, i, eye, vp, inv, eye, vp, W, H, False)
        gbufs.append(driver.GBufferSet(drv.ctx, gscene, f))
    torch.cuda.synchronize()

    rend = ctypes.c_void_p()
    ra.check(L.rgi_renderer_create(ctypes.byref(rend), W, H, ctypes.byref(ra.c_settings(settings)),
                                   ctypes.byref(gscene.scene), hs.radius))

    def render(k):
        i = k % frames_in_path
        ra.check(L.rgi_renderer_render(rend, ctypes.byref(views[i]), ctypes.byref(gbufs[i].gb), s))

    for k in range(warmup):
        render(k)
    # Pick the fastest launch variants on this workload (bit-exact with each other).
    tune_ms = (ctypes.c_float * ra.RGI_TUNE_TRIALS)()
    view_arr = (ra.View * frames_in_path)(*views)
    gb_arr = (ra.GBuffer * frames_in_path)(*[g.gb for g in gbufs])
    chosen = L.rgi_renderer_autotune(rend, view_arr, gb_arr, frames_in_path, tune_frames, tune_ms, s)
    if chosen < 0:
        ra.check(chosen)
    options = ra.tuning_of(chosen)
    tuning = {"gi_temporal_waves": options[ra.RGI_OPT_GI_TEMPORAL_WAVES],
              "gi_spatial_waves": options[ra.RGI_OPT_GI_SPATIAL_WAVES],
              "tile_order": options[ra.RGI_OPT_TILE_ORDER],
              "tile_shape": "32x8" if options[ra.RGI_OPT_TILE_SHAPE] else "16x16",
              "trial_ms": [round(float(v), 5) for v in tune_ms]}
    for k in range(warmup):
        render(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        render(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_frame = elapsed * 1e3 / steps

    def frame_for_step(step):
        i = step % frames_in_path
        p = (step - 1) % frames_in_path
        eye, vp, inv = cams[i]
        peye, pvp, _ = cams[p]
        return ra.make_frame(settings, hs, step, eye, vp, inv, peye, pvp, W, H, step > 0, step > 0), i

    frames = [frame_for_step(k) for k in range(warmup + min(steps, 32))]
    for opt, value in options.items():
        ra.check(L.rgi_context_set_option(drv.ctx, opt, value))
    for k in range(warmup):
        f, i = frames[k]
        drv.render(f, gscene, gbufs[i])
    torch.cuda.synchronize()

    # Per-pass durations with HIP events on the launch stream.
    events = {}

    def hook(name, start):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        events.setdefault(name, []).append(e)

    timed = min(steps, 32)
    for k in range(timed):
        f, i = frames[warmup + k]
        drv.render(f, gscene, gbufs[i], hook)
    torch.cuda.synchronize()
    pass_ms = {}
    for name, evs in events.items():
        pass_ms[name] = sum(evs[j].elapsed_time(evs[j + 1]) for j in range(0, len(evs), 2)) / (len(evs) // 2)
    return {"ms_per_frame": ms_per_frame, "pass_ms": pass_ms, "tuning": tuning, "options": options, "gscene": gscene,
            "drv": drv, "rend": rend, "gbufs": gbufs, "views": views}


def run_batch(cases, mode, steps, warmup, nrc=False, views_per_case=4, yaw_offsets=None,
              forms=(ra.RGI_BATCH_GRAPH, ra.RGI_BATCH_PASSES), singles=True, full=False):
    """Renderers of `cases` as one batch frame: each renderer alone (its own
    frame time), all of them eagerly one after another on one stream, then
    all of them as one rgi_batch_render call in both of its forms: each pass
    over every view in one launch (RGI_BATCH_PASSES) and one graph with each
    renderer on a stream of its own (RGI_BATCH_GRAPH). With `nrc`, every
    renderer has its own neural radiance cache. `forms` and `singles` limit
    the run to some of the batch forms, without the single renderers. With
    `full`, the renderers also run indirect specular, the denoiser and the
    composite."""
    import torch
    from restir_amd import nrc as rnrc
    L = ra.lib()
    ctx = ctypes.c_void_p()
    ra.check(L.rgi_create(ctypes.byref(ctx)))
    es = driver.stream()
    scenes = {}
    rends, frames, names, pixels, tris = [], [], [], [], []
    for c, (name, kind, size, seed, w, h, dist) in enumerate(cases):
        key = (kind, size, seed)
        if key not in scenes:
            hs = ra.HostScene(kind, seed=seed, size=size, extra_lights=4 if kind != ra.RGI_SCENE_ATRIUM else 64,
                              light_seed=1)
            scenes[key] = (hs, driver.GpuScene(hs))
        hs, gscene = scenes[key]
        table = 1 << 20 if w * h <= (1 << 21) else 1 << 21
        st = ra.Settings(grid_table_size=table, grid_mode=mode, denoiser=full, composite=full, specular_gi=full)
        yaw0 = yaw_offsets[c] if yaw_offsets else 0.0
        views, gbs = [], []
        for i in range(views_per_case):
            eye, vp, inv = ra.orbit_camera(hs, i, w, h, yaw_of=lambda k: ra.swing_yaw(k) + yaw0, distance=dist)
            f = ra.make_frame(st, hs, i, eye, vp, inv, eye, vp, w, h, False)
            gbs.append(driver.GBufferSet(ctx, gscene, f))
            views.append(ra.make_view(eye, vp, inv))
        r = ctypes.c_void_p()
        ra.check(L.rgi_renderer_create(ctypes.byref(r), w, h, ctypes.byref(ra.c_settings(st)),
                                       ctypes.byref(gscene.scene), hs.radius))
        if nrc:
            ra.check(L.rgi_renderer_set_nrc(r, ctypes.byref(rnrc.renderer_settings(hs))))
        rends.append(r)
        frames.append((views, gbs))
        names.append("%s %dx%d" % (name, w, h))
        pixels.append(w * h)
        tris.append(hs.triangle_count)
    torch.cuda.synchronize()

    def one(r, views, gbs, k):
        ra.check(L.rgi_renderer_render(r, ctypes.byref(views[k % views_per_case]),
                                       ctypes.byref(gbs[k % views_per_case].gb), es))

    per_scene, eager_ms = [], float("nan")
    for r, (views, gbs) in zip(rends, frames if singles else []):
        for k in range(warmup):
            one(r, views, gbs, k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            one(r, views, gbs, k)
        torch.cuda.synchronize()
        per_scene.append((time.perf_counter() - t0) * 1e3 / steps)
    if singles:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            for r, (views, gbs) in zip(rends, frames):
                one(r, views, gbs, k)
        torch.cuda.synchronize()
        eager_ms = (time.perf_counter() - t0) * 1e3 / steps

    batch = ctypes.c_void_p()
    arr = (ctypes.c_void_p * len(rends))(*[r.value for r in rends])
    ra.check(L.rgi_batch_create(ctypes.byref(batch), arr, len(rends)))
    side = torch.cuda.Stream()
    bs = ctypes.c_void_p(side.cuda_stream)

    def batched(k):
        va = (ra.View * len(rends))(*[v[k % views_per_case] for v, _ in frames])
        ga = (ra.GBuffer * len(rends))(*[g[k % views_per_case].gb for _, g in frames])
        ra.check(L.rgi_batch_render(batch, va, ga, bs))

    batch_ms = {m: float("nan") for m in (ra.RGI_BATCH_GRAPH, ra.RGI_BATCH_PASSES)}
    for mode in forms:
        ra.check(L.rgi_batch_set_mode(batch, mode))
        for k in range(warmup):
            batched(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            batched(k)
        torch.cuda.synchronize()
        batch_ms[mode] = (time.perf_counter() - t0) * 1e3 / steps
    L.rgi_batch_destroy(batch)
    for r in rends:
        L.rgi_renderer_destroy(r)
    L.rgi_destroy(ctx)
    return {"scenes": names, "triangles": tris, "per_scene_ms": [round(v, 4) for v in per_scene],
            "sum_of_scene_ms": round(sum(per_scene), 4), "eager_ms_per_batch_frame": round(eager_ms, 5),
            "graph_batch_ms_per_batch_frame": round(batch_ms[ra.RGI_BATCH_GRAPH], 5),
            "passes_batch_ms_per_batch_frame": round(batch_ms[ra.RGI_BATCH_PASSES], 5),
            "passes_vs_eager": round(batch_ms[ra.RGI_BATCH_PASSES] / eager_ms, 4), "pixels": sum(pixels),
            "gi_pixels_per_s": round(sum(pixels) / (batch_ms[ra.RGI_BATCH_PASSES] * 1e-3), 1), "steps": steps,
            "nrc": nrc}


def atrium_workload(args, W, H, mode):
    """The atrium workload: ~256k triangles, 2^22-cell grid, DI over 64
    point lights and the emissive triangles; the same swing path, with the
    camera inside the courtyard (0.7 scene radii from the centre); then
    traversal over its BVH (the same ray count, cosine rays from points
    uniform over its area). Returns its bench entry and the live run."""
    L = ra.lib()
    ahs = ra.HostScene(ra.RGI_SCENE_ATRIUM, seed=1, size=262144, extra_lights=64, light_seed=1)
    aset = ra.Settings(grid_table_size=1 << 22, grid_mode=mode, denoiser=False, composite=False, specular_gi=False)
    arun = run_workload(ahs, aset, W, H, args.frames, args.atrium_steps, 8, 4, distance=0.7)
    atrium = {"ms_per_frame": round(arun["ms_per_frame"], 5),
              "pass_ms": {k: round(v, 5) for k, v in sorted(arun["pass_ms"].items())},
              "triangles": ahs.triangle_count, "bvh_nodes": ahs.node_count, "di_lights": ahs.light_count,
              "grid_table_size": 1 << 22, "steps": args.atrium_steps, "camera_distance_radii": 0.7,
              "tuning": arun["tuning"]}
    atr = trace_benchmark(L, arun["drv"].ctx, arun["gscene"].scene, surface_rays(ahs, args.rays, seed=7),
                          driver.stream())
    atrium["trace"] = {"rays": atr["rays"], "ms": round(atr["ms"], 4), "mrays_per_s": round(atr["mrays_per_s"], 2),
                       "hit_rate": round(atr["hit_rate"], 4), "persistent": atr["persistent"],
                       "ms_by_variant": atr["ms_by_variant"], "watertight": atr["watertight"]}
    arun["hs"] = ahs
    return atrium, arun


def frame_variant(hs, settings, W, H, frames, steps):
    """The metric's frame with other settings (its own renderer, autotuned
    on the same path): frame time and per-pass times."""
    L = ra.lib()
    r = run_workload(hs, settings, W, H, frames, steps, 8, 4)
    out = {"ms_per_frame": round(r["ms_per_frame"], 5),
           "pass_ms": {k: round(v, 5) for k, v in sorted(r["pass_ms"].items())}, "tuning": r["tuning"]}
    L.rgi_renderer_destroy(r["rend"])
    r["drv"].close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--table", type=int, default=1 << 20)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grid-mode", choices=["atomic", "deterministic"], default="atomic")
    ap.add_argument("--reference-frames", type=int, default=4096)
    ap.add_argument("--mse-frames", type=int, default=64)
    ap.add_argument("--tune-frames", type=int, default=16)
    ap.add_argument("--calibrate-copy", action="store_true", help="run one 256 MiB device copy first")
    ap.add_argument("--atrium-steps", type=int, default=32, help="timed frames of the atrium workload (0: skip)")
    ap.add_argument("--no-batch", dest="batch", action="store_false", help="skip the uneven batch workload")
    ap.add_argument("--no-nrc", dest="nrc", action="store_false", help="skip the neural radiance cache measurements")
    ap.add_argument("--nrc-frames", type=int, default=256, help="frames rendered before the cache's MSE is measured")
    ap.add_argument("--atrium-only", action="store_true",
                    help="run only the atrium workload (--atrium-steps timed frames and its traversal) and print it: "
                         "the workload of the atrium's kernel profile")
    ap.add_argument("--no-variants", dest="variants", action="store_false",
                    help="skip the deterministic-insert and half-resolution GI frames")
    ap.add_argument("--batch-only", choices=["uneven", "views8"],
                    help="run only one batch workload (--steps timed batch frames after --warmup), pass-major form "
                         "only, and print it: the workload of the batch kernel profiles")
    ap.add_argument("--batch-full", action="store_true",
                    help="with --batch-only: the renderers also run indirect specular, the denoiser and the composite")
    ap.add_argument("--batch-all-forms", action="store_true",
                    help="with --batch-only: also time each renderer alone, all of them eagerly, and the graph form")
    ap.add_argument("--cache-frames-only", action="store_true",
                    help="render only the metric's path with the neural radiance cache on (16 warm-up + --steps "
                         "frames) and print their frame time: the workload of the cache's kernel profile")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    torch.cuda.init()
    if args.calibrate_copy:
        # A plain device copy of known size, for scaling HBM counters.
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel: 256 MiB read, 256 MiB written
        torch.cuda.synchronize()
        del src, dst
    W, H, T = args.width, args.height, args.table
    hs = ra.HostScene(ra.RGI_SCENE_CORNELL, seed=1, extra_lights=4, light_seed=1)
    if args.cache_frames_only:
        print(json.dumps(cache_frames(hs, W, H, T, args.frames, args.steps)), flush=True)
        return
    mode = ra.RGI_GRID_ATOMIC if args.grid_mode == "atomic" else ra.RGI_GRID_DETERMINISTIC
    if args.atrium_only:
        print(json.dumps(atrium_workload(args, W, H, mode)[0]), flush=True)
        return
    if args.batch_only:
        if args.batch_only == "uneven":
            cases, kw = UNEVEN_SCENES, {"nrc": args.nrc}
        else:
            cases = [("atrium_256k_view%d" % v, ra.RGI_SCENE_ATRIUM, 262144, 1, W, H, 0.7) for v in range(8)]
            kw = {"yaw_offsets": [v * math.pi / 4 for v in range(8)]}
        if args.batch_all_forms:
            out = run_batch(cases, mode, args.steps, args.warmup, full=args.batch_full, **kw)
        else:
            out = run_batch(cases, mode, args.steps, args.warmup, forms=(ra.RGI_BATCH_PASSES,), singles=False,
                            full=args.batch_full, **kw)
        keep = ("eager_ms_per_batch_frame", "graph_batch_ms_per_batch_frame", "passes_batch_ms_per_batch_frame",
                "passes_vs_eager", "sum_of_scene_ms")
        line = {"batch": args.batch_only, "full": args.batch_full, "steps": args.steps, "warmup": args.warmup}
        line.update({k: out[k] for k in keep if out[k] == out[k]})
        print(json.dumps(line), flush=True)
        return
    # The metric's frame: grid + resample + trace. The denoiser that follows
    # it in the reference is timed on its own below.
    settings = ra.Settings(grid_table_size=T, grid_mode=mode, denoiser=False, composite=False, specular_gi=False)
    run = run_workload(hs, settings, W, H, args.frames, args.steps, args.warmup, args.tune_frames)
    ms_per_frame, pass_ms, tuning, options = run["ms_per_frame"], run["pass_ms"], run["tuning"], run["options"]
    gscene, drv, rend, gbufs, views = run["gscene"], run["drv"], run["rend"], run["gbufs"], run["views"]
    L = ra.lib()
    s = driver.stream()

    # The same frames as one HIP graph per frame (captured, updated in place,
    # replayed) on a stream of their own.
    graph_stream = torch.cuda.Stream()
    gs = ctypes.c_void_p(graph_stream.cuda_stream)
    for k in range(args.warmup):
        i = k % args.frames
        ra.check(L.rgi_renderer_render_graph(rend, ctypes.byref(views[i]), ctypes.byref(gbufs[i].gb), gs))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        i = k % args.frames
        ra.check(L.rgi_renderer_render_graph(rend, ctypes.byref(views[i]), ctypes.byref(gbufs[i].gb), gs))
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    # The denoiser on the last frame's outputs: variance + 5 a-trous steps.
    outs = ra.RendererOutputs()
    ra.check(L.rgi_renderer_outputs_get(rend, ctypes.byref(outs)))
    last = ra.Frame()
    ra.check(L.rgi_renderer_last_frame(rend, ctypes.byref(last)))
    den = [torch.empty(W * H * 8, dtype=torch.uint8, device="cuda") for _ in range(2)]

    def denoise():
        ra.check(L.rgi_svgf_variance(drv.ctx, ctypes.byref(last), outs.color, outs.moments, outs.gi_receiver,
                                     outs.gi_normal, den[0].data_ptr(), s))
        for it in range(5):
            ra.check(L.rgi_svgf_atrous(drv.ctx, ctypes.byref(last), den[it & 1].data_ptr(), outs.gi_receiver,
                                       outs.gi_normal, 1 << it, den[1 - (it & 1)].data_ptr(), s))

    denoise()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        denoise()
    e1.record()
    torch.cuda.synchronize()
    denoiser_ms = e0.elapsed_time(e1) / 10

    # Indirect specular of that frame: one GGX ray per pixel (traced like the GI rays).
    spec_t = torch.empty(W * H * 8, dtype=torch.uint8, device="cuda")
    sgin = ra.GiInputs(None, outs.gi_receiver, outs.gi_normal, outs.color, outs.moments,
                       outs.di_diffuse if outs.di_diffuse else gbufs[0].gb.emission, None, outs.grid)

    def spec():
        ra.check(L.rgi_specular_gi(drv.ctx, ctypes.byref(last), ctypes.byref(gscene.scene), ctypes.byref(gbufs[0].gb),
                                   ctypes.byref(sgin), None, None, spec_t.data_ptr(), s))

    spec()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        spec()
    e1.record()
    torch.cuda.synchronize()
    specular_ms = e0.elapsed_time(e1) / 10

    # The composite of that frame (main light off, as the workload has it).
    hdr_t = torch.empty(W * H * 8, dtype=torch.uint8, device="cuda")
    ci = ra.CompositeInputs(den[1].data_ptr(), outs.di_diffuse, outs.di_specular, spec_t.data_ptr(), None)
    ci.options[:] = [1.0, 1.0, 1.0, 1.0]
    ci.specular_weight = 1.0

    def comp():
        ra.check(L.rgi_composite(drv.ctx, ctypes.byref(last), ctypes.byref(gbufs[0].gb), ctypes.byref(ci),
                                 hdr_t.data_ptr(), s))

    comp()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        comp()
    e1.record()
    torch.cuda.synchronize()
    composite_ms = e0.elapsed_time(e1) / 10

    # The metric's renderer is done. While it and its replayed graph (whose
    # parallel branch runs on streams of the graph's own) stay alive, the cache
    # frame below measures 1.66 ms instead of 1.49 ms: with more live streams
    # than the process's 4 hardware queues, the cache's training stream can
    # share a queue with the frame's stream and run behind it, not beside it.
    L.rgi_renderer_destroy(rend)
    variants = None
    if args.variants:
        # The same frame with the deterministic (bit-exact) grid insert, and
        # at the reference's default GI resolution (half the window each way,
        # settings.h:218) over the full-resolution G-buffer and DI.
        det = ra.Settings(**dict(settings.__dict__, grid_mode=ra.RGI_GRID_DETERMINISTIC))
        half = ra.Settings(**dict(settings.__dict__, gi_divisor=2))
        variants = {"deterministic_insert": frame_variant(hs, det, W, H, args.frames, args.steps),
                    "half_resolution_gi": dict(frame_variant(hs, half, W, H, args.frames, args.steps),
                                               gi_width=W // 2, gi_height=H // 2)}
    # The neural radiance cache's cost on the metric's path.
    nrc = nrc_timing(L, hs, settings, views, gbufs, gscene, W, H, args.frames, 64, ms_per_frame, s) if args.nrc else None

    # Traversal microbenchmark.
    tr = trace_benchmark(L, drv.ctx, gscene.scene, surface_rays(hs, args.rays, seed=7), s)
    tuning["trace_persistent"] = tr["persistent"]
    tuning["trace_ms"] = tr["ms_by_variant"]
    trace_ms, mrays, hit_rate, watertight = tr["ms"], tr["mrays_per_s"], tr["hit_rate"], tr["watertight"]

    dominant = max((k for k in pass_ms if k in BYTES_PER_PIXEL), key=lambda k: pass_ms[k])
    traffic = BYTES_PER_PIXEL[dominant] * W * H
    achieved = traffic / (pass_ms[dominant] * 1e-3) / 1e9
    roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "algorithmic_bytes": traffic, "avg_ms": round(pass_ms[dominant], 5)}
    measured = measured_traffic(dominant, options, W, H)
    if measured is not None:
        roofline["traffic"] = measured.pop("bytes")
        roofline["traffic_detail"] = measured

    mse = None
    if args.reference_frames > 0:
        mse, acc, acc_gi = measure_mse(L, hs, gscene, gbufs[0], views[0], W, H, T, mode, args.reference_frames,
                                       args.mse_frames, drv.ctx, s)
        if args.nrc:
            nrc["quality"] = nrc_quality(L, hs, gscene, gbufs[0], views[0], W, H, T, mode, acc, acc_gi,
                                         args.reference_frames, args.nrc_frames, drv.ctx, s)
            nrc["quality_moving_camera"] = nrc_motion_quality(L, hs, gscene, views, gbufs, T, mode, drv.ctx, s)
        del acc, acc_gi

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(hs, settings, W, H, T)

    drv.close()

    atrium = None
    if args.atrium_steps > 0:
        atrium, arun = atrium_workload(args, W, H, mode)
        if args.nrc and args.reference_frames > 0 and nrc is not None:
            # The resolved cache on the atrium (static camera, view 0).
            nrc["quality_atrium"] = resolve_quality(L, arun["hs"], arun["gscene"], arun["gbufs"][0], arun["views"][0],
                                                    1 << 22, mode, arun["drv"].ctx, s)
        L.rgi_renderer_destroy(arun["rend"])
        arun["drv"].close()

    uneven = views8 = None
    if args.batch:
        # 16 scenes of 1k to 1M triangles at 640x360 to 2560x1440, each with
        # its own neural radiance cache, as one batch frame.
        uneven = run_batch(UNEVEN_SCENES, mode, 8, 8, nrc=args.nrc)
        # The atrium seen from 8 independent cameras (yaw offsets of 45
        # degrees) at 1080p, as one batch frame.
        atrium8 = [("atrium_256k_view%d" % v, ra.RGI_SCENE_ATRIUM, 262144, 1, W, H, 0.7) for v in range(8)]
        views8 = run_batch(atrium8, mode, 8, 4, yaw_offsets=[v * math.pi / 4 for v in range(8)])

    line = {
        "metric": "GI frame ms @1080p (grid + resample + trace)",
        "value": round(ms_per_frame, 5),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural seeded Cornell box, G-buffers from primary rays)",
        "config": {"workload": "cornell_1080p_swing%d" % args.frames, "width": W, "height": H, "grid_table_size": T,
                   "frames": args.frames, "di_lights": hs.light_count, "grid_mode": args.grid_mode,
                   "parallelism": "single GPU"},
        "frames_per_s": round(1e3 / ms_per_frame, 2),
        "ms_per_frame_graph": round(graph_ms, 5),
        "pass_ms": {k: round(v, 5) for k, v in sorted(pass_ms.items())},
        "denoiser_ms": round(denoiser_ms, 5),
        "composite_ms": round(composite_ms, 5),
        "specular_gi_ms": round(specular_ms, 5),
        "mrays_per_s": round(mrays, 2),
        "trace": {"rays": args.rays, "ms": round(trace_ms, 4), "hit_rate": round(hit_rate, 4), "mode": "closest",
                  "triangle_test": "moeller_trumbore (the reference's)", "watertight": watertight},
        "mse_vs_reference": mse,
        "tuning": tuning,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "variants": variants,
        "atrium_1080p": atrium,
        "uneven_batch": uneven,
        "atrium_8_views_batch": views8,
        "neural_radiance_cache": nrc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
