"""Pins of the oracle: each check ties an oracle function to something it did
not derive from itself — a published known answer, an independent library,
a closed form, an invariant, or brute force on a small input — chosen so that
a plausible mistake (a dropped term, a wrong sign, index or constant) fails.

The reference ships no tests, golden vectors or fixtures for this path, so
these are the pins. Each test's name or docstring says which oracle function
it covers and what it is checked against; a few tests also check the host
library's exports, ABI and argument checks without a GPU.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import restir_amd as ra

import common
import orc

L = orc.lib()


def f3(*v):
    return np.array(v, np.float32)


# ---------------------------------------------------------------- hashes

def test_checksum_hash_is_xxh32_avalanche():
    """gridChecksumHash (hashgrid.sh:69-78) is XXH32's final avalanche:
    XXH32(empty input, seed) = avalanche(seed + PRIME32_5), checked against
    the xxhash library for many seeds."""
    xxhash = pytest.importorskip("xxhash")
    prime5 = 374761393
    rng = np.random.default_rng(1)
    for seed in [0, 1, 2, 0xFFFFFFFF] + rng.integers(0, 2**32, 2000).tolist():
        assert L.orc_checksum_hash((seed + prime5) & 0xFFFFFFFF) == xxhash.xxh32_intdigest(b"", seed=int(seed))


def test_pcg_is_a_permutation_with_avalanche():
    """gridHash/pcg (hashgrid.sh:61-67): an LCG step followed by the RXS-M-XS
    output permutation, both bijections on uint32. A wrong shift or constant
    breaks injectivity or the avalanche below."""
    rng = np.random.default_rng(2)
    xs = np.unique(rng.integers(0, 2**32, 1 << 18, dtype=np.uint64)).astype(np.uint32)
    ys = np.array([L.orc_pcg(int(x)) for x in xs], np.uint64)
    assert len(np.unique(ys)) == len(xs)
    # Consecutive inputs: a bijection has no collisions on a dense range either.
    dense = np.array([L.orc_pcg(i) for i in range(1 << 16)], np.uint64)
    assert len(np.unique(dense)) == 1 << 16
    # Strict avalanche: flipping any input bit flips about half the output bits.
    sample = xs[:512]
    for bit in range(32):
        flips = [bin(L.orc_pcg(int(x)) ^ L.orc_pcg(int(x) ^ (1 << bit))).count("1") for x in sample]
        assert 12.0 < np.mean(flips) < 20.0, bit


def test_next_random_is_exact_24_bit():
    """nextRandom (restir.sh:131-135): (pcg(state) >> 8) / 2^24, exactly."""
    st = ctypes.c_uint32(12345)
    for _ in range(1000):
        before = st.value
        r = L.orc_next_random(ctypes.byref(st))
        assert st.value == L.orc_pcg(before)
        assert r * 16777216.0 == float(st.value >> 8)
        assert 0.0 <= r < 1.0


# ----------------------------------------------------------- math, fp16

def ulp32(x):
    x = np.float32(abs(x))
    return float(np.spacing(x)) if x != 0 else float(np.finfo(np.float32).tiny)


def test_sincos_against_libm():
    """rgi_sincos (Cody-Waite + Taylor) within 2 ulp of the true values."""
    xs = np.concatenate([np.linspace(-64, 64, 40001, dtype=np.float32),
                         (np.float32(6.2831855) * np.linspace(0, 1, 20001, dtype=np.float32)).astype(np.float32),
                         np.array([0.0, 1e-30, -1e-7, math.pi / 4, math.pi / 2, math.pi], np.float32)])
    s, c = ctypes.c_float(), ctypes.c_float()
    worst = 0.0
    for x in xs:
        L.orc_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
        es, ec = math.sin(float(x)), math.cos(float(x))
        # Error relative to the ulp of the larger of the value and 2^-24 (absolute floor near zeros).
        worst = max(worst, abs(s.value - es) / max(ulp32(es), 2**-26), abs(c.value - ec) / max(ulp32(ec), 2**-26))
    assert worst <= 2.0, worst


def test_half_bits_match_numpy_rne():
    """RGBA16F stores: IEEE binary16 round to nearest even (numpy's conversion)."""
    rng = np.random.default_rng(3)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 100,
        (rng.random(20000) * 6e-5).astype(np.float32),             # binary16 subnormals
        np.array([65504, 65519.99, 65520, 1e9, -1e9, 0.0, -0.0, 2**-25, 2**-24, 3 * 2**-26, np.inf, -np.inf],
                 np.float32),
        # exact ties between neighbouring halves
        (np.arange(1, 2049, dtype=np.float32) + 0.5) * np.float32(2**-10),
    ])
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    got = np.array([L.orc_half_bits(float(v)) for v in vals], np.uint16)
    assert np.array_equal(got, want)
    nan = L.orc_half_bits(float("nan"))
    assert (nan & 0x7C00) == 0x7C00 and (nan & 0x3FF) != 0


# ----------------------------------------------------------- hash grid

def frame_with(**kw):
    f = ra.Frame()
    f.grid_min_cell = kw.get("min_cell", 0.0198)
    f.grid_pixel_size = kw.get("pixel_size", 0.0037)
    f.grid_cell_pixels = 6.0
    f.grid_table_size = kw.get("table", 65536)
    return f


def test_grid_cell_size_closed_form():
    """gridCellSize (hashgrid.sh:82-87) = x * 2^floor(log2(max(d*y*z/x, 1)))."""
    f = frame_with()
    rng = np.random.default_rng(4)
    cam = f3(0, 0, 0)
    for d in np.concatenate([rng.random(3000) * 50, [0.0, 1e-3, 0.5, 1.0, 1.1]]):
        p = f3(d, 0, 0)
        got = L.orc_grid_cell_size(ctypes.byref(f), orc.ptr(p), orc.ptr(cam))
        target = float(np.float32(np.float32(np.float32(d) * np.float32(f.grid_pixel_size)) * np.float32(6.0)))
        m = max(float(np.float32(target / np.float32(f.grid_min_cell))), 1.0)
        lvl = math.floor(math.log2(m))
        if abs(m / 2**round(math.log2(m)) - 1) < 1e-6:
            continue  # exactly at a power of two, where log2 rounding decides
        assert got == np.float32(f.grid_min_cell) * np.float32(2.0**lvl)


def test_grid_normal_bucket_table():
    """gridNormalBucket (hashgrid.sh:89-103): dominant axis and sign; ties
    go to x, then y."""
    cases = {(1, 0, 0): 0, (-1, 0, 0): 1, (0, 1, 0): 2, (0, -1, 0): 3, (0, 0, 1): 4, (0, 0, -1): 5,
             (0.5, 0.5, 0.5): 0, (-0.5, 0.5, 0.5): 1, (0, 0.5, 0.5): 2, (0, -0.5, 0.5): 3, (0.1, 0.2, -0.9): 5}
    for n, want in cases.items():
        v = f3(*n)
        assert L.orc_grid_normal_bucket(orc.ptr(v)) == want, n


def test_grid_key_cell_and_range():
    """gridKey (hashgrid.sh:116-127): one key per (cell, size, bucket): points
    of a cell share it, the next cell differs; slot < T; checksum never 0."""
    f = frame_with(table=1024)
    rng = np.random.default_rng(5)
    slot, cs = ctypes.c_uint32(), ctypes.c_uint32()
    n = f3(0, 1, 0)
    seen = {}
    for _ in range(2000):
        size = np.float32(0.0198 * 2 ** rng.integers(0, 6))
        cell = rng.integers(-50, 50, 3)
        inside = (cell + 0.05 + 0.9 * rng.random(3)) * float(size)
        inside = inside.astype(np.float32)
        L.orc_grid_key(ctypes.byref(f), orc.ptr(inside), orc.ptr(n), float(size),
                       ctypes.byref(slot), ctypes.byref(cs))
        assert slot.value < 1024 and cs.value != 0
        key = (tuple(cell), float(size))
        if key in seen:
            assert seen[key] == (slot.value, cs.value)
        seen[key] = (slot.value, cs.value)
    # Distinct cells give distinct checksums (32-bit, few collisions expected).
    sums = [v[1] for v in seen.values()]
    assert len(set(sums)) >= len(sums) - 1


def brute_insert(slot0, checksum, table):
    """Serial simulation of the deterministic linearization: probe rounds,
    ascending pixels within a round, then adds in pixel order."""
    grid = np.zeros(34 * table, np.uint32)
    resolved = {}
    pending = [i for i in range(len(checksum)) if checksum[i] != 0]
    for probe in range(8):
        nxt = []
        for i in pending:
            s = (int(slot0[i]) + probe) & (table - 1)
            if grid[s] == 0:
                grid[s] = checksum[i]
                resolved[i] = s
            elif grid[s] == checksum[i]:
                resolved[i] = s
            else:
                nxt.append(i)
        pending = nxt
    for i in sorted(resolved):
        s = resolved[i]
        k = grid[table + s]
        grid[table + s] += 1
        if k < 32:
            grid[2 * table + s * 32 + k] = i
    return grid, resolved


@pytest.fixture(scope="module")
def small_seq():
    hs = common.cornell(extra_lights=4)
    seq = common.oracle_sequence(hs, ra.Settings(), 64, 48, 2, yaw_step=0.2)
    return hs, seq


@pytest.mark.parametrize("table", [65536, 256, 64])
def test_grid_insert_matches_serial_linearization(small_seq, table):
    hs, seq = small_seq
    f, g, hist, out = seq[1]
    f = ra.Frame.from_buffer_copy(f)
    f.grid_table_size = table
    gb = common.host_gbuffer(g, 64, 48)
    n = 64 * 48
    slot0, cs = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    L.orc_grid_keys(ctypes.byref(f), ctypes.byref(gb), orc.ptr(slot0), orc.ptr(cs), None)
    assert np.all(slot0 < table)
    want, resolved = brute_insert(slot0, cs, table)
    got = np.zeros(34 * table, np.uint32)
    L.orc_grid_insert(ctypes.byref(f), ctypes.byref(gb), orc.ptr(got))
    assert np.array_equal(got, want)
    # Every resolved pixel is found again by a probe walk from its own key.
    for i, s in list(resolved.items())[:500]:
        found = None
        for probe in range(8):
            t = (int(slot0[i]) + probe) & (table - 1)
            if got[t] == cs[i]:
                found = t
                break
            if got[t] == 0:
                break
        assert found == s
    if table == 64:
        assert len(resolved) < int((cs != 0).sum())  # the tiny table drops pixels


def test_grid_pick_entry_reads_listed_pixels(small_seq):
    """gridPickEntry (hashgrid.sh:155-177): a pick is a listed pixel of the
    jittered key's cell, chosen by r2 over min(count, 32)."""
    hs, seq = small_seq
    f, g, hist, out = seq[1]
    grid = out["grid"]
    T = f.grid_table_size
    cam = np.array(list(f.camera_position)[:3], np.float32)
    rng = np.random.default_rng(6)
    hits = 0
    for i in rng.integers(0, 64 * 48, 400):
        if g["depth"][i] >= 0.99999:
            continue
        p = out["gi_receiver"][i, :3].copy()
        nrm = out["gi_normal"][i, :3].copy()
        r = rng.random(3).astype(np.float32)
        e = L.orc_grid_pick_entry(ctypes.byref(f), orc.ptr(grid), orc.ptr(cam), orc.ptr(p), orc.ptr(nrm),
                                  float(r[0]), float(r[1]), float(r[2]))
        if e == 0xFFFFFFFF:
            continue
        hits += 1
        slots = np.nonzero(grid[:T])[0]
        listed = {int(v) for s in slots for v in grid[2 * T + s * 32: 2 * T + s * 32 + min(int(grid[T + s]), 32)]}
        assert e in listed
    assert hits > 50


# ------------------------------------------------------------------ BVH

@pytest.mark.parametrize("kind,size", [(ra.RGI_SCENE_CORNELL, 0), (ra.RGI_SCENE_BOXFIELD, 300),
                                       (ra.RGI_SCENE_ATRIUM, 20000)])
def test_bvh_layout_invariants(kind, size):
    """The bvh.h:9-22 layout: DFS order, left child next, right child and skip
    link consistent, child bounds inside the parent, every triangle in exactly
    one leaf, records carrying the mesh's triangles."""
    hs = ra.HostScene(kind, seed=2, size=size)
    nodes = hs.bvh[: hs.node_count * 8].reshape(-1, 8)
    seen = np.zeros(hs.triangle_count, np.int32)

    def walk(i, lo, hi):
        n = nodes[i]
        assert np.all(n[0:3] >= lo - 1e-6) and np.all(n[4:7] <= hi + 1e-6)
        if n[7] >= 0:
            a, c = int(n[3]), int(n[7])
            assert c >= 1 and (c <= 16 or hs.bvh_depth >= 40)
            seen[a:a + c] += 1
            return i + 1
        right = walk(i + 1, n[0:3], n[4:7])
        assert right == int(n[3])
        end = walk(right, n[0:3], n[4:7])
        assert end == int(-n[7]) - 1
        return end

    assert walk(0, nodes[0, 0:3], nodes[0, 4:7]) == hs.node_count
    assert np.all(seen == 1)
    rec = hs.bvh[hs.node_count * 8:].reshape(-1, 16)
    mine = {tuple(np.round(r, 6)) for r in rec[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]]}
    theirs = {tuple(np.round(p, 6)) for p in hs.positions}
    assert mine == theirs


def brute_closest(hs, o, d, tmax):
    p = hs.bvh[hs.node_count * 8:].reshape(-1, 16).astype(np.float64)
    v0, v1, v2 = p[:, 0:3], p[:, 4:7], p[:, 8:11]
    e1, e2 = v1 - v0, v2 - v0
    pv = np.cross(d, e2)
    det = (e1 * pv).sum(1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / det
        s = o - v0
        u = (s * pv).sum(1) * inv
        q = np.cross(s, e1)
        v = (q * d).sum(1) * inv
        t = (e2 * q).sum(1) * inv
    ok = (np.abs(det) > 1e-12) & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0) & (t < tmax)
    return np.where(ok, t, np.inf)


@pytest.mark.parametrize("kind,size", [(ra.RGI_SCENE_CORNELL, 0), (ra.RGI_SCENE_BOXFIELD, 60)])
def test_trace_against_brute_force(kind, size):
    """traceBvh (bvh.sh:140-319): closest hit = the nearest triangle over all
    triangles (ties: any of the equally near ones); any hit = whether one
    exists."""
    hs = ra.HostScene(kind, seed=4, size=size)
    rng = np.random.default_rng(7)
    n = 3000
    o = (hs.center + (rng.random((n, 3)) - 0.5) * 1.6 * hs.radius).astype(np.float32)
    d = rng.standard_normal((n, 3))
    d = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 3] = o, d, 3.0 * hs.radius
    scene = orc.host_scene(hs)
    for any_hit in (0, 1):
        tri, dist = np.zeros(n, np.int32), np.zeros(n, np.float32)
        L.orc_trace(ctypes.byref(scene), orc.ptr(rays), n, orc.ptr(tri), orc.ptr(dist), any_hit)
        agree = 0
        for i in range(n):
            t = brute_closest(hs, o[i].astype(np.float64), d[i].astype(np.float64), 3.0 * hs.radius)
            tmin = t.min()
            if not np.isfinite(tmin):
                agree += tri[i] < 0
                continue
            if any_hit:
                agree += tri[i] >= 0
                continue
            if tri[i] >= 0 and abs(dist[i] - tmin) <= 1e-4 * max(tmin, 1.0) and t[tri[i]] <= tmin * (1 + 1e-5) + 1e-6:
                agree += 1
        # Grazing rays at triangle edges may differ between float and double.
        assert agree >= n - 3, (any_hit, n - agree)


def test_surface_view_depth_is_independent_of_uv():
    """depthBehindSurface (restir.sh:327-358) compares the ray's view depth
    with that of reconstructWorldPosition(uv, depth) re-projected. The march
    restates the latter as 1 / (c z + d), (0, 0, c, d) being the w row of the
    inverse view-projection: for a perspective camera the re-projected w does
    not depend on uv. Checked in double for random uv and depth on the
    bench's cameras, both depth conventions, to 1e-5: the float matrices are
    each other's inverses only to float precision."""
    hs = common.cornell(extra_lights=0)
    rng = np.random.default_rng(12)
    for hd in (False, True):
        for frame in range(4):
            eye = ra.orbit_camera(hs, frame, 1920, 1080, yaw_of=ra.swing_yaw, homogeneous_depth=hd)
            vp = np.array(eye[1], np.float64).reshape(4, 4).T
            inv = np.array(eye[2], np.float64).reshape(4, 4).T
            uv = rng.random((2000, 2))
            z = rng.uniform(0.2, 0.9999, 2000)
            zc = z * 2 - 1 if hd else z
            clip = np.stack([uv[:, 0] * 2 - 1, -(uv[:, 1] * 2 - 1), zc, np.ones(2000)], 1)
            wp = clip @ inv.T
            world = wp[:, :3] / wp[:, 3:4]
            w_ref = np.concatenate([world, np.ones((2000, 1))], 1) @ vp[3]
            w_fast = 1.0 / (inv[3, 2] * zc + inv[3, 3])
            assert np.allclose(w_fast, w_ref, rtol=1e-5), (hd, frame, np.abs(w_fast / w_ref - 1).max())


# -------------------------------------------------------------- denoiser

def test_expf_against_libm():
    """rgi_expf (the denoiser's exp) is within 2 ulp of the true exp over the
    whole float range, including subnormal results, overflow and underflow."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-104.0, 88.72, 60000), rng.uniform(-20.0, 0.0, 20000),
                         np.array([0.0, -0.0, 1e-30, -1e-30, 1.0, -1.0, 88.7228, -87.3, -103.9, -104.5, 89.0,
                                   float("inf"), float("-inf")])]).astype(np.float32)
    worst = 0
    for x in xs:
        got = np.float32(L.orc_expf(float(x)))
        want = np.float32(math.exp(float(x))) if float(x) < 89.0 else np.float32(np.inf)
        if not np.isfinite(want) or want == 0.0:
            assert got == want or (want == 0.0 and got <= np.float32(2.0 ** -149)), (x, got, want)
            continue
        ulp = abs(int(got.view(np.int32)) - int(want.view(np.int32)))
        worst = max(worst, ulp)
    assert worst <= 2, worst
    assert math.isnan(L.orc_expf(float("nan")))


def test_pow64_is_the_64th_power():
    """kNormalPower (svgf.sh:14): six squarings; their roundings compound to
    at most 63 half-ulps (2^-24 relative each) of x^64."""
    x = np.random.default_rng(6).random(5000).astype(np.float32)
    got = np.array([L.orc_pow64(float(v)) for v in x], np.float64)
    want = x.astype(np.float64) ** 64
    ok = want > 1e-30
    assert np.all(np.abs(got[ok] - want[ok]) <= 64 * 2.0 ** -24 * want[ok])


def svgf_frame(w, h, scale=1.0):
    f = ra.Frame()
    f.width, f.height, f.gi_width, f.gi_height, f.scene_scale = w, h, w, h, scale
    return f


def half4(rgb, a):
    out = np.zeros(rgb.shape[:-1] + (4,), np.float16)
    out[..., :3], out[..., 3] = rgb, a
    return out.view(np.uint16)


def wall(w, h, normal=(0.0, 0.0, 1.0)):
    """A flat wall z = 0 seen face on: receiver positions on a 0.01 grid."""
    rec = np.zeros((h, w, 4), np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    rec[..., 0], rec[..., 1] = xx * 0.01, yy * 0.01
    nor = np.zeros((h, w, 4), np.float32)
    nor[..., :3] = normal
    return rec, nor


def test_svgf_variance_closed_forms():
    """fs_svgf_variance.sc: with 4+ frames of history the pixel's own
    moments give the variance; with less, the mean moments of its 7 x 7
    in-image window (all on one plane, so every geometry weight is 1),
    times 4 / history; sky texels are zero."""
    w, h = 24, 16
    f = svgf_frame(w, h)
    rec, nor = wall(w, h)
    nor[0, 0, :3] = 0.0  # one sky texel
    rng = np.random.default_rng(8)
    m1 = rng.uniform(0.1, 1.0, (h, w)).astype(np.float16)
    m2 = (m1.astype(np.float32) ** 2 + rng.uniform(0.0, 0.5, (h, w))).astype(np.float16)
    history = np.where(rng.random((h, w)) < 0.5, 2.0, 6.0).astype(np.float16)
    moments = np.zeros((h, w, 4), np.float16)
    moments[..., 0], moments[..., 1], moments[..., 2] = m1, m2, history
    color = half4(rng.random((h, w, 3)).astype(np.float32), history.astype(np.float32))
    out = np.zeros((h, w, 4), np.uint16)
    L.orc_svgf_variance(ctypes.byref(f), orc.ptr(color), orc.ptr(moments.view(np.uint16)), orc.ptr(rec),
                        orc.ptr(nor), orc.ptr(out))
    got = out.view(np.float16).astype(np.float64)
    assert np.all(got[0, 0] == 0.0)
    a, b = m1.astype(np.float64), m2.astype(np.float64)
    sky = np.zeros((h, w), bool)
    sky[0, 0] = True
    for y in range(h):
        for x in range(w):
            if sky[y, x]:
                continue
            assert np.array_equal(got[y, x, :3], color.view(np.float16)[y, x, :3].astype(np.float64))
            if history[y, x] >= 4:
                want = max(b[y, x] - a[y, x] ** 2, 0.0)
            else:
                ys, xs = slice(max(y - 3, 0), y + 4), slice(max(x - 3, 0), x + 4)
                keep = ~sky[ys, xs]
                ma, mb = a[ys, xs][keep].mean(), b[ys, xs][keep].mean()
                want = max(mb - ma * ma, 0.0) * 4.0 / max(float(history[y, x]), 1.0)
            assert abs(got[y, x, 3] - want) <= 2e-3 * max(want, 1e-2), (y, x, got[y, x, 3], want)


def run_atrous(f, inp, rec, nor, step):
    out = np.zeros(inp.shape, np.uint16)
    L.orc_svgf_atrous(ctypes.byref(f), orc.ptr(inp), orc.ptr(rec), orc.ptr(nor), step, orc.ptr(out))
    return out.view(np.float16).astype(np.float64)


def test_svgf_atrous_kernel_edges_and_constants():
    """fs_svgf_atrous.sc: (1) with a huge variance the luminance weight is 1
    and an impulse spreads as the separable B3 spline (3/8, 1/4, 1/16) at
    taps `step` texels apart; (2) a constant image stays constant; (3) no
    weight crosses to a perpendicular surface (pow(cos, 64) = 0)."""
    w, h = 40, 40
    f = svgf_frame(w, h)
    rec, nor = wall(w, h)
    for step in (1, 4):
        rgb = np.zeros((h, w, 3), np.float32)
        rgb[20, 20] = 1.0
        got = run_atrous(f, half4(rgb, np.full((h, w), 60000.0, np.float32)), rec, nor, step)
        spline = {0: 0.375, 1: 0.25, 2: 0.0625}
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                want = spline[abs(dx)] * spline[abs(dy)]
                v = got[20 + dy * step, 20 + dx * step, 0]
                assert abs(v - want) <= 2e-3 * want + 1e-4, (step, dx, dy, v, want)
        mask = np.zeros((h, w), bool)
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                mask[20 + dy * step, 20 + dx * step] = True
        assert np.all(got[..., 0][~mask] == 0.0)
    const = half4(np.full((h, w, 3), 0.3, np.float32), np.full((h, w), 0.01, np.float32))
    got = run_atrous(f, const, rec, nor, 2)
    assert np.all(np.abs(got[..., :3] - float(np.float16(0.3))) <= 1e-3)
    assert np.all(got[..., 3] <= 0.01 + 1e-5)
    # Left half faces +z, right half faces +x, in different colours.
    rec2, nor2 = wall(w, h)
    nor2[:, w // 2:, :3] = (1.0, 0.0, 0.0)
    rgb = np.zeros((h, w, 3), np.float32)
    rgb[:, : w // 2] = 0.2
    rgb[:, w // 2:] = 0.8
    got = run_atrous(f, half4(rgb, np.full((h, w), 60000.0, np.float32)), rec2, nor2, 1)
    assert np.all(np.abs(got[:, : w // 2, 0] - float(np.float16(0.2))) <= 1e-3)
    assert np.all(np.abs(got[:, w // 2:, 0] - float(np.float16(0.8))) <= 1e-3)


def test_exp2f_against_libm():
    """rgi_exp2f (specularAlbedo's exp2, brdf.sh:73) within 2 ulp of 2^x."""
    rng = np.random.default_rng(9)
    xs = np.concatenate([rng.uniform(-150.0, 128.0, 40000), rng.uniform(-10.0, 0.0, 20000),
                         np.array([0.0, -1.0, 1.0, -9.28, 127.9, -126.5, -149.2, -150.5, 128.0])]).astype(np.float32)
    worst = 0
    for x in xs:
        got = np.float32(L.orc_exp2f(float(x)))
        want = np.float32(2.0 ** float(x)) if float(x) < 128.0 else np.float32(np.inf)
        if not np.isfinite(want) or want == 0.0:
            assert got == want or (want == 0.0 and got <= np.float32(2.0 ** -149)), (x, got, want)
            continue
        worst = max(worst, abs(int(got.view(np.int32)) - int(want.view(np.int32))))
    assert worst <= 2, worst


def test_specular_albedo_is_karis_fit():
    """specularAlbedo (brdf.sh:68-76) is Karis' analytic approximation of the
    split-sum environment BRDF ("Physically Based Shading on Mobile", 2014):
    r = c0 roughness + c1, a004 = min(r.x^2, 2^(-9.28 n.v)) r.x + r.y,
    (A, B) = (-1.04, 1.04) a004 + r.zw, albedo = f0 A + B."""
    c0 = np.array([-1.0, -0.0275, -0.572, 0.022])
    c1 = np.array([1.0, 0.0425, 1.04, -0.04])
    rng = np.random.default_rng(10)
    for _ in range(500):
        f0 = rng.random(3).astype(np.float32)
        rough, nv = float(rng.random()), float(rng.random())
        r = c0 * rough + c1
        a004 = min(r[0] * r[0], 2.0 ** (-9.28 * nv)) * r[0] + r[1]
        a, b = -1.04 * a004 + r[2], 1.04 * a004 + r[3]
        want = f0.astype(np.float64) * a + b
        out = np.zeros(3, np.float32)
        L.orc_specular_albedo(orc.ptr(f0), rough, nv, orc.ptr(out))
        assert np.allclose(out, want, rtol=1e-5, atol=1e-6), (f0, rough, nv, out, want)


def composite_inputs(gi, dd=None, ds=None, spec=None, shadow=None, options=(1.0, 1.0, 1.0, 1.0), spec_w=0.0):
    ci = ra.CompositeInputs()
    ci.gi = gi.ctypes.data
    ci.di_diffuse = dd.ctypes.data if dd is not None else None
    ci.di_specular = ds.ctypes.data if ds is not None else None
    ci.specular = spec.ctypes.data if spec is not None else None
    ci.shadow_map = shadow.ctypes.data if shadow is not None else None
    ci.options[:] = list(options)
    ci.specular_weight = spec_w
    return ci


def test_composite_closed_forms():
    """fs_composite.sc with the main light off and no DI: a constant GI
    irradiance c comes back as c x albedo x (1 - metallic) x occlusion plus
    the emission on every surface pixel (the same-surface upsample weights
    renormalise to 1); with options.z = 0 the albedo is not multiplied in;
    sky pixels are the sky gradient (sky.sh:20-31) when the sky is on."""
    hs = common.cornell(extra_lights=0)
    w, h = 64, 48
    s = ra.Settings(sky=True)
    eye, vp, inv = ra.orbit_camera(hs, 0, w, h)
    f = ra.make_frame(s, hs, 0, eye, vp, inv, eye, vp, w, h, False)
    g = common.oracle_gbuffer(hs, f)
    gb = common.host_gbuffer(g, w, h)
    c = np.array([0.25, 0.5, 1.0], np.float32)
    gi = np.zeros((w * h, 4), np.float16)
    gi[:, :3] = c
    gi = gi.view(np.uint16)
    for modulate in (1.0, 0.0):
        out = np.zeros((w * h, 4), np.uint16)
        ci = composite_inputs(gi, options=(1.0, 1.0, modulate, 1.0))
        L.orc_composite(ctypes.byref(f), ctypes.byref(gb), ctypes.byref(ci), orc.ptr(out))
        hdr = out.view(np.float16).astype(np.float64)
        geo = g["depth"] < 0.99999
        albedo = np.stack([(g["albedo"] >> (8 * k)) & 255 for k in range(3)], 1) / 255.0
        metallic = (g["normal"] >> 24) / 255.0
        emission = g["emission"].view(np.float16).astype(np.float64)
        scale = albedo * (1.0 - metallic)[:, None] * emission[:, 3:4] if modulate else 1.0
        want = (c.astype(np.float64) * scale + emission[:, :3])[geo] if modulate else (c + emission[:, :3])[geo]
        assert np.allclose(hdr[geo, :3], want, rtol=2e-3, atol=1e-4)
        assert np.all(hdr[geo, 3] == 1.0)
        # Sky: the gradient between horizon, zenith and ground along the view ray.
        sky = ~geo
        ys, xs = np.divmod(np.arange(w * h)[sky], w)
        ndc = np.stack([(xs + 0.5) / w * 2 - 1, -((ys + 0.5) / h * 2 - 1), np.ones(len(xs)), np.ones(len(xs))], 1)
        wp = ndc @ np.array(f.inv_view_proj, np.float64).reshape(4, 4)
        d = wp[:, :3] / wp[:, 3:4] - np.array(eye, np.float64)
        d /= np.linalg.norm(d, axis=1)[:, None]
        hgt = np.clip(d[:, 1], -1, 1)
        zen, hor, gnd = (np.array(f.sky_zenith[:3]), np.array(f.sky_horizon[:3]), np.array(f.sky_ground[:3]))
        t = np.sqrt(np.abs(hgt))[:, None]
        want_sky = np.where(hgt[:, None] >= 0, hor * (1 - t) + zen * t, hor * (1 - t) + gnd * t)
        assert np.allclose(hdr[sky, :3], want_sky, rtol=2e-3, atol=1e-3)
        assert np.all(hdr[sky, 3] == 1.0)


# ----------------------------------------------------------- specular GI

def specular_samples(n, v, roughness, count, seed):
    rng = np.random.default_rng(seed)
    r = rng.random((count, 2)).astype(np.float32)
    n3, v3 = np.asarray(n, np.float32), np.asarray(v, np.float32)
    dirs = np.zeros((count, 3), np.float32)
    g1 = np.zeros(count, np.float32)
    out_d, out_g = np.zeros(3, np.float32), np.zeros(1, np.float32)
    for i in range(count):
        L.orc_specular_sample(orc.ptr(n3), orc.ptr(v3), roughness, float(r[i, 0]), float(r[i, 1]), orc.ptr(out_d),
                              orc.ptr(out_g))
        dirs[i], g1[i] = out_d, out_g[0]
    return dirs, g1


def smith_g1(c, a):
    return 2 * c / (c + np.sqrt(a * a + (1 - a * a) * c * c))


def test_specular_vndf_sampling():
    """fs_specular.sc:139-170: the visible-normal sample (Heitz 2018) and its
    weight G1(l). (1) Near a mirror the reflected direction is the mirror
    direction. (2) E[G1(l) [n.l > 0]] over the sampler equals the integral of
    the VNDF D_v(h) = G1(v) max(v.h, 0) D(h) / (n.v) times G1(l(h)) over the
    hemisphere of h, computed by quadrature in double."""
    n = np.array([0.0, 0.0, 1.0])
    for theta in (0.3, 1.1):
        v = np.array([np.sin(theta), 0.0, np.cos(theta)])
        d, _ = specular_samples(n, v, 0.001, 16, 1)
        mirror = np.array([-v[0], -v[1], v[2]])
        assert np.all(np.abs(d - mirror) < 1e-4)
    # Tilted normal: the tangent frame of a non-axis normal.
    nt = np.array([0.3, 0.8, 0.52])
    nt /= np.linalg.norm(nt)
    vt = np.array([0.0, 1.0, 0.0])
    d, _ = specular_samples(nt, vt, 0.001, 8, 2)
    mirror = 2 * np.dot(nt, vt) * nt - vt
    assert np.all(np.abs(d - mirror) < 1e-4)

    for roughness, theta in ((0.5, 0.6), (0.8, 1.2)):
        a = roughness * roughness
        v = np.array([np.sin(theta), 0.0, np.cos(theta)])
        _, g1 = specular_samples(n, v, roughness, 60000, 3)
        th, ph = np.meshgrid((np.arange(800) + 0.5) * (np.pi / 2) / 800, (np.arange(800) + 0.5) * 2 * np.pi / 800,
                             indexing="ij")
        h = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
        nh = h[..., 2]
        D = a * a / (np.pi * (nh * nh * (a * a - 1) + 1) ** 2)
        vh = h @ v
        l = 2 * vh[..., None] * h - v
        nl = l[..., 2]
        nv = v[2]
        integrand = smith_g1(nv, a) * np.maximum(vh, 0) * D / nv * np.where(nl > 0, smith_g1(np.maximum(nl, 1e-9), a), 0)
        want = (integrand * np.sin(th)).sum() * (np.pi / 2 / 800) * (2 * np.pi / 800)
        got = float(g1.mean())
        err = 4 * g1.std() / np.sqrt(len(g1)) + 2e-3
        assert abs(got - want) <= err, (roughness, theta, got, want)


# ---------------------------------------------------------- shadow atlas

def shadow_face(d):
    """shadowFace (shadow.sh:31-45), vectorised: ties go to x, then y."""
    a = np.abs(d)
    x = (a[:, 0] >= a[:, 1]) & (a[:, 0] >= a[:, 2])
    y = ~x & (a[:, 1] >= a[:, 2])
    return np.where(x, np.where(d[:, 0] >= 0, 0, 1), np.where(y, np.where(d[:, 1] >= 0, 2, 3),
                                                             np.where(d[:, 2] >= 0, 4, 5)))


def shadow_atlas_uv(sh, p, face):
    """shadowAtlasUv (shadow.sh:52-69) in double, from the float matrices."""
    m = np.array(sh.view_proj, np.float64).reshape(6, 4, 4).transpose(0, 2, 1)[face]  # row-major per point
    clip = np.einsum("nij,nj->ni", m, np.concatenate([p, np.ones((len(p), 1))], 1))
    ndc = clip[:, :2] / clip[:, 3:4]
    obl = sh.info[3] > 0.5
    local = np.stack([ndc[:, 0] * 0.5 + 0.5, ndc[:, 1] * 0.5 + 0.5 if obl else 0.5 - ndc[:, 1] * 0.5], 1)
    row = (face >= 3).astype(np.float64)
    column = face - 3.0 * row
    if obl:
        row = 1.0 - row
    return (np.stack([column, row], 1) + local) * np.array([1.0 / 3.0, 0.5])


def texel_rays(sh, X, Y):
    """The ray of atlas texel (X, Y): the documented inverse of shadowAtlasUv
    at the texel centre, unprojected to the far plane, in double."""
    F = sh.face_size
    obl = sh.info[3] > 0.5
    column, trow = X // F, Y // F
    face = column + 3 * (1 - trow if obl else trow)
    lx = (X - column * F + 0.5) / F
    ly = (Y - trow * F + 0.5) / F
    ndc = np.stack([lx * 2 - 1, ly * 2 - 1 if obl else 1 - ly * 2, np.ones_like(lx), np.ones_like(lx)], 1)
    inv = np.array(sh.inv_view_proj, np.float64).reshape(6, 4, 4).transpose(0, 2, 1)[face]
    w = np.einsum("nij,nj->ni", inv, ndc)
    far_point = w[:, :3] / w[:, 3:4]
    return face, far_point - np.array(sh.light_position[:3], np.float64)


@pytest.mark.parametrize("obl,hd", [(False, False), (True, True)])
def test_shadow_atlas_inverts_lookup_and_matches_brute_force(obl, hd):
    """The shadow pass (renderer.cpp:527-596, fs_shadow.sc) against the
    lookup that reads it (shadowAtlasUv, shadow.sh:52-69): every texel's ray,
    looked up again, lands on that texel wherever shadowFace picks the texel's
    own face (the 4-texel borders belong to neighbouring faces). The stored
    value is the nearest triangle over all triangles along that ray, clipped
    to [near, far], over far (1 where nothing is hit)."""
    hs = common.cornell(extra_lights=0)
    light = np.array([0.1, 1.2, -0.2], np.float32)
    sh = ra.shadow_setup(light, hs.radius, 256, hd, obl)
    F = 256
    atlas = np.zeros(6 * F * F, np.float32)
    L.orc_shadow_atlas(ctypes.byref(sh), ctypes.byref(orc.host_scene(hs)), orc.ptr(atlas))
    Y, X = np.divmod(np.arange(6 * F * F), 3 * F)
    face, d = texel_rays(sh, X, Y)
    own = shadow_face(d) == face
    assert own.mean() > 0.9
    uv = shadow_atlas_uv(sh, light.astype(np.float64) + d[own], face[own])
    tx, ty = np.floor(uv[:, 0] * 3 * F).astype(int), np.floor(uv[:, 1] * 2 * F).astype(int)
    assert np.array_equal(tx, X[own]) and np.array_equal(ty, Y[own])

    far = float(sh.info[0])
    rng = np.random.default_rng(3)
    pick = rng.choice(len(X), 3000, replace=False)
    agree = 0
    for i in pick:
        dn = d[i] / np.linalg.norm(d[i])
        far_distance = np.linalg.norm(d[i])
        near_distance = far_distance * float(sh.near_plane) / far
        t = brute_closest(hs, light.astype(np.float64) + dn * near_distance, dn, far_distance - near_distance)
        want = 1.0 if not np.isfinite(t.min()) else (t.min() + near_distance) / far
        agree += abs(float(atlas[i]) - want) <= 1e-4 * max(want, 1e-3)
    assert agree >= len(pick) - 3, len(pick) - agree
    assert (atlas < 1.0).mean() > 0.9  # the light is inside the box: nearly every texel sees a wall


def test_frame_pack_carries_shadow_uniforms():
    """With shadows on, rgi_frame_pack binds the light's six face matrices and
    u_shadowInfo (renderer.cpp:527-537, 782-783): far 2.5 R, tan(half fov)
    F / (F - 8), 1 / F."""
    hs = common.cornell(extra_lights=0)
    s = ra.Settings(shadows=True, light_position=(0.1, 1.2, -0.2), shadow_face_size=2)
    eye, vp, inv = ra.orbit_camera(hs, 0, 64, 48)
    f = ra.make_frame(s, hs, 0, eye, vp, inv, eye, vp, 64, 48, False)
    sh = ra.shadow_setup((0.1, 1.2, -0.2), hs.radius, 1024)
    assert f.shadows_enabled == 1 and f.shadow_width == 3072 and f.shadow_height == 2048
    assert list(f.shadow_view_proj) == list(sh.view_proj) and list(f.shadow_info) == list(sh.info)
    assert sh.info[0] == np.float32(2.5 * np.float32(hs.radius))
    assert abs(sh.info[1] - 1024.0 / 1016.0) < 1e-6 and sh.info[2] == 1.0 / 1024.0
    off = ra.make_frame(ra.Settings(), hs, 0, eye, vp, inv, eye, vp, 64, 48, False)
    assert off.shadows_enabled == 0 and off.shadow_width == 0


# ---------------------------------------------------------------- lights

def test_light_list_cdf_and_power():
    """buildLightList (lights.cpp:86-165): pmf proportional to power
    (pi * luminance * area for triangles, 4 pi * luminance for points), CDF
    monotone, last entry exactly 1, triangle rows pointing at emissive BVH
    records."""
    hs = common.cornell(extra_lights=4)
    assert hs.light_count == 6
    lights = hs.lights.reshape(-1, 4, 4).astype(np.float64)
    lum = lambda c: 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    power = []
    for l in lights:
        if l[0, 3] > 0.5:
            power.append(4 * math.pi * lum(l[3, :3]))
        else:
            area = 0.5 * np.linalg.norm(np.cross(l[1, :3] - l[0, :3], l[2, :3] - l[0, :3]))
            power.append(math.pi * lum(l[3, :3]) * area)
            rec = hs.bvh[hs.node_count * 8 + int(l[1, 3]) * 16:][:16]
            assert rec[7] == l[3, 0] and rec[11] == l[3, 1] and rec[15] == l[3, 2]
    power = np.array(power)
    pmf = lights[:, 3, 3]
    assert np.allclose(pmf, power / power.sum(), rtol=1e-5)
    cdf = lights[:, 2, 3]
    assert np.all(np.diff(cdf) >= 0) and cdf[-1] == 1.0
    assert np.allclose(cdf, np.cumsum(pmf), atol=1e-6)


# ------------------------------------------------------------------ BRDF

def brdf(n, v, l, albedo, rough, metal):
    d, s = np.zeros(3, np.float32), np.zeros(3, np.float32)
    args = [f3(*n), f3(*v), f3(*l), f3(*albedo)]  # kept alive across the call
    L.orc_brdf(orc.ptr(args[0]), orc.ptr(args[1]), orc.ptr(args[2]), orc.ptr(args[3]), rough, metal, orc.ptr(d),
               orc.ptr(s))
    return d, s


def test_brdf_closed_forms():
    """evaluateBrdf (brdf.sh:38-62) at normal incidence: F = f0, D = 1/(pi a^2),
    G = (1/(1-k+k))^2 = 1, specular = D F / 4, diffuse = (1-F)(1-m) albedo / pi."""
    for rough in (0.3, 0.6, 1.0):
        a = rough * rough
        d, s = brdf((0, 0, 1), (0, 0, 1), (0, 0, 1), (0.5, 0.25, 0.75), rough, 0.0)
        D = 1.0 / (math.pi * a * a)
        assert np.allclose(s, D * 0.04 / 4.0, rtol=1e-5)
        assert np.allclose(d, (1 - 0.04) * np.array([0.5, 0.25, 0.75]) / math.pi, rtol=1e-5)
    # Metals: f0 = albedo, no diffuse.
    d, s = brdf((0, 0, 1), (0, 0, 1), (0, 0, 1), (0.9, 0.6, 0.3), 0.5, 1.0)
    assert np.all(d == 0) and np.allclose(s, np.array([0.9, 0.6, 0.3]) / (math.pi * 0.0625) / 4, rtol=1e-5)
    # Light below the surface: nothing.
    d, s = brdf((0, 0, 1), (0, 0, 1), (0, 0.6, -0.8), (0.5, 0.5, 0.5), 0.5, 0.0)
    assert np.all(d == 0) and np.all(s == 0)


def test_ggx_normalised():
    """The GGX distribution integrates to 1 over projected half vectors:
    the specular with F0 = 1 and no masking is checked through D alone by
    integrating D(h) cos(h) over the hemisphere numerically (inverted from
    the specular term at v = n, where G = Smith(n.l) Smith(1))."""
    rough = 0.5
    a2 = (rough * rough) ** 2
    th = np.linspace(0, math.pi / 2, 20001)
    D = a2 / (math.pi * ((np.cos(th) ** 2) * (a2 - 1) + 1) ** 2)
    integral = np.trapezoid(D * np.cos(th) * np.sin(th) * 2 * math.pi, th)
    assert abs(integral - 1.0) < 1e-4
    # And the oracle's D matches this D where masking is 1 (l = v = n, h = n).
    d, s = brdf((0, 0, 1), (0, 0, 1), (0, 0, 1), (1, 1, 1), rough, 1.0)
    assert abs(s[0] - D[0] / 4.0) < 1e-5 * D[0]


# ------------------------------------------------------------ sampling

def test_cosine_hemisphere_distribution():
    """cosineHemisphere (restir.sh:201-212): unit directions in the normal's
    hemisphere with density cos/pi, i.e. cos^2 uniform on [0, 1]."""
    rng = np.random.default_rng(8)
    for n in (f3(0, 1, 0), f3(0.6, 0.0, 0.8), f3(0.0, -0.995, 0.0998749)):
        n = (n / np.linalg.norm(n)).astype(np.float32)
        out = np.zeros(3, np.float32)
        c2 = []
        for r0, r1 in rng.random((20000, 2)).astype(np.float32):
            L.orc_cosine_hemisphere(orc.ptr(n), float(r0), float(r1), orc.ptr(out))
            assert abs(np.linalg.norm(out) - 1) < 1e-5
            c = float(np.dot(out, n))
            assert c >= -1e-6
            c2.append(c * c)
        c2 = np.sort(np.array(c2))
        ks = np.max(np.abs(c2 - np.arange(1, len(c2) + 1) / len(c2)))
        assert ks < 0.015
        assert abs(np.mean(np.sqrt(c2)) - 2.0 / 3.0) < 0.01


# ------------------------------------------------------------ passes

def test_di_single_candidate_weight_is_inverse_pdf():
    """With one candidate, no history and no taps, RIS keeps W = 1 / pdf of
    the drawn sample (pmf of its light times 1/area), or 0 when shadowed or
    unlit (cs_di_temporal.sc:52-70, 143-151)."""
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(di_candidates=1)
    f = common.frame_for(hs, s, 64, 48, 0, False)
    g = common.oracle_gbuffer(hs, f)
    hist = common.empty_history(64, 48, s.grid_table_size)
    out = common.oracle_frame(hs, f, g, hist)
    lights = hs.lights.reshape(-1, 4, 4).astype(np.float64)
    checked = 0
    for i in range(64 * 48):
        light = int(out["di_sample"][i, 3])
        W = out["di_receiver"][i, 3]
        if light < 0 or W == 0:
            continue
        l = lights[light]
        pdf = l[3, 3] if l[0, 3] > 0.5 else l[3, 3] / (0.5 * np.linalg.norm(np.cross(l[1, :3] - l[0, :3], l[2, :3] - l[0, :3])))
        assert abs(W * pdf - 1.0) < 1e-5
        checked += 1
    assert checked > 100


def test_temporal_without_taps_is_plain_ris():
    """Pairwise MIS with no neighbours hands the canonical reservoir the whole
    weight (restir.sh:822-823): results are independent of the history."""
    hs = common.cornell(extra_lights=4)
    s = ra.Settings(temporal_taps=0)
    seq = common.oracle_sequence(hs, s, 48, 32, 2, yaw_step=0.1)
    f, g, hist, out = seq[1]
    empty = common.empty_history(48, 32, s.grid_table_size)
    f0 = ra.Frame.from_buffer_copy(f)
    f0.history_valid = 0
    f0.multiple_bounces = 0
    f0.di_history_valid = 0
    f1 = ra.Frame.from_buffer_copy(f0)
    o0 = common.oracle_frame(hs, f0, g, empty)
    o1 = common.oracle_frame(hs, f1, g, hist)
    for k in ("gi_sample", "gi_receiver", "gi_normal"):
        assert np.array_equal(o0[k], o1[k])


def test_oracle_is_thread_count_independent():
    hs = common.cornell(extra_lights=4)
    s = ra.Settings()
    outs = []
    for threads in (1, 4):
        L.orc_set_threads(threads)
        seq = common.oracle_sequence(hs, s, 48, 40, 2, yaw_step=0.1)
        outs.append(seq[1][3])
    L.orc_set_threads(0)
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


# ------------------------------------------------------------ C ABI

def declared_functions():
    import re
    names = []
    for h in sorted(os.listdir(os.path.join(orc.ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(orc.ROOT, "include", h)).read()
        names += re.findall(r"^\w[\w\s\*]*?\b(rgi_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 40 and "rgi_nrc_train" in names and "rgi_batch_render" in names
    lib = ctypes.CDLL(ra.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_abi_struct_sizes_match():
    lib = ra.lib()
    for i, cls in enumerate(ra.ABI_STRUCTS):
        assert lib.rgi_abi_sizeof(i) == ctypes.sizeof(cls)


def test_nrc_config_is_checked_without_a_gpu():
    from restir_amd import nrc as rnrc
    lib = ra.lib()
    out = ctypes.c_void_p()
    for bad in ({"log2_table": 9}, {"log2_table": 23}, {"base_resolution": 1}, {"max_resolution": 8},
                {"learning_rate": 0.0}, {"beta2": 1.0}, {"aabb_min": (0, 0, 0), "aabb_max": (1, 0, 1)}):
        cfg = rnrc.config(**bad)
        assert lib.rgi_nrc_create(ctypes.byref(out), ctypes.byref(cfg)) == -1, bad
        assert out.value is None
    assert lib.rgi_nrc_train(None, None, None, 0, None, None) == -1
    assert lib.rgi_nrc_infer(None, None, 0, 0, None, None) == -1


def test_invalid_arguments_are_rejected_without_a_gpu():
    lib = ra.lib()
    assert lib.rgi_grid_clear(None, 1024, None) == -1
    assert b"NULL" in lib.rgi_last_error()
    f = ra.Frame()
    f.grid_table_size = 1000
    g = ra.GBuffer()
    assert lib.rgi_grid_insert(None, ctypes.byref(f), ctypes.byref(g), None, 0, None) == -1
    assert b"power of two" in lib.rgi_last_error()
    f.grid_table_size = 1024
    f.ray_mode = 7
    assert lib.rgi_grid_insert(None, ctypes.byref(f), ctypes.byref(g), None, 0, None) == -1


def test_cornell_scene_framing():
    hs = common.cornell()
    assert hs.triangle_count == 36
    assert abs(hs.radius - math.sqrt(3)) < 1e-5
    assert abs(hs.scene_scale - math.sqrt(3) / 1.75) < 1e-6
    em = hs.emission.sum(1) > 0
    assert em.sum() == 2 and np.all(hs.emission[em] == 17.0)


# ------------------------------------------------------ reference / error

def test_reference_error_matches_numpy():
    """cs_error.sc:59-63: relative MSE mean((L - L_ref)^2 / (L_ref^2 + 0.01))
    in luminance over geometry pixels; the FP64 tree reduction agrees with a
    plain numpy sum to rounding."""
    hs = common.cornell(extra_lights=4)
    s = ra.Settings()
    seq = common.oracle_sequence(hs, s, 64, 48, 3, yaw_step=0.0)
    n = 64 * 48
    accum = np.zeros(3 * n, np.float64)
    for f, g, hist, out in seq:
        gb = common.host_gbuffer(g, 64, 48)
        L.orc_reference_accumulate(ctypes.byref(gb), orc.ptr(out["color"]), orc.ptr(out["diffuse"]),
                                   orc.ptr(out["specular"]), orc.ptr(accum))
    f, g, hist, out = seq[-1]
    gb = common.host_gbuffer(g, 64, 48)
    res = np.zeros(5, np.float64)
    L.orc_reference_error(ctypes.byref(gb), orc.ptr(out["color"]), orc.ptr(out["diffuse"]), orc.ptr(out["specular"]),
                          orc.ptr(accum), 3, orc.ptr(res))
    geo = g["depth"] < np.float32(0.99999)
    col = out["color"].view(np.float16).astype(np.float32)[:, :3]
    alb = np.stack([(g["albedo"] >> (8 * c)) & 255 for c in range(3)], 1).astype(np.float32) / np.float32(255)
    metal = ((g["normal"] >> 24) & 255).astype(np.float32) / np.float32(255)
    img = col * (alb * (np.float32(1) - metal)[:, None])
    img = img + out["diffuse"].view(np.float16).astype(np.float32)[:, :3]
    img = img + out["specular"].view(np.float16).astype(np.float32)[:, :3]
    ref = accum.reshape(-1, 3) / 3.0
    w = np.array([0.2126, 0.7152, 0.0722])
    li, lr = img.astype(np.float64) @ w, ref @ w
    rel = ((li - lr) ** 2 / (lr ** 2 + 0.01))[geo].mean()
    mse = ((img.astype(np.float64) - ref) ** 2)[geo].mean()
    assert res[4] == geo.sum()
    assert abs(res[0] - rel) <= 1e-12 * max(rel, 1e-30) + 1e-18
    assert abs(res[1] - mse) <= 1e-12 * max(mse, 1e-30) + 1e-18
    assert abs(res[2] - li[geo].mean()) < 1e-12 and abs(res[3] - lr[geo].mean()) < 1e-12

    # The composite variant: the HDR image itself, the same statistics.
    accum = np.zeros(3 * n, np.float64)
    for f, g, hist, out in seq:
        L.orc_reference_accumulate_composite(ctypes.byref(common.host_gbuffer(g, 64, 48)), orc.ptr(out["hdr"]),
                                             orc.ptr(accum))
    f, g, hist, out = seq[-1]
    L.orc_reference_error_composite(ctypes.byref(common.host_gbuffer(g, 64, 48)), orc.ptr(out["hdr"]),
                                    orc.ptr(accum), 3, orc.ptr(res))
    img = out["hdr"].view(np.float16).astype(np.float64)[:, :3]
    ref = accum.reshape(-1, 3) / 3.0
    li, lr = img @ w, ref @ w
    rel = ((li - lr) ** 2 / (lr ** 2 + 0.01))[geo].mean()
    assert res[4] == geo.sum() and abs(res[0] - rel) <= 1e-12 * max(rel, 1e-30) + 1e-18
    assert abs(res[1] - ((img - ref) ** 2)[geo].mean()) <= 1e-12


def test_frame_pack_follows_renderer_uniforms():
    """rgi_frame_pack packs u_gridInfo, u_sceneInfo and u_lightingInfo as
    renderer.cpp:497-531, 693-731 do."""
    hs = common.cornell(extra_lights=4)
    s = ra.Settings()
    f = common.frame_for(hs, s, 320, 240, 5, True, True)
    scale = np.float32(np.float32(hs.radius) / np.float32(1.75))
    assert f.scene_scale == scale
    assert f.grid_min_cell == np.float32(0.02) * scale
    assert abs(f.grid_pixel_size - 2 * math.tan(math.radians(24)) / 240) < 1e-8
    assert f.grid_cell_pixels == 6.0 and f.world_space_reuse == 1
    assert (f.boiling_filter, f.jacobian_clamp, f.outlier_clamp) == (4.0, 10.0, 8.0)
    assert (f.lighting_surface, f.lighting_emission) == (1.0, 0.0)
    assert f.di_light_count == hs.light_count and f.frame_index == 5
    g = common.frame_for(hs, ra.Settings(restir_di=False), 320, 240, 5, True)
    assert (g.lighting_surface, g.lighting_emission, g.di_light_count) == (1.0, 1.0, 0)
    r = common.frame_for(hs, s.reference_mode(), 320, 240, 5, True)
    assert (r.temporal_taps, r.spatial_taps, r.di_spatial_taps) == (0, 0, 0)
    assert (r.boiling_filter, r.jacobian_clamp, r.outlier_clamp) == (0.0, 0.0, 0.0)


def test_constant_division_matches_ieee():
    """The kernels decode RGBA8 with x / 255 as q = x RN(1/255) plus one FMA
    correction: the same float as IEEE division for every byte and over a
    sweep of all non-negative floats; likewise x / 3. For pi the same
    refinement fails on tiny inputs, which shows the check can fail."""
    L = orc.lib()
    L.orc_div_const_mismatches.restype = ctypes.c_uint64
    L.orc_div_const_mismatches.argtypes = [ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    for b in range(256):
        bits = int(np.float32(b).view(np.uint32))
        assert L.orc_div_const_mismatches(255.0, bits, bits + 1, 1) == 0
    for c in (255.0, 3.0):
        assert L.orc_div_const_mismatches(c, 0, 0x7F800000, 251) == 0, c
    assert L.orc_div_const_mismatches(float(np.float32(np.pi)), 0, 0x0B800000, 7) > 0


def closed_mesh_scene(subdivisions=6, seed=3):
    """A closed surface: the cube [-1, 1]^3, each face split into a grid of
    quads (two triangles each), rotated by a random rotation so that no
    coordinate is round. Returns the rgi_scene over host arrays, the arrays
    (kept alive) and the vertex and edge-midpoint targets."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    tris, targets = [], set()
    g = np.linspace(-1.0, 1.0, subdivisions + 1)
    for axis in range(3):
        for side in (-1.0, 1.0):
            for i in range(subdivisions):
                for j in range(subdivisions):
                    corners = []
                    for (a, b) in ((g[i], g[j]), (g[i + 1], g[j]), (g[i + 1], g[j + 1]), (g[i], g[j + 1])):
                        p = [0.0, 0.0, 0.0]
                        p[axis], p[(axis + 1) % 3], p[(axis + 2) % 3] = side, a, b
                        corners.append(tuple(p))
                    tris.append((corners[0], corners[1], corners[2]))
                    tris.append((corners[0], corners[2], corners[3]))
    pos = (np.array(tris, np.float64).reshape(-1, 3) @ q.T).astype(np.float32).reshape(-1, 9)
    for t in pos.reshape(-1, 3, 3):
        for k in range(3):
            a, b = t[k], t[(k + 1) % 3]
            targets.add(tuple(a))
            targets.add(tuple(((a.astype(np.float64) + b) * 0.5).astype(np.float32)))
    n = len(pos)
    L = ra.lib()
    albedo = np.full((n, 3), 0.5, np.float32)
    emission = np.zeros((n, 3), np.float32)
    material = np.zeros(n, np.uint32)
    materials = np.zeros((16, 4), np.float32)
    mesh = ra.Mesh(n, pos.ctypes.data, albedo.ctypes.data, emission.ctypes.data, material.ctypes.data, 0,
                   materials.ctypes.data)
    buf = np.zeros(L.rgi_bvh_capacity(n), np.float32)
    depth = ctypes.c_uint32()
    nodes = L.rgi_bvh_build(ctypes.byref(mesh), buf.ctypes.data, ctypes.byref(depth))
    assert nodes > 0
    lights = np.zeros(16, np.float32)
    scene = ra.Scene(buf.ctypes.data, nodes, 2 * nodes, n, 0, lights.ctypes.data)
    return scene, (buf, pos, albedo, emission, material, materials, lights), np.array(sorted(targets), np.float32)


def test_watertight_triangle_test_never_leaks():
    """Rays from inside a closed, rotated, subdivided cube aimed exactly at
    every vertex and every edge midpoint (where two or more triangles meet):
    the watertight test hits a triangle for every ray, in both traversal
    forms. The reference's Moeller-Trumbore test is counted beside it (it
    may let such rays through)."""
    scene, keep, targets = closed_mesh_scene()
    L = orc.lib()
    origins = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1], [-0.55, 0.45, 0.6]], np.float32)
    rays = []
    for o in origins:
        d = targets.astype(np.float64) - o
        d /= np.linalg.norm(d, axis=1)[:, None]
        r = np.zeros((len(d), 8), np.float32)
        r[:, 0:3] = o
        r[:, 3] = 10.0
        r[:, 4:7] = d
        rays.append(r)
    rays = np.ascontiguousarray(np.concatenate(rays))
    n = len(rays)
    misses = {}
    for name, fn in (("watertight", L.orc_trace_watertight), ("moeller_trumbore", L.orc_trace)):
        for any_hit in (0, 1):
            hit = np.zeros(n, np.int32)
            dist = np.zeros(n, np.float32)
            fn(ctypes.byref(scene), orc.ptr(rays), n, orc.ptr(hit), orc.ptr(dist), any_hit)
            misses[(name, any_hit)] = int((hit < 0).sum())
    print("rays through edges and vertices: %d; misses %s" % (n, misses))
    assert misses[("watertight", 0)] == 0 and misses[("watertight", 1)] == 0


def test_watertight_agrees_with_moeller_trumbore_inside_triangles():
    """Away from edges both tests hit the same triangle at nearly the same
    distance: rays at triangle centroids of the closed cube."""
    scene, keep, _ = closed_mesh_scene(subdivisions=4, seed=8)
    pos = keep[1].reshape(-1, 3, 3).astype(np.float64)
    cent = pos.mean(axis=1)
    o = np.array([0.1, 0.05, -0.2])
    d = cent - o
    dist_true = np.linalg.norm(d, axis=1)
    d /= dist_true[:, None]
    rays = np.zeros((len(d), 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 10.0
    rays[:, 4:7] = d
    L = orc.lib()
    out = {}
    for name, fn in (("wt", L.orc_trace_watertight), ("mt", L.orc_trace)):
        hit = np.zeros(len(d), np.int32)
        dist = np.zeros(len(d), np.float32)
        fn(ctypes.byref(scene), orc.ptr(rays), len(d), orc.ptr(hit), orc.ptr(dist), 0)
        out[name] = (hit, dist)
    assert np.array_equal(out["wt"][0], out["mt"][0]) and np.all(out["wt"][0] >= 0)
    assert np.allclose(out["wt"][1], dist_true, rtol=1e-5)
    assert np.allclose(out["mt"][1], dist_true, rtol=1e-5)
