"""Pins of the oracle's resampling core by what the mathematics fixes.

The oracle restates the reference's march, sample shading, reconnection
Jacobian, pairwise MIS and DI combination (shaders/restir.sh:327-931,
cs_restir_temporal.sc, cs_restir_spatial.sc, cs_di_temporal.sc,
cs_di_spatial.sc). The GPU matches the oracle bit for bit, so these checks
pin both:

- a white furnace of two parallel emitting plates, where the irradiance at
  every point has a closed form (the point-to-parallel-rectangle form
  factor): the resampled, demodulated GI with and without temporal and
  spatial reuse must average to it;
- the pairwise MIS weights of one sample held by the canonical reservoir and
  its neighbours sum to one (restir.sh:755-766);
- the reconnection Jacobian is the ratio of the solid angles a small patch at
  the sample point subtends at the two receivers;
- the depth-buffer march against an analytic wall and occluder: the hit
  lies within the bisection's bracket of the true intersection, a ray that
  never goes behind a surface misses, and the thickness rule separates a ray
  passing just behind an occluder (a hit) from one passing far behind (a
  miss);
- ReSTIR DI with a single point light: every technique draws the same
  sample, so with MIS weights summing to one the temporal and spatial
  passes' output equals the direct light's closed form at every pixel;
- each check fails on a plausible mistake planted in the oracle
  (orc_plant): a pdf left at the ray origin, a dropped skipped-tap share, a
  neighbour density taken at the wrong receiver, a dropped Jacobian, a
  thickness ten times too large, too few bisection rounds, and DI weights
  without their MIS factors.
"""
import ctypes
import math

import numpy as np
import pytest

import restir_amd as ra

import common
import orc

L = orc.lib()

PLANT_NO_JACOBIAN, PLANT_MIS_HERE, PLANT_PDF_AT_ORIGIN = 1, 2, 3
PLANT_THICKNESS, PLANT_BISECTION = 4, 5
PLANT_DI_TEMPORAL_MIS, PLANT_DI_SPATIAL_MIS, PLANT_SKIP_SHARE = 6, 7, 8


class planted:
    """The oracle with one planted mistake for the body of the block."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        L.orc_plant(self.which)

    def __exit__(self, *a):
        L.orc_plant(0)


def f3(*v):
    return np.array(v, np.float32)


# ------------------------------------------------------------ furnace

PLATE_HALF, PLATE_GAP = 4.0, 1.0


def plates(sub=8):
    """Two parallel square plates y = 0 and y = 1 of side 8, every triangle
    with albedo 0.5 and emission 1 on both sides."""
    g = np.linspace(-PLATE_HALF, PLATE_HALF, sub + 1)
    tris = []
    for y in (0.0, PLATE_GAP):
        for i in range(sub):
            for j in range(sub):
                c = [(g[i], y, g[j]), (g[i + 1], y, g[j]), (g[i + 1], y, g[j + 1]), (g[i], y, g[j + 1])]
                tris += [(c[0], c[1], c[2]), (c[0], c[2], c[3])]
    pos = np.array(tris, np.float32).reshape(-1, 9)
    n = len(pos)
    return ra.HostScene.from_mesh(pos, np.full((n, 3), 0.5, np.float32), np.full((n, 3), 1.0, np.float32))


def form_factor(x, z):
    """Form factor from a point of one plate to the whole other plate (unit
    normal towards it), by superposition of the corner rectangles: for the
    rectangle [0, a] x [0, b] at unit height, (a / sa atan(b / sa) + b / sb
    atan(a / sb)) / (2 pi), sa = sqrt(1 + a^2), sb = sqrt(1 + b^2)."""
    def corner(a, b):
        sa, sb = np.sqrt(1 + a * a), np.sqrt(1 + b * b)
        return (a / sa * np.arctan(b / sa) + b / sb * np.arctan(a / sb)) / (2 * np.pi)
    a1, a2 = (-PLATE_HALF - x) / PLATE_GAP, (PLATE_HALF - x) / PLATE_GAP
    b1, b2 = (-PLATE_HALF - z) / PLATE_GAP, (PLATE_HALF - z) / PLATE_GAP
    return corner(a2, b2) - corner(a1, b2) - corner(a2, b1) + corner(a1, b1)


def test_form_factor_against_monte_carlo():
    rng = np.random.default_rng(1)
    for x, z in ((0.0, 0.0), (3.5, -2.0), (3.9, 3.9)):
        u = rng.random((400000, 2))
        r, phi = np.sqrt(u[:, 0]), 2 * np.pi * u[:, 1]
        dx, dy, dz = r * np.cos(phi), np.sqrt(1 - u[:, 0]), r * np.sin(phi)
        t = PLATE_GAP / dy
        inside = (np.abs(x + dx * t) <= PLATE_HALF) & (np.abs(z + dz * t) <= PLATE_HALF)
        assert abs(inside.mean() - form_factor(x, z)) < 3e-3


def furnace_ratios(taps, ray_mode, frames=32, w=128, h=96):
    """Per frame (after 8), mean demodulated GI over the surface pixels
    divided by the mean of the closed-form irradiance L_e F(x) at their
    receivers. Multiple bounces, clamps and the boiling filter are off (the
    reference's unbiased configuration); emission is the only light."""
    hs = plates()
    s = ra.Settings(restir_di=False, emissive_lighting=True, multiple_bounces=False, clamp_jacobian=False,
                    clamp_outliers=False, boiling_filter=False, denoiser=False, composite=False, specular_gi=False,
                    temporal_taps=taps[0], spatial_taps=taps[1], grid_mode=ra.RGI_GRID_DETERMINISTIC,
                    ray_mode=ray_mode)
    ratios = []
    for i, (f, g, hist, out) in enumerate(common.oracle_frames(hs, s, w, h, frames, distance=0.05)):
        surf = g["depth"] < 0.99999
        if i < 8:
            continue
        got = out["color"].view(np.float16)[:, :3].astype(np.float64)[surf]
        p = out["gi_receiver"][surf, :3].astype(np.float64)
        want = form_factor(p[:, 0], p[:, 2])
        assert np.all(np.abs(p[:, 1] - np.where(p[:, 1] > 0.5, PLATE_GAP, 0.0)) < 1e-3)  # receivers on the plates
        ratios.append(got.mean() / want.mean())
    assert surf.mean() > 0.5
    return np.array(ratios)


FURNACE_CASES = [((0, 0), ra.RGI_RAY_BVH), ((2, 0), ra.RGI_RAY_BVH), ((0, 4), ra.RGI_RAY_BVH),
                 ((2, 4), ra.RGI_RAY_BVH), ((2, 4), ra.RGI_RAY_HYBRID)]


def furnace_holds(r):
    return abs(r.mean() - 1.0) < 4e-3 and np.all(np.abs(r - 1.0) < 1e-2)


@pytest.mark.parametrize("taps,ray_mode", FURNACE_CASES)
def test_plate_furnace_matches_closed_form(taps, ray_mode):
    r = furnace_ratios(taps, ray_mode)
    assert furnace_holds(r), (r.mean(), r.min(), r.max())


@pytest.mark.parametrize("plant,taps", [(PLANT_PDF_AT_ORIGIN, (0, 0)), (PLANT_PDF_AT_ORIGIN, (2, 4)),
                                        (PLANT_SKIP_SHARE, (2, 4)), (PLANT_SKIP_SHARE, (0, 4))])
def test_plate_furnace_catches_planted_mistakes(plant, taps):
    with planted(plant):
        r = furnace_ratios(taps, ra.RGI_RAY_BVH)
    assert not furnace_holds(r), (plant, r.mean())


# -------------------------------------------------------- pairwise MIS

def far_scene():
    """One small triangle far from everything: every visibility test passes."""
    pos = np.array([[100.0, 100.0, 100.0, 100.1, 100.0, 100.0, 100.0, 100.1, 100.0]], np.float32)
    return ra.HostScene.from_mesh(pos, np.full((1, 3), 0.5, np.float32), np.zeros((1, 3), np.float32))


def pairwise_sums(trials=400, seed=5):
    """Largest |sum of MIS weights / M_total - 1| over random configurations:
    a receiver, a sample point (position and normal) facing it, 1-6 taps of
    neighbour receivers nearby (random normals, some facing away from the
    sample, some taps finding no reservoir), random confidences."""
    hs = far_scene()
    scene = orc.host_scene(hs)
    f = common.frame_for(hs, ra.Settings(ray_mode=ra.RGI_RAY_BVH), 64, 48, 0, False)
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(trials):
        rp = rng.uniform(-1, 1, 3)
        rn = rng.normal(size=3)
        rn /= np.linalg.norm(rn)
        sp = rp + rn * rng.uniform(0.3, 2.0) + rng.normal(size=3) * 0.5
        sn = rp - sp + rng.normal(size=3) * 0.3
        sn /= np.linalg.norm(sn)
        taps = int(rng.integers(1, 7))
        nbs = np.zeros((taps, 7), np.float32)
        for i in range(taps):
            q = rp + rng.normal(size=3) * 0.3
            qn = rn + rng.normal(size=3) * (0.2 if rng.random() < 0.8 else 3.0)
            qn /= np.linalg.norm(qn)
            cnt = float(rng.integers(1, 21)) if rng.random() < 0.85 else -1.0
            nbs[i] = [*q, *qn, cnt]
        cc = float(rng.integers(1, 21))
        share, mis, total = np.zeros(1, np.float32), np.zeros(taps, np.float32), np.zeros(1, np.float32)
        receiver, sample = np.r_[rp, rn].astype(np.float32), np.r_[sp, sn].astype(np.float32)
        L.orc_gi_pairwise_weights(ctypes.byref(f), ctypes.byref(scene), orc.ptr(receiver), cc, orc.ptr(sample), taps,
                                  orc.ptr(nbs), orc.ptr(share), orc.ptr(mis), orc.ptr(total))
        m_total = cc + nbs[nbs[:, 6] > 0, 6].sum()
        assert abs(total[0] - m_total) <= 1e-5 * m_total
        worst = max(worst, abs((share[0] + mis.sum()) / m_total - 1.0))
    return worst


def test_pairwise_mis_weights_sum_to_one():
    assert pairwise_sums() < 1e-5


@pytest.mark.parametrize("plant", [PLANT_MIS_HERE, PLANT_SKIP_SHARE])
def test_pairwise_mis_catches_planted_mistakes(plant):
    with planted(plant):
        assert pairwise_sums(trials=100) > 1e-2


# ------------------------------------------------------------ Jacobian

def solid_angle(origin, tri):
    """Solid angle of a triangle seen from a point (Van Oosterom and
    Strackee 1983)."""
    a, b, c = (v - origin for v in tri)
    la, lb, lc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)
    num = abs(np.dot(a, np.cross(b, c)))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2.0 * math.atan2(num, den)


def jacobian_errors(trials=300, seed=9):
    hs = far_scene()
    f = common.frame_for(hs, ra.Settings(clamp_jacobian=False), 64, 48, 0, False)
    rng = np.random.default_rng(seed)
    worst, zeros = 0.0, 0
    for _ in range(trials):
        sp = rng.uniform(-1, 1, 3)
        sn = rng.normal(size=3)
        sn /= np.linalg.norm(sn)
        here = sp + (sn + rng.normal(size=3) * 0.6) * rng.uniform(0.3, 2.0)
        there = here + rng.normal(size=3) * 0.4
        args = [v.astype(np.float32) for v in (here, there, sp, sn)]  # kept alive through the call
        j = L.orc_reconnection_jacobian(ctypes.byref(f), *(orc.ptr(v) for v in args))
        ct = np.dot(sn, (there - sp) / np.linalg.norm(there - sp))
        ch = np.dot(sn, (here - sp) / np.linalg.norm(here - sp))
        if ct <= 0.002 or ch <= 0.002:
            if ct <= 0.0005:
                assert j == 0.0  # the neighbour saw the sample from behind or edge-on
                zeros += 1
            continue
        # A small triangle in the sample's tangent plane.
        t = np.cross(sn, [0.3, 0.8, 0.5])
        t /= np.linalg.norm(t)
        b = np.cross(sn, t)
        eps = 1e-4
        tri = [sp, sp + t * eps, sp + b * eps]
        want = solid_angle(here, tri) / solid_angle(there, tri)
        worst = max(worst, abs(j / want - 1.0))
    return worst, zeros


def test_reconnection_jacobian_is_the_solid_angle_ratio():
    worst, _ = jacobian_errors()
    assert worst < 2e-3, worst


def test_reconnection_jacobian_clamp():
    hs = far_scene()
    f = common.frame_for(hs, ra.Settings(clamp_jacobian=True), 64, 48, 0, False)
    sp, sn = f3(0, 0, 0), f3(0, 0, 1)
    here, there = f3(0, 0, 0.1), f3(0, 0, 3.0)  # ratio (3 / 0.1)^2 = 900, clamped to kJacobianClamp = 10
    j = L.orc_reconnection_jacobian(ctypes.byref(f), orc.ptr(here), orc.ptr(there), orc.ptr(sp), orc.ptr(sn))
    assert j == 10.0


def test_reconnection_jacobian_catches_a_dropped_jacobian():
    with planted(PLANT_NO_JACOBIAN):
        worst, _ = jacobian_errors(trials=100)
    assert worst > 0.1


# --------------------------------------------------------------- march

MARCH_W, MARCH_H = 160, 120


def march_setup():
    """A wall z = -2 (x, y in [-2, 2]) and an occluder z = -1 over x < 0,
    seen from the origin along -z: the depth buffer shows the occluder on the
    left half of the screen and the wall on the right."""
    quads = [[(-2, -2, -2), (2, -2, -2), (2, 2, -2), (-2, 2, -2)], [(-2, -2, -1), (0, -2, -1), (0, 2, -1), (-2, 2, -1)]]
    tris = []
    for q in quads:
        tris += [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
    pos = np.array(tris, np.float32).reshape(-1, 9)
    n = len(pos)
    hs = ra.HostScene.from_mesh(pos, np.full((n, 3), 0.5, np.float32), np.zeros((n, 3), np.float32))
    eye, target = np.zeros(3, np.float32), f3(0, 0, -1)
    _, _, vp, inv = ra.camera_matrices(eye, target, 48.0, MARCH_W / MARCH_H, 0.05, 10.0)
    f = ra.make_frame(ra.Settings(), hs, 0, eye, vp, inv, eye, vp, MARCH_W, MARCH_H, False)
    g = common.oracle_gbuffer(hs, f)
    return hs, f, g


def march(f, g, origin, direction, max_distance):
    d = np.asarray(direction, np.float64)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    hit = np.zeros(3, np.float32)
    h = L.orc_screen_march(ctypes.byref(f), ctypes.byref(common.host_gbuffer(g, MARCH_W, MARCH_H)),
                           orc.ptr(np.asarray(origin, np.float32)), orc.ptr(d), max_distance, orc.ptr(hit))
    return bool(h), hit, d


def step_distance(i, sc, steps=32):
    """The march's i-th step distance (restir.sh:385-388)."""
    sp = i * 32.0 / steps
    return (0.02 + sp * sp * 0.0035) * sc


def ray_parameter(o, d, point):
    """The t at which o + d t lies on the camera ray (from the origin)
    through `point`: where the march's sample was when it read that texel."""
    u = point / np.linalg.norm(point)
    sol, *_ = np.linalg.lstsq(np.array([d, -u]).T, -o, rcond=None)
    return sol[0]


def march_findings():
    """(worst |t_hit - t_expected| over rays to the wall, in units of their
    bisection bracket; whether each miss/thickness case came out right).

    A march sample counts as behind the wall once it is 0.01 sc behind it
    along the wall's normal (restir.sh:348-356), so the crossing the
    bisection brackets is at t_true + 0.01 sc / |d . n|; five rounds leave a
    bracket of (t_i - t_(i-1)) / 32 around it, t_i the first step past it."""
    hs, f, g = march_setup()
    sc = f.scene_scale
    maxd = 3.6 * sc
    n = np.frombuffer(g["normal"][MARCH_H // 2 * MARCH_W + MARCH_W * 3 // 4].tobytes(), np.uint8)[:3] / 255.0 * 2 - 1
    n /= np.linalg.norm(n)  # the wall's normal as the G-buffer stores it
    worst = 0.0
    rng = np.random.default_rng(3)
    for _ in range(40):
        z0 = rng.uniform(-1.4, -0.6)
        o = np.array([rng.uniform(0.05, 0.45) * -z0, rng.uniform(-0.3, 0.3) * -z0, z0])
        p = np.array([rng.uniform(0.15, 0.5) * 2.0, rng.uniform(-0.3, 0.3) * 2.0, -2.0])
        hit, at, d = march(f, g, o, p - o, maxd)
        assert hit, "a ray to the wall must hit it"
        d = d.astype(np.float64)
        t_expected = np.linalg.norm(p - o) + 0.01 * sc / abs(np.dot(d, n))
        i = next(k for k in range(1, 33) if step_distance(k, sc) >= t_expected)
        bracket = (step_distance(i, sc) - step_distance(i - 1, sc)) / 32.0
        assert abs(at[2] + 2.0) < 1e-4, "the hit point is on the wall"
        worst = max(worst, abs(ray_parameter(o, d, at.astype(np.float64)) - t_expected) / bracket)
    ok = {
        # Parallel to the wall in front of it, and back towards the camera: never behind a surface.
        "parallel": not march(f, g, (0.4, 0.0, -1.5), (0, 1, 0), maxd)[0],
        "away": not march(f, g, (0.4, 0.0, -1.5), (0, 0, 1), maxd)[0],
        # Passing 0.5 behind the occluder: beyond its thickness (0.03 sc + 0.03 t), a miss.
        "far_behind": not march(f, g, (0.5, 0.0, -1.5), (-1, 0, 0), maxd)[0],
    }
    # Passing 0.04 behind it: past the 0.01 sc margin and within the
    # thickness, a hit on the occluder's edge.
    hit, at, _ = march(f, g, (0.5, 0.0, -1.04), (-1, 0, 0), maxd)
    ok["just_behind"] = hit and abs(at[2] + 1.0) < 1e-3 and abs(at[0]) < 0.05
    return worst, ok


def test_march_against_analytic_wall_and_occluder():
    worst, ok = march_findings()
    assert worst < 1.05, worst
    assert all(ok.values()), ok


@pytest.mark.parametrize("plant", [PLANT_THICKNESS, PLANT_BISECTION])
def test_march_catches_planted_mistakes(plant):
    with planted(plant):
        worst, ok = march_findings()
    assert worst >= 1.05 or not all(ok.values()), (plant, worst, ok)


# ------------------------------------------------- DI, one point light

def lit_wall(material, sub=8, seed=11):
    """A wall z = 0 of side 8 (random albedo per triangle) and a point light
    in front of it (buildLightList's seeded placement in the middle 80% of
    the bounds, which a tiny triangle in a corner stretches to z = 3)."""
    g = np.linspace(-4.0, 4.0, sub + 1)
    tris = []
    for i in range(sub):
        for j in range(sub):
            c = [(g[i], g[j], 0.0), (g[i + 1], g[j], 0.0), (g[i + 1], g[j + 1], 0.0), (g[i], g[j + 1], 0.0)]
            tris += [(c[0], c[1], c[2]), (c[0], c[2], c[3])]
    tris.append(((4.0, 4.0, 3.0), (3.99, 4.0, 3.0), (4.0, 3.99, 3.0)))
    pos = np.array(tris, np.float32).reshape(-1, 9)
    n = len(pos)
    alb = np.random.default_rng(seed).uniform(0.2, 0.9, (n, 3)).astype(np.float32)
    return ra.HostScene.from_mesh(pos, alb, np.zeros_like(alb), material=material, extra_lights=1, light_seed=seed)


def di_point_light_errors(material, frames=4, w=96, h=72):
    """Largest relative error of the DI passes' diffuse and specular output
    against the single point light's closed form, over frames with temporal
    and spatial reuse.

    The camera faces the wall head-on, so the wall's depth is the same at
    every uv: DI spatial reconstructs a neighbour at its tap's uv, not at the
    texel centre whose depth it reads (cs_di_spatial.sc), and only there is
    that point on the wall for every tap. (At a surface seen at an angle, or
    at an edge between surfaces, the reconstructed point lies off the
    surface, its shadow ray can be blocked, and the reference's pairwise
    weights there no longer sum to one: a bias of the reference itself,
    which this check steers clear of.)"""
    hs = lit_wall(material)
    assert hs.light_count == 1 and hs.lights[0, 3] > 0.5
    light = hs.lights[0].reshape(4, 4)
    lp, intensity = light[0, :3].astype(np.float64), light[3, :3].astype(np.float64)
    assert lp[2] > 0.3
    s = ra.Settings()  # ReSTIR DI: 8 candidates, temporal reuse, 3 spatial taps
    eye = f3(0.3, -0.2, 6.0)
    _, _, vp, inv = ra.camera_matrices(eye, f3(0.3, -0.2, 0.0), 48.0, w / h, 0.05 * hs.radius, 20.0)
    hist = common.empty_history(w, h, s.grid_table_size)
    worst, lit = 0.0, 0
    for i in range(frames):
        f = ra.make_frame(s, hs, i, eye, vp, inv, eye, vp, w, h, i > 0, i > 0)
        g = common.oracle_gbuffer(hs, f)
        out = common.oracle_frame(hs, f, g, hist)
        hist = common.next_history(hist, out)
        surf = g["depth"] < 0.99999
        assert surf.all()
        sc = f.scene_scale
        idx = np.nonzero(surf)[0][::3]
        dd = out["diffuse"].view(np.float16)[:, :3].astype(np.float64)
        ds = out["specular"].view(np.float16)[:, :3].astype(np.float64)
        for k in idx:
            p = out["di_receiver"][k, :3].astype(np.float64)
            a = np.frombuffer(g["albedo"][k:k + 1].tobytes(), np.uint8) / 255.0
            nm = np.frombuffer(g["normal"][k:k + 1].tobytes(), np.uint8) / 255.0
            sh = np.frombuffer(g["shading"][k:k + 1].tobytes(), np.uint8) / 255.0
            n = sh[:3] * 2 - 1
            n /= np.linalg.norm(n)
            ns = nm[:3] * 2 - 1
            ns /= np.linalg.norm(ns)
            tl = lp - p
            d2 = max(np.dot(tl, tl), (1e-4 * sc) ** 2)
            ldir = tl / math.sqrt(d2)
            if np.dot(ns, ldir) <= 0.0:
                want_d = want_s = np.zeros(3)
            else:
                v = (eye - p) / np.linalg.norm(eye - p)
                bd, bs = np.zeros(3, np.float32), np.zeros(3, np.float32)
                args = [x.astype(np.float32) for x in (n, v, ldir, a[:3])]  # kept alive through the call
                L.orc_brdf(*(orc.ptr(x) for x in args), float(a[3]), float(nm[3]), orc.ptr(bd), orc.ptr(bs))
                want_d, want_s = bd * intensity / d2, bs * intensity / d2
            lit += int(want_d.max() > 0)
            for got, want in ((dd[k], want_d), (ds[k], want_s)):
                # Relative to the value, or to binary16's smallest normal for
                # the subnormals the RGBA16F planes store below it.
                err = np.abs(got - want) / np.maximum(np.abs(want), 6.1e-5)
                worst = max(worst, float(err.max()))
    assert lit > 1000
    return worst


@pytest.mark.parametrize("material", [(0.5, 0.0), (0.3, 0.7)])
def test_di_single_point_light_closed_form(material):
    assert di_point_light_errors(material) < 3e-3


@pytest.mark.parametrize("plant", [PLANT_DI_TEMPORAL_MIS, PLANT_DI_SPATIAL_MIS])
def test_di_single_point_light_catches_planted_mistakes(plant):
    with planted(plant):
        assert di_point_light_errors((0.5, 0.0), frames=3) > 0.05
