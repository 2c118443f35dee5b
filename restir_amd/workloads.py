"""Seeded synthetic workloads of the benchmark configurations, shared by the
benchmark and the tests: the traversal microbenchmark's rays and the uneven
batch's 16 scenes. Host-side inputs only (numpy); nothing here computes a
result of the hot path."""
import numpy as np

import restir_amd as ra


def surface_rays(hs, count, seed):
    """Cosine rays from points uniform over the scene's area, both sides."""
    rng = np.random.default_rng(seed)
    p = hs.positions.reshape(-1, 3, 3).astype(np.float64)
    e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
    cr = np.cross(e1, e2)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    rays = np.zeros((count, 8), np.float32)
    chunk = 1 << 21
    for start in range(0, count, chunk):
        m = min(chunk, count - start)
        tri = rng.choice(len(area), size=m, p=area / area.sum())
        s = np.sqrt(rng.random(m))
        r1 = rng.random(m)
        pts = p[tri, 0] + e1[tri] * (1 - s)[:, None] + e2[tri] * (r1 * s)[:, None]
        n = cr[tri] / np.linalg.norm(cr[tri], axis=1)[:, None]
        n *= np.where(rng.random(m) < 0.5, -1.0, 1.0)[:, None]
        a = np.where(np.abs(n[:, 1:2]) < 0.99, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
        t = np.cross(n, a)
        t /= np.linalg.norm(t, axis=1)[:, None]
        b = np.cross(n, t)
        q0, q1 = rng.random(m), rng.random(m)
        rr, phi = np.sqrt(q0), 2 * np.pi * q1
        d = t * (rr * np.cos(phi))[:, None] + b * (rr * np.sin(phi))[:, None] + n * np.sqrt(1 - q0)[:, None]
        rays[start:start + m, 0:3] = pts + n * (1e-4 * hs.radius)
        rays[start:start + m, 3] = 4.0 * hs.scene_scale
        rays[start:start + m, 4:7] = d
    return rays


UNEVEN_SCENES = [
    # name, kind, size, seed, width, height, camera distance (scene radii;
    # box fields from 1.2 radii fill about half the view, atria from inside)
    ("cornell_s1", ra.RGI_SCENE_CORNELL, 0, 1, 640, 360, 3.17),
    ("cornell_s2", ra.RGI_SCENE_CORNELL, 0, 2, 960, 540, 3.17),
    ("boxfield_1k", ra.RGI_SCENE_BOXFIELD, 84, 1, 800, 450, 1.2),
    ("cornell_s3", ra.RGI_SCENE_CORNELL, 0, 3, 1280, 720, 3.17),
    ("boxfield_10k", ra.RGI_SCENE_BOXFIELD, 840, 1, 1280, 720, 1.2),
    ("atrium_128k", ra.RGI_SCENE_ATRIUM, 131072, 1, 640, 360, 0.7),
    ("boxfield_25k", ra.RGI_SCENE_BOXFIELD, 2100, 1, 1920, 1080, 1.2),
    ("atrium_64k", ra.RGI_SCENE_ATRIUM, 65536, 1, 1280, 720, 0.7),
    ("boxfield_50k", ra.RGI_SCENE_BOXFIELD, 4200, 1, 1920, 1080, 1.2),
    ("boxfield_100k", ra.RGI_SCENE_BOXFIELD, 8400, 1, 960, 540, 1.2),
    ("atrium_256k", ra.RGI_SCENE_ATRIUM, 262144, 1, 1920, 1080, 0.7),
    ("boxfield_500k", ra.RGI_SCENE_BOXFIELD, 42000, 1, 1600, 900, 1.2),
    ("atrium_512k", ra.RGI_SCENE_ATRIUM, 524288, 1, 1600, 900, 0.7),
    ("cornell_s4", ra.RGI_SCENE_CORNELL, 0, 4, 2560, 1440, 3.17),
    ("boxfield_1m", ra.RGI_SCENE_BOXFIELD, 83334, 1, 2560, 1440, 1.2),
    ("atrium_1m", ra.RGI_SCENE_ATRIUM, 1000000, 1, 2560, 1440, 0.7),
]
