"""Pins of the neural radiance cache's oracle (oracle/nrc.cpp) on the CPU.

The cache has no counterpart in the reference, so the oracle is pinned by
mathematics: central finite differences of its exact-mode loss against its
analytic gradients (every layer and the table), the encoding's partition of
unity and dense indexing, the closed form of Adam's first step, and BF16
rounding against torch's conversion.
"""
import ctypes

import numpy as np
import pytest

import orc
import restir_amd as ra
from restir_amd import nrc as rnrc


def small_model(rng, log2_table=10, table_scale=0.5, base=4, top=64):
    cfg = rnrc.config(log2_table=log2_table, base_resolution=base, max_resolution=top)
    res = rnrc.level_resolutions(cfg)
    mlp = np.empty(ra.RGI_NRC_MLP_PARAMS, np.float32)
    mlp[:20480] = rng.uniform(-1, 1, 20480) * np.sqrt(6.0 / 128.0)
    mlp[20480:] = rng.uniform(-1, 1, 192) * np.sqrt(6.0 / 67.0)
    table = rng.uniform(-table_scale, table_scale, 16 * (1 << log2_table) * 2).astype(np.float32)
    return orc.NrcModel(res, log2_table, [-1, -1, -1], [1, 1, 1], mlp, table)


def random_records(rng, n):
    rec = np.zeros((n, 12), np.float32)
    rec[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    nrm = rng.normal(size=(n, 3))
    rec[:, 4:7] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    rec[:, 8:11] = rng.uniform(0, 1, (n, 3))
    return rec


def random_targets(rng, n, valid=0.9):
    t = np.zeros((n, 4), np.float32)
    t[:, 0:3] = rng.uniform(0, 2, (n, 3))
    t[:, 3] = (rng.uniform(size=n) < valid).astype(np.float32)
    return t


def test_level_resolutions():
    cfg = rnrc.config()
    res = rnrc.level_resolutions(cfg)
    assert res[0] == 16 and res[-1] == 512
    assert all(b >= a for a, b in zip(res, res[1:]))
    b = np.exp((np.log(512) - np.log(16)) / 15)
    assert res == [int(np.floor(16 * b ** l + 1e-9)) for l in range(16)]


def test_encoding_partition_of_unity_and_dense_index():
    rng = np.random.default_rng(3)
    m = small_model(rng)
    L = orc.lib()
    T = 1 << m.log2_table
    for level in range(16):
        r = int(m.res[level])
        dense = (r + 1) ** 3 <= T
        for _ in range(50):
            p = rng.uniform(0, 1, 3).astype(np.float32)
            ws = 0.0
            seen = set()
            for k in range(8):
                idx, w = ctypes.c_uint32(), ctypes.c_float()
                L.orc_nrc_corner(orc.ptr(m.res), m.log2_table, level, orc.ptr(p), k, ctypes.byref(idx), ctypes.byref(w))
                assert 0 <= idx.value < T and w.value >= 0.0
                ws += w.value
                seen.add(idx.value)
                if dense:  # the corner's own coordinates, x + y (r+1) + z (r+1)^2
                    c = np.minimum(np.floor(p * r).astype(np.int64), r - 1) + np.array([k & 1, (k >> 1) & 1, k >> 2])
                    assert idx.value == c[0] + c[1] * (r + 1) + c[2] * (r + 1) ** 2
            assert abs(ws - 1.0) < 1e-6
            if dense:
                assert len(seen) == 8


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38], np.float32)])
    mine = np.array([orc.lib().orc_bf16_round(float(v)) for v in x], np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


def _loss(m, rec, tgt, den):
    """The loss with the per-sample denominators held at their base values
    (the relative L2 loss stops the gradient through its denominator)."""
    loss, _, _ = m.gradients(rec, tgt, emulate=0, denominators=den, frozen=True)
    return loss


def _central_difference(m, params, p, rec, tgt, den, base):
    """(L(w + h) - L(w - h)) / 2h at parameter p, with h the FP32 steps
    actually taken; None when a ReLU kink lies inside the interval (the two
    one-sided differences disagree)."""
    w0 = params[p]
    h = np.float32(1e-5) * max(np.float32(1.0), abs(w0))
    params[p] = w0 + h
    hp = float(params[p]) - float(w0)
    lp = _loss(m, rec, tgt, den)
    params[p] = w0 - h
    hm = float(w0) - float(params[p])
    lm = _loss(m, rec, tgt, den)
    params[p] = w0
    fwd, bwd = (lp - base) / hp, (base - lm) / hm
    if abs(fwd - bwd) > 2e-4 * max(abs(fwd), abs(bwd)) + 1e-7:
        return None
    return (lp - lm) / (hp + hm)


@pytest.mark.parametrize("seed", [5, 6])
def test_gradients_match_finite_differences(seed):
    rng = np.random.default_rng(seed)
    m = small_model(rng)
    rec = random_records(rng, 24)
    tgt = random_targets(rng, 24)
    den = np.zeros(rec.shape[0], np.float64)
    loss, mg, tg = m.gradients(rec, tgt, emulate=0, denominators=den)
    assert loss > 0 and np.all(den >= 0.01)
    # MLP weights of every layer, including the output layer.
    picks = [int(i) for l in range(5) for i in rng.integers(l * 4096, l * 4096 + 4096, 6)]
    picks += [int(i) for i in rng.integers(20480, 20672, 6)]
    checked = 0
    for p in picks:
        if abs(mg[p]) < 1e-7:
            continue
        fd = _central_difference(m, m.mlp, p, rec, tgt, den, loss)
        if fd is None:
            continue
        assert abs(fd - mg[p]) <= 1e-3 * abs(mg[p]) + 1e-8, (p, fd, mg[p])
        checked += 1
    assert checked >= 20
    # Table entries that the batch touches.
    touched = np.flatnonzero(np.abs(tg) > 1e-9)
    assert touched.size > 100
    checked = 0
    for p in rng.choice(touched, 30, replace=False):
        fd = _central_difference(m, m.table, p, rec, tgt, den, loss)
        if fd is None:
            continue
        assert abs(fd - tg[p]) <= 1e-3 * abs(tg[p]) + 1e-9, (p, fd, tg[p])
        checked += 1
    assert checked >= 20


def test_emulated_gradients_close_to_exact():
    """Rounding to BF16 between layers moves each term by about 2^-8; the
    sums of back-propagated terms cancel, so the relative difference grows
    toward the input (measured: 0.3% at W6, 6% at W1, 8-18% per table level
    on this batch). The emulated mode is the one the GPU is checked against;
    this bounds how far it strays from the exact function."""
    rng = np.random.default_rng(7)
    m = small_model(rng)
    rec = random_records(rng, 64)
    tgt = random_targets(rng, 64)
    le, mge, tge = m.gradients(rec, tgt, emulate=1)
    lx, mgx, tgx = m.gradients(rec, tgt, emulate=0)
    assert abs(le - lx) < 0.02 * lx
    assert np.linalg.norm(mge - mgx) < 0.05 * np.linalg.norm(mgx)
    assert np.linalg.norm(mge[20480:] - mgx[20480:]) < 0.01 * np.linalg.norm(mgx[20480:])
    assert np.linalg.norm(tge - tgx) < 0.25 * np.linalg.norm(tgx)


def test_adam_first_step_closed_form():
    """With zero moments, one bias-corrected step moves a parameter by
    lr * g / (|g| + eps): lr against the sign of g; zero gradients are
    skipped when asked."""
    rng = np.random.default_rng(8)
    n = 1000
    p = rng.normal(size=n).astype(np.float32)
    g = (rng.normal(size=n) * 10.0 ** rng.integers(-4, 2, n)).astype(np.float32)
    g[::7] = 0.0
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    q = p.copy()
    lr, b1, b2, eps = 0.01, 0.9, 0.99, 1e-8
    orc.lib().orc_nrc_adam(orc.ptr(q), orc.ptr(m), orc.ptr(v), orc.ptr(g), n, lr, b1, b2, eps, 1 - b1, 1 - b2, 1)
    step = q.astype(np.float64) - p
    want = -lr * g / (np.abs(g.astype(np.float64)) + eps)
    assert np.allclose(step, want, rtol=1e-5, atol=2e-7)
    assert np.all(step[::7] == 0.0) and np.all(m[::7] == 0.0)
