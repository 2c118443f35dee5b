"""GPU: rgi_grid_insert_noalloc, the deterministic grid insert over the
library's fixed device scratch, against rgi_grid_insert in
RGI_GRID_DETERMINISTIC mode (context scratch) and the oracle, bit-exact, on
calls in a row that reuse the same scratch: a 1080p-sized GI frame, then
smaller ones (a smaller table, then a smaller frame)."""
import ctypes

import numpy as np
import pytest

import restir_amd as ra

import common
import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = common.Device()
    yield d
    d.close()


def insert(dev, frame, gb, noalloc):
    T = frame.grid_table_size
    grid = dev.empty(34 * T * 4)
    ra.check(dev.L.rgi_grid_clear(grid.data_ptr(), T, dev.stream))
    if noalloc:
        ra.check(dev.L.rgi_grid_insert_noalloc(ctypes.byref(frame), ctypes.byref(gb), grid.data_ptr(), dev.stream))
    else:
        ra.check(dev.L.rgi_grid_insert(dev.ctx, ctypes.byref(frame), ctypes.byref(gb), grid.data_ptr(),
                                       ra.RGI_GRID_DETERMINISTIC, dev.stream))
    return dev.get(grid, np.zeros(34 * T, np.uint32))


def test_noalloc_insert_matches_context_insert(dev):
    dev.L.rgi_grid_insert_noalloc.argtypes = [ctypes.POINTER(ra.Frame), ctypes.POINTER(ra.GBuffer), ctypes.c_void_p,
                                              ctypes.c_void_p]
    hs = common.cornell(extra_lights=0)
    cases = [(1920, 1080, 1, 1 << 20), (1920, 1080, 1, 1024), (320, 240, 2, 1 << 16)]
    for w, h, div, T in cases:
        print("case noalloc insert %dx%d divisor %d table %d" % (w, h, div, T), flush=True)
        s = ra.Settings(gi_divisor=div, grid_table_size=T)
        f = common.frame_for(hs, s, w, h, 3, True, yaw_step=2 * np.pi / 64)
        g = common.oracle_gbuffer(hs, f)
        gb, keep = dev.gbuffer(g, w, h)
        got = insert(dev, f, gb, True)
        ctx_grid = insert(dev, f, gb, False)
        want = np.zeros(34 * T, np.uint32)
        orc.lib().orc_grid_insert(ctypes.byref(f), ctypes.byref(common.host_gbuffer(g, w, h)), orc.ptr(want))
        for part, a, b, c in zip(("checksums", "counts", "entries"), common.grid_view(got, T),
                                 common.grid_view(ctx_grid, T), common.grid_view(want, T)):
            assert np.array_equal(a, b), "noalloc vs context insert: %s differ" % part
            assert np.array_equal(a, c), "noalloc vs oracle: %s differ" % part
        assert (got[:T] != 0).sum() > 100


def test_noalloc_insert_rejects_inputs_beyond_its_scratch(dev):
    hs = common.cornell(extra_lights=0)
    dev.L.rgi_grid_insert_noalloc.argtypes = [ctypes.POINTER(ra.Frame), ctypes.POINTER(ra.GBuffer), ctypes.c_void_p,
                                              ctypes.c_void_p]
    for w, h, T in [(2048, 1025, 1 << 16), (320, 240, 1 << 23)]:
        f = common.frame_for(hs, ra.Settings(grid_table_size=T), w, h, 0, False)
        gb = ra.GBuffer(w, h, 16, 16, 16, 16, 16)
        assert dev.L.rgi_grid_insert_noalloc(ctypes.byref(f), ctypes.byref(gb), 256, None) == -1
        assert b"exceed its fixed scratch" in ra.lib().rgi_last_error()
