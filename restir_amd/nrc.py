"""The neural radiance cache (include/restir_amd_nrc.h) over torch device
buffers: a thin owner of one rgi_nrc with train / infer / gradients calls.
torch only provides the device memory and the stream; every computation is a
librestir_amd.so kernel."""
import ctypes

import restir_amd as ra


def config(**kw):
    """rgi_nrc_config with the library defaults, then the given fields."""
    c = ra.NrcConfig()
    ra.lib().rgi_nrc_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if k in ("aabb_min", "aabb_max"):
            getattr(c, k)[:] = [float(x) for x in v]
        else:
            setattr(c, k, v)
    return c


def level_resolutions(cfg):
    L = ra.lib()
    out = []
    r = ra.U32()
    for level in range(ra.RGI_NRC_LEVELS):
        ra.check(L.rgi_nrc_level_resolution(ctypes.byref(cfg), level, ctypes.byref(r)))
        out.append(r.value)
    return out


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class NeuralCache:
    def __init__(self, cfg=None, **kw):
        import torch
        self.cfg = cfg if cfg is not None else config(**kw)
        self.L = ra.lib()
        self.handle = ctypes.c_void_p()
        ra.check(self.L.rgi_nrc_create(ctypes.byref(self.handle), ctypes.byref(self.cfg)))
        self.table_floats = ra.RGI_NRC_LEVELS * (1 << self.cfg.log2_table) * 2
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    def close(self):
        if self.handle:
            self.L.rgi_nrc_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, records, targets):
        """One Adam step; returns the batch loss as a device tensor."""
        n = records.shape[0]
        ra.check(self.L.rgi_nrc_train(self.handle, records.data_ptr(), targets.data_ptr(), n, self._loss.data_ptr(),
                                      _stream()))
        return self._loss

    def infer(self, records, precision=ra.RGI_NRC_BF16):
        import torch
        n = records.shape[0]
        out = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        ra.check(self.L.rgi_nrc_infer(self.handle, records.data_ptr(), n, precision, out.data_ptr(), _stream()))
        return out

    def gradients(self, records, targets, table=True):
        import torch
        n = records.shape[0]
        mlp = torch.empty(ra.RGI_NRC_MLP_PARAMS, dtype=torch.float32, device="cuda")
        tab = torch.empty(self.table_floats, dtype=torch.float32, device="cuda") if table else None
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        ra.check(self.L.rgi_nrc_gradients(self.handle, records.data_ptr(), targets.data_ptr(), n, mlp.data_ptr(),
                                          tab.data_ptr() if tab is not None else None, loss.data_ptr(), _stream()))
        return loss, mlp, tab

    def params(self):
        import torch
        mlp = torch.empty(ra.RGI_NRC_MLP_PARAMS, dtype=torch.float32, device="cuda")
        tab = torch.empty(self.table_floats, dtype=torch.float32, device="cuda")
        ra.check(self.L.rgi_nrc_get_params(self.handle, mlp.data_ptr(), tab.data_ptr(), _stream()))
        return mlp, tab

    def average(self):
        """The moving average of the parameters (what inference reads)."""
        import torch
        mlp = torch.empty(ra.RGI_NRC_MLP_PARAMS, dtype=torch.float32, device="cuda")
        tab = torch.empty(self.table_floats, dtype=torch.float32, device="cuda")
        ra.check(self.L.rgi_nrc_get_average(self.handle, mlp.data_ptr(), tab.data_ptr(), _stream()))
        return mlp, tab

    def set_params(self, mlp=None, table=None):
        ra.check(self.L.rgi_nrc_set_params(self.handle, mlp.data_ptr() if mlp is not None else None,
                                           table.data_ptr() if table is not None else None, _stream()))

    def steps(self):
        s = ra.U32()
        ra.check(self.L.rgi_nrc_steps(self.handle, ctypes.byref(s)))
        return s.value


def renderer_settings(hs, batch=16384, warmup=4, precision=ra.RGI_NRC_BF16, resolve=0, margin=0.02, records=None,
                      **kw):
    """rgi_nrc_settings for a HostScene: the cache's bounds are the scene's
    triangle bounds, widened by `margin` of their extent. Records default to
    ReSTIR GI path vertices, or to surface pixels with `resolve` (the
    resolved image queries the cache at every pixel)."""
    if records is None:
        records = ra.RGI_NRC_RECORDS_PIXELS if resolve else ra.RGI_NRC_RECORDS_VERTICES
    st = ra.NrcSettings()
    ra.lib().rgi_nrc_settings_default(ctypes.byref(st))
    p = hs.positions.reshape(-1, 3)
    lo, hi = p.min(axis=0), p.max(axis=0)
    pad = (hi - lo) * margin + 1e-3
    st.cache = config(aabb_min=lo - pad, aabb_max=hi + pad, **kw)
    st.batch, st.warmup, st.precision, st.resolve, st.records = batch, warmup, precision, resolve, records
    return st


def renderer_state(rend):
    """(cache handle, radiance plane pointer, loss pointer, trained frames) of a renderer."""
    c, r, l, t = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ra.U32()
    ra.check(ra.lib().rgi_renderer_nrc_state(rend, ctypes.byref(c), ctypes.byref(r), ctypes.byref(l), ctypes.byref(t)))
    return c.value, r.value, l.value, t.value
