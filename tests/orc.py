"""ctypes binding of the host oracle (oracle/liboracle.so) — the checker.

Test infrastructure only: the oracle restates the reference's shader logic on
the CPU; tests compare the HIP path against it on the same inputs.
"""
import ctypes
import os
import subprocess

import numpy as np

from restir_amd import CompositeInputs, DiReservoirs, Frame, GBuffer, GiInputs, GiReservoirs, Scene, Shadow

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    L = ctypes.CDLL(ORACLE_LIB)
    P, U32, F = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float
    PF = ctypes.POINTER(Frame)
    PS = ctypes.POINTER(Scene)
    PG = ctypes.POINTER(GBuffer)
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_set_threads.restype = None
    L.orc_get_threads.restype = ctypes.c_int
    L.orc_plant.argtypes = [ctypes.c_int]
    L.orc_plant.restype = None
    L.orc_screen_march.restype = ctypes.c_int
    L.orc_screen_march.argtypes = [PF, PG, P, P, F, P]
    L.orc_reconnection_jacobian.restype = F
    L.orc_reconnection_jacobian.argtypes = [PF, P, P, P, P]
    L.orc_gi_pairwise_weights.restype = None
    L.orc_gi_pairwise_weights.argtypes = [PF, PS, P, F, P, ctypes.c_int, P, P, P, P]
    L.orc_pcg.restype = U32
    L.orc_pcg.argtypes = [U32]
    L.orc_checksum_hash.restype = U32
    L.orc_checksum_hash.argtypes = [U32]
    L.orc_next_random.restype = F
    L.orc_next_random.argtypes = [ctypes.POINTER(U32)]
    L.orc_grid_cell_size.restype = F
    L.orc_grid_cell_size.argtypes = [PF, P, P]
    L.orc_grid_normal_bucket.restype = U32
    L.orc_grid_normal_bucket.argtypes = [P]
    L.orc_grid_key.restype = None
    L.orc_grid_key.argtypes = [PF, P, P, F, ctypes.POINTER(U32), ctypes.POINTER(U32)]
    L.orc_cosine_hemisphere.restype = None
    L.orc_cosine_hemisphere.argtypes = [P, F, F, P]
    L.orc_sincos.restype = None
    L.orc_sincos.argtypes = [F, ctypes.POINTER(F), ctypes.POINTER(F)]
    L.orc_brdf.restype = None
    L.orc_brdf.argtypes = [P, P, P, P, F, F, P, P]
    L.orc_half_bits.restype = ctypes.c_uint16
    L.orc_half_bits.argtypes = [F]
    L.orc_grid_keys.restype = None
    L.orc_grid_keys.argtypes = [PF, PG, P, P, P]
    L.orc_grid_clear.restype = None
    L.orc_grid_clear.argtypes = [P, U32]
    L.orc_grid_insert.restype = None
    L.orc_grid_insert.argtypes = [PF, PG, P]
    L.orc_grid_pick_entry.restype = U32
    L.orc_grid_pick_entry.argtypes = [PF, P, P, P, P, F, F, F]
    L.orc_trace.restype = None
    L.orc_trace.argtypes = [PS, P, U32, P, P, ctypes.c_int]
    L.orc_trace_watertight.restype = None
    L.orc_trace_watertight.argtypes = [PS, P, U32, P, P, ctypes.c_int]
    L.orc_gi_temporal.restype = None
    L.orc_gi_temporal.argtypes = [PF, PS, PG, ctypes.POINTER(GiInputs), ctypes.POINTER(GiReservoirs)]
    L.orc_gi_spatial.restype = None
    L.orc_gi_spatial.argtypes = [PF, PS, PG, ctypes.POINTER(GiInputs), ctypes.POINTER(GiReservoirs), P, P]
    L.orc_di_temporal.restype = None
    L.orc_di_temporal.argtypes = [PF, PS, PG, ctypes.POINTER(DiReservoirs), ctypes.POINTER(DiReservoirs)]
    L.orc_di_spatial.restype = None
    L.orc_di_spatial.argtypes = [PF, PS, PG, ctypes.POINTER(DiReservoirs), P, P]
    L.orc_reference_accumulate.restype = None
    L.orc_reference_accumulate.argtypes = [PG, P, P, P, P]
    L.orc_reference_error.restype = None
    L.orc_reference_error.argtypes = [PG, P, P, P, P, U32, P]
    L.orc_exp2f.restype = F
    L.orc_exp2f.argtypes = [F]
    L.orc_specular_albedo.restype = None
    L.orc_specular_albedo.argtypes = [P, F, F, P]
    L.orc_composite.restype = None
    L.orc_composite.argtypes = [PF, PG, ctypes.POINTER(CompositeInputs), P]
    L.orc_reference_accumulate_composite.restype = None
    L.orc_reference_accumulate_composite.argtypes = [PG, P, P]
    L.orc_reference_error_composite.restype = None
    L.orc_reference_error_composite.argtypes = [PG, P, P, U32, P]
    L.orc_specular_gi.restype = None
    L.orc_specular_gi.argtypes = [PF, PS, PG, ctypes.POINTER(GiInputs), P, P, P]
    L.orc_specular_sample.restype = None
    L.orc_specular_sample.argtypes = [P, P, F, F, F, P, P]
    L.orc_expf.restype = F
    L.orc_expf.argtypes = [F]
    L.orc_pow64.restype = F
    L.orc_pow64.argtypes = [F]
    L.orc_svgf_variance.restype = None
    L.orc_svgf_variance.argtypes = [PF, P, P, P, P, P]
    L.orc_svgf_atrous.restype = None
    L.orc_svgf_atrous.argtypes = [PF, P, P, P, ctypes.c_int, P]
    L.orc_shadow_atlas.restype = None
    L.orc_shadow_atlas.argtypes = [ctypes.POINTER(Shadow), PS, P]
    L.orc_gbuffer_primary.restype = None
    L.orc_gbuffer_primary.argtypes = [PF, PS, P, U32, P, P, P, P, P]
    L.orc_nrc_forward.restype = None
    L.orc_nrc_forward.argtypes = [P, U32, P, P, P, P, P, U32, ctypes.c_int, P]
    L.orc_nrc_gradients.restype = None
    L.orc_nrc_gradients.argtypes = [P, U32, P, P, P, P, P, P, U32, ctypes.c_int, P, P, P, P, ctypes.c_int]
    L.orc_nrc_adam.restype = None
    L.orc_nrc_adam.argtypes = [P, P, P, P, U32, F, F, F, F, F, F, ctypes.c_int]
    L.orc_nrc_corner.restype = None
    L.orc_nrc_corner.argtypes = [P, U32, ctypes.c_int, P, ctypes.c_int, ctypes.POINTER(U32), ctypes.POINTER(F)]
    L.orc_bf16_round.restype = F
    L.orc_bf16_round.argtypes = [F]
    _lib = L
    return L


class NrcModel:
    """Host copy of a cache's shape and parameters for the oracle's NRC
    functions (oracle/nrc.cpp)."""

    def __init__(self, res, log2_table, lo, hi, mlp, table):
        self.res = np.ascontiguousarray(res, np.uint32)
        self.log2_table = int(log2_table)
        self.lo = np.ascontiguousarray(lo, np.float32)
        self.hi = np.ascontiguousarray(hi, np.float32)
        self.mlp = np.ascontiguousarray(mlp, np.float32)
        self.table = np.ascontiguousarray(table, np.float32)

    def forward(self, records, emulate=1):
        records = np.ascontiguousarray(records, np.float32)
        n = records.shape[0]
        out = np.zeros((n, 3), np.float64)
        lib().orc_nrc_forward(ptr(self.res), self.log2_table, ptr(self.lo), ptr(self.hi), ptr(self.mlp),
                              ptr(self.table), ptr(records), n, emulate, ptr(out))
        return out

    def gradients(self, records, targets, emulate=1, denominators=None, frozen=False):
        """Loss and gradients; `denominators` (float64, one per record)
        receives the loss denominators, or supplies them when `frozen`."""
        records = np.ascontiguousarray(records, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        n = records.shape[0]
        loss = np.zeros(1, np.float64)
        mg = np.zeros(self.mlp.size, np.float64)
        tg = np.zeros(self.table.size, np.float64)
        lib().orc_nrc_gradients(ptr(self.res), self.log2_table, ptr(self.lo), ptr(self.hi), ptr(self.mlp),
                                ptr(self.table), ptr(records), ptr(targets), n, emulate, ptr(loss), ptr(mg), ptr(tg),
                                ptr(denominators), int(frozen))
        return float(loss[0]), mg, tg


def ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def host_scene(hs):
    """rgi_scene over the HostScene's host arrays."""
    return hs.rgi_scene(hs.bvh.ctypes.data, hs.lights.ctypes.data)
