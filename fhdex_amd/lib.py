"""ctypes binding of libfhdx.so (the C ABI of include/fhdx.h) and device-array helpers.

The library is the product: every call here goes to the HIP kernels.  There is no CPU
fallback -- `lib()` raises if the built library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

from .params import Chemistry, Fab, Params

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfhdx.so")

# every symbol include/fhdx.h declares (tests check the library exports all of them)
EXPORTS = [
    "fhdx_abi_version", "fhdx_status_string", "fhdx_last_error", "fhdx_params_default",
    "fhdx_params_finalize", "fhdx_params_check", "fhdx_ctx_create", "fhdx_ctx_destroy",
    "fhdx_ctx_check", "fhdx_cons_to_prim", "fhdx_transport", "fhdx_fill_boundary", "fhdx_set_bc",
    "fhdx_calculate_flux", "fhdx_fill_noise", "fhdx_rk3_step_ref", "fhdx_rk3_step",
    "fhdx_rk3_step_injected", "fhdx_ctx_timing_enable", "fhdx_timing_kinds",
    "fhdx_ctx_timing_read", "fhdx_engine_bytes_needed", "fhdx_engine_create",
    "fhdx_engine_destroy", "fhdx_engine_add", "fhdx_engine_remove", "fhdx_engine_step",
    "fhdx_engine_sync", "fhdx_engine_state", "fhdx_engine_read", "fhdx_engine_errors",
    "fhdx_engine_bytes_used", "fhdx_engine_add_uniform", "fhdx_server_create",
    "fhdx_server_submit", "fhdx_server_poll", "fhdx_server_wait", "fhdx_server_stats",
    "fhdx_server_set_policy",
    "fhdx_server_destroy", "fhdx_ctx_autotune", "fhdx_ctx_set_tune", "fhdx_ctx_get_tune",
    "fhdx_evaluate_stats", "fhdx_write_plotfile", "fhdx_write_checkpoint", "fhdx_read_checkpoint",
    "fhdx_ctx_set_chemistry", "fhdx_chemistry_source", "fhdx_fill_chem_noise",
    "fhdx_rk3_stage_ref", "fhdx_rk3_step_noalloc", "fhdx_cons_to_prim_stag",
    "fhdx_calculate_flux_stag", "fhdx_rk3_step_stag",
]

_lib = None


class Fields(C.Structure):
    """``fhdx_fields`` (include/fhdx.h): the MultiFabs of WritePlotFile / WriteCheckPoint."""

    _fields_ = [(n, Fab) for n in ("cu", "cuMeans", "cuVars", "prim", "primMeans", "primVars",
                                   "spatialCross", "miscStats", "eta", "kappa")]


IO_HOST_ARRAYS = 1


class Tune(C.Structure):
    """``fhdx_tune`` (include/fhdx.h): launch shape of the fast path."""

    _fields_ = [("tile_width", C.c_int32), ("region_tiles", C.c_int32), ("z_levels", C.c_int32),
                ("noise_inline", C.c_int32), ("face_blocks_per_cu", C.c_int32),
                ("records_in_update", C.c_int32),
                ("ms_per_step", C.c_double)]

    def as_dict(self):
        return {"tile_width": self.tile_width, "region_tiles": self.region_tiles,
                "z_levels": self.z_levels, "noise_inline": self.noise_inline,
                "face_blocks_per_cu": self.face_blocks_per_cu,
                "records_in_update": self.records_in_update, "ms_per_step": self.ms_per_step}


class FhdxError(RuntimeError):
    def __init__(self, status: int, where: str):
        L = lib()
        msg = L.fhdx_status_string(status).decode()
        detail = L.fhdx_last_error().decode()
        super().__init__(f"{where}: {msg} ({status}) {detail}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m fhdex_amd.build`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(Params)
        FA = Fab * 3
        PF = C.POINTER(Fab)
        vp = C.c_void_p
        L.fhdx_abi_version.restype = C.c_int
        L.fhdx_status_string.restype = C.c_char_p
        L.fhdx_status_string.argtypes = [C.c_int]
        L.fhdx_last_error.restype = C.c_char_p
        L.fhdx_params_default.argtypes = [P]
        L.fhdx_params_finalize.argtypes = [P]
        L.fhdx_params_check.argtypes = [P]
        L.fhdx_ctx_create.argtypes = [P, C.POINTER(vp)]
        L.fhdx_ctx_destroy.argtypes = [vp]
        L.fhdx_ctx_check.argtypes = [vp, vp]
        L.fhdx_cons_to_prim.argtypes = [vp, Fab, Fab, vp]
        L.fhdx_transport.argtypes = [vp, Fab, Fab, Fab, Fab, Fab, Fab, vp]
        L.fhdx_fill_boundary.argtypes = [vp, Fab, vp]
        L.fhdx_set_bc.argtypes = [vp, Fab, Fab, vp]
        L.fhdx_calculate_flux.argtypes = [vp, Fab, Fab, Fab, Fab, Fab, Fab, Fab, FA, FA, FA, FA, FA,
                                          Fab, Fab, C.c_double, vp]
        L.fhdx_fill_noise.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_uint32, FA, FA, Fab, Fab, vp]
        L.fhdx_rk3_step_ref.argtypes = [vp, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab, FA, FA, FA, FA,
                                        FA, Fab, Fab, PF, PF, PF, PF, C.c_double, C.c_uint64,
                                        C.c_uint64, vp]
        L.fhdx_rk3_stage_ref.argtypes = [vp, C.c_int, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab, FA,
                                         FA, FA, FA, FA, Fab, Fab, FA, FA, Fab, Fab, Fab, Fab, Fab,
                                         Fab, C.c_double, vp]
        L.fhdx_rk3_step.argtypes = [vp, Fab, Fab, C.c_double, C.c_uint64, C.c_uint64, vp]
        L.fhdx_rk3_step_noalloc.argtypes = [vp, Fab, Fab, C.c_double, C.c_uint64, C.c_uint64, vp]
        L.fhdx_rk3_step_injected.argtypes = [vp, Fab, Fab, FA, FA, Fab, Fab, C.c_double, vp]
        E6 = Fab * 6
        L.fhdx_cons_to_prim_stag.argtypes = [vp, Fab, FA, Fab, FA, vp]
        L.fhdx_calculate_flux_stag.argtypes = [vp, Fab, FA, Fab, FA, Fab, Fab, Fab, Fab, Fab, FA, E6, FA,
                                               C.c_double, C.c_double, C.c_uint64, C.c_uint64, vp]
        L.fhdx_rk3_step_stag.argtypes = [vp, Fab, FA, Fab, FA, Fab, Fab, Fab, Fab, Fab, FA, E6, FA, Fab,
                                         Fab, C.c_double, C.c_uint64, C.c_uint64, vp]
        L.fhdx_ctx_timing_enable.argtypes = [vp, C.c_int]
        L.fhdx_ctx_timing_read.argtypes = [vp, C.c_int, C.c_char_p, C.c_int,
                                           C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.fhdx_engine_bytes_needed.argtypes = [P]
        L.fhdx_engine_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
        L.fhdx_engine_destroy.argtypes = [vp]
        L.fhdx_engine_add.argtypes = [vp, P, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, vp,
                                      C.POINTER(C.c_int)]
        L.fhdx_engine_add_uniform.argtypes = [vp, P, C.c_double, C.c_uint64, C.c_uint64,
                                              C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.fhdx_engine_remove.argtypes = [vp, C.c_int]
        L.fhdx_engine_step.argtypes = [vp, C.c_int, C.c_int]
        L.fhdx_engine_sync.argtypes = [vp]
        L.fhdx_engine_state.argtypes = [vp, C.c_int, PF, PF, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_int)]
        L.fhdx_engine_read.argtypes = [vp, C.c_int, vp, vp]
        L.fhdx_engine_errors.argtypes = [vp, C.POINTER(C.c_int)]
        L.fhdx_engine_bytes_used.argtypes = [vp]
        L.fhdx_ctx_autotune.argtypes = [vp, Fab, C.c_double, C.c_int, C.POINTER(Tune), C.POINTER(Tune),
                                        C.c_int, C.POINTER(C.c_int), vp]
        L.fhdx_ctx_set_tune.argtypes = [vp, C.POINTER(Tune)]
        L.fhdx_evaluate_stats.argtypes = [vp, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab,
                                          C.POINTER(C.c_double), C.c_int, C.c_int, vp]
        L.fhdx_ctx_get_tune.argtypes = [vp, C.POINTER(Tune)]
        L.fhdx_ctx_set_chemistry.argtypes = [vp, C.POINTER(Chemistry)]
        L.fhdx_chemistry_source.argtypes = [vp, Fab, Fab, Fab, C.c_double, vp]
        L.fhdx_fill_chem_noise.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_uint32, Fab, Fab, vp]
        PFd = C.POINTER(Fields)
        L.fhdx_write_plotfile.argtypes = [P, C.c_char_p, C.c_int, C.c_double, C.POINTER(C.c_double),
                                          PFd, C.c_int, C.c_int, C.c_int, vp]
        L.fhdx_write_checkpoint.argtypes = [P, C.c_char_p, C.c_int, C.c_double, C.c_int, C.c_uint64,
                                            PFd, C.c_int, vp]
        L.fhdx_read_checkpoint.argtypes = [P, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_double), C.POINTER(C.c_int),
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_int), PFd, C.c_int,
                                           vp]
        for name in EXPORTS:
            getattr(L, name).restype = C.c_int
        L.fhdx_engine_bytes_needed.restype = C.c_uint64
        L.fhdx_engine_bytes_used.restype = C.c_uint64
        L.fhdx_params_default.restype = None
        L.fhdx_params_finalize.restype = None
        L.fhdx_status_string.restype = C.c_char_p
        L.fhdx_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(status: int, where: str):
    if status != 0:
        raise FhdxError(status, where)


def fab_of_tensor(t, lo) -> Fab:
    """fhdx_fab of a contiguous float64 torch tensor shaped (ncomp, nz, ny, nx) starting at
    index `lo`."""
    assert t.dtype.is_floating_point and t.element_size() == 8 and t.is_contiguous() and t.dim() == 4
    f = Fab()
    f.p = t.data_ptr()
    nc, nz, ny, nx = t.shape
    for d, n in enumerate((nx, ny, nz)):
        f.lo[d] = lo[d]
        f.hi[d] = lo[d] + n - 1
    f.ncomp = nc
    return f


class Context:
    """Owns an fhdx_ctx (device workspace, error flag) for one parameter set."""

    def __init__(self, p: Params):
        self.p = p
        h = C.c_void_p()
        check(lib().fhdx_ctx_create(C.byref(p), C.byref(h)), "fhdx_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            lib().fhdx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, stream=None):
        check(lib().fhdx_ctx_check(self.h, stream), "fhdx_ctx_check")

    def autotune(self, cu_fab: Fab, dt: float, trial_steps: int = 2, stream=None):
        """Pick the fastest launch shape for this box (fhdx_ctx_autotune); cu is not modified.
        Returns (best, [timed candidates])."""
        t = Tune()
        trials = (Tune * 64)()
        n = C.c_int(0)
        check(lib().fhdx_ctx_autotune(self.h, cu_fab, float(dt), int(trial_steps), C.byref(t), trials, 64,
                                      C.byref(n), stream), "fhdx_ctx_autotune")
        return t, [trials[q] for q in range(n.value)]

    def set_tune(self, tile_width: int, region_tiles: int = 1, z_levels: int = 1, noise_inline: int = 1,
                 face_blocks_per_cu: int = 0, records_in_update: int = 0):
        t = Tune(tile_width, region_tiles, z_levels, noise_inline, face_blocks_per_cu,
                 records_in_update, 0.0)
        check(lib().fhdx_ctx_set_tune(self.h, C.byref(t)), "fhdx_ctx_set_tune")

    def set_chemistry(self, chem: Chemistry):
        """fhdx_ctx_set_chemistry: the reactions of a context with params.nreaction > 0."""
        check(lib().fhdx_ctx_set_chemistry(self.h, C.byref(chem)), "fhdx_ctx_set_chemistry")

    def get_tune(self) -> Tune:
        t = Tune()
        check(lib().fhdx_ctx_get_tune(self.h, C.byref(t)), "fhdx_ctx_get_tune")
        return t
