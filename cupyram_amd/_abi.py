"""ctypes binding of libcupyram_hip.so (declared in include/cupyram_hip.h).

The library is built in-tree by ``cupyram_amd.build``.  There is no fallback:
if the shared object is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcupyram_hip.so")

CPR_C64, CPR_C128 = 64, 128
VARIANTS = {"auto": 0, "thread": 1, "team": 2}
MAX_NP = 16


class CprRay(C.Structure):
    _fields_ = [
        ("dz", C.c_double), ("zstep", C.c_double), ("zmax", C.c_double), ("dr", C.c_double),
        ("k0", C.c_double), ("omega", C.c_double), ("c0", C.c_double), ("lyrw_lambda", C.c_double),
        ("rs", C.c_double), ("zs", C.c_double), ("zr", C.c_double),
        ("nz", C.c_int32), ("nsteps", C.c_int32), ("iz0", C.c_int32), ("ndr", C.c_int32),
        ("ndz", C.c_int32), ("nvr", C.c_int32), ("nvz", C.c_int32), ("np", C.c_int32),
    ]


class CprEnv(C.Structure):
    _fields_ = [
        ("n_ss", C.c_int32), ("n_rp_ss", C.c_int32), ("n_sb", C.c_int32), ("n_rp_sb", C.c_int32),
        ("n_bt", C.c_int32), ("pad_", C.c_int32),
        ("z_ss", C.c_int64), ("rp_ss", C.c_int64), ("cw", C.c_int64),
        ("z_sb", C.c_int64), ("rp_sb", C.c_int64), ("cb", C.c_int64), ("rhob", C.c_int64),
        ("attn", C.c_int64), ("rbzb", C.c_int64),
    ]


class CprConfig(C.Structure):
    _fields_ = [
        ("nrays", C.c_int32), ("precision", C.c_int32), ("variant", C.c_int32),
        ("compute_grids", C.c_int32), ("nvr_max", C.c_int32), ("nvz_max", C.c_int32),
        ("team_rows", C.c_int32), ("team_lanes", C.c_int32), ("compute_cells", C.c_int32),
    ]


RAY_DTYPE = np.dtype([
    ("dz", "f8"), ("zstep", "f8"), ("zmax", "f8"), ("dr", "f8"), ("k0", "f8"), ("omega", "f8"),
    ("c0", "f8"), ("lyrw_lambda", "f8"), ("rs", "f8"), ("zs", "f8"), ("zr", "f8"),
    ("nz", "i4"), ("nsteps", "i4"), ("iz0", "i4"), ("ndr", "i4"), ("ndz", "i4"), ("nvr", "i4"),
    ("nvz", "i4"), ("np", "i4"),
])
ENV_DTYPE = np.dtype([
    ("n_ss", "i4"), ("n_rp_ss", "i4"), ("n_sb", "i4"), ("n_rp_sb", "i4"), ("n_bt", "i4"), ("pad_", "i4"),
    ("z_ss", "i8"), ("rp_ss", "i8"), ("cw", "i8"), ("z_sb", "i8"), ("rp_sb", "i8"), ("cb", "i8"),
    ("rhob", "i8"), ("attn", "i8"), ("rbzb", "i8"),
])
assert RAY_DTYPE.itemsize == C.sizeof(CprRay)
assert ENV_DTYPE.itemsize == C.sizeof(CprEnv)

_LIB = None
_LOCK = threading.Lock()

_SYMBOLS = {
    "cpr_last_error": ([], C.c_char_p),
    "cpr_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "cpr_device_info": ([C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "cpr_pade": ([C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
    "cpr_pade_batch": ([C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                        C.c_int], C.c_int),
    "cpr_plan_create": ([C.POINTER(CprConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                         C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "cpr_plan_destroy": ([C.c_void_p], C.c_int),
    "cpr_selfs": ([C.c_void_p], C.c_int),
    "cpr_march": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "cpr_march_graph": ([C.c_void_p, C.c_int], C.c_int),
    "cpr_sync": ([C.c_void_p], C.c_int),
    "cpr_march_time": ([C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)], C.c_int),
    "cpr_read_outputs": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "cpr_read_field": ([C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "cpr_read_cells": ([C.c_void_p, C.c_void_p], C.c_int),
    "cpr_plan_device_bytes": ([C.c_void_p], C.c_int64),
    "cpr_team_shape": ([C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "cpr_plan_describe": ([C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
                          C.c_int),
    "cpr_march_noalloc": ([C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int,
                           C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    # surrogate kernels (device pointers as integers)
    "cpr_fno_gemm": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                      C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                      C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                      C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_spec_analysis": ([C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_spec_synth": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                            C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_gelu_bwd": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "cpr_fno_spec_expand": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_spec_expand_t": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_spec_fold": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_colsum": ([C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_outer_gelu_bwd": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_mse": ([C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float,
                     C.c_void_p, C.c_void_p], C.c_int),
    "cpr_fno_adam": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float,
                      C.c_float, C.c_float, C.c_float, C.c_int, C.c_float, C.c_void_p], C.c_int),
    "cpr_fno_fill": ([C.c_void_p, C.c_int64, C.c_float, C.c_void_p], C.c_int),
    "cpr_fno_cast_bf16": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "cpr_fno_cast_bf16_t": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_layer": ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_layer_bwd": ([C.c_void_p] * 9 + [C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_project": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int,
                         C.c_void_p], C.c_int),
    "cpr_fno_layer_project": ([C.c_void_p] * 8 + [C.c_float, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_lift": ([C.c_void_p] * 7 + [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p],
                     C.c_int),
    "cpr_fno_lift_modes": ([C.c_void_p] * 8 + [C.c_int, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "cpr_fno_layer_lift": ([C.c_void_p] * 7 + [C.c_int] + [C.c_void_p] * 6 + [C.c_int, C.c_void_p], C.c_int),
}


class CprError(RuntimeError):
    """Error reported by libcupyram_hip."""


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    with _LOCK:
        if _LIB is None:
            if not os.path.exists(LIB_PATH):
                raise CprError(f"{LIB_PATH} is missing: build it with `python -m cupyram_amd.build`")
            L = C.CDLL(LIB_PATH)
            for name, (args, res) in _SYMBOLS.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = res
            _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().cpr_last_error()
        raise CprError(msg.decode() if msg else f"libcupyram_hip error {rc}")


def device_count() -> int:
    n = C.c_int(0)
    lib().cpr_device_count(C.byref(n))
    return n.value


def device_info(device: int = 0):
    """(name, free bytes, total bytes) of a HIP device."""
    buf = C.create_string_buffer(256)
    free, total = C.c_int64(), C.c_int64()
    check(lib().cpr_device_info(int(device), buf, len(buf), C.byref(free), C.byref(total)))
    return buf.value.decode(errors="replace"), free.value, total.value


NOALLOC_SCRATCH_BYTES = 64 << 20


def march_noalloc(precision, rays, envs, env_data, pade, nvr_max, tll_dev, cpl_dev, stream=None):
    """cpr_march_noalloc: host inputs as for Plan, device outputs (integer
    addresses) tll [nvr_max][nrays] f64 and cpl [nvr_max][nrays] complex128."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    envs = np.ascontiguousarray(envs, ENV_DTYPE)
    env_data = np.ascontiguousarray(env_data, np.float64)
    pade = np.ascontiguousarray(pade, np.complex128)
    check(lib().cpr_march_noalloc(int(precision), len(rays), ptr(rays), ptr(envs), ptr(env_data), env_data.size,
                                  ptr(pade), int(nvr_max), C.c_void_p(tll_dev), C.c_void_p(cpl_dev),
                                  C.c_void_p(stream) if stream else None))


def team_shape(precision: int, nz: int):
    r, l = C.c_int(), C.c_int()
    check(lib().cpr_team_shape(int(precision), int(nz), C.byref(r), C.byref(l)))
    return r.value, l.value


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Plan:
    """One uploaded batch of rays on the device (owns the device buffers)."""

    def __init__(self, cfg: CprConfig, rays: np.ndarray, envs: np.ndarray, env_data: np.ndarray,
                 pade: np.ndarray, stream: int | None = None):
        self._h = C.c_void_p()
        self.cfg = cfg
        self.rays = np.ascontiguousarray(rays, RAY_DTYPE)
        envs = np.ascontiguousarray(envs, ENV_DTYPE)
        env_data = np.ascontiguousarray(env_data, np.float64)
        pade = np.ascontiguousarray(pade, np.complex128)
        check(lib().cpr_plan_create(C.byref(cfg), ptr(self.rays), ptr(envs), ptr(env_data), env_data.size,
                                    ptr(pade), C.c_void_p(stream) if stream else None, C.byref(self._h)))

    def selfs(self):
        check(lib().cpr_selfs(self._h))

    def march(self, s0, s1):
        check(lib().cpr_march(self._h, int(s0), int(s1)))

    def march_graph(self, steps_per_launch):
        check(lib().cpr_march_graph(self._h, int(steps_per_launch)))

    def sync(self):
        check(lib().cpr_sync(self._h))

    def march_time(self):
        ms, n = C.c_double(), C.c_int()
        check(lib().cpr_march_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def read_outputs(self, grids=False):
        n, nvr, nvz = self.cfg.nrays, self.cfg.nvr_max, self.cfg.nvz_max
        tll = np.zeros((n, nvr))
        cpl = np.zeros((n, nvr), np.complex128)
        tlg = np.zeros((n, nvz, nvr)) if grids else None
        cpg = np.zeros((n, nvz, nvr), np.complex128) if grids else None
        check(lib().cpr_read_outputs(self._h, ptr(tll), ptr(cpl), ptr(tlg), ptr(cpg)))
        return tll, cpl, tlg, cpg

    def read_cells(self):
        """Cell-mean TL [nrays, nvz_max, nvr_max] (plan made with compute_cells)."""
        tlc = np.zeros((self.cfg.nrays, self.cfg.nvz_max, self.cfg.nvr_max))
        check(lib().cpr_read_cells(self._h, ptr(tlc)))
        return tlc

    def read_field(self, nz_max):
        u = np.zeros((self.cfg.nrays, nz_max + 2), np.complex128)
        check(lib().cpr_read_field(self._h, ptr(u), int(nz_max)))
        return u

    def device_bytes(self):
        return int(lib().cpr_plan_device_bytes(self._h))

    def describe(self, b=0):
        v, r, l = C.c_int(), C.c_int(), C.c_int()
        check(lib().cpr_plan_describe(self._h, int(b), C.byref(v), C.byref(r), C.byref(l)))
        return v.value, r.value, l.value

    def close(self):
        if self._h:
            lib().cpr_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
