"""ctypes front-end of the CPU oracle (ram_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or the timed CPU baseline) -- never as part of the shipped march.

``resolve_ray`` restates the single-ray parameter defaults of
``CuPyRAM.get_params``/``setup`` (cupyram/cupyram.py:460-608): c0 is the mean
of the first sound-speed column, the wavelength is c0/freq, dr = np*1500/f,
dz = 0.1*1500/f, zmplt is the deepest bathymetry point, rmax is the largest
range in the inputs, rs = rmax + dr, lyrw = 20 and
zmax = max(z_sb + z_ss[-1]) + lyrw*lambda.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_NATIVE = {}


class _Ray(C.Structure):
    _fields_ = [
        ("freq", C.c_double), ("zs", C.c_double), ("zr", C.c_double),
        ("c0", C.c_double), ("lambda_", C.c_double), ("lyrw", C.c_double),
        ("dr", C.c_double), ("dz", C.c_double), ("zmax", C.c_double),
        ("rmax", C.c_double), ("zmplt", C.c_double), ("rs", C.c_double),
        ("np", C.c_int), ("ns", C.c_int), ("ndr", C.c_int), ("ndz", C.c_int),
        ("n_ss", C.c_int), ("n_rp_ss", C.c_int),
        ("z_ss", C.POINTER(C.c_double)), ("rp_ss", C.POINTER(C.c_double)),
        ("cw", C.POINTER(C.c_double)),
        ("n_sb", C.c_int), ("n_rp_sb", C.c_int),
        ("z_sb", C.POINTER(C.c_double)), ("rp_sb", C.POINTER(C.c_double)),
        ("cb", C.POINTER(C.c_double)), ("rhob", C.POINTER(C.c_double)),
        ("attn", C.POINTER(C.c_double)),
        ("n_bt", C.c_int), ("rbzb", C.POINTER(C.c_double)),
    ]


def _bind(L):
    dp = C.POINTER(C.c_double)
    L.orc_pade.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, dp, dp]
    L.orc_dims.argtypes = [C.POINTER(_Ray)] + [C.POINTER(C.c_int)] * 4
    L.orc_run.argtypes = [C.POINTER(_Ray), dp, dp, dp, dp, C.POINTER(C.c_int)]
    L.orc_run_trace.argtypes = [C.POINTER(_Ray), dp, C.POINTER(C.c_int), dp]
    L.orc_run_batch.argtypes = [C.POINTER(_Ray), C.c_int, C.POINTER(dp), C.c_int]
    L.orc_run_cells.argtypes = [C.POINTER(_Ray), dp, dp]
    L.orc_profiles.argtypes = [C.POINTER(_Ray), C.c_int, C.c_int, dp, dp, dp, dp, dp]
    L.orc_apply_factors.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, dp, dp,
                                    dp, dp, dp, dp, dp, dp]
    return L


def lib():
    """Load (building on first use if needed) liboracle: the checker build."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libram_oracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        _LIB = _bind(C.CDLL(path))
    return _LIB


def native_lib(directory):
    """The same source built with -O3 -march=native into `directory` on this
    machine (the timed CPU baseline); built on first use."""
    if directory not in _NATIVE:
        import subprocess
        subprocess.check_call(["make", "-s", "-C", _HERE, "native", f"NATIVE_DIR={directory}"])
        _NATIVE[directory] = _bind(C.CDLL(os.path.join(directory, "libram_oracle_native.so")))
    return _NATIVE[directory]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def pade(freq, c0, np_pade, ns=1, dr=50.0, ip=1):
    """FP64 Pade coefficients (pd1, pd2) -- restates cupyram/pade.py:180-273."""
    pd1 = np.zeros(np_pade, np.complex128)
    pd2 = np.zeros(np_pade, np.complex128)
    rc = lib().orc_pade(float(freq), float(c0), int(np_pade), int(ns), float(dr), int(ip), _dp(pd1), _dp(pd2))
    if rc == -1:
        raise ValueError("Laguerre method not converging")
    if rc:
        raise ValueError(f"orc_pade failed ({rc})")
    return pd1, pd2


@dataclass
class OracleRay:
    freq: float
    zs: float
    zr: float
    z_ss: np.ndarray
    rp_ss: np.ndarray
    cw: np.ndarray
    z_sb: np.ndarray
    rp_sb: np.ndarray
    cb: np.ndarray
    rhob: np.ndarray
    attn: np.ndarray
    rbzb: np.ndarray
    c0: float
    lam: float
    dr: float
    dz: float
    zmax: float
    rmax: float
    zmplt: float
    rs: float
    np_pade: int = 8
    ns: int = 1
    ndr: int = 1
    ndz: int = 1
    lyrw: float = 20.0
    _keep: list = field(default_factory=list, repr=False)

    def struct(self):
        arrs = {}
        for k in ("z_ss", "rp_ss", "z_sb", "rp_sb"):
            arrs[k] = np.ascontiguousarray(np.asarray(getattr(self, k), np.float64).ravel())
        n_ss, n_sb = arrs["z_ss"].size, arrs["z_sb"].size
        n_rss, n_rsb = arrs["rp_ss"].size, arrs["rp_sb"].size
        arrs["cw"] = np.ascontiguousarray(np.asarray(self.cw, np.float64).reshape(n_ss, n_rss))
        for k in ("cb", "rhob", "attn"):
            arrs[k] = np.ascontiguousarray(np.asarray(getattr(self, k), np.float64).reshape(n_sb, n_rsb))
        arrs["rbzb"] = np.ascontiguousarray(np.asarray(self.rbzb, np.float64).reshape(-1, 2))
        self._keep = list(arrs.values())
        return _Ray(self.freq, self.zs, self.zr, self.c0, self.lam, self.lyrw, self.dr, self.dz,
                    self.zmax, self.rmax, self.zmplt, self.rs, self.np_pade, self.ns, self.ndr, self.ndz,
                    n_ss, n_rss, _dp(arrs["z_ss"]), _dp(arrs["rp_ss"]), _dp(arrs["cw"]),
                    n_sb, n_rsb, _dp(arrs["z_sb"]), _dp(arrs["rp_sb"]), _dp(arrs["cb"]),
                    _dp(arrs["rhob"]), _dp(arrs["attn"]), arrs["rbzb"].shape[0], _dp(arrs["rbzb"]))

    def dims(self):
        s = self.struct()
        out = [C.c_int() for _ in range(4)]
        rc = lib().orc_dims(C.byref(s), *[C.byref(o) for o in out])
        if rc:
            raise ValueError("bad grid parameters")
        return tuple(o.value for o in out)  # nz, nvr, nvz, nsteps


def resolve_ray(freq, zs, zr, z_ss, rp_ss, cw, z_sb, rp_sb, cb, rhob, attn, rbzb, **kw) -> OracleRay:
    """Single-ray defaults of CuPyRAM.get_params / setup (cupyram.py:460-608)."""
    np_pade = kw.get("np", 8)
    cw = np.asarray(cw, np.float64)
    c0 = kw.get("c0", float(np.mean(cw[:, 0]) if cw.ndim > 1 else np.mean(cw)))
    lam = c0 / freq
    dr = kw.get("dr", np_pade * 1500.0 / freq)
    dz = kw.get("dz", 0.1 * 1500.0 / freq)
    rbzb = np.asarray(rbzb, np.float64).reshape(-1, 2)
    zmplt = kw.get("zmplt", float(np.max(rbzb[:, 1])))
    rmax = kw.get("rmax", float(max(np.max(rp_ss), np.max(rp_sb), np.max(rbzb[:, 0]))))
    lyrw = kw.get("lyrw", 20)
    z_ss = np.asarray(z_ss, np.float64)
    zmax = kw.get("zmax", float(np.max(np.asarray(z_sb, np.float64) + z_ss[-1]) + lyrw * lam))
    return OracleRay(freq=float(freq), zs=float(zs), zr=float(zr), z_ss=z_ss, rp_ss=np.asarray(rp_ss, np.float64),
                     cw=cw, z_sb=np.asarray(z_sb, np.float64), rp_sb=np.asarray(rp_sb, np.float64),
                     cb=np.asarray(cb, np.float64), rhob=np.asarray(rhob, np.float64),
                     attn=np.asarray(attn, np.float64), rbzb=rbzb, c0=float(c0), lam=float(lam),
                     dr=float(dr), dz=float(dz), zmax=float(zmax), rmax=float(rmax), zmplt=float(zmplt),
                     rs=float(kw.get("rs", rmax + dr)), np_pade=int(np_pade), ns=int(kw.get("ns", 1)),
                     ndr=int(kw.get("ndr", 1)), ndz=int(kw.get("ndz", 1)), lyrw=float(lyrw))


def run(ray: OracleRay, grids=True):
    """March one ray; returns dict with ranges, depths, tll, cpl, tlg, cpg, iz history."""
    nz, nvr, nvz, nsteps = ray.dims()
    s = ray.struct()
    tll = np.zeros(nvr)
    cpl = np.zeros(nvr, np.complex128)
    tlg = np.zeros((nvz, nvr)) if grids else None
    cpg = np.zeros((nvz, nvr), np.complex128) if grids else None
    izh = np.zeros(max(nsteps, 1), np.int32)
    rc = lib().orc_run(C.byref(s), _dp(tll), _dp(cpl), _dp(tlg) if grids else None,
                       _dp(cpg) if grids else None, izh.ctypes.data_as(C.POINTER(C.c_int)))
    if rc == -1:
        raise ValueError("Laguerre method not converging")
    if rc:
        raise ValueError(f"orc_run failed ({rc})")
    return {"Ranges": np.arange(1, nvr + 1) * ray.dr * ray.ndr,
            "Depths": np.arange(1, nvz + 1) * ray.dz * ray.ndz,
            "TL Line": tll, "CP Line": cpl, "TL Grid": tlg, "CP Grid": cpg,
            "iz": izh[:nsteps], "nz": nz}


def run_trace(ray: OracleRay):
    """TL line, the per-step history [nsteps, 4] = (iz, ss_ind, sb_ind, ns)
    after each range update, and the field u [nz+2] after the last step."""
    nz, nvr, _, nsteps = ray.dims()
    s = ray.struct()
    tll = np.zeros(nvr)
    trace = np.zeros((max(nsteps, 1), 4), np.int32)
    u = np.zeros(nz + 2, np.complex128)
    rc = lib().orc_run_trace(C.byref(s), _dp(tll), trace.ctypes.data_as(C.POINTER(C.c_int)), _dp(u))
    if rc == -1:
        raise ValueError("Laguerre method not converging")
    if rc:
        raise ValueError(f"orc_run_trace failed ({rc})")
    return {"TL Line": tll, "trace": trace[:nsteps], "u": u, "nz": nz}


def run_cells(ray: OracleRay):
    """TL line and cell-mean TL [nvz, nvr] (same layout as the TL Grid)."""
    nz, nvr, nvz, nsteps = ray.dims()
    s = ray.struct()
    tll = np.zeros(nvr)
    tlc = np.zeros((nvz, nvr))
    rc = lib().orc_run_cells(C.byref(s), _dp(tll), _dp(tlc))
    if rc == -1:
        raise ValueError("Laguerre method not converging")
    if rc:
        raise ValueError(f"orc_run_cells failed ({rc})")
    return {"TL Line": tll, "TL Cells": tlc}


def run_batch_tl(rays, nthreads=None, library=None):
    """TL lines for many rays on OpenMP threads (the CPU baseline); `library`
    is lib() (default) or native_lib(dir)."""
    structs = (_Ray * len(rays))(*[r.struct() for r in rays])
    outs = []
    for r in rays:
        _, nvr, _, _ = r.dims()
        outs.append(np.zeros(nvr))
    ptrs = (C.POINTER(C.c_double) * len(rays))(*[_dp(o) for o in outs])
    nthreads = nthreads or os.cpu_count() or 1
    bad = (library or lib()).orc_run_batch(structs, len(rays), ptrs, int(nthreads))
    if bad:
        raise ValueError(f"{bad} oracle rays failed")
    return outs


def profiles(ray: OracleRay, ss_ind=0, sb_ind=0):
    nz, _, _, _ = ray.dims()
    s = ray.struct()
    n = nz + 2
    alpw, alpb, rhob = np.zeros(n), np.zeros(n), np.zeros(n)
    ksqw, ksqb = np.zeros(n, np.complex128), np.zeros(n, np.complex128)
    lib().orc_profiles(C.byref(s), ss_ind, sb_ind, _dp(alpw), _dp(alpb), _dp(rhob), _dp(ksqw), _dp(ksqb))
    return alpw, alpb, rhob, ksqw, ksqb


def apply_factors(nz, iz, dz, k0, pd1, pd2, alpw, alpb, rhob, ksqw, ksqb, u):
    u = np.ascontiguousarray(u, np.complex128).copy()
    pd1 = np.ascontiguousarray(pd1, np.complex128)
    pd2 = np.ascontiguousarray(pd2, np.complex128)
    args = [np.ascontiguousarray(a, t) for a, t in ((alpw, np.float64), (alpb, np.float64),
                                                   (rhob, np.float64), (ksqw, np.complex128),
                                                   (ksqb, np.complex128))]
    lib().orc_apply_factors(nz, iz, dz, k0, len(pd1), _dp(pd1), _dp(pd2), *[_dp(a) for a in args], _dp(u))
    return u
