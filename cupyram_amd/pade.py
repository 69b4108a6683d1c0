"""FP64 Pade coefficients of the split-step march (host side).

``compute_pade_coefficients`` keeps the signature and meaning of
cupyram/pade.py:180 (PyRAM.epade): ``(pd1, pd2)`` complex128 arrays of length
``np_pade`` such that each march factor applies (1 + pd2 X)^-1 (1 + pd1 X).
The arithmetic runs in C++ inside libcupyram_hip.so (pade_host.cpp).
"""
from __future__ import annotations

import os

import numpy as np

from . import _abi


def compute_pade_coefficients(freq, c0, np_pade, ns=1, dr=50.0, ip=1):
    """Pade coefficients for frequency ``freq`` (Hz), reference speed ``c0``
    (m/s), ``np_pade`` terms, ``ns`` stability constraints, range step ``dr``
    (m); ``ip=1`` is the march operator, ``ip=2`` the self-starter operator.

    Raises ValueError when Laguerre's iteration does not converge (as
    cupyram/pade.py:146-147 does)."""
    pd1 = np.zeros(int(np_pade), np.complex128)
    pd2 = np.zeros(int(np_pade), np.complex128)
    rc = _abi.lib().cpr_pade(float(freq), float(c0), int(np_pade), int(ns), float(dr), int(ip),
                             _abi.ptr(pd1), _abi.ptr(pd2))
    if rc != 0:
        raise ValueError(_abi.lib().cpr_last_error().decode())
    return pd1, pd2


def compute_pade_batch(freq, c0, np_pade, ns, dr, ip, nthreads=None):
    """Vectorised form: arrays ``freq``, ``c0``, ``dr`` of length n ->
    array [n, 2, np_pade] complex128 (pd1, pd2)."""
    freq = np.ascontiguousarray(np.broadcast_to(np.asarray(freq, np.float64), np.shape(c0)))
    c0 = np.ascontiguousarray(c0, np.float64)
    dr = np.ascontiguousarray(np.broadcast_to(np.asarray(dr, np.float64), c0.shape))
    n = c0.size
    out = np.zeros((n, 2, int(np_pade)), np.complex128)
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    rc = _abi.lib().cpr_pade_batch(n, _abi.ptr(freq), _abi.ptr(c0), int(np_pade), int(ns), _abi.ptr(dr),
                                   int(ip), _abi.ptr(out), int(nthreads))
    if rc != 0:
        raise ValueError(_abi.lib().cpr_last_error().decode())
    return out
