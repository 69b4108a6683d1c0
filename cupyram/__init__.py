"""The reference's import names, served by the MI355X build.

``from cupyram import CuPyRAM, compute_pade_coefficients, AsyncBatcher``
(cupyram/__init__.py:1-3,15-19 of the reference) and the submodule paths
``cupyram.cupyram``, ``cupyram.pade``, ``cupyram.batching`` and
``cupyram.benchmark`` resolve to the modules of ``cupyram_amd``, so code
written against the reference imports this build unchanged.
"""
import sys as _sys

from cupyram_amd import AsyncBatcher, CuPyRAM, __version__, compute_pade_coefficients
from cupyram_amd import batching as _batching
from cupyram_amd import benchmark as _benchmark
from cupyram_amd import cupyram as _cupyram
from cupyram_amd import pade as _pade

for _name, _mod in (("cupyram", _cupyram), ("pade", _pade), ("batching", _batching), ("benchmark", _benchmark)):
    _sys.modules[f"{__name__}.{_name}"] = _mod
    globals()[_name] = _mod

__all__ = ["CuPyRAM", "compute_pade_coefficients", "AsyncBatcher"]
