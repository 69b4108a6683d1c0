"""Learned TL surrogate: a 1-D Fourier-operator network over depth trained on
the PE march's own TL fields, an ensemble router with PE fallback, and the
query domains it is trained on.

The surrogate runs on torch.  torch ships its own HIP runtime under the
soname of ROCm's (libamdhip64.so.7), and whichever copy is loaded first
serves the process: torch's starts here, before the march library binds to
it, and a process that has already run a march on ROCm's copy is told to
import this package first.
"""
import torch as _torch

from .. import _abi as _abi_mod


def _start_torch_runtime():
    if not _torch.cuda.is_available() or _torch.cuda.is_initialized():
        return
    try:
        _torch.cuda.init()
    except RuntimeError as exc:
        if _abi_mod._LIB is not None:
            raise RuntimeError("torch cannot start its HIP runtime after the march library has loaded ROCm's: "
                               "import cupyram_amd.surrogate (or torch) before the first march") from exc
        raise


_start_torch_runtime()
from .encode import Domain, Queries, encode, normalise_tl, denormalise_tl  # noqa: F401
from .fno import FNO, FNOConfig, dft_matrices  # noqa: F401
