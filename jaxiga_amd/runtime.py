"""Device selection and stream plumbing (PyTorch-ROCm owns memory and streams)."""

from __future__ import annotations

from jaxiga_amd import _lib


def device():
    """The current GPU; raises when none is visible (there is no CPU path)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("jaxiga_amd runs on a ROCm GPU (gfx950); no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def lib():
    return _lib.load()


def stream():
    return _lib.stream_handle()
