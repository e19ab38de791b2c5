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


_SIDE_STREAMS = {}


def side_stream(dev=None, name="side"):
    """A persistent second stream per device and purpose (default priority):
    for work that overlaps the current stream, and for graph capture, which
    needs a stream other than the legacy default one."""
    import torch

    dev = torch.device(dev) if dev is not None else device()
    key = (str(dev), name)
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = torch.cuda.Stream(device=dev)
    return _SIDE_STREAMS[key]
