"""Device plumbing: torch tensors in, raw pointers and the current HIP stream out.

PyTorch-ROCm provides device memory (its caching allocator) and streams; all
arithmetic on the hot path runs in libeigenfns_hip.so.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("eigenfns_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_ptr() -> int:
    """hipStream_t of the current stream.  Called once per library call, so it
    reads the raw handle directly (≈1 µs) rather than building a Stream
    object (≈9 µs, measured as the largest Python cost of a serving step)."""
    if _RAW_STREAM is not None:
        return _RAW_STREAM(torch.cuda.current_device())
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def as_block(X, dtype=torch.complex64) -> torch.Tensor:
    """Contiguous complex64 device tensor for a block of vectors."""
    if isinstance(X, np.ndarray):
        X = torch.from_numpy(np.ascontiguousarray(X))
    if not isinstance(X, torch.Tensor):
        raise TypeError(f"expected a torch.Tensor or numpy array, got {type(X).__name__}")
    if X.dtype != dtype:
        X = X.to(dtype)
    if X.device.type != "cuda":
        X = X.to(device())
    return X.contiguous()


def rows(X: torch.Tensor) -> int:
    return int(X.shape[0])


def row_len(X: torch.Tensor) -> int:
    return int(X[0].numel()) if X.shape[0] else int(np.prod(X.shape[1:]))


class Scratch:
    """Cache of byte buffers for kernel work arrays (split partials), one set per
    (device, stream): work queued on different streams never shares a buffer,
    and a buffer is allocated while its stream is current, so the caching
    allocator orders its reuse on that stream."""

    def __init__(self):
        self._bufs: dict[tuple, torch.Tensor] = {}

    def get(self, nbytes: int, tag: str = "w") -> torch.Tensor:
        dev = device()
        key = (dev.index, torch.cuda.current_stream(dev).cuda_stream, tag)
        buf = self._bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)
            self._bufs[key] = buf
        return buf


SCRATCH = Scratch()


def lib():
    return _lib.load()
