"""Raw tensor files: ``export_binary_tensor`` / ``import_binary_tensor``
(``cpp/corelib/lib/util/util_io.cppm:47-84``), the format of the reference's
``.htc`` cache blocks (``mod_tensor_caching.cppm:84-96``).

A file holds the tensor's element bytes in C order and nothing else; the
reader supplies the sizes and dtype.  Reading goes to host memory, as in the
reference; move the result to the GPU with ``.to(device)``.  A file shorter
than the requested tensor is an error here (the reference reads what is
there and leaves the rest undefined).
"""
from __future__ import annotations

import os
from typing import Sequence

import torch


def export_binary_tensor(tensor: torch.Tensor, path: str | os.PathLike) -> None:
    t = tensor.detach().contiguous().cpu()
    try:
        with open(path, "wb") as fh:
            fh.write(t.view(torch.uint8).numpy().tobytes() if t.numel() else b"")
    except OSError as e:
        raise RuntimeError("Could not open file for writing") from e


def import_binary_tensor(path: str | os.PathLike, sizes: Sequence[int], dtype: torch.dtype) -> torch.Tensor:
    t = torch.empty(tuple(int(s) for s in sizes), dtype=dtype)
    nbytes = t.numel() * t.element_size()
    try:
        with open(path, "rb") as fh:
            raw = fh.read(nbytes)
    except OSError as e:
        raise RuntimeError("Could not open file for reading") from e
    if len(raw) != nbytes:
        raise RuntimeError("Exception opening/reading file")
    if nbytes:
        t.view(-1).view(torch.uint8).copy_(torch.frombuffer(bytearray(raw), dtype=torch.uint8))
    return t
