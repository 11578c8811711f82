"""Argument normalisation shared by the operator classes."""
from __future__ import annotations

import torch

from . import hip


def as_input(a, dtype, shape=None):
    """A contiguous GPU tensor of ``dtype`` holding ``a`` (copies if needed)."""
    t = hip.asarray(a)
    if t.dtype != dtype:
        if t.is_complex() and not dtype.is_complex:
            raise TypeError("complex input where a real array is expected")
        t = t.to(dtype)
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return t.contiguous()


def output(out, dtype, shape, device=None):
    """(tensor to write, tensor to return) for an optional ``out``."""
    if out is None:
        t = torch.zeros(shape, dtype=dtype, device=device or hip.device())
        return t, t
    if tuple(out.shape) != tuple(shape):
        raise ValueError(f"out has shape {tuple(out.shape)}, expected {tuple(shape)}")
    if out.dtype == dtype and out.is_contiguous():
        return out, out
    return torch.empty(shape, dtype=dtype, device=out.device), out


def finish(tmp, out):
    if tmp is not out:
        out.copy_(tmp)
    return out
