"""Transpose driver and result decoding.

``transpose_bytestream_to_bitstream`` mirrors ``bitgen/bitstream.py:166-225``
(8 untimed + 8 timed calls recorded as the "transpose" timer row) but calls the
gfx950 kernel through the C ABI and produces uint64 LSB-first words
``[S, M, 8, W]`` with W = ceil((N + 1) / 64): one position past the input so a
match ending at the last byte is representable (the reference pads to 32
positions only and loses it when N % 32 == 0, bitstream.py:171-180).

``Bitstream`` decodes results like the reference's class (bitstream.py:27-104)
from either layout: uint64 LSB-first (hip) or uint32 MSB-first (reference).
"""
import numpy as np

from . import config as cfg
from .tool import global_timer


def words_for(n_bytes):
    return (n_bytes + 1 + 63) // 64


def transpose_bytes(input_stream_tensor, n_bytes=None, stream=None, out=None):
    """Device transpose of a uint8 tensor [..., N] -> uint64 [..., 8, W] (no timing).

    Returns (basis, staged input).  ``out`` reuses a basis tensor of the right
    shape; the input is staged (copied, zero-padded to a 16-byte row stride)
    only when it is not already in that form on the device."""
    import torch
    from .backend.hip import runtime as rt

    x = input_stream_tensor
    if x.dtype != torch.uint8:
        raise ValueError("input must be uint8")
    lead = tuple(x.shape[:-1])
    N = x.size(-1) if n_bytes is None else int(n_bytes)
    W = words_for(N)
    n_inputs = int(np.prod(lead)) if lead else 1
    stride = (N + 15) // 16 * 16
    dev = torch.device("cuda", torch.cuda.current_device())
    if x.device.type != "cuda" or x.size(-1) != stride or not x.is_contiguous() or x.data_ptr() % 16:
        padded = torch.zeros(lead + (stride,), dtype=torch.uint8, device=dev)
        padded[..., :N] = x[..., :N].to(dev)
        x = padded
    if out is not None:
        if tuple(out.shape) != lead + (8, W) or out.dtype != torch.uint64 or not out.is_contiguous():
            raise ValueError(f"out must be contiguous uint64 {lead + (8, W)}")
        basis = out
    else:
        basis = torch.empty(lead + (8, W), dtype=torch.uint64, device=dev)
    rt.init(dev.index or 0)
    rt.transpose(x.data_ptr(), N, stride, basis.data_ptr(), W, n_inputs, rt.current_stream_ptr(stream))
    return basis, x


def transpose_bytestream_to_bitstream(input_stream_tensor):
    """Reference-compatible driver: input uint8 [S, M, N] -> basis uint64 [S, M, 8, W]."""
    import torch

    if input_stream_tensor.dim() != 3:
        raise ValueError("input_stream_tensor must be [split, multi, n_bytes]")
    backend = cfg.get_config("backend", "hip")
    if backend != "hip":
        raise ValueError(f"bitgen_amd only provides the hip backend (got {backend!r})")
    N = input_stream_tensor.size(-1)
    basis, staged = transpose_bytes(input_stream_tensor, N)
    for _ in range(8):
        transpose_bytes(staged, N)
    for _ in range(8):
        with global_timer.time("transpose"):
            transpose_bytes(staged, N)
    size_mb = input_stream_tensor.numel() / 1024 / 1024
    global_timer.append_data("transpose", "input_size", size_mb)
    avg = global_timer.get_value("transpose", "avg_duration")
    global_timer.append_data("transpose", "throughput", size_mb / avg * 1024 if avg > 0 else float("nan"))
    torch.cuda.synchronize()
    return basis


class Bitstream:
    """Bit view of result words: ``data`` is a bool array [..., positions]."""

    def __init__(self, data, n_positions=None):
        data = np.asarray(data)
        shape = list(data.shape)
        if data.dtype == np.uint64:
            bits = np.unpackbits(np.ascontiguousarray(data).view(np.uint8), bitorder="little")
        elif data.dtype == np.uint32:
            # reference layout: position p = bit 31 - p%32 of word p/32
            sw = np.ascontiguousarray(data).byteswap()
            bits = np.unpackbits(sw.view(np.uint8), bitorder="big")
        elif data.dtype == np.uint8:
            bits = np.unpackbits(data, bitorder="big")
        elif data.dtype == bool:
            bits = data.reshape(-1)
        else:
            raise TypeError(f"Invalid data type: {data.dtype}")
        shape[-1] = -1
        self.data = bits.astype(bool).reshape(shape)
        if n_positions is not None:
            self.data = self.data[..., :n_positions]

    def get_count(self):
        return int(np.count_nonzero(self.data))

    def get_count_array(self):
        assert self.data.ndim == 4
        return np.count_nonzero(self.data, axis=-1).sum(axis=(0, 1))

    def get_ones_pos(self):
        return np.nonzero(self.data.reshape(-1))[0]

    def is_zero(self):
        return not np.any(self.data)

    def __eq__(self, other):
        return isinstance(other, Bitstream) and np.array_equal(self.data, other.data)

    def __repr__(self):
        return f"Bitstream(data=[{self.data.shape}])"
