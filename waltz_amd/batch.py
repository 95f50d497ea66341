"""Batched compress and decode of many fields / archives of mixed tile
shapes on the GPU, one launch per pipeline stage for the whole batch,
optionally replayed as a captured graph. The reference's loader entry point
for one tile shape, waltz_gpu_decode_batch
(python/pywaltz/gpu_dataloader.py:34-36), is exported by the same library.

    dec = BatchDecoder([(32, 32, 32), (48, 64, 96)], torch.float32)
    archives, spans = dec.compress(dec.flatten([f0, f1]), abs_error=1e-3)
    flat = dec.decode(archives, spans)                # tiles back to back
    tiles = dec.views(flat)
    packed, spans = pack([archive0, archive1])        # archives made one by one
"""
import ctypes as C

import torch

from ._lib import WaltzError, lib
from .codec import _stream

_W = {torch.float32: 4, torch.float64: 8}


def pack(archives, align=16):
    """Concatenate uint8 device archives at `align`-byte offsets; returns the
    packed device buffer and the device (offset, size) uint64 table."""
    offsets, at = [], 0
    for a in archives:
        offsets.append(at)
        at += (a.numel() + align - 1) // align * align
    device = archives[0].device if archives else torch.device("cuda")
    packed = torch.zeros(max(at, align), dtype=torch.uint8, device=device)
    for a, off in zip(archives, offsets):
        packed[off:off + a.numel()] = a
    spans = torch.tensor([v for a, off in zip(archives, offsets) for v in (off, a.numel())],
                         dtype=torch.int64).to(device)
    return packed, spans


class BatchDecoder:
    """Plan for a fixed list of tile shapes ((z, y, x) each) and dtype."""

    def __init__(self, shapes, dtype=torch.float32, device=None):
        if dtype not in _W:
            raise TypeError("float32 or float64 required")
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.dtype = dtype
        self.device = torch.device(device or "cuda")
        dims = [v for (z, y, x) in self.shapes for v in (x, y, z)]
        arr = (C.c_uint32 * len(dims))(*dims)
        self._h = lib().waltz_amd_batch_create(_W[dtype], len(self.shapes), arr)
        if not self._h:
            raise WaltzError(f"batch plan: {lib().waltz_amd_batch_error().decode(errors='replace')}")
        self.numel = int(lib().waltz_amd_batch_output_elements(self._h))
        ws = int(lib().waltz_amd_batch_workspace_bytes(self._h))
        self.workspace = torch.empty(max(ws, 1), dtype=torch.uint8, device=self.device)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().waltz_amd_batch_destroy(h)
            self._h = None

    def decode(self, packed, spans, out=None, check=True):
        """Decode every archive into `out` (flat, numel samples). With
        check=False the work is only enqueued (for graph capture); call
        check() afterwards."""
        if out is None:
            out = torch.empty(self.numel, dtype=self.dtype, device=self.device)
        if out.numel() != self.numel or out.dtype != self.dtype or not out.is_contiguous():
            raise ValueError("output must be a contiguous tensor of numel samples")
        rc = lib().waltz_amd_batch_decode(self._h, packed.data_ptr(), packed.numel(), spans.data_ptr(), out.data_ptr(),
                                          self.workspace.data_ptr(), self.workspace.numel(), _stream(self.device))
        if rc:
            raise WaltzError(f"batch decode: {lib().waltz_amd_batch_error().decode(errors='replace')}")
        if check:
            self.check()
        return out

    def check(self):
        if lib().waltz_amd_batch_check(self._h, _stream(self.device)):
            raise WaltzError(f"batch decode: {lib().waltz_amd_batch_error().decode(errors='replace')}")

    def flatten(self, fields):
        """The jobs' fields (one per shape, in order) as one flat tensor."""
        if len(fields) != len(self.shapes):
            raise ValueError("one field per job")
        for f, s in zip(fields, self.shapes):
            if tuple(f.shape) != s or f.dtype != self.dtype:
                raise ValueError("field shape or dtype differs from the plan")
        return torch.cat([f.reshape(-1) for f in fields])

    def compress(self, flat, abs_error, archives=None, spans=None, check=True):
        """Compress every job's field (flat: numel samples, jobs in order) into
        archive slots of one uint8 device buffer; returns (archives, spans)
        with spans the int64 (offset, size) table decode() takes. With
        check=False the work is only enqueued (graph capture); call
        compress_check() afterwards."""
        if flat.numel() != self.numel or flat.dtype != self.dtype or not flat.is_contiguous():
            raise ValueError("input must be a contiguous tensor of numel samples")
        cap = int(lib().waltz_amd_batch_archive_capacity(self._h))
        if archives is None:
            archives = torch.zeros(cap, dtype=torch.uint8, device=self.device)
        if spans is None:
            spans = torch.zeros(2 * len(self.shapes), dtype=torch.int64, device=self.device)
        ws = int(lib().waltz_amd_batch_compress_workspace_bytes(self._h))
        if getattr(self, "cworkspace", None) is None or self.cworkspace.numel() < ws:
            self.cworkspace = torch.empty(ws, dtype=torch.uint8, device=self.device)
        rc = lib().waltz_amd_batch_compress(self._h, flat.data_ptr(), float(abs_error), archives.data_ptr(),
                                            archives.numel(), spans.data_ptr(), self.cworkspace.data_ptr(),
                                            self.cworkspace.numel(), _stream(self.device))
        if rc:
            raise WaltzError(f"batch compress: {lib().waltz_amd_batch_error().decode(errors='replace')}")
        if check:
            self.compress_check()
        return archives, spans

    def compress_check(self):
        if lib().waltz_amd_batch_compress_check(self._h, _stream(self.device)):
            raise WaltzError(f"batch compress: {lib().waltz_amd_batch_error().decode(errors='replace')}")

    def views(self, flat):
        out, at = [], 0
        for s in self.shapes:
            n = s[0] * s[1] * s[2]
            out.append(flat[at:at + n].view(s))
            at += n
        return out
