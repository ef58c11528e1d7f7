"""xnyzip_amd -- XnYZip's compression path on AMD Instinct MI355X (gfx950).

Mirrors the reference's compressor classes (namespace XnYZip,
source/composer/block_compressor.hpp:30-39, block_compressor_rle.hpp:29-38):

    from xnyzip_amd import BlockCompressorRLE, QUANTIZER_TYPE
    frame = BlockCompressorRLE(QUANTIZER_TYPE.TRUNCATED_OCTAHEDRON, 0.01).compress(points, True)

`points` is an (N, 3) float32 array (numpy, or a torch tensor on the CPU or on
the GPU -- a GPU tensor is read in place).  `compress` returns the zstd frame
(bytes), like the reference's std::vector<uint8_t>.  All work runs through
libxnyz.so's HIP kernels; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from enum import IntEnum

import numpy as np

from . import _lib
from ._lib import Context, XnyzError  # noqa: F401

__all__ = ["QUANTIZER_TYPE", "BlockCompressor", "BlockCompressorRLE", "Context", "XnyzError",
           "compress", "quantize", "morton_to_hilbert"]


class QUANTIZER_TYPE(IntEnum):  # source/composer/utils.hpp:55-59 (ADAPTIVE not offered)
    CUBE = 0
    TRUNCATED_OCTAHEDRON = 1


_default_ctx: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = _default_ctx[device] = Context(device)
    return ctx


def _opts(quantizer, l2_bound, is_hilbert, rle, direct_threshold, frame_mode, zstd_workers):
    return _lib.Opts(int(quantizer), float(l2_bound), int(bool(is_hilbert)), int(bool(rle)),
                     int(direct_threshold), int(frame_mode), int(zstd_workers))


def compress(points, quantizer=QUANTIZER_TYPE.TRUNCATED_OCTAHEDRON, l2_bound=0.01,
             is_hilbert=False, rle=False, direct_threshold=1024,
             frame_mode=_lib.FRAME_ZSTD3, zstd_workers=0, want_prezstd=False,
             want_perm=False, ctx: Context | None = None):
    """Compress one cloud.  Returns (frame, prezstd or None, perm or None, stats).

    A torch tensor on a HIP device is compressed in place (no copy); any other
    input is converted to a contiguous float32 host array and uploaded."""
    L = _lib.lib()
    o = _opts(quantizer, l2_bound, is_hilbert, rle, direct_threshold, frame_mode, zstd_workers)
    fr, pre, st = _lib.Buf(), _lib.Buf(), _lib.Stats()
    dev_ptr, n, keep = None, 0, None
    try:
        import torch  # noqa: F401
        if hasattr(points, "is_cuda") and points.is_cuda:
            t = points.reshape(-1, 3)
            if t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError("device points must be a contiguous float32 (N, 3) tensor")
            dev_ptr, n, keep = t.data_ptr(), t.shape[0], t
            if ctx is None:
                ctx = default_context(t.device.index or 0)
    except ImportError:
        pass
    if ctx is None:
        ctx = default_context(0)
    perm = None
    if dev_ptr is not None:
        if want_perm and not rle:
            perm = np.empty(n, dtype=np.uint32)
        torch.cuda.synchronize(keep.device)  # the library works on its own stream
        rc = L.xnyz_compress_device(ctx.handle, C.c_void_p(dev_ptr), n, C.byref(o), C.byref(fr),
                                    C.byref(pre) if want_prezstd else None,
                                    C.c_void_p(perm.ctypes.data) if perm is not None else None,
                                    C.byref(st))
    else:
        a = points.detach().cpu().numpy() if hasattr(points, "detach") else points
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 3)
        n = a.shape[0]
        if want_perm and not rle:
            perm = np.empty(n, dtype=np.uint32)
        rc = L.xnyz_compress(ctx.handle, C.c_void_p(a.ctypes.data), n, C.byref(o), C.byref(fr),
                             C.byref(pre) if want_prezstd else None,
                             C.c_void_p(perm.ctypes.data) if perm is not None else None,
                             C.byref(st))
    _lib.check(rc)
    frame = _lib.take(fr)
    prez = _lib.take(pre) if want_prezstd else None
    return frame, prez, perm, st


class BlockCompressor:
    """BlockCompressor<float> (source/composer/block_compressor.hpp:30-222)."""

    rle = False

    def __init__(self, quantizer_type, l2_bound: float, is_debugging: bool = False,
                 direct_threshold: int = 1024, device: int = 0):
        self.quantizer_type = QUANTIZER_TYPE(quantizer_type)
        self.l2_bound = float(np.float32(l2_bound))
        self.is_debugging = is_debugging
        self.direct_threshold = int(direct_threshold)
        self.ctx = default_context(device)
        self.last_stats = None

    def compress(self, points, curve_type: bool) -> bytes:
        frame, _, _, st = compress(points, self.quantizer_type, self.l2_bound, curve_type,
                                   self.rle, self.direct_threshold, ctx=self.ctx)
        self.last_stats = st
        return frame


class BlockCompressorRLE(BlockCompressor):
    """BlockCompressorRLE<float> (source/composer/block_compressor_rle.hpp:29-321)."""

    rle = True


def quantize(points_device, quantizer, l2_bound, ctx: Context | None = None):
    """Shift + quantize a device tensor; returns (nodes int32 (N,3), offset float32 (3,))."""
    t = points_device.reshape(-1, 3)
    n = t.shape[0]
    nodes = np.empty((n, 3), dtype=np.int32)
    off = np.empty(3, dtype=np.float32)
    ctx = ctx or default_context(t.device.index or 0)
    import torch
    torch.cuda.synchronize(t.device)
    _lib.check(_lib.lib().xnyz_quantize_device(ctx.handle, C.c_void_p(t.data_ptr()), n,
                                                int(quantizer), float(l2_bound),
                                                C.c_void_p(nodes.ctypes.data),
                                                C.c_void_p(off.ctypes.data)))
    return nodes, off


def morton_to_hilbert(codes, iteration: int, ctx: Context | None = None) -> np.ndarray:
    a = np.ascontiguousarray(codes, dtype=np.uint64)
    out = np.empty_like(a)
    ctx = ctx or default_context(0)
    _lib.check(_lib.lib().xnyz_morton_to_hilbert(ctx.handle, C.c_void_p(a.ctypes.data), a.size,
                                                  int(iteration), C.c_void_p(out.ctypes.data)))
    return out
