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
import os
from enum import IntEnum

import numpy as np

from . import _lib
from ._lib import Context, XnyzError  # noqa: F401

__all__ = ["QUANTIZER_TYPE", "BlockCompressor", "BlockCompressorRLE", "BlockDecompressor",
           "BlockDecompressorRLE", "decompress", "Context", "XnyzError",
           "compress", "compress_batch", "compress_batch_into", "compress_batch_noalloc", "host_buffer", "search_batch", "features_batch", "mlp_train",
           "mlp_predict", "compress_selected_batch", "compress_selected_batch_into", "quantize", "morton_to_hilbert", "autotune",
           "decompress_device", "decompress_to_device", "hilbert_to_morton",
           "container_write", "container_read", "container_decompress", "fit_profile",
           "compress_with_profile_batch", "compress_with_profile_batch_into",
           "compress_global_rle", "compress_global_rle_into"]


class QUANTIZER_TYPE(IntEnum):  # source/composer/utils.hpp:55-59 (ADAPTIVE not offered)
    CUBE = 0
    TRUNCATED_OCTAHEDRON = 1


_default_ctx: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = _default_ctx[device] = Context(device)
    return ctx


def _opts(quantizer, l2_bound, is_hilbert, rle, direct_threshold, frame_mode, zstd_workers,
          gr_k=None, pfd_width=None):
    flags = (_lib.MODE_GR_K if gr_k is not None else 0) | \
        (_lib.MODE_PFD_WIDTH if pfd_width is not None else 0)
    return _lib.Opts(int(quantizer), float(l2_bound), int(bool(is_hilbert)), int(bool(rle)),
                     int(direct_threshold), int(frame_mode), int(zstd_workers), flags,
                     int(gr_k or 0), int(pfd_width or 0))


def compress(points, quantizer=QUANTIZER_TYPE.TRUNCATED_OCTAHEDRON, l2_bound=0.01,
             is_hilbert=False, rle=False, direct_threshold=1024,
             frame_mode=_lib.FRAME_ZSTD3, zstd_workers=0, want_prezstd=False,
             want_perm=False, ctx: Context | None = None, gr_k=None, pfd_width=None):
    """Compress one cloud.  Returns (frame, prezstd or None, perm or None, stats).

    A torch tensor on a HIP device is compressed in place (no copy); any other
    input is converted to a contiguous float32 host array and uploaded.
    gr_k / pfd_width force the Golomb-Rice k and the PFD width of the repos
    (or sfc-delta) stream; None keeps the reference's choice."""
    L = _lib.lib()
    o = _opts(quantizer, l2_bound, is_hilbert, rle, direct_threshold, frame_mode, zstd_workers,
              gr_k, pfd_width)
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
        _inputs_ready(keep.device)  # the library works on its own stream
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


def compress_batch(points_device, offsets, opts, want_prezstd: bool = False,
                   ctx: Context | None = None):
    """Compress many clouds in one launch set.

    points_device: contiguous float32 (N, 3) tensor on the GPU holding every
    cloud back to back; offsets: N+1-style list of cloud starts (len = count+1);
    opts: one dict per cloud with keys quantizer, l2_bound, is_hilbert, rle and
    optionally direct_threshold, frame_mode.  Returns (frames, prezstd, stats)."""
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    count = len(offsets) - 1
    if count <= 0 or len(opts) != count:
        raise ValueError("need one opts entry per cloud")
    ctx = ctx or default_context(t.device.index or 0)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    arr = (_lib.Opts * count)()
    for i, o in enumerate(opts):
        arr[i] = _opts(o["quantizer"], o["l2_bound"], o.get("is_hilbert", False),
                       o.get("rle", False), o.get("direct_threshold", 1024),
                       o.get("frame_mode", _lib.FRAME_ZSTD3), o.get("zstd_workers", 0),
                       o.get("gr_k"), o.get("pfd_width"))
    frames = (_lib.Buf * count)()
    pres = (_lib.Buf * count)() if want_prezstd else None
    stats = (_lib.Stats * count)()
    _inputs_ready(t.device)
    _lib.check(_lib.lib().xnyz_compress_batch_device(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.cast(frames, C.c_void_p),
        C.cast(pres, C.c_void_p) if pres is not None else None, C.cast(stats, C.c_void_p)))
    out_frames = [_lib.take(f) for f in frames]
    out_pres = [_lib.take(p) for p in pres] if pres is not None else None
    return out_frames, out_pres, list(stats)


def _inputs_ready(device) -> None:
    """Waits for torch's current stream, which produced the inputs: the
    library reads them on its own stream.  Not a device-wide synchronise, which
    would also wait for the library's copies still in flight
    (compress_batch_into with wait=False)."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def compress_batch_into(points_device, offsets, opts, out: np.ndarray,
                        ctx: Context | None = None, wait: bool = True):
    """compress_batch writing every cloud's output back to back into `out`
    (a uint8 numpy array, reused across calls).  Returns (offsets, stats):
    cloud s's output is out[offsets[s]:offsets[s + 1]]; offsets[0] is 0
    except for a split frame with room for its anchored layout (the host's
    blocks placed while compressed, the device's part in front of them).
    Raises XnyzError with code XNYZ_ENOSPC if `out` is too small
    (offsets[-1] then holds the size needed).  wait=False
    (xnyz_compress_batch_device_into_async: assembled streams into a pinned
    `out` from host_buffer) returns once the copy is queued; the bytes are
    in `out` after ctx.wait()."""
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    if out.dtype != np.uint8 or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous uint8 array")
    count = len(offsets) - 1
    if count <= 0 or len(opts) != count:
        raise ValueError("need one opts entry per cloud")
    ctx = ctx or default_context(t.device.index or 0)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    arr = (_lib.Opts * count)()
    for i, o in enumerate(opts):
        arr[i] = _opts(o["quantizer"], o["l2_bound"], o.get("is_hilbert", False),
                       o.get("rle", False), o.get("direct_threshold", 1024),
                       o.get("frame_mode", _lib.FRAME_ZSTD3), o.get("zstd_workers", 0),
                       o.get("gr_k"), o.get("pfd_width"))
    res = np.zeros(count + 1, dtype=np.uint64)
    stats = (_lib.Stats * count)()
    _inputs_ready(t.device)
    fn = (_lib.lib().xnyz_compress_batch_device_into if wait
          else _lib.lib().xnyz_compress_batch_device_into_async)
    rc = fn(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.c_void_p(out.ctypes.data), out.size,
        C.c_void_p(res.ctypes.data), C.cast(stats, C.c_void_p))
    _lib.check(rc)
    return res, list(stats)


def compress_batch_noalloc(points_device, offsets, opts, out: np.ndarray,
                           ctx: Context | None = None):
    """compress_batch_into without any allocation by the library during the
    call (assembled streams only; at most 2^22 points and 256 clouds; calls
    on one GPU must not overlap).  Returns (offsets, stats)."""
    t, count, off, arr = _batch_args(points_device, offsets, opts, _lib.FRAME_NONE)
    if out.dtype != np.uint8 or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous uint8 array")
    ctx = ctx or default_context(t.device.index or 0)
    res = np.zeros(count + 1, dtype=np.uint64)
    stats = (_lib.Stats * count)()
    _lib.check(_lib.lib().xnyz_compress_batch_device_noalloc(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.c_void_p(out.ctypes.data), out.size,
        C.c_void_p(res.ctypes.data), C.cast(stats, C.c_void_p)))
    return res, list(stats)


class _Pinned:
    """Owner of one xnyz_host_alloc buffer (freed with the last array
    viewing it)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _lib.check(_lib.lib().xnyz_host_alloc(nbytes, C.byref(p)))
        self.ptr = p.value

    def __del__(self):
        if self.ptr:
            _lib.lib().xnyz_host_free(C.c_void_p(self.ptr))
            self.ptr = None


def host_buffer(nbytes: int) -> np.ndarray:
    """A uint8 array of `nbytes` in pinned host memory (xnyz_host_alloc): as
    the `out` of compress_batch_into it takes the device-built blocks of a
    split frame by DMA, with no staging copy."""
    if nbytes <= 0:
        raise ValueError("nbytes must be positive")
    owner = _Pinned(int(nbytes))
    view = (C.c_uint8 * int(nbytes)).from_address(owner.ptr)
    view._owner = owner
    return np.frombuffer(view, dtype=np.uint8)


def autotune(points_device, opts: dict, reps: int = 2, ctx: Context | None = None):
    """Time compressions of one device-resident cloud over candidate run-time
    parameters (persistent grid; for FRAME_ZSTD3_MT also the host chunk size,
    the device's share of the frame and the copy-back slice) and keep the
    fastest as the context's tuning for the cloud's size class.  Returns
    (tuning dict, mean ms per call)."""
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    ctx = ctx or default_context(t.device.index or 0)
    o = _opts(opts["quantizer"], opts["l2_bound"], opts.get("is_hilbert", False),
              opts.get("rle", False), opts.get("direct_threshold", 1024),
              opts.get("frame_mode", _lib.FRAME_ZSTD3_MT), opts.get("zstd_workers", 0),
              opts.get("gr_k"), opts.get("pfd_width"))
    best = _lib.Tuning()
    ms = C.c_double()
    _inputs_ready(t.device)
    _lib.check(_lib.lib().xnyz_autotune(ctx.handle, C.c_void_p(t.data_ptr()), t.shape[0],
                                        C.byref(o), int(reps), C.byref(best), C.byref(ms)))
    return best.as_dict(), ms.value


def search_batch(points_device, offsets, opts, ctx: Context | None = None):
    """Exhaustive per-chunk mode search (curve x RLE x Golomb-Rice k x PFD
    width) over the clouds of a batch; sizes are exact assembled-stream bytes.
    opts: one dict per cloud (quantizer, l2_bound, optional direct_threshold,
    and the reference's is_hilbert / rle for `reference_bytes`).  Returns one
    dict per cloud: best (mode dict), best_bytes, reference_bytes, variants."""
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    count = len(offsets) - 1
    if count <= 0 or len(opts) != count:
        raise ValueError("need one opts entry per cloud")
    ctx = ctx or default_context(t.device.index or 0)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    arr = (_lib.Opts * count)()
    for i, o in enumerate(opts):
        arr[i] = _opts(o["quantizer"], o["l2_bound"], o.get("is_hilbert", False),
                       o.get("rle", False), o.get("direct_threshold", 1024), _lib.FRAME_NONE, 0)
    res = (_lib.SearchResult * count)()
    _inputs_ready(t.device)
    _lib.check(_lib.lib().xnyz_search_batch_device(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.cast(res, C.c_void_p)))
    out = []
    for r in res:
        out.append({"best": r.best.as_dict(), "best_bytes": r.best_bytes,
                    "reference_bytes": r.reference_bytes,
                    "variants": [dict(r.variant[v].as_dict(), bytes=r.variant_bytes[v])
                                 for v in range(4)]})
    return out


def _batch_args(points_device, offsets, opts, frame_default=_lib.FRAME_NONE):
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    count = len(offsets) - 1
    if count <= 0 or len(opts) != count:
        raise ValueError("need one opts entry per cloud")
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    arr = (_lib.Opts * count)()
    for i, o in enumerate(opts):
        arr[i] = _opts(o["quantizer"], o["l2_bound"], o.get("is_hilbert", False),
                       o.get("rle", False), o.get("direct_threshold", 1024),
                       o.get("frame_mode", frame_default), o.get("zstd_workers", 0))
    _inputs_ready(t.device)
    return t, count, off, arr


def features_batch(points_device, offsets, opts, ctx: Context | None = None) -> np.ndarray:
    """Selector features of every cloud: (count, 32) float32."""
    t, count, off, arr = _batch_args(points_device, offsets, opts)
    ctx = ctx or default_context(t.device.index or 0)
    out = np.zeros((count, _lib.MLP_FEATURES), dtype=np.float32)
    _lib.check(_lib.lib().xnyz_features_batch_device(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.c_void_p(out.ctypes.data)))
    return out


def _modes_array(modes):
    arr = (_lib.Mode * len(modes))()
    for i, m in enumerate(modes):
        arr[i] = _lib.Mode(int(m["is_hilbert"]), int(m["rle"]), int(m["gr_k"]),
                           int(m["pfd_width"]))
    return arr


def mlp_train(features, modes, epochs=30, batch=256, lr=3e-3, seed=0, params=None,
              ctx: Context | None = None):
    """Train the selector on (features, mode dicts); returns (params, final loss)."""
    x = np.ascontiguousarray(features, dtype=np.float32)
    n = x.shape[0]
    p = np.zeros(_lib.MLP_PARAMS, np.float32) if params is None else \
        np.array(params, dtype=np.float32)
    to = _lib.TrainOpts(int(epochs), int(batch), float(lr), int(params is None), int(seed))
    loss = C.c_float()
    ctx = ctx or default_context(0)
    _lib.check(_lib.lib().xnyz_mlp_train(ctx.handle, C.c_void_p(x.ctypes.data),
                                          C.cast(_modes_array(modes), C.c_void_p), n,
                                          C.byref(to), C.c_void_p(p.ctypes.data),
                                          C.byref(loss)))
    return p, loss.value


def mlp_predict(params, features, ctx: Context | None = None):
    """Predicted mode dicts and confidences."""
    x = np.ascontiguousarray(features, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    n = x.shape[0]
    out = (_lib.Mode * n)()
    conf = np.zeros(n, np.float32)
    ctx = ctx or default_context(0)
    _lib.check(_lib.lib().xnyz_mlp_predict(ctx.handle, C.c_void_p(p.ctypes.data),
                                            C.c_void_p(x.ctypes.data), n,
                                            C.cast(out, C.c_void_p),
                                            C.c_void_p(conf.ctypes.data)))
    return [m.as_dict() for m in out], conf


def compress_selected_batch(points_device, offsets, opts, params=None, min_confidence=0.0,
                            want_prezstd=False, ctx: Context | None = None):
    """Batched compression with a mode per cloud: the selector's prediction
    (params) where confident, the exhaustive search otherwise (params None:
    search everything).  Returns (frames, prezstd, stats, modes, searched)."""
    t, count, off, arr = _batch_args(points_device, offsets, opts, _lib.FRAME_ZSTD3)
    ctx = ctx or default_context(t.device.index or 0)
    p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
    frames = (_lib.Buf * count)()
    pres = (_lib.Buf * count)() if want_prezstd else None
    stats = (_lib.Stats * count)()
    modes = (_lib.Mode * count)()
    searched = np.zeros(count, np.int32)
    _lib.check(_lib.lib().xnyz_compress_selected_batch_device(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.c_void_p(p.ctypes.data) if p is not None else None,
        float(min_confidence), C.cast(frames, C.c_void_p),
        C.cast(pres, C.c_void_p) if pres is not None else None, C.cast(stats, C.c_void_p),
        C.cast(modes, C.c_void_p), C.c_void_p(searched.ctypes.data)))
    out_frames = [_lib.take(f) for f in frames]
    out_pres = [_lib.take(b) for b in pres] if pres is not None else None
    return out_frames, out_pres, list(stats), [m.as_dict() for m in modes], searched


def compress_selected_batch_into(points_device, offsets, opts, out: np.ndarray, params=None,
                                 min_confidence=0.0, ctx: Context | None = None,
                                 wait: bool = True):
    """compress_selected_batch writing the outputs back to back into `out`
    (a contiguous uint8 array, reused across calls; layout as
    compress_batch_into).  Returns (offsets, stats, modes, searched).
    wait=False (xnyz_compress_selected_batch_device_into_async: assembled
    streams into a pinned `out` from host_buffer) returns once the copy is
    queued; the bytes are in `out` after ctx.wait()."""
    t, count, off, arr = _batch_args(points_device, offsets, opts, _lib.FRAME_ZSTD3)
    if out.dtype != np.uint8 or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous uint8 array")
    ctx = ctx or default_context(t.device.index or 0)
    p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
    res = np.zeros(count + 1, dtype=np.uint64)
    stats = (_lib.Stats * count)()
    modes = (_lib.Mode * count)()
    searched = np.zeros(count, np.int32)
    fn = (_lib.lib().xnyz_compress_selected_batch_device_into if wait
          else _lib.lib().xnyz_compress_selected_batch_device_into_async)
    _lib.check(fn(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.c_void_p(p.ctypes.data) if p is not None else None,
        float(min_confidence), C.c_void_p(out.ctypes.data), out.size,
        C.c_void_p(res.ctypes.data), C.cast(stats, C.c_void_p), C.cast(modes, C.c_void_p),
        C.c_void_p(searched.ctypes.data)))
    return res, list(stats), [m.as_dict() for m in modes], searched


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


def decompress(data: bytes, quantizer, l2_bound: float, rle: bool,
               is_frame: bool = True) -> np.ndarray:
    """BlockDecompressor[RLE]<float>::decompress: a zstd frame (or, with
    is_frame=False, an assembled stream) -> (N, 3) float32 points."""
    o = _opts(quantizer, l2_bound, False, rle, 1024,
              _lib.FRAME_ZSTD3 if is_frame else _lib.FRAME_NONE, 0)
    out = _lib.Buf()
    src = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    _lib.check(_lib.lib().xnyz_decompress(C.c_void_p(src.ctypes.data), len(data), C.byref(o),
                                           C.byref(out)))
    return np.frombuffer(_lib.take(out), dtype=np.float32).reshape(-1, 3).copy()


def decompress_device(data: bytes, quantizer, l2_bound: float, rle: bool,
                      is_frame: bool = True, ctx: Context | None = None) -> np.ndarray:
    """decompress with the decoding on the GPU (xnyz_decompress_device): the
    same (N, 3) float32 points."""
    o = _opts(quantizer, l2_bound, False, rle, 1024,
              _lib.FRAME_ZSTD3 if is_frame else _lib.FRAME_NONE, 0)
    out = _lib.Buf()
    src = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    ctx = ctx or default_context(0)
    _lib.check(_lib.lib().xnyz_decompress_device(ctx.handle, C.c_void_p(src.ctypes.data),
                                                  len(data), C.byref(o), C.byref(out)))
    return np.frombuffer(_lib.take(out), dtype=np.float32).reshape(-1, 3).copy()


def decompress_to_device(data: bytes, quantizer, l2_bound: float, rle: bool, out_device,
                         is_frame: bool = True, ctx: Context | None = None) -> int:
    """Decodes on the GPU into `out_device` (a contiguous float32 (M, 3) tensor
    on the GPU, M >= the point count); returns the point count.  Raises
    XnyzError (XNYZ_ENOSPC) when M is too small."""
    import torch
    t = out_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("out_device must be a contiguous float32 (M, 3) tensor on the GPU")
    o = _opts(quantizer, l2_bound, False, rle, 1024,
              _lib.FRAME_ZSTD3 if is_frame else _lib.FRAME_NONE, 0)
    src = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    n = C.c_uint64()
    ctx = ctx or default_context(t.device.index or 0)
    _inputs_ready(t.device)
    _lib.check(_lib.lib().xnyz_decompress_to_device(
        ctx.handle, C.c_void_p(src.ctypes.data), len(data), C.byref(o),
        C.c_void_p(t.data_ptr()), t.shape[0], C.byref(n)))
    return n.value


def hilbert_to_morton(codes, iteration: int, ctx: Context | None = None) -> np.ndarray:
    """hilbertToMorton3 of each code on the GPU."""
    a = np.ascontiguousarray(codes, dtype=np.uint64)
    out = np.empty_like(a)
    ctx = ctx or default_context(0)
    _lib.check(_lib.lib().xnyz_hilbert_to_morton(ctx.handle, C.c_void_p(a.ctypes.data), a.size,
                                                  int(iteration), C.c_void_p(out.ctypes.data)))
    return out


def fit_profile(points_device, quantizer, l2_bound, is_hilbert=False,
                frame_mode=_lib.FRAME_ZSTD3, want_prezstd=False, ctx: Context | None = None):
    """BlockCompressor::fit_profile on a GPU tensor: (frame, prezstd or None,
    serialized profile bytes)."""
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    ctx = ctx or default_context(t.device.index or 0)
    o = _opts(quantizer, l2_bound, is_hilbert, False, 1024, frame_mode, 0)
    fr, pre, prof = _lib.Buf(), _lib.Buf(), _lib.Buf()
    _inputs_ready(t.device)
    _lib.check(_lib.lib().xnyz_fit_profile_device(ctx.handle, C.c_void_p(t.data_ptr()), t.shape[0],
                                                   C.byref(o), C.byref(fr),
                                                   C.byref(pre) if want_prezstd else None,
                                                   C.byref(prof)))
    return _lib.take(fr), (_lib.take(pre) if want_prezstd else None), _lib.take(prof)


def compress_with_profile_batch(points_device, offsets, opts, profile: bytes,
                                want_prezstd=False, ctx: Context | None = None):
    """BlockCompressor::compress_with_profile over a batch: (frames, prezstd, stats)."""
    t, count, off, arr = _batch_args(points_device, offsets, opts, _lib.FRAME_ZSTD3)
    ctx = ctx or default_context(t.device.index or 0)
    prof = np.frombuffer(profile, np.uint8)
    frames = (_lib.Buf * count)()
    pres = (_lib.Buf * count)() if want_prezstd else None
    stats = (_lib.Stats * count)()
    _lib.check(_lib.lib().xnyz_compress_with_profile_batch_device(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.c_void_p(prof.ctypes.data), len(profile),
        C.cast(frames, C.c_void_p), C.cast(pres, C.c_void_p) if pres is not None else None,
        C.cast(stats, C.c_void_p)))
    return ([_lib.take(f) for f in frames],
            [_lib.take(p) for p in pres] if pres is not None else None, list(stats))


def compress_with_profile_batch_into(points_device, offsets, opts, profile: bytes,
                                     out: np.ndarray, ctx: Context | None = None):
    """compress_with_profile_batch writing into `out` (layout as
    compress_batch_into).  Returns (offsets, stats)."""
    t, count, off, arr = _batch_args(points_device, offsets, opts, _lib.FRAME_ZSTD3)
    if out.dtype != np.uint8 or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous uint8 array")
    ctx = ctx or default_context(t.device.index or 0)
    prof = np.frombuffer(profile, np.uint8)
    res = np.zeros(count + 1, dtype=np.uint64)
    stats = (_lib.Stats * count)()
    _lib.check(_lib.lib().xnyz_compress_with_profile_batch_device_into(
        ctx.handle, C.c_void_p(t.data_ptr()), C.c_void_p(off.ctypes.data), count,
        C.cast(arr, C.c_void_p), C.c_void_p(prof.ctypes.data), len(profile),
        C.c_void_p(out.ctypes.data), out.size, C.c_void_p(res.ctypes.data),
        C.cast(stats, C.c_void_p)))
    return res, list(stats)


def compress_global_rle_into(points_device, quantizer, l2_bound, out: np.ndarray,
                             is_hilbert=False, frame_mode=_lib.FRAME_ZSTD3,
                             ctx: Context | None = None):
    """compress_global_rle writing into `out`.  Returns (offsets, stats): the
    output is out[offsets[0]:offsets[1]]."""
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    if out.dtype != np.uint8 or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous uint8 array")
    ctx = ctx or default_context(t.device.index or 0)
    o = _opts(quantizer, l2_bound, is_hilbert, True, 1024, frame_mode, 0)
    res = np.zeros(2, dtype=np.uint64)
    st = _lib.Stats()
    _inputs_ready(t.device)
    _lib.check(_lib.lib().xnyz_compress_global_rle_device_into(
        ctx.handle, C.c_void_p(t.data_ptr()), t.shape[0], C.byref(o),
        C.c_void_p(out.ctypes.data), out.size, C.c_void_p(res.ctypes.data), C.byref(st)))
    return res, st


def compress_global_rle(points_device, quantizer, l2_bound, is_hilbert=False,
                        frame_mode=_lib.FRAME_ZSTD3, want_prezstd=False,
                        ctx: Context | None = None):
    """The MPI driver's mapreduce mode on one GPU: (frame, prezstd or None, stats)."""
    import torch
    t = points_device.reshape(-1, 3)
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("points must be a contiguous float32 (N, 3) tensor on the GPU")
    ctx = ctx or default_context(t.device.index or 0)
    o = _opts(quantizer, l2_bound, is_hilbert, True, 1024, frame_mode, 0)
    fr, pre, st = _lib.Buf(), _lib.Buf(), _lib.Stats()
    _inputs_ready(t.device)
    _lib.check(_lib.lib().xnyz_compress_global_rle_device(
        ctx.handle, C.c_void_p(t.data_ptr()), t.shape[0], C.byref(o), C.byref(fr),
        C.byref(pre) if want_prezstd else None, C.byref(st)))
    return _lib.take(fr), (_lib.take(pre) if want_prezstd else None), st


def container_write(part_path: str, manifest_path: str, out: np.ndarray, offsets, chunks):
    """Writes a multi-chunk container (xnyz_container_write): chunk s is
    out[offsets[s]:offsets[s + 1]], described by chunks[s] (dict with
    input_offset, input_points, quantizer, l2_bound, is_hilbert, rle and
    optionally gr_k, pfd_width, direct_threshold, frame_mode).  Returns the
    chunk dicts with part_offset and compressed_size filled."""
    count = len(offsets) - 1
    if count != len(chunks):
        raise ValueError("need one chunk entry per output")
    buf = np.ascontiguousarray(out, dtype=np.uint8)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    arr = (_lib.Chunk * count)()
    for i, c in enumerate(chunks):
        arr[i] = _lib.Chunk(int(c.get("input_offset", 0)), int(c.get("input_points", 0)), 0, 0,
                            int(c["quantizer"]), float(c["l2_bound"]),
                            int(bool(c.get("is_hilbert", False))), int(bool(c.get("rle", False))),
                            int(c.get("gr_k", -1)), int(c.get("pfd_width", 8)),
                            int(c.get("direct_threshold", 1024)),
                            int(c.get("frame_mode", _lib.FRAME_ZSTD3)))
    _lib.check(_lib.lib().xnyz_container_write(os.fsencode(part_path), os.fsencode(manifest_path),
                                                C.c_void_p(buf.ctypes.data),
                                                C.c_void_p(off.ctypes.data), count,
                                                C.cast(arr, C.c_void_p)))
    return [a.as_dict() for a in arr]


def container_read(manifest_path: str):
    """The chunk dicts of a container's manifest."""
    b = _lib.Buf()
    _lib.check(_lib.lib().xnyz_container_read(os.fsencode(manifest_path), C.byref(b)))
    raw = _lib.take(b)
    n = len(raw) // C.sizeof(_lib.Chunk)
    arr = (_lib.Chunk * n).from_buffer_copy(raw) if n else []
    return [a.as_dict() for a in arr]


def container_decompress(part_path: str, manifest_path: str, gpu: bool = True,
                         ctx: Context | None = None) -> np.ndarray:
    """All points of a container, chunk after chunk (xnyz_container_decompress):
    decoded on the GPU, or on the host with gpu=False."""
    if gpu:
        ctx = ctx or default_context(0)
    b = _lib.Buf()
    _lib.check(_lib.lib().xnyz_container_decompress(ctx.handle if gpu else None,
                                                     os.fsencode(part_path),
                                                     os.fsencode(manifest_path), C.byref(b)))
    return np.frombuffer(_lib.take(b), dtype=np.float32).reshape(-1, 3).copy()


class BlockDecompressor:
    """BlockDecompressor<float> (source/composer/block_decompressor.hpp:24-192)."""

    rle = False

    def __init__(self, quantizer_type, l2_bound: float, is_debugging: bool = False):
        self.quantizer_type = QUANTIZER_TYPE(quantizer_type)
        self.l2_bound = float(np.float32(l2_bound))

    def decompress(self, compressed: bytes) -> np.ndarray:
        return decompress(compressed, self.quantizer_type, self.l2_bound, self.rle)


class BlockDecompressorRLE(BlockDecompressor):
    """BlockDecompressorRLE<float> (source/composer/block_decompressor_rle.hpp:25-207)."""

    rle = True


def quantize(points_device, quantizer, l2_bound, ctx: Context | None = None):
    """Shift + quantize a device tensor; returns (nodes int32 (N,3), offset float32 (3,))."""
    t = points_device.reshape(-1, 3)
    n = t.shape[0]
    nodes = np.empty((n, 3), dtype=np.int32)
    off = np.empty(3, dtype=np.float32)
    ctx = ctx or default_context(t.device.index or 0)
    import torch
    _inputs_ready(t.device)
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
