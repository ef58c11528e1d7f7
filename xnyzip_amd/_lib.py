"""ctypes binding of libxnyz.so (include/xnyz.h).

The library is built in-tree by `python -m xnyzip_amd.build`.  There is no
fallback: if the library or a HIP device is missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libxnyz.so")

XNYZ_OK = 0
XNYZ_EINVAL = -1
XNYZ_EUNSUPPORTED = -5
XNYZ_ENOSPC = -8
XNYZ_EIO = -9
CUBE, TRUNCATED_OCTAHEDRON = 0, 1
FRAME_ZSTD3, FRAME_ZSTD3_MT, FRAME_NONE, FRAME_ZSTD_GPU = 0, 1, 2, 3


class XnyzError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"xnyz error {code}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("quantizer", C.c_int32), ("l2_bound", C.c_float), ("is_hilbert", C.c_int32),
                ("rle", C.c_int32), ("direct_threshold", C.c_int32),
                ("frame_mode", C.c_int32), ("zstd_workers", C.c_int32),
                ("mode_flags", C.c_int32), ("gr_k", C.c_int32), ("pfd_width", C.c_int32)]


MODE_GR_K, MODE_PFD_WIDTH = 1, 2


class Mode(C.Structure):
    _fields_ = [("is_hilbert", C.c_int32), ("rle", C.c_int32), ("gr_k", C.c_int32),
                ("pfd_width", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


MLP_FEATURES, MLP_PARAMS = 32, 9196


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float),
                ("init", C.c_int32), ("seed", C.c_uint64)]


class SearchResult(C.Structure):
    _fields_ = [("best", Mode), ("best_bytes", C.c_uint64), ("reference_bytes", C.c_uint64),
                ("variant", Mode * 4), ("variant_bytes", C.c_uint64 * 4)]


class Tuning(C.Structure):
    """Run-time parameters of a context for one size class (include/xnyz.h)."""
    _fields_ = [("zstd_chunk_log", C.c_uint32), ("copy_slice_log", C.c_uint32),
                ("gpu_frame_share", C.c_float), ("persist_blocks", C.c_uint32),
                ("copy_first", C.c_uint32), ("radix_bits", C.c_uint32),
                ("radix_tile", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


SIZE_CLASSES = 33


class Buf(C.Structure):
    _fields_ = [("data", C.c_void_p), ("len", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("input_points", C.c_uint64), ("num_points", C.c_uint64),
                ("range", C.c_int32 * 3), ("is_direct", C.c_int32), ("blocks", C.c_uint64),
                ("patches", C.c_uint64), ("quad_patches", C.c_uint64),
                ("wrapped_points", C.c_uint64), ("counts_exp_golomb", C.c_int32),
                ("counts_eg_k", C.c_int32), ("entropy_repos", C.c_double),
                ("entropy_quads", C.c_double), ("entropy_counts", C.c_double),
                ("max_l2", C.c_double), ("mse", C.c_double), ("prezstd_bytes", C.c_uint64),
                ("frame_bytes", C.c_uint64), ("ms_device", C.c_double),
                ("ms_zstd", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if name == "range" else v
        return out


class Chunk(C.Structure):
    _fields_ = [("input_offset", C.c_uint64), ("input_points", C.c_uint64),
                ("part_offset", C.c_uint64), ("compressed_size", C.c_uint64),
                ("quantizer", C.c_int32), ("l2_bound", C.c_float), ("is_hilbert", C.c_int32),
                ("rle", C.c_int32), ("gr_k", C.c_int32), ("pfd_width", C.c_int32),
                ("direct_threshold", C.c_int32), ("frame_mode", C.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None
_lock = threading.Lock()

EXPORTS = ["xnyz_abi_version", "xnyz_last_error", "xnyz_create", "xnyz_destroy",
           "xnyz_compress", "xnyz_compress_device", "xnyz_compress_batch_device",
           "xnyz_compress_batch_device_into", "xnyz_compress_batch_device_into_async",
           "xnyz_wait", "xnyz_compress_batch_device_noalloc",
           "xnyz_search_batch_device", "xnyz_features_batch_device", "xnyz_mlp_train",
           "xnyz_mlp_predict", "xnyz_compress_selected_batch_device",
           "xnyz_compress_selected_batch_device_into",
           "xnyz_compress_selected_batch_device_into_async",
           "xnyz_quantize_device",
           "xnyz_morton_to_hilbert", "xnyz_hilbert_to_morton", "xnyz_decompress",
           "xnyz_decompress_device", "xnyz_decompress_to_device", "xnyz_container_write",
           "xnyz_container_read", "xnyz_container_decompress", "xnyz_fit_profile_device",
           "xnyz_compress_with_profile_batch_device",
           "xnyz_compress_with_profile_batch_device_into", "xnyz_compress_global_rle_device",
           "xnyz_compress_global_rle_device_into",
           "xnyz_graphs_enable",
           "xnyz_tuning_set", "xnyz_tuning_get", "xnyz_tuning_save", "xnyz_tuning_load",
           "xnyz_autotune", "xnyz_profile_enable", "xnyz_profile_report", "xnyz_free",
           "xnyz_host_alloc", "xnyz_host_free"]


def lib() -> C.CDLL:
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m xnyzip_amd.build`")
        L = C.CDLL(LIB_PATH)
        P, PP = C.c_void_p, C.POINTER(C.c_void_p)
        L.xnyz_abi_version.restype = C.c_int
        L.xnyz_last_error.restype = C.c_char_p
        L.xnyz_create.argtypes = [C.c_int, PP]
        L.xnyz_destroy.argtypes = [P]
        comp = [P, P, C.c_uint64, C.POINTER(Opts), C.POINTER(Buf), C.POINTER(Buf), P,
                C.POINTER(Stats)]
        L.xnyz_compress.argtypes = comp
        L.xnyz_compress_device.argtypes = comp
        L.xnyz_compress_batch_device.argtypes = [P, P, P, C.c_uint32, P, P, P, P]
        L.xnyz_compress_batch_device_into.argtypes = [P, P, P, C.c_uint32, P, P, C.c_uint64, P,
                                                      P]
        L.xnyz_compress_batch_device_into_async.argtypes = [P, P, P, C.c_uint32, P, P,
                                                            C.c_uint64, P, P]
        L.xnyz_wait.argtypes = [P]
        L.xnyz_compress_batch_device_noalloc.argtypes = [P, P, P, C.c_uint32, P, P, C.c_uint64,
                                                         P, P]
        L.xnyz_search_batch_device.argtypes = [P, P, P, C.c_uint32, P, P]
        L.xnyz_features_batch_device.argtypes = [P, P, P, C.c_uint32, P, P]
        L.xnyz_mlp_train.argtypes = [P, P, P, C.c_uint64, C.POINTER(TrainOpts), P,
                                     C.POINTER(C.c_float)]
        L.xnyz_mlp_predict.argtypes = [P, P, P, C.c_uint64, P, P]
        L.xnyz_compress_selected_batch_device.argtypes = [P, P, P, C.c_uint32, P, P, C.c_float,
                                                          P, P, P, P, P]
        L.xnyz_compress_selected_batch_device_into.argtypes = [P, P, P, C.c_uint32, P, P,
                                                               C.c_float, P, C.c_uint64, P, P,
                                                               P, P]
        L.xnyz_compress_selected_batch_device_into_async.argtypes = [
            P, P, P, C.c_uint32, P, P, C.c_float, P, C.c_uint64, P, P, P, P]
        L.xnyz_quantize_device.argtypes = [P, P, C.c_uint64, C.c_int32, C.c_float, P, P]
        L.xnyz_morton_to_hilbert.argtypes = [P, P, C.c_uint64, C.c_uint32, P]
        L.xnyz_decompress.argtypes = [P, C.c_uint64, C.POINTER(Opts), C.POINTER(Buf)]
        L.xnyz_decompress_device.argtypes = [P, P, C.c_uint64, C.POINTER(Opts), C.POINTER(Buf)]
        L.xnyz_decompress_to_device.argtypes = [P, P, C.c_uint64, C.POINTER(Opts), P, C.c_uint64,
                                                C.POINTER(C.c_uint64)]
        L.xnyz_hilbert_to_morton.argtypes = [P, P, C.c_uint64, C.c_uint32, P]
        L.xnyz_fit_profile_device.argtypes = [P, P, C.c_uint64, C.POINTER(Opts), C.POINTER(Buf),
                                              C.POINTER(Buf), C.POINTER(Buf)]
        L.xnyz_compress_with_profile_batch_device.argtypes = [P, P, P, C.c_uint32, P, P,
                                                              C.c_uint64, P, P, P]
        L.xnyz_compress_global_rle_device.argtypes = [P, P, C.c_uint64, C.POINTER(Opts),
                                                      C.POINTER(Buf), C.POINTER(Buf),
                                                      C.POINTER(Stats)]
        L.xnyz_compress_with_profile_batch_device_into.argtypes = [
            P, P, P, C.c_uint32, P, P, C.c_uint64, P, C.c_uint64, P, P]
        L.xnyz_compress_global_rle_device_into.argtypes = [
            P, P, C.c_uint64, C.POINTER(Opts), P, C.c_uint64, P, C.POINTER(Stats)]
        L.xnyz_container_write.argtypes = [C.c_char_p, C.c_char_p, P, P, C.c_uint32, P]
        L.xnyz_container_read.argtypes = [C.c_char_p, C.POINTER(Buf)]
        L.xnyz_container_decompress.argtypes = [P, C.c_char_p, C.c_char_p, C.POINTER(Buf)]
        L.xnyz_profile_enable.argtypes = [P, C.c_int]
        L.xnyz_graphs_enable.argtypes = [P, C.c_int]
        L.xnyz_tuning_set.argtypes = [P, C.c_int, C.POINTER(Tuning)]
        L.xnyz_tuning_get.argtypes = [P, C.c_int, C.POINTER(Tuning)]
        L.xnyz_tuning_save.argtypes = [P, C.c_char_p]
        L.xnyz_tuning_load.argtypes = [P, C.c_char_p]
        L.xnyz_autotune.argtypes = [P, P, C.c_uint64, C.POINTER(Opts), C.c_uint32,
                                    C.POINTER(Tuning), C.POINTER(C.c_double)]
        L.xnyz_profile_report.argtypes = [P, C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]
        L.xnyz_free.argtypes = [C.POINTER(Buf)]
        L.xnyz_host_alloc.argtypes = [C.c_uint64, PP]
        L.xnyz_host_free.argtypes = [P]
        _lib = L
        return L


def check(rc: int) -> None:
    if rc != XNYZ_OK:
        raise XnyzError(rc, lib().xnyz_last_error().decode())


def take(buf: Buf) -> bytes:
    if not buf.data:
        return b""
    out = C.string_at(buf.data, buf.len)
    lib().xnyz_free(C.byref(buf))
    return out


class Context:
    """One xnyz_ctx: a HIP device, its stream, and pooled buffers."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().xnyz_create(device, C.byref(self._h)))
        self.device = device

    def close(self):
        if self._h:
            lib().xnyz_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def wait(self) -> None:
        """Waits for the copies queued by compress_batch_into(..., wait=False)."""
        check(lib().xnyz_wait(self._h))

    def graphs(self, on: bool) -> None:
        """Replay the stage-2 launch sequence from a captured hipGraph (default on)."""
        check(lib().xnyz_graphs_enable(self._h, int(on)))

    def tuning(self, size_class: int) -> dict:
        """Run-time parameters used for calls of 2^c <= points < 2^(c+1)."""
        t = Tuning()
        check(lib().xnyz_tuning_get(self._h, int(size_class), C.byref(t)))
        return t.as_dict()

    def set_tuning(self, size_class: int = -1, **fields) -> None:
        """Change fields of one size class's parameters (-1: every class,
        starting from class 0's values)."""
        t = Tuning()
        check(lib().xnyz_tuning_get(self._h, max(0, int(size_class)), C.byref(t)))
        for k, v in fields.items():
            if k not in dict(Tuning._fields_):
                raise ValueError(f"unknown tuning field {k}")
            setattr(t, k, v)
        check(lib().xnyz_tuning_set(self._h, int(size_class), C.byref(t)))

    def save_tuning(self, path: str) -> None:
        check(lib().xnyz_tuning_save(self._h, os.fsencode(path)))

    def load_tuning(self, path: str) -> None:
        check(lib().xnyz_tuning_load(self._h, os.fsencode(path)))

    def profile(self, on: bool) -> None:
        """Enable (and clear) or disable per-kernel event timing."""
        check(lib().xnyz_profile_enable(self._h, int(on)))

    def profile_report(self) -> dict:
        import json
        need = C.c_uint64()
        check(lib().xnyz_profile_report(self._h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value + 16)
        check(lib().xnyz_profile_report(self._h, buf, len(buf), C.byref(need)))
        return json.loads(buf.value.decode())
