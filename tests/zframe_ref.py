"""Test helpers for the zstd-format frame encoder: the host driver in
tests/native/zframe_host.cpp (built here with g++ into a temporary
directory) and libzstd's decoder through ctypes."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_lib = None
_zstd = None


def host_encoder():
    global _lib
    if _lib is None:
        d = tempfile.mkdtemp(prefix="zframe_host_")
        so = os.path.join(d, "zframe_host.so")
        subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                        os.path.join(ROOT, "tests", "native", "zframe_host.cpp")], check=True)
        L = C.CDLL(so)
        L.zf_encode_frame.restype = C.c_uint64
        L.zf_encode_frame.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.zf_frame_bound.restype = C.c_uint64
        L.zf_frame_bound.argtypes = [C.c_uint64]
        _lib = L
    return _lib


def host_frame(data: bytes) -> bytes:
    L = host_encoder()
    a = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
    out = np.zeros(L.zf_frame_bound(len(data)), np.uint8)
    n = L.zf_encode_frame(a.ctypes.data, len(data), out.ctypes.data)
    return out[:n].tobytes()


def zstd():
    global _zstd
    if _zstd is None:
        for name in ("/opt/conda/lib/libzstd.so", "libzstd.so.1", "libzstd.so"):
            try:
                _zstd = C.CDLL(name)
                break
            except OSError:
                continue
        _zstd.ZSTD_decompress.restype = C.c_size_t
        _zstd.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        _zstd.ZSTD_isError.argtypes = [C.c_size_t]
        _zstd.ZSTD_getErrorName.restype = C.c_char_p
        _zstd.ZSTD_getErrorName.argtypes = [C.c_size_t]
        _zstd.ZSTD_getFrameContentSize.restype = C.c_ulonglong
        _zstd.ZSTD_getFrameContentSize.argtypes = [C.c_void_p, C.c_size_t]
    return _zstd


def unzstd(frame: bytes) -> bytes:
    z = zstd()
    f = np.frombuffer(frame, np.uint8)
    n = z.ZSTD_getFrameContentSize(f.ctypes.data, len(frame))
    assert n < (1 << 40), "frame without a content size"
    out = np.zeros(max(int(n), 1), np.uint8)
    r = z.ZSTD_decompress(out.ctypes.data, int(n), f.ctypes.data, len(frame))
    if z.ZSTD_isError(r):
        raise ValueError(z.ZSTD_getErrorName(r).decode())
    return out[:r].tobytes()
