"""ctypes view of the CPU oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product package xnyzip_amd never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libxnyz_ref.so")

CUBE, TO = 0, 1
SYM_U64, SYM_U8, SYM_INT = 0, 1, 2
_SYM_DTYPE = {SYM_U64: np.uint64, SYM_U8: np.uint8, SYM_INT: np.int32}


class OracleInfo(C.Structure):
    _fields_ = [
        ("num_points", C.c_uint64),
        ("input_points", C.c_uint64),
        ("range", C.c_int32 * 3),
        ("is_direct", C.c_int32),
        ("blocks", C.c_uint64),
        ("patches", C.c_uint64),
        ("quad_patches", C.c_uint64),
        ("counts_exp_golomb", C.c_int32),
        ("counts_eg_k", C.c_int32),
        ("max_l2", C.c_double),
        ("mse", C.c_double),
        ("seconds_total", C.c_double),
        ("seconds_stream", C.c_double),
        ("wrapped_points", C.c_uint64),
    ]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_U64P = C.POINTER(C.c_uint64)


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run: make -C oracle)")
    return C.CDLL(path)


_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = _load(LIB_PATH)
        L.oracle_compress.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, C.c_int, C.c_int,
                                      C.c_int, _PP, _U64P, _PP, _U64P, _U64P,
                                      C.POINTER(OracleInfo)]
        L.oracle_compress_mode.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_int, _PP, _U64P,
                                           C.POINTER(OracleInfo)]
        L.oracle_compress_shared.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, C.c_int, _P,
                                             C.c_uint64, _PP, _U64P, _PP, _U64P]
        L.oracle_compress_global_rle.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, C.c_int,
                                                 _PP, _U64P]
        L.oracle_compress_timed.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, C.c_int,
                                            C.c_int, C.c_int, C.c_int, _U64P,
                                            C.POINTER(OracleInfo)]
        L.oracle_decompress_buffer.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, C.c_int,
                                               _PP, _U64P]
        L.oracle_decompress_frame.argtypes = L.oracle_decompress_buffer.argtypes
        L.oracle_shift.argtypes = [_P, C.c_uint64, _P, _P]
        L.oracle_quantize.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, _P, _P]
        L.oracle_recover.argtypes = [_P, C.c_uint64, C.c_int, C.c_float, _P]
        L.oracle_morton_code.argtypes = [C.c_int32] * 3
        L.oracle_morton_code.restype = C.c_uint64
        L.oracle_decode_morton.argtypes = [C.c_uint64, _P]
        L.oracle_morton_to_hilbert3.argtypes = [C.c_uint64, C.c_uint]
        L.oracle_morton_to_hilbert3.restype = C.c_uint64
        L.oracle_hilbert_to_morton3.argtypes = [C.c_uint64, C.c_uint]
        L.oracle_hilbert_to_morton3.restype = C.c_uint64
        L.oracle_hilbert_tables.argtypes = [_P, _P, _P, _P]
        for nm in ("oracle_bitpack_u64", "oracle_gr_pack_u64", "oracle_gr_pack_u8"):
            getattr(L, nm).argtypes = [_P, C.c_uint64, _PP, _U64P, _PP, _U64P]
        L.oracle_huffman.argtypes = [_P, C.c_uint64, C.c_int, C.c_int, _PP, _U64P, _PP, _U64P]
        L.oracle_huffman_decode.argtypes = [_P, C.c_uint64, _P, C.c_uint64, C.c_uint64,
                                            C.c_int, _P]
        L.oracle_gr_unpack_u64.argtypes = [_P, C.c_uint64, _P, C.c_uint64, _P, _U64P]
        L.oracle_bitunpack_u64.argtypes = [_P, C.c_uint64, _P, C.c_uint64, _P, _U64P]
        L.oracle_zstd_version.restype = C.c_uint64
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_free.argtypes = [_P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        R = _load(REF_PATH)
        for nm in ("ref_bitpack_u64", "ref_gr_pack_u64", "ref_gr_pack_u8"):
            getattr(R, nm).argtypes = [_P, C.c_uint64, _PP, _U64P, _PP, _U64P]
        R.ref_pfd_u64.argtypes = [_P, C.c_uint64, _P, _P, _U64P]
        R.ref_huffman.argtypes = [_P, C.c_uint64, C.c_int, C.c_int, _PP, _U64P, _PP, _U64P]
        R.ref_morton_to_hilbert3.argtypes = [C.c_uint64, C.c_uint]
        R.ref_morton_to_hilbert3.restype = C.c_uint64
        R.ref_hilbert_to_morton3.argtypes = [C.c_uint64, C.c_uint]
        R.ref_hilbert_to_morton3.restype = C.c_uint64
        R.ref_free.argtypes = [_P]
        _ref = R
    return _ref


def _err() -> str:
    return lib().oracle_last_error().decode()


def _take(ptr: C.c_void_p, n: int, free) -> bytes:
    if not ptr.value:
        return b""
    out = C.string_at(ptr.value, n)
    free(ptr)
    return out


def _f32(points) -> np.ndarray:
    a = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    return a


@dataclass
class Compressed:
    pre: bytes
    frame: bytes
    perm: np.ndarray | None
    info: OracleInfo


def compress(points, quantizer: int, l2: float, hilbert: bool, rle: bool,
             direct_threshold: int = 1024, want_perm: bool = False) -> Compressed:
    a = _f32(points)
    n = a.shape[0]
    pre, fr = C.c_void_p(), C.c_void_p()
    pl, fl = C.c_uint64(), C.c_uint64()
    info = OracleInfo()
    perm = np.zeros(n, dtype=np.uint64) if (want_perm and not rle) else None
    rc = lib().oracle_compress(a.ctypes.data, n, quantizer, l2, int(hilbert), int(rle),
                               direct_threshold, C.byref(pre), C.byref(pl), C.byref(fr),
                               C.byref(fl),
                               perm.ctypes.data_as(_U64P) if perm is not None else None,
                               C.byref(info))
    if rc != 0:
        raise RuntimeError(f"oracle_compress: {_err()}")
    f = lib().oracle_free
    return Compressed(_take(pre, pl.value, f), _take(fr, fl.value, f), perm, info)


def compress_mode(points, quantizer: int, l2: float, hilbert: bool, rle: bool,
                  direct_threshold: int = 1024, gr_k: int = -1, pfd_width: int = 0) -> bytes:
    """Assembled stream with the mode overrides (gr_k >= 0 forces the
    Golomb-Rice k of the repos / sfc-delta patches; pfd_width 1..8 sets the
    PFD limit to 2^w - 2)."""
    a = _f32(points)
    pre, pl = C.c_void_p(), C.c_uint64()
    info = OracleInfo()
    rc = lib().oracle_compress_mode(a.ctypes.data, a.shape[0], quantizer, l2, int(hilbert),
                                    int(rle), direct_threshold, int(gr_k), int(pfd_width),
                                    C.byref(pre), C.byref(pl), C.byref(info))
    if rc != 0:
        raise RuntimeError(f"oracle_compress_mode: {_err()}")
    return _take(pre, pl.value, lib().oracle_free)


def compress_shared(points, quantizer: int, l2: float, hilbert: bool, profile: bytes | None = None):
    """fit_profile (profile None) or compress_with_profile (block_compressor.hpp:241-431):
    (assembled stream, serialized profile or None)."""
    a = _f32(points)
    pre, pl = C.c_void_p(), C.c_uint64()
    po, pol = C.c_void_p(), C.c_uint64()
    pin = np.frombuffer(profile, np.uint8) if profile else None
    rc = lib().oracle_compress_shared(a.ctypes.data, a.shape[0], quantizer, l2, int(hilbert),
                                      pin.ctypes.data if pin is not None else None,
                                      len(profile) if profile else 0, C.byref(pre), C.byref(pl),
                                      None if profile else C.byref(po),
                                      None if profile else C.byref(pol))
    if rc != 0:
        raise RuntimeError(f"oracle_compress_shared: {_err()}")
    stream = _take(pre, pl.value, lib().oracle_free)
    prof = None if profile else _take(po, pol.value, lib().oracle_free)
    return stream, prof


def compress_global_rle(points, quantizer: int, l2: float, hilbert: bool) -> bytes:
    """The mapreduce mode's stream on one rank (mpi_driver.cpp:365-648,
    block_compressor_rle.hpp:341-470)."""
    a = _f32(points)
    pre, pl = C.c_void_p(), C.c_uint64()
    rc = lib().oracle_compress_global_rle(a.ctypes.data, a.shape[0], quantizer, l2, int(hilbert),
                                          C.byref(pre), C.byref(pl))
    if rc != 0:
        raise RuntimeError(f"oracle_compress_global_rle: {_err()}")
    return _take(pre, pl.value, lib().oracle_free)


def compress_timed(points, quantizer: int, l2: float, hilbert: bool, rle: bool,
                   direct_threshold: int = 1024, with_zstd: bool = True):
    a = _f32(points)
    info = OracleInfo()
    nbytes = C.c_uint64()
    rc = lib().oracle_compress_timed(a.ctypes.data, a.shape[0], quantizer, l2, int(hilbert),
                                     int(rle), direct_threshold, int(with_zstd),
                                     C.byref(nbytes), C.byref(info))
    if rc != 0:
        raise RuntimeError(f"oracle_compress_timed: {_err()}")
    return nbytes.value, info


def decompress_buffer(buf: bytes, quantizer: int, l2: float, rle: bool) -> np.ndarray:
    out, n = C.c_void_p(), C.c_uint64()
    b = np.frombuffer(buf, dtype=np.uint8)
    rc = lib().oracle_decompress_buffer(b.ctypes.data, len(buf), quantizer, l2, int(rle),
                                        C.byref(out), C.byref(n))
    if rc != 0:
        raise RuntimeError(f"oracle_decompress_buffer: {_err()}")
    raw = _take(out, n.value * 12, lib().oracle_free)
    return np.frombuffer(raw, dtype=np.float32).reshape(-1, 3).copy()


def decompress_frame(frame: bytes, quantizer: int, l2: float, rle: bool) -> np.ndarray:
    out, n = C.c_void_p(), C.c_uint64()
    b = np.frombuffer(frame, dtype=np.uint8)
    rc = lib().oracle_decompress_frame(b.ctypes.data, len(frame), quantizer, l2, int(rle),
                                       C.byref(out), C.byref(n))
    if rc != 0:
        raise RuntimeError(f"oracle_decompress_frame: {_err()}")
    raw = _take(out, n.value * 12, lib().oracle_free)
    return np.frombuffer(raw, dtype=np.float32).reshape(-1, 3).copy()


def shift(points):
    a = _f32(points)
    out = np.empty_like(a)
    off = np.empty(3, dtype=np.float32)
    lib().oracle_shift(a.ctypes.data, a.shape[0], out.ctypes.data, off.ctypes.data)
    return out, off


def quantize(points, quantizer: int, l2: float):
    a = _f32(points)
    nodes = np.empty((a.shape[0], 3), dtype=np.int32)
    off = np.empty(3, dtype=np.float32)
    lib().oracle_quantize(a.ctypes.data, a.shape[0], quantizer, l2, nodes.ctypes.data,
                          off.ctypes.data)
    return nodes, off


def recover(nodes, quantizer: int, l2: float) -> np.ndarray:
    q = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1, 3)
    out = np.empty(q.shape, dtype=np.float32)
    lib().oracle_recover(q.ctypes.data, q.shape[0], quantizer, l2, out.ctypes.data)
    return out


def morton_code(x: int, y: int, z: int) -> int:
    return lib().oracle_morton_code(x, y, z)


def decode_morton(m: int):
    out = np.empty(3, dtype=np.int32)
    lib().oracle_decode_morton(m, out.ctypes.data)
    return tuple(int(v) for v in out)


def m2h(m: int, it: int) -> int:
    return lib().oracle_morton_to_hilbert3(m, it)


def h2m(h: int, it: int) -> int:
    return lib().oracle_hilbert_to_morton3(h, it)


def hilbert_tables():
    a, b = np.empty(8, np.uint8), np.empty(8, np.uint8)
    c, d = np.empty(64, np.uint8), np.empty(64, np.uint8)
    lib().oracle_hilbert_tables(a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data)
    return a, b, c.reshape(8, 8), d.reshape(8, 8)


def _packed(fn, arr, free):
    m, d = C.c_void_p(), C.c_void_p()
    ml, dl = C.c_uint64(), C.c_uint64()
    rc = fn(arr.ctypes.data, arr.size, C.byref(m), C.byref(ml), C.byref(d), C.byref(dl))
    if rc != 0:
        raise ValueError("pack failed")
    return _take(m, ml.value, free), _take(d, dl.value, free)


def bitpack_u64(v, impl: str = "oracle"):
    a = np.ascontiguousarray(v, dtype=np.uint64)
    if impl == "ref":
        return _packed(ref().ref_bitpack_u64, a, ref().ref_free)
    return _packed(lib().oracle_bitpack_u64, a, lib().oracle_free)


def gr_pack(v, dtype=np.uint64, impl: str = "oracle"):
    a = np.ascontiguousarray(v, dtype=dtype)
    if dtype == np.uint8:
        fn = ref().ref_gr_pack_u8 if impl == "ref" else lib().oracle_gr_pack_u8
    else:
        fn = ref().ref_gr_pack_u64 if impl == "ref" else lib().oracle_gr_pack_u64
    free = ref().ref_free if impl == "ref" else lib().oracle_free
    return _packed(fn, a, free)


def huffman(v, sym_type: int, mode: int = 0, impl: str = "oracle"):
    a = np.ascontiguousarray(v, dtype=_SYM_DTYPE[sym_type])
    m, d = C.c_void_p(), C.c_void_p()
    ml, dl = C.c_uint64(), C.c_uint64()
    fn = ref().ref_huffman if impl == "ref" else lib().oracle_huffman
    free = ref().ref_free if impl == "ref" else lib().oracle_free
    rc = fn(a.ctypes.data, a.size, sym_type, mode, C.byref(m), C.byref(ml), C.byref(d),
            C.byref(dl))
    if rc != 0:
        raise RuntimeError("huffman failed")
    return _take(m, ml.value, free), _take(d, dl.value, free)


def huffman_decode(meta: bytes, data: bytes, count: int, sym_type: int) -> np.ndarray:
    out = np.empty(count, dtype=_SYM_DTYPE[sym_type])
    mb = np.frombuffer(meta, np.uint8) if meta else np.zeros(1, np.uint8)
    db = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
    rc = lib().oracle_huffman_decode(mb.ctypes.data, len(meta), db.ctypes.data, len(data),
                                     count, sym_type, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"huffman_decode: {_err()}")
    return out


def gr_unpack_u64(meta: bytes, data: bytes, max_n: int) -> np.ndarray:
    out = np.empty(max(max_n, 1), dtype=np.uint64)
    n = C.c_uint64()
    mb = np.frombuffer(meta, np.uint8) if meta else np.zeros(1, np.uint8)
    db = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
    rc = lib().oracle_gr_unpack_u64(mb.ctypes.data, len(meta), db.ctypes.data, len(data),
                                    out.ctypes.data, C.byref(n))
    if rc != 0:
        raise RuntimeError(f"gr_unpack: {_err()}")
    return out[: n.value]


def bitunpack_u64(meta: bytes, data: bytes, max_n: int) -> np.ndarray:
    out = np.empty(max(max_n, 1), dtype=np.uint64)
    n = C.c_uint64()
    mb = np.frombuffer(meta, np.uint8)
    db = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
    rc = lib().oracle_bitunpack_u64(mb.ctypes.data, len(meta), db.ctypes.data, len(data),
                                    out.ctypes.data, C.byref(n))
    if rc != 0:
        raise RuntimeError(f"bitunpack: {_err()}")
    return out[: n.value]


def ref_m2h(m: int, it: int) -> int:
    return ref().ref_morton_to_hilbert3(m, it)


def ref_h2m(h: int, it: int) -> int:
    return ref().ref_hilbert_to_morton3(h, it)


def zstd_version() -> int:
    return lib().oracle_zstd_version()
