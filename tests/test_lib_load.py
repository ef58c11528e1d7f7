"""CPU checks of the product library: it loads, exports every entry point
include/xnyz.h declares, and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "xnyz.h")).read()
    decl = r"^(?:int|void|const char\s*\*)\s*(xnyz_[a-z_]+)\s*\("
    return sorted(set(re.findall(decl, src, flags=re.M)))


def test_header_matches_binding():
    from xnyzip_amd import _lib
    assert sorted(_lib.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from xnyzip_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libxnyz.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name
    L.xnyz_abi_version.restype = ctypes.c_int
    abi = re.search(r"#define XNYZ_ABI_VERSION (\d+)",
                    open(os.path.join(ROOT, "include", "xnyz.h")).read())
    assert L.xnyz_abi_version() == int(abi.group(1))


def test_argument_errors_without_gpu():
    from xnyzip_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libxnyz.so not built")
    L = _lib.lib()
    assert L.xnyz_create(0, None) == -1            # XNYZ_EINVAL: null out pointer
    assert L.xnyz_compress(None, None, 0, None, None, None, None, None) == -1
    assert b"null" in L.xnyz_last_error()
    L.xnyz_free(None)                               # no-op
    assert L.xnyz_host_alloc(0, None) == -1         # null out, zero size
    assert L.xnyz_host_free(None) == 0              # no-op
    scratch = ctypes.create_string_buffer(64)       # not from xnyz_host_alloc
    assert L.xnyz_host_free(ctypes.addressof(scratch)) == -1
    assert b"xnyz_host_alloc" in L.xnyz_last_error()
