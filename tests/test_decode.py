"""xnyz_decompress (host decoder of the product library) against the oracle's
restated decoders (block_decompressor*.hpp) on oracle-made streams.  No GPU
is needed: decoding runs on the host."""
import os

import numpy as np
import pytest

from oracle import oracle as O

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "xnyzip_amd", "lib", "libxnyz.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libxnyz.so not built")

CASES = [(q, h, rle, thr) for q in (O.TO, O.CUBE) for h in (False, True)
         for rle in (False, True) for thr in (1024, 1 << 20)]


@pytest.mark.parametrize("q,h,rle,thr", CASES)
def test_decode_matches_oracle(q, h, rle, thr):
    import xnyzip_amd as X
    rng = np.random.default_rng(21)
    pts = (rng.random((6000, 3)) * 50).astype(np.float32)
    pts[10:60] = pts[3]
    b = 0.01 if thr == 1024 else 0.2
    c = O.compress(pts, q, b, h, rle, thr)
    want = O.decompress_frame(c.frame, q, b, rle)
    got = X.decompress(c.frame, q, b, rle)
    assert np.array_equal(got, want)
    got2 = X.decompress(c.pre, q, b, rle, is_frame=False)
    assert np.array_equal(got2, want)


def test_decode_wrapped_and_eg_counts():
    import xnyzip_amd as X
    d = float(np.float32(np.float64(np.float32(1.0)) / np.sqrt(5.0)))
    pts = np.array([[0, 0, 5 * d], [d, d, 0], [3000 * d, 3000 * d, 3 * d]], np.float32)
    for rle in (False, True):
        c = O.compress(pts, O.TO, 0.5, True, rle, 0)
        assert np.array_equal(X.decompress(c.frame, O.TO, 0.5, rle),
                              O.decompress_frame(c.frame, O.TO, 0.5, rle))
    cells = (np.random.default_rng(2).random((100, 3)) * 10).astype(np.float32)
    pts = np.repeat(cells, np.arange(1, 101), axis=0)
    c = O.compress(pts, O.TO, 0.01, False, True)
    assert c.info.counts_exp_golomb == 1
    assert np.array_equal(X.decompress(c.frame, O.TO, 0.01, True),
                          O.decompress_frame(c.frame, O.TO, 0.01, True))


def test_decode_rejects_garbage():
    import xnyzip_amd as X
    with pytest.raises(X.XnyzError):
        X.decompress(b"not a zstd frame at all", 1, 0.01, False)
    c = O.compress(np.zeros((4, 3), np.float32) + np.arange(4)[:, None], O.TO, 0.01, False, False)
    with pytest.raises(X.XnyzError):
        X.decompress(c.pre[:40], 1, 0.01, False, is_frame=False)
