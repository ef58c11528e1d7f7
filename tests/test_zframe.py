"""The zstd-format block encoder (xnyzip_amd/csrc/zframe.hpp) on the host:
every frame it writes must decode with libzstd to the input, for inputs
that exercise each literal mode (raw, RLE, Huffman with FSE or direct
weights), each sequence table mode (predefined, RLE, FSE), repeat offsets,
raw blocks, partial last blocks and empty input.  The device kernels run the
same code (tests/test_gpu_zframe.py compares their frames byte for byte)."""
import numpy as np
import pytest

from oracle import oracle as O

from zframe_ref import host_frame, unzstd


def _cases():
    rng = np.random.default_rng(0)
    skew = np.r_[0.5, np.full(255, 0.5 / 255)]
    words = np.repeat(rng.integers(0, 3, 40000).astype(np.uint64), rng.integers(1, 9, 40000))
    small_alpha = rng.integers(0, 40, 70000, dtype=np.uint8)  # <= 128 weights: direct
    return {
        "empty": b"",
        "one": b"x",
        "tiny": b"abcabcabcabc",
        "zeros": bytes(200000),
        "random": rng.integers(0, 256, 150000, dtype=np.uint8).tobytes(),
        "text": b"the quick brown fox jumps over the lazy dog. " * 3000,
        "skewed": rng.choice(np.arange(256, dtype=np.uint8), 200000, p=skew).tobytes(),
        "words": words.tobytes(),
        "small_alphabet": small_alpha.tobytes(),
        "block_edge": rng.integers(0, 4, 65536 * 2 + 1, dtype=np.uint8).tobytes(),
        "periodic": (bytes(range(256)) * 700)[:170000],
    }


@pytest.mark.parametrize("name", sorted(_cases()))
def test_frames_decode(name):
    data = _cases()[name]
    f = host_frame(data)
    assert unzstd(f) == data
    if len(data) > 1000 and name != "random":
        assert len(f) < len(data)


def test_assembled_streams_decode():
    """Real assembled streams (TO and cube, RLE and not, Hilbert and Z)."""
    rng = np.random.default_rng(3)
    for i, (q, rle, hil) in enumerate([(O.TO, True, True), (O.CUBE, False, False),
                                        (O.TO, False, True)]):
        pts = (rng.random((60000, 3)) * [500, 500, 20]).astype(np.float32)
        pre = O.compress(pts, q, 0.01, hil, rle).pre
        f = host_frame(pre)
        assert unzstd(f) == pre, i
        assert len(f) < len(pre)
