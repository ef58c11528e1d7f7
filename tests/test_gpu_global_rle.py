"""Global RLE on one GPU (the MPI driver's mapreduce mode, mpi_driver.cpp:
365-648, with encode_cells_with_counts, block_compressor_rle.hpp:341-470)
against the oracle's restatement: streams on uniform, terrain and shell
clouds with duplicates, their decoding on the device, and the CLI's
--mapreduce part file and manifest read back by --unbatch."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def X():
    import xnyzip_amd
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    return xnyzip_amd


def _clouds():
    from xnyzip_amd import datasets as D
    rng = np.random.default_rng(8)
    dup = np.round(rng.random((200_000, 3)) * [30, 10, 3], 1).astype(np.float32)
    return [(dup, O.TO, 0.02, False), (D.terrain(1 << 21, seed=2), O.TO, 0.01, True),
            (D.shells(500_000, seed=3, count=50), O.CUBE, 0.05, False),
            ((rng.random((5000, 3)) * 2).astype(np.float32), O.TO, 0.1, True)]


@pytest.mark.parametrize("case", range(4))
def test_stream_and_decode(X, case):
    pts, q, b, h = _clouds()[case]
    frame, pre, st = X.compress_global_rle(torch.from_numpy(pts).cuda(), q, b, h,
                                           want_prezstd=True)
    want = O.compress_global_rle(pts, q, b, h)
    if pre != want:
        n = min(len(pre), len(want))
        first = next((i for i in range(n) if pre[i] != want[i]), n)
        raise AssertionError(f"differs at byte {first} (gpu {len(pre)}, oracle {len(want)})")
    assert not st.is_direct
    got = X.decompress_device(frame, q, b, True)
    assert np.array_equal(got, O.decompress_buffer(want, q, b, True))
    assert len(got) == len(pts)


def test_cli_mapreduce(X, tmp_path):
    from xnyzip_amd import datasets as D
    cli = os.path.join(ROOT, "xnyzip_amd", "bin", "XnYZip")
    pts = D.terrain(400_000, seed=6)
    src = tmp_path / "in.bin"
    pts.tofile(src)
    r = subprocess.run([cli, "--mapreduce", str(src), "to", "0.01", "-h", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    part = (tmp_path / "part_rank0000.bin").read_bytes()
    want = O.compress_global_rle(pts, O.TO, 0.01, True)
    assert O.decompress_frame(part, O.TO, 0.01, True).tobytes() == \
        O.decompress_buffer(want, O.TO, 0.01, True).tobytes()
    out = tmp_path / "out.bin"
    r = subprocess.run([cli, "--unbatch", str(tmp_path), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float32).reshape(-1, 3)
    assert np.array_equal(got, O.decompress_buffer(want, O.TO, 0.01, True))


def test_caller_memory_form(X):
    """compress_global_rle_into writes the allocating form's bytes."""
    from xnyzip_amd import _lib
    from xnyzip_amd import datasets as D
    pts = torch.from_numpy(D.terrain(500_000, seed=9)).cuda()
    for mode in (_lib.FRAME_NONE, _lib.FRAME_ZSTD3):
        fr, _, _ = X.compress_global_rle(pts, O.TO, 0.01, True, frame_mode=mode)
        buf = X.host_buffer(len(fr) + 64)
        res, st = X.compress_global_rle_into(pts, O.TO, 0.01, buf, True, frame_mode=mode)
        assert bytes(buf[int(res[0]):int(res[1])]) == fr, mode
        assert st.num_points > 0
