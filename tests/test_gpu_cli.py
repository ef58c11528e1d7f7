"""The XnYZip CLI drop-in (source/main.cpp): same argv, chunk flags and stdout
contract; compressed size equal to the oracle's frame; decompressed file equal
to the oracle decoder's output (in the reference's reordered point order)."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xnyzip_amd", "bin", "XnYZip")


def _run(args, cwd):
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("q,curve,mode", [("to", "-h", "-rle"), ("to", "-z", "-normal"),
                                          ("cube", "-h", "-normal"), ("cube", "-z", "-rle")])
def test_cli_matches_oracle(tmp_path, q, curve, mode):
    rng = np.random.default_rng(31)
    pts = (rng.random((40000, 3)) * 100).astype(np.float32)
    src = tmp_path / "in.bin"
    pts.tofile(src)
    out = _run([str(src), q, "0.01", curve, mode, str(tmp_path / "out.bin")], tmp_path)
    qt = O.TO if q == "to" else O.CUBE
    rle = mode == "-rle"
    ref = O.compress(pts, qt, 0.01, curve == "-h", rle, want_perm=True)
    assert f"chunk input bytes: {pts.nbytes}" in out
    assert f"chunk compressed bytes: {len(ref.frame)}" in out
    ratio = np.float32(pts.shape[0]) * np.float32(12) / np.float32(len(ref.frame))
    assert f"compression ratio: {ratio:.6f}" in out
    assert re.search(r"XnYZip compression time = [0-9.e+-]+s", out)
    assert "Encoding mode: block" in out
    dec = np.fromfile(tmp_path / "out.bin", dtype=np.float32).reshape(-1, 3)
    assert np.array_equal(dec, O.decompress_frame(ref.frame, qt, 0.01, rle))
    if rle:
        assert f"Unique points: {ref.info.num_points}" in out
    else:
        # main.cpp:166-179 against the reordered input; a min-z node with
        # z = -1 does not survive the reference's own decode (wrapped block
        # id), so the expected overflow count comes from the oracle
        ref_err = np.linalg.norm(pts[ref.perm.astype(np.int64)] - dec, axis=1)
        want = int((ref_err > np.float32(0.01)).sum())
        assert f"overflow count: {want}" in out
        err = np.fromfile(tmp_path / f"errors-{q}.bin", dtype=np.float32)
        assert err.shape[0] == pts.shape[0]
        assert (err > 0.01).sum() == want and np.sort(err)[-1 - want] <= 0.01


def test_cli_chunk_flags(tmp_path):
    rng = np.random.default_rng(32)
    pts = (rng.random((9000, 3)) * 10).astype(np.float32)
    src = tmp_path / "in.bin"
    pts.tofile(src)
    out = _run([str(src), "to", "0.01", "-z", "-rle", str(tmp_path / "o.bin"),
                "--offset=12000", "--count=36000"], tmp_path)
    assert "Chunk mode: file_size=108000 offset=12000 count=36000 points=3000" in out
    ref = O.compress(pts[1000:4000], O.TO, 0.01, False, True)
    assert f"chunk compressed bytes: {len(ref.frame)}" in out
    # count 0 or past the end: the rest of the file
    out = _run([str(src), "to", "0.01", "-z", "-rle", str(tmp_path / "o.bin"),
                "--offset=96000", "--count=0"], tmp_path)
    assert "points=1000" in out
    r = subprocess.run([CLI, str(src), "to", "0.01", "-z", "-rle", "x", "--offset=7"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
