"""The multi-chunk container on the GPU: a learned batch of selector-workload
chunks written into a container and decoded back (on the device and on the
host) to the oracle's points for each chunk's chosen mode; and the CLI's
--batch / --unbatch round trip."""
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


def test_learned_batch_round_trip(X, tmp_path):
    from xnyzip_amd import datasets as D
    from xnyzip_amd import selector
    pts, offs, opts = D.selector_workload(8_000_000)
    mid = next(i for i, o in enumerate(offs) if o >= len(pts) // 2)
    pick = [0, 1, 2, mid, mid + 1, mid + 2]
    clouds = [np.ascontiguousarray(pts[offs[i]:offs[i + 1]]) for i in pick]
    copts = [dict(opts[i], frame_mode=(0 if k % 2 else 2)) for k, i in enumerate(pick)]
    sub = np.concatenate(clouds)
    soffs = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).tolist()
    out = np.zeros(len(sub) * 13 + (1 << 20), np.uint8)
    ro, st, modes, searched = X.compress_selected_batch_into(
        torch.from_numpy(sub).cuda(), soffs, copts, out, params=selector.load_params(),
        min_confidence=0.5)
    chunks = [dict(o, input_offset=12 * soffs[k], input_points=len(c), is_hilbert=m["is_hilbert"],
                   rle=m["rle"], gr_k=m["gr_k"], pfd_width=m["pfd_width"])
              for k, (c, o, m) in enumerate(zip(clouds, copts, modes))]
    part, man = str(tmp_path / "part_rank0000.bin"), str(tmp_path / "manifest_rank0000.jsonl")
    X.container_write(part, man, out, [int(v) for v in ro], chunks)
    want = []
    for c, o, m in zip(clouds, copts, modes):
        pre = O.compress_mode(c, o["quantizer"], o["l2_bound"], bool(m["is_hilbert"]),
                              bool(m["rle"]), 1024, m["gr_k"], m["pfd_width"])
        want.append(O.decompress_buffer(pre, o["quantizer"], o["l2_bound"], bool(m["rle"])))
    want = np.concatenate(want)
    for gpu in (True, False):
        got = X.container_decompress(part, man, gpu=gpu)
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), gpu


def test_cli_batch_round_trip(X, tmp_path):
    from xnyzip_amd import datasets as D
    cli = os.path.join(ROOT, "xnyzip_amd", "bin", "XnYZip")
    pts = D.terrain(300_000, seed=4)
    src = tmp_path / "in.bin"
    pts.tofile(src)
    outdir = tmp_path / "c"
    outdir.mkdir()
    for curve, rle in (("-h", "-rle"), ("-auto", "-auto")):
        r = subprocess.run([cli, "--batch", str(src), "to", "0.01", curve, rle, str(outdir),
                            "--chunk-points=100000"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "chunks: 3" in r.stdout
        dec = tmp_path / "out.bin"
        r = subprocess.run([cli, "--unbatch", str(outdir), str(dec)], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = np.fromfile(dec, np.float32).reshape(-1, 3)
        man = str(outdir / "manifest_rank0000.jsonl")
        want = []
        for c in X.container_read(man):
            a = c["input_offset"] // 12
            chunk = pts[a:a + c["input_points"]]
            ref = O.compress(chunk, O.TO, 0.01, bool(c["is_hilbert"]), bool(c["rle"]))
            want.append(O.decompress_buffer(ref.pre, O.TO, 0.01, bool(c["rle"])))
            if curve == "-h":
                assert c["is_hilbert"] == 1 and c["rle"] == 1
        want = np.concatenate(want)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), curve
