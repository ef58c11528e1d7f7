"""The measured workloads through the HIP path, against the oracle.

* terrain, 50M points (TO, L2 0.01, Hilbert, RLE -- the throughput
  configuration): the assembled stream's SHA-256 against the oracle's, and
  the device decoder's points on it against the oracle decoder's,
  committed in tests/golden/workloads.json by tests/golden/make_workloads.py
  (the oracle needs about two minutes of one core for this size); the input's
  SHA-256 is checked first so the generator is known to give the same bytes.
* terrain, the first 2^22 points of that cloud, at every other configuration
  of the terrain workload (TO and cube x L2 0.001 / 0.01 / 0.1, Hilbert, RLE):
  stream and zstd frame against the oracle run live.
* shells, 100M points in 2^20-point chunks (TO, L2 0.05, Z-order): the first
  three chunks as one batch, each stream against the oracle.
* the batched uneven workload's first 64 clouds (log-uniform sizes, terrain
  and shells, L2 {0.001, 0.01, 0.1}, TO, Z-order, RLE) as one batch; cloud
  29 (a single small shell) is in direct mode.
* the selector workload: chunks of the terrain half and of the shell half
  through compress_selected_batch with the packaged selector parameters,
  each chunk's stream against the oracle's for the mode chosen.

The oracle runs on host threads beside the GPU work (ctypes releases the
interpreter lock during the oracle's calls)."""
import hashlib
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "workloads.json")


@pytest.fixture(scope="module")
def X():
    import xnyzip_amd
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    return xnyzip_amd


@pytest.fixture(scope="module")
def pool():
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        yield ex


@pytest.fixture(scope="module")
def terrain50m():
    from xnyzip_amd import datasets as D
    return D.terrain(50_000_000, seed=0)


def _sha(b) -> str:
    return hashlib.sha256(b).hexdigest()


def _first_diff(a: bytes, b: bytes) -> int:
    n = min(len(a), len(b))
    x = np.frombuffer(a[:n], np.uint8) != np.frombuffer(b[:n], np.uint8)
    return int(np.argmax(x)) if x.any() else n


def _same(gpu: bytes, ref: bytes, what: str) -> None:
    if gpu != ref:
        raise AssertionError(f"{what}: assembled stream differs at byte {_first_diff(gpu, ref)} "
                             f"(gpu {len(gpu)} B, oracle {len(ref)} B)")


def test_terrain_50m_stream(X, terrain50m):
    g = json.load(open(GOLDEN))["terrain-50M"]
    pts = terrain50m
    assert pts.shape[0] == g["points"]
    assert _sha(pts.tobytes()) == g["input_sha256"], "generator output changed"
    dev = torch.from_numpy(pts).cuda()
    _, pre, _, st = X.compress(dev, O.TO, 0.01, True, True, want_prezstd=True, frame_mode=2)
    assert len(pre) == g["stream_bytes"]
    assert st.num_points == g["num_points"]
    assert _sha(pre) == g["stream_sha256"], \
        f"50M terrain stream differs from the oracle's ({len(pre)} B)"
    # the device decoder on that stream: the oracle decoder's points
    del dev
    out = torch.empty((g["points"], 3), dtype=torch.float32, device="cuda")
    m = X.decompress_to_device(pre, O.TO, 0.01, True, out, is_frame=False)
    assert m == g["decoded_points"]
    assert _sha(out[:m].cpu().numpy().astype("<f4").tobytes()) == g["decoded_sha256"], \
        "50M terrain decode differs from the oracle decoder's points"


C2 = [(q, b) for q in (O.TO, O.CUBE) for b in (0.001, 0.01, 0.1)]


def test_terrain_prefix_all_configs(X, terrain50m, pool):
    pts = np.ascontiguousarray(terrain50m[:1 << 22])
    refs = {c: pool.submit(O.compress, pts, c[0], c[1], True, True) for c in C2}
    dev = torch.from_numpy(pts).cuda()
    for q, b in C2:
        frame, pre, _, st = X.compress(dev, q, b, True, True, want_prezstd=True)
        ref = refs[(q, b)].result()
        _same(pre, ref.pre, f"terrain 2^22 q={q} b={b}")
        assert frame == ref.frame, f"terrain 2^22 q={q} b={b}: frame"
        assert st.num_points == ref.info.num_points and st.blocks == ref.info.blocks
        assert st.wrapped_points == ref.info.wrapped_points
        assert abs(st.max_l2 - ref.info.max_l2) <= 1e-12 * max(1.0, ref.info.max_l2)


def test_shell_chunks(X, pool):
    from xnyzip_amd import datasets as D
    pts = D.shells(100_000_000, seed=0)
    chunk = 1 << 20
    sub = np.ascontiguousarray(pts[:3 * chunk])
    del pts
    offs = [0, chunk, 2 * chunk, 3 * chunk]
    refs = [pool.submit(O.compress, sub[a:b], O.TO, 0.05, False, False)
            for a, b in zip(offs, offs[1:])]
    opts = [dict(quantizer=O.TO, l2_bound=0.05, is_hilbert=False, rle=False)] * 3
    frames, pres, st = X.compress_batch(torch.from_numpy(sub).cuda(), offs, opts,
                                        want_prezstd=True)
    for i, r in enumerate(refs):
        ref = r.result()
        _same(pres[i], ref.pre, f"shell chunk {i}")
        assert frames[i] == ref.frame, f"shell chunk {i}: frame"
        assert not st[i].is_direct


def test_uneven_clouds(X, pool):
    from xnyzip_amd import datasets as D
    clouds = list(pool.map(D.uneven_cloud, range(64)))
    opts = [dict(D.uneven_opts(i), is_hilbert=False, rle=True) for i in range(64)]
    refs = [pool.submit(O.compress, c, o["quantizer"], o["l2_bound"], False, True)
            for c, o in zip(clouds, opts)]
    pts = np.concatenate(clouds)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).tolist()
    frames, pres, st = X.compress_batch(torch.from_numpy(pts).cuda(), offs, opts,
                                        want_prezstd=True)
    direct = 0
    for i, r in enumerate(refs):
        ref = r.result()
        _same(pres[i], ref.pre, f"uneven cloud {i} ({len(clouds[i])} points)")
        assert frames[i] == ref.frame, f"uneven cloud {i}: frame"
        assert st[i].is_direct == ref.info.is_direct
        direct += st[i].is_direct
    assert direct >= 1, "no cloud of the batch took direct mode"


def test_selector_chunks(X, pool):
    from xnyzip_amd import datasets as D
    from xnyzip_amd import selector
    pts, offs, opts = D.selector_workload(64_000_000)
    mid = next(i for i, o in enumerate(offs) if o >= len(pts) // 2)
    pick = list(range(0, 8)) + list(range(mid, mid + 8))
    clouds = [np.ascontiguousarray(pts[offs[i]:offs[i + 1]]) for i in pick]
    copts = [dict(opts[i], frame_mode=2) for i in pick]
    del pts
    sub = np.concatenate(clouds)
    soffs = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).tolist()
    dev = torch.from_numpy(sub).cuda()
    res = X.search_batch(dev, soffs, [dict(o) for o in copts])
    _, pres, st, modes, searched = X.compress_selected_batch(
        dev, soffs, copts, params=selector.load_params(), min_confidence=0.5,
        want_prezstd=True)
    refs = [pool.submit(O.compress_mode, c, o["quantizer"], o["l2_bound"],
                        bool(m["is_hilbert"]), bool(m["rle"]), 1024, m["gr_k"], m["pfd_width"])
            for c, o, m in zip(clouds, copts, modes)]
    total = best = 0
    for i, r in enumerate(refs):
        _same(pres[i], r.result(), f"selector chunk {pick[i]} mode {modes[i]}")
        var = 2 * modes[i]["is_hilbert"] + modes[i]["rle"]
        assert len(pres[i]) == res[i]["variants"][var]["bytes"]
        total += len(pres[i])
        best += res[i]["best_bytes"]
    print(f"selector chunks: {total} B vs exhaustive optimum {best} B "
          f"({100.0 * (total / best - 1):+.4f}%), fell back {int(np.sum(searched & 1))}")
    assert total <= 1.005 * best
