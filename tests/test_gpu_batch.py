"""Batched compression: many clouds of uneven size and mixed options in one
launch set.  Every cloud's assembled stream and frame must equal the oracle's
for that cloud compressed alone."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def X():
    import xnyzip_amd
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    return xnyzip_amd


def _uniform(seed, n, scale=100.0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale).astype(np.float32)


def _clouds():
    d = float(np.float32(np.float64(np.float32(1.0)) / np.sqrt(5.0)))
    wrapped = _uniform(40, 3000, scale=3000 * d)
    wrapped[:, 2] += 1.0
    wrapped[5] = [d, d, 1.0]
    wrapped[6] = [0.0, 0.0, 3.0]
    dup = np.round(_uniform(41, 20000) / 5, 1).astype(np.float32)
    cases = [
        (_uniform(1, 20000), dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=True, rle=True)),
        (_uniform(2, 1), dict(quantizer=O.CUBE, l2_bound=0.01, is_hilbert=False, rle=False)),
        (_uniform(3, 9000), dict(quantizer=O.CUBE, l2_bound=0.3, is_hilbert=True, rle=False,
                                 direct_threshold=1 << 20)),
        (dup, dict(quantizer=O.TO, l2_bound=0.02, is_hilbert=False, rle=True)),
        (wrapped, dict(quantizer=O.TO, l2_bound=0.5, is_hilbert=True, rle=False,
                       direct_threshold=0)),
        (_uniform(4, 4096), dict(quantizer=O.TO, l2_bound=0.05, is_hilbert=False, rle=True,
                                 direct_threshold=1 << 20)),
        (_uniform(5, 70000, 20000.0), dict(quantizer=O.TO, l2_bound=0.001, is_hilbert=True,
                                           rle=False)),
        (_uniform(6, 12345), dict(quantizer=O.CUBE, l2_bound=0.01, is_hilbert=False, rle=True)),
    ]
    return cases


def _check_batch(X, cases, frame_mode=0):
    pts = np.concatenate([c for c, _ in cases])
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    opts = [dict(o, frame_mode=frame_mode) for _, o in cases]
    frames, pres, stats = X.compress_batch(torch.from_numpy(pts).cuda(), offs, opts,
                                           want_prezstd=True)
    for i, (c, o) in enumerate(cases):
        ref = O.compress(c, o["quantizer"], o["l2_bound"], o["is_hilbert"], o["rle"],
                         o.get("direct_threshold", 1024))
        assert pres[i] == ref.pre, f"cloud {i}: assembled stream differs"
        if frame_mode == 0:
            assert frames[i] == ref.frame, f"cloud {i}: frame differs"
        else:
            back = O.decompress_frame(frames[i], o["quantizer"], o["l2_bound"], o["rle"])
            assert np.array_equal(back, O.decompress_buffer(ref.pre, o["quantizer"],
                                                            o["l2_bound"], o["rle"]))
        assert stats[i].input_points == len(c)
    return frames


def test_mixed_batch(X):
    _check_batch(X, _clouds())


def test_batch_all_plain_and_all_rle(X):
    base = _clouds()
    _check_batch(X, [(c, dict(o, rle=False)) for c, o in base])
    _check_batch(X, [(c, dict(o, rle=True)) for c, o in base])


def test_many_small_clouds(X):
    rng = np.random.default_rng(7)
    cases = []
    for i in range(300):
        n = int(rng.integers(1, 3000))
        cases.append((_uniform(100 + i, n, scale=float(rng.uniform(1, 200))),
                      dict(quantizer=int(i % 2), l2_bound=[0.001, 0.01, 0.1][i % 3],
                           is_hilbert=bool(i % 5 < 2), rle=bool(i % 7 < 4))))
    _check_batch(X, cases)


def test_chunked_frames_decode(X):
    """The host-threaded frame mode writes one valid zstd frame per cloud."""
    big = _uniform(9, 600000)
    frame, pre, _, _ = X.compress(torch.from_numpy(big).cuda(), O.TO, 0.01, True, True,
                                  frame_mode=1, want_prezstd=True)
    ref = O.compress(big, O.TO, 0.01, True, True)
    assert pre == ref.pre and len(pre) > 2 * (1 << 20)
    assert np.array_equal(O.decompress_frame(frame, O.TO, 0.01, True),
                          O.decompress_frame(ref.frame, O.TO, 0.01, True))
    _check_batch(X, _clouds()[:3], frame_mode=1)


def test_batch_into_caller_buffer(X):
    """The caller-memory entry point writes the same bytes as the allocating
    one, back to back, and reports the size needed when the buffer is short."""
    from xnyzip_amd import _lib
    cases = _clouds()
    pts = torch.from_numpy(np.concatenate([c for c, _ in cases])).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    for mode in (0, 1, 2):
        opts = [dict(o, frame_mode=mode) for _, o in cases]
        frames, _, _ = X.compress_batch(pts, offs, opts)
        buf = np.zeros(sum(len(f) for f in frames) + 100, dtype=np.uint8)
        res, stats = X.compress_batch_into(pts, offs, opts, buf)
        for i, f in enumerate(frames):
            assert bytes(buf[res[i]:res[i + 1]]) == f, (mode, i)
            assert stats[i].frame_bytes == len(f)
        small = np.zeros(10, dtype=np.uint8)
        with pytest.raises(_lib.XnyzError) as e:
            X.compress_batch_into(pts, offs, opts, small)
        assert e.value.code == _lib.XNYZ_ENOSPC
        assert "%d bytes needed" % res[-1] in str(e.value)


def test_pinned_streams_and_async_calls(X):
    """Assembled streams into pinned memory: a batch is packed on the device
    (segments at byte offsets of any alignment) and copied once.  The
    asynchronous form returns before the copy lands and gives the same bytes
    after ctx.wait(): calls on both of the context's stream buffers in
    flight, a synchronous call between them, one output buffer reused (the
    later call's bytes land last)."""
    from xnyzip_amd import _lib
    rng = np.random.default_rng(11)
    small = [(_uniform(200 + i, int(rng.integers(1, 2000))),
              dict(quantizer=int(i % 2), l2_bound=0.01, is_hilbert=i % 3 == 0, rle=i % 2 == 1))
             for i in range(90)]
    batches = [_clouds(), small, _clouds()[:1]]
    ctx = X.Context(0)
    ref = []
    for b in batches:
        pts = torch.from_numpy(np.concatenate([c for c, _ in b])).cuda()
        offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in b])]).tolist()
        opts = [dict(o, frame_mode=_lib.FRAME_NONE) for _, o in b]
        want = [O.compress(c, o["quantizer"], o["l2_bound"], o["is_hilbert"], o["rle"],
                           o.get("direct_threshold", 1024)).pre for c, o in b]
        ref.append((pts, offs, opts, want))

    def check(buf, res, want):
        assert len(res) == len(want) + 1
        for i, w in enumerate(want):
            assert bytes(buf[int(res[i]):int(res[i + 1])]) == w, i

    for pts, offs, opts, want in ref:  # synchronous
        buf = X.host_buffer(sum(map(len, want)) + 64)
        res, _ = X.compress_batch_into(pts, offs, opts, buf, ctx=ctx)
        check(buf, res, want)
    cap = max(sum(map(len, w)) for _, _, _, w in ref) + 64
    bufs = [X.host_buffer(cap) for _ in range(3)]
    got = []
    for k, (bi, bb) in enumerate([(0, 0), (1, 1), (2, 2), (1, None), (2, 0)]):
        pts, offs, opts, want = ref[bi]
        if bb is None:
            plain = np.zeros(cap, np.uint8)
            res, _ = X.compress_batch_into(pts, offs, opts, plain, ctx=ctx)
            check(plain, res, want)
            continue
        res, _ = X.compress_batch_into(pts, offs, opts, bufs[bb], ctx=ctx, wait=False)
        got.append((bb, res.copy(), want))
    ctx.wait()
    for bb, res, want in got[1:]:  # buffer 0's first call was overwritten
        check(bufs[bb], res, want)
    pts, offs, opts, _ = ref[0]
    with pytest.raises(_lib.XnyzError) as e:
        X.compress_batch_into(pts, offs, opts, np.zeros(cap, np.uint8), ctx=ctx, wait=False)
    assert e.value.code == _lib.XNYZ_EINVAL
    with pytest.raises(_lib.XnyzError) as e:
        X.compress_batch_into(pts, offs, [dict(o, frame_mode=_lib.FRAME_ZSTD3) for o in opts], bufs[0],
                              ctx=ctx, wait=False)
    assert e.value.code == _lib.XNYZ_EINVAL
    ctx.close()


def test_graph_replay_matches_eager(X):
    """Stage 2 replayed from a captured hipGraph gives the eager bytes on
    every replay, with other work (another launch shape, eager calls and
    synchronisations) in between."""
    ctx = X.Context(0)
    cases = _clouds()
    pts = torch.from_numpy(np.concatenate([c for c, _ in cases])).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    opts = [dict(o, frame_mode=2) for _, o in cases]
    other = torch.from_numpy(_uniform(77, 50000)).cuda()
    oopt = [dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=True, rle=True, frame_mode=2)]
    ctx.graphs(False)
    eager, _, _ = X.compress_batch(pts, offs, opts, ctx=ctx)
    eager_other, _, _ = X.compress_batch(other, [0, 50000], oopt, ctx=ctx)
    ctx.graphs(True)
    for rep in range(4):
        got, _, _ = X.compress_batch(pts, offs, opts, ctx=ctx)
        assert got == eager, rep
        torch.cuda.synchronize()
        got_o, _, _ = X.compress_batch(other, [0, 50000], oopt, ctx=ctx)  # recaptures
        assert got_o == eager_other, rep
        if rep == 1:
            ctx.graphs(False)
            assert X.compress_batch(pts, offs, opts, ctx=ctx)[0] == eager
            ctx.graphs(True)
    ctx.close()
