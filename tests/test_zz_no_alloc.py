"""xnyz_compress_batch_device_noalloc against xnyz_compress_batch_device_into:
the same assembled streams and offsets on two calls in a row (the second
smaller), no change in the device's free memory across a call, the limits
rejected before any launch, and the context's own buffers intact after."""
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


def _batch(clouds):
    pts = torch.from_numpy(np.concatenate([c for c, _ in clouds])).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in clouds])]).tolist()
    return pts, offs, [dict(o, frame_mode=2) for _, o in clouds]


def _big_batch():
    from xnyzip_amd import datasets as D
    ter = D.terrain(1 << 21, seed=3)
    return [(ter[:1_500_000], dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=True, rle=True)),
            (ter[1_500_000:], dict(quantizer=O.CUBE, l2_bound=0.1, is_hilbert=False, rle=False)),
            (_uniform(1, 300_000), dict(quantizer=O.TO, l2_bound=0.01, rle=False)),
            (_uniform(2, 20_000), dict(quantizer=O.TO, l2_bound=0.3, is_hilbert=True, rle=True,
                                       direct_threshold=1 << 20)),
            (D.shells(400_000, seed=4, count=10), dict(quantizer=O.TO, l2_bound=0.05))]


def _small_batch():
    return [(_uniform(5, 7_000), dict(quantizer=O.CUBE, l2_bound=0.02, is_hilbert=True, rle=True)),
            (_uniform(6, 3_000, 20.0), dict(quantizer=O.TO, l2_bound=0.1, rle=False))]


def _check_same(X, ctx, clouds):
    pts, offs, opts = _batch(clouds)
    ref = np.zeros(64 << 20, np.uint8)
    ro, rst = X.compress_batch_into(pts, offs, opts, ref, ctx=ctx)
    out = np.zeros(64 << 20, np.uint8)
    no, nst = X.compress_batch_noalloc(pts, offs, opts, out, ctx=ctx)
    assert list(no) == list(ro)
    assert np.array_equal(out[:int(no[-1])], ref[:int(ro[-1])])
    for a, b in zip(nst, rst):
        assert (a.num_points, a.blocks, a.patches, a.prezstd_bytes) == \
            (b.num_points, b.blocks, b.patches, b.prezstd_bytes)
    return no, out


def test_two_calls_match_the_allocating_form(X):
    ctx = X.Context(0)
    big = _big_batch()
    no, out = _check_same(X, ctx, big)
    # spot check against the oracle: the direct-mode cloud
    c, o = big[3]
    want = O.compress(c, o["quantizer"], o["l2_bound"], True, True, 1 << 20).pre
    assert bytes(out[int(no[3]):int(no[4])]) == want
    _check_same(X, ctx, _small_batch())
    ctx.close()


def test_no_device_memory_taken(X):
    ctx = X.Context(0)
    pts, offs, opts = _batch(_big_batch())
    out = np.zeros(64 << 20, np.uint8)
    first, _ = X.compress_batch_noalloc(pts, offs, opts, out, ctx=ctx)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info(0)
    again, _ = X.compress_batch_noalloc(pts, offs, opts, out, ctx=ctx)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info(0)
    assert list(first) == list(again)
    assert free1 == free0, f"device free memory moved by {free0 - free1} bytes"
    ctx.close()


def test_limits_rejected(X):
    from xnyzip_amd import _lib
    ctx = X.Context(0)
    out = np.zeros(1 << 20, np.uint8)
    pts = torch.zeros(((1 << 22) + 1, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.XnyzError) as e:
        X.compress_batch_noalloc(pts, [0, pts.shape[0]], [dict(quantizer=1, l2_bound=0.01)],
                                 out, ctx=ctx)
    assert e.value.code == _lib.XNYZ_EUNSUPPORTED
    small = torch.from_numpy(_uniform(7, 1000)).cuda()
    with pytest.raises(_lib.XnyzError) as e:
        X.compress_batch_noalloc(small, [0, 1000], [dict(quantizer=1, l2_bound=0.01,
                                                         frame_mode=0)], out, ctx=ctx)
    assert e.value.code == _lib.XNYZ_EUNSUPPORTED
    offs = list(range(0, 258)) + [1000]
    with pytest.raises(_lib.XnyzError) as e:
        X.compress_batch_noalloc(small, offs, [dict(quantizer=1, l2_bound=0.01)] * 258, out,
                                 ctx=ctx)
    assert e.value.code == _lib.XNYZ_EUNSUPPORTED
    with pytest.raises(_lib.XnyzError) as e:  # too small an output: size needed reported
        X.compress_batch_noalloc(small, [0, 1000], [dict(quantizer=1, l2_bound=0.01)],
                                 np.zeros(16, np.uint8), ctx=ctx)
    assert e.value.code == _lib.XNYZ_ENOSPC
    ctx.close()


def test_context_buffers_intact(X):
    ctx = X.Context(0)
    pts, offs, opts = _batch(_small_batch())
    a = np.zeros(1 << 22, np.uint8)
    ra, _ = X.compress_batch_into(pts, offs, opts, a, ctx=ctx)
    X.compress_batch_noalloc(*_batch(_big_batch()), np.zeros(64 << 20, np.uint8), ctx=ctx)
    b = np.zeros(1 << 22, np.uint8)
    rb, _ = X.compress_batch_into(pts, offs, opts, b, ctx=ctx)
    assert list(ra) == list(rb) and np.array_equal(a[:int(ra[-1])], b[:int(rb[-1])])
    ctx.close()
