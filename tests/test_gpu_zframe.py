"""Frames built on the device (XNYZ_FRAME_ZSTD_GPU): byte-identical to the
host run of the same encoder on the oracle's assembled stream, and decoded
by libzstd back to that stream."""
import numpy as np
import pytest

from oracle import oracle as O

from zframe_ref import host_frame, unzstd

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GPU = 3


@pytest.fixture(scope="module")
def X():
    import xnyzip_amd
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    return xnyzip_amd


def _uniform(seed, n, scale=100.0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale).astype(np.float32)


@pytest.mark.parametrize("n,q,rle,hil", [(1, O.TO, False, False), (5000, O.CUBE, True, False),
                                         (300000, O.TO, True, True),
                                         (700000, O.TO, False, False)])
def test_single_cloud(X, n, q, rle, hil):
    c = _uniform(n, n)
    frame, pre, _, st = X.compress(torch.from_numpy(c).cuda(), q, 0.01, hil, rle,
                                   frame_mode=GPU, want_prezstd=True)
    ref = O.compress(c, q, 0.01, hil, rle).pre
    assert pre == ref
    assert unzstd(frame) == ref
    assert frame == host_frame(ref)
    assert st.frame_bytes == len(frame)
    back = X.decompress(frame, q, 0.01, rle)
    assert np.array_equal(back, O.decompress_buffer(ref, q, 0.01, rle))


def test_mixed_frame_modes_in_one_batch(X):
    clouds = [_uniform(10 + i, n) for i, n in enumerate([40000, 7, 120000, 3000, 90000])]
    modes = [GPU, 0, 1, 2, GPU]
    opts = [dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=bool(i % 2), rle=bool(i % 3 == 0),
                 frame_mode=m) for i, m in enumerate(modes)]
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).tolist()
    frames, pres, _ = X.compress_batch(pts, offs, opts, want_prezstd=True)
    buf = np.zeros(sum(len(f) for f in frames) + 64, np.uint8)
    res, _ = X.compress_batch_into(pts, offs, opts, buf)
    for i, (c, o) in enumerate(zip(clouds, opts)):
        ref = O.compress(c, o["quantizer"], o["l2_bound"], o["is_hilbert"], o["rle"])
        assert pres[i] == ref.pre, i
        assert bytes(buf[res[i]:res[i + 1]]) == frames[i], i
        if o["frame_mode"] == GPU:
            assert frames[i] == host_frame(ref.pre), i
        elif o["frame_mode"] == 0:
            assert frames[i] == ref.frame, i
        elif o["frame_mode"] == 2:
            assert frames[i] == ref.pre, i
        if o["frame_mode"] != 2:
            assert unzstd(frames[i]) == ref.pre, i


def test_repeat_calls(X):
    c = torch.from_numpy(_uniform(99, 250000)).cuda()
    a = X.compress(c, O.TO, 0.01, True, True, frame_mode=GPU)[0]
    for _ in range(3):
        assert X.compress(c, O.TO, 0.01, True, True, frame_mode=GPU)[0] == a
