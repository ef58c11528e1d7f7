"""Shared codebooks (fit_profile / compress_with_profile,
block_compressor.hpp:241-431) on the GPU against the oracle's restatement
(whose build_complete and load_meta + get_meta are pinned against the
reference's own encoder in tests/test_oracle_ref.py and the golden vectors):
the fitted chunk's stream and profile, every other chunk's stream with that
profile, a small-range chunk forced into block mode, decoding, and RLE
refused as in the reference's driver."""
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


def _same(a, b, what):
    if a != b:
        n = min(len(a), len(b))
        first = next((i for i in range(n) if a[i] != b[i]), n)
        raise AssertionError(f"{what}: differs at byte {first} (gpu {len(a)} B, oracle {len(b)} B)")


@pytest.mark.parametrize("h", [False, True])
def test_fit_then_batch_with_profile(X, h):
    from xnyzip_amd import datasets as D
    pts = D.shells(400_000, seed=3, count=200)
    chunk = 100_000
    chunks = [np.ascontiguousarray(pts[i:i + chunk]) for i in range(0, len(pts), chunk)]
    # a small cloud whose range would take direct mode: forced into blocks
    chunks.append((np.random.default_rng(5).random((3000, 3)) * 5).astype(np.float32))
    dev0 = torch.from_numpy(chunks[0]).cuda()
    frame0, pre0, prof = X.fit_profile(dev0, O.TO, 0.05, h, want_prezstd=True)
    want0, wprof = O.compress_shared(chunks[0], O.TO, 0.05, h)
    _same(pre0, want0, "fitted chunk")
    assert prof == wprof
    rest = chunks[1:]
    sub = torch.from_numpy(np.concatenate(rest)).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c in rest])]).tolist()
    opts = [dict(quantizer=O.TO, l2_bound=0.05, is_hilbert=h)] * len(rest)
    frames, pres, st = X.compress_with_profile_batch(sub, offs, opts, prof, want_prezstd=True)
    for i, c in enumerate(rest):
        want, _ = O.compress_shared(c, O.TO, 0.05, h, profile=prof)
        _same(pres[i], want, f"chunk {i + 1}")
        assert not st[i].is_direct
        back = X.decompress_device(frames[i], O.TO, 0.05, False)
        assert np.array_equal(back, O.decompress_buffer(want, O.TO, 0.05, False))
    # every chunk embeds the same codebook bytes (the profile's metas as a
    # fresh encoder writes them after load_meta)
    metas = [_metas(p) for p in pres]
    assert all(m == metas[0] for m in metas)


def _metas(stream):
    """(repos meta, quads meta) of a block-mode stream."""
    import struct
    at, out = 30, []
    for k in range(6):
        dn, mn = struct.unpack_from("<QQ", stream, at)
        meta = stream[at + 16:at + 16 + mn]
        if k in (3, 5):
            out.append(meta)
        at += 16 + mn + dn
    return out


def test_profile_from_another_workload(X):
    from xnyzip_amd import datasets as D
    ter = D.terrain(200_000, seed=1)
    she = D.shells(200_000, seed=2, count=100)
    _, _, prof = X.fit_profile(torch.from_numpy(ter).cuda(), O.CUBE, 0.01)
    fr, pres, _ = X.compress_with_profile_batch(torch.from_numpy(she).cuda(), [0, len(she)],
                                                [dict(quantizer=O.CUBE, l2_bound=0.01)], prof,
                                                want_prezstd=True)
    want, _ = O.compress_shared(she, O.CUBE, 0.01, False, profile=prof)
    _same(pres[0], want, "shells with the terrain profile")


def test_rle_refused(X):
    from xnyzip_amd import _lib
    pts = torch.from_numpy((np.random.default_rng(1).random((1000, 3)) * 10).astype(np.float32))
    _, _, prof = X.fit_profile(pts.cuda(), O.TO, 0.05)
    with pytest.raises(_lib.XnyzError) as e:
        X.compress_with_profile_batch(pts.cuda(), [0, 1000],
                                      [dict(quantizer=O.TO, l2_bound=0.05, rle=True)], prof)
    assert e.value.code == _lib.XNYZ_EUNSUPPORTED


def test_caller_memory_form(X):
    """compress_with_profile_batch_into writes the allocating form's bytes
    back to back (pinned and pageable memory), and reports the size needed
    when the buffer is short."""
    from xnyzip_amd import _lib
    from xnyzip_amd import datasets as D
    pts = D.shells(300_000, seed=8, count=150)
    offs = [0, 100_000, 170_000, 300_000]
    dev = torch.from_numpy(pts).cuda()
    _, _, prof = X.fit_profile(dev[:100_000], O.TO, 0.05)
    for mode in (_lib.FRAME_NONE, _lib.FRAME_ZSTD3):
        opts = [dict(quantizer=O.TO, l2_bound=0.05, frame_mode=mode)] * 3
        frames, _, _ = X.compress_with_profile_batch(dev, offs, opts, prof)
        for buf in (X.host_buffer(sum(map(len, frames)) + 64),
                    np.zeros(sum(map(len, frames)) + 64, np.uint8)):
            res, _ = X.compress_with_profile_batch_into(dev, offs, opts, prof, buf)
            for i, f in enumerate(frames):
                assert bytes(buf[int(res[i]):int(res[i + 1])]) == f, (mode, i)
        with pytest.raises(_lib.XnyzError) as e:
            X.compress_with_profile_batch_into(dev, offs, opts, prof, np.zeros(8, np.uint8))
        assert e.value.code == _lib.XNYZ_ENOSPC
