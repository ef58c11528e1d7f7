"""Exhaustive per-chunk mode search and the mode overrides.

The search's sizes come from a model (Golomb-Rice and PFD costs out of the
device histograms); every size it reports is checked here against the
oracle's real stream for that mode, and the optimum against a brute force
over the whole (k, width) grid on a small cloud."""
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


def _shell(seed, n, r=30.0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return (v * (r + rng.normal(0, 1.0, (n, 1))) + 100.0).astype(np.float32)


def _cases():
    dup = np.round(_uniform(21, 15000) / 4, 1).astype(np.float32)
    return [
        (_uniform(1, 30000), dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=False, rle=True)),
        (_uniform(2, 8000, 2000.0), dict(quantizer=O.TO, l2_bound=0.05, is_hilbert=True,
                                         rle=False)),
        (_shell(3, 20000), dict(quantizer=O.TO, l2_bound=0.05, is_hilbert=False, rle=True)),
        (dup, dict(quantizer=O.CUBE, l2_bound=0.02, is_hilbert=True, rle=True)),
        (_uniform(4, 5000, 20.0), dict(quantizer=O.TO, l2_bound=0.1, is_hilbert=False,
                                       rle=False)),  # direct mode
        (_uniform(5, 1), dict(quantizer=O.CUBE, l2_bound=0.01, is_hilbert=False, rle=False)),
    ]


def _oracle_len(c, o, m):
    return len(O.compress_mode(c, o["quantizer"], o["l2_bound"], bool(m["is_hilbert"]),
                               bool(m["rle"]), o.get("direct_threshold", 1024),
                               gr_k=m["gr_k"], pfd_width=m["pfd_width"]))


@pytest.fixture(scope="module")
def searched(X):
    cases = _cases()
    pts = np.concatenate([c for c, _ in cases])
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    res = X.search_batch(torch.from_numpy(pts).cuda(), offs, [o for _, o in cases])
    return cases, res


def test_reference_bytes(searched):
    cases, res = searched
    for i, ((c, o), r) in enumerate(zip(cases, res)):
        ref = O.compress(c, o["quantizer"], o["l2_bound"], o["is_hilbert"], o["rle"])
        assert r["reference_bytes"] == len(ref.pre), f"cloud {i}"


def test_every_variant_size_is_real(searched):
    cases, res = searched
    for i, ((c, o), r) in enumerate(zip(cases, res)):
        for v in r["variants"]:
            assert v["bytes"] == _oracle_len(c, o, v), f"cloud {i} variant {v}"
        assert r["best_bytes"] == min(v["bytes"] for v in r["variants"])
        assert r["best_bytes"] <= r["reference_bytes"]


def test_optimum_is_exhaustive():
    """Brute force over k 0..31 and width 1..8 with the oracle's real streams."""
    import xnyzip_amd as X
    c = _uniform(11, 3000, 100.0)
    o = dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=False, rle=False)
    r = X.search_batch(torch.from_numpy(c).cuda(), [0, len(c)], [o])[0]
    v = r["variants"][0]
    sizes = {}
    for w in range(1, 9):
        for k in range(32):
            sizes[(k, w)] = len(O.compress_mode(c, O.TO, 0.01, False, False, gr_k=k,
                                                pfd_width=w))
    assert v["bytes"] == min(sizes.values())
    assert sizes[(v["gr_k"], v["pfd_width"])] == v["bytes"]


def test_override_streams_match_oracle(X, searched):
    """The device path honours the overrides byte for byte, and the stream
    decodes to the same points as the reference mode's."""
    cases, res = searched
    for i, ((c, o), r) in enumerate(zip(cases, res)):
        m = r["best"]
        _, pre, _, _ = X.compress(torch.from_numpy(c).cuda(), o["quantizer"], o["l2_bound"],
                                  bool(m["is_hilbert"]), bool(m["rle"]), frame_mode=2,
                                  want_prezstd=True, gr_k=m["gr_k"], pfd_width=m["pfd_width"])
        want = O.compress_mode(c, o["quantizer"], o["l2_bound"], bool(m["is_hilbert"]),
                               bool(m["rle"]), gr_k=m["gr_k"], pfd_width=m["pfd_width"])
        assert pre == want, f"cloud {i}"
        plain = O.compress(c, o["quantizer"], o["l2_bound"], bool(m["is_hilbert"]),
                           bool(m["rle"])).pre
        a = X.decompress(pre, o["quantizer"], o["l2_bound"], bool(m["rle"]), is_frame=False)
        b = O.decompress_buffer(plain, o["quantizer"], o["l2_bound"], bool(m["rle"]))
        assert np.array_equal(a, b), f"cloud {i}"


def test_batch_with_overrides(X):
    cases = _cases()[:4]
    pts = np.concatenate([c for c, _ in cases])
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    modes = [(12, 5), (None, 2), (14, None), (17, 8)]
    opts = [dict(o, gr_k=k, pfd_width=w, frame_mode=2) for (_, o), (k, w) in zip(cases, modes)]
    _, pres, _ = X.compress_batch(torch.from_numpy(pts).cuda(), offs, opts, want_prezstd=True)
    for i, ((c, o), (k, w)) in enumerate(zip(cases, modes)):
        want = O.compress_mode(c, o["quantizer"], o["l2_bound"], o["is_hilbert"], o["rle"],
                               gr_k=-1 if k is None else k, pfd_width=w or 0)
        assert pres[i] == want, f"cloud {i}"
