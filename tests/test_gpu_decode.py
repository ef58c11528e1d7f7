"""The decoder on the device (xnyz_decompress_device, xnyz_decompress_to_device)
against the oracle's decoder (a restatement of block_decompressor.hpp and
block_decompressor_rle.hpp): the same points, bit for bit, in the same
order, on every mode, on the wrapped z = -1 node, on streams with Golomb-Rice
patches, long Huffman codes and Exp-Golomb counts, on the hand-written stream
images, and on a 2^22-point terrain sample; plus hilbertToMorton3 on the
device against the oracle on every 18-bit block code."""
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


def _same_points(got, want, what):
    assert got.shape == want.shape, f"{what}: {got.shape} vs {want.shape}"
    if not np.array_equal(got.view(np.uint32), want.view(np.uint32)):
        bad = np.nonzero((got.view(np.uint32) != want.view(np.uint32)).any(axis=1))[0]
        raise AssertionError(f"{what}: {len(bad)} points differ, first at {bad[0]}: "
                             f"{got[bad[0]]} vs {want[bad[0]]}")


def _check(X, pts, q, b, h, rle, thr=1024):
    ref = O.compress(pts, q, b, h, rle, thr)
    want = O.decompress_buffer(ref.pre, q, b, rle)
    got = X.decompress_device(ref.pre, q, b, rle, is_frame=False)
    _same_points(got, want, f"q={q} b={b} h={h} rle={rle} thr={thr}")
    return ref, got


def test_hilbert_to_morton_block_codes(X):
    codes = np.arange(1 << 18, dtype=np.uint64)
    got = X.hilbert_to_morton(codes, 20)
    want = np.array([O.h2m(int(c), 20) for c in codes], dtype=np.uint64)
    assert np.array_equal(got, want)


def test_hilbert_to_morton_any_width(X):
    rng = np.random.default_rng(9)
    for it in range(22):
        codes = rng.integers(0, 2**63, 2000, dtype=np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, 2000, dtype=np.uint64)
        got = X.hilbert_to_morton(codes, it)
        want = np.array([O.h2m(int(c), it) for c in codes], dtype=np.uint64)
        assert np.array_equal(got, want), it


MODES = [(q, h, rle) for q in (O.TO, O.CUBE) for h in (False, True) for rle in (False, True)]


@pytest.mark.parametrize("q,h,rle", MODES)
def test_block_mode(X, q, h, rle):
    _check(X, _uniform(1, 30000), q, 0.01, h, rle)


@pytest.mark.parametrize("q,h,rle", MODES)
def test_direct_mode(X, q, h, rle):
    ref, _ = _check(X, _uniform(2, 20000), q, 0.3, h, rle, thr=1 << 20)
    assert ref.info.is_direct


@pytest.mark.parametrize("rle", [False, True])
def test_duplicates(X, rle):
    rng = np.random.default_rng(3)
    pts = np.round(rng.random((40000, 3)) * 20, 1).astype(np.float32)
    _check(X, pts, O.TO, 0.02, True, rle)


@pytest.mark.parametrize("h", [False, True])
@pytest.mark.parametrize("span", [100, 3000, 20000])
def test_wrapped_min_z_point(X, h, span):
    d = float(np.float32(np.float64(np.float32(1.0)) / np.sqrt(5.0)))
    pts = _uniform(4, 5000, scale=span * d)
    pts[17] = [d, d, 0.0]
    pts[:, 2] += 1.0
    pts[17, 2] = 1.0
    pts[200] = [0.0, 0.0, 3.0]
    for rle in (False, True):
        for thr in (0, 1 << 20):
            _check(X, pts, O.TO, 0.5, h, rle, thr)


def test_sparse_patches_and_long_codes(X):
    # a sparse cloud: most sub-block keys are Golomb-Rice patches and the
    # Huffman alphabet is wide with skewed counts (codes longer than the
    # 11-bit lookup)
    rng = np.random.default_rng(5)
    pts = (rng.standard_normal((60000, 3)) ** 3 * 50 + 500).astype(np.float32)
    ref, _ = _check(X, pts, O.TO, 0.01, False, False)
    assert ref.info.patches > 1000


def test_exp_golomb_counts(X):
    rng = np.random.default_rng(6)
    base = (rng.random((3000, 3)) * 30).astype(np.float32)
    reps = rng.integers(1, 4000, 3000)
    pts = np.repeat(base, reps, axis=0)
    ref, _ = _check(X, pts, O.CUBE, 0.05, False, True)
    assert ref.info.counts_exp_golomb


def test_hand_written_streams(X):
    import hand_streams as H
    for name, (pts, rle, thr, stream) in H.CASES.items():
        want = O.decompress_buffer(stream, O.TO, H.B1, rle)
        got = X.decompress_device(stream, O.TO, H.B1, rle, is_frame=False)
        _same_points(got, want, name)
        assert np.array_equal(np.sort(got, axis=0), np.sort(pts, axis=0)), name


def test_frame_input_and_device_output(X):
    pts = _uniform(7, 50000)
    ref = O.compress(pts, O.TO, 0.01, True, True)
    want = O.decompress_frame(ref.frame, O.TO, 0.01, True)
    got = X.decompress_device(ref.frame, O.TO, 0.01, True)
    _same_points(got, want, "frame")
    out = torch.zeros((len(want), 3), dtype=torch.float32, device="cuda")
    n = X.decompress_to_device(ref.frame, O.TO, 0.01, True, out)
    assert n == len(want)
    _same_points(out.cpu().numpy(), want, "device output")
    from xnyzip_amd import _lib
    small = torch.zeros((len(want) - 1, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.XnyzError) as e:
        X.decompress_to_device(ref.frame, O.TO, 0.01, True, small)
    assert e.value.code == _lib.XNYZ_ENOSPC


@pytest.mark.parametrize("q,b", [(O.TO, 0.01), (O.CUBE, 0.001)])
def test_terrain_sample(X, q, b):
    from xnyzip_amd import datasets as D
    pts = D.terrain(1 << 22, seed=0)
    dev = torch.from_numpy(pts).cuda()
    _, pre, _, st = X.compress(dev, q, b, True, True, want_prezstd=True, frame_mode=2)
    want = O.decompress_buffer(pre, q, b, True)
    got = X.decompress_device(pre, q, b, True, is_frame=False)
    _same_points(got, want, f"terrain 2^22 q={q} b={b}")
    assert len(got) == len(pts)
