"""HIP path vs the CPU oracle: assembled streams and zstd frames must be
byte-identical for the same input and options (the frame with the same
libzstd, single-shot level 3).  Every call goes through the C ABI."""
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


def _check(X, pts, q, b, h, rle, thr=1024, device=True):
    ref = O.compress(pts, q, b, h, rle, thr)
    src = torch.from_numpy(np.ascontiguousarray(pts)).cuda() if device else pts
    frame, pre, _, st = X.compress(src, q, b, h, rle, thr, want_prezstd=True)
    assert st.num_points == ref.info.num_points
    assert st.is_direct == ref.info.is_direct
    assert st.blocks == ref.info.blocks
    assert st.patches == ref.info.patches
    if pre != ref.pre:
        n = min(len(pre), len(ref.pre))
        first = next((i for i in range(n) if pre[i] != ref.pre[i]), n)
        raise AssertionError(f"assembled stream differs at byte {first} "
                             f"(gpu {len(pre)} B, oracle {len(ref.pre)} B)")
    assert frame == ref.frame
    if rle:
        assert abs(st.max_l2 - ref.info.max_l2) <= 1e-12 * max(1.0, ref.info.max_l2)
    return st, ref


def _uniform(seed, n, scale=100.0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale).astype(np.float32)


MODES = [(q, h, rle) for q in (O.TO, O.CUBE) for h in (False, True) for rle in (False, True)]


@pytest.mark.parametrize("q,h,rle", MODES)
def test_block_mode(X, q, h, rle):
    pts = _uniform(1, 20000)
    st, _ = _check(X, pts, q, 0.01, h, rle)
    assert not st.is_direct


@pytest.mark.parametrize("q,h,rle", MODES)
def test_direct_mode(X, q, h, rle):
    pts = _uniform(2, 20000)
    st, _ = _check(X, pts, q, 0.3, h, rle, thr=1 << 20)
    assert st.is_direct


@pytest.mark.parametrize("q,h,rle", MODES)
def test_duplicates_and_host_input(X, q, h, rle):
    rng = np.random.default_rng(3)
    pts = np.round(rng.random((30000, 3)) * 20, 1).astype(np.float32)  # heavy duplication
    _check(X, pts, q, 0.02, h, rle, device=False)


@pytest.mark.parametrize("h", [False, True])
@pytest.mark.parametrize("span", [100, 3000, 20000])
def test_wrapped_min_z_point(X, h, span):
    d = float(np.float32(np.float64(np.float32(1.0)) / np.sqrt(5.0)))
    pts = _uniform(4, 5000, scale=span * d)
    pts[17] = [d, d, 0.0]      # min z on an odd x/y lattice node: node z = -1
    pts[:, 2] += 1.0
    pts[17, 2] = 1.0
    pts[200] = [0.0, 0.0, 3.0]
    for rle in (False, True):
        st, ref = _check(X, pts, O.TO, 0.5, h, rle, thr=0)
        assert st.wrapped_points == ref.info.wrapped_points >= 1


@pytest.mark.parametrize("rle", [False, True])
def test_wide_keys(X, rle):
    # ~2^50 blocks: (seg, id, quad, sfc) needs more than 64 bits
    pts = _uniform(5, 20000, scale=20000.0)
    st, _ = _check(X, pts, O.TO, 0.001, True, rle)
    assert not st.is_direct


def _grid(step, n, origin=0):
    g = origin + np.arange(n) * step
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


# cube bound whose divisor is exactly 1.0f: integer coordinates are the nodes
CUBE_B1 = float(np.float32(np.sqrt(3.0) / 2.0))


@pytest.mark.parametrize("rle", [False, True])
@pytest.mark.parametrize("direct", [False, True])
def test_tile_spans_beyond_the_window(X, direct, rle):
    """Tiles whose Golomb-Rice runs exceed the section writer's LDS window
    (over 66 bits per element), which are written straight to memory.  Block
    mode: a dense grid of Morton deltas 512 sets k = 9; the last 4096 cells
    are pairs one sub-block diagonal apart in blocks of their own (quotients
    of 512).  Direct mode: the last 4096 points lie 2^14 apart on each axis
    (Morton deltas near 2^42, k near 35: codes longer than 32 bits)."""
    dense = _grid(8, 160)
    if direct:
        sparse = _grid(16384, 16, origin=16384)
    else:
        bx, by, bz = np.meshgrid(np.arange(16), np.arange(16), np.arange(16, 24), indexing="ij")
        base = np.stack([bx.ravel(), by.ravel(), bz.ravel()], 1) * 128
        sparse = np.concatenate([base, base + 63])
    pts = np.concatenate([dense, sparse]).astype(np.float32)
    st, ref = _check(X, pts, O.CUBE, CUBE_B1, False, rle, thr=(1 << 20) if direct else 1024)
    assert st.is_direct == direct
    assert ref.info.patches > 3_000_000


def test_large_counts(X):
    pts = _uniform(6, 3000)
    heavy = np.repeat(pts[:3], [9000, 5000, 300], axis=0)  # counts above 4096
    _check(X, np.concatenate([pts, heavy]), O.TO, 0.01, False, True)


def test_exp_golomb_counts(X):
    # 100 cells with 100 distinct counts: the Huffman meta outweighs a quarter
    # of the data, so the counts stream switches to Exp-Golomb
    cells = _uniform(7, 100)
    pts = np.repeat(cells, np.arange(1, 101), axis=0)
    st, ref = _check(X, pts, O.TO, 0.01, True, True)
    assert st.counts_exp_golomb == ref.info.counts_exp_golomb == 1


def test_high_entropy_counts(X):
    # > 1024 distinct counts: entropy above 10 bits -> Exp-Golomb on the first build
    cells = _uniform(8, 1400)
    pts = np.repeat(cells, np.arange(1, 1401), axis=0)
    st, _ = _check(X, pts, O.CUBE, 0.01, False, True)
    assert st.counts_exp_golomb == 1


@pytest.mark.parametrize("n", [1, 2, 3, 64, 4097])
def test_small_clouds(X, n):
    pts = _uniform(9, n)
    for rle in (False, True):
        _check(X, pts, O.TO, 0.01, True, rle)
        _check(X, pts, O.CUBE, 0.01, False, rle, thr=1 << 20)


def test_signed_zero_minimum(X):
    a = _uniform(10, 1000)
    a[5] = [0.0, 1.0, 1.0]
    a[9] = [-0.0, 2.0, 2.0]
    _check(X, a, O.TO, 0.01, False, False)
    b = a.copy()
    b[5, 0], b[9, 0] = np.float32(-0.0), np.float32(0.0)
    _check(X, b, O.TO, 0.01, False, False)


def test_quantizer_parity_near_ties(X):
    # points placed on E/O decision boundaries (midplanes) of the lattice
    rng = np.random.default_rng(11)
    b = 0.01
    d = np.float32(np.float64(np.float32(2 * b)) / np.sqrt(5.0))
    k = rng.integers(0, 5000, (20000, 3)).astype(np.float64)
    frac = rng.choice([0.5, 0.25, 0.75, 1.0, 1.5], (20000, 3))
    pts = ((k + frac) * d).astype(np.float32)
    pts[0] = 0
    t = torch.from_numpy(pts).cuda()
    for q in (O.TO, O.CUBE):
        nodes, off = X.quantize(t, q, b)
        ref, roff = O.quantize(pts, q, b)
        assert np.array_equal(nodes, ref) and np.array_equal(off, roff)


def test_hilbert_parity(X):
    codes = np.arange(1 << 18, dtype=np.uint64)
    got = X.morton_to_hilbert(codes, 20)
    want = np.array([O.m2h(int(m), 20) for m in range(1 << 18)], dtype=np.uint64)
    assert np.array_equal(got, want)
    rng = np.random.default_rng(12)
    for it in range(0, 22):
        c = rng.integers(0, 2**63, 4000, dtype=np.uint64)
        assert np.array_equal(X.morton_to_hilbert(c, it),
                              np.array([O.m2h(int(m), it) for m in c], dtype=np.uint64))


def test_plain_1m_uniform(X):
    """The 1M-point plain workload: U[0,100)^3 seed 0, TO, L2 0.01, Z-order, RLE."""
    from xnyzip_amd import datasets
    pts = datasets.uniform(1_000_000, seed=0)
    st, ref = _check(X, pts, O.TO, 0.01, False, True)
    assert st.num_points == ref.info.num_points


@pytest.mark.parametrize("scale,b,thr", [(100.0, 0.01, 1024), (100.0, 0.3, 1 << 20),
                                         (20000.0, 0.001, 1024)])
def test_permutation(X, scale, b, thr):
    """The reference reorders the caller's points (ords / sfc order); the
    index payload must give the same cell at every position (the order of
    points sharing one cell is the sort's own and carries no meaning)."""
    pts = _uniform(14, 30000, scale=scale)
    pts[1000:1100] = pts[7]  # duplicates
    ref = O.compress(pts, O.TO, b, True, False, thr, want_perm=True)
    _, _, perm, _ = X.compress(torch.from_numpy(pts).cuda(), O.TO, b, True, False, thr,
                               want_perm=True)
    assert sorted(perm.tolist()) == list(range(len(pts)))
    nodes, _ = O.quantize(pts, O.TO, b)
    assert np.array_equal(nodes[perm.astype(np.int64)], nodes[ref.perm.astype(np.int64)])


def test_repeat_calls_identical(X):
    pts = torch.from_numpy(_uniform(13, 200000)).cuda()
    outs = [X.compress(pts, O.TO, 0.01, True, True, want_prezstd=True)[1] for _ in range(3)]
    assert outs[0] == outs[1] == outs[2]


@pytest.mark.parametrize("name", sorted(__import__("hand_streams").CASES))
def test_hand_written_streams(X, name):
    """The byte images written out by hand from the format (tests/hand_streams.py)."""
    import hand_streams as H
    pts, rle, thr, want = H.CASES[name]
    _, pre, _, _ = X.compress(torch.from_numpy(pts).cuda(), O.TO, H.B1, False, rle, thr,
                              want_prezstd=True)
    assert pre == want


def test_quantizer_known_answers(X):
    """Half-way, tie, q >= 2^23 and floor-below-integer points, derived by hand."""
    import hand_streams as H
    for q, b, rows in ((O.TO, H.B1, H.TO_KNOWN), (O.CUBE, H.B1_CUBE, H.CUBE_KNOWN)):
        pts = torch.tensor([p for p, _ in rows], dtype=torch.float32).cuda()
        nodes, off = X.quantize(pts, q, b)
        assert (off == 0).all()
        assert nodes.tolist() == [n for _, n in rows]
        # the compression pipeline's own quantizer on the same points
        _check(X, np.array([p for p, _ in rows], np.float32), q, b, False, False, thr=0)
