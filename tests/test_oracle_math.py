"""Pin the oracle where no reference output exists: closed forms, textbook
routines, brute force on small inputs, and round trips through the restated
decoders (block_decompressor*.hpp)."""
import heapq
from collections import Counter

import numpy as np
import pytest

from oracle import oracle as O


# --------------------------------------------------------------------------
# Quantizers (truncated_octahedron_quantizer.hpp:21-55, cube_quantizer.hpp:15-27)

def _to_div(b):
    return np.float32(np.float64(np.float32(2) * np.float32(b)) / np.sqrt(5.0))


def _expected_to_nodes(q):
    """Nearest point of the BCC lattice 2Z^3 u (2Z^3+1) to q, by exhaustive
    search over the 8 even and 8 odd corners around q, in float64."""
    q64 = q.astype(np.float64)
    best = np.full(q.shape[0], np.inf)
    node = np.zeros(q.shape, dtype=np.int64)
    second = np.full(q.shape[0], np.inf)
    for parity in (0, 1):
        base = 2 * np.floor((q64 - parity) / 2) + parity
        for dx in (0, 2):
            for dy in (0, 2):
                for dz in (0, 2):
                    cand = base + np.array([dx, dy, dz])
                    d = ((q64 - cand) ** 2).sum(axis=1)
                    better = d < best
                    second = np.where(better, best, np.minimum(second, d))
                    best = np.where(better, d, best)
                    node[better] = cand[better]
    odd = (node[:, 0] & 1) == 1
    out = node.copy()
    out[odd, 0] = (node[odd, 0] + 1) >> 1
    out[odd, 1] = (node[odd, 1] + 1) >> 1
    out[~odd, 0] = node[~odd, 0] >> 1
    out[~odd, 1] = node[~odd, 1] >> 1
    return out, second - best


@pytest.mark.parametrize("b", [0.01, 0.05, 0.3])
def test_to_quantizer_is_nearest_bcc_node(oracle, b):
    rng = np.random.default_rng(1)
    pts = (rng.random((20000, 3)) * 100).astype(np.float32)
    pts[0] = 0  # min at the origin: shifting is then the identity
    nodes, off = oracle.quantize(pts, O.TO, b)
    assert (off == 0).all()
    q = pts / _to_div(b)  # IEEE float32 division, as the reference
    exp, gap = _expected_to_nodes(q)
    clear = gap > 1e-3
    assert clear.mean() > 0.99
    assert (nodes[clear] == exp[clear]).all()


def test_to_quantizer_known_answers(oracle):
    d = float(_to_div(0.5))
    # (origin, then q = (0.4, 0.4, 0.4)): E=(0,0,0) beats O=(1,1,1)
    # (q = (1, 1, 0.1)): O=(1,1,1) beats E; x odd -> x=(1+1)>>1=1
    # (q = (3.2, 2.8, 5)): O=(3,3,5) -> (2, 2, 5)
    pts = np.array([[0, 0, 0], [0.4, 0.4, 0.4], [1, 1, 0.1], [3.2, 2.8, 5]],
                   dtype=np.float64) * d
    nodes, _ = oracle.quantize(pts.astype(np.float32), O.TO, 0.5)
    assert nodes.tolist() == [[0, 0, 0], [0, 0, 0], [1, 1, 1], [2, 2, 5]]


def test_cube_quantizer_is_floor(oracle):
    rng = np.random.default_rng(2)
    pts = (rng.random((5000, 3)) * 50).astype(np.float32)
    pts[0] = 0
    nodes, _ = oracle.quantize(pts, O.CUBE, 0.02)
    d = np.float32(np.float64(np.float32(0.04)) / np.sqrt(3.0))
    assert (nodes == np.floor(pts / d).astype(np.int32)).all()


def test_shift_first_occurrence_of_signed_zero(oracle):
    # shifting.hpp:19-21: '<' keeps the first of equal minima (+0 vs -0)
    a = np.array([[0.0, 1, 1], [-0.0, 2, 2]], dtype=np.float32)
    _, off = oracle.shift(a)
    assert not np.signbit(off[0])
    _, off = oracle.shift(a[::-1].copy())
    assert np.signbit(off[0])


def test_recover_inverts_quantize_within_bound(oracle):
    rng = np.random.default_rng(3)
    pts = (rng.random((4000, 3)) * 10).astype(np.float32)
    for qt, b in ((O.TO, 0.01), (O.CUBE, 0.01)):
        nodes, off = oracle.quantize(pts, qt, b)
        rec = oracle.recover(nodes, qt, b) + off
        ok = nodes[:, 2] != -1
        err = np.linalg.norm(rec[ok].astype(np.float64) - pts[ok], axis=1)
        assert err.max() <= b * 1.0001


# --------------------------------------------------------------------------
# Morton (z_order_curve.hpp:14-47) -- textbook bit interleave

def _interleave(x, y, z):
    m = 0
    for i in range(21):
        m |= ((x >> i) & 1) << (3 * i)
        m |= ((y >> i) & 1) << (3 * i + 1)
        m |= ((z >> i) & 1) << (3 * i + 2)
    return m


def test_morton_textbook_and_round_trip(oracle):
    rng = np.random.default_rng(4)
    for x, y, z in rng.integers(0, 1 << 21, (300, 3)).tolist():
        m = oracle.morton_code(x, y, z)
        assert m == _interleave(x, y, z)
        assert oracle.decode_morton(m) == (x, y, z)
    # an int of -1 is taken as uint32 and masked to 21 ones
    assert oracle.morton_code(0, 0, -1) == _interleave(0, 0, (1 << 21) - 1)


# --------------------------------------------------------------------------
# Hilbert (hilbert.cpp)

# hilbert.cpp:374-393 hard-codes tables for builds that define
# DONT_CONSTRUCT_PERMUTATIONS.  The reference does not define it, so the
# constructed tables (hilbert.cpp:138-341) are the ones in use; they differ
# from the hard-coded block, which is therefore dead data.  The constructed
# tables are pinned by the golden transform vectors (test_oracle_golden.py);
# here their structure is checked.
HARD_H2M = [[0, 4, 1, 5, 2, 6, 3, 7], [0, 2, 4, 6, 1, 3, 5, 7], [3, 7, 2, 6, 1, 5, 0, 4],
            [6, 4, 2, 0, 7, 5, 3, 1], [3, 1, 7, 5, 2, 0, 6, 4], [0, 1, 2, 3, 4, 5, 6, 7],
            [5, 4, 7, 6, 1, 0, 3, 2], [6, 4, 2, 0, 7, 5, 3, 1]]


def _rotation_perms():
    """The 24 proper rotations of the cube as permutations of the octant
    index (x<<2 | y<<1 | z), by brute force over signed axis permutations."""
    import itertools
    out = set()
    for axes in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), int)
            for r, (a, s) in enumerate(zip(axes, signs)):
                m[r, a] = s
            if round(np.linalg.det(m)) != 1:
                continue
            p = []
            for i in range(8):
                v = np.array([(i >> 2) & 1, (i >> 1) & 1, i & 1]) * 2 - 1
                w = (m @ v + 1) // 2
                p.append(int(w[0] << 2 | w[1] << 1 | w[2]))
            out.add(tuple(p))
    return out


def test_hilbert_tables_structure(oracle):
    m2h_perm, h2m_perm, m2h_perms, h2m_perms = oracle.hilbert_tables()
    assert h2m_perm.tolist() == [0, 1, 3, 2, 6, 7, 5, 4]  # pattern A, a Gray code
    assert [int(h2m_perm[m2h_perm[i]]) for i in range(8)] == list(range(8))
    rots = _rotation_perms()
    assert len(rots) == 24
    for i in range(8):
        assert tuple(h2m_perms[i].tolist()) in rots
        assert [int(h2m_perms[i][m2h_perms[i][j]]) for j in range(8)] == list(range(8))
    assert h2m_perms.tolist() != HARD_H2M
    # the block-mode shortcut relies on zero digits staying zero
    assert m2h_perm[0] == 0 and m2h_perms[0][0] == 0


def test_hilbert_round_trip_and_adjacency(oracle):
    rng = np.random.default_rng(5)
    for it in (1, 5, 20):
        for m in rng.integers(0, 1 << (3 * it + 3) if it < 20 else 1 << 63, 200,
                              dtype=np.uint64).tolist():
            assert oracle.h2m(oracle.m2h(m, it), it) == m
    # a curve of iteration 2 covers 8x8x8 cells; consecutive indices must be
    # face neighbours and every cell must be visited once
    cells = [oracle.decode_morton(oracle.h2m(h, 2)) for h in range(512)]
    assert len(set(cells)) == 512
    for a, b in zip(cells, cells[1:]):
        assert sum(abs(i - j) for i, j in zip(a, b)) == 1


# --------------------------------------------------------------------------
# Encoders: hand-computed known answers

def test_bitpack_known_answer(oracle):
    # width = bitlen(3) = 2; values LSB-first: 1 -> 1,0  2 -> 0,1  3 -> 1,1
    m, d = oracle.bitpack_u64([1, 2, 3])
    assert m == bytes([3, 0, 0, 0, 2]) and d == bytes([0b111001])
    m, d = oracle.bitpack_u64([0, 0])
    assert m == bytes([2, 0, 0, 0, 1]) and d == bytes([0])
    with pytest.raises(ValueError):
        oracle.bitpack_u64([])


def test_golomb_rice_known_answer(oracle):
    # mean 5 -> k = floor(log2 5) = 2; 5 = 1*4 + 1: '1','0', then 1 LSB-first: '1','0'
    m, d = oracle.gr_pack([5])
    assert m == bytes([1, 0, 0, 0, 2]) and d == bytes([0b0101])
    assert oracle.gr_pack([]) == (b"", b"")


def test_exp_golomb_known_answer(oracle):
    # select k over 0..16: values [0, 0, 0, 1] zigzag -> [0, 0, 0, 2]; k=0 costs
    # 1+1+1+3 = 6, k=1 costs 2+2+2+4 = 10: k = 0.  Codes: '1','1','1','011'
    m, d = oracle.huffman(np.array([0, 0, 0, 1], np.int32), O.SYM_INT, mode=2)
    assert m == bytes([0xE1, 0xC1, 0x01, 0x01, 0x00])
    assert d == bytes([0b11101100])


def _textbook_huffman_cost(freqs):
    if len(freqs) == 1:
        return sum(freqs)
    h = list(freqs)
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a, b = heapq.heappop(h), heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


def test_huffman_cost_is_optimal(oracle):
    rng = np.random.default_rng(6)
    for _ in range(20):
        v = np.minimum(rng.geometric(rng.uniform(0.02, 0.6), 3000), 255).astype(np.uint64)
        m, d = oracle.huffman(v, O.SYM_U64)
        k = int.from_bytes(m[:4], "big")
        lens = {}
        for i in range(k):
            e = m[4 + 13 * i: 4 + 13 * (i + 1)]
            lens[int.from_bytes(e[:8], "big")] = e[8]
        cnt = Counter(v.tolist())
        bits = sum(cnt[s] * lens[s] for s in cnt)
        assert bits == _textbook_huffman_cost(list(cnt.values()))
        assert len(d) == (bits + 7) // 8


# --------------------------------------------------------------------------
# End to end: compress -> restated decoder

def _cloud(seed, n, scale=100.0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale).astype(np.float32)


CONFIGS = [(qt, h, rle, thr) for qt in (O.TO, O.CUBE) for h in (False, True)
           for rle in (False, True) for thr in (1024, 1 << 20)]


@pytest.mark.parametrize("qt,h,rle,thr", CONFIGS)
def test_round_trip(oracle, qt, h, rle, thr):
    b = 0.01 if thr == 1024 else 0.3
    pts = _cloud(11, 3000)
    pts[100:140] = pts[100]  # duplicates
    c = oracle.compress(pts, qt, b, h, rle, thr, want_perm=True)
    assert c.info.is_direct == (thr > 1024)
    assert oracle.decompress_buffer(c.pre, qt, b, rle).shape[0] == pts.shape[0]
    dec = oracle.decompress_frame(c.frame, qt, b, rle)
    nodes, off = oracle.quantize(pts, qt, b)
    ok = nodes[:, 2] != -1
    if rle:
        exp = Counter(map(tuple, (oracle.recover(nodes[ok], qt, b) + off).tolist()))
        got = Counter(map(tuple, dec.tolist()))
        assert all(got[k] >= v for k, v in exp.items())
        assert sum(got.values()) == pts.shape[0]
    else:
        perm = c.perm.astype(np.int64)
        assert sorted(perm.tolist()) == list(range(pts.shape[0]))
        err = np.linalg.norm(dec.astype(np.float64) - pts[perm], axis=1)
        assert err[ok[perm]].max() <= b * 1.0001


def _wrapped_cloud(span):
    # the min-z point sits on an odd x/y lattice coordinate, so the odd
    # sub-lattice wins and its node z is -1 (the "1-point overflow" case)
    d = float(_to_div(0.5))
    pts = np.array([[0, 0, 5 * d], [d, d, 0], [span * d, span * d, 3 * d]], np.float32)
    return pts


def _wrapped_decode(node, rng_xyz):
    """Closed form of block_quantizer.hpp:99-105 (encode, size_t arithmetic)
    followed by :172-201 (decode, then the int cast) for one node, with
    Python integers reduced mod 2^64 as size_t is."""
    M = 1 << 64
    nx, ny = rng_xyz[0] // 128 + 1, rng_xyz[1] // 128 + 1
    c = [(v % M) // 64 for v in node]
    d = [(v % M) % 64 for v in node]
    id_ = (c[0] // 2 + (c[1] // 2) * nx + (c[2] // 2) * nx * ny) % M
    quad = (c[0] & 1) | (c[1] & 1) << 1 | (c[2] & 1) << 2
    pb = [(id_ % nx * 2) * 64, (id_ // nx % ny * 2) * 64, (id_ // nx // ny * 2) * 64]
    out = []
    for a in range(3):
        v = (pb[a] + ((quad >> a) & 1) * 64 + d[a]) % M
        out.append(v - M if v >= 1 << 63 else v)
    return [((v + (1 << 31)) % (1 << 32)) - (1 << 31) for v in out]


@pytest.mark.parametrize("span", [100, 3000, 4000, 20000])
def test_wrapped_min_z_point(oracle, span):
    pts = _wrapped_cloud(span)
    nodes, _ = oracle.quantize(pts, O.TO, 0.5)
    assert nodes[1].tolist()[2] == -1
    c = oracle.compress(pts, O.TO, 0.5, False, False, 0, want_perm=True)
    assert c.info.wrapped_points == 1 and not c.info.is_direct
    dec = oracle.decompress_frame(c.frame, O.TO, 0.5, False)
    perm = c.perm.astype(np.int64).tolist()
    got = dec[perm.index(1)]
    want_node = _wrapped_decode(nodes[1].tolist(), list(c.info.range))
    want = oracle.recover(np.array([want_node], np.int32), O.TO, 0.5)[0]
    assert np.array_equal(got, want)
    # the wrapped node survives when nx*ny is 1 or a power of two, as here
    # for spans 100 and 4000 (nx*ny = 1, 256), and not for 3000 (144)
    assert (want_node == nodes[1].tolist()) == (span in (100, 4000))


def test_tiny_clouds_and_errors(oracle):
    for n in (1, 2, 3):
        pts = _cloud(n, n)
        for rle in (False, True):
            c = oracle.compress(pts, O.TO, 0.01, True, rle, 1024)
            assert oracle.decompress_frame(c.frame, O.TO, 0.01, rle).shape == (n, 3)
    with pytest.raises(RuntimeError):
        oracle.compress(np.zeros((0, 3), np.float32), O.TO, 0.01, False, False)


def test_frame_decodes_to_buffer(oracle):
    import ctypes
    pts = _cloud(12, 5000)
    c = oracle.compress(pts, O.TO, 0.01, True, True, 1024)
    lib = ctypes.CDLL("/opt/conda/lib/libzstd.so")
    lib.ZSTD_decompress.restype = ctypes.c_size_t
    out = ctypes.create_string_buffer(len(c.pre))
    r = lib.ZSTD_decompress(out, ctypes.c_size_t(len(c.pre)), c.frame, ctypes.c_size_t(len(c.frame)))
    assert r == len(c.pre) and out.raw == c.pre


def test_global_rle_matches_rle_compress_on_a_cube(oracle):
    """The mapreduce stream (mpi_driver.cpp:365-648 on one rank) is the RLE
    stream in block mode with every axis given the largest range: on a cloud
    whose three ranges agree, it is the plain RLE compress() stream."""
    rng = np.random.default_rng(12)
    pts = np.round(rng.random((20000, 3)) * 40, 1).astype(np.float32)
    pts[0] = [0, 0, 0]
    pts[1] = [40, 40, 40]
    nodes, _ = oracle.quantize(pts, O.CUBE, 0.05)
    assert len(set(nodes.max(axis=0).tolist())) == 1
    assert oracle.compress_global_rle(pts, O.CUBE, 0.05, True) == \
        oracle.compress(pts, O.CUBE, 0.05, True, True, 0).pre


def test_global_rle_counts_every_point(oracle):
    rng = np.random.default_rng(13)
    pts = np.round(rng.random((30000, 3)) * [50, 20, 5], 0).astype(np.float32)
    pre = oracle.compress_global_rle(pts, O.TO, 0.3, False)
    import struct
    n = struct.unpack_from("<Q", pre)[0]
    rx, ry, rz = struct.unpack_from("<iii", pre, 16)
    assert rx == ry == rz and pre[29] == 0  # one range, block mode
    nodes, _ = oracle.quantize(pts, O.TO, 0.3)
    assert n == len({tuple(x) for x in nodes.tolist()})
    assert rx == max(0, int(nodes.max()))
    back = oracle.decompress_buffer(pre, O.TO, 0.3, True)
    assert len(back) == len(pts)
