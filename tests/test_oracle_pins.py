"""Pins of the oracle that do not go through the oracle's own decoder:

* quantizer known answers at the half-way and tie points (round half away
  from zero, dE <= dO, q >= 2^23, cube floor just below an integer), derived
  by hand from truncated_octahedron_quantizer.hpp:31-50 and
  cube_quantizer.hpp:20-21 (tests/hand_streams.py);
* assembled streams of 2-, 3- and 4-point clouds in block and direct mode,
  with and without RLE, written out byte by byte from the format
  (tests/hand_streams.py);
* key generation, block runs and the LCP streams (block_quantizer.hpp:84-150)
  recomputed by brute force in Python integers from the quantized nodes, and
  compared with the sections of the oracle's stream as parsed and decoded by
  the small format readers below (written from the encoders:
  bitpacking_encoder.hpp:17-51, golomb_rice_encoder.hpp:12-148,
  huffman_encoder.hpp:126-170 and :365-396)."""
import struct

import numpy as np
import pytest

import hand_streams as H
from oracle import oracle as O

M64 = 1 << 64


@pytest.mark.parametrize("name", sorted(H.CASES))
def test_hand_written_streams(oracle, name):
    pts, rle, thr, want = H.CASES[name]
    got = oracle.compress(pts, O.TO, H.B1, False, rle, thr).pre
    assert got == want


def test_bound_gives_unit_divisors():
    two_b = np.float32(2) * np.float32(H.B1)
    assert np.float32(np.float64(two_b) / np.sqrt(5.0)) == np.float32(1.0)
    two_c = np.float32(2) * np.float32(H.B1_CUBE)
    assert np.float32(np.float64(two_c) / np.sqrt(3.0)) == np.float32(1.0)


def test_to_half_way_and_tie_known_answers(oracle):
    pts = np.array([p for p, _ in H.TO_KNOWN], np.float32)
    nodes, off = oracle.quantize(pts, O.TO, H.B1)
    assert (off == 0).all()
    assert nodes.tolist() == [n for _, n in H.TO_KNOWN]


def test_cube_floor_known_answers(oracle):
    pts = np.array([p for p, _ in H.CUBE_KNOWN], np.float32)
    nodes, off = oracle.quantize(pts, O.CUBE, H.B1_CUBE)
    assert (off == 0).all()
    assert nodes.tolist() == [n for _, n in H.CUBE_KNOWN]


# --------------------------------------------------------------------------
# Format readers

class Reader:
    def __init__(self, b: bytes):
        self.b, self.p = b, 0

    def take(self, n):
        v = self.b[self.p:self.p + n]
        assert len(v) == n, "stream ends early"
        self.p += n
        return v

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def section(self):
        dn, mn = self.u64(), self.u64()
        return self.take(mn), self.take(dn)


def bits_lsb(data: bytes):
    for byte in data:
        for i in range(8):
            yield (byte >> i) & 1


def bits_msb(data: bytes):
    for byte in data:
        for i in range(7, -1, -1):
            yield (byte >> i) & 1


def bitunpack(meta, data):
    n, w = struct.unpack("<I", meta[:4])[0], meta[4]
    it = bits_lsb(data)
    return [sum(next(it) << b for b in range(w)) for _ in range(n)]


def gr_unpack(meta, data):
    if not meta:
        return []
    n, k = struct.unpack("<I", meta[:4])[0], meta[4]
    it = bits_lsb(data)
    out = []
    for _ in range(n):
        q = 0
        while next(it):
            q += 1
        out.append((q << k) | sum(next(it) << b for b in range(k)))
    return out


def huffman_decode(meta, data, count, sym_bytes, signed=False):
    if meta[:2] == b"\xe1\xc1":  # Exp-Golomb
        k = meta[4]
        it = bits_msb(data)
        out = []
        for _ in range(count):
            z = 0
            while next(it) == 0:
                z += 1
            n1 = 1
            for _ in range(z):
                n1 = (n1 << 1) | next(it)
            n = ((n1 - 1) << k) | sum(next(it) << b for b in range(k))
            out.append((n >> 1) ^ -(n & 1) if signed else n)
        return out
    k = struct.unpack(">I", meta[:4])[0]
    codes, p = {}, 4
    for _ in range(k):
        raw = meta[p:p + sym_bytes]
        sym = int.from_bytes(raw, "big" if sym_bytes == 8 else "little", signed=signed)
        ln, bits = meta[p + sym_bytes], struct.unpack(">I", meta[p + sym_bytes + 1:p + sym_bytes + 5])[0]
        codes[(ln, bits)] = sym
        p += sym_bytes + 5
    it = bits_msb(data)
    out = []
    for _ in range(count):
        ln = bits = 0
        while (ln, bits) not in codes:
            bits = (bits << 1) | next(it)
            ln += 1
            assert ln <= 32, "no code matches"
        out.append(codes[(ln, bits)])
    return out


def parse(stream: bytes, rle: bool):
    r = Reader(stream)
    n = r.u64()
    assert r.u64() == 0
    rng = (r.i32(), r.i32(), r.i32())
    curve, direct = r.take(1)[0], r.take(1)[0]
    out = dict(n=n, range=rng, curve=curve, direct=direct)
    if not direct:
        out["blk"] = bitunpack(*r.section())
        out["cnt"] = bitunpack(*r.section())
    out["patches"] = gr_unpack(*r.section())
    meta, data = r.section()
    out["symbols"] = huffman_decode(meta, data, n, 8)
    if not direct:
        assert r.section() == (b"", b"")  # quads patches: always empty
        meta, data = r.section()
        out["quads"] = huffman_decode(meta, data, n, 1)
    if rle:
        meta, data = r.section()
        out["counts"] = huffman_decode(meta, data, n, 4, signed=True)
    out["offset"] = struct.unpack("<fff", r.take(12))
    assert r.p == len(stream)
    return out


# --------------------------------------------------------------------------
# Brute force of block_quantizer.hpp:84-150 in Python integers

def _interleave(x, y, z):
    m = 0
    for i in range(21):
        m |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return m


def block_key(node, rx, ry):
    """(id, reid) of one node: sub-blocks of 64, blocks of 128 per axis."""
    nx, ny = rx // 128 + 1, ry // 128 + 1
    ux, uy, uz = (v % M64 for v in node)  # int -> size_t
    cx, dx, cy, dy, cz, dz = ux // 64, ux % 64, uy // 64, uy % 64, uz // 64, uz % 64
    ident = (cx // 2 + (cy // 2) * nx + (cz // 2) * nx * ny) % M64
    return ident, _interleave(dx, dy, dz) | (cx & 1) << 60 | (cy & 1) << 61 | (cz & 1) << 62


def lcp_bruteforce(nodes, rx, ry, rz):
    recs = sorted(block_key(n, rx, ry) for n in nodes)
    blkst, blkcnt, quads, repos = [], [], [], []
    prev_id, prev_q, prev_re = None, 0, 0
    for ident, reid in recs:
        quad, re = reid >> 60, reid & ((1 << 60) - 1)
        if ident != prev_id:
            blkst.append(ident)
            blkcnt.append(0)
            prev_id, prev_q, prev_re = ident, 0, 0
        elif quad != prev_q:
            prev_re = 0
        blkcnt[-1] += 1
        quads.append((quad - prev_q) % 256)
        repos.append((re - prev_re) % M64)
        prev_q, prev_re = quad, re
    return blkst, blkcnt, quads, repos, recs


def _deltas(v):
    return [v[0]] + [(v[i] - v[i - 1]) % M64 for i in range(1, len(v))]


def _clouds():
    rng = np.random.default_rng(11)
    d = float(np.float32(np.float64(np.float32(2) * np.float32(0.5)) / np.sqrt(5.0)))
    wrapped = (rng.random((300, 3)) * 40 * d).astype(np.float32)
    wrapped[7] = [d, d, 0.0]  # min-z point on an odd x/y node: node z = -1
    wrapped[:, 2] += 1.0
    wrapped[7, 2] = 1.0
    wrapped[8] = [0.0, 0.0, 3.0]
    return [
        ((rng.random((500, 3)) * 300).astype(np.float32), O.TO, 0.05),     # several blocks
        ((rng.random((400, 3)) * 30).astype(np.float32), O.CUBE, 0.01),    # dense, duplicates
        (np.round(rng.random((600, 3)) * 50, 0).astype(np.float32), O.TO, 0.02),
        (wrapped, O.TO, 0.5),                                              # z = -1 node
    ]


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("rle", [False, True])
def test_block_streams_match_bruteforce(oracle, case, rle):
    pts, q, b = _clouds()[case]
    nodes, _ = oracle.quantize(pts, q, b)
    nodes = [tuple(map(int, n)) for n in nodes]
    cells = sorted(set(nodes)) if rle else nodes
    rx = max(0, max(n[0] for n in cells))
    ry = max(0, max(n[1] for n in cells))
    rz = max(0, max(n[2] for n in cells))
    ref = oracle.compress(pts, q, b, False, rle, 0)  # direct_threshold 0: block mode
    s = parse(ref.pre, rle)
    assert s["n"] == len(cells) and s["range"] == (rx, ry, rz) and s["direct"] == 0
    blkst, blkcnt, quads, repos, recs = lcp_bruteforce(cells, rx, ry, rz)
    assert sum(blkcnt) == len(cells)
    assert all(a < b for a, b in zip(blkst, blkst[1:])), "block ids strictly increase"
    assert s["blk"] == _deltas(blkst)
    assert s["cnt"] == _deltas(blkcnt)
    assert s["symbols"] == [v + 1 if v <= 254 else 0 for v in repos]
    assert s["patches"] == [v for v in repos if v > 254]
    assert s["quads"] == [v + 1 for v in quads]
    if rle:
        from collections import Counter
        c = Counter(nodes)
        # counts follow the cells in (block id, sub-block key) order
        order = sorted(set(nodes), key=lambda n: block_key(n, rx, ry))
        assert s["counts"] == [c[n] for n in order]
    if case == 3:
        assert any(n[2] == -1 for n in nodes)


@pytest.mark.parametrize("name", sorted(H.CASES))
def test_hand_streams_parse(name):
    pts, rle, _, want = H.CASES[name]
    s = parse(want, rle)
    assert s["n"] == (len(set(map(tuple, pts.tolist()))) if rle else len(pts))
