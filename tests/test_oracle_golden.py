"""Pin the oracle's encoders and Hilbert transform on golden vectors.

tests/golden/encoders.json holds outputs of the reference's own code
(source/encoder/*.hpp and hilbert-curves-cpp/hilbert.cpp compiled as-is into
oracle/_ref by tests/golden/make_golden.py).  These checks need no reference
checkout.
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "encoders.json")


@pytest.fixture(scope="module")
def fx():
    with open(GOLDEN) as f:
        return json.load(f)


def test_bitpack_golden(oracle, fx):
    assert fx["bitpack_u64"]
    for c in fx["bitpack_u64"]:
        m, d = oracle.bitpack_u64(c["values"])
        assert (m.hex(), d.hex()) == (c["meta"], c["data"])
        back = oracle.bitunpack_u64(m, d, len(c["values"]))
        assert back.tolist() == c["values"]


def test_golomb_rice_golden(oracle, fx):
    for c in fx["gr_u64"]:
        m, d = oracle.gr_pack(c["values"], np.uint64)
        assert (m.hex(), d.hex()) == (c["meta"], c["data"])
        if c["values"]:
            assert oracle.gr_unpack_u64(m, d, len(c["values"])).tolist() == c["values"]
    for c in fx["gr_u8"]:
        m, d = oracle.gr_pack(c["values"], np.uint8)
        assert (m.hex(), d.hex()) == (c["meta"], c["data"])


def test_pfd_golden(fx):
    # p_for_delta_encoder.hpp:27-43: v > 254 -> patch v and delta 0, else v+1
    for c in fx["pfd_u64"]:
        v = c["values"]
        assert [0 if x > 254 else x + 1 for x in v] == c["deltas"]
        assert [x for x in v if x > 254] == c["patches"]


def test_huffman_golden(oracle, fx):
    modes = set()
    for c in fx["huffman"]:
        m, d = oracle.huffman(c["values"], c["sym_type"], c["mode"])
        assert m.hex() == c["meta"], (c["sym_type"], c["mode"], len(c["values"]))
        assert d.hex() == c["data"]
        back = oracle.huffman_decode(m, d, len(c["values"]), c["sym_type"])
        assert back.tolist() == c["values"]
        modes.add((c["sym_type"], c["mode"], m[:2] == b"\xe1\xc1"))
    # the fixture exercises Huffman and Exp-Golomb metas for the counts stream
    assert (2, 1, True) in modes and (2, 1, False) in modes


def test_hilbert_golden(oracle, fx):
    for m, it, h, back in fx["hilbert"]:
        assert oracle.m2h(m, it) == h
        assert oracle.h2m(m, it) == back


def test_hilbert_block_mode_map_golden(oracle, fx):
    full = np.array([oracle.m2h(m, 20) for m in range(1 << 18)], dtype="<u8")
    assert hashlib.sha256(full.tobytes()).hexdigest() == fx["hilbert18_sha256"]
    assert int(full.max()) == fx["hilbert18_max"] < (1 << 18)
