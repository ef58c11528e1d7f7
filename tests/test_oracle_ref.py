"""Randomised differential checks: oracle restatement vs the reference's own
encoder and Hilbert sources (compiled into oracle/_ref from the reference
checkout).  Skipped where oracle/_ref is not built (e.g. on a GPU box)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


@pytest.mark.parametrize("seed", range(6))
def test_encoders_match_reference(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 3000))
    kinds = [
        rng.integers(0, 2**18, n, dtype=np.uint64),
        np.minimum(rng.geometric(0.02, n), 2**40).astype(np.uint64),
        rng.integers(0, 2**63, n, dtype=np.uint64) >> rng.integers(0, 63, n).astype(np.uint64),
    ]
    for v in kinds:
        assert oracle.bitpack_u64(v) == oracle.bitpack_u64(v, impl="ref")
        if int(v.astype(object).sum()) < 2**64:
            assert oracle.gr_pack(v) == oracle.gr_pack(v, impl="ref")
        sym = np.where(v > 254, 0, v + 1).astype(np.uint64)
        assert oracle.huffman(sym, O.SYM_U64) == oracle.huffman(sym, O.SYM_U64, impl="ref")
    q = (1 + rng.binomial(7, rng.random(), n)).astype(np.uint8)
    assert oracle.huffman(q, O.SYM_U8) == oracle.huffman(q, O.SYM_U8, impl="ref")
    assert oracle.gr_pack(q, np.uint8) == oracle.gr_pack(q, np.uint8, impl="ref")
    for c in (rng.geometric(rng.uniform(0.05, 0.95), n),
              rng.integers(1, int(rng.integers(2, 5000)), n)):
        c = c.astype(np.int32)
        for mode in (0, 1, 2):
            assert oracle.huffman(c, O.SYM_INT, mode) == \
                oracle.huffman(c, O.SYM_INT, mode, impl="ref")


def test_hilbert_matches_reference(oracle):
    rng = np.random.default_rng(7)
    for it in range(0, 21):
        for m in rng.integers(0, 2**63, 200, dtype=np.uint64).tolist():
            assert oracle.m2h(m, it) == O.ref_m2h(m, it)
            assert oracle.h2m(m, it) == O.ref_h2m(m, it)
    for m in range(0, 1 << 18, 97):
        assert oracle.m2h(m, 20) == O.ref_m2h(m, 20)


@pytest.mark.parametrize("seed", range(4))
def test_shared_codebooks_match_reference(oracle, seed):
    """build_complete over 0..255 (fit_profile) and its meta reloaded by a
    fresh encoder (compress_with_profile): meta bytes and data bits."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 5000))
    sym = np.where(rng.random(n) < 0.2, 0, 1 + rng.geometric(0.1, n).clip(0, 254))
    q = (1 + rng.binomial(7, rng.random(), n)).astype(np.uint8)
    for mode in (3, 4):
        assert oracle.huffman(sym.astype(np.uint64), O.SYM_U64, mode) == \
            oracle.huffman(sym.astype(np.uint64), O.SYM_U64, mode, impl="ref")
        assert oracle.huffman(q, O.SYM_U8, mode) == oracle.huffman(q, O.SYM_U8, mode, impl="ref")
