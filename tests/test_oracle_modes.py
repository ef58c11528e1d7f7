"""Mode overrides in the oracle: a forced Golomb-Rice k or a lower PFD limit
changes the stream but not what the reference's decoder recovers from it
(the k is read from the GR meta, golomb_rice_encoder.hpp:150-193; the PFD
decoder only tests for symbol 0, p_for_delta_encoder.hpp:45-57)."""
import numpy as np
import pytest

from oracle import oracle as O


def _pts(seed, n, scale):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale).astype(np.float32)


# Direct mode codes sfc deltas near 2^24 here: small k would make the unary
# parts gigabytes long, so that case forces k near the rule's.
@pytest.mark.parametrize("rle,hil,scale,thr,ks", [
    (False, False, 100.0, 1024, (0, 3, 9, 30, -1)),
    (True, True, 100.0, 1024, (0, 3, 9, 30, -1)),
    (False, True, 20.0, 1 << 20, (22, 20, 25, 30, -1))])
def test_override_streams_decode_like_plain(rle, hil, scale, thr, ks):
    c = _pts(3, 4000, scale)
    plain = O.compress(c, O.TO, 0.01, hil, rle, thr).pre
    want = O.decompress_buffer(plain, O.TO, 0.01, rle)
    sizes = set()
    for k, w in zip(ks, (1, 4, 8, 2, 5)):
        s = O.compress_mode(c, O.TO, 0.01, hil, rle, thr, gr_k=k, pfd_width=w)
        sizes.add(len(s))
        assert np.array_equal(O.decompress_buffer(s, O.TO, 0.01, rle), want)
    assert len(sizes) > 1  # the overrides do change the stream


def test_defaults_are_the_reference():
    c = _pts(4, 3000, 100.0)
    for rle in (False, True):
        assert O.compress_mode(c, O.TO, 0.01, False, rle) == O.compress(c, O.TO, 0.01, False,
                                                                        rle).pre


def test_width8_rule_k_is_the_reference():
    """Forcing the reference's own k at width 8 reproduces its stream."""
    c = _pts(5, 3000, 100.0)
    plain = O.compress(c, O.TO, 0.01, False, False).pre
    # the GR meta of section 3 holds k: header 30 B, two BitPacker sections
    off = 30
    for _ in range(2):
        ds = int.from_bytes(plain[off:off + 8], "little")
        ms = int.from_bytes(plain[off + 8:off + 16], "little")
        off += 16 + ms + ds
    ms = int.from_bytes(plain[off + 8:off + 16], "little")
    assert ms == 5
    k = plain[off + 16 + 4]
    assert O.compress_mode(c, O.TO, 0.01, False, False, gr_k=k, pfd_width=8) == plain
