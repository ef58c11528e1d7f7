"""Run-time parameters (xnyz_tuning): the per-size-class table, split frames
whose tail blocks the device encodes, the persistent grid size, and the
autotuner.  Every setting must leave the assembled stream byte-identical to
the oracle's and the frame decodable to it."""
import numpy as np
import pytest

from oracle import oracle as O

from zframe_ref import host_frame, unzstd

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MT, NONE = 1, 2
DEFAULT = {"zstd_chunk_log": 20, "copy_slice_log": 24, "gpu_frame_share": 0.0,
           "persist_blocks": 2048, "copy_first": 1, "radix_bits": 8, "radix_tile": 4096}


@pytest.fixture(scope="module")
def X():
    import xnyzip_amd
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    return xnyzip_amd


@pytest.fixture
def ctx(X):
    c = X.Context(0)
    yield c
    c.close()


def _uniform(seed, n, scale=100.0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale).astype(np.float32)


def _size_class(n):
    return max(0, int(n).bit_length() - 1)


def test_table_set_get_save_load(X, ctx, tmp_path):
    from xnyzip_amd import _lib
    for c in (0, 17, 32):
        assert ctx.tuning(c) == DEFAULT
    ctx.set_tuning(21, zstd_chunk_log=18, gpu_frame_share=0.25, persist_blocks=640)
    t21 = ctx.tuning(21)
    assert t21["zstd_chunk_log"] == 18 and t21["persist_blocks"] == 640
    assert abs(t21["gpu_frame_share"] - 0.25) < 1e-7
    assert ctx.tuning(20) == DEFAULT
    path = str(tmp_path / "tuning.txt")
    ctx.save_tuning(path)
    other = X.Context(0)
    try:
        other.load_tuning(path)
        for c in range(_lib.SIZE_CLASSES):
            a, b = ctx.tuning(c), other.tuning(c)
            assert a["zstd_chunk_log"] == b["zstd_chunk_log"], c
            assert a["persist_blocks"] == b["persist_blocks"], c
            assert a["copy_first"] == b["copy_first"], c
            assert abs(a["gpu_frame_share"] - b["gpu_frame_share"]) < 1e-4, c
        # A malformed file or an out-of-range row changes nothing.
        bad = tmp_path / "bad.txt"
        bad.write_text("# comment\n3 20 24 0.1 2048 1 8 4096\n4 20 24\n")
        with pytest.raises(_lib.XnyzError) as e:
            other.load_tuning(str(bad))
        assert e.value.code == _lib.XNYZ_EINVAL and other.tuning(3) == DEFAULT
        bad.write_text("3 20 24 0.1 2048 0 8 4096\n5 30 24 0.0 2048 1 8 4096\n")
        with pytest.raises(_lib.XnyzError) as e:
            other.load_tuning(str(bad))
        assert e.value.code == _lib.XNYZ_EINVAL and other.tuning(3) == DEFAULT
        with pytest.raises(_lib.XnyzError) as e:
            other.load_tuning(str(tmp_path / "missing.txt"))
        assert e.value.code == _lib.XNYZ_EIO
    finally:
        other.close()
    for field, value in (("zstd_chunk_log", 15), ("copy_slice_log", 31),
                         ("gpu_frame_share", 0.95), ("persist_blocks", 63), ("copy_first", 2),
                         ("radix_bits", 9), ("radix_tile", 1024)):
        with pytest.raises(_lib.XnyzError) as e:
            ctx.set_tuning(7, **{field: value})
        assert e.value.code == _lib.XNYZ_EINVAL
    assert ctx.tuning(7) == DEFAULT
    with pytest.raises(_lib.XnyzError):
        ctx.tuning(33)


def _blocks(frame, pos):
    """(offset, last flag) of every block from `pos` to the frame's last."""
    out = []
    while True:
        h = int.from_bytes(frame[pos:pos + 3], "little")
        last, kind, size = h & 1, (h >> 1) & 3, h >> 3
        out.append((pos, last))
        pos += 3 + (1 if kind == 1 else size)
        if last:
            return out, pos


@pytest.mark.parametrize("share,chunk_log,rle,first", [(0.3, 16, True, 1), (0.6, 18, False, 0),
                                                       (0.9, 17, True, 0), (0.4, 20, True, 1)])
def test_split_frame(X, ctx, share, chunk_log, rle, first):
    """The device encodes the stream's head as the frame's first blocks and
    host threads the rest: the frame decodes to the oracle's stream, its
    header carries the whole length, and the first blocks equal the device
    encoder's host twin run on the head (last-block flag cleared)."""
    n = 1_500_000
    c = _uniform(77, n)
    ref = O.compress(c, O.TO, 0.01, True, rle)
    pre = ref.pre
    ctx.set_tuning(_size_class(n), zstd_chunk_log=chunk_log, gpu_frame_share=share,
                   copy_slice_log=20, copy_first=first)
    frame, _, _, st = X.compress(torch.from_numpy(c).cuda(), O.TO, 0.01, True, rle,
                                 frame_mode=MT, ctx=ctx)
    chunk = 1 << chunk_log
    head = -(-int(len(pre) * float(np.float32(share))) // chunk) * chunk
    assert head >= 1 << 20 and head + chunk <= len(pre), "case must split"
    assert st.frame_bytes == len(frame)
    assert frame[:5] == b"\x28\xb5\x2f\xfd\xe0"
    assert int.from_bytes(frame[5:13], "little") == len(pre)
    assert unzstd(frame) == pre
    dev = bytearray(host_frame(pre[:head]))
    blocks, end = _blocks(bytes(dev), 13)
    dev[5:13] = len(pre).to_bytes(8, "little")
    dev[blocks[-1][0]] &= 0xFE
    assert frame[:end] == bytes(dev[:end]) and end == len(dev)
    host_blocks, fend = _blocks(frame, end)
    assert fend == len(frame) and len(host_blocks) >= 1
    assert np.array_equal(O.decompress_frame(frame, O.TO, 0.01, rle),
                          O.decompress_frame(ref.frame, O.TO, 0.01, rle))
    # Repeat calls give the same bytes (copy stream, events and scratch reused).
    again, _, _, _ = X.compress(torch.from_numpy(c).cuda(), O.TO, 0.01, True, rle,
                                frame_mode=MT, ctx=ctx)
    assert again == frame


def test_split_frame_into(X, ctx):
    """Split frames into caller memory.  With room for both parts' bounds the
    layout is anchored: the host's blocks sit at the device part's bound
    and the frame starts past offset 0; in a tight buffer it starts at 0.
    Pinned buffers from xnyz_host_alloc take the device's blocks by DMA.
    Every layout gives the same bytes, nothing outside the frame is
    written, and a buffer one byte short is refused."""
    from xnyzip_amd import _lib
    n = 1_500_000
    c = _uniform(86, n)
    ctx.set_tuning(_size_class(n), zstd_chunk_log=17, gpu_frame_share=0.5, copy_slice_log=20)
    dev = torch.from_numpy(c).cuda()
    opts = [dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=True, rle=True, frame_mode=MT)]
    page = np.zeros(13 * n, dtype=np.uint8)
    offs, _ = X.compress_batch_into(dev, [0, n], opts, page, ctx=ctx)
    assert offs[0] > 0, "13 bytes per point leave room for the anchored layout"
    want = bytes(page[int(offs[0]):int(offs[1])])
    assert unzstd(want) == O.compress(c, O.TO, 0.01, True, True).pre
    frame, _, _, _ = X.compress(dev, O.TO, 0.01, True, True, frame_mode=MT, ctx=ctx)
    assert frame == want
    size = len(want)
    pinned = X.host_buffer(13 * n + 4096)
    for at, cap in ((0, 13 * n), (4096, 13 * n), (8, size + 100), (0, size)):
        pinned[:] = 0xA5
        view = pinned[at:at + cap]
        o2, _ = X.compress_batch_into(dev, [0, n], opts, view, ctx=ctx)
        a, b = int(o2[0]), int(o2[1])
        assert b - a == size and bytes(view[a:b]) == want, (at, cap)
        assert (a > 0) == (cap == 13 * n), (at, cap)
        assert not np.any(pinned[:at + a] != 0xA5) and not np.any(pinned[at + b:] != 0xA5)
    tight = np.zeros(size + 100, dtype=np.uint8)
    o3, _ = X.compress_batch_into(dev, [0, n], opts, tight, ctx=ctx)
    assert o3[0] == 0 and bytes(tight[:size]) == want
    pinned[:] = 0xA5
    with pytest.raises(_lib.XnyzError) as e:
        X.compress_batch_into(dev, [0, n], opts, pinned[:size - 1], ctx=ctx)
    assert e.value.code == _lib.XNYZ_ENOSPC and not np.any(pinned[:size - 1] != 0xA5)
    del view, pinned


def test_split_needs_room(X, ctx):
    """A stream too short to split (device share under 1 MiB) takes the host
    path."""
    n = 200_000
    c = _uniform(78, n)
    ref = O.compress(c, O.TO, 0.01, False, True)
    ctx.set_tuning(_size_class(n), gpu_frame_share=0.5, zstd_chunk_log=16)
    frame, _, _, _ = X.compress(torch.from_numpy(c).cuda(), O.TO, 0.01, False, True,
                                frame_mode=MT, ctx=ctx)
    assert len(ref.pre) < 2 << 20
    assert unzstd(frame) == ref.pre


@pytest.mark.parametrize("blocks", [64, 1000, 16384])
def test_persist_blocks_same_stream(X, ctx, blocks):
    n = 400_000
    c = _uniform(79, n)
    ctx.set_tuning(-1, persist_blocks=blocks)
    for rle in (False, True):
        _, pre, _, _ = X.compress(torch.from_numpy(c).cuda(), O.TO, 0.01, True, rle,
                                  frame_mode=NONE, want_prezstd=True, ctx=ctx)
        assert pre == O.compress(c, O.TO, 0.01, True, rle).pre, (blocks, rle)


@pytest.mark.parametrize("bits,tile", [(6, 4096), (7, 2048), (8, 2048), (6, 2048)])
def test_radix_shapes_same_stream(X, ctx, bits, tile):
    """Every radix digit width and tile size sorts to the same order: the
    streams equal the oracle's for narrow keys (a uniform cloud, a cloud
    with duplicates), wide keys (~2^50 blocks: a u64 + u32 key) and a
    non-RLE cloud whose permutation payload rides along."""
    ctx.set_tuning(-1, radix_bits=bits, radix_tile=tile)
    dup = np.repeat(_uniform(81, 5000), 7, axis=0)
    cases = [(_uniform(82, 300_000), O.TO, 0.01, True, True),
             (dup, O.CUBE, 0.02, False, True),
             (_uniform(83, 20000, scale=20000.0), O.TO, 0.001, True, False),
             (_uniform(84, 20000, scale=20000.0), O.TO, 0.001, True, True),
             (_uniform(85, 123_457), O.CUBE, 0.05, True, False)]
    for i, (c, q, l2, hil, rle) in enumerate(cases):
        _, pre, _, _ = X.compress(torch.from_numpy(c).cuda(), q, l2, hil, rle,
                                  frame_mode=NONE, want_prezstd=True, ctx=ctx)
        assert pre == O.compress(c, q, l2, hil, rle).pre, (bits, tile, i)


def test_autotune(X, ctx):
    n = 1_200_000
    c = _uniform(80, n)
    dev = torch.from_numpy(c).cuda()
    opts = dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=True, rle=True, frame_mode=MT)
    best, ms = X.autotune(dev, opts, reps=1, ctx=ctx)
    assert ms > 0
    got = ctx.tuning(_size_class(n))
    assert got["radix_bits"] == best["radix_bits"] and got["radix_tile"] == best["radix_tile"]
    assert got["zstd_chunk_log"] == best["zstd_chunk_log"]
    assert got["persist_blocks"] == best["persist_blocks"]
    assert abs(got["gpu_frame_share"] - best["gpu_frame_share"]) < 1e-7
    assert ctx.tuning(_size_class(n) + 1) == DEFAULT
    frame, _, _, _ = X.compress(dev, O.TO, 0.01, True, True, frame_mode=MT, ctx=ctx)
    assert unzstd(frame) == O.compress(c, O.TO, 0.01, True, True).pre
