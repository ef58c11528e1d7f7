"""The multi-chunk container on the host (no GPU): part file + JSON-lines
manifest in the shape of the MPI driver's (source/mpi_driver.cpp:709-714,
:902-914), with each chunk's mode.  Streams come from the oracle; the
container's host decode must give the oracle decoder's points, chunk after
chunk."""
import json

import numpy as np

from oracle import oracle as O


def _cases():
    rng = np.random.default_rng(2)
    return [
        ((rng.random((3000, 3)) * 50).astype(np.float32),
         dict(quantizer=O.TO, l2_bound=0.01, is_hilbert=True, rle=True, frame_mode=0)),
        ((rng.random((2000, 3)) * 10).astype(np.float32),
         dict(quantizer=O.CUBE, l2_bound=0.05, is_hilbert=False, rle=False, frame_mode=2)),
        ((rng.random((500, 3)) * 3).astype(np.float32),
         dict(quantizer=O.TO, l2_bound=0.2, is_hilbert=False, rle=True, frame_mode=2,
              direct_threshold=1 << 20)),
    ]


def test_write_read_decompress(tmp_path):
    import xnyzip_amd as X
    outs, chunks, want, at = [], [], [], 0
    for pts, o in _cases():
        thr = o.get("direct_threshold", 1024)
        ref = O.compress(pts, o["quantizer"], o["l2_bound"], o["is_hilbert"], o["rle"], thr)
        outs.append(ref.frame if o["frame_mode"] == 0 else ref.pre)
        want.append(O.decompress_buffer(ref.pre, o["quantizer"], o["l2_bound"], o["rle"]))
        chunks.append(dict(o, input_offset=at * 12, input_points=len(pts),
                           direct_threshold=thr))
        at += len(pts)
    offs = np.concatenate([[0], np.cumsum([len(b) for b in outs])])
    buf = np.frombuffer(b"".join(outs), np.uint8)
    part, man = str(tmp_path / "part_rank0000.bin"), str(tmp_path / "manifest_rank0000.jsonl")
    written = X.container_write(part, man, buf, offs, chunks)
    assert [c["compressed_size"] for c in written] == [len(b) for b in outs]
    lines = [json.loads(l) for l in open(man)]
    assert [l["block_id"] for l in lines] == [0, 1, 2]
    assert [l["part_offset"] for l in lines] == list(offs[:-1])
    assert [l["input_size"] for l in lines] == [12 * c["input_points"] for c in chunks]
    assert [l["rle"] for l in lines] == [1, 0, 1]
    assert [l["frame"] for l in lines] == ["zstd", "stream", "stream"]
    back = X.container_read(man)
    for b, c in zip(back, chunks):
        assert (b["quantizer"], b["is_hilbert"], b["rle"], b["input_points"]) == \
            (c["quantizer"], int(c["is_hilbert"]), int(c["rle"]), c["input_points"])
        assert np.float32(b["l2_bound"]) == np.float32(c["l2_bound"])
    got = X.container_decompress(part, man, gpu=False)
    assert np.array_equal(got, np.concatenate(want))


def test_reference_manifest_line(tmp_path):
    """A line as the MPI driver writes it (no mode fields): TO, zstd frames."""
    import xnyzip_amd as X
    man = tmp_path / "m.jsonl"
    man.write_text('{"block_id":3,"input_offset":1200,"input_size":600,"rank":2,'
                   '"part_offset":77,"compressed_size":55}\n')
    (c,) = X.container_read(str(man))
    assert (c["input_offset"], c["input_points"], c["part_offset"], c["compressed_size"]) == \
        (1200, 50, 77, 55)
    assert c["quantizer"] == O.TO and c["frame_mode"] == 0 and c["rle"] == 0


def test_mapreduce_manifest_line(tmp_path):
    """The MPI driver's mapreduce line (mpi_driver.cpp:626-635): spaces after
    the colons, no input_offset, "mode": "mapreduce" -> RLE."""
    import xnyzip_amd as X
    man = tmp_path / "m.jsonl"
    man.write_text('{"rank": 0, "mode": "mapreduce", "input_bytes": 1200, '
                   '"input_blocks_owned": 2, "owned_cells": 40, "part_offset": 0, '
                   '"compressed_size": 99}\n')
    (c,) = X.container_read(str(man))
    assert (c["input_offset"], c["input_points"], c["compressed_size"], c["rle"]) == (0, 100, 99, 1)
