"""Writes tests/golden/workloads.json: the oracle's assembled stream for the
50M-point terrain (TO, L2 0.01, Hilbert, RLE) as a SHA-256, with the input's
SHA-256 and the SHA-256 of the points the oracle's decoder returns for that
stream, so a GPU test can check the HIP compressor and decoder at that size
without running the oracle (a few minutes of one core) on the GPU host.

    python tests/golden/make_workloads.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main() -> int:
    from oracle import oracle as O
    from xnyzip_amd import datasets as D
    pts = D.terrain(50_000_000, seed=0)
    pre = O.compress_mode(pts, O.TO, 0.01, True, True)
    # num_points of an RLE stream is its first u64 (the unique cells)
    entry = {"generator": "datasets.terrain(50_000_000, seed=0)",
             "options": "TO, L2 0.01, Hilbert, RLE, direct_threshold 1024",
             "points": int(pts.shape[0]),
             "input_sha256": hashlib.sha256(pts.tobytes()).hexdigest(),
             "num_points": int.from_bytes(pre[:8], "little"),
             "stream_bytes": len(pre),
             "stream_sha256": hashlib.sha256(pre).hexdigest()}
    back = O.decompress_buffer(pre, O.TO, 0.01, True)
    entry["decoded_points"] = int(back.shape[0])
    entry["decoded_sha256"] = hashlib.sha256(back.astype("<f4").tobytes()).hexdigest()
    path = os.path.join(HERE, "workloads.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    data["terrain-50M"] = entry
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
        f.write("\n")
    print(json.dumps(entry, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
