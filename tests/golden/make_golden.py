"""Write tests/golden/encoders.json from the reference's own code.

Runs in the container that holds the reference checkout: it loads
oracle/_ref/libxnyz_ref.so, which `make -C oracle ref` compiles from the
reference's encoder headers (source/encoder/*.hpp) and
source/quantizer/hilbert-curves-cpp/hilbert.cpp as they lie.  The fixture
holds inputs and the reference's outputs only (hex strings), so the oracle can
be pinned on machines without the reference.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "encoders.json")


def u64_cases(rng):
    cases = [
        [0], [1], [0, 0, 0], [5, 3, 7, 255, 256, 1 << 40],
        [2**64 - 1, 0, 1], list(range(1, 40)), [3] * 17,
    ]
    cases.append(rng.integers(0, 300, 500).tolist())
    cases.append(rng.integers(0, 2**62, 50, dtype=np.uint64).tolist())
    cases.append((rng.geometric(0.01, 400) + 254).tolist())
    return cases


def huffman_cases(rng):
    out = []
    # repos-like PFD deltas (0 = patch, 1..255)
    out.append((O.SYM_U64, 0, np.minimum(rng.geometric(0.05, 3000), 255).tolist()))
    out.append((O.SYM_U64, 0, rng.integers(0, 256, 2000).tolist()))
    out.append((O.SYM_U64, 0, [7] * 10))
    out.append((O.SYM_U64, 0, [1, 2]))
    out.append((O.SYM_U64, 0, [0, 0, 9, 9, 9, 4]))
    # quads-like (1..8)
    out.append((O.SYM_U8, 0, rng.integers(1, 9, 1500).tolist()))
    out.append((O.SYM_U8, 0, (1 + rng.binomial(7, 0.2, 1000)).tolist()))
    out.append((O.SYM_U8, 0, [1] * 33))
    # counts-like, double build (RLE counts)
    out.append((O.SYM_INT, 1, rng.geometric(0.7, 2500).tolist()))
    out.append((O.SYM_INT, 1, [1] * 100))
    out.append((O.SYM_INT, 1, rng.integers(1, 3000, 1200).tolist()))   # entropy > 10
    out.append((O.SYM_INT, 1, rng.integers(1, 40, 60).tolist()))       # meta-heavy
    out.append((O.SYM_INT, 1, (rng.geometric(0.2, 4000) * 3).tolist()))
    out.append((O.SYM_INT, 2, rng.geometric(0.3, 700).tolist()))
    out.append((O.SYM_INT, 0, [-5, 3, 3, -5, 0, 2**31 - 1, -2**31]))
    # shared codebooks (own generator, so the cases above keep their values):
    # build_complete over 0..255 (mode 3) and its meta reloaded (mode 4)
    r2 = np.random.default_rng(7)
    for mode in (3, 4):
        out.append((O.SYM_U64, mode, np.minimum(r2.geometric(0.05, 3000), 255).tolist()))
        out.append((O.SYM_U8, mode, (1 + r2.binomial(7, 0.3, 2000)).tolist()))
    out.append((O.SYM_U64, 3, [2, 2, 2]))
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref`")
    rng = np.random.default_rng(20261203)
    fx = {"source": "reference encoder headers + hilbert.cpp via oracle/_ref",
          "bitpack_u64": [], "gr_u64": [], "gr_u8": [], "pfd_u64": [],
          "huffman": [], "hilbert": []}
    for v in u64_cases(rng):
        m, d = O.bitpack_u64(v, impl="ref")
        fx["bitpack_u64"].append({"values": [int(x) for x in v], "meta": m.hex(), "data": d.hex()})
        # A u64 sum that wraps drives the reference's k to 0 and its unary
        # run to ~2^64 bits; such inputs are kept out of the GR cases.
        if sum(int(x) for x in v) < 2**64:
            m, d = O.gr_pack(v, np.uint64, impl="ref")
            fx["gr_u64"].append({"values": [int(x) for x in v], "meta": m.hex(),
                                 "data": d.hex()})
        a = np.ascontiguousarray(v, dtype=np.uint64)
        deltas = np.empty_like(a)
        patches = np.empty_like(a)
        npat = O.C.c_uint64()
        O.ref().ref_pfd_u64(a.ctypes.data, a.size, deltas.ctypes.data, patches.ctypes.data,
                            O.C.byref(npat))
        fx["pfd_u64"].append({"values": [int(x) for x in v],
                              "deltas": [int(x) for x in deltas],
                              "patches": [int(x) for x in patches[: npat.value]]})
    for v in ([0], [255, 254, 0, 1], rng.integers(0, 256, 300).tolist(), [9] * 12):
        m, d = O.gr_pack(v, np.uint8, impl="ref")
        fx["gr_u8"].append({"values": [int(x) for x in v], "meta": m.hex(), "data": d.hex()})
    fx["gr_u64"].append({"values": [], "meta": "", "data": ""})
    for st, mode, v in huffman_cases(rng):
        m, d = O.huffman(v, st, mode, impl="ref")
        fx["huffman"].append({"sym_type": st, "mode": mode, "values": [int(x) for x in v],
                              "meta": m.hex(), "data": d.hex()})
    trip = []
    for it in list(range(1, 21)):
        for m in rng.integers(0, 2**63, 12, dtype=np.uint64).tolist() + [0, 1, 7, 8, 63]:
            trip.append([int(m), it, int(O.ref_m2h(int(m), it)), int(O.ref_h2m(int(m), it))])
    fx["hilbert"] = trip
    # The whole block-mode map: mortonToHilbert3(m, 20) for every 18-bit m.
    full = np.array([O.ref_m2h(m, 20) for m in range(1 << 18)], dtype="<u8")
    fx["hilbert18_sha256"] = hashlib.sha256(full.tobytes()).hexdigest()
    fx["hilbert18_max"] = int(full.max())
    with open(OUT, "w") as f:
        json.dump(fx, f, separators=(",", ":"))
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
