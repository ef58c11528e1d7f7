"""Build the run-time tuning table for this GPU and host.

    python -m xnyzip_amd.tune [--out FILE] [--sizes 20,22,24,25,26] [--reps 2]

For each size class 2^c in --sizes, compresses a synthetic terrain cloud of
1.5 * 2^c points (TO quantizer, L2 0.01, Hilbert, RLE, host-threaded zstd
frames) under candidate run-time parameters (libxnyz's xnyz_autotune) and
writes the fastest setting per class to FILE in the format
xnyz_tuning_load reads; every other class takes the setting of the nearest
tuned class.  Load the table with
Context.load_tuning(FILE); DEFAULT_TABLE is the one measured on MI355X with
16 host cores.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                             "tuning_mi355x.txt")


def fill_untuned(ctx, tuned) -> None:
    """Give every class outside `tuned` the setting of the nearest tuned
    class (the smaller one on ties)."""
    from xnyzip_amd import _lib
    if not tuned:
        return
    for c in range(_lib.SIZE_CLASSES):
        if c in tuned:
            continue
        near = min(tuned, key=lambda t: (abs(t - c), t))
        ctx.set_tuning(c, **ctx.tuning(near))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--sizes", default="20,22,24,25,26",
                    help="size classes (log2 of the point count) to tune")
    ap.add_argument("--reps", type=int, default=4)
    args = ap.parse_args(argv)

    import torch

    import xnyzip_amd as X
    from xnyzip_amd import _lib, datasets

    ctx = X.Context(0)
    tuned = []
    opts = dict(quantizer=1, l2_bound=0.01, is_hilbert=True, rle=True,
                frame_mode=_lib.FRAME_ZSTD3_MT)
    for c in [int(v) for v in args.sizes.split(",") if v]:
        n = (1 << c) + (1 << (c - 1))
        dev = torch.from_numpy(datasets.terrain(n, seed=c)).to("cuda:0")
        t0 = time.perf_counter()
        best, ms = X.autotune(dev, opts, reps=args.reps, ctx=ctx)
        print(f"class {c} ({n} points): {best} {ms:.2f} ms/call "
              f"(tuned in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        tuned.append(c)
        del dev
    fill_untuned(ctx, tuned)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    ctx.save_tuning(args.out)
    ctx.close()
    print(f"wrote {args.out}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
