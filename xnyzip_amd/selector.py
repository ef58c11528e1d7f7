"""Training data, training and evaluation of the learned mode selector.

    python -m xnyzip_amd.selector --chunks 200000 --out xnyzip_amd/data/selector_params.npy

Chunks are random windows of synthetic pools (terrain and vesicle-like
shells, both from xnyzip_amd.datasets) with sizes 2^U(lo, hi) points, a
random quantizer and an L2 bound log-uniform in [0.001, 0.1].  Each chunk's
label is the exhaustive search's optimum (curve, RLE, Golomb-Rice k, PFD
width; xnyz_search_batch_device), its input the 32 device features.  The
MLP is trained by the library's HIP kernels (xnyz_mlp_train).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
PARAMS_PATH = os.path.join(PKG, "data", "selector_params.npy")


def load_params(path: str = PARAMS_PATH) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (9196,):
        raise ValueError(f"{path}: not a selector parameter blob")
    return p


def make_pools(n_terrain: int, n_shells: int, seed: int):
    from . import datasets
    return [datasets.terrain(n_terrain, seed=seed), datasets.shells(n_shells, seed=seed + 1)]


def sample_chunks(rng, pool_sizes, count, lo, hi):
    """(pool, start, n) windows with n = 2^U(lo, hi)."""
    out = []
    for _ in range(count):
        p = int(rng.integers(len(pool_sizes)))
        n = int(2.0 ** rng.uniform(lo, hi))
        n = max(1, min(n, pool_sizes[p]))
        s = int(rng.integers(0, pool_sizes[p] - n + 1))
        out.append((p, s, n))
    return out


def random_opts(rng, count):
    return [dict(quantizer=int(rng.integers(2)), l2_bound=float(10 ** rng.uniform(-3, -1)))
            for _ in range(count)]


def batches(items, max_points):
    """Groups of (chunk, ...) tuples holding at most max_points points each
    (a larger chunk forms a group of its own)."""
    b, tot = [], 0
    for it in items:
        n = it[0][2]
        if b and tot + n > max_points:
            yield b
            b, tot = [], 0
        b.append(it)
        tot += n
    if b:
        yield b


def gather(pools_dev, chunk_list):
    import torch
    pts = torch.cat([pools_dev[p][s:s + n] for p, s, n in chunk_list]).contiguous()
    offs = np.concatenate([[0], np.cumsum([n for _, _, n in chunk_list])]).tolist()
    return pts, offs


def label_chunks(X, pools_dev, chunks, opts, max_points=1 << 25, log=None):
    """Features and search results of every chunk, batch by batch."""
    feats, res = [], []
    done = 0
    t0 = time.time()
    for b in batches(list(zip(chunks, opts)), max_points):
        cl = [c for c, _ in b]
        ol = [o for _, o in b]
        pts, offs = gather(pools_dev, cl)
        feats.append(X.features_batch(pts, offs, ol))
        res.extend(X.search_batch(pts, offs, ol))
        done += len(b)
        if log:
            log(f"labelled {done}/{len(chunks)} chunks ({time.time() - t0:.0f} s)")
    return np.concatenate(feats), res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--chunks", type=int, default=200000)
    ap.add_argument("--eval-chunks", type=int, default=2000)
    ap.add_argument("--log2-min", type=float, default=12.0)
    ap.add_argument("--log2-max", type=float, default=20.0)
    ap.add_argument("--pool", type=int, default=1 << 24, help="points per pool")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=PARAMS_PATH)
    args = ap.parse_args()

    import torch

    import xnyzip_amd as X

    def log(msg):
        print(f"[selector] {msg}", flush=True)

    rng = np.random.default_rng(args.seed)
    pools = make_pools(args.pool, args.pool, args.seed)
    pools_dev = [torch.from_numpy(p).cuda() for p in pools]
    sizes = [len(p) for p in pools]
    log(f"pools ready: {sizes}")
    chunks = sample_chunks(rng, sizes, args.chunks, args.log2_min, args.log2_max)
    opts = random_opts(rng, args.chunks)
    F, R = label_chunks(X, pools_dev, chunks, opts, log=log)
    labels = [r["best"] for r in R]
    t = time.time()
    params, loss = X.mlp_train(F, labels, epochs=args.epochs, batch=256, lr=3e-3,
                               seed=args.seed)
    log(f"trained {args.epochs} epochs on {len(labels)} chunks in {time.time() - t:.1f} s, "
        f"final loss {loss:.4f}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params.astype(np.float32), allow_pickle=False)

    # held-out chunks from fresh pools
    erng = np.random.default_rng(args.seed + 1000)
    epools = make_pools(args.pool // 4, args.pool // 4, args.seed + 1000)
    epools_dev = [torch.from_numpy(p).cuda() for p in epools]
    echunks = sample_chunks(erng, [len(p) for p in epools], args.eval_chunks, args.log2_min,
                            args.log2_max)
    eopts = random_opts(erng, args.eval_chunks)
    EF, ER = label_chunks(X, epools_dev, echunks, eopts)
    pred, conf = X.mlp_predict(params, EF)
    acc = {k: float(np.mean([p[k] == r["best"][k] for p, r in zip(pred, ER)]))
           for k in ["gr_k", "pfd_width", "rle", "is_hilbert"]}
    log(f"held-out accuracy {acc}")
    gain = sum(r["reference_bytes"] for r in ER) / sum(r["best_bytes"] for r in ER) - 1
    log(f"reference mode (Z-order, no RLE, k rule, width 8) vs exhaustive optimum: {gain:+.4%}")
    for mc in (0.0, 0.3, 0.5, 0.7):
        sel = best = 0
        fell = 0
        for b in batches(list(zip(echunks, eopts, ER)), 1 << 25):
            pts, offs = gather(epools_dev, [c for c, _, _ in b])
            ol = [dict(o, frame_mode=2) for _, o, _ in b]
            _, pres, _, _, searched = X.compress_selected_batch(pts, offs, ol, params=params,
                                                                min_confidence=mc,
                                                                want_prezstd=True)
            sel += sum(len(p) for p in pres)
            best += sum(r["best_bytes"] for _, _, r in b)
            fell += int(np.sum(searched & 1))
        log(f"min_confidence {mc}: stream bytes vs exhaustive optimum {sel / best - 1:+.4%}, "
            f"fell back on {fell}/{len(echunks)} chunks")
    return 0


if __name__ == "__main__":
    sys.exit(main())
