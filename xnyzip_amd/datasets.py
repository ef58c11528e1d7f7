"""Synthetic point clouds for the workloads the compressor is measured on.

All generators are seeded and deterministic (numpy PCG64), return (N, 3)
float32 arrays of interleaved x, y, z, and run on the host.

* uniform(n, seed, scale): i.i.d. U[0, scale)^3 -- the 1M-point plain case.
* terrain(n, seed): LiDAR-like ground: x, y uniform over 10 km x 10 km,
  z = 6-octave fBm heightfield (amplitude 200 m, Hurst 0.7) + N(0, 0.05),
  5% of points lifted by U(0, 30) (vegetation), emitted in 1000-point scan
  lines (sorted by y within each line).
* shells(n, seed, count): `count` spherical shells, centres U[0, 2000)^3,
  radii U(20, 60), points uniform on each surface with N(0, 1) radial
  thickness; each shell's points contiguous, shells in random order.
* uneven_cloud(i) / uneven_opts(i): cloud i of the batched uneven workload
  (size log-uniform in [1K, 4M] from seed i; terrain for even i, shells for
  odd i; L2 bound {0.001, 0.01, 0.1}[i mod 3]).
* selector_workload(n): half terrain, half shells, cut into chunks of
  2^U(16, 20) points with L2 bounds {0.001, 0.01, 0.1} in turn.
"""
from __future__ import annotations

import os
import threading

import numpy as np


def uniform(n: int, seed: int = 0, scale: float = 100.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3), dtype=np.float32) * np.float32(scale)).astype(np.float32)


def _value_noise_octave(x, y, freq, rng_seed):
    """Bilinear value noise on a lattice of spacing 1/freq (coords in [0,1))."""
    rng = np.random.default_rng(rng_seed)
    g = int(np.ceil(freq)) + 2
    lat = rng.standard_normal((g, g)).astype(np.float32)
    fx, fy = x * freq, y * freq
    ix, iy = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
    tx, ty = (fx - ix).astype(np.float32), (fy - iy).astype(np.float32)
    tx = tx * tx * (3 - 2 * tx)
    ty = ty * ty * (3 - 2 * ty)
    a = lat[ix, iy] * (1 - tx) + lat[ix + 1, iy] * tx
    b = lat[ix, iy + 1] * (1 - tx) + lat[ix + 1, iy + 1] * tx
    return a * (1 - ty) + b * ty


def terrain(n: int, seed: int = 0, extent: float = 10000.0, amplitude: float = 200.0,
            hurst: float = 0.7, octaves: int = 6, chunk: int = 1 << 20,
            threads: int = 0) -> np.ndarray:
    """Chunks of `chunk` points are drawn from independent streams
    (seed, chunk index), so the result does not depend on `threads`."""
    from concurrent.futures import ThreadPoolExecutor
    out = np.empty((n, 3), dtype=np.float32)
    base_freq = 4.0
    norm = sum((2.0 ** (-hurst * o)) for o in range(octaves))
    starts = list(range(0, n, chunk))

    def one(ci: int) -> None:
        s = starts[ci]
        rng = np.random.default_rng([seed, ci])
        m = min(chunk, n - s)
        u = rng.random((m, 2))
        # scan lines of 1000 points: sort each line along y
        lines = (m + 999) // 1000
        pad = lines * 1000 - m
        if pad:
            u = np.concatenate([u, rng.random((pad, 2))])
        u = u.reshape(lines, 1000, 2)
        order = np.argsort(u[:, :, 1], axis=1)
        u = np.take_along_axis(u, order[:, :, None], axis=1).reshape(-1, 2)[:m]
        x, y = u[:, 0], u[:, 1]
        z = np.zeros(m, dtype=np.float32)
        for o in range(octaves):
            z += (2.0 ** (-hurst * o)) * _value_noise_octave(x, y, base_freq * 2 ** o, seed * 131 + o)
        z *= np.float32(amplitude / norm)
        z += rng.normal(0.0, 0.05, m).astype(np.float32)
        veg = rng.random(m) < 0.05
        z[veg] += rng.uniform(0.0, 30.0, int(veg.sum())).astype(np.float32)
        out[s:s + m, 0] = (x * extent).astype(np.float32)
        out[s:s + m, 1] = (y * extent).astype(np.float32)
        out[s:s + m, 2] = z

    workers = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(one, range(len(starts))))
    return out


def shells(n: int, seed: int = 0, count: int = 2000, box: float = 2000.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, box, (count, 3))
    radii = rng.uniform(20, 60, count)
    sizes = np.full(count, n // count)
    sizes[: n - sizes.sum()] += 1
    order = rng.permutation(count)
    out = np.empty((n, 3), dtype=np.float32)
    pos = 0
    for k in order:
        m = int(sizes[k])
        v = rng.standard_normal((m, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        r = radii[k] + rng.standard_normal(m)
        out[pos:pos + m] = (centres[k] + v * r[:, None]).astype(np.float32)
        pos += m
    return out


def uneven_size(i: int) -> int:
    rng = np.random.default_rng(i)
    return int(np.exp(rng.uniform(np.log(1e3), np.log(4e6))))


def uneven_cloud(i: int, threads: int = 4) -> np.ndarray:
    n = uneven_size(i)
    if i % 2 == 0:
        return terrain(n, seed=i, threads=threads)
    return shells(n, seed=i, count=max(1, min(2000, n // 1000)))


def uneven_workload(count: int, threads: int = 0, progress=None):
    """The first `count` uneven clouds back to back in one array (filled in
    place by a thread pool over the clouds, so the whole set is never held
    twice), their point offsets and options.  progress(done, count) is
    called after every 256th finished cloud."""
    from concurrent.futures import ThreadPoolExecutor
    sizes = [uneven_size(i) for i in range(count)]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    out = np.empty((int(offs[-1]), 3), dtype=np.float32)

    done = [0]
    lock = threading.Lock()

    def one(i: int) -> None:
        out[offs[i]:offs[i + 1]] = uneven_cloud(i, threads=1)
        with lock:
            done[0] += 1
            if progress is not None and done[0] % 256 == 0:
                progress(done[0], count)

    # largest clouds first, so the pool does not end on one big cloud
    order = sorted(range(count), key=lambda i: -sizes[i])
    with ThreadPoolExecutor(max_workers=threads or min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(one, order))
    return out, offs.tolist(), [uneven_opts(i) for i in range(count)]


def uneven_opts(i: int) -> dict:
    return dict(quantizer=1, l2_bound=[0.001, 0.01, 0.1][i % 3])


def selector_workload(n: int):
    """(points, chunk offsets, per-chunk opts) of the selector workload."""
    half = n // 2
    pts = np.concatenate([terrain(half, seed=0), shells(half, seed=1)])
    rng = np.random.default_rng(0)
    offs = [0]
    while offs[-1] < len(pts):
        offs.append(min(len(pts), offs[-1] + int(2.0 ** rng.uniform(16, 20))))
    opts = [dict(quantizer=1, l2_bound=[0.001, 0.01, 0.1][i % 3])
            for i in range(len(offs) - 1)]
    return pts, offs, opts
