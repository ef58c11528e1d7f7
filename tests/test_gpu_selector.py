"""Learned mode selector: features, the HIP MLP kernels (against a numpy
FP32 model of the same FP16-operand arithmetic), training, and batched
compression with selected / searched modes (byte-exact against the oracle
for the chosen mode)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def X():
    import xnyzip_amd
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    return xnyzip_amd


def _uniform(seed, n, scale=100.0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale).astype(np.float32)


def _f16(a):
    return a.astype(np.float16).astype(np.float32)


def _np_logits(P, F):
    mean, istd = P[0:32], P[32:64]
    o = 64
    W1 = P[o:o + 2048].reshape(32, 64); o += 2048
    b1 = P[o:o + 64]; o += 64
    W2 = P[o:o + 4096].reshape(64, 64); o += 4096
    b2 = P[o:o + 64]; o += 64
    W3 = P[o:o + 2816].reshape(64, 44); o += 2816
    b3 = P[o:o + 44]
    xn = _f16((F - mean) * istd)
    h1 = _f16(np.maximum(xn @ _f16(W1) + b1, 0))
    h2 = _f16(np.maximum(h1 @ _f16(W2) + b2, 0))
    return h2 @ _f16(W3) + b3


HEADS = [(0, 32), (32, 40), (40, 42), (42, 44)]


def test_features_match_exact_statistics(X):
    clouds = [_uniform(1, 40000), np.round(_uniform(2, 30000) / 3, 1).astype(np.float32),
              _uniform(3, 5000, 20.0)]
    opts = [dict(quantizer=O.TO, l2_bound=0.01), dict(quantizer=O.CUBE, l2_bound=0.05),
            dict(quantizer=O.TO, l2_bound=0.1)]
    pts = np.concatenate(clouds)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).tolist()
    F = X.features_batch(torch.from_numpy(pts).cuda(), offs, opts)
    assert F.shape == (3, 32) and np.isfinite(F).all()
    for c, o, f in zip(clouds, opts, F):
        ref = O.compress(c, o["quantizer"], o["l2_bound"], False, True)
        assert f[0] == pytest.approx(np.log2(len(c)), abs=1e-5)
        for a in range(3):
            assert f[1 + a] == pytest.approx(np.log2(1 + ref.info.range[a]), abs=1e-5)
        u_est = 2.0 ** f[5]
        assert abs(u_est / ref.info.num_points - 1) < 0.2  # HyperLogLog, 256 registers
        assert 0 <= f[15] <= 1 and 0 <= f[16] <= 1 and 0 <= f[17] <= 1


def test_predict_matches_numpy_model(X):
    rng = np.random.default_rng(5)
    P = (rng.normal(size=9196) * 0.3).astype(np.float32)
    P[0:32] = rng.normal(size=32).astype(np.float32)
    P[32:64] = rng.uniform(0.5, 2, 32).astype(np.float32)
    F = rng.normal(size=(3000, 32)).astype(np.float32)
    modes, conf = X.mlp_predict(P, F)
    Z = _np_logits(P, F)
    keys = ["gr_k", "pfd_width", "rle", "is_hilbert"]
    checked = 0
    for i, m in enumerate(modes):
        c = 1.0
        for h, (a, b) in enumerate(HEADS):
            z = Z[i, a:b]
            top = np.sort(z)[::-1]
            p = np.exp(z - z.max())
            c = min(c, 1.0 / p.sum())
            if top[0] - top[1] > 1e-2:
                want = int(np.argmax(z)) + (1 if keys[h] == "pfd_width" else 0)
                assert m[keys[h]] == want, (i, h)
                checked += 1
        assert conf[i] == pytest.approx(c, abs=2e-3)
    assert checked > 4 * 3000 * 0.9


def test_training_learns_a_rule(X):
    rng = np.random.default_rng(6)
    F = rng.normal(size=(8192, 32)).astype(np.float32) * 3 + 1
    g = F > 1
    modes = [dict(gr_k=int(4 * a + 2 * b + c), pfd_width=int(1 + d + 2 * e), rle=int(f),
                  is_hilbert=int(gg))
             for a, b, c, d, e, f, gg in zip(g[:, 0], g[:, 1], g[:, 2], g[:, 3], g[:, 4],
                                             g[:, 5], g[:, 6] ^ g[:, 7])]
    p1, l1 = X.mlp_train(F, modes, epochs=1, lr=3e-3, seed=1)
    p2, l2 = X.mlp_train(F, modes, epochs=40, lr=3e-3, params=p1, seed=2)
    assert l2 < 0.5 * l1
    pred, _ = X.mlp_predict(p2, F)
    for key in ["gr_k", "pfd_width", "rle", "is_hilbert"]:
        acc = np.mean([p[key] == m[key] for p, m in zip(pred, modes)])
        assert acc > 0.9, key


def _clouds():
    return [(_uniform(11, 20000), dict(quantizer=O.TO, l2_bound=0.01)),
            (np.round(_uniform(12, 15000) / 4, 1).astype(np.float32),
             dict(quantizer=O.CUBE, l2_bound=0.02)),
            (_uniform(13, 6000, 20.0), dict(quantizer=O.TO, l2_bound=0.1)),
            (_uniform(14, 9000, 3000.0), dict(quantizer=O.TO, l2_bound=0.05))]


def test_selected_batch_streams(X):
    cases = _clouds()
    pts = torch.from_numpy(np.concatenate([c for c, _ in cases])).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    opts = [dict(o, frame_mode=0) for _, o in cases]
    # search only (params None): the chosen modes are the search's optimum
    fr, pre, st, modes, searched = X.compress_selected_batch(pts, offs, opts, want_prezstd=True)
    res = X.search_batch(pts, offs, [o for _, o in cases])
    assert list(searched) == [1] * len(cases)
    for i, ((c, o), m, r) in enumerate(zip(cases, modes, res)):
        assert m == r["best"], i
        assert len(pre[i]) == r["best_bytes"], i
        want = O.compress_mode(c, o["quantizer"], o["l2_bound"], bool(m["is_hilbert"]),
                               bool(m["rle"]), gr_k=m["gr_k"], pfd_width=m["pfd_width"])
        assert pre[i] == want, i
        back = O.decompress_frame(fr[i], o["quantizer"], o["l2_bound"], bool(m["rle"]))
        assert np.array_equal(back, O.decompress_buffer(want, o["quantizer"], o["l2_bound"],
                                                        bool(m["rle"])))
    # with a selector: confident clouds take the prediction, the rest the search
    F = X.features_batch(pts, offs, [o for _, o in cases])
    # np.repeat repeats rows in place, so each label repeats in place too
    params, _ = X.mlp_train(np.repeat(F, 64, axis=0),
                            [r["best"] for r in res for _ in range(64)], epochs=30, seed=3)
    fr2, pre2, _, modes2, searched2 = X.compress_selected_batch(pts, offs, opts, params=params,
                                                                min_confidence=0.5,
                                                                want_prezstd=True)
    pred, conf = X.mlp_predict(params, F)
    for i, ((c, o), m) in enumerate(zip(cases, modes2)):
        var = 2 * m["is_hilbert"] + m["rle"]
        if searched2[i] & 1:
            assert conf[i] < 0.5 and m == res[i]["best"]
        else:
            assert conf[i] >= 0.5
            assert (m["is_hilbert"], m["rle"]) == (pred[i]["is_hilbert"], pred[i]["rle"])
            if not searched2[i] & 2:
                assert (m["gr_k"], m["pfd_width"]) == (pred[i]["gr_k"], pred[i]["pfd_width"])
        # k and width are the exact optimum of the chosen curve x RLE variant
        assert len(pre2[i]) == res[i]["variants"][var]["bytes"], i
        want = O.compress_mode(c, o["quantizer"], o["l2_bound"], bool(m["is_hilbert"]),
                               bool(m["rle"]), gr_k=m["gr_k"], pfd_width=m["pfd_width"])
        assert pre2[i] == want, i


def _mixed_chunks():
    """Terrain and shell chunks, a small direct-mode shell, and a TO cloud in
    direct mode whose minimum-z point takes the wrapped z = -1 node (its
    sfc delta is near 2^63: a Golomb-Rice k below ~40 would cost more than
    2^23 bytes)."""
    from xnyzip_amd import datasets as D
    ter = D.terrain(1 << 18, seed=5)
    she = D.shells(1 << 18, seed=6, count=40)
    d = float(np.float32(np.float64(np.float32(1.0)) / np.sqrt(5.0)))
    w = _uniform(21, 3000, scale=100 * d)
    w[17] = [d, d, 0.0]
    w[:, 2] += 1.0
    w[17, 2] = 1.0
    w[200] = [0.0, 0.0, 3.0]
    return [(ter[:100000], dict(quantizer=O.TO, l2_bound=0.01)),
            (ter[100000:160000], dict(quantizer=O.CUBE, l2_bound=0.1)),
            (she[:70000], dict(quantizer=O.TO, l2_bound=0.05)),
            (she[70000:72000], dict(quantizer=O.TO, l2_bound=0.1)),
            (_uniform(13, 6000, 20.0), dict(quantizer=O.TO, l2_bound=0.1)),
            (w, dict(quantizer=O.TO, l2_bound=0.5, direct_threshold=1 << 20))]


def test_packaged_selector_streams(X):
    """The packaged selector parameters (as used by the selector benchmark) on
    mixed chunks, through the caller-memory entry point reusing one context
    after search-only calls: every stream is the oracle's for the chosen mode,
    k / width are the variant's exact optimum, and no cloud blows up."""
    from xnyzip_amd import selector
    params = selector.load_params()
    cases = _mixed_chunks()
    pts = torch.from_numpy(np.concatenate([c for c, _ in cases])).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    sopts = [dict(o) for _, o in cases]
    opts = [dict(o, frame_mode=2) for _, o in cases]
    ctx = X.Context(0)
    res = X.search_batch(pts, offs, sopts, ctx=ctx)
    _, _, _, _, s0 = X.compress_selected_batch(pts, offs, opts, ctx=ctx)
    assert list(s0) == [1] * len(cases)
    out = np.zeros(64 << 20, np.uint8)
    for rep in range(2):
        ro, st, modes, searched = X.compress_selected_batch_into(
            pts, offs, opts, out, params=params, min_confidence=0.5, ctx=ctx)
        total = best = 0
        for i, ((c, o), m) in enumerate(zip(cases, modes)):
            pre = bytes(out[int(ro[i]):int(ro[i + 1])])
            var = 2 * m["is_hilbert"] + m["rle"]
            assert len(pre) == res[i]["variants"][var]["bytes"], (rep, i)
            assert len(pre) <= 1.02 * res[i]["reference_bytes"], (rep, i, m)
            want = O.compress_mode(c, o["quantizer"], o["l2_bound"], bool(m["is_hilbert"]),
                                   bool(m["rle"]), o.get("direct_threshold", 1024),
                                   gr_k=m["gr_k"], pfd_width=m["pfd_width"])
            assert pre == want, (rep, i)
            total += len(pre)
            best += res[i]["best_bytes"]
        print(f"selected {total} B vs search optimum {best} B "
              f"({100.0 * (total / best - 1):+.3f}%), searched {list(searched)}")
    assert st[-1].is_direct and st[-1].wrapped_points == 1


def test_selected_batch_async_matches_sync(X):
    """The asynchronous caller-memory form (copy queued on the copy stream,
    bytes in `out` after ctx.wait()): three calls back to back, alternating
    between two batches so that both of the context's stream buffers carry
    a copy in flight, give the synchronous form's bytes, offsets and modes;
    the second batch is the smaller."""
    from xnyzip_amd import selector
    params = selector.load_params()
    cases = _mixed_chunks()
    pts = torch.from_numpy(np.concatenate([c for c, _ in cases])).cuda()
    offs = np.concatenate([[0], np.cumsum([len(c) for c, _ in cases])]).tolist()
    from xnyzip_amd import _lib
    opts = [dict(o, frame_mode=_lib.FRAME_NONE) for _, o in cases]
    sub = cases[1:4]
    pts2 = torch.from_numpy(np.concatenate([c for c, _ in sub])).cuda()
    offs2 = np.concatenate([[0], np.cumsum([len(c) for c, _ in sub])]).tolist()
    opts2 = opts[1:4]
    ctx = X.Context(0)
    want = []
    for p, o, op in ((pts, offs, opts), (pts2, offs2, opts2)):
        buf = np.zeros(32 << 20, np.uint8)
        ro, _, modes, searched = X.compress_selected_batch_into(
            p, o, op, buf, params=params, min_confidence=0.5, ctx=ctx)
        want.append((bytes(buf[:int(ro[-1])]), ro.tolist(), modes, list(searched)))
    outs = [X.host_buffer(32 << 20) for _ in range(3)]
    got = []
    for k, (p, o, op) in enumerate(((pts, offs, opts), (pts2, offs2, opts2), (pts, offs, opts))):
        outs[k][:] = 0xA5
        ro, _, modes, searched = X.compress_selected_batch_into(
            p, o, op, outs[k], params=params, min_confidence=0.5, ctx=ctx, wait=False)
        got.append((ro.tolist(), modes, list(searched)))
    ctx.wait()
    for k in range(3):
        w = want[k % 2]
        ro, modes, searched = got[k]
        assert ro == w[1] and modes == w[2] and searched == w[3], k
        assert bytes(outs[k][:ro[-1]]) == w[0], k
