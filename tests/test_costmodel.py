"""The learned cost model: FP32 training kernels against a float64 numpy
restatement of the same network, BF16-MFMA inference against a numpy
emulation of BF16 rounding, window features against values read back from
the forest, and scheduling invariance (every schedule gives bit-identical
trees and BC).

Tolerances: FP32 forward within 1e-5 relative of float64; FP32 gradients
within 1e-3 relative (float32 sums over a few hundred samples); MFMA output
within 1e-3 of the numpy BF16 emulation (FP32 accumulation order differs) and
within 5e-2 of the FP32 path (BF16 has an 8-bit mantissa).
"""
import numpy as np
import pytest

from sbc_testutil import karate_path

sbc = pytest.importorskip("streamingbc_amd")
from streamingbc_amd.schedule import LearnedSchedule, StaticSchedule, run_stream, spearman  # noqa: E402

NF, NI, NH = 12, 16, 64


def unpack(p):
    p = p.astype(np.float64)
    o = 0
    W1 = p[o:o + NI * NH].reshape(NI, NH); o += NI * NH
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1[:NF], b1, W2, b2, w3, b3


def np_forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = X.astype(np.float64) @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3, (z1, h1, z2, h2)


def np_grad(p, X, y):
    """Gradient of mean((out - y)^2) by hand-written backpropagation."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = X.astype(np.float64)
    out, (z1, h1, z2, h2) = np_forward(p, X)
    n = X.shape[0]
    d = 2.0 * (out - y) / n
    gw3 = h2.T @ d
    gb3 = d.sum()
    dz2 = np.outer(d, w3) * (z2 > 0)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (z1 > 0)
    gW1 = np.zeros((NI, NH))
    gW1[:NF] = X.T @ dz1
    gb1 = dz1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def np_forward_bf16(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = bf16(X).astype(np.float64)
    z1 = x @ bf16(W1).astype(np.float64) + b1
    h1 = bf16(np.maximum(z1, 0).astype(np.float32)).astype(np.float64)
    z2 = h1 @ bf16(W2).astype(np.float64) + b2
    return np.maximum(z2, 0) @ w3 + b3


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((n, NF), dtype=np.float32)
    X[:, 0] = rng.random(n) < 0.7
    y = (2.0 * X[:, 0] * (1 + X[:, 3] + X[:, 5] ** 2) + np.sin(3 * X[:, 9])).astype(np.float32)
    return X, y


def test_spearman_matches_scipy():
    scipy_stats = pytest.importorskip("scipy.stats")
    rng = np.random.default_rng(3)
    a = rng.random(200)
    b = a + 0.3 * rng.random(200)
    assert abs(spearman(a, b) - scipy_stats.spearmanr(a, b).correlation) < 1e-12


def test_bf16_emulation_rounding():
    # ties go to the even 8-bit mantissa; other cases to nearest
    x = np.array([1.0, 1 + 2 ** -8, 1 + 2 ** -9, 1 + 3 * 2 ** -9, 1 + 3 * 2 ** -8, -2.5], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -2.5], dtype=np.float32)
    assert np.array_equal(bf16(x), want)


def test_numpy_gradient_is_a_derivative():
    # pins the float64 backpropagation itself with central differences
    rng = np.random.default_rng(5)
    p = rng.standard_normal(NI * NH + NH + NH * NH + NH + NH + 1) * 0.3
    p[NF * NH:NI * NH] = 0
    X, y = _data(32, 9)
    g = np_grad(p, X, y)
    live = np.flatnonzero(np.r_[np.arange(NI * NH) < NF * NH, np.ones(p.size - NI * NH, bool)])
    for q in rng.choice(live, 40, replace=False):
        e = np.zeros_like(p)
        e[q] = 1e-6
        f1 = np.mean((np_forward(p + e, X)[0] - y) ** 2)
        f0 = np.mean((np_forward(p - e, X)[0] - y) ** 2)
        assert abs((f1 - f0) / 2e-6 - g[q]) <= 1e-5 * (1 + abs(g[q]))


@pytest.mark.gpu
def test_forward_and_gradients_match_numpy():
    m = sbc.CostModel(seed=3)
    p = m.params
    assert p.shape == (NI * NH + NH + NH * NH + NH + NH + 1,)
    assert np.all(p[NF * NH:NI * NH] == 0)
    X, y = _data(300, 1)
    out = m.predict(X, mfma=False)
    ref, _ = np_forward(p, X)
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-5)
    g, loss = m.gradients(X, y)
    gr = np_grad(p, X, y)
    assert abs(loss - np.mean((ref - y) ** 2)) <= 1e-4 * (1 + loss)
    scale = np.abs(gr).max()
    assert np.abs(g - gr).max() <= 1e-3 * scale, np.abs(g - gr).max() / scale
    assert np.all(g[NF * NH:NI * NH] == 0)


@pytest.mark.gpu
def test_mfma_inference_matches_bf16_emulation():
    m = sbc.CostModel(seed=4)
    X, y = _data(4096, 2)
    m.train(X, y, epochs=3, batch=256, lr=3e-3, seed=2)
    p = m.params
    for n in (1, 15, 17, 64, 1000, 4096):   # ragged tails of the 16-row tiles
        out = m.predict(X[:n], mfma=True)
        emu = np_forward_bf16(p, X[:n])
        fp = m.predict(X[:n], mfma=False)
        assert np.abs(out - emu).max() <= 1e-3 * (1 + np.abs(emu).max())
        assert np.abs(out - fp).max() <= 5e-2 * (1 + np.abs(fp).max())


@pytest.mark.gpu
def test_training_reduces_loss():
    m = sbc.CostModel(seed=5)
    X, y = _data(8192, 3)
    _, l0 = m.gradients(X, y)
    l1 = m.train(X, y, epochs=20, batch=512, lr=3e-3, seed=1)
    assert l1 < 0.05 * l0, (l0, l1)
    # same seed, same data: training is deterministic
    m2 = sbc.CostModel(seed=5)
    assert m2.train(X, y, epochs=20, batch=512, lr=3e-3, seed=1) == l1
    assert np.array_equal(m.params, m2.params)


@pytest.mark.gpu
def test_window_features_match_forest_state():
    hg = sbc.read_dimacs(karate_path())
    ctx = sbc.Context(hg.nv)
    g = sbc.Graph(hg)
    f = sbc.Forest(hg.nv)
    f.init_static(g, ctx)
    u, v = hg.insertions(12, 4)
    hist = np.linspace(0, 3, f.k).astype(np.float32)
    X = sbc.window_features(f, g, ctx, u, v, hist)
    lv, sg, dl = f.trees()
    deg = hg.degree()
    LONG_MAX = np.iinfo(np.int64).max
    for i in range(len(u)):
        for r in range(f.k):
            x = X[i, r]
            assert x[11] == np.float32(np.float32(hist[r]) / np.float32(10))
            a, b = lv[r, u[i]], lv[r, v[i]]
            if a == b:
                assert np.all(x[:11] == 0)
                continue
            s, pv = (u[i], v[i]) if b < a else (v[i], u[i])
            assert x[0] == 1
            if LONG_MAX not in (a, b):
                gap = abs(int(a) - int(b))
                assert x[1] == (gap >= 2) and abs(x[2] - min(gap, 16) / 16) < 1e-7
                assert abs(x[3] - min(lv[r, s], 16) / 16) < 1e-7 and abs(x[4] - min(lv[r, pv], 16) / 16) < 1e-7
                assert abs(x[7] - np.log1p(sg[r, pv]) / 20) < 1e-6 and abs(x[8] - np.log1p(sg[r, s]) / 20) < 1e-6
            assert abs(x[5] - np.log1p(deg[s]) / 10) < 1e-6 and abs(x[6] - np.log1p(deg[pv]) / 10) < 1e-6
            assert abs(x[9] - np.log1p(dl[r, s]) / 10) < 1e-5 and abs(x[10] - np.log1p(dl[r, pv]) / 10) < 1e-5


def _stream_case():
    hg = sbc.gen_rmat(12, 16 << 12, seed=21, keep_lcc=True)
    roots = hg.pick_roots(96, 5)
    u, v = hg.insertions(16 * 24, 13)
    return hg, roots, u, v


def _run(hg, roots, u, v, sched, window=16):
    ctx = sbc.Context(hg.nv)
    g = sbc.Graph(hg)
    f = sbc.Forest(hg.nv, roots)
    f.init_static(g, ctx)
    _, rec = run_stream(f, g, ctx, u, v, window, sched, sbc.BatchConfig(window=window, block_threads=1024))
    f.recompute_bc(ctx)
    return f, rec


@pytest.mark.gpu
@pytest.mark.parametrize("device,light_max", [(False, 64.0), (True, 64.0), (True, 0.0)])
def test_learned_schedule_is_result_invariant_and_predictive(device, light_max):
    hg, roots, u, v = _stream_case()
    print(f"[schedule] nv={hg.nv} k={len(roots)} n={len(u)} device={device} light_max={light_max}", flush=True)
    f0, r0 = _run(hg, roots, u, v, StaticSchedule())
    sched = LearnedSchedule(train_after=8, epochs=20, batch=512, lr=3e-3, light_max=light_max, device=device,
                            eval_every=1)
    f1, r1 = _run(hg, roots, u, v, sched)
    assert sched.trained
    assert len(sched.eval_pred) == 24 - 8
    # reordering and binning roots changed nothing: trees, item records and
    # BC are bit-identical
    for a, b in zip(f0.trees(), f1.trees()):
        assert np.array_equal(a, b)
    assert np.array_equal(f0.bc(), f1.bc())
    for k in ("kase", "gap", "processed", "scanned", "affected", "moved"):
        assert np.array_equal(r0[k], r1[k])
    rep = sched.report()
    print(rep, flush=True)
    # held-out windows: per-root window work is ranked far better than chance
    assert rep["root_window_spearman"] > 0.5
    assert rep["rmse_log1p"] < 1.5


@pytest.mark.gpu
def test_device_schedule_matches_host_schedule():
    """streamingBCScheduleWindow computes the features, the MFMA predictions,
    the per-root costs and the heavy count on the device; they agree with the
    host-side pipeline (window_features + predict + numpy ordering), and a
    batched call that uses the device order gives the same trees and BC."""
    hg, roots, u, v = _stream_case()
    ctx = sbc.Context(hg.nv)
    g = sbc.Graph(hg)
    f = sbc.Forest(hg.nv, roots)
    f.init_static(g, ctx)
    sched = LearnedSchedule(train_after=4, epochs=8, batch=512, lr=3e-3, light_max=64.0, device=False)
    run_stream(f, g, ctx, u[:64], v[:64], 16, sched, sbc.BatchConfig(window=16, block_threads=1024))
    assert sched.trained
    uu, vv = u[64:80], v[64:80]
    X = sbc.window_features(f, g, ctx, uu, vv, sched._hist)
    p_host = sched.model.predict(X, mfma=True)
    heavy, p_dev = sbc.schedule_window(f, g, ctx, sched.model, uu, vv, sched._hist, 64.0, 4.0, predictions=True)
    assert np.array_equal(p_host, p_dev)
    work = np.expm1(np.maximum(p_host.astype(np.float64), 0.0)).sum(axis=0) + 4.0 * len(uu)
    lo = int(np.count_nonzero(work >= 64.0 * (1 + 1e-5)))
    hi = int(np.count_nonzero(work >= 64.0 * (1 - 1e-5)))
    print(f"[schedule] heavy device={heavy} host=[{lo},{hi}]", flush=True)
    assert lo <= heavy <= hi
    # the order stays in the context for the next batched call
    f.insert_batched(g, ctx, uu, vv, sbc.BatchConfig(window=16, block_threads=1024, heavy_roots=heavy,
                                                     use_context_order=True))
    f.recompute_bc(ctx)
    # same stream without any schedule
    g2 = sbc.Graph(hg)
    f2 = sbc.Forest(hg.nv, roots)
    f2.init_static(g2, ctx)
    f2.insert_batched(g2, ctx, u[:80], v[:80], sbc.BatchConfig(window=16, block_threads=1024))
    f2.recompute_bc(ctx)
    for a, b in zip(f.trees(), f2.trees()):
        assert np.array_equal(a, b)
    assert np.array_equal(f.bc(), f2.bc())
    # a second batched call without a fresh schedule is refused
    with pytest.raises(sbc.SbcError):
        f.insert_batched(g, ctx, u[80:96], v[80:96], sbc.BatchConfig(window=16, block_threads=1024,
                                                                     use_context_order=True))
