"""Autotuner: the cache (host logic, no GPU) and, on the GPU, that tuning
windows are real work whose results match an untuned run bit for bit."""
import json

import numpy as np
import pytest

sbc = pytest.importorskip("streamingbc_amd")
from streamingbc_amd.autotune import DEFAULT_CANDIDATES, HANDOVER, _store, autotune, cache_key, tuning_windows  # noqa: E402


class _NoDevice:
    """Stands in for a forest: a cache hit must not touch the device."""
    nv, k = 1000, 64

    def insert_batched(self, *a, **k):
        raise AssertionError("device touched on a cache hit")


def test_cache_hit_skips_the_sweep(tmp_path):
    path = str(tmp_path / "tune.json")
    key = cache_key(1000, 5000, 64, 32)
    _store(path, key, {"block_threads": 512, "grid_blocks": 0, "table": []})
    _store(path, cache_key(1, 2, 3, 4), {"block_threads": 256, "grid_blocks": 7})
    data = json.load(open(path))
    assert set(data) == {key, cache_key(1, 2, 3, 4)}
    res, used = autotune(_NoDevice(), None, None, np.zeros(64), np.zeros(64), 32, 5000, cache=path)
    assert used == 0 and res.from_cache and res.block_threads == 512
    cfg = res.config(32, recompute_bc=True)
    assert cfg.window == 32 and cfg.block_threads == 512 and cfg.recompute_bc


def test_candidates_are_valid_block_sizes():
    for c in DEFAULT_CANDIDATES:
        assert c["block_threads"] in (64, 128, 256, 512, 1024) and c["grid_blocks"] >= 0


@pytest.mark.gpu
@pytest.mark.parametrize("ordered", [False, True])
def test_tuning_windows_are_real_work(tmp_path, ordered):
    hg = sbc.gen_rmat(12, 16 << 12, seed=3)
    roots = hg.pick_roots(64, 2)
    W = 8
    u, v = hg.insertions(W * (tuning_windows(2) + 4), 5)
    # a root order turns on the heavy/light stage
    order = np.random.default_rng(4).permutation(64).astype(np.int32) if ordered else None

    def fresh():
        ctx = sbc.Context(hg.nv)
        g = sbc.Graph(hg)
        f = sbc.Forest(hg.nv, roots)
        f.init_static(g, ctx)
        return ctx, g, f

    ctx, g, f = fresh()
    path = str(tmp_path / "tune.json")
    res, used = autotune(f, g, ctx, u, v, W, hg.ne, reps=2, cache=path,
                         base=sbc.BatchConfig(window=W, root_order=order))
    assert used <= tuning_windows(2, ordered) * W and used % W == 0 and not res.from_cache
    stages = [r["stage"] for r in res.table]
    split = ["heavy/light split"] * 2 if ordered else []
    assert stages == ["team size"] * 3 + split + ["hand-over thresholds"] * len(HANDOVER)
    assert all(r["ms_per_window"] > 0 for r in res.table)
    assert res.heavy_roots in ((-1, 64 - 16) if ordered else (-1,))
    assert (res.park_min, res.park_scan) in [(0, 0)] + list(HANDOVER)
    f.insert_batched(g, ctx, u[used:], v[used:], res.config(W))
    f.recompute_bc(ctx)
    ctx2, g2, f2 = fresh()
    f2.insert_batched(g2, ctx2, u, v, sbc.BatchConfig(window=W, block_threads=256))
    f2.recompute_bc(ctx2)
    for a, b in zip(f.trees(), f2.trees()):
        assert np.array_equal(a, b)
    assert np.array_equal(f.bc(), f2.bc())
    # second call: served from the cache
    res2, used2 = autotune(f, g, ctx, u, v, W, hg.ne, cache=path)
    assert res2.from_cache and used2 == 0 and res2.block_threads == res.block_threads
    assert (res2.heavy_roots, res2.park_min, res2.park_scan) == (res.heavy_roots, res.park_min, res.park_scan)
