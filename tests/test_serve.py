"""Serving loop: admission by HBM footprint (host logic, runs without a GPU)
and, on the GPU, that a graph served in a mix with others ends bit-identical
to the same graph served alone."""
import numpy as np
import pytest

sbc = pytest.importorskip("streamingbc_amd")
from streamingbc_amd.serve import Server, poisson_arrivals, summarize  # noqa: E402


def test_resident_footprint_counts_forest_graph_and_scratch():
    hg = sbc.gen_rmat(10, 8192, seed=1)
    k = 32
    fp = sbc.resident_footprint(hg, k, 256)
    deg = hg.degree()
    slots = int(np.sum(deg + 4 + deg // 16))
    # forest state plus at least the CSR slots (column + sequence per slot)
    assert fp >= sbc.forest_footprint(hg.nv, k) + 8 * slots
    # hub tables and their scratch appear only with vertices above 512 edges
    assert deg.max() <= 512
    assert fp < sbc.forest_footprint(hg.nv, k) + 8 * slots + 64 * hg.nv
    hub = sbc.gen_chung_lu(1 << 14, 2.0, 16.0, seed=3)
    assert hub.degree().max() > 512
    fh = sbc.resident_footprint(hub, k, 256)
    # lookup-bucket scratch: 4 workgroups per CU, 512 * 256 term slots of 24 bytes each
    assert fh - sbc.resident_footprint(hub, k, 128) == 128 * 4 * 512 * 256 * 24
    assert sbc.resident_footprint(hg, 0, 256) == -1


def test_admission_refuses_over_budget_before_touching_the_device():
    hg = sbc.gen_rmat(10, 8192, seed=1)
    need = sbc.resident_footprint(hg, 64, 256)
    srv = Server(budget_bytes=need - 1, num_cu=256)
    with pytest.raises(sbc.SbcError) as e:
        srv.admit("a", hg, np.arange(64))
    assert e.value.status == sbc.SBC_ERR_BUDGET
    assert srv.used == 0 and not srv.resident
    with pytest.raises(sbc.SbcError) as e:
        srv.admit("a", hg, np.zeros(0, dtype=np.int64))
    assert e.value.status == sbc.SBC_ERR_ARG


def test_poisson_arrivals_and_summary():
    t = poisson_arrivals(1000.0, 20000, 11)
    assert np.all(np.diff(t) > 0)
    assert abs(t[-1] / 20000 - 1e-3) < 5e-5            # mean gap 1/rate
    assert np.array_equal(t, poisson_arrivals(1000.0, 20000, 11))
    s = summarize({"a": np.array([0.001, 0.002, 0.003]), "b": np.array([0.010])}, 2.0, {"a": [2, 1], "b": [1]})
    assert s["insertions"] == 4 and s["sustained_ins_per_s"] == 2.0
    assert s["graphs"]["a"]["windows"] == 2 and s["graphs"]["a"]["mean_window"] == 1.5
    assert s["p50_ms"] == 2.5


def _graphs():
    out = {}
    for sc, k in ((10, 32), (11, 48), (12, 64)):
        hg = sbc.gen_chung_lu(1 << sc, 2.3, 8.0, seed=sc)
        out[f"cl{sc}"] = (hg, hg.pick_roots(k, sc))
    return out


@pytest.mark.gpu
def test_alone_vs_mix_bit_identical():
    gs = _graphs()
    n = 96
    arr = {}
    for i, (name, (hg, roots)) in enumerate(gs.items()):
        u, v = hg.insertions(n, 11)
        arr[name] = (poisson_arrivals(4000.0, n, 11 + i), u, v)
    print("[serve] alone", flush=True)
    alone = {}
    for name, (hg, roots) in gs.items():
        srv = Server(budget_bytes=8 << 30, max_window=16)
        srv.admit(name, hg, roots)
        srv.serve({name: arr[name]})
        r = srv.resident[name]
        alone[name] = (srv.bc(name), r.f.trees())
        srv.evict(name)
    print("[serve] mix", flush=True)
    srv = Server(budget_bytes=8 << 30, max_window=16)
    for name, (hg, roots) in gs.items():
        srv.admit(name, hg, roots)
    res = srv.serve(arr)
    assert res["insertions"] == 3 * n and res["p99_ms"] >= res["p50_ms"] > 0
    for name in gs:
        bc = srv.bc(name)
        assert np.array_equal(bc, alone[name][0]), name
        for a, b in zip(srv.resident[name].f.trees(), alone[name][1]):
            assert np.array_equal(a, b), name
        assert srv.resident[name].applied == n
        # every window ran as a replay of its graph (graphs are built without
        # stream capture: no other thread's work can invalidate them)
        built, launched = srv.resident[name].ctx.graph_counts()
        assert built >= 1 and launched == len(srv.resident[name].windows), name


def test_submit_refuses_a_graph_above_the_whole_budget():
    hg = sbc.gen_rmat(10, 8192, seed=1)
    need = sbc.resident_footprint(hg, 16, 256)
    srv = Server(budget_bytes=need - 1, num_cu=256)
    with pytest.raises(sbc.SbcError) as e:
        srv.submit("a", hg, np.arange(16))
    assert e.value.status == sbc.SBC_ERR_BUDGET
    assert srv.used == 0 and not srv.resident and not srv.pending


@pytest.mark.gpu
def test_queued_graph_admitted_while_serving():
    """Three graphs, a budget that holds two: the third is queued and
    admitted during serve() once a graph whose stream has ended is evicted;
    every graph ends bit-identical to the same graph served alone."""
    gs = _graphs()
    names = list(gs)
    fps = {n: sbc.resident_footprint(hg, len(r), 0) for n, (hg, r) in gs.items()}
    budget = fps[names[1]] + fps[names[2]]
    assert fps[names[0]] + fps[names[1]] <= budget < sum(fps.values())
    counts = {names[0]: 24, names[1]: 64, names[2]: 64}
    arr = {}
    for i, (name, (hg, roots)) in enumerate(gs.items()):
        u, v = hg.insertions(counts[name], 17)
        arr[name] = (poisson_arrivals(2000.0, counts[name], 5 + i), u, v)
    alone = {}
    for name, (hg, roots) in gs.items():
        srv = Server(budget_bytes=8 << 30, max_window=16)
        srv.admit(name, hg, roots)
        srv.serve({name: arr[name]})
        alone[name] = srv.bc(name)
        srv.evict(name)
    srv = Server(budget_bytes=budget, max_window=16)
    got = [srv.submit(name, hg, roots) for name, (hg, roots) in gs.items()]
    assert got == [True, True, False] and len(srv.pending) == 1
    results = {}
    res = srv.serve(arr, evict_when_done=True, results=results)
    print(f"[serve] queued graph admitted at {srv.admitted_at.get(names[2])} s; p99 {res['p99_ms']} ms", flush=True)
    assert names[2] in srv.admitted_at and not srv.pending and not srv.resident and srv.used == 0
    assert res["insertions"] == sum(counts.values())
    for name in gs:
        assert np.array_equal(results[name], alone[name]), name
