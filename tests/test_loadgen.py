"""Host logic of the serving workload: client population, request streams."""
import numpy as np

from fastatari import loadgen as LG


def test_client_population():
    specs = LG.serve_clients(0xFA7A0005)
    assert len(specs) == 256
    n = np.array([s.n for s in specs])
    assert n.min() >= 64 and n.max() < 2048
    # E[64 * 32**u] = 64 * 31 / ln 32 = 572.5
    assert 450 < n.mean() < 700
    r = np.array([s.rate_mult for s in specs])
    assert r.min() >= 0.5 and r.max() <= 2.0
    assert set(s.cls for s in specs) == {0, 1, 2, 3}
    assert specs == LG.serve_clients(0xFA7A0005)


def test_client_stream_is_deterministic_and_restarts_one_percent():
    spec = LG.ClientSpec(1, 2000, 1.0, 1234)
    a = LG.ClientStream(spec, record=np.arange(0, 2000, 10))
    b = LG.ClientStream(spec)
    total = 0
    for _ in range(50):
        (a1, r1), (a2, r2) = a.next(), b.next()
        assert np.array_equal(a1, a2) and np.array_equal(r1, r2)
        assert a1.max() < 18
        total += int(r1.sum())
    assert 0.006 < total / (50 * 2000) < 0.014
    assert len(a.history) == 50 and a.history[0][0].shape == (200,)


def test_summarize():
    res = LG.LoadResult(service_ms=[1.0, 2.0, 3.0, 100.0], e2e_ms=[1.0, 2.0, 3.0, 200.0], cls=[0, 0, 1, 2],
                        env_frames=4000, requests=4, wall_s=2.0)
    s = LG.summarize(res)
    assert s["fps"] == 2000.0 and s["requests"] == 4
    assert s["p50_ms"] == 2.5 and s["per_class"][0]["requests"] == 2


def test_jump_ahead_matches_the_sequential_stream():
    spec = LG.ClientSpec(2, 300, 1.0, 0xABCDEF)
    seq = LG.ClientStream(spec)
    steps = [seq.next() for _ in range(7)]
    envs = np.array([0, 5, 299])
    a, r = LG.ClientStream(spec).inputs_at(7, envs)
    for s in range(7):
        assert np.array_equal(a[s], steps[s][0][envs])
        assert np.array_equal(r[s], steps[s][1][envs])
