"""GPU tests of the env-step server: clients start from the ROM's reset state,
steps and restarts match the oracle bit-exactly, a client's results are the
same alone and in a loaded mix, admission follows the memory budget, and a
client cannot have two requests in flight."""
import numpy as np
import pytest

from gpu_util import describe_state_diff

pytestmark = pytest.mark.gpu

ROM_SEED = 0xFA7A0004      # the four mixed ROMs
SEED = 0xFA7A0005


@pytest.fixture(scope="module")
def W():
    from fastatari import workloads
    return workloads


@pytest.fixture(scope="module")
def LG():
    from fastatari import loadgen
    return loadgen


@pytest.fixture(scope="module")
def roms(W):
    return W.mix_roms(ROM_SEED)


@pytest.fixture(scope="module")
def snaps(orc, roms):
    out = []
    for rom in roms:
        b = orc.Batch(rom, 1, nthreads=1)
        b.reset()
        out.append((b.states()[0].copy(), b.counters()[0].copy(), b.obs()[0].copy()))
    return out


def _compare(got, ref, idx, where):
    g = got["states"][idx]
    r = ref.states()
    assert np.array_equal(g, r), f"{where}: {describe_state_diff(g, r, idx)}"
    rc = ref.counters()
    assert np.array_equal(got["cpu_cycles"][idx], rc[:, 0]), f"{where}: cpu_cycles"
    assert np.array_equal(got["frame_cycles"][idx], rc[:, 1]), f"{where}: frame_cycles"
    assert np.array_equal(got["status"][idx].astype(np.uint64), rc[:, 3]), f"{where}: status"


def _replay(orc, rom, snap, stream, frames):
    idx = stream.record
    ref = orc.Batch(rom, len(idx), nthreads=16)
    ref.reset()
    fs = np.full(len(idx), frames, np.uint8)
    for a, r in stream.history:
        ref.restart(r, snap[0], snap[1])
        ref.run_frameskip(a, fs)
    return ref


def test_open_gives_reset_state(fa_mod, roms, snaps):
    srv = fa_mod.EnvServer(roms, budget_bytes=1 << 30)
    for c in range(4):
        cid, obs = srv.open(c, 100)
        d = srv.download(cid)
        assert (d["states"] == snaps[c][0]).all(), f"class {c}"
        assert (d["cpu_cycles"] == snaps[c][1][0]).all()
        assert (d["frame_cycles"] == snaps[c][1][1]).all()
        assert (obs == snaps[c][2]).all()
    srv.close_server()


@pytest.mark.parametrize("cls", [0, 1, 2, 3])
def test_steps_and_restarts_match_oracle(fa_mod, orc, W, LG, roms, snaps, cls):
    srv = fa_mod.EnvServer(roms, budget_bytes=1 << 30)
    n = 150
    cid, _ = srv.open(cls, n)
    spec = LG.ClientSpec(cls, n, 1.0, SEED + cls)
    stream = LG.ClientStream(spec, frames=2, record=np.arange(n))
    stream.threshold = np.uint64(int(0.1 * 2 ** 64))       # 10 % restarts to hit many envs
    for _ in range(8):
        a, r = stream.next()
        srv.submit(cid, a, r, 2)
        obs, ms = srv.wait(cid)
        assert ms > 0
    ref = _replay(orc, roms[cls], snaps[cls], stream, 2)
    _compare(srv.download(cid), ref, np.arange(n), f"class {cls}")
    assert np.array_equal(obs, ref.obs())
    srv.close_server()


def test_client_alone_vs_mix(fa_mod, orc, LG, roms, snaps):
    """Client 0 stepped alone and stepped inside a closed-loop mix of 24
    clients gives identical states and counters; every client of the mix also
    matches the oracle on a sample of its envs."""
    specs = LG.serve_clients(SEED, n_clients=24, lo=64, ratio=8.0)
    steps = 6

    def run(which):
        srv = fa_mod.EnvServer(roms, budget_bytes=1 << 30)
        cids, streams = [], []
        for i in which:
            s = specs[i]
            cid, _ = srv.open(s.cls, s.n)
            cids.append(cid)
            streams.append(LG.ClientStream(s, frames=1, record=np.arange(0, s.n, 7)))
        LG.run_load(srv, [specs[i] for i in which], cids, streams, duration_s=120.0, max_steps=steps)
        out = [srv.download(c) for c in cids]
        srv.close_server()
        return out, streams

    alone, _ = run([0])
    mix, streams = run(list(range(len(specs))))
    for k in ("states", "cpu_cycles", "frame_cycles", "status"):
        assert np.array_equal(alone[0][k], mix[0][k]), f"client 0 {k} differs alone vs in the mix"
    for i, (d, st) in enumerate(zip(mix, streams)):
        assert len(st.history) == steps
        ref = _replay(orc, roms[specs[i].cls], snaps[specs[i].cls], st, 1)
        _compare(d, ref, st.record, f"client {i}")


def test_open_loop_load_matches_oracle(fa_mod, orc, LG, roms, snaps):
    """Poisson arrivals at 0.5-2x an assumed saturation rate, overlapping
    launches on several streams per class; sampled envs of every client match
    the oracle afterwards."""
    specs = LG.serve_clients(SEED + 1, n_clients=32, lo=64, ratio=16.0)
    srv = fa_mod.EnvServer(roms, budget_bytes=1 << 30)
    cids, streams = [], []
    for s in specs:
        cid, _ = srv.open(s.cls, s.n)
        cids.append(cid)
        streams.append(LG.ClientStream(s, frames=1, record=np.arange(0, s.n, 11)))
    sat = np.array([40.0 if s.cls == 2 else 200.0 for s in specs])
    res = LG.run_load(srv, specs, cids, streams, duration_s=3.0, sat_rate=sat)
    summ = LG.summarize(res)
    assert summ["requests"] > 0 and summ["p50_ms"] > 0
    smp = srv.samples()
    assert len(smp["ms"]) == summ["requests"] and smp["launches"] > 0
    for i, s in enumerate(specs):
        ref = _replay(orc, roms[s.cls], snaps[s.cls], streams[i], 1)
        _compare(srv.download(cids[i]), ref, streams[i].record, f"client {i}")
    srv.close_server()


def test_admission_against_budget(fa_mod, roms):
    probe = fa_mod.EnvServer(roms[:2], budget_bytes=1 << 30)
    chunk = 65536 * probe.memory()["bytes_per_slot"]
    probe.close_server()
    srv = fa_mod.EnvServer(roms[:2], budget_bytes=chunk)
    ids = [srv.open(0, 2048)[0] for _ in range(32)]
    assert srv.memory()["used_bytes"] == chunk
    with pytest.raises(fa_mod.Busy):
        srv.open(0, 64)           # the class-0 chunk is full
    with pytest.raises(fa_mod.Busy):
        srv.open(1, 64)           # a class-1 chunk would exceed the budget
    srv.close(ids[5])
    cid, _ = srv.open(0, 2000)    # the freed range is reused
    assert srv.memory()["used_bytes"] == chunk
    with pytest.raises(ValueError):
        srv.open(0, 0)
    srv.close_server()


def test_second_submit_while_in_flight_is_busy(fa_mod, roms):
    srv = fa_mod.EnvServer(roms, budget_bytes=1 << 30)
    cid, _ = srv.open(2, 256)     # the no-WSYNC ROM: about 10 ms per frame
    a = np.zeros(256, np.uint8)
    srv.submit(cid, a, None, 30)
    with pytest.raises(fa_mod.Busy):
        srv.submit(cid, a, None, 1)
    obs, ms = srv.wait(cid)
    assert ms > 0
    with pytest.raises(ValueError):
        srv.submit(cid, np.zeros(10, np.uint8))
    srv.close_server()


def _replay_counts(orc, roms, snaps, specs, steps, records, frames=1):
    """Oracle replay of the recorded envs of every client from its completed
    step count (inputs regenerated by jumping the client's streams ahead);
    one oracle Batch per ROM class, clients with fewer steps padded with
    zero-frame steps repeating their last action (no state change)."""
    from fastatari import loadgen as LG
    out = {}
    for c in range(len(roms)):
        members = [i for i, sp in enumerate(specs) if sp.cls == c and steps[i] > 0 and len(records[i])]
        for i, sp in enumerate(specs):
            if sp.cls == c and steps[i] == 0:
                out[i] = None
        if not members:
            continue
        ins = {i: LG.ClientStream(specs[i]).inputs_at(int(steps[i]), records[i]) for i in members}
        rows = [len(records[i]) for i in members]
        n = sum(rows)
        ref = orc.Batch(roms[c], n, nthreads=16)
        ref.reset()
        for s in range(int(max(steps[i] for i in members))):
            a = np.zeros(n, np.uint8)
            r = np.zeros(n, np.uint8)
            fs = np.zeros(n, np.uint8)
            off = 0
            for i, m in zip(members, rows):
                acts, rst = ins[i]
                if s < len(acts):
                    a[off:off + m], r[off:off + m], fs[off:off + m] = acts[s], rst[s], frames
                else:
                    a[off:off + m] = acts[-1]
                off += m
            ref.restart(r, snaps[c][0], snaps[c][1])
            ref.run_frameskip(a, fs)
        st, cn = ref.states(), ref.counters()
        off = 0
        for i, m in zip(members, rows):
            out[i] = (st[off:off + m], cn[off:off + m])
            off += m
    return out


def test_config5_full_scale_sampled(fa_mod, orc, LG, roms, snaps):
    """The serving configuration at full scale: 256 clients (~146K envs,
    64-2047 each) on the four ROMs, native load driver, closed loop then
    Poisson arrivals at 0.5-2x saturation, 1 % restarts per step; 1 % of the
    envs (every 100th of each client) replayed on the oracle bit-exactly."""
    specs = LG.serve_clients(SEED)
    srv = fa_mod.EnvServer(roms, budget_bytes=64 << 30)
    cids = [srv.open(sp.cls, sp.n)[0] for sp in specs]
    steps = np.zeros(len(specs), np.uint64)
    closed = srv.run_load(cids, specs, 0.6, steps=steps)
    assert closed["requests"] > 0
    sat = np.maximum(steps, 1) / closed["wall_s"]
    opened = srv.run_load(cids, specs, 1.0, sat=sat, steps=steps, arrival_seed=3)
    assert opened["requests"] > 0 and (opened["service_ms"] > 0).all()
    records = [np.arange(0, sp.n, 100) for sp in specs]
    ref = _replay_counts(orc, roms, snaps, specs, steps, records)
    checked = 0
    for i, (cid, sp) in enumerate(zip(cids, specs)):
        d = srv.download(cid)
        if ref[i] is None:
            assert (d["states"] == snaps[sp.cls][0]).all()
            continue
        st, cn = ref[i]
        idx = records[i]
        assert np.array_equal(d["states"][idx], st), f"client {i}: {describe_state_diff(d['states'][idx], st, idx)}"
        assert np.array_equal(d["cpu_cycles"][idx], cn[:, 0]), f"client {i} cpu_cycles"
        assert np.array_equal(d["frame_cycles"][idx], cn[:, 1]), f"client {i} frame_cycles"
        assert not d["status"].any()
        checked += len(idx)
    assert checked >= 1000
    srv.close_server()


def test_native_driver_alone_vs_mix(fa_mod, LG, roms):
    """The native driver: a client's states after S steps are the same when it
    runs alone and when it runs inside a closed-loop mix of 40 clients."""
    specs = LG.serve_clients(SEED + 7, n_clients=40, lo=64, ratio=16.0)
    S = 5

    def run(which):
        srv = fa_mod.EnvServer(roms, budget_bytes=1 << 30)
        cids = [srv.open(specs[i].cls, specs[i].n)[0] for i in which]
        steps = np.zeros(len(which), np.uint64)
        srv.run_load(cids, [specs[i] for i in which], 120.0, steps=steps, max_steps=S)
        assert (steps == S).all()
        out = [srv.download(c) for c in cids]
        srv.close_server()
        return out

    alone = run([3])
    mix = run(list(range(len(specs))))
    for k in ("states", "cpu_cycles", "frame_cycles", "status"):
        assert np.array_equal(alone[0][k], mix[3][k]), f"{k} differs alone vs in the mix"
