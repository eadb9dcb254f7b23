"""GPU parity of heterogeneous batches: four ROM classes (2 KB Fixed WSYNC,
8 KB F8, 16 KB F6 no-WSYNC, 32 KB F4) with uneven env counts, K steps per
graph launch, per-env frameskip -- every env bit-exact against the oracle
(set_action + fs x run_frame per step) at every launch boundary, for every
launch shape the autotuner can pick."""
import numpy as np
import pytest

from gpu_util import assert_same

pytestmark = pytest.mark.gpu

SEED = 0xFA7A0004


@pytest.fixture(scope="module")
def W():
    from fastatari import workloads
    return workloads


@pytest.fixture(scope="module")
def roms(W):
    return W.mix_roms(SEED)


def _refs(orc, roms, counts):
    refs = []
    for rom, n in zip(roms, counts):
        r = orc.Batch(rom, n, nthreads=16)
        r.reset()
        refs.append(r)
    return refs


def _inputs(W, n, k, launch, fs_mode="per_env"):
    acts = W.ActionStream(SEED + 17 * launch, n).frames(k)
    if fs_mode == "per_env":
        fs = np.broadcast_to(W.frameskips(SEED, n), (k, n)).copy()
    else:
        fs = np.stack([W.frameskips(SEED + 1000 * launch + s, n) for s in range(k)])
    return acts, fs


def _run_refs(refs, ranges, acts, fs):
    for (first, cnt), ref in zip(ranges, refs):
        for s in range(acts.shape[0]):
            ref.run_frameskip(acts[s, first:first + cnt], fs[s, first:first + cnt])


def _check(mix, refs, ranges, where):
    for c, ((first, cnt), ref) in enumerate(zip(ranges, refs)):
        assert_same(mix, ref, f"{where} class {c}", gpu_rows=slice(first, first + cnt))


def test_reset_and_k8_frameskip(fa_mod, orc, W, roms):
    counts = [37, 300, 101, 64]
    mix = fa_mod.BatchMixGpu(roms, counts)
    refs = _refs(orc, roms, counts)
    ranges = [mix.class_range(c) for c in range(4)]
    assert ranges == [(0, 37), (37, 300), (337, 101), (438, 64)]
    obs = mix.reset()
    for (first, cnt), ref in zip(ranges, refs):
        assert np.array_equal(obs[first:first + cnt], ref.obs())
    _check(mix, refs, ranges, "reset")
    for launch in range(2):
        acts, fs = _inputs(W, mix.n, 8, launch, "per_env" if launch == 0 else "per_step")
        obs = mix.step(acts, fs, obs_every=(launch == 1))
        _run_refs(refs, ranges, acts, fs)
        _check(mix, refs, ranges, f"launch {launch}")
        last = obs[-1] if launch == 1 else obs
        for (first, cnt), ref in zip(ranges, refs):
            assert np.array_equal(last[first:first + cnt], ref.obs()), f"obs launch {launch}"


def test_obs_every_step(fa_mod, orc, W, roms):
    counts = [8, 70, 9, 20]
    mix = fa_mod.BatchMixGpu(roms, counts)
    refs = _refs(orc, roms, counts)
    ranges = [mix.class_range(c) for c in range(4)]
    mix.reset()
    acts, fs = _inputs(W, mix.n, 4, 5, "per_step")
    obs = mix.step(acts, fs, obs_every=True)
    for s in range(4):
        for (first, cnt), ref in zip(ranges, refs):
            ref.run_frameskip(acts[s, first:first + cnt], fs[s, first:first + cnt])
            assert np.array_equal(obs[s, first:first + cnt], ref.obs()), f"step {s}"


def test_graph_replays_with_eager_work_between(fa_mod, orc, W, roms):
    """The captured graph is replayed several times on fresh inputs written
    into the batch buffers, with eager copies between replays; every replay is
    checked against the oracle."""
    counts = [64, 128, 64, 64]
    mix = fa_mod.BatchMixGpu(roms, counts)
    refs = _refs(orc, roms, counts)
    ranges = [mix.class_range(c) for c in range(4)]
    mix.reset()
    for launch in range(3):
        acts, fs = _inputs(W, mix.n, 8, 10 + launch, "per_step")
        mix.write_inputs(acts, fs)
        mix.launch(8, use_fs=True)
        ms = mix.sync()
        assert ms > 0
        obs = mix.read_obs()
        _run_refs(refs, ranges, acts, fs)
        _check(mix, refs, ranges, f"replay {launch}")
        for (first, cnt), ref in zip(ranges, refs):
            assert np.array_equal(obs[first:first + cnt], ref.obs())


@pytest.mark.parametrize("block,sorted_,k_split", [(64, False, 1), (128, True, 3), (256, False, 0),
                                                   (64, True, 2)])
def test_launch_shapes_agree(fa_mod, orc, W, roms, block, sorted_, k_split):
    counts = [100, 200, 90, 70]
    mix = fa_mod.BatchMixGpu(roms, counts)
    for c in range(4):
        mix.set_class_config(c, block, sorted_, k_split)
        cfg = mix.class_config(c)
        assert cfg == {"block": block, "lanes": 0, "sorted": sorted_, "k_split": k_split}
    refs = _refs(orc, roms, counts)
    ranges = [mix.class_range(c) for c in range(4)]
    mix.reset()
    acts, fs = _inputs(W, mix.n, 8, 20, "per_step")
    mix.step(acts, fs)
    _run_refs(refs, ranges, acts, fs)
    _check(mix, refs, ranges, f"block {block} sorted {sorted_} k_split {k_split}")


@pytest.mark.parametrize("block,lanes,sorted_,k_split", [(64, 8, False, 0), (64, 16, True, 3), (64, 32, False, 1),
                                                         (128, 16, False, 0)])
def test_spread_launch_shapes_agree(fa_mod, orc, W, roms, block, lanes, sorted_, k_split):
    """Fewer envs per workgroup than its lanes (task lists of 8 / 16 / 32
    envs, the last one partial), also with opcode grouping, split launches,
    the graph path and obs of every step."""
    counts = [100, 200, 90, 70]
    mix = fa_mod.BatchMixGpu(roms, counts)
    for c in range(4):
        mix.set_class_config(c, block, sorted_, k_split)
        mix.set_class_lanes(c, lanes)
        assert mix.class_config(c) == {"block": block, "lanes": lanes, "sorted": sorted_, "k_split": k_split}
    refs = _refs(orc, roms, counts)
    ranges = [mix.class_range(c) for c in range(4)]
    mix.reset()
    acts, fs = _inputs(W, mix.n, 8, 40, "per_step")
    obs = mix.step(acts, fs, obs_every=True)
    for s in range(8):
        for (first, cnt), ref in zip(ranges, refs):
            ref.run_frameskip(acts[s, first:first + cnt], fs[s, first:first + cnt])
            assert np.array_equal(obs[s, first:first + cnt], ref.obs()), f"obs step {s}"
    _check(mix, refs, ranges, f"lanes {lanes} block {block}")
    for launch in range(2):
        acts, fs = _inputs(W, mix.n, 8, 41 + launch, "per_step")
        mix.write_inputs(acts, fs)
        mix.launch(8, use_fs=True)
        mix.sync()
        _run_refs(refs, ranges, acts, fs)
        _check(mix, refs, ranges, f"lanes {lanes} graph replay {launch}")
    with pytest.raises(ValueError):
        mix.set_class_lanes(0, 24)


def test_tune_keeps_state_and_parity(fa_mod, orc, W, roms):
    counts = [256, 1024, 256, 256]
    mix = fa_mod.BatchMixGpu(roms, counts)
    refs = _refs(orc, roms, counts)
    ranges = [mix.class_range(c) for c in range(4)]
    mix.reset()
    acts, fs = _inputs(W, mix.n, 8, 30, "per_env")
    mix.step(acts, fs)
    _run_refs(refs, ranges, acts, fs)
    before = mix.download_states_bytes().copy()
    cnt_before = mix.counters()
    acts, fs = _inputs(W, mix.n, 8, 31, "per_env")
    res = mix.tune(acts, fs)
    assert len(res) == 4
    for r in res:
        assert r["block"] in (64, 128, 256) and r["tuned_ms"] > 0 and r["default_ms"] > 0
        assert r["lanes"] in (0, 8, 16, 32) and (r["lanes"] == 0 or r["block"] == 64)
    assert np.array_equal(mix.download_states_bytes(), before), "tuning changed env state"
    cnt_after = mix.counters()
    for k in cnt_before:
        assert np.array_equal(cnt_before[k], cnt_after[k]), f"tuning changed {k}"
    mix.step(acts, fs)
    _run_refs(refs, ranges, acts, fs)
    _check(mix, refs, ranges, "after tune")


def test_full_size_sampled(fa_mod, orc, W, roms):
    """Config-4 shape: 1K / 40K / 12K / 3K envs, K = 8 steps with per-env
    frameskip 1-4, two graph launches; a seeded sample of envs per class is
    replayed on the oracle (envs are independent, so a sample run alone sees
    the same inputs as in the full batch)."""
    counts = list(W.MIX_COUNTS)
    mix = fa_mod.BatchMixGpu(roms, counts)
    mix.reset()
    ranges = [mix.class_range(c) for c in range(4)]
    rng = np.random.default_rng(4)
    picks = [np.sort(rng.choice(cnt, size=min(cnt, 48), replace=False)) for cnt in counts]
    refs = _refs(orc, roms, [len(p) for p in picks])
    for launch in range(2):
        acts, fs = _inputs(W, mix.n, 8, 40 + launch, "per_env")
        mix.step(acts, fs)
        for (first, cnt), ref, p in zip(ranges, refs, picks):
            for s in range(8):
                ref.run_frameskip(acts[s, first + p], fs[s, first + p])
    for c, ((first, cnt), ref, p) in enumerate(zip(ranges, refs, picks)):
        assert_same(mix, ref, f"full size class {c}", env_ids=first + p, gpu_rows=first + p)
    st = mix.counters()["status"]
    assert not st.any(), "an env reported an emulation error"
