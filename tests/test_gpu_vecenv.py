"""GPU parity of the env adapter BatchAtariGpuEnv (src/rl4burn_env.rs:268-380):
f32 observations equal the oracle's RAM / 255 in IEEE single precision, bit
for bit, after reset and after every step (broadcast action held for
`frameskip` frames, as the reference's step_multi with the action repeated,
or one action per env); the full packed state, counters and status equal
the oracle's; reward 0, terminated False, truncated from the step count."""
import numpy as np
import pytest

from gpu_util import assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def W():
    from fastatari import workloads
    return workloads


@pytest.fixture(scope="module")
def fa():
    import fastatari
    fastatari.load_library()
    return fastatari


def _f32(ram: np.ndarray) -> np.ndarray:
    return ram.reshape(-1).astype(np.float32) / np.float32(255.0)


def _same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    return a.dtype == np.float32 and np.array_equal(a.view(np.uint32), b.astype(np.float32).view(np.uint32))


def test_obs_to_f32_exact_for_every_byte(fa):
    from fastatari.vecenv import obs_to_f32
    src = np.tile(np.arange(256, dtype=np.uint8), 3)
    d_in, d_out = fa.DeviceBuffer(src.nbytes), fa.DeviceBuffer(src.nbytes * 4)
    d_in.upload(src)
    obs_to_f32(d_in.ptr, d_out.ptr, src.size)
    got = d_out.download().view(np.float32)
    want = src.astype(np.float32) / np.float32(255.0)
    assert _same_bits(got, want)
    assert got[0] == 0.0 and got[255] == 1.0


def test_f32_obs_of_a_large_batch(fa, W):
    """8,192 envs (4 MB of f32 observations, copied out of pinned host memory
    in several parts): every observation is the env's RAM / 255 in float32."""
    n = 8192
    env = fa.BatchAtariGpuEnv(W.rom_for(2), n).with_frameskip(2)

    def want():
        ram = env.batch().download_states_bytes()[:, 8:136].reshape(-1)
        return ram.astype(np.float32) / np.float32(255.0)

    assert _same_bits(env.reset(), want())
    for a in (3, 9, 17):
        assert _same_bits(env.step(a).observation, want()), f"action {a}"
    st = env.step_each(W.ActionStream(11, n).next())
    assert _same_bits(st.observation, want())


@pytest.mark.parametrize("cfg,frameskip", [(2, 4), (3, 2)])
def test_broadcast_steps_match_oracle(fa, orc, W, cfg, frameskip):
    rom = W.rom_for(cfg)
    n = 384
    env = fa.BatchAtariGpuEnv(rom, n).with_frameskip(frameskip)
    ref = orc.Batch(rom, n, nthreads=16)
    obs = env.reset()
    ref.reset()
    assert obs.shape == (n * 128,) and _same_bits(obs, _f32(ref.obs()))
    assert env.step_count() == 0
    for s, action in enumerate([0, 1, 17, 5, 200, 300 + 3, 12]):
        st = env.step(action)
        ref.run(np.full((frameskip, n), action & 0xFF, np.uint8))
        assert _same_bits(st.observation, _f32(ref.obs())), f"step {s} obs"
        assert st.reward == 0.0 and st.terminated is False and st.truncated is False
        assert_same(env.batch(), ref, f"step {s}")
    assert env.step_count() == 7


def test_frameskip_one_and_truncation(fa, orc, W):
    rom = W.rom_for(1)
    n = 200
    env = fa.BatchAtariGpuEnv(rom, n).with_frameskip(1).with_max_steps(3)
    ref = orc.Batch(rom, n, nthreads=16)
    env.reset()
    ref.reset()
    flags = []
    for s in range(5):
        st = env.step(s)
        ref.run(np.full((1, n), s, np.uint8))
        assert _same_bits(st.observation, _f32(ref.obs()))
        flags.append(st.truncated)
    assert flags == [False, False, True, True, True]
    env.reset()
    assert env.step_count() == 0 and env.step(0).truncated is False
    unlimited = fa.BatchAtariGpuEnv(rom, 8).with_max_steps(0)
    unlimited.reset()
    assert not any(unlimited.step(0).truncated for _ in range(4))


def test_per_env_actions_match_oracle(fa, orc, W):
    rom = W.rom_for(4, 3)
    n, frameskip = 300, 3
    env = fa.BatchAtariGpuEnv(rom, n).with_frameskip(frameskip)
    ref = orc.Batch(rom, n, nthreads=16)
    env.reset()
    ref.reset()
    acts = W.ActionStream(W.config_seed(4), n)
    for s in range(5):
        a = acts.next()
        st = env.step_each(a)
        ref.run(np.repeat(a[None, :], frameskip, axis=0))
        assert _same_bits(st.observation, _f32(ref.obs())), f"step {s}"
        assert_same(env.batch(), ref, f"step {s}")


def test_device_step_and_sorted_match_host(fa, W):
    rom = W.rom_for(2)
    n = 256
    host = fa.BatchAtariGpuEnv(rom, n)
    dev = fa.BatchAtariGpuEnv(rom, n).with_sorted(True)
    host.reset()
    dev.reset()
    acts = W.ActionStream(W.config_seed(2), n)
    d_act, d_obs = fa.DeviceBuffer(n), fa.DeviceBuffer(n * 128 * 4)
    for s in range(4):
        a = acts.next()
        want = host.step_each(a).observation
        d_act.upload(a)
        trunc = dev.step_device(d_obs.ptr, d_actions=d_act.ptr)
        dev.sync()
        assert trunc is False
        assert np.array_equal(d_obs.download().view(np.float32), want), f"step {s}"
    want = host.step(7).observation
    dev.step_device(d_obs.ptr, action=7)
    dev.sync()
    assert np.array_equal(d_obs.download().view(np.float32), want)
    assert dev.action_space() == {"kind": "discrete", "n": 18}
    assert dev.observation_space()["shape"] == (n * 128,)


def _score(ram: np.ndarray, addrs, bcd: bool) -> np.ndarray:
    s = np.zeros(ram.shape[0], np.int64)
    for a in addrs:
        b = ram[:, a].astype(np.int64)
        s = s * 100 + (b >> 4) * 10 + (b & 15) if bcd else s * 256 + b
    return s


@pytest.mark.parametrize("bcd", [False, True])
def test_ram_reader_rewards_and_termination(fa, orc, W, bcd):
    """Per-env reward = change of a multi-byte score in RAM (binary or packed
    BCD), terminated = a masked lives field at 0, against the same reading of
    the oracle's RAM; the batch reward is their sum, terminated their AND."""
    rom = W.rom_for(1)
    n, fs = 256, 2
    addrs, lives_addr, lives_shift, lives_mask = [0x00, 0x01, 0x05], 0x00, 4, 3
    env = (fa.BatchAtariGpuEnv(rom, n).with_frameskip(fs)
           .with_reader(addrs, bcd=bcd, lives_addr=lives_addr, lives_shift=lives_shift, lives_mask=lives_mask))
    ref = orc.Batch(rom, n, nthreads=16)
    env.reset()
    ref.reset()
    ram = ref.obs().reshape(n, 128)
    prev = _score(ram, addrs, bcd)
    rew, term = env.rewards()
    assert np.all(rew == 0.0)
    assert np.array_equal(term, ((ram[:, lives_addr] >> lives_shift) & lives_mask) == 0)
    acts = W.ActionStream(W.config_seed(1), n)
    nonzero = 0
    for s in range(6):
        a = acts.next()
        st = env.step_each(a)
        ref.run(np.repeat(a[None, :], fs, axis=0))
        ram = ref.obs().reshape(n, 128)
        sc = _score(ram, addrs, bcd)
        want = (sc - prev).astype(np.float32)
        prev = sc
        rew, term = env.rewards()
        assert _same_bits(rew, want), f"step {s} rewards"
        wterm = ((ram[:, lives_addr] >> lives_shift) & lives_mask) == 0
        assert np.array_equal(term, wterm), f"step {s} terminated"
        assert 0 < wterm.sum() < n
        assert st.reward == np.float32(rew.astype(np.float64).sum())
        assert st.terminated == bool(wterm.all())
        nonzero += int((rew != 0).sum())
    assert nonzero > 0
    env.without_reader()
    st = env.step(0)
    assert st.reward == 0.0 and st.terminated is False
