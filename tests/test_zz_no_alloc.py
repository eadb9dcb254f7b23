"""fa_step_noalloc (the entry point that allocates nothing) against
fa_step_device on the same states and actions, bit-exact: states, obs and
status, over two calls in a row -- the second on fewer envs -- and a third
call on another ROM (a 2 KB Fixed ROM, mirrored into the window on the
device), so the shared scratch array is reused with other contents."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa():
    import fastatari
    fastatari.load_library()
    return fastatari


def _device(fa, host):
    h = np.ascontiguousarray(host)
    b = fa.DeviceBuffer(max(h.nbytes, 1))
    b.upload(h)
    return b


def _run_pair(fa, env, d_states, d_status, acts, n, k):
    """k frames on the first n envs: the handle (all its envs) via
    fa_step_device, the caller's states via fa_step_noalloc."""
    from fastatari import env as fe
    N = env.num_envs()
    d_act_h = _device(fa, acts)                                   # [k, N]
    d_act_n = _device(fa, np.ascontiguousarray(acts[:, :n]))      # [k, n]
    d_obs_h, d_obs_n = fa.DeviceBuffer(N * 128), fa.DeviceBuffer(n * 128)
    env.step_device(d_act_h.ptr, k, d_obs_h.ptr)
    env.sync()
    fe.step_noalloc(env._rom, d_states.ptr, n, k, d_act_n.ptr, d_obs_n.ptr, d_status.ptr)
    got = d_states.download(n * 222).reshape(n, 222)              # synchronous: waits for the null stream
    want = env.download_states_bytes()[:n]
    assert np.array_equal(got, want), f"states differ (n={n}, k={k})"
    assert np.array_equal(d_obs_n.download().reshape(n, 128), d_obs_h.download().reshape(N, 128)[:n])
    st = d_status.download(n * 4).view(np.uint32)
    assert np.array_equal(st, env.counters()["status"][:n].astype(np.uint32))


@pytest.mark.parametrize("cfg", [3, 2])
def test_noalloc_matches_step_device(fa, cfg):
    from fastatari import workloads as W
    rom = W.rom_for(cfg)
    N = 3000
    env = fa.BatchAtariGpu(rom, N)
    env.reset()
    st0 = env.download_states_bytes()
    d_states = _device(fa, st0)
    d_status = _device(fa, np.zeros(N, np.uint32))
    acts = W.ActionStream(W.config_seed(cfg), N)
    _run_pair(fa, env, d_states, d_status, np.stack([acts.next() for _ in range(3)]), N, 3)
    n2 = 1234
    _run_pair(fa, env, d_states, d_status, np.stack([acts.next() for _ in range(2)]), n2, 2)


def test_noalloc_fixed_2k_rom_and_limits(fa):
    from fastatari import env as fe
    from fastatari import workloads as W
    rom = W.rom_for(4, 0)                                         # 2 KB Fixed, WSYNC
    assert len(rom) == 2048
    N = 700
    env = fa.BatchAtariGpu(rom, N)
    env.reset()
    d_states = _device(fa, env.download_states_bytes())
    d_status = _device(fa, np.zeros(N, np.uint32))
    acts = W.ActionStream(W.config_seed(4), N)
    _run_pair(fa, env, d_states, d_status, np.stack([acts.next() for _ in range(2)]), N, 2)
    with pytest.raises(ValueError):
        fe.step_noalloc(rom, d_states.ptr, fe.NOALLOC_MAX_ENVS + 1, 1)
    with pytest.raises(ValueError):
        fe.step_noalloc(rom, d_states.ptr, N, 0)
