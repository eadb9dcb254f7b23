"""GPU parity of the rendering console (fa_render_kernel) with the rendering
oracle: 160 x 192 palette-index frames, 256-byte render states and counters
after every launch, bit for bit.  Sizes: 1,000 envs x 50 frames on the
config-1 ROM and on a graphics kernel, 256 envs x 10 frames on the other ROM
classes and on frames cut by run_frame's cycle cap, ragged workgroups, state
round trips, device-side actions, error statuses, and a 65,536-env run checked
through a size-independent property (identical inputs give identical frames,
equal to one oracle console's)."""
import numpy as np
import pytest

from test_render_cpu import _illegal_rom, _vsync_held_rom

pytestmark = pytest.mark.gpu


def _render():
    from fastatari.render import BatchAtariRender
    return BatchAtariRender


def _check(tag, g, frames, r, n):
    s1, s2 = g.download_states(), r.states()
    if not np.array_equal(s1, s2):
        i = int(np.nonzero((s1 != s2).any(1))[0][0])
        pytest.fail(f"{tag}: env {i} state differs at offsets {np.nonzero(s1[i] != s2[i])[0][:16].tolist()}")
    c = g.counters()
    c1 = np.stack([c["cpu_cycles"], c["frame_cycles"], c["instructions"], c["status"].astype(np.uint64)], 1)
    assert np.array_equal(c1, r.counters()), f"{tag}: counters differ"
    f2 = r.frames()
    if not np.array_equal(frames, f2):
        i = int(np.nonzero((frames != f2).reshape(n, -1).any(1))[0][0])
        ys, xs = np.nonzero(frames[i] != f2[i])
        pytest.fail(f"{tag}: env {i} frame differs at {len(ys)} pixels, first row {ys[0]} x {xs[0]}")


def run_parity(orc, rom, n, chunks, block=None, seed=2):
    print(f"render parity: n={n} chunks={chunks} block={block} rom={len(rom)} B", flush=True)
    g = _render()(rom, n)
    if block:
        g.set_block_size(block)
    r = orc.RBatch(rom, n, nthreads=16)
    frames = g.reset()
    r.reset()
    _check("reset", g, frames, r, n)
    rng = np.random.default_rng(seed)
    for j, k in enumerate(chunks):
        a = rng.integers(0, 18, size=(k, n), dtype=np.uint8)
        frames = g.step_multi(a.reshape(-1), k)
        r.run(a)
        _check(f"launch {j} (k={k})", g, frames, r, n)
    return g, frames


def test_config1_rom_1000_envs_50_frames(orc):
    from fastatari import workloads as W
    run_parity(orc, W.rom_for(1), 1000, [1, 1, 8, 40])


def test_graphics_rom_1000_envs_50_frames(orc):
    from fastatari import workloads as W
    g, frames = run_parity(orc, W.gen_render_rom(100), 1000, [1, 1, 8, 40])
    assert len(np.unique(frames)) >= 5


@pytest.mark.parametrize("name", ["config2_f8", "config3_f6_branchy", "mix_fixed2k", "mix_f4", "graphics_no_vsync",
                                  "graphics_seed_103"])
def test_rom_classes_256_envs(orc, name):
    from fastatari import workloads as W
    mix = W.mix_roms(W.config_seed(4))
    rom = {"config2_f8": W.rom_for(2), "config3_f6_branchy": W.rom_for(3), "mix_fixed2k": mix[0], "mix_f4": mix[3],
           "graphics_no_vsync": W.gen_render_rom(7, vsync=False), "graphics_seed_103": W.gen_render_rom(103)}[name]
    run_parity(orc, rom, 256, [1, 3, 6])


@pytest.mark.parametrize("block", [128, 192, 256])
def test_block_sizes_and_ragged_workgroups(orc, block):
    from fastatari import workloads as W
    run_parity(orc, W.gen_render_rom(104), 300, [2, 3], block=block)


def test_illegal_opcode_and_held_vsync(orc):
    g, _ = run_parity(orc, _illegal_rom(), 70, [2, 2, 3])
    assert (g.counters()["status"] == 1).all()
    g, _ = run_parity(orc, _vsync_held_rom(), 3, [1])
    assert (g.counters()["status"] == 2).all()


def test_state_and_frame_round_trip(orc):
    """Downloaded states and frames, uploaded into a fresh handle, continue
    exactly as the handle they came from (here mid-picture: frames cut by the
    cycle cap)."""
    from fastatari import workloads as W
    rom = W.gen_render_rom(8, vsync=False)
    n = 130
    a = np.random.default_rng(5).integers(0, 18, size=(6, n), dtype=np.uint8)
    g1 = _render()(rom, n)
    g1.reset()
    g1.step_multi(a[:3].reshape(-1), 3)
    g2 = _render()(rom, n)
    g2.upload_states(g1.download_states())
    g2.upload_frames(g1.framebuffer())
    f1 = g1.step_multi(a[3:].reshape(-1), 3)
    f2 = g2.step_multi(a[3:].reshape(-1), 3)
    assert np.array_equal(f1, f2)
    assert np.array_equal(g1.download_states(), g2.download_states())
    bad = g1.download_states()
    bad[0, 176] = 200                       # pos_p0 above 159
    with pytest.raises(Exception):
        g2.upload_states(bad)


def test_device_actions_match_host_actions():
    from fastatari import DeviceBuffer
    from fastatari import workloads as W
    rom = W.gen_render_rom(9)
    n, k = 200, 4
    a = np.random.default_rng(6).integers(0, 18, size=(k, n), dtype=np.uint8)
    g1, g2 = _render()(rom, n), _render()(rom, n)
    g1.reset()
    g2.reset()
    f1 = g1.step_multi(a.reshape(-1), k)
    buf = DeviceBuffer(a.size)
    buf.upload(a.reshape(-1))
    g2.step_device(buf.ptr, k)
    g2.sync()
    assert np.array_equal(f1, g2.framebuffer())
    assert np.array_equal(g1.download_states(), g2.download_states())


def test_argument_errors_and_empty_batch():
    from fastatari import workloads as W
    rom = W.rom_for(1)
    g = _render()(rom, 8)
    with pytest.raises(ValueError):
        g.step_multi(np.zeros(7, np.uint8), 1)
    with pytest.raises(ValueError):
        g.set_block_size(96)
    e = _render()(rom, 0)
    assert e.reset().shape == (0, 192, 160)
    assert e.step_multi(np.zeros(0, np.uint8), 2).shape == (0, 192, 160)


def test_65536_envs_identical_inputs_give_the_oracle_frame(orc):
    """Full size: every env gets the same actions, so every frame, state and
    counter must equal env 0's, and env 0 must equal one oracle console."""
    from fastatari import workloads as W
    rom = W.gen_render_rom(100)
    n, k = 65536, 3
    print(f"render 65536 envs: k={k}", flush=True)
    g = _render()(rom, n)
    g.reset()
    acts = [3, 11, 0]
    frames = g.step_multi(np.repeat(np.array(acts, np.uint8), n), k)
    for lo in range(0, n, 8192):
        assert (frames[lo:lo + 8192] == frames[0]).all()
    s = g.download_states()
    assert (s == s[0]).all()
    ref = orc.RBatch(rom, 1)
    ref.reset()
    ref.run(np.array(acts, np.uint8).reshape(k, 1))
    assert np.array_equal(frames[0], ref.frames()[0])
    assert np.array_equal(s[0], ref.states()[0])
