"""GPU parity: the HIP frame step against the oracle, bit-exact on full packed
state (222 B), cumulative CPU cycles, last-frame cycles, instruction counts
and error status -- every frame for the per-frame path, at every launch
boundary for step_multi (test shapes of the reference's tests/test_cuda.rs,
tests/test_cuda_persistent.rs and tests/test_cuda_sorted.rs, on the seeded
synthetic ROMs)."""
import numpy as np
import pytest

from gpu_util import assert_same
from romkit import branch_hotspot_rom, debt_rom, io_rom, jsr_rom, lane_read_rom, poll_rom, stack_rom

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


def _pair(fa, orc, rom, n, sorted_=False, nthreads=16):
    gpu = fa.BatchAtariGpu(rom, n)
    gpu.set_sorted(sorted_)
    ref = orc.Batch(rom, n, nthreads=nthreads)
    obs = gpu.reset()
    ref.reset()
    assert np.array_equal(obs, ref.obs())
    return gpu, ref


@pytest.mark.parametrize("which", ["c1", "c2", "c3", "mix0", "mix3"])
def test_reset_matches_oracle(fa, orc, W, which):
    rom = {"c1": lambda: W.rom_for(1), "c2": lambda: W.rom_for(2), "c3": lambda: W.rom_for(3),
           "mix0": lambda: W.rom_for(4, 0), "mix3": lambda: W.rom_for(4, 3)}[which]()
    gpu, ref = _pair(fa, orc, rom, 130)
    assert_same(gpu, ref, f"reset {which}")


def test_config1_every_frame(fa, orc, W):
    """Config 1: 1,000 envs x 500 frames, per-frame step, everything compared
    every frame (tests/test_cuda.rs:75-116 shape)."""
    rom = W.rom_for(1)
    n, frames = 1000, 500
    gpu, ref = _pair(fa, orc, rom, n)
    acts = W.ActionStream(W.config_seed(1), n)
    for f in range(frames):
        a = acts.next()
        obs = gpu.step(a).reshape(n, 128)
        ref.run(a[None, :])
        assert np.array_equal(obs, ref.obs()), f"obs frame {f}"
        assert_same(gpu, ref, f"frame {f}")


@pytest.mark.parametrize("sorted_", [False, True])
def test_config2_step_multi_k16(fa, orc, W, sorted_):
    """Config 2 (F8) with step_multi K=16, compared at every launch boundary,
    and the per-frame path over the same frames."""
    rom = W.rom_for(2)
    n, launches, k = 2048, 6, 16
    gpu, ref = _pair(fa, orc, rom, n, sorted_)
    acts = W.ActionStream(W.config_seed(2), n)
    for L in range(launches):
        a = acts.frames(k)
        obs = gpu.step_multi(a, k).reshape(n, 128)
        ref.run(a)
        assert np.array_equal(obs, ref.obs()), f"obs launch {L}"
        assert_same(gpu, ref, f"launch {L}")


def test_config2_per_frame(fa, orc, W):
    rom = W.rom_for(2)
    n, frames = 1024, 64
    gpu, ref = _pair(fa, orc, rom, n)
    acts = W.ActionStream(W.config_seed(2), n)
    for f in range(frames):
        a = acts.next()
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")


@pytest.mark.parametrize("sorted_", [False, True])
def test_config3_branchy(fa, orc, W, sorted_):
    """Config 3 (F6, no WSYNC, all 151 opcodes, data-dependent branches),
    sorted and unsorted (tests/test_cuda_sorted.rs shape), every frame past
    the first 30 after reset, over which the waves' divergence grows to its
    steady level."""
    rom = W.rom_for(3)
    n, frames = 1024, 48
    gpu, ref = _pair(fa, orc, rom, n, sorted_)
    acts = W.ActionStream(W.config_seed(3), n)
    for f in range(frames):
        a = acts.next()
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")


@pytest.mark.parametrize("cfg,frames,k,checks", [(2, 1000, 16, 5), (3, 300, 10, 6)])
def test_full_size_sampled(fa, orc, W, cfg, frames, k, checks):
    """Configs 2 and 3 at their full size (65,536 envs; 1,000 F8 frames as
    step_multi K = 16, 300 branchy frames), every env on the GPU and a seeded
    sample of 1,024 envs replayed on the oracle with the same action columns:
    state, counters, status and obs equal at `checks` launch boundaries spread
    over the run (the last one at the end)."""
    rom = W.rom_for(cfg)
    n = 65536
    gpu = fa.BatchAtariGpu(rom, n)
    gpu.reset()
    acts = W.ActionStream(W.config_seed(cfg), n).frames(frames)       # [frames, n]
    sample = np.sort(np.random.default_rng(cfg).choice(n, 1024, replace=False))
    ref = orc.Batch(rom, len(sample), nthreads=16)
    ref.reset()
    d_act, d_obs = fa.DeviceBuffer(acts.nbytes), fa.DeviceBuffer(n * 128)
    d_act.upload(acts)
    bounds = [0] + [min(frames, (frames * (i + 1) // checks + k - 1) // k * k) for i in range(checks)]
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        for f in range(lo, hi, k):
            gpu.step_device(d_act.at(f * n), min(k, hi - f), d_obs.ptr)
        gpu.sync()
        ref.run(acts[lo:hi, sample])
        assert_same(gpu, ref, f"config {cfg} after frame {hi}", env_ids=sample, gpu_rows=sample)
        assert np.array_equal(d_obs.download().reshape(n, 128)[sample], ref.obs())


@pytest.mark.parametrize("case", ["identical", "few_differ"])
def test_branchy_uniform_waves(fa, orc, W, case):
    """The branchy ROM -- every documented opcode: JSR/RTS, BRK/RTI, JMP (ind),
    stack pairs, indexed and indirect modes, decimal ADC/SBC, F6 bank switches
    through hotspot reads -- with every env given the same actions: whole waves
    follow one control path, so each instruction runs in the wave-uniform mode
    with its scalar control state.  In "few_differ" one env in every 64 has its
    own actions, so its wave leaves that mode at the first branch on the
    differing data and runs on the per-lane paths from there."""
    rom = W.rom_for(3)
    n, frames = 512, 10
    gpu, ref = _pair(fa, orc, rom, n)
    base = W.ActionStream(W.config_seed(3), 1)
    other = W.ActionStream(W.config_seed(3) + 7, n)
    for f in range(frames):
        a = np.full(n, base.next()[0], np.uint8)
        o = other.next()
        if case == "few_differ":
            a[::64] = o[::64]
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"{case} frame {f}")


@pytest.mark.parametrize("member", [0, 1, 2, 3])
def test_mix_frameskip(fa, orc, W, member):
    """Each ROM of the heterogeneous mix with per-env frameskip 1-4."""
    rom = W.rom_for(4, member)
    n = 512
    gpu, ref = _pair(fa, orc, rom, n)
    acts = W.ActionStream(W.config_seed(4) + member, n)
    fs = W.frameskips(W.config_seed(4) + member, n)
    for step in range(4):
        a = acts.next()
        gpu.step_frameskip(a, fs)
        ref.run_frameskip(a, fs)
        assert_same(gpu, ref, f"member {member} step {step}")


def test_upload_states_roundtrip(fa, orc, W):
    """Upload oracle states taken mid-run (RAM seeded per env), continue both."""
    rom = W.rom_for(3)
    n = 256
    ref = orc.Batch(rom, n, nthreads=8)
    ref.reset()
    acts = W.ActionStream(99, n)
    ref.run(acts.frames(3))
    st = ref.states()
    rng = np.random.default_rng(5)
    st[:, 8:8 + 0x5F] = rng.integers(0, 256, (n, 0x5F), dtype=np.uint8)   # seed data RAM per env
    ref.set_states(st)
    gpu = fa.BatchAtariGpu(rom, n)
    gpu.upload_states(st)
    cyc = ref.counters()[:, 0].copy()
    gpu.upload_cpu_cycles(cyc)
    # instruction counters restart from zero on the GPU: compare deltas
    ins0 = ref.counters()[:, 2].copy()
    for f in range(5):
        a = acts.next()
        gpu.step(a)
        ref.run(a[None, :])
        g = gpu.download_states_bytes()
        assert np.array_equal(g, ref.states()), f"frame {f}"
        c = gpu.counters()
        rc = ref.counters()
        assert np.array_equal(c["cpu_cycles"], rc[:, 0])
        assert np.array_equal(c["instructions"], rc[:, 2] - ins0)


def _rom_with(prog, at=0xF000):
    rom = bytearray([0xEA] * 4096)
    rom[at - 0xF000: at - 0xF000 + len(prog)] = bytes(prog)
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)


def test_illegal_opcode_sets_status(fa, orc):
    """Reference panics on an illegal opcode (src/cpu.rs:1139-1141): both
    sides latch status 1 with PC past the opcode and freeze the env."""
    rom = _rom_with([0x78, 0xA2, 0x05, 0xCA, 0xD0, 0xFD, 0x02])   # SEI; LDX #5; loop DEX/BNE; .byte $02
    gpu, ref = _pair(fa, orc, rom, 64)
    assert_same(gpu, ref, "after reset")
    assert int(gpu.counters()["status"][0]) == 1
    a = np.zeros(64, np.uint8)
    gpu.step(a)
    ref.run(a[None, :])
    assert_same(gpu, ref, "frozen")


def test_frame_cap_without_vsync(fa, orc):
    """A ROM that never raises VSYNC ends each frame at the 100,000-cycle cap
    (src/atari.rs:592-595)."""
    rom = _rom_with([0xE6, 0x80, 0x4C, 0x00, 0xF0])              # INC $80; JMP $F000
    gpu, ref = _pair(fa, orc, rom, 64)
    for f in range(3):
        a = np.full(64, f, np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
    assert int(gpu.counters()["frame_cycles"][0]) > 100000


@pytest.mark.parametrize("sorted_", [False, True])
def test_frame_caps_with_staggered_lanes(fa, orc, sorted_):
    """Lanes that agree on everything but the frame's cycle count (TIA clocks
    staggered, realigned by the first WSYNC) run the wave-uniform mode
    together; the caps (src/atari.rs:592-595: loop B at 100,000 cycles, and the
    held-VSYNC cap of loop A) still end each lane's frame at its own count."""
    rom = _rom_with([0x85, 0x02, 0xE6, 0x80, 0x4C, 0x00, 0xF0])   # STA WSYNC; INC $80; JMP $F000
    n = 128
    gpu, ref = _pair(fa, orc, rom, n, sorted_=sorted_)
    s = ref.states().copy()
    for i in range(n):
        s[i, 152:154] = np.frombuffer(np.uint16(3 * ((i * 7) % 76)).tobytes(), np.uint8)   # TIA clock
        if i >= 64:
            s[i, 158] = 2                                            # VSYNC held: loop A
    ref.set_states(s)
    gpu.upload_states(s)
    spread = []
    for f in range(3):
        a = np.zeros(n, np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
        spread.append(len(set(ref.counters()[:64, 1].tolist())))
    assert spread[0] > 1                                             # the lanes' frame counts differed
    c = ref.counters()
    assert (c[64:, 3] == 2).all()                                    # loop A hit its cap: VSYNC held


def test_hotspot_mid_instruction_switch(fa, orc):
    """Fetches through the F8 hotspots switch banks mid-instruction
    (src/atari.rs:504-507): the phantom read of NOP at $FFF7 selects bank 0,
    the JMP opcode comes from bank 0 at $FFF8 and its operand fetch at $FFF9
    switches back to bank 1, so the target mixes both banks ($FF80, not
    $FF11)."""
    rom = bytearray([0xEA] * 8192)
    for b in range(2):
        o = b * 4096
        rom[o: o + 6] = bytes([0xA2, 0xFF, 0x9A, 0x4C, 0xF0, 0xFF])      # LDX #$FF; TXS; JMP $FFF0
        rom[o + 0xFF8: o + 0xFFB] = bytes([0x4C, 0x11, 0xFF] if b == 0 else [0xEA, 0x80, 0xFF])
        rom[o + 0xFFC: o + 0x1000] = bytes([0x00, 0xF0, 0x00, 0xF0])
        rom[o + 0xF80: o + 0xF83] = bytes([0x4C, 0x00, 0xF1])             # JMP $F100
        rom[o + 0xF11: o + 0xF16] = bytes([0xE6, 0x90, 0x4C, 0x00, 0xF1])  # INC $90; JMP $F100
        rom[o + 0x100: o + 0x10D] = bytes([0xA9, 0x02, 0x85, 0x00, 0xE6, 0x80 + b, 0xA9, 0x00, 0x85, 0x00,
                                           0x4C, 0x00, 0xF0])
    gpu, ref = _pair(fa, orc, bytes(rom), 64)
    for f in range(3):
        a = np.full(64, 1, np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
    assert ref.states()[0, 8 + 0x10] == 0          # $90 untouched: JMP target came from bank 1


@pytest.mark.parametrize("sorted_", [False, True])
def test_jsr_outside_the_rom_window(fa, orc, sorted_):
    """JSR run from RAM and from TIA space (the per-lane general path): the
    high operand byte is read after both pushes (src/cpu.rs:983-1005).  JSR
    at $F0 with SP = $F3 pushes its PCL onto the operand ($F2 -> target
    $F2C0); with SP = $F2 the PCH lands there ($00 -> $00C0); SP = $F1 and
    $FF leave it ($D0C0).  In TIA space (PC = $0000, opcode and low operand
    from CXM0P / CXM1P), SP = $2C makes the first push a CXCLR write, so the
    high byte read from CXP0FB is 0 after the pushes."""
    rom = jsr_rom()
    n = 160
    gpu, ref = _pair(fa, orc, rom, n, sorted_=sorted_)
    s = ref.states().copy()
    case = np.arange(n) % 5
    sp_of = {0: 0xF3, 1: 0xF2, 2: 0xF1, 3: 0xFF, 4: 0xF3}
    for i in range(n):
        s[i, 8 + 0x10] = sp_of[int(case[i])]                 # RAM $90
        if case[i] == 4:
            s[i, 3] = 0x2C                                   # SP
            s[i, 4:6] = 0                                    # PC = $0000
            s[i, 160:163] = [0x20, 0xC0, 0xD0]               # CXM0P, CXM1P, CXP0FB: JSR $D0C0
    ref.set_states(s)
    gpu.upload_states(s)
    expect = {0: (1, 0, 0), 1: (0, 1, 0), 2: (0, 0, 1), 3: (0, 0, 1), 4: (1, 0, 0)}
    for f in range(1, 4):
        a = np.zeros(n, np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
        st = ref.states()
        for c, (m1, m2, m3) in expect.items():
            row = st[case == c]
            extra = 1 if c == 4 else 0                       # the TIA-space JSR of the first frame -> $82
            assert (row[:, 8 + 1] == m1 * f).all(), f"case {c} $81"
            assert (row[:, 8 + 2] == m2 * f + extra).all(), f"case {c} $82"
            assert (row[:, 8 + 3] == m3 * f).all(), f"case {c} $83"


def test_host_obs_of_a_large_batch(fa, W):
    """Host-side steps at 20,000 envs (2.56 MB of obs, copied out of pinned
    host memory in several parts): reset, step, step_multi and
    step_frameskip return exactly each env's RAM (state bytes 8..135)."""
    n = 20000
    gpu = fa.BatchAtariGpu(W.rom_for(2), n)
    acts = W.ActionStream(W.config_seed(2), n)

    def ram():
        return gpu.download_states_bytes()[:, 8:136]

    assert np.array_equal(gpu.reset().reshape(n, 128), ram())
    for f in range(3):
        assert np.array_equal(gpu.step(acts.next()).reshape(n, 128), ram()), f"step {f}"
    obs = gpu.step_multi(np.stack([acts.next() for _ in range(4)]), 4)
    assert np.array_equal(obs.reshape(n, 128), ram())
    obs = gpu.step_frameskip(acts.next(), W.frameskips(7, n))
    assert np.array_equal(obs.reshape(n, 128), ram())
    assert not gpu.counters()["status"].any()


def test_bad_lengths_raise(fa, W):
    gpu = fa.BatchAtariGpu(W.rom_for(1), 8)
    gpu.reset()
    with pytest.raises(ValueError):
        gpu.step(np.zeros(7, np.uint8))
    with pytest.raises(ValueError):
        gpu.step_multi(np.zeros(8 * 3, np.uint8), 2)
    with pytest.raises(ValueError):
        fa.BatchAtariGpu(bytes(5000), 8)
    st = gpu.download_states_bytes()
    st[0, 152] = 1                                                   # clock not a multiple of 3
    with pytest.raises(fa.FastAtariError):
        gpu.upload_states(st)
    assert gpu.num_envs() == 8


@pytest.mark.parametrize("sorted_", [False, True])
def test_stack_instructions_at_every_stack_pointer(fa, orc, W, sorted_):
    """PHA / PHP / PLA / PLP (straight-line stage) and JSR / RTS / BRK / RTI
    (slow pass) with the stack pointer swept over the whole stack page by a
    per-env walk, so that pushes and pulls reach the TIA registers at
    $0100-$017F: every frame equal to the oracle."""
    rom = stack_rom()
    n, frames = 256, 96
    gpu, ref = _pair(fa, orc, rom, n, sorted_)
    acts = W.ActionStream(0x57AC, n)
    sps = set()
    for f in range(frames):
        a = acts.next()
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
        sps.update(int(v) for v in ref.states()[:, 8 + 0x10])   # RAM $90: this frame's stack pointer
    assert len([v for v in sps if v < 0x80]) > 64, "the walk should reach the TIA half of the stack page"


@pytest.mark.parametrize("sorted_", [False, True])
def test_branch_dummy_reads_switch_banks(fa, orc, sorted_):
    """A taken branch's dummy reads select F8 banks: the page-cross fix-up
    read of a branch inside the window (straight-line stage) and the PC read
    of a branch past it (slow pass); the targets come from the switched-to
    bank (markers $80 and $82 count, $81 and $83 stay 0)."""
    gpu, ref = _pair(fa, orc, branch_hotspot_rom(), 64, sorted_)
    for f in range(4):
        a = np.zeros(64, np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
    ram = ref.states()[0, 8:12]
    assert ram[0] > 0 and ram[2] > 0 and ram[1] == 0 and ram[3] == 0, ram


@pytest.mark.parametrize("sorted_", [False, True])
def test_tia_and_riot_registers_at_per_env_addresses(fa, orc, W, sorted_):
    """Every TIA read register (collision latches, paddles against the
    counter, INPT4/5, the constant ones, through two mirrors) and every RIOT
    register (ports, DDRs, INTIM, INSTAT and mirrors; timer writes of every
    interval and port / DDR writes) at a per-env index, with per-env VBLANK
    paddle dumps: the per-lane data read's single load and register fix-up
    and the RIOT paths, every frame equal to the oracle."""
    rom = io_rom()
    n, frames = 512, 64
    gpu, ref = _pair(fa, orc, rom, n, sorted_)
    acts = W.ActionStream(0x10A0, n)
    for f in range(frames):
        a = acts.next()
        obs = gpu.step(a).reshape(n, 128)
        ref.run(a[None, :])
        assert np.array_equal(obs, ref.obs()), f"obs frame {f}"
        assert_same(gpu, ref, f"frame {f}")


@pytest.mark.parametrize("sorted_", [False, True])
@pytest.mark.parametrize("end,acts", [("frame", "varied"), ("frame", "same"), ("endless_b", "varied"),
                                      ("endless_a", "varied")])
def test_timer_polls(fa, orc, W, end, acts, sorted_):
    """RIOT timer polls (a load of INTIM / INSTAT and a branch back to it, on
    every flag and load kind, INTIM mirrors, per-env flags, a page-crossing
    branch) in the wave-uniform mode, with lanes realigned and their frame
    cycle counts spread, at per-env timer phases on the per-lane path, and in
    polls that run into loop B's and loop A's caps: every frame equal to the
    oracle."""
    rom = poll_rom(end)
    n = 256
    gpu, ref = _pair(fa, orc, rom, n, sorted_)
    stream = W.ActionStream(0x5011, n)
    for f in range(2 if end == "endless_a" else 4):
        a = stream.next()
        if acts == "same":
            a = np.full(n, a[0], np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
    c = ref.counters()
    if end == "endless_b":
        assert (c[:, 1] > 100000).all()
    if end == "endless_a":
        assert (c[:, 3] == 2).all()


@pytest.mark.parametrize("sorted_", [False, True])
@pytest.mark.parametrize("acts", ["varied", "same"])
def test_per_env_reads_outside_ram(fa, orc, W, acts, sorted_):
    """Reads at per-env addresses while the control flow is the same on every
    env: a ROM table of the current bank, TIA mirrors and the flat array's
    padding, the RIOT ports, (zp),Y with the pointer in ROM, a page crossing
    on some envs only, and hotspot reads (bank switch) on some envs only;
    every frame equal to the oracle."""
    rom = lane_read_rom()
    n = 256
    gpu, ref = _pair(fa, orc, rom, n, sorted_)
    stream = W.ActionStream(0x7A11, n)
    for f in range(8):
        a = stream.next()
        if acts == "same":
            a = np.full(n, a[0], np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")


@pytest.mark.parametrize("sorted_", [False, True])
@pytest.mark.parametrize("acts", ["varied", "same"])
def test_page_crossings_on_some_envs(fa, orc, W, acts, sorted_):
    """Indexed reads that cross a page on some envs only (a cycle more there)
    while the control flow stays the same, followed by WSYNC, a second
    crossing, an INTIM read, RESP0 and timer writes, a line end before the
    next WSYNC, a timer poll and a VSYNC write: every frame equal to the
    oracle."""
    rom = debt_rom()
    n = 256
    gpu, ref = _pair(fa, orc, rom, n, sorted_)
    stream = W.ActionStream(0x0DE8, n)
    for f in range(8):
        a = stream.next()
        if acts == "same":
            a = np.full(n, a[0], np.uint8)
        gpu.step(a)
        ref.run(a[None, :])
        assert_same(gpu, ref, f"frame {f}")
