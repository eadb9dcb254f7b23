"""Independent pins of the oracle's frame-loop caps, the HeadlessBus
flat-array read quirk, the action table and JSR's operand order.

Each expectation here comes from outside the oracle's code: closed-form cycle
arithmetic of a two-instruction loop, the documented 2600 address map and
port bits, ALE's published 18-action order, and the 6502's documented JSR
bus sequence.  A mistake in the oracle (a >= for a >, a cap tested before the
cycles are added, a swapped port bit, an operand read before the pushes)
fails one of them."""
import numpy as np
import pytest

from romkit import jsr_rom


def _rom_with(prog, at=0xF000):
    rom = bytearray([0xEA] * 4096)
    rom[at - 0xF000: at - 0xF000 + len(prog)] = bytes(prog)
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)


def _first_total_above(cap, start_cycles, pattern):
    """Cycle totals grow by `pattern` repeated; returns (total, instructions)
    at the first total strictly above `cap` (the frame loop tests its cap
    after every instruction)."""
    total, k = start_cycles, 0
    # closed form: whole periods first, then step through at most one period
    per = sum(pattern)
    periods = max(0, (cap - total) // per - 1)
    total += periods * per
    k += periods * len(pattern)
    i = 0
    while total <= cap:
        total += pattern[i % len(pattern)]
        i += 1
        k += 1
    return total, k


def test_loop_b_cap_closed_form(orc):
    """src/atari.rs:592-595: loop B breaks once the frame's cycles exceed
    100,000, tested after each instruction.  `INC $80; JMP $F000` (5 + 3
    cycles) from power-on ends its first frame at the first total above
    100,000 (= 100,005) with $80 counting the INCs; the second frame starts at
    the JMP, so its totals run 3, 8, 11, ... and it ends at 100,003."""
    rom = _rom_with([0xE6, 0x80, 0x4C, 0x00, 0xF0])
    c = orc.Console(rom)
    t1, k1 = _first_total_above(100000, 0, [5, 3])
    assert t1 == 100005 and k1 % 2 == 1                      # the cap lands right after an INC
    assert c.step(0, 1) == t1
    incs = (k1 + 1) // 2
    assert int(c.ram()[0]) == incs & 0xFF
    t2, k2 = _first_total_above(100000, 0, [3, 5])
    assert t2 == 100003 and k2 % 2 == 1                      # ... and then right after a JMP
    assert c.step(0, 1) == t2
    incs += k2 // 2
    assert int(c.ram()[0]) == incs & 0xFF
    cpu_cycles, last, instr, err = c.counters()
    assert (cpu_cycles, last, instr, err) == (t1 + t2, t2, k1 + k2, 0)


def test_loop_a_cap_closed_form(orc):
    """VSYNC held forever: the reference's loop A (src/atari.rs:588-590) never
    returns; the oracle and the kernel stop it at the first total above 10^6
    cycles with status 2 (a documented deviation).  `LDA #2; STA VSYNC` ends
    frame 1 after 5 cycles (VSYNC rises in loop B); frame 2 then loops
    `INC $80; JMP` with VSYNC held."""
    rom = _rom_with([0xA9, 0x02, 0x85, 0x00, 0xE6, 0x80, 0x4C, 0x04, 0xF0])
    c = orc.Console(rom)
    assert c.step(0, 1) == 5
    t, k = _first_total_above(1000000, 0, [5, 3])
    assert t == 1000005
    assert c.step(0, 1) == t
    cpu_cycles, last, instr, err = c.counters()
    assert (last, err) == (t, 2)
    assert int(c.ram()[0]) == ((k + 1) // 2) & 0xFF
    # a frozen env runs no further frames
    assert c.step(0, 1) == 0


def _padding_addresses():
    """13-bit addresses that HeadlessBus::read serves from its flat array
    (src/atari.rs:503-507: (a & 0x1080) != 0 and (a & 0x1280) != 0x0280)
    outside the regions the array holds: RAM at $80-$FF and $180-$1FF, the
    ROM window at $1000-$1FFF."""
    a = np.arange(0x2000)
    flat = ((a & 0x1080) != 0) & ((a & 0x1280) != 0x0280)
    held = ((a >= 0x80) & (a < 0x100)) | ((a >= 0x180) & (a < 0x200)) | (a >= 0x1000)
    return a[flat & ~held]


def test_flat_array_padding_reads_zero_after_random_writes(orc):
    """Brute force over the whole 13-bit space: after random writes to every
    address, every flat-array byte outside RAM ($80-$FF, $180-$1FF) and the
    ROM window stays 0, reads of those addresses return 0 (on the console
    they would mirror RAM: $0480 is RAM $80), RAM reads at $80 / $180 agree
    with each other, and a write to any RAM mirror ((a & 0x280) == 0x080)
    lands in both."""
    rom = _rom_with([0x4C, 0x00, 0xF0])
    c = orc.Console(rom)
    rng = np.random.default_rng(11)
    pad = _padding_addresses()
    assert len(pad) > 0 and 0x480 in pad
    for rnd in range(3):
        for a in rng.permutation(0x2000):
            a = int(a)
            if (a & 0x1280) == 0x0280 and (a & 0x10):
                continue                                      # timer writes: irrelevant here, keep it running
            c.bus_write(a | (int(rng.integers(0, 8)) << 13), int(rng.integers(1, 256)))
        mem = c.mem()
        assert not mem[pad].any(), f"round {rnd}: a write reached the padding"
        for a in pad[rng.choice(len(pad), 256, replace=False)]:
            assert c.bus_read(int(a)) == 0
        assert np.array_equal(mem[0x80:0x100], mem[0x180:0x200])
        assert np.array_equal(mem[0x80:0x100], c.ram())
    for a in rng.choice([x for x in range(0x2000) if (x & 0x1280) == 0x0080], 64, replace=False):
        v = int(rng.integers(0, 256))
        c.bus_write(int(a), v)
        assert c.bus_read(0x80 | (int(a) & 0x7F)) == v and c.bus_read(0x180 | (int(a) & 0x7F)) == v


# ALE's published minimal-to-full action order (Arcade Learning Environment,
# "Action" enum: PLAYER_A_NOOP ... PLAYER_A_DOWNLEFTFIRE).
ALE_ACTIONS = ["NOOP", "FIRE", "UP", "RIGHT", "LEFT", "DOWN", "UPRIGHT", "UPLEFT", "DOWNRIGHT", "DOWNLEFT",
               "UPFIRE", "RIGHTFIRE", "LEFTFIRE", "DOWNFIRE", "UPRIGHTFIRE", "UPLEFTFIRE", "DOWNRIGHTFIRE",
               "DOWNLEFTFIRE"]
# Atari 2600 port map: SWCHA bits 7/6/5/4 = player-0 right/left/down/up,
# INPT4 bit 7 = player-0 fire; all active low.
SWCHA_BIT = {"RIGHT": 0x80, "LEFT": 0x40, "DOWN": 0x20, "UP": 0x10}


@pytest.mark.parametrize("idx", range(20))
def test_action_table_matches_ale_order_and_port_map(orc, idx):
    """set_action (src/atari.rs:571-582) through the bus the ROM reads: SWCHA
    at $0280 and INPT4 at $000C.  Expected values come from the action's ALE
    name and the console's port map; indices past 17 are NOOP
    (Action::from_index).  The paddle position (200 left / 50 right / 128)
    is fixed only by the reference's text (src/atari.rs:580); its effect on
    INPT0 is checked through the paddle-counter comparison."""
    name = ALE_ACTIONS[idx] if idx < 18 else "NOOP"
    swcha = 0xFF
    for word, bit in SWCHA_BIT.items():
        if word in name:
            swcha &= ~bit
    inpt4 = 0x00 if "FIRE" in name else 0x80
    c = orc.Console(_rom_with([0x4C, 0x00, 0xF0]))
    c.step(idx, 0)
    assert c.bus_read(0x0280) == swcha, name
    assert c.bus_read(0x000C) == inpt4, name
    st = c.state()
    paddle = 200 if "LEFT" in name else (50 if "RIGHT" in name else 128)
    assert int(st[170]) == paddle
    # INPT0 reads 0x80 once the paddle counter reaches the position
    s = st.copy()
    s[172:174] = np.frombuffer(np.uint16(paddle - 1).tobytes(), np.uint8)
    c.set_state(s)
    assert c.bus_read(0x0008) == 0x00
    s[172:174] = np.frombuffer(np.uint16(paddle).tobytes(), np.uint8)
    c.set_state(s)
    assert c.bus_read(0x0008) == 0x80


def test_jsr_reads_high_byte_after_the_pushes(orc):
    """The 6502's JSR bus sequence (opcode, low operand, stack read, push PCH,
    push PCL, high operand; src/cpu.rs:983-1005): run from RAM at $F0 with the
    stack pointer at $F3 its PCL push ($F2) overwrites the high operand before
    it is read, at $F2 the PCH push ($00) does, at $F1 / $FF neither does.
    From TIA space with SP = $2C the PCH push is a CXCLR write, so the high
    operand read from CXP0FB afterwards is 0."""
    rom = jsr_rom()
    expect = {0xF3: 1, 0xF2: 2, 0xF1: 3, 0xFF: 3}             # marker byte $80 + m
    for sp, m in expect.items():
        c = orc.Console(rom)
        c.reset_frame()
        s = c.state()
        s[8 + 0x10] = sp
        c.set_state(s)
        for f in range(1, 3):
            c.step(0, 1)
            ram = c.ram()
            got = [int(ram[k]) for k in (1, 2, 3)]
            want = [f if k == m else 0 for k in (1, 2, 3)]
            assert got == want, f"SP ${sp:02X} frame {f}: markers {got}"
    c = orc.Console(rom)
    c.reset_frame()
    s = c.state()
    s[8 + 0x10] = 0xF3
    s[3] = 0x2C
    s[4:6] = 0
    s[160:163] = [0x20, 0xC0, 0xD0]
    c.set_state(s)
    c.step(0, 1)
    ram = c.ram()
    assert [int(ram[k]) for k in (1, 2, 3)] == [1, 1, 0]
    assert not c.state()[160:168].any()                      # the CXCLR push cleared the latches
