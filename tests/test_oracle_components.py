"""Pins the oracle's RIOT timer, TIA timing/registers and bus decode.

* PIA timer (src/pia.rs:107-166): bulk tick_n(n) must equal n single-cycle
  tick() calls (src/pia.rs:80-103) whenever the value stays above 0 and no
  underflow happens inside the n cycles -- exhaustively for TIM1T/TIM8T/TIM64T
  and sampled for T1024T.  (The bulk path treats the value reaching 0 inside
  full prescaler periods as an underflow, one period early; that quirk is the
  reference's and is kept.)
  Around underflow the bulk path is compared with an independent transcription
  of the reference's second statement of it (cuda/src/pia.cuh:6-57).
* TIA (src/headless_tia.rs): 76 CPU cycles per scanline, 262 lines per frame,
  the WSYNC skip lands exactly on the line end, RESPx/HMOVE position maths.
* Bus (src/atari.rs:495-553): the 2600 address map (A12 = ROM, A7=0 = TIA,
  A9=1 = RIOT I/O, else RIOT RAM) plus the flat-array padding documented at
  src/atari.rs:386-395 (RAM mirrors outside $80-$FF/$180-$1FF read 0).
"""
import random

import numpy as np
import pytest

SEL = (1, 8, 64, 1024)


def _pia_states(tsel, rng, k):
    tval = rng.integers(0, 256, k)
    tps = rng.integers(1, tsel + 1, k)
    return np.stack([tval, tps, np.full(k, tsel), np.zeros(k, int)], 1)


@pytest.mark.parametrize("tsel", SEL)
def test_tick_n_equals_n_ticks_without_underflow(orc, tsel):
    rng = np.random.default_rng(tsel)
    if tsel <= 64:
        tv, tp, nn = np.meshgrid(np.arange(256), np.arange(1, tsel + 1), np.arange(1, 77), indexing="ij")
        st = np.stack([tv.ravel(), tp.ravel(), np.full(tv.size, tsel), np.zeros(tv.size, int)], 1)
        n = nn.ravel()
    else:
        st = _pia_states(tsel, rng, 200000)
        n = rng.integers(1, 77, len(st))
    bulk, single, uf = orc.pia_sweep(st, n)
    # The bulk path counts the value *reaching* 0 during full prescaler
    # periods as an underflow (src/pia.rs:151-161), one period before tick()
    # does, so the two statements agree while the value stays above 0.
    ok = ~uf & (single[:, 0] > 0)
    assert ok.sum() > len(ok) // 2
    assert np.array_equal(bulk[ok], single[ok])
    early = ~uf & (single[:, 0] == 0) & (bulk[:, 3] == 1)
    assert np.all(bulk[early, 0] <= 0xFF)


def _cuda_pia_tick_n(tval, tps, tsel, tuf, n):
    """Independent transcription of cuda/src/pia.cuh:6-57 (u16 arithmetic)."""
    if n == 0:
        return tval, tps, tsel, tuf
    if tval == 0 and tuf:
        return (0xFF - (n - 1)) & 0xFF, tps, tsel, tuf
    remaining = n
    if tps > 0:
        if remaining < tps:
            return tval, tps - remaining, tsel, tuf
        remaining -= tps
        tps = 0
        if tval == 0:
            tuf = 1
            tval = 0xFF
            if remaining > 0:
                tval = (0xFF - (remaining - 1)) & 0xFF
            return tval, tps, tsel, tuf
        tval -= 1
    if remaining == 0:
        return tval, tsel, tsel, tuf
    ps = tsel
    full, left = remaining // ps, remaining % ps
    if full >= tval:
        remaining -= (tval * ps) & 0xFFFF
        tuf = 1
        tval = 0xFF
        if remaining > 0:
            tval = (0xFF - (remaining - 1)) & 0xFF
    else:
        tval -= full
        tps = ps - left
    return tval, tps, tsel, tuf


def test_tick_n_matches_second_statement(orc):
    rng = np.random.default_rng(7)
    k = 60000
    tsel = rng.choice(SEL, k)
    tval = np.where(rng.random(k) < 0.3, 0, rng.integers(0, 256, k))
    tps = np.array([rng.integers(0, s + 1) for s in tsel])
    tuf = rng.integers(0, 2, k)
    n = rng.integers(0, 77, k)
    st = np.stack([tval, tps, tsel, tuf], 1)
    bulk, _, _ = orc.pia_sweep(st, n)
    for i in range(k):
        exp = _cuda_pia_tick_n(int(tval[i]), int(tps[i]), int(tsel[i]), int(tuf[i]), int(n[i]))
        assert tuple(int(x) for x in bulk[i]) == exp, (st[i], n[i])


def _fresh_state(orc):
    c = orc.Console(bytes(4096))
    return c.state()


def test_tia_76_cycles_per_line_and_262_lines(orc):
    st = _fresh_state(orc)
    rng = random.Random(3)
    ops, tot = [], 0
    while tot < 262 * 76 * 2 + 5:
        n = rng.randint(2, 7)
        ops.append((0 << 24) | n)
        tot += n
    out, _ = orc.component_run(st, ops)
    clock = int(out[152]) | (int(out[153]) << 8)
    line = int(out[154]) | (int(out[155]) << 8)
    # one crossing per tick_n: lines advance once per 76 cycles (228 clocks)
    assert clock == (tot * 3) % 228
    assert line == (tot // 76) % 262
    assert out[157] == 1            # frame_complete latched at the 262nd line


@pytest.mark.parametrize("clock", range(0, 228, 3))
def test_wsync_skip_lands_on_line_end(orc, clock):
    st = _fresh_state(orc)
    st[152], st[153] = clock & 0xFF, clock >> 8
    st[156] = 1
    out, res = orc.component_run(st, [1 << 24])
    assert int(res[0]) * 3 + clock == 228
    assert out[152] == 0 and out[153] == 0 and out[154] == 1 and out[156] == 0


@pytest.mark.parametrize("clock", range(0, 228, 3))
def test_resp_and_hmove_positions(orc, clock):
    st = _fresh_state(orc)
    st[152], st[153] = clock & 0xFF, clock >> 8
    hm = random.Random(clock).randrange(256)
    ops = [(2 << 24) | (0x10 << 8), (2 << 24) | (0x20 << 8) | hm, (2 << 24) | (0x2A << 8)]
    out, _ = orc.component_run(st, ops)
    pos = int(out[176]) | (int(out[177]) << 8)
    # RESP0: horizontal position = colour clock minus 68 HBLANK clocks plus
    # the 5-clock start delay, modulo 160; HMOVE moves left by the signed
    # high nibble of HMP0.
    start = (clock - 68 + 5) % 160
    shift = (hm >> 4) - 16 if hm & 0x80 else hm >> 4
    assert pos == (start - shift) % 160


def test_rmw_on_grp0_shifts_twice(orc):
    """INC GRP0 reads $1B (TIA read register $B -> $80), writes the old value
    back, then the new one: GRP0 is written twice, so GRP0_old holds the
    phantom value (src/cpu.rs:507 with src/headless_tia.rs:263)."""
    rom = bytearray(4096)
    # SEI; LDA #$55; STA GRP0; INC GRP0; LDA #2; STA VSYNC; JMP *
    prog = [0x78, 0xA9, 0x55, 0x85, 0x1B, 0xE6, 0x1B, 0xA9, 0x02, 0x85, 0x00, 0x4C, 0x0B, 0xF0]
    rom[: len(prog)] = bytes(prog)
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    c = orc.Console(bytes(rom))
    c.step(0)
    st = c.state()
    assert st[192] == 0x81 and st[194] == 0x80
    cyc, frame_cyc, instr, err = c.counters()
    assert instr == 6 and err == 0
    assert frame_cyc == cyc == 2 + 2 + 3 + 5 + 2 + 3


def test_bus_read_map(orc):
    rom = bytes((i * 7 + 3) & 0xFF for i in range(4096))
    c = orc.Console(rom)
    for i in range(128):
        c.bus_write(0x80 + i, i ^ 0x5A)
    st = c.state()
    for i in range(128):
        assert st[8 + i] == i ^ 0x5A
    tia_vals = {0xC: 0x80, 0xD: 0x80, 0xA: 0x80, 0xB: 0x80}
    for a in range(0x10000):
        v = c.bus_read(a)
        m = a & 0x1FFF
        if m & 0x1000:
            assert v == rom[m & 0xFFF], hex(a)
        elif not (m & 0x80):
            # TIA read registers by addr & 0x0F (collisions 0, paddles: counter
            # 0 >= paddle 0 -> 0x80, INPT4/5 high, INPT2/3 high)
            exp = tia_vals.get(m & 0xF, 0x80 if (m & 0xF) in (8, 9) else 0)
            assert v == exp, hex(a)
        elif m & 0x200:
            pass  # RIOT I/O, checked below
        elif m < 0x200:
            assert v == (m & 0x7F) ^ 0x5A, hex(a)
        else:
            assert v == 0, hex(a)   # flat-array padding (src/atari.rs:386-395)


def test_bus_riot_io_and_write_map(orc):
    c = orc.Console(bytes(4096))
    assert c.bus_read(0x280) == 0xFF          # SWCHA: inputs high, DDR 0
    assert c.bus_read(0x282) == 0x3F          # SWCHB default switches
    c.bus_write(0x281, 0xF0)                  # SWACNT
    c.bus_write(0x280, 0x0A)                  # SWCHA output
    assert c.bus_read(0x280) == (0xFF & 0x0F) | (0x0A & 0xF0)
    c.bus_write(0x296, 10)                    # TIM64T
    st = c.state()
    assert (st[142], st[144] | st[145] << 8, st[146] | st[147] << 8, st[148]) == (10, 64, 64, 0)
    # writes to $0200-$027F reach the RIOT too (shared write decode of
    # src/atari.rs:545 and cuda/src/memory_bus.cuh:71-88)
    c.bus_write(0x203, 0x77)
    assert c.state()[140] == 0x77
    # RAM write through any mirror lands in both flat mirrors
    c.bus_write(0x0C85, 0x99)
    mem = c.mem()
    assert mem[0x85] == 0x99 and mem[0x185] == 0x99 and c.state()[8 + 5] == 0x99
    # INTIM read clears the underflow flag, INSTAT does not
    st = c.state()
    st[142], st[148] = 0, 1
    c.set_state(st)
    assert c.bus_read(0x285) == 0x80 and c.state()[148] == 1
    c.bus_read(0x284)
    assert c.state()[148] == 0


@pytest.mark.parametrize("size,hot", [(8192, (0x1FF8, 2)), (16384, (0x1FF6, 4)), (32768, (0x1FF4, 8))])
def test_bankswitch_hotspots(orc, size, hot):
    rom = bytes((i >> 12) * 16 + (i & 0xF) for i in range(size))
    c = orc.Console(rom)
    st = c.state()
    assert st[217] == size // 4096 - 1        # initial bank = last (src/atari.rs:412-417)
    base, nb = hot
    for b in range(nb):
        v = c.bus_read(0xF000 | (base + b) & 0xFFF)   # read hotspot: switch before the read
        assert c.state()[217] == b and v == rom[b * 4096 + ((base + b) & 0xFFF)]
    for b in reversed(range(nb)):
        c.bus_write(0x1000 | (base + b) & 0xFFF, 0)
        assert c.state()[217] == b
    assert c.bus_read(0x1000 | ((base - 1) & 0xFFF)) is not None
    assert c.state()[217] == 0


# TIA registers on which a read-modify-write's first (old-value) write leaves
# a trace in the state; the per-lane write-back of the HIP kernel skips that
# write on every other TIA / RIOT register.
LASTING_TIA = {0x00, 0x01, 0x1B, 0x1C, 0x2A}       # VSYNC, VBLANK, GRP0, GRP1, HMOVE


def test_double_write_equals_last_write_outside_the_lasting_registers(orc):
    """Writing v then r leaves the state writing r alone leaves, on every TIA
    register except VSYNC / VBLANK / GRP0 / GRP1 / HMOVE and on every RIOT
    port and timer register (src/headless_tia.rs:200-298, src/pia.rs:202-228);
    on each of those five some pair (v, r) leaves a different state."""
    from fastatari import workloads
    c = orc.Console(workloads.rom_for(3))
    states = []
    rng = random.Random(11)
    for f in range(4):                         # mid-run states of the all-opcode ROM
        c.step(f)
        st = c.state()
        states.append(st)
        mid = st.copy()                        # mid-frame: VSYNC off, a later line, frame flag clear
        mid[154], mid[155] = rng.randrange(1, 256), 0
        mid[157], mid[158] = 0, 0
        states.append(mid)
    seen_diff = set()
    regs = [(2, a) for a in range(0x40)] + [(6, a) for a in (0x280, 0x281, 0x282, 0x283, 0x294, 0x295, 0x296, 0x297)]
    for st in states:
        for kind, addr in regs:
            for _ in range(24):
                v, r = rng.randrange(256), rng.randrange(256)
                both, _ = orc.component_run(st, [(kind << 24) | (addr << 8) | v, (kind << 24) | (addr << 8) | r])
                last, _ = orc.component_run(st, [(kind << 24) | (addr << 8) | r])
                if kind == 2 and addr in LASTING_TIA:
                    if not np.array_equal(both, last):
                        seen_diff.add(addr)
                else:
                    assert np.array_equal(both, last), f"kind {kind} addr {addr:#x} v {v:#x} r {r:#x}"
    assert seen_diff == LASTING_TIA
