"""The rendering oracle (src/tia.rs, src/atari.rs:107-318) pinned against
what the console hardware fixes: playfield bit order and reflection, score
and priority modes, player bit order / reflection / NUSIZ copies and sizes,
missile and ball widths, HMOVE direction, the collision-latch bit map, the
HMOVE blanking of the first 8 pixels, VBLANK, the INPT4 latch, line and frame
geometry, and the per-access clocking of the rendering bus (3 TIA clocks per
CPU cycle, writes landing after the second).  Expected values come from the
2600 documentation (Stella Programmer's Guide), not from the oracle's code.

Parity unpinned (stated only by the reference's text): the player RESPx
position formula (src/tia.rs:536-545, a wrap-around of the 228-clock line
that is not the hardware's) and the missile / ball RESMx / RESBL offset (+5,
src/tia.rs:548-550); the visible-area resizer constants (src/tia.rs:76-80)."""
import numpy as np
import pytest

BK, PF, C0, C1 = 0x80, 0x1E, 0x44, 0x88
LINE = 40                     # a visible scanline (visible_top starts at 32)
ROW = LINE - 32


def _rom():
    rom = bytearray([0xEA] * 4096)
    rom[0x000:0x003] = bytes([0x4C, 0x00, 0xF0])
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)


@pytest.fixture()
def con(orc):
    c = orc.RConsole(_rom())
    s = c.state()
    s[152:154] = 0                                   # TIA clock 0
    s[154:156] = np.frombuffer(np.uint16(LINE).tobytes(), np.uint8)
    c.set_state(s)
    c.tia_write(0x09, BK)                            # COLUBK
    c.tia_write(0x08, PF)                            # COLUPF
    c.tia_write(0x06, C0)                            # COLUP0
    c.tia_write(0x07, C1)                            # COLUP1
    return c


def _set(c, **fields):
    offs = {"pos_p0": 176, "pos_p1": 178, "pos_m0": 180, "pos_m1": 182, "pos_bl": 184}
    s = c.state()
    for k, v in fields.items():
        s[offs[k]:offs[k] + 2] = np.frombuffer(np.uint16(v).tobytes(), np.uint8)
    c.set_state(s)


def _line(c):
    """Render one full line from clock 0 and return its row."""
    s = c.state()
    assert int(s[152]) | int(s[153]) << 8 == 0
    scan = int(s[154]) | int(s[155]) << 8
    c.tia_ticks(228)
    return c.frame()[scan - 32].copy()


def _pf_pixels(bit, reflect):
    """Pixels lit by playfield bit `bit` (0..19 in display order): 4 pixels in
    the left half; the right half repeats it, or mirrors it with reflection."""
    lit = set(range(4 * bit, 4 * bit + 4))
    lit |= set(range(160 - 4 * bit - 4, 160 - 4 * bit)) if reflect else set(range(80 + 4 * bit, 84 + 4 * bit))
    return lit


@pytest.mark.parametrize("reflect", [0, 1])
def test_playfield_bit_order_and_reflection(orc, reflect):
    """PF0 bits 4..7, then PF1 bits 7..0, then PF2 bits 0..7, left to right,
    4 pixels each; CTRLPF bit 0 mirrors the right half."""
    order = [(0x0D, 1 << (4 + k)) for k in range(4)] + [(0x0E, 0x80 >> k) for k in range(8)] + \
            [(0x0F, 1 << k) for k in range(8)]
    for bit, (reg, val) in enumerate(order):
        c = orc.RConsole(_rom())
        s = c.state()
        s[152:154] = 0
        s[154:156] = np.frombuffer(np.uint16(LINE).tobytes(), np.uint8)
        c.set_state(s)
        c.tia_write(0x09, BK)
        c.tia_write(0x08, PF)
        c.tia_write(0x0A, reflect)
        c.tia_write(reg, val)
        row = _line(c)
        want = np.full(160, BK, np.uint8)
        want[sorted(_pf_pixels(bit, reflect))] = PF
        assert np.array_equal(row, want), f"playfield bit {bit}"


def test_score_mode_colours_each_half_with_its_player(con):
    con.tia_write(0x0A, 0x02)                        # CTRLPF score
    for reg, v in ((0x0D, 0xF0), (0x0E, 0xFF), (0x0F, 0xFF)):
        con.tia_write(reg, v)
    row = _line(con)
    assert (row[:80] == C0).all() and (row[80:] == C1).all()


@pytest.mark.parametrize("priority", [0, 1])
def test_playfield_priority(con, priority):
    """Players draw over the playfield unless CTRLPF bit 2 is set."""
    _set(con, pos_p0=8)                              # (a state load drops a queued playfield write)
    con.tia_write(0x0A, 0x04 if priority else 0)
    con.tia_write(0x0D, 0xF0)                        # playfield on pixels 0..15 and 80..95
    con.tia_write(0x1B, 0xFF)                        # GRP0: pixels 8..15
    row = _line(con)
    assert (row[8:16] == (PF if priority else C0)).all()
    assert (row[0:8] == PF).all() and (row[16:80] == BK).all()


@pytest.mark.parametrize("mode,copies,width", [(0, [0], 1), (1, [0, 16], 1), (2, [0, 32], 1), (3, [0, 16, 32], 1),
                                               (4, [0, 64], 1), (5, [0], 2), (6, [0, 32, 64], 1), (7, [0], 4)])
@pytest.mark.parametrize("reflect", [0, 1])
def test_player_bits_copies_and_sizes(con, mode, copies, width, reflect):
    """GRP0 bit 7 is the leftmost pixel (REFP0 bit 3 mirrors the byte); NUSIZ0
    modes 0-7 draw one, two (close / medium / wide), three copies, or one copy
    of double / quadruple width."""
    p, g = 20, 0xC4                                  # pixels 0, 1, 5 of the 8
    _set(con, pos_p0=p)
    con.tia_write(0x04, mode)
    con.tia_write(0x0B, 0x08 if reflect else 0)
    con.tia_write(0x1B, g)
    row = _line(con)
    bits = [(g >> (7 - k)) & 1 for k in range(8)]
    if reflect:
        bits = bits[::-1]
    want = np.full(160, BK, np.uint8)
    for c in copies:
        for k, b in enumerate(bits):
            if b:
                for w in range(width):
                    want[(p + c + k * width + w) % 160] = C0
    assert np.array_equal(row, want)


@pytest.mark.parametrize("size", [0, 1, 2, 3])
def test_missile_and_ball_widths(con, size):
    """NUSIZ0 bits 4-5 and CTRLPF bits 4-5 give 1, 2, 4, 8 pixels."""
    _set(con, pos_m0=30, pos_bl=100)
    con.tia_write(0x04, size << 4)
    con.tia_write(0x0A, size << 4)
    con.tia_write(0x1D, 0x02)                        # ENAM0
    con.tia_write(0x1F, 0x02)                        # ENABL
    row = _line(con)
    w = 1 << size
    want = np.full(160, BK, np.uint8)
    want[30:30 + w] = C0
    want[100:100 + w] = PF
    assert np.array_equal(row, want)


@pytest.mark.parametrize("hm,move", [(0x10, -1), (0x70, -7), (0xF0, 1), (0x80, 8), (0x00, 0)])
def test_hmove_direction_and_first_8_pixels(con, hm, move):
    """HMxx's upper nibble is a signed motion, positive to the left; HMOVE
    applies it to every object, and the line it is struck on shows the
    background colour in its first 8 pixels."""
    _set(con, pos_p0=40, pos_m1=50, pos_bl=60)
    con.tia_write(0x20, hm)
    con.tia_write(0x23, hm)
    con.tia_write(0x24, hm)
    con.tia_write(0x2A, 0)
    s = con.state()
    pos = [int(s[o]) | int(s[o + 1]) << 8 for o in (176, 182, 184)]
    assert pos == [(40 + move) % 160, (50 + move) % 160, (60 + move) % 160]
    con.tia_write(0x0D, 0xF0)                        # playfield over pixels 0..15
    row = _line(con)
    assert (row[:8] == BK).all() and (row[8:16] == PF).all()
    row2 = _line(con)                                # the next line: no blanking
    assert (row2[:16] == PF).all()


def test_collision_latch_bit_map(con):
    """CXM0P D7 M0-P1 / D6 M0-P0, CXM1P D7 M1-P0 / D6 M1-P1, CXP0FB D7 P0-PF /
    D6 P0-BL, CXP1FB D7 P1-PF / D6 P1-BL, CXM0FB D7 M0-PF / D6 M0-BL, CXM1FB
    D7 M1-PF / D6 M1-BL, CXBLPF D7 BL-PF, CXPPMM D7 P0-P1 / D6 M0-M1; CXCLR
    clears them."""
    # everything on one pixel column (x = 100, outside the playfield's PF0 part)
    _set(con, pos_p0=100, pos_p1=100, pos_m0=100, pos_m1=100, pos_bl=100)
    con.tia_write(0x1B, 0x80)
    con.tia_write(0x1C, 0x80)
    con.tia_write(0x1D, 0x02)
    con.tia_write(0x1E, 0x02)
    con.tia_write(0x1F, 0x02)
    con.tia_write(0x0E, 0xFF)                        # PF1 covers pixels 96..127 on the right half
    _line(con)
    cx = con.state()[160:168]
    assert list(cx) == [0xC0, 0xC0, 0xC0, 0xC0, 0xC0, 0xC0, 0x80, 0xC0]
    con.tia_write(0x2C, 0)
    assert not con.state()[160:168].any()
    # only M0 over P1: CXM0P bit 7 alone
    con.tia_write(0x1C, 0x80)
    con.tia_write(0x1B, 0x00)
    con.tia_write(0x1B, 0x00)                        # GRP0 and its old copy clear
    con.tia_write(0x1E, 0)
    con.tia_write(0x1F, 0)
    con.tia_write(0x0E, 0)
    _line(con)
    cx = con.state()[160:168]
    assert list(cx) == [0x80, 0, 0, 0, 0, 0, 0, 0]


def test_vblank_blanks_and_inpt4_latch(orc, con):
    """VBLANK bit 1 turns the output off (the row keeps its old pixels);
    VBLANK bit 6 latches INPT4 low once the button is pressed."""
    f = con.frame()
    f[ROW] = 0x33
    con.set_frame(f)
    con.tia_write(0x01, 0x02)
    row = _line(con)
    assert (row == 0x33).all()
    con.tia_write(0x01, 0x40)                        # latch on, blanking off
    con.step(1, 0)                                   # FIRE held
    con.tia_ticks(1)
    con.step(0, 0)                                   # released
    con.tia_ticks(1)
    s = con.state()
    assert s[232] == 0 and s[234] == 1               # inpt4_latch, input_latch_enabled
    con.tia_write(0x01, 0x00)
    assert con.state()[234] == 0


def test_line_and_frame_geometry(orc):
    """228 TIA clocks per line, 262 lines per frame; the rendering bus gives
    3 TIA clocks per CPU cycle (a NOP's 2 cycles = 6 clocks) and lands a write
    after the second clock of its cycle (STA RESBL at clock c sets the ball
    from clock c + 2 of that cycle)."""
    c = orc.RConsole(_rom())
    s = c.state()
    assert int(s[152]) == 6                          # power-on: two reset-vector reads = 6 clocks
    s[152:156] = 0
    c.set_state(s)
    c.tia_ticks(228)
    s = c.state()
    assert (int(s[152]), int(s[154])) == (0, 1)
    c.tia_ticks(228 * 261)
    s = c.state()
    assert (int(s[152]), int(s[154]), int(s[157])) == (0, 0, 1)


def test_bus_cycle_clocks(orc):
    """One bus read = 3 TIA clocks and 1 RIOT clock; a write = 2 clocks, the
    write, 1 clock, 1 RIOT clock."""
    c = orc.RConsole(_rom())
    s = c.state()
    s[152:154] = 0
    s[154:156] = 0
    c.set_state(s)
    c.bus_read(0xF000)
    assert int(c.state()[152]) == 3
    c.bus_write(0x0014, 0)                           # RESBL at clock 3 + 2 = 5
    s = c.state()
    assert int(s[152]) == 6
    assert int(s[184]) | int(s[185]) << 8 == (5 - 68 + 5) % 160
