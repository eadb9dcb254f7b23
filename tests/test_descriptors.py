"""The per-opcode descriptor table that steers the kernels (fa_descriptor_table,
csrc/fa_decode.hpp), checked on the CPU against the documented instruction set
of tests/opcode_table.py: instruction length and base cycles per addressing
mode, and -- for every compute-class opcode and for PHA / PHP / PLA / PLP --
the one-hot words d2 / d3 select the source, the result, the destination
register, the stack-pointer step and the status bits that the instruction's
textbook definition gives, on random registers and operands (binary mode;
decimal ADC / SBC is checked end-to-end by the GPU parity tests); branches and
JMP carry their own d3 bit, the other stack / jump instructions none.  The
register selectors d4 .. d7 (v_perm byte selectors over A | X << 8 | Y << 16 |
SP << 24) are what the per-lane stage uses for the source, index and
destination registers and the stack-pointer step: the textbook check runs
through them, the index selectors are checked against the addressing modes,
and d1's slow-pass bit marks exactly JSR, RTS, RTI, BRK and JMP (ind).
"""
import ctypes

import numpy as np
import pytest

from opcode_table import OPCODES

LENGTH = {"imp": 1, "acc": 1, "imm": 2, "zp": 2, "zpx": 2, "zpy": 2, "izx": 2, "izy": 2, "rel": 2,
          "abs": 3, "abx": 3, "aby": 3, "ind": 3}
STACK_JUMP = {"JMP", "JSR", "RTS", "RTI", "BRK", "PHA", "PHP", "PLA", "PLP"}
BRANCH = {"BPL", "BMI", "BVC", "BVS", "BCC", "BCS", "BNE", "BEQ"}

# word layout (csrc/fa_decode.hpp)
D_LEN_SHIFT, D_CYC_SHIFT, D_IMM, D_CIN1_BIT, D_INV_BIT, D_DEC = 2, 4, 1 << 7, 26, 25, 1 << 29
G_SRC, G_DST, G_R_LOG, G_R_ARI, G_R_SHL, G_R_SHR, G_R_INC, G_R_SRC = 0, 4, 8, 9, 10, 11, 12, 13
G_CIN_SHL, G_CIN_SHR, G_DELTA, G_IX_X, G_IX_Y, G_IZX, G_IND = 14, 15, 16, 24, 25, 26, 27
G_BIT, G_C_SUM, G_C_B7, G_C_B0 = 28, 29, 30, 31
H_TAKE, H_SET, H_M_AND, H_M_XOR, H_M_LD = 8, 16, 24, 25, 26
H_BR, H_JMP, H_PUSH, H_PULL, H_PREG = 27, 28, 29, 30, 31
PUSH_PULL = {"PHA", "PHP", "PLA", "PLP"}
D1_SLOW = 1 << 15
WORDS = 8
FC, FZ, FI, FD, FV, FN = 0x01, 0x02, 0x04, 0x08, 0x40, 0x80


@pytest.fixture(scope="module")
def table():
    import fastatari
    L = fastatari.load_library()
    t = np.zeros(256 * WORDS, np.uint32)
    assert L.fa_descriptor_table(t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 256 * WORDS) == 0
    assert L.fa_descriptor_table(t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 1024) == -1
    return t.reshape(256, WORDS).astype(np.int64)


def perm(s0, s1, sel):
    """v_perm_b32 for the selector bytes the table uses: 0-3 pick a byte of
    s1, 4-7 a byte of s0, 0x0C gives 0."""
    src = [(s1 >> (8 * k)) & 0xFF for k in range(4)] + [(s0 >> (8 * k)) & 0xFF for k in range(4)]
    out = 0
    for i in range(4):
        b = (sel >> (8 * i)) & 0xFF
        assert b < 8 or b == 0x0C, hex(sel)
        out |= (src[b] if b < 8 else 0) << (8 * i)
    return out


def _bit(w, k):
    return -((w >> k) & 1)          # all ones (as int64) when set


def lanes_select(d, a, x, y, sp, p, v_in):
    """The per-lane compute stage's selections, evaluated with the table's
    words: returns (a, x, y, sp, p, r)."""
    d0, _, g, h, ssel, _, dsel, spstep = (int(w) for w in d)
    m = 0xFFFFFFFF
    push, pull, preg = _bit(h, H_PUSH), _bit(h, H_PULL), _bit(h, H_PREG)
    regs = a | (x << 8) | (y << 16) | (sp << 24)
    s = perm(0, regs, ssel) | ((p | 0x30) & preg & push)
    # D_RD or a pull: memory operand
    v = v_in if ((d0 & (1 << 14)) or pull) else (v_in if (d0 & D_IMM) else s)
    lg = (a & v & _bit(h, H_M_AND)) | ((a ^ v) & _bit(h, H_M_XOR)) | (v & _bit(h, H_M_LD))
    vv = v ^ (0xFF & _bit(d0, D_INV_BIT))
    cin = p & FC
    sm = s + vv + (cin | ((d0 >> D_CIN1_BIT) & 1))
    ar = sm & 0xFF
    shl = ((v << 1) | (cin & _bit(g, G_CIN_SHL))) & 0xFF
    shr = (v >> 1) | ((cin << 7) & _bit(g, G_CIN_SHR))
    inc = (v + ((g >> G_DELTA) & 0xFF)) & 0xFF
    r = ((lg & _bit(g, G_R_LOG)) | (ar & _bit(g, G_R_ARI)) | (shl & _bit(g, G_R_SHL)) | (shr & _bit(g, G_R_SHR))
         | (inc & _bit(g, G_R_INC)) | (s & _bit(g, G_R_SRC))) & m
    mb = _bit(g, G_BIT)
    nv = (v & mb) | (r & ~mb)
    vf = ((((~(s ^ vv)) & (s ^ ar)) >> 1) & FV & ~mb) | (v & FV & mb)
    cf = (((sm >> 8) & _bit(g, G_C_SUM)) | ((v >> 7) & _bit(g, G_C_B7)) | (v & _bit(g, G_C_B0))) & FC
    cand = (nv & FN) | vf | (FZ if r == 0 else 0) | cf
    p2 = ((p & ~(h & 0xFF)) | (cand & (h >> H_TAKE)) | ((h >> H_SET) & 0xFF) | (v & 0xCF & preg & pull)) & 0xFF
    rg = perm(r & m, (regs + spstep) & m, dsel)
    return (rg & 0xFF, (rg >> 8) & 0xFF, (rg >> 16) & 0xFF, rg >> 24, p2, r & 0xFF)


def textbook(mn, mode, a, x, y, sp, p, v):
    """Binary-mode effect of one instruction on (a, x, y, sp, p) and the byte it
    writes (None when it writes nothing)."""
    def zn(q, r):
        return (q & ~(FN | FZ)) | (r & 0x80) | (FZ if r == 0 else 0)
    c = p & FC
    out = None
    if mode == "acc":
        v = a
    if mn in ("ORA", "AND", "EOR"):
        a = {"ORA": a | v, "AND": a & v, "EOR": a ^ v}[mn]
        p = zn(p, a)
    elif mn in ("LDA", "LDX", "LDY"):
        if mn == "LDA": a = v
        elif mn == "LDX": x = v
        else: y = v
        p = zn(p, v)
    elif mn in ("ADC", "SBC"):
        w = v if mn == "ADC" else v ^ 0xFF
        t = a + w + c
        r = t & 0xFF
        ov = (~(a ^ w) & (a ^ r) & 0x80) != 0
        p = zn((p & ~(FC | FV)) | (FC if t > 0xFF else 0) | (FV if ov else 0), r)
        a = r
    elif mn in ("CMP", "CPX", "CPY"):
        reg = {"CMP": a, "CPX": x, "CPY": y}[mn]
        p = zn((p & ~FC) | (FC if reg >= v else 0), (reg - v) & 0xFF)
    elif mn == "BIT":
        p = (p & ~(FN | FV | FZ)) | (v & 0xC0) | (FZ if (a & v) == 0 else 0)
    elif mn in ("STA", "STX", "STY"):
        out = {"STA": a, "STX": x, "STY": y}[mn]
    elif mn in ("ASL", "ROL", "LSR", "ROR", "INC", "DEC"):
        if mn in ("ASL", "ROL"):
            r, nc = ((v << 1) | (c if mn == "ROL" else 0)) & 0xFF, v >> 7
        elif mn in ("LSR", "ROR"):
            r, nc = (v >> 1) | ((c << 7) if mn == "ROR" else 0), v & 1
        else:
            r, nc = (v + (1 if mn == "INC" else 0xFF)) & 0xFF, None
        if nc is not None:
            p = (p & ~FC) | nc
        p = zn(p, r)
        if mode == "acc":
            a = r
        else:
            out = r
    elif mn in ("INX", "INY", "DEX", "DEY"):
        d = 1 if mn[:2] == "IN" else 0xFF
        if mn[2] == "X":
            x = (x + d) & 0xFF
            p = zn(p, x)
        else:
            y = (y + d) & 0xFF
            p = zn(p, y)
    elif mn in ("TAX", "TAY", "TXA", "TYA", "TSX", "TXS"):
        src = {"A": a, "X": x, "Y": y, "S": sp}[mn[1]]
        dst = mn[2]
        if dst == "X": x = src
        elif dst == "Y": y = src
        elif dst == "A": a = src
        else: sp = src
        if mn != "TXS":
            p = zn(p, src)
    elif mn in ("CLC", "SEC", "CLI", "SEI", "CLV", "CLD", "SED"):
        bit = {"C": FC, "I": FI, "V": FV, "D": FD}[mn[2]]
        p = (p | bit) if mn[0] == "S" else (p & ~bit)
    elif mn in ("PHA", "PHP"):
        out = a if mn == "PHA" else (p | 0x30)
        sp = (sp - 1) & 0xFF
    elif mn in ("PLA", "PLP"):
        sp = (sp + 1) & 0xFF
        if mn == "PLA":
            a = v
            p = zn(p, v)
        else:
            p = v & 0xCF            # B and U are not stored in P
    else:
        assert mn == "NOP", mn
    return a, x, y, sp, p, out


def test_lengths_and_cycles_match_the_instruction_set(table):
    for op in range(256):
        d0 = int(table[op, 0])
        if op not in OPCODES:
            assert d0 & 3 == 3, f"{op:#04x} should be illegal"
            continue
        mn, mode, cyc = OPCODES[op]
        assert (d0 >> D_LEN_SHIFT) & 3 == LENGTH[mode], f"{mn} {mode} length"
        assert (d0 >> D_CYC_SHIFT) & 7 == cyc, f"{mn} {mode} cycles"
        cls = 1 if mn in BRANCH else (2 if mn in STACK_JUMP else 0)
        assert d0 & 3 == cls, f"{mn} class"
        if cls == 0:
            assert bool(d0 & D_DEC) == (mn in ("ADC", "SBC")), f"{mn} decimal flag"


def test_addressing_words_match_the_modes(table):
    for op, (mn, mode, _) in OPCODES.items():
        if mn in BRANCH or mn in STACK_JUMP:
            continue
        g = int(table[op, 2])
        want = {"zpx": {G_IX_X}, "abx": {G_IX_X}, "zpy": {G_IX_Y}, "aby": {G_IX_Y},
                "izx": {G_IND, G_IZX}, "izy": {G_IND, G_IX_Y}}.get(mode, set())
        got = {k for k in (G_IX_X, G_IX_Y, G_IZX, G_IND) if (g >> k) & 1}
        assert got == want, f"{mn} {mode}"


def test_index_selectors_match_the_modes(table):
    regs = 0x44332211                      # A = $11, X = $22, Y = $33, SP = $44
    for op in range(256):
        d4, d5, d6, d7 = (int(w) for w in table[op, 4:8])
        if op not in OPCODES:
            assert (d4, d5, d6, d7) == (0x0C0C0C0C, 0x0C0C0C0C, 0x03020100, 0), f"{op:#04x}"
            continue
        mn, mode, _ = OPCODES[op]
        ix = perm(0, regs, d5)
        if mn in BRANCH or mn in STACK_JUMP:
            assert ix == 0, f"{mn} {mode}"
            continue
        want_ea = {"zpx": 0x22, "abx": 0x22, "zpy": 0x33, "aby": 0x33, "izy": 0x33}.get(mode, 0)
        want_zp = 0x22 if mode == "izx" else 0
        assert ix == want_ea | (want_zp << 8), f"{mn} {mode}: {ix:#x}"


def test_slow_pass_bit(table):
    for op in range(256):
        d1 = int(table[op, 1])
        mn, mode = OPCODES[op][:2] if op in OPCODES else (None, None)
        slow = mn in ("JSR", "RTS", "RTI", "BRK") or (mn == "JMP" and mode == "ind")
        assert bool(d1 & D1_SLOW) == slow, f"{op:#04x}"


def test_branch_and_jump_words(table):
    for op, (mn, mode, _) in OPCODES.items():
        g, h = int(table[op, 2]), int(table[op, 3])
        if mn in BRANCH:
            assert (g, h) == (0, 1 << H_BR), mn
        elif mn == "JMP" and mode == "abs":
            assert (g, h) == (0, 1 << H_JMP), mn
        elif mn in STACK_JUMP and mn not in PUSH_PULL:
            assert (g, h) == (0, 0), f"{mn} {mode}"
        else:
            assert h >> H_BR & 3 == 0, f"{mn} {mode}"


def test_one_hot_words_give_the_textbook_effect(table):
    rng = np.random.default_rng(7)
    for op, (mn, mode, _) in sorted(OPCODES.items()):
        if mn in BRANCH or (mn in STACK_JUMP and mn not in PUSH_PULL):
            continue
        for _ in range(400):
            a, x, y, sp, v = (int(q) for q in rng.integers(0, 256, 5))
            p = int(rng.integers(0, 256)) & ~FD            # binary mode
            if mode in ("imm", "zp", "zpx", "zpy", "abs", "abx", "aby", "izx", "izy") or mn in ("PLA", "PLP"):
                operand = v
            else:
                operand = None
            got = lanes_select(table[op], a, x, y, sp, p, operand if operand is not None else 0)
            want = textbook(mn, mode, a, x, y, sp, p, v)
            assert got[:5] == want[:5], f"{mn} {mode}: regs/p {got[:5]} != {want[:5]} (a={a} x={x} y={y} sp={sp} p={p} v={v})"
            if want[5] is not None:
                assert got[5] == want[5], f"{mn} {mode}: written byte"
