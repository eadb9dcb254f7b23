"""Known-answer tests of the oracle's 6502 core, ported from the reference's
tests/cycle_counts.rs:36-366 (same programs, same expected values), run on the
FlatMemory bus (src/cpu.rs:69-90) with the reset vector at $8000."""
import pytest


def _cpu(orc, code=b"", base=0x8000, vec=0x8000, pokes=()):
    f = orc.Flat()
    f.load(base, bytes(code))
    f.poke(0xFFFC, vec & 0xFF)
    f.poke(0xFFFD, vec >> 8)
    for a, v in pokes:
        f.poke(a, v)
    f.reset()
    assert f.pc == vec
    return f


def run_one(orc, code):
    f = _cpu(orc, code)
    cy, _ = f.step()
    return f, cy


def run_n(orc, code, n):
    f = _cpu(orc, code)
    tot = 0
    for _ in range(n):
        cy, _ = f.step()
        tot += cy
    return f, tot


@pytest.mark.parametrize("code,cycles", [
    ([0xEA], 2), ([0xAA], 2), ([0xE8], 2), ([0x0A], 2),        # cycle_counts.rs:38-60
    ([0x69, 0x10], 2),                                          # :71-75
    ([0xB5, 0x10], 4),                                          # :101-105
    ([0xAD, 0x00, 0x90], 4),                                    # :109-113
])
def test_single_instruction_cycles(orc, code, cycles):
    _, cy = run_one(orc, code)
    assert cy == cycles


def test_lda_immediate(orc):                                    # :64-69
    f, cy = run_one(orc, [0xA9, 0x42])
    assert cy == 2 and f.a == 0x42


def test_lda_zero_page(orc):                                    # :79-91
    f = _cpu(orc, [0xA5, 0x42], pokes=[(0x42, 0xBE)])
    cy, _ = f.step()
    assert cy == 3 and f.a == 0xBE


@pytest.mark.parametrize("code,n,cycles", [
    ([0xE6, 0x50], 1, 5),                                       # :93-97 INC zp
    ([0xEE, 0x00, 0x90], 1, 6),                                 # :149-153 INC abs
    ([0xA2, 0x01, 0xBD, 0x00, 0x90], 2, 2 + 4),                 # :157-163
    ([0xA2, 0xFF, 0xBD, 0xFF, 0x90], 2, 2 + 5),                 # :167-173
    ([0xA0, 0x01, 0xB9, 0x00, 0x90], 2, 2 + 4),                 # :177-182
    ([0xA0, 0xFF, 0xB9, 0xFF, 0x90], 2, 2 + 5),                 # :186-191
    ([0xA2, 0xFF, 0x9D, 0xFF, 0x90], 2, 2 + 5),                 # :195-200
    ([0x18, 0xB0, 0x10], 2, 2 + 2),                             # :260-265
    ([0x08, 0x28], 2, 3 + 4),                                   # :320-324
    ([0x0E, 0x00, 0x90], 1, 6),                                 # :328-332
    ([0xA2, 0xFF, 0x1E, 0xFF, 0x90], 2, 2 + 7),                 # :336-341
])
def test_sequence_cycles(orc, code, n, cycles):
    _, tot = run_n(orc, code, n)
    assert tot == cycles


def test_jmp_absolute(orc):                                     # :115-126
    f = _cpu(orc, [0x4C, 0x00, 0x90])
    cy, _ = f.step()
    assert cy == 3 and f.pc == 0x9000


def test_jsr_rts(orc):                                          # :128-147
    f = _cpu(orc, [0x20, 0x10, 0x80], pokes=[(0x8010, 0x60)])
    c1, _ = f.step()
    assert c1 == 6 and f.pc == 0x8010
    c2, _ = f.step()
    assert c2 == 6 and f.pc == 0x8003


@pytest.mark.parametrize("ptr,cycles", [((0x00, 0x90), 5), ((0xFF, 0x90), 6)])
def test_lda_indirect_y(orc, ptr, cycles):                      # :204-238
    f = _cpu(orc, [0xA0, 0x01, 0xB1, 0x20], pokes=[(0x20, ptr[0]), (0x21, ptr[1])])
    c1, _ = f.step()
    c2, _ = f.step()
    assert c1 == 2 and c2 == cycles


def test_lda_indirect_x(orc):                                   # :242-256
    f = _cpu(orc, [0xA2, 0x01, 0xA1, 0x20], pokes=[(0x21, 0x00), (0x22, 0x90)])
    f.step()
    c, _ = f.step()
    assert c == 6


def test_branch_taken_no_cross(orc):                            # :269-275
    f, tot = run_n(orc, [0x38, 0xB0, 0x02], 2)
    assert tot == 2 + 3 and f.pc == 0x8005


def test_branch_taken_page_cross(orc):                          # :279-308
    f = _cpu(orc, b"", vec=0x80FD, pokes=[(0x80FD, 0x38), (0x80FE, 0xB0), (0x80FF, 0xFB)])
    c1, _ = f.step()
    assert c1 == 2 and f.pc == 0x80FE
    c2, _ = f.step()
    assert c2 == 4 and f.pc == 0x80FB


def test_pha_pla(orc):                                          # :312-318
    f, tot = run_n(orc, [0xA9, 0x42, 0x48, 0xA9, 0x00, 0x68], 4)
    assert tot == 2 + 3 + 2 + 4 and f.a == 0x42


def test_brk_rti(orc):                                          # :345-366
    f = _cpu(orc, [0x00], pokes=[(0xFFFE, 0x00), (0xFFFF, 0x90), (0x9000, 0x40)])
    c1, _ = f.step()
    assert c1 == 7 and f.pc == 0x9000
    c2, _ = f.step()
    assert c2 == 6 and f.pc == 0x8002
