"""The documented NMOS 6502 instruction set, written from the MCS6500 family
hardware/programming manuals (mnemonic, addressing mode, base cycle count) and
the manual's cycle-by-cycle bus tables.  This is an independent statement used
to pin the oracle: it is not derived from the reference's Rust or CUDA code.
"""

# opcode -> (mnemonic, mode, base cycles)
OPCODES = {
    0x00: ("BRK", "imp", 7), 0x01: ("ORA", "izx", 6), 0x05: ("ORA", "zp", 3), 0x06: ("ASL", "zp", 5),
    0x08: ("PHP", "imp", 3), 0x09: ("ORA", "imm", 2), 0x0A: ("ASL", "acc", 2), 0x0D: ("ORA", "abs", 4),
    0x0E: ("ASL", "abs", 6), 0x10: ("BPL", "rel", 2), 0x11: ("ORA", "izy", 5), 0x15: ("ORA", "zpx", 4),
    0x16: ("ASL", "zpx", 6), 0x18: ("CLC", "imp", 2), 0x19: ("ORA", "aby", 4), 0x1D: ("ORA", "abx", 4),
    0x1E: ("ASL", "abx", 7), 0x20: ("JSR", "abs", 6), 0x21: ("AND", "izx", 6), 0x24: ("BIT", "zp", 3),
    0x25: ("AND", "zp", 3), 0x26: ("ROL", "zp", 5), 0x28: ("PLP", "imp", 4), 0x29: ("AND", "imm", 2),
    0x2A: ("ROL", "acc", 2), 0x2C: ("BIT", "abs", 4), 0x2D: ("AND", "abs", 4), 0x2E: ("ROL", "abs", 6),
    0x30: ("BMI", "rel", 2), 0x31: ("AND", "izy", 5), 0x35: ("AND", "zpx", 4), 0x36: ("ROL", "zpx", 6),
    0x38: ("SEC", "imp", 2), 0x39: ("AND", "aby", 4), 0x3D: ("AND", "abx", 4), 0x3E: ("ROL", "abx", 7),
    0x40: ("RTI", "imp", 6), 0x41: ("EOR", "izx", 6), 0x45: ("EOR", "zp", 3), 0x46: ("LSR", "zp", 5),
    0x48: ("PHA", "imp", 3), 0x49: ("EOR", "imm", 2), 0x4A: ("LSR", "acc", 2), 0x4C: ("JMP", "abs", 3),
    0x4D: ("EOR", "abs", 4), 0x4E: ("LSR", "abs", 6), 0x50: ("BVC", "rel", 2), 0x51: ("EOR", "izy", 5),
    0x55: ("EOR", "zpx", 4), 0x56: ("LSR", "zpx", 6), 0x58: ("CLI", "imp", 2), 0x59: ("EOR", "aby", 4),
    0x5D: ("EOR", "abx", 4), 0x5E: ("LSR", "abx", 7), 0x60: ("RTS", "imp", 6), 0x61: ("ADC", "izx", 6),
    0x65: ("ADC", "zp", 3), 0x66: ("ROR", "zp", 5), 0x68: ("PLA", "imp", 4), 0x69: ("ADC", "imm", 2),
    0x6A: ("ROR", "acc", 2), 0x6C: ("JMP", "ind", 5), 0x6D: ("ADC", "abs", 4), 0x6E: ("ROR", "abs", 6),
    0x70: ("BVS", "rel", 2), 0x71: ("ADC", "izy", 5), 0x75: ("ADC", "zpx", 4), 0x76: ("ROR", "zpx", 6),
    0x78: ("SEI", "imp", 2), 0x79: ("ADC", "aby", 4), 0x7D: ("ADC", "abx", 4), 0x7E: ("ROR", "abx", 7),
    0x81: ("STA", "izx", 6), 0x84: ("STY", "zp", 3), 0x85: ("STA", "zp", 3), 0x86: ("STX", "zp", 3),
    0x88: ("DEY", "imp", 2), 0x8A: ("TXA", "imp", 2), 0x8C: ("STY", "abs", 4), 0x8D: ("STA", "abs", 4),
    0x8E: ("STX", "abs", 4), 0x90: ("BCC", "rel", 2), 0x91: ("STA", "izy", 6), 0x94: ("STY", "zpx", 4),
    0x95: ("STA", "zpx", 4), 0x96: ("STX", "zpy", 4), 0x98: ("TYA", "imp", 2), 0x99: ("STA", "aby", 5),
    0x9A: ("TXS", "imp", 2), 0x9D: ("STA", "abx", 5), 0xA0: ("LDY", "imm", 2), 0xA1: ("LDA", "izx", 6),
    0xA2: ("LDX", "imm", 2), 0xA4: ("LDY", "zp", 3), 0xA5: ("LDA", "zp", 3), 0xA6: ("LDX", "zp", 3),
    0xA8: ("TAY", "imp", 2), 0xA9: ("LDA", "imm", 2), 0xAA: ("TAX", "imp", 2), 0xAC: ("LDY", "abs", 4),
    0xAD: ("LDA", "abs", 4), 0xAE: ("LDX", "abs", 4), 0xB0: ("BCS", "rel", 2), 0xB1: ("LDA", "izy", 5),
    0xB4: ("LDY", "zpx", 4), 0xB5: ("LDA", "zpx", 4), 0xB6: ("LDX", "zpy", 4), 0xB8: ("CLV", "imp", 2),
    0xB9: ("LDA", "aby", 4), 0xBA: ("TSX", "imp", 2), 0xBC: ("LDY", "abx", 4), 0xBD: ("LDA", "abx", 4),
    0xBE: ("LDX", "aby", 4), 0xC0: ("CPY", "imm", 2), 0xC1: ("CMP", "izx", 6), 0xC4: ("CPY", "zp", 3),
    0xC5: ("CMP", "zp", 3), 0xC6: ("DEC", "zp", 5), 0xC8: ("INY", "imp", 2), 0xC9: ("CMP", "imm", 2),
    0xCA: ("DEX", "imp", 2), 0xCC: ("CPY", "abs", 4), 0xCD: ("CMP", "abs", 4), 0xCE: ("DEC", "abs", 6),
    0xD0: ("BNE", "rel", 2), 0xD1: ("CMP", "izy", 5), 0xD5: ("CMP", "zpx", 4), 0xD6: ("DEC", "zpx", 6),
    0xD8: ("CLD", "imp", 2), 0xD9: ("CMP", "aby", 4), 0xDD: ("CMP", "abx", 4), 0xDE: ("DEC", "abx", 7),
    0xE0: ("CPX", "imm", 2), 0xE1: ("SBC", "izx", 6), 0xE4: ("CPX", "zp", 3), 0xE5: ("SBC", "zp", 3),
    0xE6: ("INC", "zp", 5), 0xE8: ("INX", "imp", 2), 0xE9: ("SBC", "imm", 2), 0xEA: ("NOP", "imp", 2),
    0xEC: ("CPX", "abs", 4), 0xED: ("SBC", "abs", 4), 0xEE: ("INC", "abs", 6), 0xF0: ("BEQ", "rel", 2),
    0xF1: ("SBC", "izy", 5), 0xF5: ("SBC", "zpx", 4), 0xF6: ("INC", "zpx", 6), 0xF8: ("SED", "imp", 2),
    0xF9: ("SBC", "aby", 4), 0xFD: ("SBC", "abx", 4), 0xFE: ("INC", "abx", 7),
}

WRITE_OPS = {"STA", "STX", "STY"}
RMW_OPS = {"ASL", "LSR", "ROL", "ROR", "INC", "DEC"}
BRANCHES = {"BPL", "BMI", "BVC", "BVS", "BCC", "BCS", "BNE", "BEQ"}
# Non-control-flow opcodes (everything except branches, jumps, calls,
# returns and BRK) -- used by the workload generators.
CONTROL = BRANCHES | {"JMP", "JSR", "RTS", "RTI", "BRK"}


def kind(mn: str) -> str:
    if mn in WRITE_OPS:
        return "W"
    if mn in RMW_OPS:
        return "M"
    return "R"


def expected_trace(op, pc, mem, x, y, sp, p):
    """Per-cycle bus accesses (kind, address) from the manual's tables for the
    instruction at `pc`; `mem(addr)` supplies operand / pointer bytes.
    Written-value checks are left to the caller."""
    mn, mode, _ = OPCODES[op]
    k = kind(mn)
    t = [("R", pc)]
    b1 = mem((pc + 1) & 0xFFFF)
    b2 = mem((pc + 2) & 0xFFFF)

    def data(addr):
        if k == "R":
            t.append(("R", addr))
        elif k == "W":
            t.append(("W", addr))
        else:
            t.extend([("R", addr), ("W", addr), ("W", addr)])

    if mode in ("imp", "acc"):
        t.append(("R", (pc + 1) & 0xFFFF))
        s = 0x100 | sp
        if mn in ("PHA", "PHP"):
            t.append(("W", s))
        elif mn in ("PLA", "PLP"):
            t += [("R", s), ("R", 0x100 | ((sp + 1) & 0xFF))]
        elif mn == "RTS":
            lo = mem(0x100 | ((sp + 1) & 0xFF))
            hi = mem(0x100 | ((sp + 2) & 0xFF))
            t += [("R", s), ("R", 0x100 | ((sp + 1) & 0xFF)), ("R", 0x100 | ((sp + 2) & 0xFF)),
                  ("R", (hi << 8) | lo)]
        elif mn == "RTI":
            t += [("R", s)] + [("R", 0x100 | ((sp + i) & 0xFF)) for i in (1, 2, 3)]
        elif mn == "BRK":
            t += [("W", s), ("W", 0x100 | ((sp - 1) & 0xFF)), ("W", 0x100 | ((sp - 2) & 0xFF)),
                  ("R", 0xFFFE), ("R", 0xFFFF)]
    elif mode == "imm":
        t.append(("R", (pc + 1) & 0xFFFF))
    elif mode == "zp":
        t.append(("R", (pc + 1) & 0xFFFF))
        data(b1)
    elif mode in ("zpx", "zpy"):
        t += [("R", (pc + 1) & 0xFFFF), ("R", b1)]
        data((b1 + (x if mode == "zpx" else y)) & 0xFF)
    elif mode == "abs":
        t += [("R", (pc + 1) & 0xFFFF)]
        if mn == "JSR":
            t += [("R", 0x100 | sp), ("W", 0x100 | sp), ("W", 0x100 | ((sp - 1) & 0xFF)),
                  ("R", (pc + 2) & 0xFFFF)]
        else:
            t += [("R", (pc + 2) & 0xFFFF)]
            if mn != "JMP":
                data((b2 << 8) | b1)
    elif mode in ("abx", "aby"):
        t += [("R", (pc + 1) & 0xFFFF), ("R", (pc + 2) & 0xFFFF)]
        base = (b2 << 8) | b1
        eff = (base + (x if mode == "abx" else y)) & 0xFFFF
        wrong = (base & 0xFF00) | (eff & 0xFF)
        if k != "R" or (base & 0xFF00) != (eff & 0xFF00):
            t.append(("R", wrong))
        data(eff)
    elif mode == "izx":
        t += [("R", (pc + 1) & 0xFFFF), ("R", b1)]
        zp = (b1 + x) & 0xFF
        t += [("R", zp), ("R", (zp + 1) & 0xFF)]
        data((mem((zp + 1) & 0xFF) << 8) | mem(zp))
    elif mode == "izy":
        t += [("R", (pc + 1) & 0xFFFF), ("R", b1), ("R", (b1 + 1) & 0xFF)]
        base = (mem((b1 + 1) & 0xFF) << 8) | mem(b1)
        eff = (base + y) & 0xFFFF
        if k != "R" or (base & 0xFF00) != (eff & 0xFF00):
            t.append(("R", (base & 0xFF00) | (eff & 0xFF)))
        data(eff)
    elif mode == "ind":
        ptr = (b2 << 8) | b1
        t += [("R", (pc + 1) & 0xFFFF), ("R", (pc + 2) & 0xFFFF), ("R", ptr),
              ("R", (ptr & 0xFF00) | ((ptr + 1) & 0xFF))]
    elif mode == "rel":
        t.append(("R", (pc + 1) & 0xFFFF))
        flag = {"BPL": (0x80, 0), "BMI": (0x80, 1), "BVC": (0x40, 0), "BVS": (0x40, 1),
                "BCC": (0x01, 0), "BCS": (0x01, 1), "BNE": (0x02, 0), "BEQ": (0x02, 1)}[mn]
        taken = bool(p & flag[0]) == bool(flag[1])
        if taken:
            nxt = (pc + 2) & 0xFFFF
            off = b1 - 256 if b1 >= 128 else b1
            tgt = (nxt + off) & 0xFFFF
            t.append(("R", nxt))
            if (nxt & 0xFF00) != (tgt & 0xFF00):
                t.append(("R", (nxt & 0xFF00) | (tgt & 0xFF)))
    return t
