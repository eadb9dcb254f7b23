"""A minimal 6502 assembler used to synthesise the seeded benchmark ROMs.

Instructions are emitted by mnemonic + addressing mode; labels resolve
absolute (JMP/JSR/abs operands) and relative (branch) fixups at `link()`.
"""
from __future__ import annotations

# (mnemonic, mode) -> opcode, for the 151 documented NMOS 6502 opcodes.
_OPS = {
    "ADC": {"imm": 0x69, "zp": 0x65, "zpx": 0x75, "abs": 0x6D, "abx": 0x7D, "aby": 0x79, "izx": 0x61, "izy": 0x71},
    "AND": {"imm": 0x29, "zp": 0x25, "zpx": 0x35, "abs": 0x2D, "abx": 0x3D, "aby": 0x39, "izx": 0x21, "izy": 0x31},
    "ASL": {"acc": 0x0A, "zp": 0x06, "zpx": 0x16, "abs": 0x0E, "abx": 0x1E},
    "BCC": {"rel": 0x90}, "BCS": {"rel": 0xB0}, "BEQ": {"rel": 0xF0}, "BMI": {"rel": 0x30},
    "BNE": {"rel": 0xD0}, "BPL": {"rel": 0x10}, "BVC": {"rel": 0x50}, "BVS": {"rel": 0x70},
    "BIT": {"zp": 0x24, "abs": 0x2C},
    "BRK": {"imp": 0x00},
    "CLC": {"imp": 0x18}, "CLD": {"imp": 0xD8}, "CLI": {"imp": 0x58}, "CLV": {"imp": 0xB8},
    "CMP": {"imm": 0xC9, "zp": 0xC5, "zpx": 0xD5, "abs": 0xCD, "abx": 0xDD, "aby": 0xD9, "izx": 0xC1, "izy": 0xD1},
    "CPX": {"imm": 0xE0, "zp": 0xE4, "abs": 0xEC},
    "CPY": {"imm": 0xC0, "zp": 0xC4, "abs": 0xCC},
    "DEC": {"zp": 0xC6, "zpx": 0xD6, "abs": 0xCE, "abx": 0xDE},
    "DEX": {"imp": 0xCA}, "DEY": {"imp": 0x88},
    "EOR": {"imm": 0x49, "zp": 0x45, "zpx": 0x55, "abs": 0x4D, "abx": 0x5D, "aby": 0x59, "izx": 0x41, "izy": 0x51},
    "INC": {"zp": 0xE6, "zpx": 0xF6, "abs": 0xEE, "abx": 0xFE},
    "INX": {"imp": 0xE8}, "INY": {"imp": 0xC8},
    "JMP": {"abs": 0x4C, "ind": 0x6C},
    "JSR": {"abs": 0x20},
    "LDA": {"imm": 0xA9, "zp": 0xA5, "zpx": 0xB5, "abs": 0xAD, "abx": 0xBD, "aby": 0xB9, "izx": 0xA1, "izy": 0xB1},
    "LDX": {"imm": 0xA2, "zp": 0xA6, "zpy": 0xB6, "abs": 0xAE, "aby": 0xBE},
    "LDY": {"imm": 0xA0, "zp": 0xA4, "zpx": 0xB4, "abs": 0xAC, "abx": 0xBC},
    "LSR": {"acc": 0x4A, "zp": 0x46, "zpx": 0x56, "abs": 0x4E, "abx": 0x5E},
    "NOP": {"imp": 0xEA},
    "ORA": {"imm": 0x09, "zp": 0x05, "zpx": 0x15, "abs": 0x0D, "abx": 0x1D, "aby": 0x19, "izx": 0x01, "izy": 0x11},
    "PHA": {"imp": 0x48}, "PHP": {"imp": 0x08}, "PLA": {"imp": 0x68}, "PLP": {"imp": 0x28},
    "ROL": {"acc": 0x2A, "zp": 0x26, "zpx": 0x36, "abs": 0x2E, "abx": 0x3E},
    "ROR": {"acc": 0x6A, "zp": 0x66, "zpx": 0x76, "abs": 0x6E, "abx": 0x7E},
    "RTI": {"imp": 0x40}, "RTS": {"imp": 0x60},
    "SBC": {"imm": 0xE9, "zp": 0xE5, "zpx": 0xF5, "abs": 0xED, "abx": 0xFD, "aby": 0xF9, "izx": 0xE1, "izy": 0xF1},
    "SEC": {"imp": 0x38}, "SED": {"imp": 0xF8}, "SEI": {"imp": 0x78},
    "STA": {"zp": 0x85, "zpx": 0x95, "abs": 0x8D, "abx": 0x9D, "aby": 0x99, "izx": 0x81, "izy": 0x91},
    "STX": {"zp": 0x86, "zpy": 0x96, "abs": 0x8E},
    "STY": {"zp": 0x84, "zpx": 0x94, "abs": 0x8C},
    "TAX": {"imp": 0xAA}, "TAY": {"imp": 0xA8}, "TSX": {"imp": 0xBA}, "TXA": {"imp": 0x8A},
    "TXS": {"imp": 0x9A}, "TYA": {"imp": 0x98},
}

SIZES = {"imp": 1, "acc": 1, "imm": 2, "zp": 2, "zpx": 2, "zpy": 2, "izx": 2, "izy": 2, "rel": 2,
         "abs": 3, "abx": 3, "aby": 3, "ind": 3}

# All (opcode -> (mnemonic, mode)) pairs.
OPCODES = {op: (mn, md) for mn, modes in _OPS.items() for md, op in modes.items()}
assert len(OPCODES) == 151


def opcode(mn: str, mode: str) -> int:
    return _OPS[mn][mode]


class Asm:
    """Assembles into one 4 KB bank window mapped at `origin` (e.g. $F000)."""

    def __init__(self, origin: int = 0xF000, size: int = 4096, fill: int = 0xEA):
        self.origin = origin
        self.size = size
        self.buf = bytearray([fill] * size)
        self.pc = origin
        self.labels: dict[str, int] = {}
        self.fix: list[tuple[int, str, str]] = []   # (buffer offset, label, kind)

    # -- layout
    def org(self, addr: int):
        self.pc = addr

    def here(self) -> int:
        return self.pc

    def label(self, name: str):
        assert name not in self.labels, name
        self.labels[name] = self.pc

    def free(self) -> int:
        return self.origin + self.size - self.pc

    def _put(self, *bs: int):
        for b in bs:
            off = self.pc - self.origin
            if not 0 <= off < self.size:
                raise OverflowError(f"assembler overflow at ${self.pc:04X}")
            self.buf[off] = b & 0xFF
            self.pc += 1

    # -- instructions
    def op(self, mn: str, mode: str = "imp", arg=None):
        code = _OPS[mn][mode]
        n = SIZES[mode]
        if n == 1:
            self._put(code)
        elif mode == "rel":
            self._put(code)
            if isinstance(arg, str):
                self.fix.append((self.pc - self.origin, arg, "rel"))
                self._put(0)
            else:
                self._put(arg)
        elif n == 2:
            self._put(code, arg)
        else:
            self._put(code)
            if isinstance(arg, str):
                self.fix.append((self.pc - self.origin, arg, "abs"))
                self._put(0, 0)
            else:
                self._put(arg & 0xFF, arg >> 8)

    def raw(self, *bs: int):
        self._put(*bs)

    def word_at(self, addr: int, value: int):
        off = addr - self.origin
        self.buf[off] = value & 0xFF
        self.buf[off + 1] = (value >> 8) & 0xFF

    def link(self, extra: dict[str, int] | None = None) -> bytes:
        labels = dict(self.labels)
        if extra:
            labels.update(extra)
        for off, name, kind in self.fix:
            tgt = labels[name]
            if kind == "abs":
                self.buf[off] = tgt & 0xFF
                self.buf[off + 1] = (tgt >> 8) & 0xFF
            else:
                nxt = self.origin + off + 1
                d = tgt - nxt
                if not -128 <= d <= 127:
                    raise ValueError(f"branch to {name} out of range ({d})")
                self.buf[off] = d & 0xFF
        return bytes(self.buf)
