"""Small hand-assembled ROMs shared by the oracle and GPU tests."""


def jsr_rom():
    """Fixed 4 KB ROM whose frame loop runs `JSR $D0C0` stored in RAM at $F0
    with the stack pointer taken from RAM $90.  Where the JSR lands depends on
    whether its high operand byte ($F2) is read before or after the two
    pushes (src/cpu.rs:983-1005 reads it after): each target increments its
    own marker byte ($81: $F2C0, $82: $00C0 -> $F180, $83: $D0C0)."""
    rom = bytearray([0xEA] * 4096)

    def put(addr, prog):
        rom[addr - 0xF000: addr - 0xF000 + len(prog)] = bytes(prog)

    put(0xF000, [0x78, 0xD8, 0xA2, 0xFF, 0x9A,                     # SEI; CLD; LDX #$FF; TXS
                 0xA9, 0x4C, 0x85, 0xC0, 0xA9, 0x80, 0x85, 0xC1,   # RAM $C0: JMP $F180
                 0xA9, 0xF1, 0x85, 0xC2,
                 0xA9, 0xFF, 0x85, 0x90,                           # SP for the JSR: $FF
                 0x4C, 0x00, 0xF1])                                # JMP $F100
    put(0xF100, [0xE6, 0x80,                                       # INC $80
                 0xA9, 0x02, 0x85, 0x00, 0xA9, 0x00, 0x85, 0x00,   # VSYNC on / off: the frame boundary
                 0xA6, 0x90, 0x9A,                                 # LDX $90; TXS
                 0xA9, 0x20, 0x85, 0xF0, 0xA9, 0xC0, 0x85, 0xF1,   # $F0: JSR $D0C0
                 0xA9, 0xD0, 0x85, 0xF2,
                 0x4C, 0xF0, 0x00])                                # JMP $00F0
    put(0xF180, [0xE6, 0x82, 0x4C, 0x00, 0xF1])                   # INC $82; JMP $F100
    put(0xF2C0, [0xE6, 0x81, 0x4C, 0x00, 0xF1])                   # INC $81; JMP $F100
    put(0xF0C0, [0xE6, 0x83, 0x4C, 0x00, 0xF1])                   # INC $83; JMP $F100
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)
