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


def stack_rom():
    """Fixed 4 KB ROM for the stack instructions at every stack pointer: each
    frame takes SP from a per-env walk (SWCHA EOR $90 + $1D, written back to
    $90) and runs PHA, PHP, PHA, PLA, PLP, PLA there, so over the frames the
    pushes and pulls sweep the stack page, including $0100-$017F, where they
    write and read TIA registers (VSYNC, WSYNC, RESPx, HMOVE, ... on a push;
    collision and input latches on a pull).  Then, with SP moved into the RAM
    half ($90 | $88), it runs JSR / RTS (PHP / PLP and TSX inside the
    subroutine) and BRK / RTI (the handler pushes and pulls A), whose pushes
    land on the zero-page RAM they mirror, and restores SP = $FF."""
    rom = bytearray([0xEA] * 4096)

    def put(addr, prog):
        rom[addr - 0xF000: addr - 0xF000 + len(prog)] = bytes(prog)

    put(0xF000, [0x78, 0xD8, 0xA2, 0xFF, 0x9A,                     # SEI; CLD; LDX #$FF; TXS
                 0xA9, 0x00, 0x85, 0x90,                           # $90 = 0
                 0x4C, 0x00, 0xF1])                                # JMP $F100
    put(0xF100, [0xA9, 0x02, 0x85, 0x00, 0xA9, 0x00, 0x85, 0x00,   # VSYNC on / off: the frame boundary
                 0xAD, 0x80, 0x02, 0x45, 0x90, 0x69, 0x1D, 0x85, 0x90,   # LDA SWCHA; EOR $90; ADC #$1D; STA $90
                 0xAA, 0x9A,                                       # TAX; TXS
                 0xA5, 0x91, 0x48,                                 # LDA $91; PHA
                 0x08,                                             # PHP
                 0xA5, 0x90, 0x48,                                 # LDA $90; PHA
                 0x68, 0x85, 0x93,                                 # PLA; STA $93
                 0x28,                                             # PLP
                 0x68, 0x85, 0x94,                                 # PLA; STA $94
                 0xA5, 0x90, 0x09, 0x88, 0xAA, 0x9A,               # LDA $90; ORA #$88; TAX; TXS
                 0x20, 0x00, 0xF2,                                 # JSR $F200
                 0x00, 0xEA,                                       # BRK (+ padding byte)
                 0xE6, 0x91,                                       # INC $91
                 0xA2, 0xFF, 0x9A,                                 # LDX #$FF; TXS
                 0x4C, 0x00, 0xF1])                                # JMP $F100
    put(0xF200, [0xBA, 0x86, 0x95, 0x08, 0x28, 0x60])             # TSX; STX $95; PHP; PLP; RTS
    put(0xF300, [0x48, 0x8A, 0x85, 0x96, 0x68, 0x40])             # PHA; TXA; STA $96; PLA; RTI
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    rom[0xFFE], rom[0xFFF] = 0x00, 0xF3
    return bytes(rom)


def branch_hotspot_rom():
    """F8 ROM whose taken branches reach the bank-switch hotspots through
    their dummy reads (src/cpu.rs: a taken branch reads PC, a page-crossing
    one then reads the target's low byte in PC's page).  Bank 1: BNE at $FF6E
    (inside the window) to $FEF8 crosses from page $FF to $FE, so its fix-up
    read is $FFF8 and the target comes from bank 0 (marker $80; bank 1's
    $FEF8 would bump $81).  Bank 0: BNE at $FFF7 (past the window) reads PC =
    $FFF9 when taken, so its target $FF80 comes from bank 1 (marker $82;
    bank 0's $FF80 would bump $83).  Each frame runs both, then VSYNC."""
    rom = bytearray([0xEA] * 8192)
    for b in range(2):
        o = b * 4096

        def put(addr, prog):
            rom[o + addr - 0xF000: o + addr - 0xF000 + len(prog)] = bytes(prog)

        put(0xF000, [0xA2, 0xFF, 0x9A, 0x4C, 0x00, 0xF1])                  # LDX #$FF; TXS; JMP $F100
        put(0xF100, [0xA9, 0x02, 0x85, 0x00, 0xA9, 0x00, 0x85, 0x00,       # VSYNC on / off
                     0xAD, 0xF9, 0xFF,                                     # LDA $FFF9: select bank 1
                     0x4C, 0x6C, 0xFF])                                    # JMP $FF6C
        put(0xFF6C, [0xA2, 0x01, 0xD0, 0x88])                             # LDX #1; BNE -$78 (-> $FEF8)
        put(0xFEF8, [0xE6, 0x80 + b, 0x4C, 0xF5, 0xFF])                   # INC $80 / $81; JMP $FFF5
        put(0xFFF5, [0xA2, 0x01, 0xD0, 0x87])                             # LDX #1; BNE -$79 (-> $FF80)
        put(0xFF80, [0xE6, 0x83 - b, 0x4C, 0x00, 0xF1])                   # INC $83 / $82; JMP $F100
        rom[o + 0xFFC: o + 0x1000] = bytes([0x00, 0xF0, 0x00, 0xF0])
    return bytes(rom)


def io_rom():
    """Fixed 4 KB ROM whose TIA / RIOT accesses use a per-env index: X = a
    per-env walk ($90 + SWCHA's joystick nibble + 7 each frame) & 15 selects the TIA read register ($00,X and its $30
    mirror, and $08,X: paddles, INPT4/5 and the constant registers), the
    RIOT register ($0280,X: ports, DDRs, INTIM, INSTAT and their mirrors;
    $0284,X every line), the timer interval written ($0294,X: TIM1T ..
    T1024T for X < 12, the port / DDR registers above), and VBLANK gets a
    per-env byte (bit 7 dumps the paddles).  A frame is VSYNC, those
    accesses, then 32 WSYNC lines reading INTIM / INSTAT and the paddles."""
    rom = bytearray([0xEA] * 4096)

    def put(addr, prog):
        rom[addr - 0xF000: addr - 0xF000 + len(prog)] = bytes(prog)

    put(0xF000, [0x78, 0xD8, 0xA2, 0xFF, 0x9A, 0x4C, 0x00, 0xF1])  # SEI; CLD; LDX #$FF; TXS; JMP $F100
    put(0xF100, [0xA9, 0x02, 0x85, 0x00, 0xA9, 0x00, 0x85, 0x00,   # VSYNC on / off
                 0xAD, 0x80, 0x02, 0x4A, 0x4A, 0x4A, 0x4A,         # LDA SWCHA; LSR x 4
                 0x65, 0x90, 0x69, 0x07, 0x85, 0x90,               # ADC $90; ADC #7; STA $90
                 0x29, 0x0F, 0xAA,                                 # AND #$0F; TAX
                 0xB5, 0x00, 0x85, 0x80,                           # LDA $00,X; STA $80
                 0xB5, 0x30, 0x85, 0x81,                           # LDA $30,X; STA $81
                 0xBD, 0x80, 0x02, 0x85, 0x82,                     # LDA $0280,X; STA $82
                 0xA5, 0x90, 0x85, 0x01,                           # LDA $90; STA VBLANK
                 0x9D, 0x94, 0x02,                                 # STA $0294,X
                 0xA0, 0x20,                                       # LDY #$20
                 0x85, 0x02,                                       # STA WSYNC
                 0xBD, 0x84, 0x02, 0x45, 0x83, 0x85, 0x83,         # LDA $0284,X; EOR $83; STA $83
                 0xB5, 0x08, 0x45, 0x84, 0x85, 0x84,               # LDA $08,X; EOR $84; STA $84
                 0x88, 0xD0, 0xEE,                                 # DEY; BNE (to STA WSYNC)
                 0x4C, 0x00, 0xF1])                                # JMP $F100
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)


def poll_rom(end="frame"):
    """Fixed 4 KB ROM of RIOT timer polls: a load of INTIM / INSTAT followed by
    a branch back to it, the way a kernel waits out VBLANK and overscan.  One
    frame runs: a poll inside loop A (VSYNC held); a per-env delay (SWCHA's
    low bits) and a poll at per-env timer phases; then, with the lanes
    realigned by WSYNC and the timer rewritten, BIT INSTAT / BPL, LDA INTIM /
    BNE, LDX INTIM / BPL, BIT INTIM / BVC, LDY (an INTIM mirror) / BNE,
    LDA INTIM / BMI after an underflow; polls whose branch flag is per env
    (BIT / BNE with a per-env A; LDA / BCC) and one whose branch crosses a
    page.  `end` = "endless_b" or "endless_a" adds a BIT INSTAT / BMI loop
    that never leaves, after VSYNC drops (the frame ends at loop B's cap) or
    while VSYNC is held (loop A's cap, status VSYNC held)."""
    rom = bytearray([0xEA] * 4096)

    def put(addr, prog):
        rom[addr - 0xF000: addr - 0xF000 + len(prog)] = bytes(prog)

    endless = [0xA9, 0x01, 0x8D, 0x94, 0x02,                      # LDA #1; STA TIM1T
               0x2C, 0x85, 0x02, 0x10, 0xFB,                      # BIT INSTAT; BPL *-3
               0x2C, 0x85, 0x02, 0x30, 0xFB]                      # BIT INSTAT; BMI *-3 (flag stays set)
    put(0xF000, [0x78, 0xD8, 0xA2, 0xFF, 0x9A, 0x4C, 0x00, 0xF1])  # SEI; CLD; LDX #$FF; TXS; JMP $F100
    put(0xF100, [0xA9, 0x02, 0x85, 0x00,                          # VSYNC on
                 0xA9, 0x02, 0x8D, 0x96, 0x02,                    # LDA #2; STA TIM64T
                 0x2C, 0x85, 0x02, 0x10, 0xFB]                    # BIT INSTAT; BPL *-3 (loop A)
                + (endless if end == "endless_a" else []) +
                [0x85, 0x02, 0xA9, 0x00, 0x85, 0x00,              # STA WSYNC; VSYNC off
                 0xE6, 0x80,                                      # INC $80
                 0xAD, 0x80, 0x02, 0x29, 0x03, 0xAA,              # LDA SWCHA; AND #3; TAX
                 0xCA, 0x10, 0xFD,                                # DEX; BPL *-1: per-env delay
                 0xA9, 0x14, 0x8D, 0x96, 0x02,                    # LDA #20; STA TIM64T
                 0x2C, 0x85, 0x02, 0x10, 0xFB,                    # BIT INSTAT; BPL (per-env phases)
                 0x85, 0x02,                                      # STA WSYNC
                 0xA9, 0x1E, 0x8D, 0x95, 0x02,                    # LDA #30; STA TIM8T
                 0xAD, 0x84, 0x02, 0xD0, 0xFB,                    # LDA INTIM; BNE
                 0xA9, 0x05, 0x8D, 0x94, 0x02,                    # LDA #5; STA TIM1T
                 0xAE, 0x84, 0x02, 0x10, 0xFB,                    # LDX INTIM; BPL
                 0xA9, 0x03, 0x8D, 0x97, 0x02,                    # LDA #3; STA T1024T
                 0x2C, 0x84, 0x02, 0x50, 0xFB,                    # BIT INTIM; BVC
                 0xA9, 0x10, 0x8D, 0x95, 0x02,                    # LDA #16; STA TIM8T
                 0xAC, 0x8C, 0x02, 0xD0, 0xFB,                    # LDY $028C (INTIM); BNE
                 0xA9, 0x01, 0x8D, 0x94, 0x02,                    # LDA #1; STA TIM1T
                 0x2C, 0x85, 0x02, 0x10, 0xFB,                    # BIT INSTAT; BPL (underflow)
                 0xAD, 0x84, 0x02, 0x30, 0xFB,                    # LDA INTIM; BMI
                 0xA9, 0x20, 0x8D, 0x95, 0x02,                    # LDA #32; STA TIM8T
                 0xAD, 0x80, 0x02, 0x09, 0x01,                    # LDA SWCHA; ORA #1
                 0x2C, 0x84, 0x02, 0xD0, 0xFB,                    # BIT INTIM; BNE (Z per env)
                 0x38, 0xAD, 0x84, 0x02, 0x90, 0xFB,              # SEC; LDA INTIM; BCC (C per env)
                 0xA9, 0x04, 0x8D, 0x95, 0x02,                    # LDA #4; STA TIM8T
                 0x4C, 0xFD, 0xF2])                               # JMP $F2FD
    put(0xF2FD, [0x2C, 0x85, 0x02, 0x10, 0xFB]                    # BIT INSTAT; BPL across $F300
                + (endless if end == "endless_b" else []) +
                [0x4C, 0x00, 0xF1])                               # JMP $F100
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)


def lane_read_rom():
    """F8 ROM (both banks the same code, different data) whose reads use a
    per-env index while the control flow stays the same on every env: X is a
    per-env walk ($90 + SWCHA, written back each frame) and Y = X & $3F.  Per
    frame, after selecting bank 1 at a fixed address: LDA $F200,X (a ROM
    table of the current bank), LDA $0400,X (TIA mirrors for X < $80, the
    flat array's padding above), LDA $0280,X & 3 ports through TAX of X & 3
    (the RIOT ports), LDA ($92),Y with the pointer in ROM ($F1C0), LDA
    $F2F0,X (a page crossing on some envs only), then LDA $FFC0,X, which
    reaches the hotspots $FFF8 / $FFF9 (bank switch) on some envs and wraps to
    $0000+ on others.  Each value is folded into its own RAM byte."""
    rom = bytearray([0xEA] * 8192)
    for b in range(2):
        o = b * 4096

        def put(addr, prog):
            rom[o + addr - 0xF000: o + addr - 0xF000 + len(prog)] = bytes(prog)

        put(0xF000, [0x78, 0xD8, 0xA2, 0xFF, 0x9A,                     # SEI; CLD; LDX #$FF; TXS
                     0xA9, 0xC0, 0x85, 0x92, 0xA9, 0xF1, 0x85, 0x93,   # ($92) = $F1C0
                     0x4C, 0x00, 0xF1])                                # JMP $F100
        put(0xF100, [0xA9, 0x02, 0x85, 0x00, 0x85, 0x02,               # VSYNC on; STA WSYNC
                     0xA9, 0x00, 0x85, 0x00,                           # VSYNC off
                     0xAD, 0xF9, 0xFF,                                 # LDA $FFF9: bank 1
                     0xAD, 0x80, 0x02, 0x45, 0x90, 0x69, 0x1D, 0x85, 0x90,   # LDA SWCHA; EOR $90; ADC #$1D; STA $90
                     0xAA,                                             # TAX
                     0xBD, 0x00, 0xF2, 0x45, 0x81, 0x85, 0x81,         # LDA $F200,X; EOR $81; STA $81
                     0xBD, 0x00, 0x04, 0x45, 0x82, 0x85, 0x82,         # LDA $0400,X; EOR $82; STA $82
                     0x8A, 0x29, 0x03, 0xA8,                           # TXA; AND #3; TAY
                     0xB9, 0x80, 0x02, 0x45, 0x86, 0x85, 0x86,         # LDA $0280,Y; EOR $86; STA $86
                     0x8A, 0x29, 0x3F, 0xA8,                           # TXA; AND #$3F; TAY
                     0xB1, 0x92, 0x45, 0x83, 0x85, 0x83,               # LDA ($92),Y; EOR $83; STA $83
                     0xBD, 0xF0, 0xF2, 0x45, 0x85, 0x85, 0x85,         # LDA $F2F0,X; EOR $85; STA $85
                     0x85, 0x02,                                       # STA WSYNC
                     0xBD, 0xC0, 0xFF, 0x45, 0x84, 0x85, 0x84,         # LDA $FFC0,X; EOR $84; STA $84
                     0x4C, 0x00, 0xF1])                                # JMP $F100
        for i in range(256):                                           # per-bank data
            rom[o + 0x200 + i] = (i * 7 + 0x35 * (b + 1)) & 0xFF
            rom[o + 0x300 + i] = (i * 13 + 0x11 * (b + 3)) & 0xFF
        for i in range(64):
            rom[o + 0x1C0 + i] = (i * 5 + 0x60 * (b + 1)) & 0xFF
        rom[o + 0xFFC: o + 0x1000] = bytes([0x00, 0xF0, 0x00, 0xF0])
    return bytes(rom)


def debt_rom():
    """Fixed 4 KB ROM of indexed reads whose page crossing differs between
    envs (LDA $F2F0,X with a per-env X: a cycle more on the envs with X >=
    $10) while the control flow is the same on every env, each followed by
    something that sees the cycle or not: two crossings then WSYNC; an INTIM
    read then WSYNC; a RESP0 write; a timer write; 40 INC (a line end before
    the next WSYNC); a timer poll (BIT INSTAT / BPL); and a VSYNC write at the
    next frame's start."""
    rom = bytearray([0xEA] * 4096)

    def put(addr, prog):
        rom[addr - 0xF000: addr - 0xF000 + len(prog)] = bytes(prog)

    cross = [0xBD, 0xF0, 0xF2]                                         # LDA $F2F0,X
    wsync = [0x85, 0x02]
    prog = ([0xA9, 0x02, 0x85, 0x00, 0x85, 0x02, 0xA9, 0x00, 0x85, 0x00,   # VSYNC on; WSYNC; VSYNC off
             0xAD, 0x80, 0x02, 0x45, 0x90, 0x69, 0x1D, 0x85, 0x90, 0xAA,   # X = per-env walk
             0xA9, 0xC8, 0x8D, 0x96, 0x02]                             # LDA #200; STA TIM64T
            + cross + [0x85, 0x81] + cross + [0x45, 0x81, 0x85, 0x81] + wsync
            + cross + [0xAD, 0x84, 0x02, 0x85, 0x82] + wsync           # LDA INTIM; STA $82
            + cross + [0x85, 0x10] + wsync                             # STA RESP0
            + cross + [0xA9, 0x64, 0x8D, 0x96, 0x02] + wsync           # LDA #100; STA TIM64T
            + cross + [0xE6, 0x83] * 40 + wsync                        # 40 x INC $83
            + [0xA9, 0x03, 0x8D, 0x95, 0x02] + cross                   # LDA #3; STA TIM8T; LDA $F2F0,X
            + [0x2C, 0x85, 0x02, 0x10, 0xFB] + wsync                   # BIT INSTAT; BPL *-3
            + cross + [0x85, 0x84]                                     # STA $84
            + [0x4C, 0x00, 0xF1])                                      # JMP $F100 (VSYNC next)
    assert len(prog) < 0x1F0
    put(0xF000, [0x78, 0xD8, 0xA2, 0xFF, 0x9A, 0x4C, 0x00, 0xF1])      # SEI; CLD; LDX #$FF; TXS; JMP $F100
    put(0xF100, prog)
    for i in range(0x100):
        rom[0x2F0 + i] = (i * 11 + 3) & 0xFF
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)
