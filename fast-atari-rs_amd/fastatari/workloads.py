"""Seeded synthetic workloads: ROM generators, action and frameskip streams.

Four ROM classes are generated from a 64-bit seed:

* ``wsync``  -- a game-like frame: VSYNC, a TIM64T=43 vertical blank, 192
  visible lines of L~U[4,12] random non-control-flow instructions each ended
  by ``STA WSYNC``, an input read of SWCHA/INPT4 XOR-ed into RAM, 30 overscan
  lines and ``JMP`` back.  Fixed (2 KB / 4 KB) or bankswitched (F8 / F4) with
  a bank switch (``LDA $1FF8+b``) every `lines_per_bank` lines through
  trampolines mirrored in every bank.
* ``branchy`` -- no WSYNC: each frame sets T1024T=29 and loops over a chain of
  random basic blocks (8-24 instructions from all 151 official opcodes:
  JSR/RTS, BRK/RTI, stack pairs, indexed and indirect modes) ending in
  data-dependent BNE/BMI/BCS on RAM, hopping banks (F6) between segments,
  until INSTAT reports the timer underflow.

RAM map used by every generator:
  $80-$DE data (random writes land here), $DF a loop counter, $E0-$EF eight
  indirect pointers
  (written only by the frame prologue), $F0-$FF stack (SP starts at $FF).

Randomness: SplitMix64.  Config c uses master seed ``0xFA7A0000 + c``; env i's
action stream is seeded with ``splitmix64(seed ^ (i * 0x9E3779B97F4A7C15))`` and
each action is ``(u64 * 18) >> 64``.
"""
from __future__ import annotations

import numpy as np

from .asm6502 import OPCODES, SIZES, Asm, opcode

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15

DATA_LO, DATA_HI = 0x80, 0xDE   # $DF is the generators' own loop counter
LOOP_CTR = 0xDF
PTR_BASE = 0xE0
N_PTR = 8

VSYNC, VBLANK, WSYNC = 0x00, 0x01, 0x02
INPT4 = 0x0C
SWCHA, INTIM, INSTAT, TIM64T, T1024T = 0x280, 0x284, 0x285, 0x296, 0x297

# TIA write registers a random store / RMW may touch without changing the
# frame structure (everything except VSYNC, WSYNC).
SAFE_TIA_WRITES = [0x01, 0x03] + list(range(0x04, 0x2D))


def splitmix64(x: int) -> int:
    x = (x + GOLDEN) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


class Rng:
    """SplitMix64 stream (the generators' only source of randomness)."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def u64(self) -> int:
        self.s = (self.s + GOLDEN) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        return (self.u64() * n) >> 64

    def range(self, lo: int, hi: int) -> int:       # inclusive
        return lo + self.below(hi - lo + 1)

    def choice(self, seq):
        return seq[self.below(len(seq))]

    def coin(self, p: float = 0.5) -> bool:
        return self.u64() < int(p * (1 << 64))


def config_seed(c: int) -> int:
    return 0xFA7A0000 + c


# --------------------------------------------------------------------------
# action / frameskip streams
# --------------------------------------------------------------------------
def _splitmix_vec(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _mulhi(x: np.ndarray, m: int) -> np.ndarray:
    """(x * m) >> 64 for uint64 x and a small m."""
    hi = x >> np.uint64(32)
    lo = x & np.uint64(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        return (hi * np.uint64(m) + ((lo * np.uint64(m)) >> np.uint64(32))) >> np.uint64(32)


class ActionStream:
    """Per-env SplitMix64 action streams; `next()` yields one uint8 action
    per env (uniform over the 18 actions)."""

    def __init__(self, seed: int, n: int, env_offset: int = 0):
        ids = np.arange(env_offset, env_offset + n, dtype=np.uint64)
        with np.errstate(over="ignore"):
            self.state = _splitmix_vec(np.uint64(seed) ^ (ids * np.uint64(GOLDEN)))

    def next_u64(self) -> np.ndarray:
        with np.errstate(over="ignore"):
            self.state = self.state + np.uint64(GOLDEN)
            z = self.state
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def next(self) -> np.ndarray:
        return _mulhi(self.next_u64(), 18).astype(np.uint8)

    def frames(self, k: int) -> np.ndarray:
        return np.stack([self.next() for _ in range(k)])


def frameskips(seed: int, n: int, lo: int = 1, hi: int = 4) -> np.ndarray:
    s = ActionStream(seed ^ 0x5F5F5F5F, n)
    return (lo + _mulhi(s.next_u64(), hi - lo + 1)).astype(np.uint8)


# --------------------------------------------------------------------------
# random instruction synthesis
# --------------------------------------------------------------------------
CONTROL = {"BPL", "BMI", "BVC", "BVS", "BCC", "BCS", "BNE", "BEQ", "JMP", "JSR", "RTS", "RTI", "BRK"}
STACK = {"PHA", "PHP", "PLA", "PLP", "TXS"}
WRITES = {"STA", "STX", "STY"}
RMWS = {"ASL", "LSR", "ROL", "ROR", "INC", "DEC"}
READ_ONLY_RIOT = [SWCHA, 0x281, 0x282, 0x283, INSTAT]


class Emitter:
    """Emits random instructions with addressing kept inside the RAM map so
    that no random access can change control flow (no hotspot access, no
    VSYNC/WSYNC write, no stack or pointer corruption)."""

    def __init__(self, asm: Asm, rng: Rng, allow_intim: bool):
        self.a = asm
        self.r = rng
        self.allow_intim = allow_intim

    def _data_zp(self):
        return self.r.range(DATA_LO, DATA_HI)

    def _read_abs(self):
        r = self.r.below(10)
        if r < 4:
            return self.r.range(0x80, 0xFF)                 # RAM
        if r < 6:
            return 0x100 | self.r.range(0x80, 0xFF)         # RAM stack mirror
        if r < 7:
            return self.r.range(0x00, 0x3F)                 # TIA read registers (mirrored)
        if r < 8:
            return self.r.range(0x0480, 0x04FF)             # RAM mirror outside the flat array
        riot = READ_ONLY_RIOT + ([INTIM] if self.allow_intim else [])
        return self.r.choice(riot)

    def _write_abs(self, rmw: bool):
        r = self.r.below(8)
        if r < 4:
            return self._data_zp()
        if r < 6:
            return 0x100 | self._data_zp()
        if r < 7:
            return self.r.choice(SAFE_TIA_WRITES) | self.r.choice([0x00, 0x40])
        return 0x0C00 | self._data_zp()                     # far RAM mirror (writes reach RAM)

    def emit(self, mn: str, mode: str) -> int:
        """Emit one random instance of (mn, mode); returns instructions emitted."""
        a, r = self.a, self.r
        n = 1
        is_w = mn in WRITES
        is_m = mn in RMWS
        if mn in ("ADC", "SBC") and r.coin(0.5):
            a.op(r.choice(["SED", "CLD"]))
            n += 1
        if mode in ("imp", "acc"):
            a.op(mn, mode)
        elif mode == "imm":
            a.op(mn, mode, r.below(256))
        elif mode == "zp":
            if (is_w or is_m) and r.coin(0.15):
                a.op(mn, mode, r.choice(SAFE_TIA_WRITES))
            elif is_w or is_m:
                a.op(mn, mode, self._data_zp())
            else:
                a.op(mn, mode, r.range(0x80, 0xFF) if r.coin(0.85) else r.range(0x00, 0x3F))
        elif mode in ("zpx", "zpy"):
            idx = "LDX" if mode == "zpx" else "LDY"
            if mn in ("LDX", "STX"):
                idx = "LDY"
            eff = self._data_zp() if (is_w or is_m) else r.below(256)
            k = r.below(256)
            a.op(idx, "imm", k)
            a.op(mn, mode, (eff - k) & 0xFF)
            n += 1
        elif mode == "abs":
            if is_w or is_m:
                a.op(mn, mode, self._write_abs(is_m))
            else:
                a.op(mn, mode, self._read_abs())
        elif mode in ("abx", "aby"):
            idx = "LDX" if mode == "abx" else "LDY"
            if mn == "LDX":
                idx = "LDY"
            if mn == "LDY":
                idx = "LDX"
            if is_w or is_m:
                eff = (0x100 if r.coin() else 0) | self._data_zp()
                k = r.range(0, (eff & 0xFF) - 0x80)
                base = eff - k
            else:
                base = r.choice([0x80, 0x100, 0x180, 0x0480, 0x00]) + r.below(128)
                k = r.below(256)
                if base < 0x80:
                    base &= 0x3F
                if not self.allow_intim and 0x200 <= base + k < 0x300:
                    k = 0          # keep indexed reads off the RIOT I/O (INTIM clears the timer flag)
            a.op(idx, "imm", k)
            a.op(mn, mode, base)
            n += 1
        elif mode == "izx":
            slot = PTR_BASE + 2 * r.below(N_PTR)
            base = r.below(256)
            a.op("LDX", "imm", (slot - base) & 0xFF)
            a.op(mn, mode, base)
            n += 1
        elif mode == "izy":
            slot = PTR_BASE + 2 * r.below(N_PTR)
            k = r.range(0, 0x1F) if is_w else r.below(0x80)
            a.op("LDY", "imm", k)
            a.op(mn, mode, slot)
            n += 1
        else:
            raise ValueError(mode)
        return n


def _pool(names_excluded) -> list[tuple[str, str]]:
    return [OPCODES[op] for op in sorted(OPCODES) if OPCODES[op][0] not in names_excluded]


POOL_LINEAR = _pool(CONTROL | STACK)          # configs with WSYNC lines


def _reset_code(a: Asm, rng: Rng):
    """SEI; CLD; LDX #$FF; TXS; clear $80-$FF; set the pointer slots."""
    a.op("SEI")
    a.op("CLD")
    a.op("LDX", "imm", 0xFF)
    a.op("TXS")
    a.op("LDA", "imm", 0)
    a.op("LDX", "imm", 0x7F)
    a.label("clr")
    a.op("STA", "zpx", 0x80)
    a.op("DEX")
    a.op("BPL", "rel", "clr")
    for j in range(N_PTR):
        lo = 0x80 + 12 * j
        a.op("LDA", "imm", lo)
        a.op("STA", "zp", PTR_BASE + 2 * j)
        a.op("LDA", "imm", rng.choice([0x00, 0x01]))
        a.op("STA", "zp", PTR_BASE + 2 * j + 1)
    for i in range(8):
        a.op("LDA", "imm", rng.below(256))
        a.op("STA", "zp", DATA_LO + 11 * i)


def _pointer_prologue(a: Asm):
    """Re-derive four pointers from RAM each frame (env-dependent addresses)."""
    for j in range(4):
        a.op("LDA", "zp", DATA_LO + 5 * j)
        a.op("AND", "imm", 0x3F)
        a.op("ORA", "imm", 0x80)
        a.op("STA", "zp", PTR_BASE + 2 * j)


def _vectors(a: Asm, reset: int, irq: int):
    top = a.origin + a.size
    a.word_at(top - 4, reset)
    a.word_at(top - 2, irq)


SCHEMES = {2048: ("fixed", 1, 0), 4096: ("fixed", 1, 0), 8192: ("F8", 2, 0x1FF8),
           16384: ("F6", 4, 0x1FF6), 32768: ("F4", 8, 0x1FF4)}


def gen_wsync_rom(seed: int, rom_size: int = 4096, lines: int = 192, lines_per_bank: int = 16,
                  unique_lines: int | None = None) -> bytes:
    """Game-like WSYNC ROM (configs 1, 2 and the Fixed / F8 / F4 members of the
    heterogeneous mix).  For 2 KB ROMs `unique_lines` visible lines are
    generated and looped `lines // unique_lines` times."""
    rng = Rng(seed)
    kind, nbanks, hot = SCHEMES[rom_size]
    origin = 0xF800 if rom_size == 2048 else 0xF000
    size = min(rom_size, 4096)
    banks = [Asm(origin, size) for _ in range(nbanks)]
    first = nbanks - 1
    reserve_top = origin + size - 16        # vectors + hotspots
    nseg_est = (lines + lines_per_bank - 1) // lines_per_bank if nbanks > 1 else 0
    tramp_base = reserve_top - 6 * nseg_est  # trampoline slots, mirrored in every bank
    for a in banks:
        a.org(origin)

    em = [Emitter(a, rng, allow_intim=True) for a in banks]
    main = banks[first]
    _reset_code(main, rng)
    main.label("frame")
    main.op("LDA", "imm", 2)
    main.op("STA", "zp", VSYNC)
    for _ in range(3):
        main.op("STA", "zp", WSYNC)
    main.op("LDA", "imm", 0)
    main.op("STA", "zp", VSYNC)
    main.op("LDA", "imm", 43)
    main.op("STA", "abs", TIM64T)
    _pointer_prologue(main)
    main.label("vwait")
    main.op("BIT", "abs", INSTAT)
    main.op("BPL", "rel", "vwait")
    main.op("STA", "zp", WSYNC)

    def line_body(e: Emitter):
        L = e.r.range(4, 12)
        n = 0
        while n < L:
            mn, mode = e.r.choice(POOL_LINEAR)
            n += e.emit(mn, mode)
        e.a.op("STA", "zp", WSYNC)

    tramp_slot = 0
    labels_global = {}
    if nbanks == 1:
        if unique_lines:
            reps = lines // unique_lines
            main.op("LDA", "imm", reps)
            main.op("STA", "zp", LOOP_CTR)
            main.label("vis")
            for _ in range(unique_lines):
                line_body(em[first])
            main.op("DEC", "zp", LOOP_CTR)
            main.op("BEQ", "rel", "vis_done")
            main.op("JMP", "abs", "vis")
            main.label("vis_done")
        else:
            for _ in range(lines):
                line_body(em[first])
        cur = first
    else:
        nseg = (lines + lines_per_bank - 1) // lines_per_bank
        cur = first
        seq = [first] + [(first - 1 - s) % nbanks for s in range(nseg - 1)]
        for s in range(nseg):
            b = seq[s]
            if s > 0:
                # land here after the trampoline of transition s-1
                banks[b].label(f"seg{s}")
            for _ in range(min(lines_per_bank, lines - s * lines_per_bank)):
                line_body(em[b])
            nxt = seq[s + 1] if s + 1 < nseg else first
            t = tramp_base + 6 * tramp_slot
            tramp_slot += 1
            banks[b].op("JMP", "abs", t)
            # trampoline: LDA hotspot in the source bank, JMP in the destination
            src_pc = banks[b].pc
            banks[b].org(t)
            banks[b].op("LDA", "abs", (hot + nxt) | 0xE000)
            banks[b].org(src_pc)
            dst_pc = banks[nxt].pc
            banks[nxt].org(t + 3)
            banks[nxt].op("JMP", "abs", f"seg{s + 1}" if s + 1 < nseg else "post")
            banks[nxt].org(dst_pc)
            cur = nxt
        main.label("post")
    main.op("LDA", "abs", SWCHA)
    main.op("EOR", "zp", DATA_LO)
    main.op("STA", "zp", DATA_LO)
    main.op("LDA", "zp", INPT4)
    main.op("EOR", "zp", DATA_LO + 1)
    main.op("STA", "zp", DATA_LO + 1)
    main.op("LDX", "imm", 30)
    main.label("ovs")
    main.op("STA", "zp", WSYNC)
    main.op("DEX")
    main.op("BNE", "rel", "ovs")
    main.op("JMP", "abs", "frame")
    for a in banks:
        if a.pc > tramp_base:
            raise OverflowError(f"ROM overflow: ${a.pc:04X} > ${tramp_base:04X}")
    # segment labels live in other banks; resolve across banks
    for a in banks:
        labels_global.update(a.labels)
    out = b"".join(a.link(labels_global) for a in banks)
    rom = bytearray(out)
    reset = main.labels.get("reset_entry", origin)
    for b in range(nbanks):
        base = b * size
        for off, val in ((size - 4, reset & 0xFF), (size - 3, reset >> 8),
                         (size - 2, reset & 0xFF), (size - 1, reset >> 8)):
            rom[base + off] = val
    return bytes(rom)


def gen_branchy_rom(seed: int, rom_size: int = 16384, segs_per_bank: int = 6, blocks_per_seg: int = 4,
                    timer64: int = 29) -> bytes:
    """Divergence-heavy no-WSYNC ROM (config 3; the F6 member of the mix)."""
    rng = Rng(seed)
    kind, nbanks, hot = SCHEMES[rom_size]
    origin, size = 0xF000, 4096
    banks = [Asm(origin, size) for _ in range(nbanks)]
    first = nbanks - 1
    reserve_top = origin + size - 16
    tramp_base = reserve_top - 6 * segs_per_bank * nbanks
    handler = tramp_base - 16                    # BRK handler, same address in every bank
    for a in banks:
        a.org(handler)
        # BRK handler: bump a RAM counter without disturbing A or flags
        a.op("PHP")
        a.op("PHA")
        a.op("LDA", "zp", DATA_LO + 9)
        a.op("EOR", "imm", 0x5A)
        a.op("STA", "zp", DATA_LO + 9)
        a.op("PLA")
        a.op("PLP")
        a.op("RTI")
        a.org(origin)

    pool_all = [OPCODES[op] for op in sorted(OPCODES)]
    subs: dict[int, list[str]] = {b: [] for b in range(nbanks)}
    em = [Emitter(a, rng, allow_intim=False) for a in banks]

    # per-bank subroutines (JSR targets)
    for b in range(nbanks):
        a = banks[b]
        for k in range(4):
            name = f"sub{b}_{k}"
            a.label(name)
            for _ in range(rng.range(3, 8)):
                mn, mode = rng.choice(POOL_LINEAR)
                em[b].emit(mn, mode)
            a.op("RTS")
            subs[b].append(name)

    main = banks[first]
    main.label("reset_entry")
    _reset_code(main, rng)
    main.label("frame")
    main.op("LDA", "imm", 2)
    main.op("STA", "zp", VSYNC)
    main.op("LDA", "imm", 0)
    main.op("STA", "zp", VSYNC)
    main.op("LDA", "imm", timer64)
    main.op("STA", "abs", T1024T)
    main.op("LDA", "abs", SWCHA)
    main.op("EOR", "zp", DATA_LO)
    main.op("ADC", "zp", DATA_LO + 3)
    main.op("STA", "zp", DATA_LO)
    main.op("LDA", "zp", INPT4)
    main.op("EOR", "zp", DATA_LO + 1)
    main.op("STA", "zp", DATA_LO + 1)
    _pointer_prologue(main)
    main.label("chain")

    nseg = segs_per_bank * nbanks
    seq = [first] + [(first - 1 - s) % nbanks for s in range(nseg - 1)]
    tramp_slot = 0
    blk = 0

    def block(e: Emitter, b: int, skip_label: str):
        a, r = e.a, e.r
        L = r.range(8, 24)
        n = 0
        pushed = []
        while n < L:
            mn, mode = r.choice(pool_all)
            if mn in ("BPL", "BMI", "BVC", "BVS", "BCC", "BCS", "BNE", "BEQ"):
                # short forward skip over one padding instruction
                a.op(mn, "rel", 1)
                a.op("INX")
                n += 2
            elif mn == "JMP" and mode == "abs":
                lab = f"j{b}_{a.pc:04X}"
                a.op("JMP", "abs", lab)
                a.label(lab)
                n += 1
            elif mn == "JMP":
                lab = f"ji{b}_{a.pc:04X}"
                ptr = f"jp{b}_{a.pc:04X}"
                a.op("JMP", "ind", ptr)
                a.op("JMP", "abs", lab)          # skipped: the pointer targets `lab`
                a.label(ptr)
                if (a.pc & 0xFF) == 0xFF:
                    a.raw(0xEA)
                    a.labels[ptr] = a.pc
                a.fix.append((a.pc - a.origin, lab, "abs"))
                a.raw(0, 0)
                a.label(lab)
                n += 1
            elif mn == "JSR":
                a.op("JSR", "abs", r.choice(subs[b]))
                n += 1
            elif mn == "BRK":
                a.op("BRK")
                a.raw(r.below(256))
                n += 1
            elif mn in ("RTS", "RTI"):
                continue
            elif mn in ("PHA", "PHP"):
                if len(pushed) < 2:
                    a.op(mn)
                    pushed.append("PLA" if mn == "PHA" else "PLP")
                    n += 1
            elif mn in ("PLA", "PLP"):
                if pushed:
                    a.op(pushed.pop())
                    n += 1
            elif mn == "TXS":
                a.op("TSX")
                a.op("TXS")
                n += 2
            else:
                n += e.emit(mn, mode)
        while pushed:
            a.op(pushed.pop())
        # data-dependent exit: skip the next block on a RAM byte
        br = r.choice(["BNE", "BMI", "BCS"])
        if br == "BCS":
            a.op("LDA", "zp", e._data_zp())
            a.op("CMP", "zp", e._data_zp())
        else:
            a.op(r.choice(["LDA", "LDX", "LDY"]), "zp", e._data_zp())
        a.op(br, "rel", skip_label)

    for s in range(nseg):
        b = seq[s]
        a = banks[b]
        if s > 0:
            a.label(f"seg{s}")
        for k in range(blocks_per_seg):
            a.label(f"blk{blk}")
            block(em[b], b, f"blk{blk + 2}" if k + 2 < blocks_per_seg else f"end{s}")
            blk += 1
        a.label(f"end{s}")
        blk += 0
        nxt = seq[s + 1] if s + 1 < nseg else first
        t = tramp_base + 6 * tramp_slot
        tramp_slot += 1
        if s + 1 < nseg:
            a.op("JMP", "abs", t)
            src_pc = a.pc
            a.org(t)
            a.op("LDA", "abs", (hot + nxt) | 0xE000)
            a.org(src_pc)
            dst_pc = banks[nxt].pc
            banks[nxt].org(t + 3)
            banks[nxt].op("JMP", "abs", f"seg{s + 1}")
            banks[nxt].org(dst_pc)
        else:
            a.op("JMP", "abs", t)
            src_pc = a.pc
            a.org(t)
            a.op("LDA", "abs", (hot + first) | 0xE000)
            a.org(src_pc)
            dst_pc = main.pc
            main.org(t + 3)
            main.op("JMP", "abs", "poll")
            main.org(dst_pc)
        # relabel block counter so that skip targets inside the next segment start fresh
        blk = (blk // blocks_per_seg) * blocks_per_seg
    main.label("poll")
    main.op("BIT", "abs", INSTAT)
    main.op("BMI", "rel", "poll_done")
    main.op("JMP", "abs", "chain")
    main.label("poll_done")
    main.op("JMP", "abs", "frame")
    for a in banks:
        if a.pc > handler:
            raise OverflowError(f"ROM overflow: ${a.pc:04X} > ${handler:04X}")
    labels_global = {}
    for a in banks:
        for k, v in a.labels.items():
            if k.startswith("seg") or k in ("poll", "frame", "chain", "reset_entry"):
                labels_global[k] = v
    out = bytearray()
    for a in banks:
        lab = dict(labels_global)
        lab.update(a.labels)
        out += a.link(lab)
    reset = main.labels["reset_entry"]
    for b in range(nbanks):
        base = b * size
        for off, val in ((size - 4, reset & 0xFF), (size - 3, reset >> 8),
                         (size - 2, handler & 0xFF), (size - 1, handler >> 8)):
            out[base + off] = val
    return bytes(out)


# TIA graphics registers (write addresses)
NUSIZ0, NUSIZ1, COLUP0, COLUP1, COLUPF, COLUBK, CTRLPF, REFP0, REFP1 = 0x04, 0x05, 0x06, 0x07, 0x08, 0x09, 0x0A, 0x0B, 0x0C
PF0, PF1, PF2, RESP0, RESP1, RESM0, RESM1, RESBL = 0x0D, 0x0E, 0x0F, 0x10, 0x11, 0x12, 0x13, 0x14
GRP0, GRP1, ENAM0, ENAM1, ENABL, HMP0, HMP1, HMM0, HMM1, HMBL = 0x1B, 0x1C, 0x1D, 0x1E, 0x1F, 0x20, 0x21, 0x22, 0x23, 0x24
VDELP0, VDELP1, VDELBL, RESMP0, RESMP1, HMOVE, HMCLR, CXCLR = 0x25, 0x26, 0x27, 0x28, 0x29, 0x2A, 0x2B, 0x2C
FRAME_CTR = 0xD0
LINE_REGS = [GRP0, GRP1, GRP0, GRP1, COLUP0, COLUP1, COLUPF, COLUBK, PF0, PF1, PF2, PF1, PF2, ENAM0, ENAM1,
             ENABL, NUSIZ0, NUSIZ1, CTRLPF, REFP0, REFP1, VDELP0, VDELP1, VDELBL, RESMP0, RESMP1, HMP0, HMP1,
             HMM0, HMM1, HMBL]


def gen_render_rom(seed: int, bodies: int = 8, vsync: bool = True) -> bytes:
    """A 4 KB Fixed graphics kernel for the rendering console: every frame
    positions the five objects with delay loops whose counts come from RAM
    (RESP0 / RESP1 / RESM0 / RESM1 / RESBL at env-dependent clocks, then
    HMxx + HMOVE), runs 192 visible lines of `bodies` random line bodies
    (stores of RAM-derived or immediate values to the graphics registers,
    mid-line RESPx, HMOVE after WSYNC, a VBLANK pulse inside a line, RAM
    updates), reads the collision latches, SWCHA and the latched INPT4 into
    RAM, and clears the latches after every second frame.  vsync=False leaves out VSYNC, so frames end
    at run_frame's 100,000-cycle cap, anywhere in the picture."""
    rng = Rng(seed)
    a = Asm(0xF000, 4096)
    _reset_code(a, rng)
    a.label("frame")
    if vsync:
        a.op("LDA", "imm", 2)
        a.op("STA", "zp", VSYNC)
        for _ in range(3):
            a.op("STA", "zp", WSYNC)
        a.op("LDA", "imm", 0)
        a.op("STA", "zp", VSYNC)
    a.op("LDA", "imm", 0x42)                 # blank, input latch on
    a.op("STA", "zp", VBLANK)
    # object positions: delay loops of RAM-derived length
    for j, reg in enumerate((RESP0, RESP1, RESM0, RESM1, RESBL)):
        a.op("STA", "zp", WSYNC)
        a.op("LDA", "zp", DATA_LO + 3 * j)
        a.op("AND", "imm", 0x0F)
        a.op("TAX")
        a.op("INX")
        a.label(f"pos{j}")
        a.op("DEX")
        a.op("BNE", "rel", f"pos{j}")
        a.op("STA", "zp", reg)
    for j, reg in enumerate((HMP0, HMP1, HMM0, HMM1, HMBL)):
        a.op("LDA", "zp", DATA_LO + 3 * j + 1)
        a.op("STA", "zp", reg)
    a.op("STA", "zp", WSYNC)
    a.op("STA", "zp", HMOVE)
    for _ in range(28):
        a.op("STA", "zp", WSYNC)
    a.op("LDA", "imm", 0)
    a.op("STA", "zp", VBLANK)
    # 192 visible lines: `bodies` line bodies looped
    a.op("LDA", "imm", 192 // bodies)
    a.op("STA", "zp", LOOP_CTR)
    a.op("LDY", "imm", 0)
    a.label("vis")
    for b in range(bodies):
        a.op("STA", "zp", WSYNC)
        if rng.coin(0.25):
            a.op("STA", "zp", HMOVE)
        for _ in range(rng.range(3, 7)):
            reg = rng.choice(LINE_REGS)
            src = rng.below(4)
            if src == 0:
                a.op("LDA", "imm", rng.below(256))
            elif src == 1:
                a.op("LDA", "zp", rng.range(DATA_LO, DATA_HI))
            elif src == 2:
                a.op("LDA", "aby", rng.range(DATA_LO, DATA_LO + 0x20))
            else:
                a.op("TYA")
                a.op("EOR", "zp", rng.range(DATA_LO, DATA_HI))
            a.op("STA", "zp", reg)
        r = rng.below(8)
        if r == 0:
            a.op("STA", "zp", rng.choice([RESP0, RESP1, RESM0, RESM1, RESBL]))
        elif r == 1:
            a.op("LDA", "imm", 2)
            a.op("STA", "zp", VBLANK)
            a.op("NOP")
            a.op("LDA", "imm", 0)
            a.op("STA", "zp", VBLANK)
        elif r == 2:
            a.op("STA", "zp", HMCLR)
        a.op("INC", "zp", rng.range(DATA_LO, DATA_LO + 0x20))
        a.op("INY")
    a.op("DEC", "zp", LOOP_CTR)
    a.op("BEQ", "rel", "vis_done")
    a.op("JMP", "abs", "vis")
    a.label("vis_done")
    a.op("LDA", "imm", 0x42)
    a.op("STA", "zp", VBLANK)
    # collisions, joystick and the latched fire button into RAM
    for j, reg in enumerate(range(0x00, 0x08)):
        a.op("LDA", "zp", reg)
        a.op("EOR", "zp", DATA_LO + j)
        a.op("STA", "zp", DATA_LO + j)
    a.op("LDA", "abs", SWCHA)
    a.op("EOR", "zp", DATA_LO + 8)
    a.op("STA", "zp", DATA_LO + 8)
    a.op("LDA", "zp", INPT4)
    a.op("ADC", "zp", DATA_LO + 9)
    a.op("STA", "zp", DATA_LO + 9)
    a.op("INC", "zp", FRAME_CTR)             # latches cleared after even frames only
    a.op("LDA", "zp", FRAME_CTR)
    a.op("LSR", "acc")
    a.op("BCS", "rel", "keep_cx")
    a.op("STA", "zp", CXCLR)
    a.label("keep_cx")
    a.op("LDX", "imm", 24)
    a.label("ovs")
    a.op("STA", "zp", WSYNC)
    a.op("DEX")
    a.op("BNE", "rel", "ovs")
    a.op("JMP", "abs", "frame")
    rom = bytearray(a.link())
    for off, val in ((0xFFC, 0x00), (0xFFD, 0xF0), (0xFFE, 0x00), (0xFFF, 0xF0)):
        rom[off] = val
    return bytes(rom)


# --------------------------------------------------------------------------
# the five benchmark configurations
# --------------------------------------------------------------------------
def rom_for(config: int, member: int = 0) -> bytes:
    """ROM of a config (for config 4/5, `member` picks one of the four ROMs)."""
    s = config_seed(config)
    if config == 1:
        return gen_wsync_rom(s, 4096)
    if config == 2:
        return gen_wsync_rom(s, 8192, lines_per_bank=16)
    if config == 3:
        return gen_branchy_rom(s, 16384)
    if config in (4, 5):
        s4 = config_seed(4)
        return mix_roms(s4)[member]
    raise ValueError(config)


def mix_roms(seed: int) -> list[bytes]:
    """The heterogeneous set: 2 KB Fixed WSYNC, 8 KB F8, 16 KB F6 no-WSYNC,
    32 KB F4 switching banks every 8 scanlines."""
    return [
        gen_wsync_rom(seed + 0x100, 2048, unique_lines=48),
        gen_wsync_rom(seed + 0x200, 8192, lines_per_bank=16),
        gen_branchy_rom(seed + 0x300, 16384),
        gen_wsync_rom(seed + 0x400, 32768, lines_per_bank=8),
    ]


MIX_COUNTS = (1000, 40000, 12000, 3000)
