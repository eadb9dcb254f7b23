// fa_uniform.hpp -- the wave-uniform control mode of the frame loop.
//
// When every active lane of a wave holds the same control state -- PC, stack
// pointer, bank, TIA line timing, VSYNC and the RIOT timer (the frame loop's
// cycle count may differ, see u_enter) -- that state is kept once per wave in
// scalar registers and every
// decision that depends on it (opcode dispatch, the bus region of an address,
// TIA / RIOT side effects, line and frame ends) is a scalar branch.  Only what
// a program computes on -- A, X, Y, P, the paddle counter and RAM -- stays per
// lane.  ROMs whose control flow does not depend on the env's data (WSYNC
// kernels driven by the timer) run whole frames this way.
//
// An instruction whose effect on the control state would differ between lanes
// -- a branch taken by some lanes only, a page crossing on some lanes only, TXS
// / RTS / RTI with per-lane values, a VSYNC or timer write of per-lane values, a
// write at a per-lane address outside RAM, a read at per-lane addresses that
// reach a hotspot or a RIOT timer register -- is declined before anything changes: the
// frame loop writes the scalar state back to every lane and runs that
// instruction on the per-lane path, and re-enters this mode once the lanes agree
// again.  Behaviour is the reference's CPU emulator throughout
// (src/cpu.rs:245-1146, src/headless_tia.rs:153-298, src/pia.rs:107-228,
// src/atari.rs:459-614), in the same access order as Emu::execute.
#pragma once

#include "fa_emu.hpp"

namespace fa {

// Wave-uniform control state (scalar registers).
struct UCtl {
    uint32_t pc, sp, bank;
    uint32_t clock, line, wsync, fcomp, vsync;   // wsync bit 1: see u_loop_bit
    uint32_t tval, tps, tsel, tsh, tuf;
    uint32_t tlim;           // timer fast-path budget (u_tlim)
    uint32_t cycles, in_a;   // run_frame's cycle count (with cspread: a bound, see u_enter) and loop-A flag
    uint32_t c0, cspread;    // cycles at entry; non-zero when the lanes' counts differ
    uint32_t skip, instr;    // WSYNC-skip cycles and instructions run since entry (the CPU's cycles: cycles - c0 - skip)
};

enum : uint32_t { U_DONE = 0, U_EXIT = 1 };

// Cycles one iteration of the frame loop can add: a WSYNC skip (at most 76) or
// one instruction (at most 7).
constexpr uint32_t kIterCycles = 80u;

// The first active lane's v, opaque to the optimiser: after a test that v
// equals it on every lane it must not be replaced by the per-lane v.
__device__ __forceinline__ uint32_t wave_first(uint32_t v) {
    uint32_t s = __builtin_amdgcn_readfirstlane(v);
    __asm__ volatile("" : "+s"(s));
    return s;
}
__device__ __forceinline__ uint32_t first_lane(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t active_lanes() { return __ballot(1); }
// v equals s on every active lane.
__device__ __forceinline__ bool same_on_all(uint32_t v, uint32_t s) { return __ballot(v != s) == 0; }

// ---------------------------------------------------------------- entry / exit
// Cheap pre-test run before every per-lane instruction: PC and TIA clock agree.
__device__ __forceinline__ bool u_maybe(const Emu &e) {
    const uint32_t w = e.pc | (e.clock << 16);
    return same_on_all(w, wave_first(w));
}

// Bit 1 of UCtl::wsync: VSYNC no longer agrees with the frame loop (loop A
// runs while VSYNC is held, loop B until it rises), set when VSYNC is written
// and at entry, so the loop tests one word for both this and WSYNC.
__device__ __forceinline__ uint32_t u_loop_bit(const UCtl &u) {
    return (((u.vsync >> 1) & 1u) != u.in_a) ? 2u : 0u;
}

// Full test; on success fills u from the common state.  Fields are compared at
// the widths they have in the state image, so agreement is exact.  The frame's
// cycle count may differ between lanes by up to 127 (an env restarted from the
// reset snapshot runs its first frame at another TIA phase, realigned by the
// first WSYNC): it only feeds the frame-loop caps, so the mode then carries an
// upper bound of every lane's count (the first lane's + 127), adds the cycles
// it runs to each lane's own count on leaving, and leaves while the bound is
// still at least one iteration (kIterCycles) short of the cap that applies, so
// that the caps never separate the lanes inside the mode.
__device__ __forceinline__ bool u_enter(const Emu &e, uint32_t cycles, bool in_a, UCtl &u) {
    const uint32_t w0 = e.pc | (e.sp << 16) | (e.bank << 24);
    const uint32_t w1 = e.clock | (e.line << 16);
    const uint32_t w2 = e.tval | (e.tps << 16);
    const uint32_t w3 = e.tsel | (e.tuf << 16) | (e.wsync << 24);
    const uint32_t w4 = e.fcomp | (e.vsync << 8) | ((uint32_t)in_a << 16);
    const uint32_t s0 = wave_first(w0), s1 = wave_first(w1), s2 = wave_first(w2), s3 = wave_first(w3),
                   s4 = wave_first(w4), s5 = wave_first(cycles);
    const bool diff = (w0 != s0) | (w1 != s1) | (w2 != s2) | (w3 != s3) | (w4 != s4) | (cycles - s5 + 127u > 254u);
    if (__ballot(diff) != 0) return false;
    u.cspread = __ballot(cycles != s5) != 0 ? 1u : 0u;
    u.cycles = u.cspread ? s5 + 127u : s5;
    if (u.cspread && u.cycles > LOOPB_CAP - kIterCycles && (!(s4 & 0x10000u) || u.cycles > LOOPA_CAP - kIterCycles))
        return false;
    u.c0 = u.cycles;
    u.pc = s0 & 0xFFFFu;
    u.sp = (s0 >> 16) & 0xFFu;
    u.bank = s0 >> 24;
    u.clock = s1 & 0xFFFFu;
    u.line = s1 >> 16;
    u.tval = s2 & 0xFFFFu;
    u.tps = s2 >> 16;
    u.tsel = s3 & 0xFFFFu;
    u.tsh = 31u - __builtin_clz(u.tsel | 1u);
    u.tuf = (s3 >> 16) & 0xFFu;
    u.wsync = s3 >> 24;
    u.fcomp = s4 & 0xFFu;
    u.vsync = (s4 >> 8) & 0xFFu;
    u.in_a = (s4 >> 16) & 1u;
    u.wsync |= u_loop_bit(u);
    u.skip = 0;
    u.instr = 0;
    return true;
}

__device__ __forceinline__ void u_leave(const UCtl &u, Emu &e, uint32_t &cycles, bool &in_a, uint32_t &cpu_cyc,
                                        uint32_t &instr) {
    e.pc = u.pc; e.sp = u.sp; e.bank = u.bank;
    e.clock = u.clock; e.line = u.line; e.wsync = u.wsync & 1u; e.fcomp = u.fcomp; e.vsync = u.vsync;
    e.tval = u.tval; e.tps = u.tps; e.tsel = u.tsel; e.tsh = u.tsh; e.tuf = u.tuf;
    cycles += u.cycles - u.c0;
    in_a = u.in_a != 0u;
    cpu_cyc += u.cycles - u.c0 - u.skip;
    instr += u.instr;
}

// ---------------------------------------------------------------- TIA timing
// HeadlessTia::tick_n / skip_to_scanline_end / end_scanline
// (src/headless_tia.rs:153-198); the paddle counter is per lane.
__device__ __forceinline__ void u_end_scanline(UCtl &u, Emu &e) {
    u.clock -= 228u;
    u.line += 1u;
    u.wsync &= 2u;
    e.pcount += (uint32_t)(e.pdump == 0u) & (uint32_t)(e.pcount < 256u);
    if (u.line >= 262u) { u.line = 0; u.fcomp = 1; }
}
__device__ __forceinline__ void u_tia_tick(UCtl &u, Emu &e, uint32_t n) {
    u.clock += 3u * n;
    if (__builtin_expect(u.clock >= 228u, 0)) u_end_scanline(u, e);
}
__device__ __forceinline__ uint32_t u_tia_skip(UCtl &u, Emu &e) {
    const uint32_t rem = u.clock >= 228u ? 0u : 228u - u.clock;
    const uint32_t cyc = (rem + 2u) / 3u;
    if (cyc) { u.clock = 228u; u_end_scanline(u, e); }
    return cyc;
}

// Cycles a timer advance may take by the fast path alone (prescaler count
// minus the cycles, no value change): the prescaler count, or 0 in the
// post-underflow state (value 0 with the underflow flag set), where every
// cycle decrements the value.  Kept in step with tval / tps / tuf, so the
// per-instruction test is one compare.
__device__ __forceinline__ uint32_t u_tlim(const UCtl &u) { return (u.tval == 0u && u.tuf) ? 0u : u.tps; }

// What the frame loop does after an instruction of `cyc` cycles: the counts,
// the TIA clock, and Pia::tick_n (its fast branch inline, the rest out of line).
__device__ __forceinline__ void u_step(UCtl &u, Emu &e, uint32_t cyc) {
    u.cycles += cyc;
    u.instr += 1u;
    u_tia_tick(u, e, cyc);
    if (__builtin_expect(u.tlim > cyc, 1)) {
        u.tps -= cyc;
        u.tlim -= cyc;
    } else {
        timer_tick_slow(u, cyc);
        u.tlim = u_tlim(u);
    }
}

// ---------------------------------------------------------------- bus
// Side effects of a read at a wave-uniform 13-bit address whose value is not
// used: a hotspot switches the bank, INTIM clears the underflow flag.
__device__ __forceinline__ void u_touch(UCtl &u, const Emu &e, uint32_t a13) {
    const uint32_t hd = (a13 & 0xFFFu) - e.hot_lo;
    if ((a13 & 0x1000u) && hd < e.hot_n) u.bank = hd;
    if ((a13 & 0x1287u) == 0x284u) {
        u.tuf = 0;
        u.tlim = u.tps;
    }
}
// ROM read at a uniform address (HeadlessBus::read: the hotspot first).
__device__ __forceinline__ uint32_t u_rom_read(UCtl &u, const Emu &e, uint32_t a13) {
    u_touch(u, e, a13);
    return first_lane(ld8(ROM0 + (u.bank << 12) + (a13 & 0xFFFu)));
}
// TIA / RIOT / flat-array padding read at a uniform address (src/atari.rs:509-519).
__device__ __forceinline__ uint32_t u_io_read(UCtl &u, Emu &e, uint32_t a13) {
    if ((a13 & 0x280u) == 0x280u) {
        const uint32_t r = a13 & 7u;
        if (r == 4u) { u.tuf = 0; u.tlim = u.tps; return u.tval & 0xFFu; }   // INTIM
        if (r == 5u) return u.tuf ? 0x80u : 0u;                 // INSTAT
        return e.pia_read(a13);                                 // ports: per lane
    }
    if (!(a13 & 0x80u)) return e.tia_read(a13);
    return 0u;
}
// TIA / RIOT write at a uniform address (src/atari.rs:541-551).  The caller
// has checked that VSYNC and timer writes carry a wave-uniform value.
__device__ __forceinline__ void u_io_write(UCtl &u, Emu &e, uint32_t a13, uint32_t v) {
    if (a13 & 0x200u) {
        if (a13 & 0x10u) {
            timer_set(u, a13, first_lane(v));
            u.tlim = u.tps;
        } else {
            e.pia_write(a13, v);
        }
        return;
    }
    const uint32_t r = a13 & 0x3Fu;
    if (r == 0x00u) {
        const uint32_t vu = first_lane(v);
        if (!(u.vsync & 2u) && (vu & 2u)) { u.line = 0; u.fcomp = 1; }
        u.vsync = vu;
        u.wsync = (u.wsync & 1u) | u_loop_bit(u);
    } else if (r == 0x01u) {
        // VBLANK: bit 7 dumps the paddle capacitor (src/headless_tia.rs:232-240)
        const bool dump = (v & 0x80u) != 0u;
        e.pcount = pick(dump, 0u, e.pcount);
        e.pdump = dump ? 1u : 0u;
        st8(e.img + S_VBLANK, v);
    } else if (r == 0x02u) {
        u.wsync |= 1u;
    } else if (r >= 0x10u && r <= 0x14u) {
        st16(e.img + S_POS + 2u * (r - 0x10u), resp_pos_of(u.clock));
    } else {
        e.tia_write(a13, v);          // per-lane registers only
    }
}

// ---------------------------------------------------------------- compute
// The uniform loop contains no branch on a per-lane value: once the optimiser
// merges the tails of different instructions, a per-lane branch would make the
// scalar state it joins per-lane too.  Per-lane choices below are selects.

// Decimal-mode ADC / SBC (src/cpu.rs:155-166, 180-191) on the lanes whose D
// flag is set, as Emu::decimal, written with selects.
template <bool SUB>
__device__ __forceinline__ void u_decimal(Emu &e, uint32_t a0, uint32_t v, uint32_t p0) {
    const bool on = (p0 & FD) != 0u;
    const uint32_t c = p0 & FC;
    if constexpr (!SUB) {
        uint32_t lo = (a0 & 0xFu) + (v & 0xFu) + c;
        lo += lo > 9u ? 6u : 0u;
        uint32_t hi = (a0 >> 4) + (v >> 4) + (lo > 0xFu ? 1u : 0u);
        const uint32_t inter = ((hi << 4) | (lo & 0xFu)) & 0xFFu;
        uint32_t np = e.p & ~(FN | FV | FC);                  // Z stays the binary one
        np |= (hi & 8u) ? FN : 0u;
        np |= (~(a0 ^ v) & (a0 ^ inter) & 0x80u) ? FV : 0u;
        hi += hi > 9u ? 6u : 0u;
        np |= hi > 0xFu ? FC : 0u;
        e.p = pick(on, np, e.p);
        e.a = pick(on, ((hi & 0xFu) << 4) | (lo & 0xFu), e.a);
    } else {
        // all flags are the binary ones; only the accumulator differs
        const uint32_t borrow = c ^ 1u;
        int32_t lo = (int32_t)(a0 & 0xFu) - (int32_t)(v & 0xFu) - (int32_t)borrow;
        const uint32_t lb = lo < 0 ? 1u : 0u;
        lo -= lb ? 6 : 0;
        int32_t hi = (int32_t)(a0 >> 4) - (int32_t)(v >> 4) - (int32_t)lb;
        hi -= hi < 0 ? 6 : 0;
        e.a = pick(on, (((uint32_t)hi & 0xFu) << 4) | ((uint32_t)lo & 0xFu), e.a);
    }
}

// The register / flag part of one compute-class instruction on every lane
// (Emu::fast_uniform's compute stage with the descriptor known at compile
// time).  Returns the value a store or read-modify-write writes back.
template <uint32_t D0, uint32_t D1>
__device__ __forceinline__ uint32_t u_alu(Emu &e, uint32_t s, uint32_t v) {
    constexpr uint32_t kind = (D0 >> D_KIND_SHIFT) & 7u;
    constexpr uint32_t dst = (D1 >> D1_DST_SHIFT) & 7u;
    const uint32_t p0 = e.p, a0 = e.a;
    uint32_t r = s;
    uint32_t q = e.p;
    if constexpr (kind == K_LOGIC) {
        r = (D0 & D_M_LD) ? v : ((D0 & D_M_XOR) ? ((D0 & D_M_AND) ? (e.a | v) : (e.a ^ v)) : (e.a & v));
    } else if constexpr (kind == K_ARITH) {
        const uint32_t vv = (D0 & D_INV) ? (v ^ 0xFFu) : v;
        const uint32_t sum = s + vv + ((D0 & D_CIN1) ? 1u : (e.p & FC));
        r = sum & 0xFFu;
        q = (q & ~FC) | ((sum >> 8) & 1u);
        if constexpr ((D0 & D_SETV) != 0u) q = (q & ~FV) | ((~(s ^ vv) & (s ^ r) & 0x80u) >> 1);
    } else if constexpr (kind == K_RMW) {
        constexpr uint32_t rop = (D1 >> D1_RMW_SHIFT) & 7u;
        if constexpr (rop < 2u) {
            r = ((v << 1) | (rop == 1u ? (e.p & FC) : 0u)) & 0xFFu;
            q = (q & ~FC) | (v >> 7);
        } else if constexpr (rop < 4u) {
            r = (v >> 1) | (rop == 3u ? ((e.p & FC) << 7) : 0u);
            q = (q & ~FC) | (v & 1u);
        } else {
            r = (v + (rop == 5u ? 1u : 0xFFu)) & 0xFFu;
        }
    } else if constexpr (kind == K_MOVE) {
        constexpr uint32_t dl = (D1 >> D1_DELTA_SHIFT) & 3u;
        r = (s + (dl == 1u ? 1u : (dl == 2u ? 0xFFu : 0u))) & 0xFFu;
    } else if constexpr (kind == K_BIT) {
        q = (q & ~(FN | FV | FZ)) | (v & 0xC0u) | ((e.a & v) == 0u ? FZ : 0u);
    } else if constexpr (kind == K_FLAGS) {
        constexpr uint32_t fm = (D1 >> D1_FMASK_SHIFT) & 0xFFu;
        q = (q & ~fm) | (D1 >> D1_FVAL_SHIFT);
    }
    if constexpr ((D0 & D_SETNZ) != 0u) q = (q & ~(FN | FZ)) | (r & 0x80u) | (r == 0u ? FZ : 0u);
    e.p = q;
    if constexpr (dst == R_A) e.a = r;
    else if constexpr (dst == R_X) e.x = r;
    else if constexpr (dst == R_Y) e.y = r;
    if constexpr ((D0 & D_DEC) != 0u) {
        if (__ballot((p0 & FD) != 0u) != 0) u_decimal<(D0 & D_INV) != 0u>(e, a0, v, p0);
    }
    return r;
}

// A data read at per-lane addresses none of which has a side effect (no
// hotspot, no RIOT timer register), as the per-lane stage makes it: one load
// from the ROM window in the current bank, a RAM mirror, the TIA register's
// image byte (fixed by tia_fix) or a zero byte (the flat array's padding);
// the RIOT ports (Pia::read, src/pia.rs:182-200) from the lane's port bytes.
// Per-lane choices are selects (see u_alu's section note).
__device__ __forceinline__ uint32_t u_read_lanes(const UCtl &u, const Emu &e, uint32_t a13) {
    const bool rom = (a13 & 0x1000u) != 0u;
    const bool ram = (a13 & 0x1E80u) == 0x0080u;
    const bool tia = (a13 & 0x1080u) == 0u;
    const uint32_t r4 = a13 & 0x0Fu;
    uint32_t ad = pick(tia, e.img + S_CX + Emu::tia_lut(r4), ZERO_BYTE);
    ad = pick(ram, e.ram_off(a13), ad);
    ad = pick(rom, ROM0 + (u.bank << 12) + (a13 & 0xFFFu), ad);
    uint32_t m = ld8(ad);
    m = pick(tia, e.tia_fix(r4, m), m);
    const bool riot = (a13 & 0x1280u) == 0x0280u;
    if (__builtin_expect(__ballot(riot) != 0, 0)) {
        // port A at S_PA_OUT (output, DDR, input), port B three bytes on
        const uint32_t pb = e.img + S_PA_OUT + 3u * ((a13 >> 1) & 1u);
        const uint32_t ddr = ld8(pb + 1u);
        const uint32_t port = (ld8(pb + 2u) & ~ddr & 0xFFu) | (ld8(pb) & ddr);
        m = pick(riot, pick((a13 & 1u) != 0u, ddr, port), m);
    }
    return m;
}

// Memory operand at a per-lane address: every lane must access RAM (a plain
// read: any region whose read has no side effect, u_read_lanes), no
// discarded fix-up read may have a side effect, and the page crossing (a
// cycle) must agree across lanes.
template <uint32_t D0, uint32_t D1>
__device__ __forceinline__ bool u_mem_lane(UCtl &u, Emu &e, uint32_t s, uint32_t base, uint32_t ea, uint32_t pcn,
                                           uint32_t &cyc) {
    constexpr bool rd = (D0 & D_RD) != 0u, wr = (D0 & D_WR) != 0u;
    constexpr bool dalways = (D0 & D_DUMMY_ALWAYS) != 0u, dcross = (D0 & D_DUMMY_CROSS) != 0u;
    constexpr bool any = rd && !wr;
    const uint32_t a13 = ea & 0x1FFFu;
    const bool cross = ((base ^ ea) & 0xFF00u) != 0u;
    const bool ram = rd ? (a13 & 0x1E80u) == 0x0080u : (a13 & 0x1280u) == 0x0080u;
    bool bad = !ram;
    if constexpr (any) {
        // a hotspot switches the bank; INTIM clears the underflow flag (and the
        // timer registers are wave-uniform state here)
        bad = ((((a13 & 0x1000u) != 0u) & (((a13 & 0xFFFu) - e.hot_lo) < e.hot_n))) | ((a13 & 0x1284u) == 0x0284u);
    }
    if constexpr (dalways || dcross) {
        // a discarded fix-up read at a hotspot or INTIM has a side effect
        const uint32_t fa = ((base & 0xFF00u) | (ea & 0xFFu)) & 0x1FFFu;
        const bool eff = (((fa & 0x1000u) != 0u) & (((fa & 0xFFFu) - e.hot_lo) < e.hot_n)) |
                         ((fa & 0x1287u) == 0x284u);
        bad = bad | ((dalways | cross) & eff);
    }
    if (__ballot(bad) != 0) return false;
    if constexpr (dcross) {
        const uint64_t cm = __ballot(cross);
        if (cm != 0 && cm != active_lanes()) return false;
        cyc += cm != 0 ? 1u : 0u;
    }
    const uint32_t off = e.ram_off(a13);
    uint32_t v = s;
    if constexpr (rd) {
        if (!any || __ballot(!ram) == 0) v = ld8(off);
        else v = u_read_lanes(u, e, a13);
    }
    const uint32_t r = u_alu<D0, D1>(e, s, v);
    if constexpr (wr) st8(off, r);
    u.pc = pcn;
    return true;
}

// Memory operand at a wave-uniform address: any bus region, side effects in
// the reference's order (fix-up read, data read, phantom write, final write).
template <uint32_t D0, uint32_t D1>
__device__ __forceinline__ bool u_mem_uni(UCtl &u, Emu &e, uint32_t s, uint32_t base, uint32_t ea, uint32_t pcn,
                                          uint32_t &cyc) {
    constexpr bool rd = (D0 & D_RD) != 0u, wr = (D0 & D_WR) != 0u, rmw = (D0 & D_RMWMEM) != 0u;
    constexpr bool dalways = (D0 & D_DUMMY_ALWAYS) != 0u, dcross = (D0 & D_DUMMY_CROSS) != 0u;
    const uint32_t a13 = ea & 0x1FFFu;
    const bool cross = ((base ^ ea) & 0xFF00u) != 0u;
    const bool rom = (a13 & 0x1000u) != 0u;
    const bool ram_r = (a13 & 0x1E80u) == 0x0080u;
    const bool ram_w = (a13 & 0x1280u) == 0x0080u;
    if constexpr (wr) {
        // VSYNC and timer writes change wave-uniform state: the value must agree.
        const bool io_w = !rom && !ram_w;
        const bool ctl = io_w && ((a13 & 0x200u) ? (a13 & 0x10u) != 0u : (a13 & 0x3Fu) == 0u);
        if (ctl) {
            if (rmw) return false;
            if (!same_on_all(s, wave_first(s))) return false;
        }
    }
    if (dalways || (dcross && cross)) u_touch(u, e, ((base & 0xFF00u) | (ea & 0xFFu)) & 0x1FFFu);
    if (dcross && cross) cyc += 1u;
    if ((!rd || ram_r) && (!wr || ram_w)) {
        const uint32_t off = e.ram_off(a13);
        const uint32_t v = rd ? ld8(off) : s;
        const uint32_t r = u_alu<D0, D1>(e, s, v);
        if constexpr (wr) st8(off, r);
        u.pc = pcn;
        return true;
    }
    uint32_t v = s;
    if constexpr (rd) {
        if (rom) v = u_rom_read(u, e, a13);
        else if (ram_r) v = ld8(e.ram_off(a13));
        else v = u_io_read(u, e, a13);
    }
    const uint32_t r = u_alu<D0, D1>(e, s, v);
    if constexpr (wr) {
        if (rom) {
            u_touch(u, e, a13);                       // hotspot; the data goes nowhere
        } else if (ram_w) {
            st8(e.ram_off(a13), r);
        } else {
            if constexpr (rmw) u_io_write(u, e, a13, v);
            u_io_write(u, e, a13, r);
        }
    }
    u.pc = pcn;
    return true;
}

// A poll of the RIOT timer: the load at u.pc (BIT / LDA / LDX / LDY abs of
// INTIM or INSTAT, the caller has checked the address) followed, at pcn, by a
// branch back to it on a flag that the load sets from the value alone (N; V
// after BIT; Z after LDA / LDX / LDY).  While the branch would be taken, whole
// iterations -- the load, then the branch -- run here on the scalar state
// with the steps the frame loop takes after each instruction: the value is
// known before the read (INTIM: the timer value, INSTAT: the underflow flag),
// the read's one side effect (INTIM clears the underflow flag) is made, and
// the flags and register the load writes per lane are written again by the
// load that follows.  Iterations end an iteration short of the frame loop's
// cap tests, so those tests fail on every instruction run here as they would
// in the loop.  Returns with u.pc still at the load, which the caller then
// runs as usual.
template <bool BIT>
__device__ __forceinline__ void u_poll(UCtl &u, Emu &e, uint32_t a13, uint32_t pcn) {
    constexpr uint32_t kLoad = 4u, kBranch = 3u;          // abs load; branch taken, same page
    if (((pcn & 0x1FFFu) - 0x1000u) >= 0xFF0u) return;    // the branch's fetch bound (u_run)
    if ((u.pc ^ (pcn + 2u)) & 0xFF00u) return;            // the taken branch crosses a page
    const uint32_t o = ROM0 + (u.bank << 12) + (pcn & 0xFFFu);
    const uint32_t bop = first_lane(ld8(o)), boff = first_lane(ld8(o + 1u));
    if ((bop & 0x1Fu) != 0x10u || boff != 0xFBu) return;   // not a branch back to the load
    // branch opcodes: bits 7-6 the flag (N, V, C, Z), bit 5 the value taken on
    const uint32_t fl = bop >> 6;
    if (fl == 2u || fl == (BIT ? 3u : 1u)) return;         // C; Z after BIT, V after a load: per lane
    // taken when ((v & m) != 0) == want
    const uint32_t m = fl == 0u ? 0x80u : (fl == 1u ? 0x40u : 0xFFu);
    const uint32_t want = ((bop >> 5) & 1u) ^ (fl == 3u ? 1u : 0u);
    const bool intim = (a13 & 1u) == 0u;
    constexpr uint32_t kIter = kLoad + kBranch, kLimit = LOOPB_CAP - kIterCycles;
    while (u.cycles <= kLimit - kIter) {
        const uint32_t v = intim ? (u.tval & 0xFFu) : (u.tuf ? 0x80u : 0u);
        if (((v & m) != 0u ? 1u : 0u) != want) return;
        if (intim) {
            u.tuf = 0;
            u.tlim = u.tps;
        }
        // k iterations that end no line (clock + 21 k <= 227), take the
        // timer's fast path at every instruction (tlim > 7 k) and stay below
        // the limit change only the counts, the TIA clock and the prescaler
        // count, and leave the value read (and so the branch) as it is
        uint32_t k = (227u - u.clock) / (3u * kIter);
        k = __builtin_elementwise_min(k, (u.tlim - 1u) / kIter);
        k = __builtin_elementwise_min(k, (kLimit - u.cycles) / kIter);
        if (u.tlim != 0u && k != 0u) {
            u.cycles += kIter * k;
            u.instr += 2u * k;
            u.clock += 3u * kIter * k;
            u.tps -= kIter * k;
            u.tlim -= kIter * k;
            continue;
        }
        u_step(u, e, kLoad);
        u_step(u, e, kBranch);
    }
}

template <uint32_t D0, uint32_t D1>
__device__ __forceinline__ bool u_compute(UCtl &u, Emu &e, uint32_t b1, uint32_t b2, uint32_t pcn, uint32_t &cyc) {
    constexpr uint32_t srcsel = (D1 >> D1_SRC_SHIFT) & 7u;
    constexpr uint32_t dst = (D1 >> D1_DST_SHIFT) & 7u;
    const uint32_t s = srcsel == R_A ? e.a : srcsel == R_X ? e.x : srcsel == R_Y ? e.y : u.sp;
    if constexpr ((D0 & D_MEM) == 0u) {
        if constexpr (dst == R_SP) {                     // TXS: the stack pointer is wave-uniform
            const uint32_t su = wave_first(s);
            if (!same_on_all(s, su)) return false;
            u.sp = su;
        } else {
            u_alu<D0, D1>(e, s, (D0 & D_IMM) ? b1 : s);
        }
        u.pc = pcn;
        return true;
    } else {
        constexpr uint32_t ind = (D0 >> D_IND_SHIFT) & 3u;
        constexpr uint32_t ixsel = (D0 >> D_IDX_SHIFT) & 3u;
        constexpr bool zp = (D0 & D_ZP) != 0u;
        if constexpr (ind == 0u) {
            const uint32_t base = zp ? b1 : (b1 | (b2 << 8));
            uint32_t iu = 0;
            if constexpr (ixsel != 0u) {
                const uint32_t idx = ixsel == 1u ? e.x : e.y;
                iu = wave_first(idx);
                if (!same_on_all(idx, iu)) {
                    const uint32_t ea = (base + idx) & (zp ? 0xFFu : 0xFFFFu);
                    return u_mem_lane<D0, D1>(u, e, s, base, ea, pcn, cyc);
                }
            }
            constexpr uint32_t kind = (D0 >> D_KIND_SHIFT) & 7u;
            constexpr bool poll = ixsel == 0u && !zp && (D0 & D_RD) != 0u && (D0 & D_WR) == 0u &&
                                  (kind == K_BIT || (kind == K_LOGIC && (D0 & D_M_LD) != 0u));
            if constexpr (poll) {
                // INTIM / INSTAT, the addresses u_io_read answers from the timer
                if (__builtin_expect((base & 0x1286u) == 0x284u, 0)) u_poll<kind == K_BIT>(u, e, base & 0x1FFFu, pcn);
            }
            const uint32_t ea = (base + iu) & (zp ? 0xFFu : 0xFFFFu);
            return u_mem_uni<D0, D1>(u, e, s, base, ea, pcn, cyc);
        } else {
            // (zp,X) / (zp),Y: the pointer bytes come from RAM, so the address is per lane
            const uint32_t zpa = ind == 1u ? ((b1 + e.x) & 0xFFu) : b1;
            const uint32_t zp1 = (zpa + 1u) & 0xFFu;
            if (__ballot(!(zpa & zp1 & 0x80u)) != 0) return false;     // a pointer byte in the TIA
            const uint32_t base = ld8(e.ram_off(zpa)) | (ld8(e.ram_off(zp1)) << 8);
            const uint32_t ea = ind == 2u ? ((base + e.y) & 0xFFFFu) : base;
            return u_mem_lane<D0, D1>(u, e, s, base, ea, pcn, cyc);
        }
    }
}

// ---------------------------------------------------------------- branches, stack, jumps
template <uint32_t D1>
__device__ __forceinline__ bool u_branch(UCtl &u, Emu &e, uint32_t b1, uint32_t pcn, uint32_t &cyc) {
    constexpr uint32_t fm = (D1 >> D1_FMASK_SHIFT) & 0xFFu;
    constexpr bool fv = (D1 >> D1_FVAL_SHIFT) != 0u;
    const uint64_t tk = __ballot(((e.p & fm) != 0u) == fv);
    if (tk == 0) { u.pc = pcn; return true; }
    if (tk != active_lanes()) return false;           // lanes disagree: per-lane path
    // the phantom read of PC (pcn lies below the hotspot area) has no effect
    const uint32_t np = (pcn + b1 - ((b1 & 0x80u) << 1)) & 0xFFFFu;
    if ((np ^ pcn) & 0xFF00u) {
        u_touch(u, e, ((pcn & 0xFF00u) | (np & 0xFFu)) & 0x1FFFu);
        cyc += 2u;
    } else {
        cyc += 1u;
    }
    u.pc = np;
    return true;
}

template <uint32_t D1>
__device__ __forceinline__ bool u_special(UCtl &u, Emu &e, uint32_t b1, uint32_t b2, uint32_t pcn) {
    constexpr uint32_t spec = (D1 >> D1_SPEC_SHIFT) & 15u;
    const uint32_t sp = u.sp;
    const uint32_t s1 = (sp + 1u) & 0xFFu, s2 = (sp + 2u) & 0xFFu, s3 = (sp + 3u) & 0xFFu;
    const uint32_t m1 = (sp - 1u) & 0xFFu, m2 = (sp - 2u) & 0xFFu;
    if constexpr (spec == SP_JMP) {
        u.pc = b1 | (b2 << 8);
    } else if constexpr (spec == SP_JMPI) {
        const uint32_t ptr = b1 | (b2 << 8);
        if (!(ptr & 0x1000u)) return false;              // pointer outside the ROM window
        const uint32_t lo = u_rom_read(u, e, ptr & 0x1FFFu);
        const uint32_t hi = u_rom_read(u, e, ((ptr & 0xFF00u) | ((ptr + 1u) & 0xFFu)) & 0x1FFFu);
        u.pc = lo | (hi << 8);
    } else if constexpr (spec == SP_JSR) {
        if (!(sp & m1 & 0x80u)) return false;            // stack bytes in the TIA
        const uint32_t ret = (pcn - 1u) & 0xFFFFu;
        st8(e.ram_off(sp), ret >> 8);
        st8(e.ram_off(m1), ret & 0xFFu);
        u.sp = m2;
        u.pc = b1 | (b2 << 8);
    } else if constexpr (spec == SP_RTS) {
        if (!(s1 & s2 & 0x80u)) return false;
        const uint32_t ad = ld8(e.ram_off(s1)) | (ld8(e.ram_off(s2)) << 8);
        const uint32_t au = wave_first(ad);
        if (!same_on_all(ad, au)) return false;
        u_touch(u, e, au & 0x1FFFu);
        u.sp = s2;
        u.pc = (au + 1u) & 0xFFFFu;
    } else if constexpr (spec == SP_RTI) {
        if (!(s1 & s2 & s3 & 0x80u)) return false;
        const uint32_t ad = ld8(e.ram_off(s2)) | (ld8(e.ram_off(s3)) << 8);
        const uint32_t au = wave_first(ad);
        if (!same_on_all(ad, au)) return false;
        e.p = ld8(e.ram_off(s1)) & ~(FB | FU);
        u.sp = s3;
        u.pc = au;
    } else if constexpr (spec == SP_BRK) {
        if (!(sp & m1 & m2 & 0x80u)) return false;
        const uint32_t ret = (pcn + 1u) & 0xFFFFu;       // the padding byte is skipped
        st8(e.ram_off(sp), ret >> 8);
        st8(e.ram_off(m1), ret & 0xFFu);
        st8(e.ram_off(m2), e.p | FB | FU);
        u.sp = (sp - 3u) & 0xFFu;
        e.p |= FI;
        const uint32_t vb = ROM0 + (u.bank << 12);       // $FFFE/$FFFF are never hotspots
        u.pc = first_lane(ld8(vb + 0xFFEu)) | (first_lane(ld8(vb + 0xFFFu)) << 8);
    } else if constexpr (spec == SP_PHA || spec == SP_PHP) {
        if (!(sp & 0x80u)) return false;
        st8(e.ram_off(sp), spec == SP_PHA ? e.a : (e.p | FB | FU));
        u.sp = m1;
        u.pc = pcn;
    } else {                                             // PLA, PLP
        if (!(s1 & 0x80u)) return false;
        const uint32_t v = ld8(e.ram_off(s1));
        u.sp = s1;
        if constexpr (spec == SP_PLA) {
            e.a = v;
            e.set_zn(v);
        } else {
            e.p = v & ~(FB | FU);
        }
        u.pc = pcn;
    }
    return true;
}

// One instruction with its descriptor known at compile time.  Returns its
// cycle count, or 0 (nothing changed) when it needs the per-lane path.
template <uint64_t D>
__device__ __forceinline__ uint32_t u_op(UCtl &u, Emu &e, uint32_t b1, uint32_t b2) {
    constexpr uint32_t D0 = (uint32_t)D, D1 = (uint32_t)(D >> 32);
    constexpr uint32_t cls = D0 & 3u;
    const uint32_t pcn = (u.pc + ((D0 >> D_LEN_SHIFT) & 3u)) & 0xFFFFu;
    uint32_t cyc = (D0 >> D_CYC_SHIFT) & 7u;
    bool ok;
    if constexpr (cls == CLS_ILLEGAL) ok = false;
    else if constexpr (cls == CLS_BRANCH) ok = u_branch<D1>(u, e, b1, pcn, cyc);
    else if constexpr (cls == CLS_SPECIAL) ok = u_special<D1>(u, e, b1, b2, pcn);
    else ok = u_compute<D0, D1>(u, e, b1, b2, pcn, cyc);
    return ok ? cyc : 0u;
}

// ---------------------------------------------------------------- frame loop
// HeadlessAtari::run_frame's loop (src/atari.rs:584-614) on the uniform state.
// Returns U_DONE when the frame ended (st set as the per-lane loop would), or
// U_EXIT at the top of an iteration whose instruction needs the per-lane path
// (or, with the lanes' cycle counts spread, whose caps are near).
__device__ __forceinline__ uint32_t u_run(UCtl &u, Emu &e, uint32_t &st) {
    // The optimiser may merge the entry values with per-lane values on the
    // path where the entry test failed; taking every field through an opaque
    // readfirstlane here keeps the whole loop in scalar registers.
    u.pc = wave_first(u.pc); u.sp = wave_first(u.sp); u.bank = wave_first(u.bank);
    u.clock = wave_first(u.clock); u.line = wave_first(u.line); u.wsync = wave_first(u.wsync);
    u.fcomp = wave_first(u.fcomp); u.vsync = wave_first(u.vsync);
    u.tval = wave_first(u.tval); u.tps = wave_first(u.tps); u.tsel = wave_first(u.tsel);
    u.tsh = wave_first(u.tsh); u.tuf = wave_first(u.tuf);
    u.tlim = u_tlim(u);
    u.cycles = wave_first(u.cycles); u.in_a = wave_first(u.in_a); u.cspread = wave_first(u.cspread);
    u.skip = 0;
    u.instr = 0;
    // opcode and operand bytes at PC, loaded one iteration ahead so that the
    // LDS latency overlaps the previous instruction's timing work (the address
    // stays inside the LDS allocation for any PC and bank)
    uint32_t f0, f1, f2;
    {
        const uint32_t o = ROM0 + (u.bank << 12) + (u.pc & 0xFFFu);
        f0 = ld8(o); f1 = ld8(o + 1u); f2 = ld8(o + 2u);
    }
    while (true) {
        // one test for two rare events: WSYNC (bit 0), and VSYNC disagreeing
        // with the loop (bit 1): loop A runs while VSYNC is held and leaves
        // when it drops; loop B ends the frame when VSYNC rises
        if (__builtin_expect(u.wsync != 0u, 0)) {
            if (u.wsync & 2u) {
                if (!u.in_a) return U_DONE;            // VSYNC rose in loop B: the frame ends
                // VSYNC dropped: loop A is over, and loop B's cap may be near
                if (u.cspread && u.cycles > LOOPB_CAP - kIterCycles) return U_EXIT;
                u.in_a = 0;
                u.wsync &= 1u;
            }
            if (u.wsync) {
                const uint32_t s = u_tia_skip(u, e);
                if (s == 0u) { st = ENV_VSYNC_HELD; return U_DONE; }
                timer_tick(u, s);
                u.tlim = u_tlim(u);
                u.cycles += s;
                u.skip += s;
                goto caps;
            }
        }
        {
            // opcode and operands straight from the LDS ROM image: below the
            // hotspot / vector area these fetches have no side effects
            if (__builtin_expect(((u.pc & 0x1FFFu) - 0x1000u) >= 0xFF0u, 0)) return U_EXIT;
            const uint32_t op = first_lane(f0), b1 = first_lane(f1), b2 = first_lane(f2);
            // a handler that declines (0 cycles) leaves from its own case, so
            // the common path after the dispatch carries no test for it
            uint32_t cyc = 0;
            switch (op) {
#define FA_UCASE(OPC)                                                   \
            case OPC:                                                   \
                cyc = u_op<descriptor_of(OPC)>(u, e, b1, b2);           \
                if (__builtin_expect(cyc == 0u, 0)) return U_EXIT;      \
                break;
                FA_FOR_EACH_OPCODE(FA_UCASE)
#undef FA_UCASE
            default: return U_EXIT;
            }
            u_step(u, e, cyc);
            const uint32_t o = ROM0 + (u.bank << 12) + (u.pc & 0xFFFu);
            f0 = ld8(o); f1 = ld8(o + 1u); f2 = ld8(o + 2u);
        }
        // the frame-loop caps (loop A's is the larger)
        static_assert(LOOPA_CAP > LOOPB_CAP, "cap order");
    caps:
        if (__builtin_expect(u.cycles > LOOPB_CAP - kIterCycles, 0)) {
            if (u.cspread) {
                // no lane has passed its cap yet; the per-lane loop goes on
                if (!u.in_a || u.cycles > LOOPA_CAP - kIterCycles) return U_EXIT;
            } else if (u.cycles > LOOPB_CAP) {
                if (!u.in_a) return U_DONE;
                if (u.cycles > LOOPA_CAP) { st = ENV_VSYNC_HELD; return U_DONE; }
            }
        }
    }
}

}  // namespace fa
