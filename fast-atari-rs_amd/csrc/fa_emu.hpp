// fa_emu.hpp -- device-side Atari 2600 emulator, one wave64 lane per env.
//
// Behaviour follows the reference's CPU emulator, which is the parity target:
//   6502 core      src/cpu.rs:245-1146 (every bus access with a side effect,
//                  in the reference's order; side-effect-free phantom reads
//                  of RAM/TIA/ROM-outside-hotspots are elided)
//   headless TIA   src/headless_tia.rs:153-298
//   RIOT / PIA     src/pia.rs:107-228
//   bus            src/atari.rs:459-553 (HeadlessBus read/write decode)
//   frame loop     src/atari.rs:584-614
//
// Instructions run through one shared pipeline (fetch, descriptor decode,
// effective address, data read, compute, write-back) steered by the per-opcode
// descriptors of fa_decode.hpp with predicated selects, so lanes executing
// different opcodes stay on the same code path (branches, JMP and the
// push / pull instructions included); only TIA / RIOT register writes, RIOT
// reads, decimal-mode arithmetic and JSR / RTS / RTI / BRK / JMP (ind) branch.
//
// LDS layout of one workgroup (byte offsets into fa_lds):
//   [0, 4096)                     opcode descriptors (256 x 16 bytes)
//   [4096, 4096 + rom_bytes)      ROM image: the 4 KB window of a Fixed ROM,
//                                 or every 4 KB bank of an F8/F6/F4 ROM
//   [.., + bs * 222)              the workgroup's AtariStateGpu images (AoS,
//                                 222-byte stride), staged with 16-byte loads
//   [.., + bs)                    one discard byte per lane (predicated stores)
// Hot fields (registers, TIA timing, RIOT timer, bank) live in VGPRs for the
// whole launch; RAM and the rarely touched TIA/RIOT registers stay in the
// lane's LDS image.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fa_decode.hpp"

namespace fa {

constexpr uint32_t DESC0 = 0;        // LDS offset of the descriptor table
constexpr uint32_t ROM0 = kDescBytes; // LDS offset of the ROM image
// A byte that always reads 0: the second word of an undocumented opcode's
// descriptor ($02).
constexpr uint32_t ZERO_BYTE = DESC0 + 0x02u * 32u + 4u;

// Byte offsets inside the 222-byte AtariStateGpu (src/cuda_env.rs:12-38).
enum : uint32_t {
    S_A = 0, S_X = 1, S_Y = 2, S_SP = 3, S_PC = 4, S_P = 6, S_RAM = 8,
    S_PA_OUT = 136, S_PA_DDR = 137, S_PA_IN = 138, S_PB_OUT = 139, S_PB_DDR = 140, S_PB_IN = 141,
    S_TVAL = 142, S_TPS = 144, S_TSEL = 146, S_TUF = 148,
    S_CLOCK = 152, S_LINE = 154, S_WSYNC = 156, S_FCOMP = 157, S_VSYNC = 158, S_VBLANK = 159,
    S_CX = 160,
    S_INPT4 = 168, S_INPT5 = 169, S_PAD0 = 170, S_PAD1 = 171, S_PCOUNT = 172, S_PDUMP = 174,
    S_POS = 176,
    S_HM = 186,
    S_GRP0 = 192, S_GRP1 = 193, S_GRP0_OLD = 194, S_GRP1_OLD = 195,
    S_ENAM0 = 196, S_ENAM1 = 197, S_ENABL = 198, S_ENABL_OLD = 199,
    S_VDELP0 = 200, S_VDELP1 = 201, S_VDELBL = 202, S_RESMP0 = 203, S_RESMP1 = 204,
    S_NUSIZ0 = 205, S_NUSIZ1 = 206, S_COLUP0 = 207, S_COLUP1 = 208, S_COLUPF = 209, S_COLUBK = 210,
    S_CTRLPF = 211, S_REFP0 = 212, S_REFP1 = 213, S_PF0 = 214, S_PF1 = 215, S_PF2 = 216,
    S_BANK = 217, S_SCHEME = 218,
    STATE_BYTES = 222,
};

enum : uint32_t { ENV_OK = 0, ENV_ILLEGAL = 1, ENV_VSYNC_HELD = 2 };
constexpr uint32_t LOOPB_CAP = 100000u;     // src/atari.rs:594
constexpr uint32_t LOOPA_CAP = 1000000u;    // uncapped in the reference

// 6502 status bits (src/cpu.rs:17-26)
enum : uint32_t { FC = 0x01, FZ = 0x02, FI = 0x04, FD = 0x08, FB = 0x10, FU = 0x20, FV = 0x40, FN = 0x80 };

extern __shared__ __attribute__((aligned(16))) uint8_t fa_lds[];

__device__ __forceinline__ uint32_t ld8(uint32_t off) { return fa_lds[off]; }
__device__ __forceinline__ void st8(uint32_t off, uint32_t v) { fa_lds[off] = (uint8_t)v; }
__device__ __forceinline__ uint32_t ld32(uint32_t off) {
    return *reinterpret_cast<const uint32_t *>(fa_lds + off);
}
__device__ __forceinline__ uint4 ld128(uint32_t off) {
    return *reinterpret_cast<const uint4 *>(fa_lds + off);
}
__device__ __forceinline__ uint32_t ld16(uint32_t off) {
    return (uint32_t)fa_lds[off] | ((uint32_t)fa_lds[off + 1] << 8);
}
__device__ __forceinline__ void st16(uint32_t off, uint32_t v) {
    fa_lds[off] = (uint8_t)v;
    fa_lds[off + 1] = (uint8_t)(v >> 8);
}

// All-ones when c holds.  Selections are written as AND/OR of masks so the
// compiler emits straight-line VALU code instead of exec-mask branches
// (chains of ?: are otherwise turned into switches and short-circuit && / ||
// into branches).
__device__ __forceinline__ uint32_t mk(bool c) { return 0u - (uint32_t)c; }
__device__ __forceinline__ uint32_t pick(bool c, uint32_t t, uint32_t f) { return (t & mk(c)) | (f & ~mk(c)); }
// All-ones when bit k of w is set (one v_bfe_i32).
__device__ __forceinline__ uint32_t bm(uint32_t w, uint32_t k) { return (uint32_t)((int32_t)(w << (31u - k)) >> 31); }

// ---------------------------------------------------------------- RIOT timer
// Pia::tick_n (src/pia.rs:107-166) with the prescaler division done by shifts
// (select is 1, 8, 64 or 1024; tsh = log2(select)).  T is any holder of the
// timer fields: the per-lane Emu, or the wave-uniform control state.
template <class T>
__device__ __forceinline__ void timer_tick_slow(T &t, uint32_t n) {
    if (n == 0) return;
    if (t.tval == 0 && t.tuf) {
        t.tval = (0x100u - n) & 0xFFu;
        return;
    }
    uint32_t rem = n;
    if (t.tps > 0) {
        if (rem < t.tps) { t.tps -= rem; return; }
        rem -= t.tps;
        t.tps = 0;
        if (t.tval == 0) {
            t.tuf = 1;
            t.tval = rem ? ((0x100u - rem) & 0xFFu) : 0xFFu;
            return;
        }
        t.tval -= 1;
    }
    if (rem == 0) { t.tps = t.tsel; return; }
    const uint32_t full = rem >> t.tsh;
    const uint32_t left = rem & (t.tsel - 1u);
    if (full >= t.tval) {
        rem -= t.tval << t.tsh;
        t.tuf = 1;
        t.tval = rem ? ((0x100u - rem) & 0xFFu) : 0xFFu;
    } else {
        t.tval -= full;
        t.tps = t.tsel - left;
    }
}
template <class T>
__device__ __forceinline__ void timer_tick(T &t, uint32_t n) {
    const bool post = t.tval == 0 && t.tuf;
    if (__builtin_expect(!post && t.tps > n, 1)) { t.tps -= n; return; }   // common case: inside a prescaler period
    timer_tick_slow(t, n);
}
// Timer write (Pia::write with addr & 0x10, src/pia.rs:202-212).
template <class T>
__device__ __forceinline__ void timer_set(T &t, uint32_t addr, uint32_t v) {
    t.tval = v;
    t.tuf = 0;
    t.tsh = (0x0A060300u >> ((addr & 3u) * 8u)) & 0xFFu;   // 1, 8, 64, 1024
    t.tsel = 1u << t.tsh;
    t.tps = t.tsel;
}

// (clock - 68 + 5) rem_euclid 160 for clock < 228 (src/headless_tia.rs:257-261)
__device__ __forceinline__ uint32_t resp_pos_of(uint32_t clock) {
    uint32_t t = clock + 97u;
    t = t >= 160u ? t - 160u : t;
    return t >= 160u ? t - 160u : t;
}

struct Emu {
    // CPU
    uint32_t a, x, y, sp, p, pc;
    // TIA timing / frame detection
    uint32_t clock, line, wsync, fcomp, vsync, pcount, pdump;
    // RIOT interval timer; tsh = log2(select)
    uint32_t tval, tps, tsel, tsh, tuf;
    // cartridge
    uint32_t bank;
    // addressing: LDS offsets of this lane's image and discard byte; hotspot window
    uint32_t img, junk;
    uint32_t hot_lo, hot_n;

    // ------------------------------------------------------------ state I/O
    __device__ __forceinline__ void load(uint32_t img_off, uint32_t junk_off, uint32_t hot_lo_, uint32_t hot_n_) {
        img = img_off;
        junk = junk_off;
        hot_lo = hot_lo_;
        hot_n = hot_n_;
        a = ld8(img + S_A); x = ld8(img + S_X); y = ld8(img + S_Y); sp = ld8(img + S_SP);
        pc = ld16(img + S_PC); p = ld8(img + S_P);
        tval = ld16(img + S_TVAL); tps = ld16(img + S_TPS); tsel = ld16(img + S_TSEL); tuf = ld8(img + S_TUF);
        tsh = 31u - __builtin_clz(tsel | 1u);
        clock = ld16(img + S_CLOCK); line = ld16(img + S_LINE);
        wsync = ld8(img + S_WSYNC); fcomp = ld8(img + S_FCOMP); vsync = ld8(img + S_VSYNC);
        pcount = ld16(img + S_PCOUNT); pdump = ld8(img + S_PDUMP);
        bank = ld8(img + S_BANK);
    }
    __device__ __forceinline__ void store() const {
        st8(img + S_A, a); st8(img + S_X, x); st8(img + S_Y, y); st8(img + S_SP, sp);
        st16(img + S_PC, pc); st8(img + S_P, p);
        st16(img + S_TVAL, tval); st16(img + S_TPS, tps); st16(img + S_TSEL, tsel); st8(img + S_TUF, tuf);
        st16(img + S_CLOCK, clock); st16(img + S_LINE, line);
        st8(img + S_WSYNC, wsync); st8(img + S_FCOMP, fcomp); st8(img + S_VSYNC, vsync);
        st16(img + S_PCOUNT, pcount); st8(img + S_PDUMP, pdump);
        st8(img + S_BANK, bank);
    }

    // HeadlessAtari::set_action (src/atari.rs:571-582) on action index 0..17.
    __device__ __forceinline__ void set_action(uint32_t act) {
        // bit 0 up, 1 down, 2 left, 3 right, 4 fire (src/atari.rs:60-79)
        constexpr uint32_t kDec[18] = {0x00, 0x10, 0x01, 0x08, 0x04, 0x02, 0x09, 0x05, 0x0A,
                                       0x06, 0x11, 0x18, 0x14, 0x12, 0x19, 0x15, 0x1A, 0x16};
        uint32_t d = act < 18 ? kDec[act] : 0u;
        uint32_t swcha = 0xFFu & ~(((d & 1u) << 4) | ((d & 2u) << 4) | ((d & 4u) << 4) | ((d & 8u) << 4));
        st8(img + S_PA_IN, swcha);
        st8(img + S_INPT4, (d & 0x10u) ? 0u : 1u);
        st8(img + S_PAD0, (d & 4u) ? 200u : ((d & 8u) ? 50u : 128u));
    }

    // ------------------------------------------------------------ RIOT
    __device__ __forceinline__ void pia_tick(uint32_t n) { timer_tick(*this, n); }

    // Pia::read for I/O addresses (src/pia.rs:169-199)
    __device__ __forceinline__ uint32_t pia_read(uint32_t addr) {
        switch (addr & 7u) {
        case 0: { uint32_t d = ld8(img + S_PA_DDR); return (ld8(img + S_PA_IN) & ~d & 0xFFu) | (ld8(img + S_PA_OUT) & d); }
        case 1: return ld8(img + S_PA_DDR);
        case 2: { uint32_t d = ld8(img + S_PB_DDR); return (ld8(img + S_PB_IN) & ~d & 0xFFu) | (ld8(img + S_PB_OUT) & d); }
        case 3: return ld8(img + S_PB_DDR);
        case 4: tuf = 0; return tval & 0xFFu;
        case 5: return tuf ? 0x80u : 0u;
        default: return 0u;
        }
    }

    // Pia::write for I/O addresses (src/pia.rs:202-228)
    __device__ __forceinline__ void pia_write(uint32_t addr, uint32_t v) {
        if (addr & 0x10u) {
            timer_set(*this, addr, v);
            return;
        }
        // SWCHA, SWACNT, SWCHB, SWBCNT -> port_a_output, port_a_ddr, port_b_output, port_b_ddr
        uint32_t r = addr & 3u;
        st8(img + S_PA_OUT + r + (r >> 1), v);
    }

    // ------------------------------------------------------------ TIA
    // HeadlessTia::tick_n + end_scanline (src/headless_tia.rs:153-198)
    __device__ __forceinline__ void end_scanline() {
        clock -= 228u;
        line += 1u;
        wsync = 0;
        if (!pdump && pcount < 256u) pcount += 1u;
        if (line >= 262u) { line = 0; fcomp = 1; }
    }
    __device__ __forceinline__ void tia_tick(uint32_t n) {
        // one line crossing at most (n <= 7); end_scanline written as masks
        clock += 3u * n;
        const uint32_t w = mk(clock >= 228u);
        clock -= 228u & w;
        line += 1u & w;
        wsync &= ~w;
        pcount += (uint32_t)(pdump == 0u) & (uint32_t)(pcount < 256u) & w;
        const uint32_t wrap = mk(line >= 262u);
        line &= ~wrap;
        fcomp |= 1u & wrap;
    }
    // skip_to_scanline_end (src/headless_tia.rs:163-180)
    __device__ __forceinline__ uint32_t tia_skip() {
        uint32_t rem = clock >= 228u ? 0u : 228u - clock;
        uint32_t cyc = (rem + 2u) / 3u;
        if (cyc) { clock = 228u; end_scanline(); }
        return cyc;
    }

    // HeadlessTia::read (src/headless_tia.rs:205-223)
    __device__ __forceinline__ uint32_t tia_read(uint32_t addr) {
        uint32_t r = addr & 0x0Fu;
        if (r < 8u) return ld8(img + S_CX + r);
        if (r < 10u) return pcount >= ld8(img + S_PAD0 + (r - 8u)) ? 0x80u : 0u;
        if (r < 12u) return 0x80u;
        if (r < 14u) return ld8(img + S_INPT4 + (r - 12u)) ? 0x80u : 0u;
        return 0u;
    }


    // HeadlessTia::write (src/headless_tia.rs:225-298)
    __device__ __forceinline__ void tia_write(uint32_t addr, uint32_t v) {
        uint32_t r = addr & 0x3Fu;
        switch (r) {
        case 0x00:
            if (!(vsync & 2u) && (v & 2u)) { line = 0; fcomp = 1; }
            vsync = v;
            break;
        case 0x01:
            if (v & 0x80u) { pcount = 0; pdump = 1; }
            else if (pdump) pdump = 0;
            st8(img + S_VBLANK, v);
            break;
        case 0x02: wsync = 1; break;
        case 0x04: case 0x05: case 0x06: case 0x07: case 0x08: case 0x09: case 0x0A:
            st8(img + S_NUSIZ0 + (r - 0x04u), v);          // NUSIZ0..CTRLPF are contiguous
            break;
        case 0x0B: case 0x0C: st8(img + S_REFP0 + (r - 0x0Bu), (v >> 3) & 1u); break;
        case 0x0D: case 0x0E: case 0x0F: st8(img + S_PF0 + (r - 0x0Du), v); break;
        case 0x10: case 0x11: case 0x12: case 0x13: case 0x14:
            st16(img + S_POS + 2u * (r - 0x10u), resp_pos_of(clock));
            break;
        case 0x1B:
            st8(img + S_GRP0_OLD, ld8(img + S_GRP0));
            st8(img + S_GRP0, v);
            break;
        case 0x1C:
            st8(img + S_GRP1_OLD, ld8(img + S_GRP1));
            st8(img + S_GRP1, v);
            st8(img + S_ENABL_OLD, ld8(img + S_ENABL));
            break;
        case 0x1D: case 0x1E: case 0x1F: st8(img + S_ENAM0 + (r - 0x1Du), (v >> 1) & 1u); break;
        case 0x20: case 0x21: case 0x22: case 0x23: case 0x24: st8(img + S_HM + (r - 0x20u), v); break;
        case 0x25: case 0x26: case 0x27: st8(img + S_VDELP0 + (r - 0x25u), v & 1u); break;
        case 0x28: case 0x29: st8(img + S_RESMP0 + (r - 0x28u), (v >> 1) & 1u); break;
        case 0x2A:
            // HMOVE: pos - ((hm as i8) >> 4), rem_euclid 160
            for (uint32_t i = 0; i < 5u; ++i) {
                int32_t off = ((int32_t)(int8_t)ld8(img + S_HM + i)) >> 4;
                int32_t q = (int32_t)ld16(img + S_POS + 2u * i) - off;
                q += q < 0 ? 160 : 0;
                q -= q >= 160 ? 160 : 0;
                st16(img + S_POS + 2u * i, (uint32_t)q);
            }
            break;
        case 0x2B:
            for (uint32_t i = 0; i < 5u; ++i) st8(img + S_HM + i, 0);
            break;
        case 0x2C:
            for (uint32_t i = 0; i < 8u; ++i) st8(img + S_CX + i, 0);
            break;
        default: break;
        }
    }

    // HeadlessTia::read / write for register numbers that differ between
    // lanes (the per-lane paths).  tia_read / tia_write above branch on the
    // register, which is cheap when it is wave-uniform but, on a per-lane
    // register, runs the whole case chain under exec masks whenever any lane
    // touches the TIA.  Here the single-byte registers are one computed load
    // or one predicated store with selects; only the multi-byte ones (RESPx,
    // GRP0 / GRP1, HMOVE / HMCLR / CXCLR) branch, on a ballot.
    // LDS offset of the image byte behind TIA read register addr & 15: CXxx at
    // S_CX + r (r < 8), paddles at S_PAD0 + r - 8, INPT4/5 at S_INPT4 + r - 12
    // (registers 14 and 15 read as 0 and point anywhere in the image).
    __device__ __forceinline__ uint32_t tia_read_off(uint32_t addr) const {
        const uint32_t r = addr & 0x0Fu;
        const uint32_t off = pick(r < 8u, S_CX + r, pick(r < 10u, S_PAD0 - 8u + r, S_INPT4 - 12u + r));
        return img + pick(r < 14u, off, S_CX);
    }
    // The register's value from that byte.
    __device__ __forceinline__ uint32_t tia_read_val(uint32_t addr, uint32_t x) const {
        const uint32_t r = addr & 0x0Fu;
        const uint32_t pad = pcount >= x ? 0x80u : 0u;
        const uint32_t inp = x ? 0x80u : 0u;
        return pick(r < 8u, x, pick(r < 10u, pad, pick(r < 12u, 0x80u, pick(r < 14u, inp, 0u))));
    }
    __device__ __forceinline__ uint32_t tia_read_lanes(uint32_t addr) {
        return tia_read_val(addr, ld8(tia_read_off(addr)));
    }
    __device__ __forceinline__ void tia_write_lanes(uint32_t addr, uint32_t v) {
        const uint32_t r = addr & 0x3Fu;
        // single-byte registers: NUSIZ0..PF2 (4-15), ENAM0/ENAM1/ENABL (0x1D-0x1F),
        // HMxx (0x20-0x24), VDELxx / RESMPx (0x25-0x29) sit at S_ offsets linear in r
        const bool g1 = (r - 4u) < 12u, g2 = (r - 0x1Du) < 3u, g3 = (r - 0x20u) < 5u, g4 = (r - 0x25u) < 5u;
        const bool vbl = r == 0x01u;
        const uint32_t off = pick(g1, 201u + r, pick(g2, 167u + r, pick(g3, 154u + r, pick(g4, 163u + r, S_VBLANK))));
        const bool sh3 = (r - 0x0Bu) < 2u;                          // REFP0/1: bit 3
        const bool sh1 = g2 | ((r - 0x28u) < 2u);                  // ENAMx/ENABL, RESMPx: bit 1
        const bool bit0 = (r - 0x25u) < 3u;                         // VDELxx: bit 0
        const uint32_t val = pick(sh3, (v >> 3) & 1u, pick(sh1, (v >> 1) & 1u, pick(bit0, v & 1u, v)));
        st8((g1 | g2 | g3 | g4 | vbl) ? img + off : junk, val);
        // VSYNC: a rising bit 1 starts a frame; VBLANK bit 7 dumps the paddle; WSYNC
        const bool vs = r == 0x00u;
        const bool rise = vs & !(vsync & 2u) & ((v & 2u) != 0u);
        line = pick(rise, 0u, line);
        fcomp = pick(rise, 1u, fcomp);
        vsync = pick(vs, v, vsync);
        const bool dump = (v & 0x80u) != 0u;
        pcount = pick(vbl & dump, 0u, pcount);
        pdump = pick(vbl, dump ? 1u : 0u, pdump);
        wsync = pick(r == 0x02u, 1u, wsync);
        // RESP0..RESBL: position from the clock at the start of the instruction
        const bool resp = (r - 0x10u) < 5u;
        if (__builtin_expect(__ballot(resp) != 0, 0)) {
            const uint32_t pos = resp_pos_of(clock);
            const uint32_t po = img + S_POS + 2u * (r - 0x10u);
            st8(resp ? po : junk, pos);
            st8(resp ? po + 1u : junk, pos >> 8);
        }
        const bool gp0 = r == 0x1Bu, gp1 = r == 0x1Cu;
        if (__builtin_expect(__ballot(gp0 | gp1) != 0, 0)) {
            // GRP0: grp0_old = grp0; GRP1: grp1_old = grp1 and enabl_old = enabl
            const uint32_t g = ld8(img + (gp0 ? S_GRP0 : S_GRP1));
            st8((gp0 | gp1) ? img + (gp0 ? S_GRP0_OLD : S_GRP1_OLD) : junk, g);
            st8((gp0 | gp1) ? img + (gp0 ? S_GRP0 : S_GRP1) : junk, v);
            const uint32_t en = ld8(img + S_ENABL);
            st8(gp1 ? img + S_ENABL_OLD : junk, en);
        }
        if (__builtin_expect(__ballot((r - 0x2Au) < 3u) != 0, 0)) {
            if (r == 0x2Au) {
                // HMOVE: pos - ((hm as i8) >> 4), rem_euclid 160
                for (uint32_t i = 0; i < 5u; ++i) {
                    int32_t o = ((int32_t)(int8_t)ld8(img + S_HM + i)) >> 4;
                    int32_t q = (int32_t)ld16(img + S_POS + 2u * i) - o;
                    q += q < 0 ? 160 : 0;
                    q -= q >= 160 ? 160 : 0;
                    st16(img + S_POS + 2u * i, (uint32_t)q);
                }
            } else if (r == 0x2Bu) {
                for (uint32_t i = 0; i < 5u; ++i) st8(img + S_HM + i, 0);
            } else if (r == 0x2Cu) {
                for (uint32_t i = 0; i < 8u; ++i) st8(img + S_CX + i, 0);
            }
        }
    }

    // ------------------------------------------------------------ bus
    __device__ __forceinline__ uint32_t ram_off(uint32_t a) const { return img + S_RAM + (a & 0x7Fu); }

    // A ROM access at a hotspot switches the bank (HeadlessBus::check_bankswitch,
    // src/atari.rs:459-492); hot_n = 0 for Fixed ROMs.
    __device__ __forceinline__ void hotspot_if(bool pred, uint32_t a) {
        const uint32_t hd = (a & 0xFFFu) - hot_lo;
        bank = (pred && (a & 0x1000u) && hd < hot_n) ? hd : bank;
    }
    // TIA / RIOT I/O and the flat-array padding (src/atari.rs:509-519)
    __device__ __forceinline__ uint32_t io_read(uint32_t a) {
        if ((a & 0x280u) == 0x280u) return pia_read(a);
        if (!(a & 0x80u)) return tia_read_lanes(a);
        return 0u;
    }
    __device__ __forceinline__ void io_write(uint32_t a, uint32_t v) {
        if (a & 0x200u) pia_write(a, v);
        else tia_write_lanes(a, v);
    }
    // HeadlessBus::read (src/atari.rs:497-520): ROM window and the two RAM
    // mirrors come from LDS, everything else from the TIA / RIOT.
    __device__ __forceinline__ uint32_t bus_read(uint32_t addr) {
        const uint32_t a = addr & 0x1FFFu;
        const bool rom = (a & 0x1000u) != 0u;
        hotspot_if(true, a);
        const bool ram = (a & 0x1E80u) == 0x0080u;
        uint32_t v = ld8(rom ? ROM0 + (bank << 12) + (a & 0xFFFu) : ram_off(a));
        if (!rom && !ram) v = io_read(a);
        return v;
    }
    // Side effects of a read whose value is discarded: a hotspot switches the
    // bank, INTIM clears the underflow flag; nothing else has any.
    __device__ __forceinline__ void touch_if(bool pred, uint32_t addr) {
        const uint32_t a = addr & 0x1FFFu;
        hotspot_if(pred, a);
        tuf = (pred && (a & 0x1287u) == 0x284u) ? 0u : tuf;
    }
    // HeadlessBus::write (src/atari.rs:523-552)
    __device__ __forceinline__ void bus_write(uint32_t addr, uint32_t v) {
        const uint32_t a = addr & 0x1FFFu;
        const bool rom = (a & 0x1000u) != 0u;
        hotspot_if(true, a);
        const bool ram = (a & 0x1280u) == 0x0080u;
        st8(ram ? ram_off(a) : junk, v);
        if (!rom && !ram) io_write(a, v);
    }
    // Zero page and stack page: bit 7 selects RAM, otherwise the TIA (under a
    // ballot: the stack of JSR / RTS / RTI / BRK rarely reaches $0100-$017F).
    __device__ __forceinline__ uint32_t zp_read(uint32_t a8) {
        uint32_t v = ld8(ram_off(a8));
        if (__builtin_expect(__ballot(!(a8 & 0x80u)) != 0, 0)) {
            if (!(a8 & 0x80u)) v = tia_read_lanes(a8);
        }
        return v;
    }
    __device__ __forceinline__ void zp_write(uint32_t a8, uint32_t v) {
        st8((a8 & 0x80u) ? ram_off(a8) : junk, v);
        if (__builtin_expect(__ballot(!(a8 & 0x80u)) != 0, 0)) {
            if (!(a8 & 0x80u)) tia_write_lanes(a8, v);
        }
    }
    __device__ __forceinline__ void push(uint32_t v) { zp_write(sp, v); sp = (sp - 1u) & 0xFFu; }
    __device__ __forceinline__ uint32_t pull() { sp = (sp + 1u) & 0xFFu; return zp_read(sp); }

    // ------------------------------------------------------------ fetch
    // Opcode and operand bytes plus the descriptor.  Inside the ROM window
    // and below the hotspot / vector area the fetches have no side effects
    // and come straight from LDS; elsewhere they go through the bus in the
    // reference's order (opcode, then operands or the phantom read of PC).
    // JSR's high operand byte is read here for its side effects (an INTIM
    // read, a hotspot) and read again after the two pushes (execute_slow()).
    __device__ __forceinline__ void fetch(uint32_t &b1, uint32_t &b2, uint4 &d, uint4 &sl, uint32_t &op) {
        if ((pc & 0x1000u) && (pc & 0xFFFu) < 0xFF0u) {
            const uint32_t o = ROM0 + (bank << 12) + (pc & 0xFFFu);
            op = ld8(o);
            b1 = ld8(o + 1u);
            b2 = ld8(o + 2u);
            d = ld128(DESC0 + op * 32u);
            sl = ld128(DESC0 + op * 32u + 16u);
        } else {
            op = bus_read(pc);
            d = ld128(DESC0 + op * 32u);
            sl = ld128(DESC0 + op * 32u + 16u);
            const uint32_t d0 = d.x;
            const uint32_t len = (d0 >> D_LEN_SHIFT) & 3u;
            const bool illegal = (d0 & 3u) == CLS_ILLEGAL;
            b1 = len >= 2u ? bus_read((pc + 1u) & 0xFFFFu) : 0u;
            touch_if(len == 1u && !illegal, (pc + 1u) & 0xFFFFu);
            b2 = len == 3u ? bus_read((pc + 2u) & 0xFFFFu) : 0u;
        }
    }

    // ------------------------------------------------------------ execute
    __device__ __forceinline__ void set_zn(uint32_t v) {
        p = (p & ~(FZ | FN)) | (v & 0x80u) | (v == 0u ? FZ : 0u);
    }

    // Decimal-mode ADC / SBC (src/cpu.rs:155-166, 180-191) for the lanes in
    // `dec`, given the binary-mode operand `v`, the pre-instruction A and P and
    // the status already updated by the binary path (ADC keeps its Z; SBC keeps
    // every binary flag, only the accumulator differs).  Both forms are computed
    // and selected per lane.
    __device__ __forceinline__ void decimal_lanes(bool dec, bool sub, uint32_t a0, uint32_t v, uint32_t p0) {
        const uint32_t c = p0 & FC;
        uint32_t lo = (a0 & 0xFu) + (v & 0xFu) + c;
        lo += lo > 9u ? 6u : 0u;
        uint32_t hi = (a0 >> 4) + (v >> 4) + (lo > 0xFu ? 1u : 0u);
        const uint32_t inter = ((hi << 4) | (lo & 0xFu)) & 0xFFu;
        uint32_t np = p & ~(FN | FV | FC);
        np |= (hi & 8u) ? FN : 0u;
        np |= (~(a0 ^ v) & (a0 ^ inter) & 0x80u) ? FV : 0u;
        hi += hi > 9u ? 6u : 0u;
        np |= hi > 0xFu ? FC : 0u;
        const uint32_t a_add = ((hi & 0xFu) << 4) | (lo & 0xFu);
        int32_t slo = (int32_t)(a0 & 0xFu) - (int32_t)(v & 0xFu) - (int32_t)(c ^ 1u);
        const uint32_t lb = slo < 0 ? 1u : 0u;
        slo -= lb ? 6 : 0;
        int32_t shi = (int32_t)(a0 >> 4) - (int32_t)(v >> 4) - (int32_t)lb;
        shi -= shi < 0 ? 6 : 0;
        const uint32_t a_sub = (((uint32_t)shi & 0xFu) << 4) | ((uint32_t)slo & 0xFu);
        a = pick(dec, pick(sub, a_sub, a_add), a);
        p = pick(dec & !sub, np, p);
    }

    // TIA read registers from the lane's image: register r4 (addr & 15) is
    // read from the byte at S_CX + tia_lut(r4) and fixed by tia_fix.
    // CXxx (0-7) at S_CX + r, the paddle values (8, 9) at S_PAD0 = S_CX + 10,
    // INPT4 / INPT5 (12, 13) at S_INPT4 = S_CX + 8; 10, 11, 14, 15 read
    // constants (the byte is ignored).
    __device__ __forceinline__ static uint32_t tia_lut(uint32_t r4) {
        return pick(r4 < 8u, r4, __builtin_amdgcn_ubfe(0x009800BAu, (r4 & 7u) << 2, 4u));
    }
    // 8, 9: 0x80 once the paddle counter has reached the value; 10, 11: 0x80;
    // 12, 13: 0x80 while the input latch is set; 14, 15: 0 (src/headless_tia.rs:205-223).
    // pcount is a 16-bit count and m a byte, so pcount >= m is the sign bit of
    // pcount - m cleared, and m != 0 is min(m, 1).
    __device__ __forceinline__ uint32_t tia_fix(uint32_t r4, uint32_t m) const {
        const uint32_t bits = 2u | (((pcount - m) >> 31) ^ 1u) | (__builtin_elementwise_min(m, 1u) << 2);
        return pick(r4 < 8u, m, ((bits >> ((r4 >> 1) & 3u)) & 1u) << 7);
    }

    // One instruction of every lane whose PC lies in the cartridge window, in
    // one straight-line pass steered by the one-hot words d2 / d3 and the
    // register selectors d4 .. d7 (A, X, Y, SP as one word for the step): loads,
    // stores, ALU, read-modify-write, transfers, flag instructions,
    // conditional branches, JMP, PHA / PHP / PLA / PLP.  Only TIA / RIOT
    // writes, RIOT reads, zero-page pointers in the TIA and decimal mode
    // branch, each on a 64-bit ballot of the lanes that need it.  Bus
    // accesses are those of src/cpu.rs in the reference's order; the phantom
    // write of a read-modify-write is made only where it can be seen (TIA /
    // RIOT), since on RAM the final write replaces it and on ROM both writes
    // select the same bank.  A branch's target is computed as an effective
    // address (PC + offset) whose page-cross fix-up read is an indexed read's;
    // the taken branch's phantom read of PC has no side effect for a branch
    // that starts below $xFF0 of the window (below the hotspots, outside the
    // RIOT).  A stack access is a data access at $0100 + SP (RAM or TIA, as
    // zero page).  Lanes of any other kind (d2 = d3 = 0; branches outside
    // the window, whose H_BR is masked off by fz) change nothing here;
    // illegal opcodes return 0 cycles.
    __device__ __forceinline__ uint32_t compute_lanes(uint32_t b1, uint32_t b2, uint32_t d0, uint32_t d1, uint32_t g,
                                                      uint32_t h, const uint4 &sl, bool fz) {
        uint32_t cyc = (d0 >> D_CYC_SHIFT) & 7u;
        const uint32_t mbr = bm(h, H_BR) & mk(fz);
        const uint32_t mpush = bm(h, H_PUSH), mpull = bm(h, H_PULL), mpreg = bm(h, H_PREG);
        const uint32_t regs = a | (x << 8) | (y << 16) | (sp << 24);
        const uint32_t s = __builtin_amdgcn_perm(0u, regs, sl.x) | ((p | FB | FU) & mpreg & mpush);
        const uint32_t ix = __builtin_amdgcn_perm(0u, regs, sl.y);   // byte 0: the address's index; byte 1: (zp,X)'s X
        // effective address (every lane computes one; only lanes with a data access use it)
        const uint32_t zpm = bm(d0, 8u);                         // D_ZP: 8-bit address arithmetic
        const bool ind = (g & (1u << G_IND)) != 0u;
        const uint32_t zp = (b1 + (ix >> 8)) & 0xFFu;
        const uint32_t zq = (zp + 1u) & 0xFFu;
        uint32_t lo = ld8(ram_off(zp)), hi = ld8(ram_off(zq));
        if (__builtin_expect(__ballot(ind & (((zp & zq) & 0x80u) == 0u)) != 0, 0)) {
            // pointer bytes in the TIA read its registers
            if (ind & !(zp & 0x80u)) lo = tia_read_lanes(zp);
            if (ind & !(zq & 0x80u)) hi = tia_read_lanes(zq);
        }
        const uint32_t mind = mk(ind), mstk = mpush | mpull;
        uint32_t base = ((lo | (hi << 8)) & mind) | ((b1 | ((b2 << 8) & ~zpm)) & ~mind);
        base = ((0x100u | ((sp + (1u & mpull)) & 0xFFu)) & mstk) | (base & ~mstk);
        const uint32_t jbase = base;
        base = (pc & mbr) | (base & ~mbr);
        const uint32_t rel = (b1 - ((b1 & 0x80u) << 1)) & mbr;
        const uint32_t ea = (base + ((ix & 0xFFu) | rel)) & (0xFFFFu ^ (0xFF00u & zpm));
        const bool taken = (mbr != 0u) & (((p & (d1 >> D1_FMASK_SHIFT) & 0xFFu) != 0u) == ((d1 >> D1_FVAL_SHIFT) != 0u));
        const bool cross = ((base ^ ea) & 0xFF00u) != 0u;
        const bool cross_fix = (((d0 & D_DUMMY_CROSS) != 0u) | taken) & cross;
        touch_if(((d0 & D_DUMMY_ALWAYS) != 0u) | cross_fix, (base & 0xFF00u) | (ea & 0xFFu));
        cyc += (cross_fix ? 1u : 0u) + (taken ? 1u : 0u);
        pc = pick(taken, ea, pick((h & (1u << H_JMP)) != 0u, jbase, pc));
        // operand: immediate, register or the data read.  The read is one
        // load from the ROM window, a RAM mirror, the TIA register's image
        // byte or a zero byte (the flat array's padding); the RIOT branches.
        uint32_t v = (d0 & D_IMM) ? b1 : s;
        {
            const bool rd = ((d0 & D_RD) != 0u) | (mpull != 0u);
            const uint32_t a13 = ea & 0x1FFFu;
            const bool rom = (a13 & 0x1000u) != 0u;
            hotspot_if(rd, a13);
            const bool ram = (a13 & 0x1E80u) == 0x0080u;
            const bool tia = (a13 & 0x1080u) == 0u;
            const uint32_t r4 = a13 & 0x0Fu;
            uint32_t ad = pick(tia, img + S_CX + tia_lut(r4), ZERO_BYTE);
            ad = pick(ram, ram_off(a13), ad);
            ad = pick(rom, ROM0 + (bank << 12) + (a13 & 0xFFFu), ad);
            uint32_t m = ld8(ad);
            m = pick(tia, tia_fix(r4, m), m);
            v = pick(rd, m, v);
            const bool riot = rd & ((a13 & 0x1280u) == 0x0280u);
            if (__builtin_expect(__ballot(riot) != 0, 0)) {
                if (riot) v = pia_read(a13);
            }
        }
        // result candidates and the selected one
        const uint32_t p0 = p, a0 = a;
        const uint32_t lg = (a & v & bm(h, H_M_AND)) | ((a ^ v) & bm(h, H_M_XOR)) | (v & bm(h, H_M_LD));
        const uint32_t vv = v ^ (0xFFu & bm(d0, 25u));           // D_INV
        const uint32_t cin = p & FC;
        const uint32_t sum = s + vv + (cin | ((d0 >> 26) & 1u));  // D_CIN1
        const uint32_t ar = sum & 0xFFu;
        const uint32_t shl = ((v << 1) | (cin & bm(g, G_CIN_SHL))) & 0xFFu;
        const uint32_t shr = (v >> 1) | ((cin << 7) & bm(g, G_CIN_SHR));
        const uint32_t inc = (v + (g >> G_DELTA)) & 0xFFu;       // bits above the delta fall off with the mask
        const uint32_t r = (lg & bm(g, G_R_LOG)) | (ar & bm(g, G_R_ARI)) | (shl & bm(g, G_R_SHL)) |
                           (shr & bm(g, G_R_SHR)) | (inc & bm(g, G_R_INC)) | (s & bm(g, G_R_SRC));
        // status: N, V, Z, C candidates, then the descriptor's clear / take / set
        // masks; PLP takes every flag from the pulled byte
        const uint32_t mbit = bm(g, G_BIT);
        const uint32_t nv = (v & mbit) | (r & ~mbit);
        const uint32_t vf = (((~(s ^ vv) & (s ^ ar)) >> 1) & FV & ~mbit) | (v & FV & mbit);
        const uint32_t cf = (((sum >> 8) & bm(g, G_C_SUM)) | ((v >> 7) & bm(g, G_C_B7)) | (v & bm(g, G_C_B0))) & FC;
        const uint32_t cand = (nv & FN) | vf | (r == 0u ? FZ : 0u) | cf;
        p = (p & ~(h & 0xFFu)) | (cand & (h >> H_TAKE)) | ((h >> H_SET) & 0xFFu) | (v & ~(FB | FU) & mpreg & mpull);
        const uint32_t rg = __builtin_amdgcn_perm(r, regs + sl.w, sl.z);
        a = rg & 0xFFu;
        x = (rg >> 8) & 0xFFu;
        y = (rg >> 16) & 0xFFu;
        sp = rg >> 24;
        const bool dec = ((d0 & D_DEC) != 0u) & ((p0 & FD) != 0u);
        if (__builtin_expect(__ballot(dec) != 0, 0)) decimal_lanes(dec, (d0 & D_INV) != 0u, a0, v, p0);
        // write-back
        {
            const bool wr = ((d0 & D_WR) != 0u) | (mpush != 0u);
            const uint32_t a13 = ea & 0x1FFFu;
            hotspot_if(wr, a13);
            const bool rom = (a13 & 0x1000u) != 0u;
            const bool ram = (a13 & 0x1280u) == 0x0080u;
            st8(wr & ram ? ram_off(a13) : junk, r);
            const bool io = wr & !rom & !ram;
            if (__builtin_expect(__ballot(io) != 0, 0)) {
                // The read-modify-write's first write (the old value) leaves a
                // trace only on VSYNC (edge), VBLANK (paddle dump), GRP0 / GRP1
                // (the old-copy shift) and HMOVE (motion); on every other TIA
                // and RIOT register the second write sets the final state alone.
                const uint32_t tr = a13 & 0x3Fu;
                const bool lasting = !(a13 & 0x200u) & ((tr <= 0x01u) | (tr == 0x1Bu) | (tr == 0x1Cu) | (tr == 0x2Au));
                const bool phantom = io & ((d0 & D_RMWMEM) != 0u) & lasting;
                if (__builtin_expect(__ballot(phantom) != 0, 0)) {
                    if (phantom) io_write(a13, v);
                }
                if (io) io_write(a13, r);
            }
        }
        return cyc;
    }

    // Executes one decoded instruction (pc already points past it).  Returns
    // the cycle count, or 0 for an illegal opcode (src/cpu.rs:1139-1141).
    __device__ __forceinline__ uint32_t execute(uint32_t b1, uint32_t b2, const uint4 &d, const uint4 &sl, bool fz) {
        // JSR, RTS, RTI, BRK, JMP (ind), and branches outside the window
        const bool slow = ((d.y & D1_SLOW) != 0u) | (((d.w >> H_BR) & 1u) > (uint32_t)fz);
        uint32_t cyc = compute_lanes(b1, b2, d.x, d.y, d.z, d.w, sl, fz);
        if (__builtin_expect(__ballot(slow) != 0, 0)) {
            if (slow) cyc = execute_slow(b1, b2, d);
        }
        return cyc;
    }

    // The instructions compute_lanes leaves: every bus access in the
    // reference's order (src/cpu.rs:949-1137).
    __device__ __forceinline__ uint32_t execute_slow(uint32_t b1, uint32_t b2, const uint4 &d) {
        const uint32_t d0 = d.x, d1 = d.y;
        uint32_t cyc = (d0 >> D_CYC_SHIFT) & 7u;
        const uint32_t cls = d0 & 3u;
        if (cls == CLS_BRANCH) {
            const uint32_t fm = (d1 >> D1_FMASK_SHIFT) & 0xFFu;
            const bool taken = ((p & fm) != 0u) == ((d1 >> D1_FVAL_SHIFT) != 0u);
            const uint32_t np = (pc + b1 - ((b1 & 0x80u) << 1)) & 0xFFFFu;
            const bool cross = ((np ^ pc) & 0xFF00u) != 0u;
            touch_if(taken, pc);                                    // phantom read of PC
            touch_if(taken && cross, (pc & 0xFF00u) | (np & 0xFFu)); // before the high-byte fix
            cyc += taken ? (cross ? 2u : 1u) : 0u;
            pc = taken ? np : pc;
            return cyc;
        }
        // JSR, RTS, RTI, BRK, JMP (ind) (src/cpu.rs:949-1137); JMP absolute
        // and the pushes / pulls run in compute_lanes; phantom stack reads
        // have no side effect
        switch ((d1 >> D1_SPEC_SHIFT) & 15u) {
        case SP_JMPI: {
            const uint32_t ptr = b1 | (b2 << 8);
            const uint32_t lo = bus_read(ptr);
            const uint32_t hi = bus_read((ptr & 0xFF00u) | ((ptr + 1u) & 0xFFu));
            pc = lo | (hi << 8);
            break;
        }
        case SP_JSR: {
            // The reference reads the high operand byte after the two pushes
            // (src/cpu.rs:996-1003).  The pushes (stack page: RAM or TIA
            // registers) can change only a RAM byte or a TIA register there,
            // and reading either has no side effect, so fetch()'s read stands
            // for the side effects of a RIOT / ROM byte and the value is
            // taken again from RAM or the TIA here.
            const uint32_t ret = (pc - 1u) & 0xFFFFu;       // address of the high operand byte
            push(ret >> 8);
            push(ret & 0xFFu);
            const uint32_t a13 = ret & 0x1FFFu;
            uint32_t hi = b2;
            if (__builtin_expect(__ballot((a13 & 0x1000u) == 0u) != 0, 0)) {   // some lane's JSR lies outside the ROM
                const bool ram = (a13 & 0x1E80u) == 0x0080u, tia = (a13 & 0x1080u) == 0u;
                const uint32_t m = ld8(ram ? ram_off(a13) : tia_read_off(a13));
                hi = ram ? m : (tia ? tia_read_val(a13, m) : b2);
            }
            pc = b1 | (hi << 8);
            break;
        }
        case SP_RTS: {
            const uint32_t lo = pull();
            const uint32_t hi = pull();
            const uint32_t ad = lo | (hi << 8);
            touch_if(true, ad);
            pc = (ad + 1u) & 0xFFFFu;
            break;
        }
        case SP_RTI: {
            p = pull() & ~(FB | FU);
            const uint32_t lo = pull();
            const uint32_t hi = pull();
            pc = lo | (hi << 8);
            break;
        }
        case SP_BRK: {
            pc = (pc + 1u) & 0xFFFFu;                       // skip the padding byte
            push(pc >> 8);
            push(pc & 0xFFu);
            push(p | FB | FU);
            p |= FI;
            const uint32_t lo = bus_read(0xFFFEu);
            const uint32_t hi = bus_read(0xFFFFu);
            pc = lo | (hi << 8);
            break;
        }
        default: break;
        }
        return cyc;
    }

    // ------------------------------------------------------------ fast paths
    // Side effects of a discarded read at a13 (13-bit address): a hotspot or INTIM.
    __device__ __forceinline__ bool touch_has_effect(uint32_t a13) const {
        return ((a13 & 0x1000u) && ((a13 & 0xFFFu) - hot_lo) < hot_n) || (a13 & 0x1287u) == 0x284u;
    }

    // Convergent fast path: every active lane executes the same opcode, whose
    // descriptor (u0, u1) is wave-uniform (SGPRs), so each descriptor test
    // below is a scalar branch and only the work of this opcode's stages is
    // issued.  TIA / RIOT accesses are made in place (exact, in order).
    // Returns false, having changed nothing, when the lane needs the general
    // path: a hotspot or INTIM reached by a phantom read, a bankswitch data
    // access, a zero-page pointer or stack byte in the TIA, JMP (ind) or an
    // illegal opcode.
    __device__ __forceinline__ bool fast_uniform(uint32_t u0, uint32_t u1, uint32_t b1, uint32_t b2, uint32_t pc0,
                                                 uint32_t &cyc) {
        const uint32_t cls = u0 & 3u;
        const uint32_t pcn = (pc0 + ((u0 >> D_LEN_SHIFT) & 3u)) & 0xFFFFu;
        cyc = (u0 >> D_CYC_SHIFT) & 7u;
        if (cls == CLS_BRANCH) {
            const uint32_t fm = (u1 >> D1_FMASK_SHIFT) & 0xFFu;
            const bool taken = ((p & fm) != 0u) == ((u1 >> D1_FVAL_SHIFT) != 0u);
            const uint32_t np = (pcn + b1 - ((b1 & 0x80u) << 1)) & 0xFFFFu;
            const bool bcross = ((np ^ pcn) & 0xFF00u) != 0u;
            if (taken && bcross && touch_has_effect(((pcn & 0xFF00u) | (np & 0xFFu)) & 0x1FFFu)) return false;
            pc = pick(taken, np, pcn);
            cyc += (uint32_t)taken + (uint32_t)(taken && bcross);
            return true;
        }
        if (cls == CLS_SPECIAL) {
            const uint32_t spec = (u1 >> D1_SPEC_SHIFT) & 15u;
            const uint32_t s1 = (sp + 1u) & 0xFFu, s2 = (sp + 2u) & 0xFFu, s3 = (sp + 3u) & 0xFFu;
            const uint32_t m1 = (sp - 1u) & 0xFFu, m2 = (sp - 2u) & 0xFFu;
            switch (spec) {
            case SP_JMP: pc = b1 | (b2 << 8); return true;
            case SP_JSR: {
                if (!(sp & m1 & 0x80u)) return false;
                const uint32_t ret = (pcn - 1u) & 0xFFFFu;
                st8(ram_off(sp), ret >> 8);
                st8(ram_off(m1), ret & 0xFFu);
                sp = m2;
                pc = b1 | (b2 << 8);
                return true;
            }
            case SP_RTS: {
                if (!(s1 & s2 & 0x80u)) return false;
                const uint32_t ad = ld8(ram_off(s1)) | (ld8(ram_off(s2)) << 8);
                if (touch_has_effect(ad & 0x1FFFu)) return false;
                sp = s2;
                pc = (ad + 1u) & 0xFFFFu;
                return true;
            }
            case SP_RTI: {
                if (!(s1 & s2 & s3 & 0x80u)) return false;
                p = ld8(ram_off(s1)) & ~(FB | FU);
                pc = ld8(ram_off(s2)) | (ld8(ram_off(s3)) << 8);
                sp = s3;
                return true;
            }
            case SP_BRK: {
                if (!(sp & m1 & m2 & 0x80u)) return false;
                const uint32_t ret = (pcn + 1u) & 0xFFFFu;
                st8(ram_off(sp), ret >> 8);
                st8(ram_off(m1), ret & 0xFFu);
                st8(ram_off(m2), p | FB | FU);
                sp = (sp - 3u) & 0xFFu;
                p |= FI;
                const uint32_t vb = ROM0 + (bank << 12);     // $FFFE/$FFFF are never hotspots
                pc = ld8(vb + 0xFFEu) | (ld8(vb + 0xFFFu) << 8);
                return true;
            }
            case SP_PHA: case SP_PHP:
                if (!(sp & 0x80u)) return false;
                st8(ram_off(sp), spec == SP_PHA ? a : (p | FB | FU));
                sp = m1;
                pc = pcn;
                return true;
            case SP_PLA: case SP_PLP: {
                if (!(s1 & 0x80u)) return false;
                const uint32_t v = ld8(ram_off(s1));
                sp = s1;
                if (spec == SP_PLA) { a = v; set_zn(v); } else { p = v & ~(FB | FU); }
                pc = pcn;
                return true;
            }
            default: return false;                              // JMP (ind)
            }
        }
        if (cls == CLS_ILLEGAL) return false;

        // ---- compute class
        const uint32_t srcsel = (u1 >> D1_SRC_SHIFT) & 7u;
        const uint32_t s = srcsel == R_A ? a : srcsel == R_X ? x : srcsel == R_Y ? y : sp;
        uint32_t v = (u0 & D_IMM) ? b1 : s;
        uint32_t a13 = 0;
        bool io = false, ramw = false;
        if (u0 & D_MEM) {
            const uint32_t ind = (u0 >> D_IND_SHIFT) & 3u;
            const uint32_t ixsel = (u0 >> D_IDX_SHIFT) & 3u;
            uint32_t base, ea;
            if (ind) {
                const uint32_t zp = ind == 1u ? ((b1 + x) & 0xFFu) : b1;
                const uint32_t zp1 = (zp + 1u) & 0xFFu;
                if (!(zp & zp1 & 0x80u)) return false;          // pointer byte in the TIA
                base = ld8(ram_off(zp)) | (ld8(ram_off(zp1)) << 8);
                ea = ind == 2u ? (base + y) & 0xFFFFu : base;
            } else if (u0 & D_ZP) {
                base = b1;
                ea = ixsel == 0u ? b1 : ((b1 + (ixsel == 1u ? x : y)) & 0xFFu);
            } else {
                base = b1 | (b2 << 8);
                ea = ixsel == 0u ? base : ((base + (ixsel == 1u ? x : y)) & 0xFFFFu);
            }
            a13 = ea & 0x1FFFu;
            if (u0 & (D_DUMMY_ALWAYS | D_DUMMY_CROSS)) {
                const bool cross = ((base ^ ea) & 0xFF00u) != 0u;
                const bool fix = (u0 & D_DUMMY_ALWAYS) || cross;
                if (fix && touch_has_effect(((base & 0xFF00u) | (ea & 0xFFu)) & 0x1FFFu)) return false;
                if (u0 & D_DUMMY_CROSS) cyc += (uint32_t)cross;
            }
            const bool rom = (a13 & 0x1000u) != 0u;
            if (rom && ((a13 & 0xFFFu) - hot_lo) < hot_n) return false;   // bankswitch access
            if (u0 & D_RD) {
                const bool ram = (a13 & 0x1E80u) == 0x0080u;
                v = ld8(rom ? ROM0 + (bank << 12) + (a13 & 0xFFFu) : ram_off(a13));
                io = !rom && !ram;
                if (io) v = io_read(a13);
            }
            if (u0 & D_WR) { ramw = (a13 & 0x1280u) == 0x0080u; io = !rom && !ramw; }
        }
        // ---- compute (scalar branches on the kind)
        const uint32_t kind = (u0 >> D_KIND_SHIFT) & 7u;
        const uint32_t p0 = p, a0 = a;
        uint32_t r = s;
        uint32_t q = p;
        if (kind == K_LOGIC) {
            r = u0 & D_M_LD ? v : (u0 & D_M_XOR ? (u0 & D_M_AND ? (a | v) : (a ^ v)) : (a & v));
        } else if (kind == K_ARITH) {
            const uint32_t vv = (u0 & D_INV) ? (v ^ 0xFFu) : v;
            const uint32_t sum = s + vv + ((u0 & D_CIN1) ? 1u : (p & FC));
            r = sum & 0xFFu;
            q = (q & ~FC) | ((sum >> 8) & 1u);
            if (u0 & D_SETV) q = (q & ~FV) | ((~(s ^ vv) & (s ^ r) & 0x80u) >> 1);
        } else if (kind == K_RMW) {
            const uint32_t rop = (u1 >> D1_RMW_SHIFT) & 7u;
            if (rop < 2u) { r = ((v << 1) | (rop == 1u ? (p & FC) : 0u)) & 0xFFu; q = (q & ~FC) | (v >> 7); }
            else if (rop < 4u) { r = (v >> 1) | (rop == 3u ? ((p & FC) << 7) : 0u); q = (q & ~FC) | (v & 1u); }
            else r = (v + (rop == 5u ? 1u : 0xFFu)) & 0xFFu;
        } else if (kind == K_MOVE) {
            const uint32_t dl = (u1 >> D1_DELTA_SHIFT) & 3u;
            r = (s + (dl == 1u ? 1u : (dl == 2u ? 0xFFu : 0u))) & 0xFFu;
        } else if (kind == K_BIT) {
            q = (q & ~(FN | FV | FZ)) | (v & 0xC0u) | ((a & v) == 0u ? FZ : 0u);
        } else if (kind == K_FLAGS) {
            const uint32_t fm = (u1 >> D1_FMASK_SHIFT) & 0xFFu;
            q = (q & ~fm) | (u1 >> D1_FVAL_SHIFT);
        }
        if (u0 & D_SETNZ) q = (q & ~(FN | FZ)) | (r & 0x80u) | (r == 0u ? FZ : 0u);
        p = q;
        const uint32_t dst = (u1 >> D1_DST_SHIFT) & 7u;
        if (dst == R_A) a = r;
        else if (dst == R_X) x = r;
        else if (dst == R_Y) y = r;
        else if (dst == R_SP) sp = r;
        if ((u0 & D_DEC) && (p0 & FD)) decimal_lanes(true, (u0 & D_INV) != 0u, a0, v, p0);
        if (u0 & D_WR) {
            st8(ramw ? ram_off(a13) : junk, r);                 // ROM / I/O targets hit the discard byte
            if (io) {
                if (u0 & D_RMWMEM) io_write(a13, v);
                io_write(a13, r);
            }
        }
        pc = pcn;
        return true;
    }

    // The three ROM bytes at PC in the current bank (opcode and operands of
    // the next instruction when PC is inside the ROM window; read ahead of
    // exec() so their LDS latency overlaps other work).
    __device__ __forceinline__ void fetch_ahead(uint32_t &op, uint32_t &b1, uint32_t &b2) const {
        const uint32_t o = ROM0 + (bank << 12) + (pc & 0xFFFu);
        op = ld8(o);
        b1 = ld8(o + 1u);
        b2 = ld8(o + 2u);
    }

    // Advance PC and execute one instruction, given fetch_ahead()'s bytes for
    // the current PC and bank.  Returns the cycle count, or 0 for an illegal
    // opcode.  Lanes whose PC is inside the ROM window (below the hotspot /
    // vector area) use those bytes, as fetching there has no side effects;
    // the others fetch through the bus.
    // Without GROUP every wave takes the general path (execute()): waves whose
    // lanes agree on their control state run in the wave-uniform mode
    // instead, so a per-opcode convergent path here would rarely run, and its
    // merge with the general path costs every divergent instruction register
    // copies and a long back edge (measured: branchy ROM 17.2 -> 16.1 ms per
    // frame, F8 0.90 -> 0.86 ms without it).
    // With GROUP (set_sorted), a wave whose lanes hold at most two distinct
    // opcodes runs each opcode group on the convergent path in turn: a scalar
    // dispatch on the group's opcode into the fast path specialised with that
    // opcode's compile-time descriptor (64-bit ballot + readlane find the
    // groups); with three or more, the lanes share one pass of the general
    // path, which costs less than one convergent pass per opcode.  Anything
    // the fast path declines goes through the general path, which replays
    // the reference's bus accesses one by one.  `grouped` (GROUP only, may be
    // null) reports whether the groups covered the wave.
    template <bool GROUP>
    __device__ __forceinline__ uint32_t exec(uint32_t op, uint32_t b1, uint32_t b2, bool *grouped = nullptr) {
        const uint32_t pc0 = pc;
        const bool fz = (pc0 & 0x1000u) && ((pc0 & 0xFFFu) < 0xFF0u);
        uint32_t cyc = 0;
        bool done = false;
        if constexpr (GROUP) {
            const uint32_t key = fz ? op : 0x100u;
            const uint64_t all = __ballot(1);
            // The lane tests compare against an opaque copy of the group's key:
            // were it the key itself, the optimiser would substitute the
            // per-lane key for it where the two are equal, and the dispatch
            // below would become a VALU compare tree under exec masks instead
            // of scalar branches.
            const uint32_t k1 = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)__builtin_ctzll(all));
            uint32_t k1c = k1;
            __asm__ volatile("" : "+s"(k1c));
            const uint64_t g1 = __ballot(key == k1c);
            const uint64_t r2 = all & ~g1;
            uint32_t k2 = k1;
            uint64_t g2 = 0;
            if (r2 != 0) {
                k2 = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)__builtin_ctzll(r2));
                uint32_t k2c = k2;
                __asm__ volatile("" : "+s"(k2c));
                g2 = __ballot(key == k2c);
            }
            if (grouped) *grouped = (g1 | g2) == all;
            if ((g1 | g2) == all) {
#pragma nounroll
                for (uint32_t q = 0; q < (r2 != 0 ? 2u : 1u); ++q) {
                    // the scalar dispatch on the group's opcode encloses the
                    // per-lane group test (a uniform region around a divergent one)
                    uint32_t lk = q ? k2 : k1;
                    __asm__ volatile("" : "+s"(lk));
                    const bool mine = key == lk;
                    if (lk < 0x100u) {
                        switch (lk) {
#define FA_GCASE(OPC)                                                                             \
                        case OPC: {                                                               \
                            constexpr uint64_t D = descriptor_of(OPC);                            \
                            if (mine) done = fast_uniform((uint32_t)D, (uint32_t)(D >> 32), b1, b2, pc0, cyc); \
                            break;                                                                \
                        }
                            FA_FOR_EACH_OPCODE(FA_GCASE)
#undef FA_GCASE
                        default: break;
                        }
                    }
                }
            }
        }
        if (!done) {
            uint32_t gb1 = b1, gb2 = b2, gop = op;
            uint4 gd = ld128(DESC0 + op * 32u);
            uint4 gs = ld128(DESC0 + op * 32u + 16u);
            if (__builtin_expect(!fz, 0)) fetch(gb1, gb2, gd, gs, gop);
            pc = (pc0 + ((gd.x >> D_LEN_SHIFT) & 3u)) & 0xFFFFu;
            cyc = execute(gb1, gb2, gd, gs, fz);
        }
        return cyc;
    }
};

}  // namespace fa
