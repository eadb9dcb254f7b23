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
// LDS layout of one workgroup (byte offsets into fa_lds):
//   [0, rom_bytes)                ROM image: the 4 KB window of a Fixed ROM,
//                                 or every 4 KB bank of an F8/F6/F4 ROM
//   [rom_bytes, + bs * 222)       the workgroup's AtariStateGpu images (AoS,
//                                 222-byte stride), staged with 16-byte loads
// Hot fields (registers, TIA timing, RIOT timer, bank) live in VGPRs for the
// whole launch; RAM and the rarely touched TIA/RIOT registers stay in the
// lane's LDS image.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fa {

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
__device__ __forceinline__ uint32_t ld16(uint32_t off) {
    return (uint32_t)fa_lds[off] | ((uint32_t)fa_lds[off + 1] << 8);
}
__device__ __forceinline__ void st16(uint32_t off, uint32_t v) {
    fa_lds[off] = (uint8_t)v;
    fa_lds[off + 1] = (uint8_t)(v >> 8);
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
    // addressing: LDS offset of this lane's image; hotspot window
    uint32_t img;
    uint32_t hot_lo, hot_n;

    // ------------------------------------------------------------ state I/O
    __device__ __forceinline__ void load(uint32_t img_off, uint32_t hot_lo_, uint32_t hot_n_) {
        img = img_off;
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
    // Pia::tick_n (src/pia.rs:107-166) with the prescaler division done by
    // shifts (select is 1, 8, 64 or 1024).
    __device__ __forceinline__ void pia_tick(uint32_t n) {
        if (n == 0) return;
        if (tval == 0 && tuf) {
            tval = (0x100u - n) & 0xFFu;
            return;
        }
        uint32_t rem = n;
        if (tps > 0) {
            if (rem < tps) { tps -= rem; return; }
            rem -= tps;
            tps = 0;
            if (tval == 0) {
                tuf = 1;
                tval = rem ? ((0x100u - rem) & 0xFFu) : 0xFFu;
                return;
            }
            tval -= 1;
        }
        if (rem == 0) { tps = tsel; return; }
        uint32_t full = rem >> tsh;
        uint32_t left = rem & (tsel - 1u);
        if (full >= tval) {
            rem -= tval << tsh;
            tuf = 1;
            tval = rem ? ((0x100u - rem) & 0xFFu) : 0xFFu;
        } else {
            tval -= full;
            tps = tsel - left;
        }
    }

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
            tval = v;
            tuf = 0;
            tsh = (0x0A060300u >> ((addr & 3u) * 8u)) & 0xFFu;   // 1, 8, 64, 1024
            tsel = 1u << tsh;
            tps = tsel;
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
        clock += 3u * n;
        if (clock >= 228u) end_scanline();
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

    // (clock - 68 + 5) rem_euclid 160 for clock < 228 (src/headless_tia.rs:257-261)
    __device__ __forceinline__ uint32_t resp_pos() const {
        uint32_t t = clock + 97u;
        t = t >= 160u ? t - 160u : t;
        return t >= 160u ? t - 160u : t;
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
            st16(img + S_POS + 2u * (r - 0x10u), resp_pos());
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

    // ------------------------------------------------------------ bus
    __device__ __forceinline__ void hotspot(uint32_t lo12) {
        uint32_t d = lo12 - hot_lo;
        if (d < hot_n) bank = d;
    }
    __device__ __forceinline__ uint32_t ram_off(uint32_t a) const { return img + S_RAM + (a & 0x7Fu); }

    // HeadlessBus::read (src/atari.rs:497-520)
    __device__ __forceinline__ uint32_t bus_read(uint32_t addr) {
        uint32_t a = addr & 0x1FFFu;
        if (a & 0x1000u) {
            hotspot(a & 0xFFFu);
            return ld8((bank << 12) + (a & 0xFFFu));
        }
        if (a & 0x80u) {
            if (!(a & 0x200u)) return a < 0x200u ? ld8(ram_off(a)) : 0u;
            return pia_read(a);
        }
        return tia_read(a);
    }
    // Side effects of a read whose value is discarded: a hotspot switches the
    // bank, INTIM clears the underflow flag.  RAM, TIA and the other RIOT
    // registers have none.
    __device__ __forceinline__ void touch(uint32_t addr) {
        uint32_t a = addr & 0x1FFFu;
        if (a & 0x1000u) hotspot(a & 0xFFFu);
        else if ((a & 0x287u) == 0x284u) tuf = 0;
    }
    // HeadlessBus::write (src/atari.rs:523-552)
    __device__ __forceinline__ void bus_write(uint32_t addr, uint32_t v) {
        uint32_t a = addr & 0x1FFFu;
        if (a & 0x1000u) { hotspot(a & 0xFFFu); return; }
        if ((a & 0x280u) == 0x80u) { st8(ram_off(a), v); return; }
        if (a & 0x200u) { pia_write(a, v); return; }
        tia_write(a, v);
    }
    // Zero page ($00-$FF): bit 7 selects RAM, otherwise the TIA.
    __device__ __forceinline__ uint32_t zp_read(uint32_t a8) {
        return (a8 & 0x80u) ? ld8(ram_off(a8)) : tia_read(a8);
    }
    __device__ __forceinline__ void zp_write(uint32_t a8, uint32_t v) {
        if (a8 & 0x80u) st8(ram_off(a8), v);
        else tia_write(a8, v);
    }
    // Stack page ($100 | sp) decodes like the zero page.
    __device__ __forceinline__ void push(uint32_t v) { zp_write(sp, v); sp = (sp - 1u) & 0xFFu; }
    __device__ __forceinline__ uint32_t pull() { sp = (sp + 1u) & 0xFFu; return zp_read(sp); }

    // ------------------------------------------------------------ ALU
    __device__ __forceinline__ void set_zn(uint32_t v) {
        p = (p & ~(FZ | FN)) | (v & 0x80u) | (v == 0u ? FZ : 0u);
    }
    __device__ __forceinline__ void setf(uint32_t f, bool on) { p = on ? (p | f) : (p & ~f); }

    // ADC, src/cpu.rs:152-175
    __device__ __forceinline__ void adc(uint32_t v) {
        uint32_t c = p & FC;
        if (p & FD) {
            uint32_t lo = (a & 0xFu) + (v & 0xFu) + c;
            if (lo > 9u) lo += 6u;
            uint32_t hi = (a >> 4) + (v >> 4) + (lo > 0xFu ? 1u : 0u);
            uint32_t bin = (a + v + c) & 0xFFu;
            uint32_t inter = ((hi << 4) | (lo & 0xFu)) & 0xFFu;
            uint32_t np = p & ~(FZ | FN | FV | FC);
            np |= bin == 0u ? FZ : 0u;
            np |= (hi & 8u) ? FN : 0u;
            np |= (~(a ^ v) & (a ^ inter) & 0x80u) ? FV : 0u;
            if (hi > 9u) hi += 6u;
            np |= hi > 0xFu ? FC : 0u;
            p = np;
            a = ((hi & 0xFu) << 4) | (lo & 0xFu);
        } else {
            uint32_t s = a + v + c;
            uint32_t r = s & 0xFFu;
            uint32_t np = p & ~(FZ | FN | FV | FC);
            np |= s > 0xFFu ? FC : 0u;
            np |= (~(a ^ v) & (a ^ r) & 0x80u) ? FV : 0u;
            a = r;
            p = np;
            set_zn(a);
        }
    }
    // SBC, src/cpu.rs:177-200
    __device__ __forceinline__ void sbc(uint32_t v) {
        uint32_t borrow = (p & FC) ? 0u : 1u;
        int32_t d = (int32_t)a - (int32_t)v - (int32_t)borrow;
        uint32_t r = (uint32_t)d & 0xFFu;
        uint32_t np = p & ~(FV | FC);
        np |= d >= 0 ? FC : 0u;
        np |= ((a ^ v) & (a ^ r) & 0x80u) ? FV : 0u;
        if (p & FD) {
            int32_t lo = (int32_t)(a & 0xFu) - (int32_t)(v & 0xFu) - (int32_t)borrow;
            uint32_t lo_b = 0;
            if (lo < 0) { lo -= 6; lo_b = 1; }
            int32_t hi = (int32_t)(a >> 4) - (int32_t)(v >> 4) - (int32_t)lo_b;
            if (hi < 0) hi -= 6;
            p = np;
            set_zn(r);
            a = (((uint32_t)hi & 0xFu) << 4) | ((uint32_t)lo & 0xFu);
        } else {
            p = np;
            a = r;
            set_zn(a);
        }
    }
    __device__ __forceinline__ void cmp(uint32_t reg, uint32_t v) {
        setf(FC, reg >= v);
        set_zn((reg - v) & 0xFFu);
    }

    // Read operations of every read mode (aaa / cc opcode fields).
    __device__ __forceinline__ void read_op(uint32_t op, uint32_t v) {
        uint32_t cc = op & 3u, aaa = op >> 5;
        if (cc == 1u) {
            switch (aaa) {
            case 0: a |= v; set_zn(a); break;
            case 1: a &= v; set_zn(a); break;
            case 2: a ^= v; set_zn(a); break;
            case 3: adc(v); break;
            case 5: a = v; set_zn(a); break;
            case 6: cmp(a, v); break;
            case 7: sbc(v); break;
            default: break;
            }
        } else if (cc == 2u) {
            x = v; set_zn(x);
        } else {
            switch (aaa) {
            case 1: p = (p & ~(FZ | FV | FN)) | ((a & v) == 0u ? FZ : 0u) | (v & 0xC0u); break;   // BIT
            case 5: y = v; set_zn(y); break;
            case 6: cmp(y, v); break;
            default: cmp(x, v); break;                                                            // CPX
            }
        }
    }
    __device__ __forceinline__ uint32_t store_reg(uint32_t op) const {
        uint32_t cc = op & 3u;
        return cc == 1u ? a : (cc == 2u ? x : y);
    }
    __device__ __forceinline__ uint32_t rmw_op(uint32_t op, uint32_t v) {
        uint32_t r;
        switch (op >> 5) {
        case 0: setf(FC, v & 0x80u); r = (v << 1) & 0xFFu; break;                 // ASL
        case 1: r = ((v << 1) | (p & FC)) & 0xFFu; setf(FC, v & 0x80u); break;    // ROL
        case 2: setf(FC, v & 1u); r = v >> 1; break;                               // LSR
        case 3: r = (v >> 1) | ((p & FC) << 7); setf(FC, v & 1u); break;         // ROR
        case 6: r = (v - 1u) & 0xFFu; break;                                       // DEC
        default: r = (v + 1u) & 0xFFu; break;                                      // INC
        }
        set_zn(r);
        return r;
    }
    __device__ __forceinline__ bool branch_taken(uint32_t op) const {
        // bits 7-6 select N, V, C, Z; bit 5 is the value tested for
        constexpr uint32_t kFlag = (FN << 0) | (FV << 8) | (FC << 16) | (FZ << 24);
        uint32_t f = (kFlag >> ((op >> 6) * 8u)) & 0xFFu;
        return ((p & f) != 0u) == ((op & 0x20u) != 0u);
    }

    // ------------------------------------------------------------ CPU step
    // Executes the instruction at pc.  With FAST, the opcode and both operand
    // bytes were prefetched from the current bank (pc is inside the ROM
    // window, away from the hotspots, so those reads and the phantom read of
    // PC have no side effect).  Returns the cycle count, or 0 for an illegal
    // opcode (PC already advanced past it, src/cpu.rs:1139-1141).
    template <bool FAST>
    __device__ __forceinline__ uint32_t opnd(uint32_t pre) {
        uint32_t v = FAST ? pre : bus_read(pc);
        pc = (pc + 1u) & 0xFFFFu;
        return v;
    }
    template <bool FAST>
    __device__ __forceinline__ void phantom_pc() {
        if (!FAST) touch(pc);
    }

    template <bool FAST>
    __device__ __forceinline__ uint32_t step(uint32_t op, uint32_t b1, uint32_t b2) {
        pc = (pc + 1u) & 0xFFFFu;
        uint32_t v, ad, eff;
        switch (op) {
        // implied / accumulator (src/cpu.rs:257-330)
        case 0xEA: phantom_pc<FAST>(); return 2;
        case 0x18: case 0x38: case 0x58: case 0x78: case 0xB8: case 0xD8: case 0xF8: {
            phantom_pc<FAST>();
            // bits 7-6 pick C, I, V, D; bit 5 the value (CLV has no set form)
            constexpr uint32_t kF = FC | (FI << 8) | (FV << 16) | (FD << 24);
            uint32_t f = (kF >> ((op >> 6) * 8u)) & 0xFFu;
            setf(f, (op & 0x20u) && op != 0xB8u);
            return 2;
        }
        case 0xAA: phantom_pc<FAST>(); x = a; set_zn(x); return 2;
        case 0xA8: phantom_pc<FAST>(); y = a; set_zn(y); return 2;
        case 0x8A: phantom_pc<FAST>(); a = x; set_zn(a); return 2;
        case 0x98: phantom_pc<FAST>(); a = y; set_zn(a); return 2;
        case 0xBA: phantom_pc<FAST>(); x = sp; set_zn(x); return 2;
        case 0x9A: phantom_pc<FAST>(); sp = x; return 2;
        case 0xE8: phantom_pc<FAST>(); x = (x + 1u) & 0xFFu; set_zn(x); return 2;
        case 0xC8: phantom_pc<FAST>(); y = (y + 1u) & 0xFFu; set_zn(y); return 2;
        case 0xCA: phantom_pc<FAST>(); x = (x - 1u) & 0xFFu; set_zn(x); return 2;
        case 0x88: phantom_pc<FAST>(); y = (y - 1u) & 0xFFu; set_zn(y); return 2;
        case 0x0A: case 0x2A: case 0x4A: case 0x6A: phantom_pc<FAST>(); a = rmw_op(op, a); return 2;

        // immediate (src/cpu.rs:336-356)
        case 0x69: case 0x29: case 0xC9: case 0xE0: case 0xC0: case 0x49:
        case 0xA9: case 0xA2: case 0xA0: case 0x09: case 0xE9:
            v = opnd<FAST>(b1);
            read_op(op, v);
            return 2;

        // zero page read / write / RMW (src/cpu.rs:362-410, 498-521)
        case 0xA5: case 0xA6: case 0xA4: case 0x65: case 0x25: case 0xC5:
        case 0xE4: case 0xC4: case 0x45: case 0x05: case 0xE5: case 0x24:
            ad = opnd<FAST>(b1);
            read_op(op, zp_read(ad));
            return 3;
        case 0x85: case 0x86: case 0x84:
            ad = opnd<FAST>(b1);
            zp_write(ad, store_reg(op));
            return 3;
        case 0x06: case 0x46: case 0x26: case 0x66: case 0xE6: case 0xC6:
            ad = opnd<FAST>(b1);
            v = zp_read(ad);
            if (!(ad & 0x80u)) tia_write(ad, v);                 // phantom write (RAM: same value)
            zp_write(ad, rmw_op(op, v));
            return 5;

        // zero page indexed (src/cpu.rs:415-493, 526-553); the phantom read of
        // the base address is RAM or TIA and has no side effect
        case 0xB5: case 0xB4: case 0x75: case 0x35: case 0xD5: case 0x55:
        case 0x15: case 0xF5: case 0xB6:
            ad = (opnd<FAST>(b1) + (op == 0xB6u ? y : x)) & 0xFFu;
            read_op(op, zp_read(ad));
            return 4;
        case 0x95: case 0x94: case 0x96:
            ad = (opnd<FAST>(b1) + (op == 0x96u ? y : x)) & 0xFFu;
            zp_write(ad, store_reg(op));
            return 4;
        case 0x16: case 0x56: case 0x36: case 0x76: case 0xF6: case 0xD6:
            ad = (opnd<FAST>(b1) + x) & 0xFFu;
            v = zp_read(ad);
            if (!(ad & 0x80u)) tia_write(ad, v);
            zp_write(ad, rmw_op(op, v));
            return 6;

        // absolute (src/cpu.rs:559-644)
        case 0xAD: case 0xAE: case 0xAC: case 0x6D: case 0x2D: case 0xCD:
        case 0xEC: case 0xCC: case 0x4D: case 0x0D: case 0xED: case 0x2C:
            ad = opnd<FAST>(b1);
            ad |= opnd<FAST>(b2) << 8;
            read_op(op, bus_read(ad));
            return 4;
        case 0x8D: case 0x8E: case 0x8C:
            ad = opnd<FAST>(b1);
            ad |= opnd<FAST>(b2) << 8;
            bus_write(ad, store_reg(op));
            return 4;
        case 0x0E: case 0x4E: case 0x2E: case 0x6E: case 0xEE: case 0xCE:
            ad = opnd<FAST>(b1);
            ad |= opnd<FAST>(b2) << 8;
            v = bus_read(ad);
            bus_write(ad, v);
            bus_write(ad, rmw_op(op, v));
            return 6;

        // absolute indexed (src/cpu.rs:649-788)
        case 0xBD: case 0xBC: case 0x7D: case 0x3D: case 0xDD: case 0x5D:
        case 0x1D: case 0xFD:
        case 0xB9: case 0xBE: case 0x79: case 0x39: case 0xD9: case 0x59:
        case 0x19: case 0xF9: {
            ad = opnd<FAST>(b1);
            ad |= opnd<FAST>(b2) << 8;
            uint32_t idx = ((op & 0x1Fu) == 0x19u || op == 0xBEu) ? y : x;
            eff = (ad + idx) & 0xFFFFu;
            uint32_t cross = (ad ^ eff) & 0xFF00u;
            if (cross) touch((ad & 0xFF00u) | (eff & 0xFFu));
            read_op(op, bus_read(eff));
            return cross ? 5 : 4;
        }
        case 0x9D: case 0x99:
            ad = opnd<FAST>(b1);
            ad |= opnd<FAST>(b2) << 8;
            eff = (ad + (op == 0x99u ? y : x)) & 0xFFFFu;
            touch((ad & 0xFF00u) | (eff & 0xFFu));
            bus_write(eff, a);
            return 5;
        case 0x1E: case 0x5E: case 0x3E: case 0x7E: case 0xFE: case 0xDE:
            ad = opnd<FAST>(b1);
            ad |= opnd<FAST>(b2) << 8;
            eff = (ad + x) & 0xFFFFu;
            touch((ad & 0xFF00u) | (eff & 0xFFu));
            v = bus_read(eff);
            bus_write(eff, v);
            bus_write(eff, rmw_op(op, v));
            return 7;

        // (zp,X) (src/cpu.rs:793-841); base and pointer reads are zero page
        case 0xA1: case 0x61: case 0x21: case 0xC1: case 0x41: case 0x01: case 0xE1: case 0x81: {
            uint32_t zp = (opnd<FAST>(b1) + x) & 0xFFu;
            ad = zp_read(zp) | (zp_read((zp + 1u) & 0xFFu) << 8);
            if (op == 0x81u) bus_write(ad, a);
            else read_op(op, bus_read(ad));
            return 6;
        }
        // (zp),Y (src/cpu.rs:846-901)
        case 0xB1: case 0x71: case 0x31: case 0xD1: case 0x51: case 0x11: case 0xF1: {
            uint32_t zp = opnd<FAST>(b1);
            ad = zp_read(zp) | (zp_read((zp + 1u) & 0xFFu) << 8);
            eff = (ad + y) & 0xFFFFu;
            uint32_t cross = (ad ^ eff) & 0xFF00u;
            if (cross) touch((ad & 0xFF00u) | (eff & 0xFFu));
            read_op(op, bus_read(eff));
            return cross ? 6 : 5;
        }
        case 0x91: {
            uint32_t zp = opnd<FAST>(b1);
            ad = zp_read(zp) | (zp_read((zp + 1u) & 0xFFu) << 8);
            eff = (ad + y) & 0xFFFFu;
            touch((ad & 0xFF00u) | (eff & 0xFFu));
            bus_write(eff, a);
            return 6;
        }

        // branches (src/cpu.rs:906-944)
        case 0x90: case 0xB0: case 0xF0: case 0x30: case 0xD0: case 0x10: case 0x50: case 0x70: {
            uint32_t off = opnd<FAST>(b1);
            if (!branch_taken(op)) return 2;
            phantom_pc<FAST>();
            uint32_t np = (pc + off - ((off & 0x80u) << 1)) & 0xFFFFu;
            uint32_t c = 3;
            if ((np ^ pc) & 0xFF00u) {
                touch((pc & 0xFF00u) | (np & 0xFFu));
                c = 4;
            }
            pc = np;
            return c;
        }

        // JMP abs / JMP (ind) (src/cpu.rs:949-978)
        case 0x4C: {
            uint32_t lo = opnd<FAST>(b1);
            uint32_t hi = FAST ? b2 : bus_read(pc);
            pc = lo | (hi << 8);
            return 3;
        }
        case 0x6C: {
            ad = opnd<FAST>(b1);
            ad |= opnd<FAST>(b2) << 8;
            uint32_t lo = bus_read(ad);
            uint32_t hi = bus_read((ad & 0xFF00u) | ((ad + 1u) & 0xFFu));
            pc = lo | (hi << 8);
            return 5;
        }
        // JSR (src/cpu.rs:983-1005): the phantom stack read has no side effect
        case 0x20: {
            uint32_t lo = opnd<FAST>(b1);
            push(pc >> 8);
            push(pc & 0xFFu);
            uint32_t hi = FAST ? b2 : bus_read(pc);
            pc = lo | (hi << 8);
            return 6;
        }
        // RTS (src/cpu.rs:1010-1028)
        case 0x60: {
            phantom_pc<FAST>();
            uint32_t lo = pull();
            uint32_t hi = pull();
            ad = lo | (hi << 8);
            touch(ad);
            pc = (ad + 1u) & 0xFFFFu;
            return 6;
        }
        // RTI (src/cpu.rs:1033-1051)
        case 0x40: {
            phantom_pc<FAST>();
            p = pull() & ~(FB | FU);
            uint32_t lo = pull();
            uint32_t hi = pull();
            pc = lo | (hi << 8);
            return 6;
        }
        // BRK (src/cpu.rs:1056-1080)
        case 0x00: {
            phantom_pc<FAST>();
            pc = (pc + 1u) & 0xFFFFu;
            push(pc >> 8);
            push(pc & 0xFFu);
            push(p | FB | FU);
            p |= FI;
            uint32_t lo = bus_read(0xFFFEu);
            uint32_t hi = bus_read(0xFFFFu);
            pc = lo | (hi << 8);
            return 7;
        }
        // stack (src/cpu.rs:1085-1137)
        case 0x48: phantom_pc<FAST>(); push(a); return 3;
        case 0x08: phantom_pc<FAST>(); push(p | FB | FU); return 3;
        case 0x68: phantom_pc<FAST>(); a = pull(); set_zn(a); return 4;
        case 0x28: phantom_pc<FAST>(); p = pull() & ~(FB | FU); return 4;

        default:
            return 0;
        }
    }

    // PC is inside the ROM window and pc+2 stays below the hotspot/vector
    // area, so the opcode and operand fetches have no side effects.
    __device__ __forceinline__ bool fast_fetch() const {
        return (pc & 0x1000u) && ((pc & 0xFFFu) < 0xFF0u);
    }
    __device__ __forceinline__ uint32_t rom_off() const { return (bank << 12) + (pc & 0xFFFu); }
};

}  // namespace fa
