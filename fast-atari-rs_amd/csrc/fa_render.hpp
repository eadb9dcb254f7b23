// fa_render.hpp -- the rendering console on the GPU: one wave64 lane per env,
// a cycle-stepped 6502 and the pixel-rendering TIA.
//
// Behaviour follows the reference's rendering console: the pixel TIA
// (src/tia.rs:98-675), the rendering bus with its per-access clocking
// (src/atari.rs:107-224: a read returns the value then clocks the TIA three
// times and the RIOT once; a write clocks the TIA twice, writes, clocks the
// TIA once more and the RIOT once) and Atari::{new, set_action, run_frame}
// (src/atari.rs:226-318), with the CPU of src/cpu.rs:245-1146.
//
// Every loop iteration is one bus cycle (or one WSYNC stall cycle) of every
// lane: the instruction's next bus access comes from a micro-program (one
// 32-bit micro-op per cycle, chosen by the opcode's addressing class and the
// cycle index), so lanes at different opcodes share one code path -- one
// access site and one TIA clock site -- and only the micro-op's data differs.
// The register / flag effect of an instruction runs once, on the cycle its
// micro-program marks.
#pragma once

#include <stdint.h>

#include "fa_decode.hpp"

namespace fa {
namespace rnd {

// ---------------------------------------------------------------- micro-ops
// bits 0-3: address source
enum : uint32_t {
    AS_PC = 0, AS_OPND = 1, AS_ZPI = 2, AS_PTRX0 = 3, AS_PTRX1 = 4, AS_PTR1 = 5, AS_ABS = 6, AS_WRONG = 7,
    AS_EA = 8, AS_STACK = 9, AS_VEC0 = 10, AS_VEC1 = 11, AS_JIND1 = 12, AS_BRWRONG = 13,
};
constexpr uint32_t U_W = 1u << 4;                 // write cycle (else read)
// bits 5-7: read destination / write source
enum : uint32_t { DS_NONE = 0, DS_OPND = 1, DS_LO = 2, DS_HI = 3, DS_DAT = 4, DS_P = 5 };
enum : uint32_t { WS_STORE = 0, WS_DAT = 1, WS_RES = 2, WS_PCH = 3, WS_PCL = 4, WS_PPUSH = 5 };
constexpr uint32_t U_PCINC = 1u << 8;             // PC + 1 after the cycle
constexpr uint32_t U_SPDEC = 1u << 9;             // SP - 1 after the cycle (push)
constexpr uint32_t U_SPINC = 1u << 10;            // SP + 1 before the cycle (pull)
constexpr uint32_t U_IFCROSS = 1u << 11;          // cycle only when the indexed address crosses a page
constexpr uint32_t U_LAST = 1u << 12;             // the instruction's last cycle
constexpr uint32_t U_LAST_NT = 1u << 13;          // last when the branch is not taken
constexpr uint32_t U_LAST_NC = 1u << 14;          // last when the taken branch stays on its page
constexpr uint32_t U_ALU = 1u << 15;              // the register / flag effect follows this cycle
constexpr uint32_t U_SETI = 1u << 16;             // set I after this cycle (BRK)
// bits 17-19: the PC update of the last cycle
enum : uint32_t { F_NONE = 0, F_JMP = 1, F_RTS = 2, F_BRANCH = 3, F_JMPI = 4 };

constexpr uint32_t rd(uint32_t as, uint32_t ds, uint32_t fl = 0) { return as | (ds << 5) | fl; }
constexpr uint32_t wr(uint32_t as, uint32_t ws, uint32_t fl = 0) { return as | U_W | (ws << 5) | fl; }
constexpr uint32_t fin(uint32_t f) { return f << 17; }

// micro-programs (cycles after the opcode fetch), by id
enum : uint32_t {
    P_IMP, P_IMM, P_ZP_R, P_ZP_W, P_ZP_M, P_ZPI_R, P_ZPI_W, P_ZPI_M, P_ABS_R, P_ABS_W, P_ABS_M,
    P_ABI_R, P_ABI_W, P_ABI_M, P_IZX_R, P_IZX_W, P_IZY_R, P_IZY_W, P_REL, P_JMP, P_JMPI, P_JSR,
    P_RTS, P_RTI, P_BRK, P_PHA, P_PHP, P_PL, P_ILLEGAL, P_COUNT,
};
constexpr uint32_t kProgLen = 8;                  // micro-ops per program (at most 6 used)

struct Prog { uint32_t u[kProgLen]; };

constexpr Prog prog_of(uint32_t id) {
    const uint32_t OPND = rd(AS_PC, DS_OPND, U_PCINC), LO = rd(AS_PC, DS_LO, U_PCINC), HI = rd(AS_PC, DS_HI, U_PCINC);
    switch (id) {
    case P_IMP: return {{rd(AS_PC, DS_NONE, U_LAST | U_ALU)}};
    case P_IMM: return {{rd(AS_PC, DS_DAT, U_PCINC | U_LAST | U_ALU)}};
    case P_ZP_R: return {{OPND, rd(AS_OPND, DS_DAT, U_LAST | U_ALU)}};
    case P_ZP_W: return {{OPND, wr(AS_OPND, WS_STORE, U_LAST)}};
    case P_ZP_M: return {{OPND, rd(AS_OPND, DS_DAT, U_ALU), wr(AS_OPND, WS_DAT), wr(AS_OPND, WS_RES, U_LAST)}};
    case P_ZPI_R: return {{OPND, rd(AS_OPND, DS_NONE), rd(AS_ZPI, DS_DAT, U_LAST | U_ALU)}};
    case P_ZPI_W: return {{OPND, rd(AS_OPND, DS_NONE), wr(AS_ZPI, WS_STORE, U_LAST)}};
    case P_ZPI_M: return {{OPND, rd(AS_OPND, DS_NONE), rd(AS_ZPI, DS_DAT, U_ALU), wr(AS_ZPI, WS_DAT),
                           wr(AS_ZPI, WS_RES, U_LAST)}};
    case P_ABS_R: return {{LO, HI, rd(AS_ABS, DS_DAT, U_LAST | U_ALU)}};
    case P_ABS_W: return {{LO, HI, wr(AS_ABS, WS_STORE, U_LAST)}};
    case P_ABS_M: return {{LO, HI, rd(AS_ABS, DS_DAT, U_ALU), wr(AS_ABS, WS_DAT), wr(AS_ABS, WS_RES, U_LAST)}};
    case P_ABI_R: return {{LO, HI, rd(AS_WRONG, DS_NONE, U_IFCROSS), rd(AS_EA, DS_DAT, U_LAST | U_ALU)}};
    case P_ABI_W: return {{LO, HI, rd(AS_WRONG, DS_NONE), wr(AS_EA, WS_STORE, U_LAST)}};
    case P_ABI_M: return {{LO, HI, rd(AS_WRONG, DS_NONE), rd(AS_EA, DS_DAT, U_ALU), wr(AS_EA, WS_DAT),
                           wr(AS_EA, WS_RES, U_LAST)}};
    case P_IZX_R: return {{OPND, rd(AS_OPND, DS_NONE), rd(AS_PTRX0, DS_LO), rd(AS_PTRX1, DS_HI),
                           rd(AS_ABS, DS_DAT, U_LAST | U_ALU)}};
    case P_IZX_W: return {{OPND, rd(AS_OPND, DS_NONE), rd(AS_PTRX0, DS_LO), rd(AS_PTRX1, DS_HI),
                           wr(AS_ABS, WS_STORE, U_LAST)}};
    case P_IZY_R: return {{OPND, rd(AS_OPND, DS_LO), rd(AS_PTR1, DS_HI), rd(AS_WRONG, DS_NONE, U_IFCROSS),
                           rd(AS_EA, DS_DAT, U_LAST | U_ALU)}};
    case P_IZY_W: return {{OPND, rd(AS_OPND, DS_LO), rd(AS_PTR1, DS_HI), rd(AS_WRONG, DS_NONE),
                           wr(AS_EA, WS_STORE, U_LAST)}};
    case P_REL: return {{rd(AS_PC, DS_OPND, U_PCINC | U_LAST_NT), rd(AS_PC, DS_NONE, U_LAST_NC | fin(F_BRANCH)),
                         rd(AS_BRWRONG, DS_NONE, U_LAST | fin(F_BRANCH))}};
    case P_JMP: return {{LO, rd(AS_PC, DS_HI, U_LAST | fin(F_JMP))}};
    case P_JMPI: return {{LO, HI, rd(AS_ABS, DS_DAT), rd(AS_JIND1, DS_HI, U_LAST | fin(F_JMPI))}};
    case P_JSR: return {{LO, rd(AS_STACK, DS_NONE), wr(AS_STACK, WS_PCH, U_SPDEC), wr(AS_STACK, WS_PCL, U_SPDEC),
                         rd(AS_PC, DS_HI, U_LAST | fin(F_JMP))}};
    case P_RTS: return {{rd(AS_PC, DS_NONE), rd(AS_STACK, DS_NONE), rd(AS_STACK, DS_LO, U_SPINC),
                         rd(AS_STACK, DS_HI, U_SPINC), rd(AS_ABS, DS_NONE, U_LAST | fin(F_RTS))}};
    case P_RTI: return {{rd(AS_PC, DS_NONE), rd(AS_STACK, DS_NONE), rd(AS_STACK, DS_P, U_SPINC),
                         rd(AS_STACK, DS_LO, U_SPINC), rd(AS_STACK, DS_HI, U_SPINC | U_LAST | fin(F_JMP))}};
    case P_BRK: return {{rd(AS_PC, DS_NONE, U_PCINC), wr(AS_STACK, WS_PCH, U_SPDEC), wr(AS_STACK, WS_PCL, U_SPDEC),
                         wr(AS_STACK, WS_PPUSH, U_SPDEC | U_SETI), rd(AS_VEC0, DS_LO),
                         rd(AS_VEC1, DS_HI, U_LAST | fin(F_JMP))}};
    case P_PHA: return {{rd(AS_PC, DS_NONE), wr(AS_STACK, WS_STORE, U_SPDEC | U_LAST)}};
    case P_PHP: return {{rd(AS_PC, DS_NONE), wr(AS_STACK, WS_PPUSH, U_SPDEC | U_LAST)}};
    case P_PL: return {{rd(AS_PC, DS_NONE), rd(AS_STACK, DS_NONE), rd(AS_STACK, DS_DAT, U_SPINC | U_LAST | U_ALU)}};
    default: return {{}};
    }
}

// Per-opcode render decode word: program id (bits 0-4), index register for
// ZPI / ABI / IZY-style addressing (bit 5: 1 = Y), store register (bits 6-7:
// 0 A, 1 X, 2 Y), special effect (bits 8-10: 1 PLA, 2 PLP).
enum : uint32_t { RX_IDXY = 1u << 5, RX_SREG_SHIFT = 6, RX_SPEC_SHIFT = 8 };
enum : uint32_t { SX_NONE = 0, SX_PLA = 1, SX_PLP = 2 };

constexpr uint32_t rdesc_of(uint32_t code) {
    using namespace detail;
    for (const Op &o : kOps) {
        if (o.code != code) continue;
        const Mn m = o.mn;
        const Mode md = o.mode;
        const bool store = m == STA || m == STX || m == STY;
        const bool rmw = (m == ASL || m == LSR || m == ROL || m == ROR || m == INC || m == DEC) && md != ACC;
        const uint32_t k = store ? 1 : (rmw ? 2 : 0);       // read / write / rmw
        uint32_t p = P_IMP, x = 0;
        switch (md) {
        case IMP: case ACC: p = P_IMP; break;
        case IMM: p = P_IMM; break;
        case ZP: p = P_ZP_R + k; break;
        case ZPX: p = P_ZPI_R + k; break;
        case ZPY: p = P_ZPI_R + k; x |= RX_IDXY; break;
        case ABS: p = P_ABS_R + k; break;
        case ABX: p = P_ABI_R + k; break;
        case ABY: p = P_ABI_R + k; x |= RX_IDXY; break;
        case IZX: p = P_IZX_R + (k ? 1 : 0); break;
        case IZY: p = P_IZY_R + (k ? 1 : 0); x |= RX_IDXY; break;
        case REL: p = P_REL; break;
        case IND: p = P_JMPI; break;
        }
        if (m == JMP && md == ABS) p = P_JMP;
        if (m == JSR) p = P_JSR;
        if (m == RTS) p = P_RTS;
        if (m == RTI) p = P_RTI;
        if (m == BRK) p = P_BRK;
        if (m == PHA) p = P_PHA;
        if (m == PHP) p = P_PHP;
        if (m == PLA) { p = P_PL; x |= SX_PLA << RX_SPEC_SHIFT; }
        if (m == PLP) { p = P_PL; x |= SX_PLP << RX_SPEC_SHIFT; }
        const uint32_t sreg = m == STX ? 1u : (m == STY ? 2u : 0u);
        return p | x | (sreg << RX_SREG_SHIFT);
    }
    return P_ILLEGAL;
}

// LDS tables of one workgroup: micro-programs, then per opcode {rdesc, d0,
// g, h}: for compute-class opcodes g / h are the one-hot selection words of
// the per-lane compute stage (fa_decode.hpp d2 / d3); for branches h holds
// d1's tested flag mask and taken-when-set bit (D1_FMASK_SHIFT,
// D1_FVAL_SHIFT).
constexpr uint32_t kProgBytes = P_COUNT * kProgLen * 4;
constexpr uint32_t kOpBytes = 256 * 16;
constexpr uint32_t kTableBytes = kProgBytes + kOpBytes;
// LDS bytes per lane: RAM, the open 16-pixel group, and the six objects'
// 160-pixel line masks (5 words each).
constexpr uint32_t kMaskWords = 5;
constexpr uint32_t kLaneBytes = 128 + 16 + 6 * kMaskWords * 4;

// The sets of objects on one pixel (bit 0 P0, 1 P1, 2 M0, 3 M1, 4 BL, 5 PF)
// in which objects i and j overlap, as a 64-bit mask over the 64 sets.
constexpr uint64_t overlap_sets(uint32_t i, uint32_t j) {
    uint64_t m = 0;
    for (uint32_t o = 0; o < 64; ++o)
        if (((o >> i) & 1u) && ((o >> j) & 1u)) m |= 1ull << o;
    return m;
}

inline void build_tables(uint32_t *out) {
    for (uint32_t p = 0; p < P_COUNT; ++p) {
        const Prog g = prog_of(p);
        for (uint32_t i = 0; i < kProgLen; ++i) out[p * kProgLen + i] = g.u[i];
    }
    uint32_t *ops = out + kProgBytes / 4;
    for (uint32_t c = 0; c < 256; ++c) {
        const uint64_t d = descriptor_of(c);
        const uint32_t d0 = (uint32_t)d, d1 = (uint32_t)(d >> 32);
        uint32_t g = 0, h = 0;
        describe_lanes(d0, d1, g, h);
        // stack / jump instructions run micro-programs here: the frame
        // kernel's straight-line stack and jump bits do not apply
        if ((d0 & 3u) == CLS_SPECIAL) g = h = 0;
        if ((d0 & 3u) == CLS_BRANCH) h = d1 & 0xFFFF0000u;
        ops[4 * c] = rdesc_of(c);
        ops[4 * c + 1] = d0;
        ops[4 * c + 2] = g;
        ops[4 * c + 3] = h;
    }
}

}  // namespace rnd
}  // namespace fa
