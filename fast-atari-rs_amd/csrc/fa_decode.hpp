// fa_decode.hpp -- per-opcode micro-op descriptors for the uniform pipeline.
//
// Every documented NMOS 6502 opcode (151) is described by two 32-bit words
// that steer one shared fetch / address / read / compute / write pipeline
// with predicated selects instead of a per-opcode switch, plus two words with
// the same decode as one-hot masks for the per-lane path and, in the table
// the kernels stage, four words of register selectors (32 bytes an opcode).  Cycle counts and
// the order of bus accesses are those of the reference's Cpu::step
// (src/cpu.rs:245-1146); the table itself is derived from the instruction
// set (mnemonic x addressing mode), not from the reference's code.
#pragma once

#include <stdint.h>

namespace fa {

// ---- d0
enum : uint32_t {
    D_CLS_SHIFT = 0,        // 2 bits: class
    D_LEN_SHIFT = 2,        // 2 bits: instruction length 1..3
    D_CYC_SHIFT = 4,        // 3 bits: base cycles
    D_IMM = 1u << 7,        // operand is the immediate byte
    D_ZP = 1u << 8,         // 8-bit address arithmetic (zero page)
    D_IDX_SHIFT = 9,        // 2 bits: index 0 none, 1 X, 2 Y
    D_IND_SHIFT = 11,       // 2 bits: 0 none, 1 (zp,X), 2 (zp),Y
    D_MEM = 1u << 13,       // memory operand (an effective address exists)
    D_RD = 1u << 14,        // data read
    D_WR = 1u << 15,        // final data write
    D_RMWMEM = 1u << 16,    // phantom write of the old value before the final write
    D_DUMMY_ALWAYS = 1u << 17,   // fix-up read at (base & $FF00) | (ea & $FF), always
    D_DUMMY_CROSS = 1u << 18,    // fix-up read and +1 cycle only on a page cross
    D_KIND_SHIFT = 19,      // 3 bits: compute kind
    D_M_AND = 1u << 22, D_M_XOR = 1u << 23, D_M_LD = 1u << 24,   // logic masks
    D_INV = 1u << 25,       // arith: operand inverted (SBC, compares)
    D_CIN1 = 1u << 26,      // arith: carry-in forced to 1 (compares)
    D_KEEP = 1u << 27,      // arith: result not written (compares)
    D_SETV = 1u << 28,      // arith: V updated (ADC, SBC)
    D_DEC = 1u << 29,       // arith: decimal mode applies (ADC, SBC)
    D_SETNZ = 1u << 30,
    D_SETC = 1u << 31,
};
enum : uint32_t { CLS_COMPUTE = 0, CLS_BRANCH = 1, CLS_SPECIAL = 2, CLS_ILLEGAL = 3 };
enum : uint32_t { K_LOGIC = 0, K_ARITH = 1, K_BIT = 2, K_STORE = 3, K_RMW = 4, K_MOVE = 5, K_FLAGS = 6, K_NOP = 7 };

// ---- d1
enum : uint32_t {
    D1_SRC_SHIFT = 0,       // 3 bits: 0 A, 1 X, 2 Y, 3 SP
    D1_DST_SHIFT = 3,       // 3 bits: 0 A, 1 X, 2 Y, 3 SP, 4 none
    D1_RMW_SHIFT = 6,       // 3 bits: 0 ASL, 1 ROL, 2 LSR, 3 ROR, 4 DEC, 5 INC
    D1_DELTA_SHIFT = 9,     // 2 bits: 0 none, 1 +1, 2 -1
    D1_SPEC_SHIFT = 11,     // 4 bits: special id
    D1_SLOW = 1u << 15,     // descriptor table only: JSR, RTS, RTI, BRK, JMP (ind) (the per-lane slow pass)
    D1_FMASK_SHIFT = 16,    // 8 bits: flags written (K_FLAGS) / tested (branch)
    D1_FVAL_SHIFT = 24,     // 8 bits: value written / branch taken when set
};
enum : uint32_t { R_A = 0, R_X = 1, R_Y = 2, R_SP = 3, R_NONE = 4 };
enum : uint32_t { SP_JMP = 0, SP_JMPI = 1, SP_JSR = 2, SP_RTS = 3, SP_RTI = 4, SP_BRK = 5,
                  SP_PHA = 6, SP_PHP = 7, SP_PLA = 8, SP_PLP = 9 };

// ---- d2, d3: the same decode as one-hot bits and ready-made flag masks, for
// the per-lane path, where every selection is an AND with a sign-extended bit
// (one v_bfe_i32) instead of a compare on a field.  Derived from d0 / d1.
enum : uint32_t {
    G_SRC = 0,              // 4 bits: source register A, X, Y, SP (one-hot)
    G_DST = 4,              // 4 bits: destination register A, X, Y, SP (one-hot; none: 0)
    G_R_LOG = 8,            // result: logic (AND / EOR / ORA / loads / BIT's A & v)
    G_R_ARI = 9,            // result: binary add with carry
    G_R_SHL = 10,           // result: ASL / ROL
    G_R_SHR = 11,           // result: LSR / ROR
    G_R_INC = 12,           // result: operand + delta (INC, DEC, INX .. DEY, transfers)
    G_R_SRC = 13,           // result: the source register (stores)
    G_CIN_SHL = 14,         // ROL: carry into bit 0
    G_CIN_SHR = 15,         // ROR: carry into bit 7
    G_DELTA = 16,           // 8 bits: delta of G_R_INC (0, 1 or $FF)
    G_IX_X = 24,            // X added to the base address
    G_IX_Y = 25,            // Y added to the base (absolute / zero-page Y, (zp),Y)
    G_IZX = 26,             // (zp,X): X added to the zero-page pointer address
    G_IND = 27,             // base address read from a zero-page pointer
    G_BIT = 28,             // BIT: N and V from the operand
    G_C_SUM = 29,           // new carry: bit 8 of the sum
    G_C_B7 = 30,            // new carry: operand bit 7 (ASL / ROL)
    G_C_B0 = 31,            // new carry: operand bit 0 (LSR / ROR)
};
enum : uint32_t {
    H_CLR = 0,              // 8 bits: status bits this instruction writes
    H_TAKE = 8,             // 8 bits: status bits taken from the computed N, V, Z, C
    H_SET = 16,             // 8 bits: status bits forced to 1 (SEC, SED, SEI)
    H_M_AND = 24, H_M_XOR = 25, H_M_LD = 26,   // logic masks (d0's D_M_*, BIT: AND)
    H_BR = 27,              // conditional branch: target = PC + offset (d1: flag mask, taken value)
    H_JMP = 28,             // JMP absolute: PC = the operand
    H_PUSH = 29,            // PHA / PHP: the source written at $0100 + SP, SP - 1
    H_PULL = 30,            // PLA / PLP: SP + 1, the result read from $0100 + SP
    H_PREG = 31,            // PHP / PLP: the status register is the source / destination
};
// ---- d4 .. d7 (descriptor table only): the per-lane stage keeps A, X, Y and
// SP as the bytes of one word (A | X << 8 | Y << 16 | SP << 24) for the
// step, and selects from it with v_perm_b32 (byte selectors 0-3: a register,
// 4: the result, 0x0C: zero):
//   d4  the source register into byte 0 (none: 0)
//   d5  byte 0: the index added to the address (X / Y / none), byte 1: the
//       index added to a (zp,X) pointer address (X / none)
//   d6  the result into the destination register's byte, the others kept
//   d7  added to the word before d6: SP + 1 (pulls) or SP - 1 (pushes) in the
//       top byte, which wraps within the byte
// All four follow from d2 / d3 (build_descriptors).
constexpr uint32_t kDescWords = 8;                  // words per opcode
constexpr uint32_t kDescBytes = 256u * kDescWords * 4u;

// d2 and d3 of an opcode from its d0 and d1.
constexpr void describe_lanes(uint32_t d0, uint32_t d1, uint32_t &d2, uint32_t &d3) {
    d2 = 0;
    d3 = 0;
    constexpr uint32_t NZ = 0x82;
    if ((d0 & 3u) == CLS_BRANCH) { d3 = 1u << H_BR; return; }
    if ((d0 & 3u) == CLS_SPECIAL) {
        switch ((d1 >> D1_SPEC_SHIFT) & 15u) {
        case SP_JMP: d3 = 1u << H_JMP; break;
        case SP_PHA: d2 = (1u << G_SRC) | (1u << G_R_SRC); d3 = 1u << H_PUSH; break;
        case SP_PHP: d2 = 1u << G_R_SRC; d3 = (1u << H_PUSH) | (1u << H_PREG); break;
        case SP_PLA:
            d2 = (1u << (G_DST + R_A)) | (1u << G_R_LOG);
            d3 = (1u << H_PULL) | (1u << H_M_LD) | (NZ << H_CLR) | (NZ << H_TAKE);
            break;
        case SP_PLP: d2 = 1u << G_R_LOG; d3 = (1u << H_PULL) | (1u << H_PREG) | (1u << H_M_LD) | (0xFFu << H_CLR); break;
        default: break;
        }
        return;
    }
    if ((d0 & 3u) != CLS_COMPUTE) return;
    const uint32_t kind = (d0 >> D_KIND_SHIFT) & 7u;
    const uint32_t src = (d1 >> D1_SRC_SHIFT) & 7u, dst = (d1 >> D1_DST_SHIFT) & 7u;
    const uint32_t rop = (d1 >> D1_RMW_SHIFT) & 7u, dl = (d1 >> D1_DELTA_SHIFT) & 3u;
    const uint32_t ind = (d0 >> D_IND_SHIFT) & 3u, idx = (d0 >> D_IDX_SHIFT) & 3u;
    constexpr uint32_t N = 0x80, V = 0x40, Z = 0x02, C = 0x01;
    d2 |= 1u << (G_SRC + (src & 3u));
    if (dst < 4u) d2 |= 1u << (G_DST + dst);
    uint32_t delta = 0;
    switch (kind) {
    case K_LOGIC: d2 |= 1u << G_R_LOG; break;
    case K_ARITH: d2 |= 1u << G_R_ARI; break;
    case K_BIT: d2 |= (1u << G_R_LOG) | (1u << G_BIT); d3 |= 1u << H_M_AND; break;
    case K_RMW:
        if (rop < 2u) d2 |= (1u << G_R_SHL) | (rop == 1u ? 1u << G_CIN_SHL : 0u);
        else if (rop < 4u) d2 |= (1u << G_R_SHR) | (rop == 3u ? 1u << G_CIN_SHR : 0u);
        else { d2 |= 1u << G_R_INC; delta = rop == 5u ? 1u : 0xFFu; }
        break;
    case K_MOVE: d2 |= 1u << G_R_INC; delta = dl == 1u ? 1u : (dl == 2u ? 0xFFu : 0u); break;
    default: d2 |= 1u << G_R_SRC; break;
    }
    d2 |= delta << G_DELTA;
    if (ind == 1u) d2 |= (1u << G_IND) | (1u << G_IZX);
    else if (ind == 2u) d2 |= (1u << G_IND) | (1u << G_IX_Y);
    else if (idx == 1u) d2 |= 1u << G_IX_X;
    else if (idx == 2u) d2 |= 1u << G_IX_Y;
    uint32_t take = 0;
    if (d0 & D_SETNZ) take |= N | Z;
    if (d0 & D_SETV) take |= V;
    if (kind == K_BIT) take |= N | V | Z;
    if (d0 & D_SETC) {
        take |= C;
        d2 |= kind == K_ARITH ? 1u << G_C_SUM : (rop < 2u ? 1u << G_C_B7 : 1u << G_C_B0);
    }
    const uint32_t fm = kind == K_FLAGS ? (d1 >> D1_FMASK_SHIFT) & 0xFFu : 0u;
    const uint32_t fv = kind == K_FLAGS ? (d1 >> D1_FVAL_SHIFT) & 0xFFu : 0u;
    d3 |= ((take | fm) << H_CLR) | (take << H_TAKE) | (fv << H_SET);
    if (d0 & D_M_AND) d3 |= 1u << H_M_AND;
    if (d0 & D_M_XOR) d3 |= 1u << H_M_XOR;
    if (d0 & D_M_LD) d3 |= 1u << H_M_LD;
}

namespace detail {

enum Mode : uint8_t { IMP, ACC, IMM, ZP, ZPX, ZPY, ABS, ABX, ABY, IZX, IZY, REL, IND };
enum Mn : uint8_t {
    ADC, AND, ASL, BCC, BCS, BEQ, BIT, BMI, BNE, BPL, BRK, BVC, BVS, CLC, CLD, CLI, CLV, CMP, CPX, CPY,
    DEC, DEX, DEY, EOR, INC, INX, INY, JMP, JSR, LDA, LDX, LDY, LSR, NOP, ORA, PHA, PHP, PLA, PLP, ROL,
    ROR, RTI, RTS, SBC, SEC, SED, SEI, STA, STX, STY, TAX, TAY, TSX, TXA, TXS, TYA,
};
struct Op { uint8_t code; Mn mn; Mode mode; };

// The documented instruction set: opcode, mnemonic, addressing mode.
constexpr Op kOps[151] = {
    {0x00, BRK, IMP}, {0x01, ORA, IZX}, {0x05, ORA, ZP}, {0x06, ASL, ZP}, {0x08, PHP, IMP}, {0x09, ORA, IMM},
    {0x0A, ASL, ACC}, {0x0D, ORA, ABS}, {0x0E, ASL, ABS}, {0x10, BPL, REL}, {0x11, ORA, IZY}, {0x15, ORA, ZPX},
    {0x16, ASL, ZPX}, {0x18, CLC, IMP}, {0x19, ORA, ABY}, {0x1D, ORA, ABX}, {0x1E, ASL, ABX}, {0x20, JSR, ABS},
    {0x21, AND, IZX}, {0x24, BIT, ZP}, {0x25, AND, ZP}, {0x26, ROL, ZP}, {0x28, PLP, IMP}, {0x29, AND, IMM},
    {0x2A, ROL, ACC}, {0x2C, BIT, ABS}, {0x2D, AND, ABS}, {0x2E, ROL, ABS}, {0x30, BMI, REL}, {0x31, AND, IZY},
    {0x35, AND, ZPX}, {0x36, ROL, ZPX}, {0x38, SEC, IMP}, {0x39, AND, ABY}, {0x3D, AND, ABX}, {0x3E, ROL, ABX},
    {0x40, RTI, IMP}, {0x41, EOR, IZX}, {0x45, EOR, ZP}, {0x46, LSR, ZP}, {0x48, PHA, IMP}, {0x49, EOR, IMM},
    {0x4A, LSR, ACC}, {0x4C, JMP, ABS}, {0x4D, EOR, ABS}, {0x4E, LSR, ABS}, {0x50, BVC, REL}, {0x51, EOR, IZY},
    {0x55, EOR, ZPX}, {0x56, LSR, ZPX}, {0x58, CLI, IMP}, {0x59, EOR, ABY}, {0x5D, EOR, ABX}, {0x5E, LSR, ABX},
    {0x60, RTS, IMP}, {0x61, ADC, IZX}, {0x65, ADC, ZP}, {0x66, ROR, ZP}, {0x68, PLA, IMP}, {0x69, ADC, IMM},
    {0x6A, ROR, ACC}, {0x6C, JMP, IND}, {0x6D, ADC, ABS}, {0x6E, ROR, ABS}, {0x70, BVS, REL}, {0x71, ADC, IZY},
    {0x75, ADC, ZPX}, {0x76, ROR, ZPX}, {0x78, SEI, IMP}, {0x79, ADC, ABY}, {0x7D, ADC, ABX}, {0x7E, ROR, ABX},
    {0x81, STA, IZX}, {0x84, STY, ZP}, {0x85, STA, ZP}, {0x86, STX, ZP}, {0x88, DEY, IMP}, {0x8A, TXA, IMP},
    {0x8C, STY, ABS}, {0x8D, STA, ABS}, {0x8E, STX, ABS}, {0x90, BCC, REL}, {0x91, STA, IZY}, {0x94, STY, ZPX},
    {0x95, STA, ZPX}, {0x96, STX, ZPY}, {0x98, TYA, IMP}, {0x99, STA, ABY}, {0x9A, TXS, IMP}, {0x9D, STA, ABX},
    {0xA0, LDY, IMM}, {0xA1, LDA, IZX}, {0xA2, LDX, IMM}, {0xA4, LDY, ZP}, {0xA5, LDA, ZP}, {0xA6, LDX, ZP},
    {0xA8, TAY, IMP}, {0xA9, LDA, IMM}, {0xAA, TAX, IMP}, {0xAC, LDY, ABS}, {0xAD, LDA, ABS}, {0xAE, LDX, ABS},
    {0xB0, BCS, REL}, {0xB1, LDA, IZY}, {0xB4, LDY, ZPX}, {0xB5, LDA, ZPX}, {0xB6, LDX, ZPY}, {0xB8, CLV, IMP},
    {0xB9, LDA, ABY}, {0xBA, TSX, IMP}, {0xBC, LDY, ABX}, {0xBD, LDA, ABX}, {0xBE, LDX, ABY}, {0xC0, CPY, IMM},
    {0xC1, CMP, IZX}, {0xC4, CPY, ZP}, {0xC5, CMP, ZP}, {0xC6, DEC, ZP}, {0xC8, INY, IMP}, {0xC9, CMP, IMM},
    {0xCA, DEX, IMP}, {0xCC, CPY, ABS}, {0xCD, CMP, ABS}, {0xCE, DEC, ABS}, {0xD0, BNE, REL}, {0xD1, CMP, IZY},
    {0xD5, CMP, ZPX}, {0xD6, DEC, ZPX}, {0xD8, CLD, IMP}, {0xD9, CMP, ABY}, {0xDD, CMP, ABX}, {0xDE, DEC, ABX},
    {0xE0, CPX, IMM}, {0xE1, SBC, IZX}, {0xE4, CPX, ZP}, {0xE5, SBC, ZP}, {0xE6, INC, ZP}, {0xE8, INX, IMP},
    {0xE9, SBC, IMM}, {0xEA, NOP, IMP}, {0xEC, CPX, ABS}, {0xED, SBC, ABS}, {0xEE, INC, ABS}, {0xF0, BEQ, REL},
    {0xF1, SBC, IZY}, {0xF5, SBC, ZPX}, {0xF6, INC, ZPX}, {0xF8, SED, IMP}, {0xF9, SBC, ABY}, {0xFD, SBC, ABX},
    {0xFE, INC, ABX},
};

constexpr uint32_t kC = 0x01, kZ = 0x02, kI = 0x04, kD = 0x08, kV = 0x40, kN = 0x80;

// Returns {d0, d1} for one documented opcode.
constexpr void describe(const Op &o, uint32_t &d0, uint32_t &d1) {
    const Mn m = o.mn;
    const Mode md = o.mode;
    const bool store = m == STA || m == STX || m == STY;
    const bool rmw = m == ASL || m == LSR || m == ROL || m == ROR || m == INC || m == DEC;
    uint32_t len = 1, cyc = 2;
    d0 = 0;
    d1 = (R_NONE << D1_DST_SHIFT);
    // addressing
    switch (md) {
    case IMP: case ACC: len = 1; break;
    case IMM: len = 2; d0 |= D_IMM; break;
    case ZP: len = 2; d0 |= D_ZP | D_MEM; break;
    case ZPX: len = 2; d0 |= D_ZP | D_MEM | (1u << D_IDX_SHIFT); break;
    case ZPY: len = 2; d0 |= D_ZP | D_MEM | (2u << D_IDX_SHIFT); break;
    case ABS: len = 3; d0 |= D_MEM; break;
    case ABX: len = 3; d0 |= D_MEM | (1u << D_IDX_SHIFT); break;
    case ABY: len = 3; d0 |= D_MEM | (2u << D_IDX_SHIFT); break;
    case IZX: len = 2; d0 |= D_MEM | (1u << D_IND_SHIFT); break;
    case IZY: len = 2; d0 |= D_MEM | (2u << D_IND_SHIFT); break;
    case REL: len = 2; break;
    case IND: len = 3; break;
    }
    // cycles by mode and access kind (src/cpu.rs:245-1146)
    const int acc = store ? 1 : (rmw && md != ACC ? 2 : 0);   // 0 read, 1 write, 2 rmw
    switch (md) {
    case IMP: case ACC: case IMM: cyc = 2; break;
    case ZP: cyc = acc == 2 ? 5 : 3; break;
    case ZPX: case ZPY: case ABS: cyc = acc == 2 ? 6 : 4; break;
    case ABX: case ABY: cyc = acc == 2 ? 7 : (acc == 1 ? 5 : 4); break;
    case IZX: cyc = 6; break;
    case IZY: cyc = acc == 1 ? 6 : 5; break;
    case REL: cyc = 2; break;
    case IND: cyc = 5; break;
    }
    if (md == ABX || md == ABY || md == IZY) d0 |= (acc == 0 ? D_DUMMY_CROSS : D_DUMMY_ALWAYS);
    if (d0 & D_MEM) {
        if (acc != 1) d0 |= D_RD;
        if (acc != 0) d0 |= D_WR;
        if (acc == 2) d0 |= D_RMWMEM;
    }
    uint32_t cls = CLS_COMPUTE, kind = K_NOP;
    auto dst = [&](uint32_t r) { d1 = (d1 & ~(7u << D1_DST_SHIFT)) | (r << D1_DST_SHIFT); };
    auto src = [&](uint32_t r) { d1 = (d1 & ~(7u << D1_SRC_SHIFT)) | (r << D1_SRC_SHIFT); };
    switch (m) {
    case ORA: kind = K_LOGIC; d0 |= D_M_AND | D_M_XOR | D_SETNZ; dst(R_A); break;
    case AND: kind = K_LOGIC; d0 |= D_M_AND | D_SETNZ; dst(R_A); break;
    case EOR: kind = K_LOGIC; d0 |= D_M_XOR | D_SETNZ; dst(R_A); break;
    case LDA: kind = K_LOGIC; d0 |= D_M_LD | D_SETNZ; dst(R_A); break;
    case LDX: kind = K_LOGIC; d0 |= D_M_LD | D_SETNZ; dst(R_X); break;
    case LDY: kind = K_LOGIC; d0 |= D_M_LD | D_SETNZ; dst(R_Y); break;
    case ADC: kind = K_ARITH; d0 |= D_SETV | D_DEC | D_SETNZ | D_SETC; src(R_A); dst(R_A); break;
    case SBC: kind = K_ARITH; d0 |= D_INV | D_SETV | D_DEC | D_SETNZ | D_SETC; src(R_A); dst(R_A); break;
    case CMP: kind = K_ARITH; d0 |= D_INV | D_CIN1 | D_KEEP | D_SETNZ | D_SETC; src(R_A); break;
    case CPX: kind = K_ARITH; d0 |= D_INV | D_CIN1 | D_KEEP | D_SETNZ | D_SETC; src(R_X); break;
    case CPY: kind = K_ARITH; d0 |= D_INV | D_CIN1 | D_KEEP | D_SETNZ | D_SETC; src(R_Y); break;
    case BIT: kind = K_BIT; break;
    case STA: kind = K_STORE; src(R_A); break;
    case STX: kind = K_STORE; src(R_X); break;
    case STY: kind = K_STORE; src(R_Y); break;
    case ASL: case ROL: case LSR: case ROR: case DEC: case INC: {
        kind = K_RMW;
        uint32_t r = m == ASL ? 0 : m == ROL ? 1 : m == LSR ? 2 : m == ROR ? 3 : m == DEC ? 4 : 5;
        d1 |= r << D1_RMW_SHIFT;
        d0 |= D_SETNZ | (r < 4 ? D_SETC : 0u);
        if (md == ACC) { src(R_A); dst(R_A); }
        break;
    }
    case TAX: kind = K_MOVE; src(R_A); dst(R_X); d0 |= D_SETNZ; break;
    case TAY: kind = K_MOVE; src(R_A); dst(R_Y); d0 |= D_SETNZ; break;
    case TXA: kind = K_MOVE; src(R_X); dst(R_A); d0 |= D_SETNZ; break;
    case TYA: kind = K_MOVE; src(R_Y); dst(R_A); d0 |= D_SETNZ; break;
    case TSX: kind = K_MOVE; src(R_SP); dst(R_X); d0 |= D_SETNZ; break;
    case TXS: kind = K_MOVE; src(R_X); dst(R_SP); break;
    case INX: kind = K_MOVE; src(R_X); dst(R_X); d1 |= 1u << D1_DELTA_SHIFT; d0 |= D_SETNZ; break;
    case INY: kind = K_MOVE; src(R_Y); dst(R_Y); d1 |= 1u << D1_DELTA_SHIFT; d0 |= D_SETNZ; break;
    case DEX: kind = K_MOVE; src(R_X); dst(R_X); d1 |= 2u << D1_DELTA_SHIFT; d0 |= D_SETNZ; break;
    case DEY: kind = K_MOVE; src(R_Y); dst(R_Y); d1 |= 2u << D1_DELTA_SHIFT; d0 |= D_SETNZ; break;
    case CLC: kind = K_FLAGS; d1 |= kC << D1_FMASK_SHIFT; break;
    case SEC: kind = K_FLAGS; d1 |= (kC << D1_FMASK_SHIFT) | (kC << D1_FVAL_SHIFT); break;
    case CLI: kind = K_FLAGS; d1 |= kI << D1_FMASK_SHIFT; break;
    case SEI: kind = K_FLAGS; d1 |= (kI << D1_FMASK_SHIFT) | (kI << D1_FVAL_SHIFT); break;
    case CLV: kind = K_FLAGS; d1 |= kV << D1_FMASK_SHIFT; break;
    case CLD: kind = K_FLAGS; d1 |= kD << D1_FMASK_SHIFT; break;
    case SED: kind = K_FLAGS; d1 |= (kD << D1_FMASK_SHIFT) | (kD << D1_FVAL_SHIFT); break;
    case NOP: kind = K_NOP; break;
    case BPL: cls = CLS_BRANCH; d1 |= kN << D1_FMASK_SHIFT; break;
    case BMI: cls = CLS_BRANCH; d1 |= (kN << D1_FMASK_SHIFT) | (1u << D1_FVAL_SHIFT); break;
    case BVC: cls = CLS_BRANCH; d1 |= kV << D1_FMASK_SHIFT; break;
    case BVS: cls = CLS_BRANCH; d1 |= (kV << D1_FMASK_SHIFT) | (1u << D1_FVAL_SHIFT); break;
    case BCC: cls = CLS_BRANCH; d1 |= kC << D1_FMASK_SHIFT; break;
    case BCS: cls = CLS_BRANCH; d1 |= (kC << D1_FMASK_SHIFT) | (1u << D1_FVAL_SHIFT); break;
    case BNE: cls = CLS_BRANCH; d1 |= kZ << D1_FMASK_SHIFT; break;
    case BEQ: cls = CLS_BRANCH; d1 |= (kZ << D1_FMASK_SHIFT) | (1u << D1_FVAL_SHIFT); break;
    case JMP: cls = CLS_SPECIAL; d1 |= (md == IND ? SP_JMPI : SP_JMP) << D1_SPEC_SHIFT; cyc = md == IND ? 5 : 3; break;
    case JSR: cls = CLS_SPECIAL; d1 |= SP_JSR << D1_SPEC_SHIFT; cyc = 6; d0 &= ~(D_MEM | D_RD | D_WR); break;
    case RTS: cls = CLS_SPECIAL; d1 |= SP_RTS << D1_SPEC_SHIFT; cyc = 6; break;
    case RTI: cls = CLS_SPECIAL; d1 |= SP_RTI << D1_SPEC_SHIFT; cyc = 6; break;
    case BRK: cls = CLS_SPECIAL; d1 |= SP_BRK << D1_SPEC_SHIFT; cyc = 7; break;
    case PHA: cls = CLS_SPECIAL; d1 |= SP_PHA << D1_SPEC_SHIFT; cyc = 3; break;
    case PHP: cls = CLS_SPECIAL; d1 |= SP_PHP << D1_SPEC_SHIFT; cyc = 3; break;
    case PLA: cls = CLS_SPECIAL; d1 |= SP_PLA << D1_SPEC_SHIFT; cyc = 4; break;
    case PLP: cls = CLS_SPECIAL; d1 |= SP_PLP << D1_SPEC_SHIFT; cyc = 4; break;
    }
    if (cls != CLS_COMPUTE) d0 &= ~(D_MEM | D_RD | D_WR | D_RMWMEM | D_DUMMY_ALWAYS | D_DUMMY_CROSS);
    d0 |= (cls << D_CLS_SHIFT) | (len << D_LEN_SHIFT) | (cyc << D_CYC_SHIFT) | (kind << D_KIND_SHIFT);
}

}  // namespace detail

// Descriptor of one opcode, usable at compile time: {d0, d1} packed as
// d0 | d1 << 32 (CLS_ILLEGAL with length 1 for undocumented opcodes).
constexpr uint64_t descriptor_of(uint32_t code) {
    for (const detail::Op &o : detail::kOps) {
        if (o.code == code) {
            uint32_t d0 = 0, d1 = 0;
            detail::describe(o, d0, d1);
            return (uint64_t)d0 | ((uint64_t)d1 << 32);
        }
    }
    return (uint64_t)((CLS_ILLEGAL << D_CLS_SHIFT) | (1u << D_LEN_SHIFT));
}

// The 151 documented opcodes, for code generation over the instruction set.
#define FA_FOR_EACH_OPCODE(X)                                                                              \
    X(0x00) X(0x01) X(0x05) X(0x06) X(0x08) X(0x09) X(0x0A) X(0x0D) X(0x0E) X(0x10) X(0x11) X(0x15)       \
    X(0x16) X(0x18) X(0x19) X(0x1D) X(0x1E) X(0x20) X(0x21) X(0x24) X(0x25) X(0x26) X(0x28) X(0x29)       \
    X(0x2A) X(0x2C) X(0x2D) X(0x2E) X(0x30) X(0x31) X(0x35) X(0x36) X(0x38) X(0x39) X(0x3D) X(0x3E)       \
    X(0x40) X(0x41) X(0x45) X(0x46) X(0x48) X(0x49) X(0x4A) X(0x4C) X(0x4D) X(0x4E) X(0x50) X(0x51)       \
    X(0x55) X(0x56) X(0x58) X(0x59) X(0x5D) X(0x5E) X(0x60) X(0x61) X(0x65) X(0x66) X(0x68) X(0x69)       \
    X(0x6A) X(0x6C) X(0x6D) X(0x6E) X(0x70) X(0x71) X(0x75) X(0x76) X(0x78) X(0x79) X(0x7D) X(0x7E)       \
    X(0x81) X(0x84) X(0x85) X(0x86) X(0x88) X(0x8A) X(0x8C) X(0x8D) X(0x8E) X(0x90) X(0x91) X(0x94)       \
    X(0x95) X(0x96) X(0x98) X(0x99) X(0x9A) X(0x9D) X(0xA0) X(0xA1) X(0xA2) X(0xA4) X(0xA5) X(0xA6)       \
    X(0xA8) X(0xA9) X(0xAA) X(0xAC) X(0xAD) X(0xAE) X(0xB0) X(0xB1) X(0xB4) X(0xB5) X(0xB6) X(0xB8)       \
    X(0xB9) X(0xBA) X(0xBC) X(0xBD) X(0xBE) X(0xC0) X(0xC1) X(0xC4) X(0xC5) X(0xC6) X(0xC8) X(0xC9)       \
    X(0xCA) X(0xCC) X(0xCD) X(0xCE) X(0xD0) X(0xD1) X(0xD5) X(0xD6) X(0xD8) X(0xD9) X(0xDD) X(0xDE)       \
    X(0xE0) X(0xE1) X(0xE4) X(0xE5) X(0xE6) X(0xE8) X(0xE9) X(0xEA) X(0xEC) X(0xED) X(0xEE) X(0xF0)       \
    X(0xF1) X(0xF5) X(0xF6) X(0xF8) X(0xF9) X(0xFD) X(0xFE)

// Fills table[4 * op + k] = dk (k = 0..3) for all 256 opcodes; undocumented
// opcodes are CLS_ILLEGAL with length 1.
inline void build_descriptors(uint32_t *table) {
    for (int i = 0; i < 256; ++i) {
        uint32_t *t = table + kDescWords * i;
        t[0] = (CLS_ILLEGAL << D_CLS_SHIFT) | (1u << D_LEN_SHIFT);
        t[1] = t[2] = t[3] = 0;
        t[4] = t[5] = 0x0C0C0C0Cu;
        t[6] = 0x03020100u;
        t[7] = 0;
    }
    for (const detail::Op &o : detail::kOps) {
        uint32_t d0 = 0, d1 = 0, d2 = 0, d3 = 0;
        detail::describe(o, d0, d1);
        describe_lanes(d0, d1, d2, d3);
        uint32_t *t = table + kDescWords * o.code;
        t[0] = d0;
        t[1] = d1 | (((d0 & 3u) == CLS_SPECIAL && d3 == 0u) ? D1_SLOW : 0u);
        t[2] = d2;
        t[3] = d3;
        // d4 .. d7: v_perm selectors over the register word
        uint32_t ss = 0x0C0C0C0Cu, sd = 0x03020100u;
        for (uint32_t k = 0; k < 4u; ++k) {
            if (d2 & (1u << (G_SRC + k))) ss = 0x0C0C0C00u | k;
            if (d2 & (1u << (G_DST + k))) sd = (sd & ~(0xFFu << (8u * k))) | (4u << (8u * k));
        }
        const uint32_t ie = (d2 & (1u << G_IX_X)) ? 1u : ((d2 & (1u << G_IX_Y)) ? 2u : 0x0Cu);
        const uint32_t iz = (d2 & (1u << G_IZX)) ? 1u : 0x0Cu;
        t[4] = ss;
        t[5] = 0x0C0C0000u | (iz << 8) | ie;
        t[6] = sd;
        t[7] = (d3 & (1u << H_PULL)) ? 0x01000000u : ((d3 & (1u << H_PUSH)) ? 0xFF000000u : 0u);
    }
}

}  // namespace fa
