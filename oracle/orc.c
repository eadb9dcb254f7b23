/*
 * ORACLE -- test infrastructure only (see orc.h for scope and citations).
 * Plain C restatement of the reference CPU emulator's headless path.
 */
#define _POSIX_C_SOURCE 200809L
#include "orc.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------- flags */
/* src/cpu.rs:17-26 */
#define ORC_C 0x01
#define ORC_Z 0x02
#define ORC_I 0x04
#define ORC_D 0x08
#define ORC_B 0x10
#define ORC_U 0x20
#define ORC_V 0x40
#define ORC_N 0x80

static inline void orc_setf(orc_cpu *c, uint8_t f, int on)
{
    if (on) c->p |= f; else c->p &= (uint8_t)~f;
}
/* Status::set_zn, src/cpu.rs:48-51 */
static inline void orc_zn(orc_cpu *c, uint8_t v)
{
    orc_setf(c, ORC_Z, v == 0);
    orc_setf(c, ORC_N, (v & 0x80) != 0);
}

/* ---------------------------------------------------------------- ALU */
/* src/cpu.rs:152-175 */
static inline void orc_adc(orc_cpu *c, uint8_t val)
{
    uint8_t carry = (c->p & ORC_C) ? 1 : 0;
    if (c->p & ORC_D) {
        uint8_t lo = (uint8_t)((c->a & 0x0F) + (val & 0x0F) + carry);
        if (lo > 9) lo = (uint8_t)(lo + 6);
        uint8_t hi = (uint8_t)((c->a >> 4) + (val >> 4) + (lo > 0x0F ? 1 : 0));
        uint16_t bin = (uint16_t)(c->a + val + carry);
        orc_setf(c, ORC_Z, (uint8_t)bin == 0);
        orc_setf(c, ORC_N, (hi & 0x08) != 0);
        orc_setf(c, ORC_V, ((uint8_t)(~(c->a ^ val)) & (c->a ^ (uint8_t)((hi << 4) | (lo & 0x0F))) & 0x80) != 0);
        if (hi > 9) hi = (uint8_t)(hi + 6);
        orc_setf(c, ORC_C, hi > 0x0F);
        c->a = (uint8_t)(((hi & 0x0F) << 4) | (lo & 0x0F));
    } else {
        uint16_t s = (uint16_t)(c->a + val + carry);
        uint8_t s8 = (uint8_t)s;
        orc_setf(c, ORC_C, s > 0xFF);
        orc_setf(c, ORC_V, ((uint8_t)(~(c->a ^ val)) & (c->a ^ s8) & 0x80) != 0);
        c->a = s8;
        orc_zn(c, c->a);
    }
}

/* src/cpu.rs:177-200 */
static inline void orc_sbc(orc_cpu *c, uint8_t val)
{
    uint8_t borrow = (c->p & ORC_C) ? 0 : 1;
    if (c->p & ORC_D) {
        uint8_t lo = (uint8_t)((c->a & 0x0F) - (val & 0x0F) - borrow);
        uint8_t lo_borrow = 0;
        if ((int8_t)lo < 0) { lo = (uint8_t)(lo - 6); lo_borrow = 1; }
        uint8_t hi = (uint8_t)((c->a >> 4) - (val >> 4) - lo_borrow);
        if ((int8_t)hi < 0) hi = (uint8_t)(hi - 6);
        int16_t bin = (int16_t)((int16_t)c->a - (int16_t)val - (int16_t)borrow);
        orc_setf(c, ORC_C, bin >= 0);
        orc_setf(c, ORC_Z, (uint8_t)bin == 0);
        orc_setf(c, ORC_N, ((uint8_t)bin & 0x80) != 0);
        orc_setf(c, ORC_V, ((c->a ^ val) & (c->a ^ (uint8_t)bin) & 0x80) != 0);
        c->a = (uint8_t)(((hi & 0x0F) << 4) | (lo & 0x0F));
    } else {
        int16_t d = (int16_t)((int16_t)c->a - (int16_t)val - (int16_t)borrow);
        uint8_t d8 = (uint8_t)d;
        orc_setf(c, ORC_C, d >= 0);
        orc_setf(c, ORC_V, ((c->a ^ val) & (c->a ^ d8) & 0x80) != 0);
        c->a = d8;
        orc_zn(c, c->a);
    }
}

/* src/cpu.rs:203-207 */
static inline void orc_cmp(orc_cpu *c, uint8_t reg, uint8_t val)
{
    orc_setf(c, ORC_C, reg >= val);
    orc_zn(c, (uint8_t)(reg - val));
}

/* src/cpu.rs:210-241 */
static inline uint8_t orc_asl(orc_cpu *c, uint8_t v)
{
    orc_setf(c, ORC_C, (v & 0x80) != 0);
    uint8_t r = (uint8_t)(v << 1);
    orc_zn(c, r);
    return r;
}
static inline uint8_t orc_lsr(orc_cpu *c, uint8_t v)
{
    orc_setf(c, ORC_C, (v & 0x01) != 0);
    uint8_t r = (uint8_t)(v >> 1);
    orc_zn(c, r);
    return r;
}
static inline uint8_t orc_rol(orc_cpu *c, uint8_t v)
{
    uint8_t oc = (c->p & ORC_C) ? 1 : 0;
    orc_setf(c, ORC_C, (v & 0x80) != 0);
    uint8_t r = (uint8_t)((v << 1) | oc);
    orc_zn(c, r);
    return r;
}
static inline uint8_t orc_ror(orc_cpu *c, uint8_t v)
{
    uint8_t oc = (c->p & ORC_C) ? 1 : 0;
    orc_setf(c, ORC_C, (v & 0x01) != 0);
    uint8_t r = (uint8_t)((v >> 1) | (oc << 7));
    orc_zn(c, r);
    return r;
}

/* The "read" operations shared by every read addressing mode: the inner
 * `match opcode` arms of src/cpu.rs:342-355, 371-390, 427-436, 573-591,
 * 668-677, 758-767, 812-820, 869-877.  The operation is the opcode's
 * aaa/cc field (bits 7-5 and 1-0). */
static inline void orc_read_op(orc_cpu *c, uint8_t op, uint8_t v)
{
    if ((op & 0x03) == 0x01) {           /* ORA AND EOR ADC - LDA CMP SBC */
        switch (op >> 5) {
        case 0: c->a |= v; orc_zn(c, c->a); break;
        case 1: c->a &= v; orc_zn(c, c->a); break;
        case 2: c->a ^= v; orc_zn(c, c->a); break;
        case 3: orc_adc(c, v); break;
        case 5: c->a = v; orc_zn(c, c->a); break;
        case 6: orc_cmp(c, c->a, v); break;
        case 7: orc_sbc(c, v); break;
        default: break;                  /* STA has no read op */
        }
        return;
    }
    if ((op & 0x03) == 0x02) {           /* LDX */
        c->x = v; orc_zn(c, c->x);
        return;
    }
    switch (op >> 5) {                    /* cc = 00 */
    case 1:                               /* BIT: src/cpu.rs:383-388 */
        orc_setf(c, ORC_Z, (c->a & v) == 0);
        orc_setf(c, ORC_V, (v & 0x40) != 0);
        orc_setf(c, ORC_N, (v & 0x80) != 0);
        break;
    case 5: c->y = v; orc_zn(c, c->y); break;     /* LDY */
    case 6: orc_cmp(c, c->y, v); break;            /* CPY */
    case 7: orc_cmp(c, c->x, v); break;            /* CPX */
    default: break;
    }
}

/* STA / STX / STY source register (src/cpu.rs:402-407, 453-457, 605-610). */
static inline uint8_t orc_store_reg(const orc_cpu *c, uint8_t op)
{
    switch (op & 0x03) {
    case 0x01: return c->a;
    case 0x02: return c->x;
    default: return c->y;
    }
}

/* RMW operations (src/cpu.rs:510-518 and the three other RMW groups). */
static inline uint8_t orc_rmw(orc_cpu *c, uint8_t op, uint8_t v)
{
    uint8_t r;
    switch (op >> 5) {
    case 0: return orc_asl(c, v);
    case 1: return orc_rol(c, v);
    case 2: return orc_lsr(c, v);
    case 3: return orc_ror(c, v);
    case 6: r = (uint8_t)(v - 1); orc_zn(c, r); return r;
    default: r = (uint8_t)(v + 1); orc_zn(c, r); return r;   /* 7: INC */
    }
}

/* src/cpu.rs:912-922 */
static inline int orc_branch_taken(const orc_cpu *c, uint8_t op)
{
    switch (op) {
    case 0x90: return !(c->p & ORC_C);
    case 0xB0: return (c->p & ORC_C) != 0;
    case 0xF0: return (c->p & ORC_Z) != 0;
    case 0x30: return (c->p & ORC_N) != 0;
    case 0xD0: return !(c->p & ORC_Z);
    case 0x10: return !(c->p & ORC_N);
    case 0x50: return !(c->p & ORC_V);
    default: return (c->p & ORC_V) != 0;  /* 0x70 */
    }
}

/* ---------------------------------------------------------------- PIA */
/* Pia::default, src/pia.rs:46-62 */
void orc_pia_default(orc_pia *p)
{
    memset(p, 0, sizeof(*p));
    p->pa_in = 0xFF;
    p->pb_in = 0x3F;
    p->tps = 1;
    p->tsel = 1;
}

/* Pia::tick (src/pia.rs:80-103) -- not on the headless path; used by the
 * tests as an independent per-cycle statement of the timer. */
void orc_pia_tick(orc_pia *p)
{
    if (p->tval == 0 && p->tuf) { p->tval = 0xFF; return; }
    p->tps = (uint16_t)(p->tps - 1);
    if (p->tps == 0) {
        p->tps = p->tuf ? 1 : p->tsel;
        if (p->tval == 0) { p->tuf = 1; p->tval = 0xFF; }
        else p->tval = (uint16_t)(p->tval - 1);
    }
}

/* Pia::tick_n, src/pia.rs:107-166 */
void orc_pia_tick_n(orc_pia *p, uint16_t n)
{
    if (n == 0) return;
    if (p->tval == 0 && p->tuf) {
        p->tval = (uint16_t)((0xFFu - (uint32_t)(n - 1)) & 0xFF);
        return;
    }
    uint32_t rem = n;
    if (p->tps > 0) {
        if (rem < p->tps) { p->tps = (uint16_t)(p->tps - rem); return; }
        rem -= p->tps;
        p->tps = 0;
        if (p->tval == 0) {
            p->tuf = 1;
            p->tval = 0xFF;
            if (rem > 0) p->tval = (uint16_t)((0xFFu - (rem - 1)) & 0xFF);
            return;
        }
        p->tval = (uint16_t)(p->tval - 1);
    }
    if (rem == 0) { p->tps = p->tsel; return; }
    uint32_t ps = p->tsel;
    uint32_t full = rem / ps;
    uint32_t left = rem % ps;
    if (full >= p->tval) {
        rem -= (uint32_t)p->tval * ps;
        p->tuf = 1;
        p->tval = 0xFF;
        if (rem > 0) p->tval = (uint16_t)((0xFFu - (rem - 1)) & 0xFF);
    } else {
        p->tval = (uint16_t)(p->tval - full);
        p->tps = (uint16_t)(ps - left);
    }
}

/* Pia::read, src/pia.rs:169-199 */
uint8_t orc_pia_read(orc_pia *p, uint16_t addr)
{
    if ((addr & 0x200) == 0) return p->ram[addr & 0x7F];
    switch (addr & 0x07) {
    case 0: return (uint8_t)((p->pa_in & ~p->pa_ddr) | (p->pa_out & p->pa_ddr));
    case 1: return p->pa_ddr;
    case 2: return (uint8_t)((p->pb_in & ~p->pb_ddr) | (p->pb_out & p->pb_ddr));
    case 3: return p->pb_ddr;
    case 4: p->tuf = 0; return (uint8_t)p->tval;
    case 5: return p->tuf ? 0x80 : 0x00;
    default: return 0;
    }
}

/* Pia::write, src/pia.rs:202-228 */
void orc_pia_write(orc_pia *p, uint16_t addr, uint8_t val)
{
    if ((addr & 0x200) == 0) { p->ram[addr & 0x7F] = val; return; }
    if (addr & 0x10) {
        static const uint16_t sel[4] = {1, 8, 64, 1024};
        p->tval = val;
        p->tuf = 0;
        p->tsel = sel[addr & 3];
        p->tps = p->tsel;
        return;
    }
    switch (addr & 0x03) {
    case 0: p->pa_out = val; break;
    case 1: p->pa_ddr = val; break;
    case 2: p->pb_out = val; break;
    default: p->pb_ddr = val; break;
    }
}

/* ---------------------------------------------------------------- TIA */
#define ORC_SCANLINES 262   /* tia::SCANLINES_PER_FRAME, src/tia.rs:70 */

/* HeadlessTia::default, src/headless_tia.rs:90-118 */
void orc_tia_default(orc_tia *t)
{
    memset(t, 0, sizeof(*t));
    t->inpt4 = 1;
    t->inpt5 = 1;
}

/* src/headless_tia.rs:183-198 */
static inline void orc_tia_end_scanline(orc_tia *t)
{
    t->clock = (uint16_t)(t->clock - 228);
    t->scanline = (uint16_t)(t->scanline + 1);
    t->wsync = 0;
    if (!t->paddle_dumped && t->paddle_counter < 256) t->paddle_counter++;
    if (t->scanline >= ORC_SCANLINES) {
        t->scanline = 0;
        t->frame_complete = 1;
    }
}

/* src/headless_tia.rs:153-158 */
void orc_tia_tick_n(orc_tia *t, uint8_t n)
{
    t->clock = (uint16_t)(t->clock + (uint16_t)n * 3);
    if (t->clock >= 228) orc_tia_end_scanline(t);
}

/* src/headless_tia.rs:163-180 */
uint16_t orc_tia_skip_to_scanline_end(orc_tia *t)
{
    uint16_t rem = t->clock >= 228 ? 0 : (uint16_t)(228 - t->clock);
    uint16_t cyc = (uint16_t)((rem + 2) / 3);
    if (cyc > 0) {
        t->clock = 228;
        orc_tia_end_scanline(t);
    }
    return cyc;
}

/* src/headless_tia.rs:200-203 */
static inline uint16_t orc_hmove(uint16_t pos, uint8_t hm)
{
    int off = ((int8_t)hm) >> 4;
    int r = ((int)pos - off) % 160;
    if (r < 0) r += 160;
    return (uint16_t)r;
}

static inline uint16_t orc_resp(uint16_t clock)
{
    int r = ((int)(int16_t)clock - 68 + 5) % 160;
    if (r < 0) r += 160;
    return (uint16_t)r;
}

/* src/headless_tia.rs:205-223 */
uint8_t orc_tia_read(const orc_tia *t, uint16_t addr)
{
    switch (addr & 0x0F) {
    case 0x0: return t->cxm0p;
    case 0x1: return t->cxm1p;
    case 0x2: return t->cxp0fb;
    case 0x3: return t->cxp1fb;
    case 0x4: return t->cxm0fb;
    case 0x5: return t->cxm1fb;
    case 0x6: return t->cxblpf;
    case 0x7: return t->cxppmm;
    case 0x8: return t->paddle_counter >= t->paddle0 ? 0x80 : 0x00;
    case 0x9: return t->paddle_counter >= t->paddle1 ? 0x80 : 0x00;
    case 0xA: return 0x80;
    case 0xB: return 0x80;
    case 0xC: return t->inpt4 ? 0x80 : 0x00;
    case 0xD: return t->inpt5 ? 0x80 : 0x00;
    default: return 0;
    }
}

/* src/headless_tia.rs:225-298 */
void orc_tia_write(orc_tia *t, uint16_t addr, uint8_t v)
{
    switch (addr & 0x3F) {
    case 0x00:
        if ((t->vsync & 0x02) == 0 && (v & 0x02) != 0) {
            t->scanline = 0;
            t->frame_complete = 1;
        }
        t->vsync = v;
        break;
    case 0x01:
        if (v & 0x80) { t->paddle_counter = 0; t->paddle_dumped = 1; }
        else if (t->paddle_dumped) t->paddle_dumped = 0;
        t->vblank = v;
        break;
    case 0x02: t->wsync = 1; break;
    case 0x04: t->nusiz0 = v; break;
    case 0x05: t->nusiz1 = v; break;
    case 0x06: t->colup0 = v; break;
    case 0x07: t->colup1 = v; break;
    case 0x08: t->colupf = v; break;
    case 0x09: t->colubk = v; break;
    case 0x0A: t->ctrlpf = v; break;
    case 0x0B: t->refp0 = (v & 0x08) != 0; break;
    case 0x0C: t->refp1 = (v & 0x08) != 0; break;
    case 0x0D: t->pf0 = v; break;
    case 0x0E: t->pf1 = v; break;
    case 0x0F: t->pf2 = v; break;
    case 0x10: t->pos_p0 = orc_resp(t->clock); break;
    case 0x11: t->pos_p1 = orc_resp(t->clock); break;
    case 0x12: t->pos_m0 = orc_resp(t->clock); break;
    case 0x13: t->pos_m1 = orc_resp(t->clock); break;
    case 0x14: t->pos_bl = orc_resp(t->clock); break;
    case 0x1B: t->grp0_old = t->grp0; t->grp0 = v; break;
    case 0x1C: t->grp1_old = t->grp1; t->grp1 = v; t->enabl_old = t->enabl; break;
    case 0x1D: t->enam0 = (v & 0x02) != 0; break;
    case 0x1E: t->enam1 = (v & 0x02) != 0; break;
    case 0x1F: t->enabl = (v & 0x02) != 0; break;
    case 0x20: t->hmp0 = v; break;
    case 0x21: t->hmp1 = v; break;
    case 0x22: t->hmm0 = v; break;
    case 0x23: t->hmm1 = v; break;
    case 0x24: t->hmbl = v; break;
    case 0x25: t->vdelp0 = (v & 0x01) != 0; break;
    case 0x26: t->vdelp1 = (v & 0x01) != 0; break;
    case 0x27: t->vdelbl = (v & 0x01) != 0; break;
    case 0x28: t->resmp0 = (v & 0x02) != 0; break;
    case 0x29: t->resmp1 = (v & 0x02) != 0; break;
    case 0x2A:
        t->pos_p0 = orc_hmove(t->pos_p0, t->hmp0);
        t->pos_p1 = orc_hmove(t->pos_p1, t->hmp1);
        t->pos_m0 = orc_hmove(t->pos_m0, t->hmm0);
        t->pos_m1 = orc_hmove(t->pos_m1, t->hmm1);
        t->pos_bl = orc_hmove(t->pos_bl, t->hmbl);
        break;
    case 0x2B: t->hmp0 = t->hmp1 = t->hmm0 = t->hmm1 = t->hmbl = 0; break;
    case 0x2C:
        t->cxm0p = t->cxm1p = t->cxp0fb = t->cxp1fb = 0;
        t->cxm0fb = t->cxm1fb = t->cxblpf = t->cxppmm = 0;
        break;
    default: break;   /* 0x03, audio 0x15-0x1A, 0x2D-0x3F */
    }
}

/* ---------------------------------------------------------------- bus */
/* BankScheme::detect, src/atari.rs:97-104 */
int orc_scheme_detect(uint32_t n)
{
    if (n <= 4096) return ORC_FIXED;
    if (n <= 8192) return ORC_F8;
    if (n <= 16384) return ORC_F6;
    return ORC_F4;
}

/* HeadlessBus::load_bank, src/atari.rs:433-449 */
static void orc_load_bank(orc_bus *b, uint32_t bank)
{
    if (b->scheme == ORC_FIXED) {
        for (uint32_t i = 0; i < 4096; i++) b->mem[0x1000 + i] = b->rom[i % b->rom_len];
        return;
    }
    memcpy(&b->mem[0x1000], &b->rom[bank * 4096], 4096);
}

/* HeadlessBus::sync_ram_to_mem, src/atari.rs:454-457 */
static void orc_sync_ram(orc_bus *b)
{
    memcpy(&b->mem[0x80], b->pia.ram, 128);
    memcpy(&b->mem[0x180], b->pia.ram, 128);
}

/* HeadlessBus::check_bankswitch, src/atari.rs:460-492 */
static inline void orc_check_bankswitch(orc_bus *b, uint16_t addr)
{
    uint16_t a = addr & 0x1FFF;
    int nb;
    switch (b->scheme) {
    case ORC_F8: if (a < 0x1FF8 || a > 0x1FF9) return; nb = a - 0x1FF8; break;
    case ORC_F6: if (a < 0x1FF6 || a > 0x1FF9) return; nb = a - 0x1FF6; break;
    case ORC_F4: if (a < 0x1FF4 || a > 0x1FFB) return; nb = a - 0x1FF4; break;
    default: return;
    }
    if ((uint32_t)nb != b->bank) {
        b->bank = (uint32_t)nb;
        orc_load_bank(b, (uint32_t)nb);
    }
}

/* HeadlessBus::read, src/atari.rs:497-520 */
uint8_t orc_bus_read(orc_bus *b, uint16_t addr)
{
    uint16_t a = addr & 0x1FFF;
    if ((a & 0x1080) != 0 && (a & 0x1280) != 0x0280) {
        if ((a & 0x1000) && b->scheme != ORC_FIXED) orc_check_bankswitch(b, addr);
        return b->mem[a];
    }
    if ((a & 0x1280) == 0x0280) return orc_pia_read(&b->pia, a);
    if ((a & 0x1080) == 0) return orc_tia_read(&b->tia, a);
    return 0;
}

/* HeadlessBus::write, src/atari.rs:523-552 */
void orc_bus_write(orc_bus *b, uint16_t addr, uint8_t val)
{
    uint16_t a = addr & 0x1FFF;
    if (a & 0x1000) {
        if (b->scheme != ORC_FIXED) orc_check_bankswitch(b, addr);
        return;
    }
    if ((a & 0x80) && !(a & 0x200)) {
        b->pia.ram[a & 0x7F] = val;
        b->mem[(a | 0x80) & 0x1FF] = val;
        b->mem[((a | 0x80) & 0x1FF) ^ 0x100] = val;
        return;
    }
    if (a & 0x200) { orc_pia_write(&b->pia, a, val); return; }
    orc_tia_write(&b->tia, a, val);
}

/* ---------------------------------------------------------------- CPU */
#define STEP_FN orc_step_headless
#define MEMT orc_bus
#define RD(a) orc_bus_read(m, (uint16_t)(a))
#define WR(a, v) orc_bus_write(m, (uint16_t)(a), (uint8_t)(v))
#include "orc_cpu_step.inc"
#undef STEP_FN
#undef MEMT
#undef RD
#undef WR

static inline uint8_t orc_flat_rd(orc_flat *f, uint16_t a)
{
    uint8_t v = f->ram[a];
    if (f->ntrace < 64) f->trace[f->ntrace++] = ((uint32_t)a << 8) | v;
    return v;
}
static inline void orc_flat_wr(orc_flat *f, uint16_t a, uint8_t v)
{
    f->ram[a] = v;
    if (f->ntrace < 64) f->trace[f->ntrace++] = (1u << 24) | ((uint32_t)a << 8) | v;
}
#define STEP_FN orc_step_flat
#define MEMT orc_flat
#define RD(a) orc_flat_rd(m, (uint16_t)(a))
#define WR(a, v) orc_flat_wr(m, (uint16_t)(a), (uint8_t)(v))
#include "orc_cpu_step.inc"
#undef STEP_FN
#undef MEMT
#undef RD
#undef WR

int orc_cpu_step_bus(orc_cpu *c, orc_bus *b) { return orc_step_headless(c, b); }

/* ---------------------------------------------------------------- console */
/* HeadlessBus::new + HeadlessAtari::new + Cpu::reset
 * (src/atari.rs:410-430, 562-569; src/cpu.rs:107-134) */
void orc_atari_init(orc_atari *e, const uint8_t *rom, uint32_t rom_len)
{
    memset(e, 0, sizeof(*e));
    orc_bus *b = &e->bus;
    orc_tia_default(&b->tia);
    orc_pia_default(&b->pia);
    b->rom = rom;
    b->rom_len = rom_len;
    b->scheme = orc_scheme_detect(rom_len);
    static const uint32_t first_bank[4] = {0, 1, 3, 7};
    b->bank = first_bank[b->scheme];
    orc_load_bank(b, b->bank);
    orc_sync_ram(b);
    e->cpu.sp = 0xFD;
    uint8_t lo = orc_bus_read(b, 0xFFFC);
    uint8_t hi = orc_bus_read(b, 0xFFFD);
    e->cpu.pc = (uint16_t)((hi << 8) | lo);
    e->cpu.sp = 0xFD;
    e->cpu.p = 0;
    e->cpu.a = e->cpu.x = e->cpu.y = 0;
}

/* HeadlessAtari::set_action + Action::{from_index, decode}
 * (src/atari.rs:34-80, 571-582).  Direction bits per action index:
 * up, down, left, right, fire. */
void orc_atari_set_action(orc_atari *e, uint32_t action)
{
    /* (up, down, left, right, fire) for indices 0..17, src/atari.rs:60-79 */
    static const uint8_t dec[18][5] = {
        {0,0,0,0,0}, {0,0,0,0,1}, {1,0,0,0,0}, {0,0,0,1,0}, {0,0,1,0,0}, {0,1,0,0,0},
        {1,0,0,1,0}, {1,0,1,0,0}, {0,1,0,1,0}, {0,1,1,0,0}, {1,0,0,0,1}, {0,0,0,1,1},
        {0,0,1,0,1}, {0,1,0,0,1}, {1,0,0,1,1}, {1,0,1,0,1}, {0,1,0,1,1}, {0,1,1,0,1},
    };
    const uint8_t *d = dec[action < 18 ? action : 0];
    uint8_t swcha = 0xFF;
    if (d[0]) swcha &= (uint8_t)~0x10;
    if (d[1]) swcha &= (uint8_t)~0x20;
    if (d[2]) swcha &= (uint8_t)~0x40;
    if (d[3]) swcha &= (uint8_t)~0x80;
    e->bus.pia.pa_in = swcha;
    e->bus.tia.inpt4 = !d[4];
    e->bus.tia.paddle0 = d[2] ? 200 : (d[3] ? 50 : 128);
}

/* HeadlessAtari::run_one_cycle, src/atari.rs:600-614.  Returns 0, or -1 on
 * an illegal opcode. */
static inline int orc_run_one(orc_atari *e, uint64_t *cycles)
{
    orc_bus *b = &e->bus;
    if (b->tia.wsync) {
        uint16_t sk = orc_tia_skip_to_scanline_end(&b->tia);
        orc_pia_tick_n(&b->pia, sk);
        *cycles += sk;
    } else {
        int c = orc_step_headless(&e->cpu, b);
        if (c < 0) return -1;
        e->instructions++;
        *cycles += (uint64_t)c;
        orc_tia_tick_n(&b->tia, (uint8_t)c);
        orc_pia_tick_n(&b->pia, (uint16_t)c);
    }
    return 0;
}

/* HeadlessAtari::run_frame, src/atari.rs:584-598 */
uint64_t orc_atari_run_frame(orc_atari *e)
{
    if (e->error) { e->last_frame_cycles = 0; return 0; }
    orc_bus *b = &e->bus;
    b->tia.frame_complete = 0;
    uint64_t cycles = 0;
    while (b->tia.vsync & 0x02) {
        if (orc_run_one(e, &cycles)) { e->error = ORC_ERR_ILLEGAL; goto out; }
        if (cycles > ORC_LOOPA_CAP) { e->error = ORC_ERR_VSYNC_HELD; goto out; }
    }
    while (!(b->tia.vsync & 0x02)) {
        if (orc_run_one(e, &cycles)) { e->error = ORC_ERR_ILLEGAL; goto out; }
        if (cycles > 100000) break;
    }
out:
    e->last_frame_cycles = cycles;
    return cycles;
}

static inline void put16(uint8_t *o, int off, uint16_t v)
{
    o[off] = (uint8_t)v;
    o[off + 1] = (uint8_t)(v >> 8);
}
static inline uint16_t get16(const uint8_t *o, int off)
{
    return (uint16_t)(o[off] | (o[off + 1] << 8));
}

/* AtariStateGpu::from_headless, src/cuda_env.rs:43-94, written with the
 * byte offsets of the #[repr(C)] struct at src/cuda_env.rs:12-38. */
void orc_atari_pack(const orc_atari *e, uint8_t *o)
{
    const orc_cpu *c = &e->cpu;
    const orc_tia *t = &e->bus.tia;
    const orc_pia *p = &e->bus.pia;
    memset(o, 0, ORC_STATE_BYTES);
    o[0] = c->a; o[1] = c->x; o[2] = c->y; o[3] = c->sp;
    put16(o, 4, c->pc); o[6] = c->p;
    memcpy(o + 8, p->ram, 128);
    o[136] = p->pa_out; o[137] = p->pa_ddr; o[138] = p->pa_in;
    o[139] = p->pb_out; o[140] = p->pb_ddr; o[141] = p->pb_in;
    put16(o, 142, p->tval); put16(o, 144, p->tps); put16(o, 146, p->tsel);
    o[148] = p->tuf;
    put16(o, 152, t->clock); put16(o, 154, t->scanline);
    o[156] = t->wsync; o[157] = t->frame_complete; o[158] = t->vsync; o[159] = t->vblank;
    o[160] = t->cxm0p; o[161] = t->cxm1p; o[162] = t->cxp0fb; o[163] = t->cxp1fb;
    o[164] = t->cxm0fb; o[165] = t->cxm1fb; o[166] = t->cxblpf; o[167] = t->cxppmm;
    o[168] = t->inpt4; o[169] = t->inpt5; o[170] = t->paddle0; o[171] = t->paddle1;
    put16(o, 172, t->paddle_counter); o[174] = t->paddle_dumped;
    put16(o, 176, t->pos_p0); put16(o, 178, t->pos_p1); put16(o, 180, t->pos_m0);
    put16(o, 182, t->pos_m1); put16(o, 184, t->pos_bl);
    o[186] = t->hmp0; o[187] = t->hmp1; o[188] = t->hmm0; o[189] = t->hmm1; o[190] = t->hmbl;
    o[192] = t->grp0; o[193] = t->grp1; o[194] = t->grp0_old; o[195] = t->grp1_old;
    o[196] = t->enam0; o[197] = t->enam1; o[198] = t->enabl; o[199] = t->enabl_old;
    o[200] = t->vdelp0; o[201] = t->vdelp1; o[202] = t->vdelbl;
    o[203] = t->resmp0; o[204] = t->resmp1; o[205] = t->nusiz0; o[206] = t->nusiz1;
    o[207] = t->colup0; o[208] = t->colup1; o[209] = t->colupf; o[210] = t->colubk;
    o[211] = t->ctrlpf; o[212] = t->refp0; o[213] = t->refp1;
    o[214] = t->pf0; o[215] = t->pf1; o[216] = t->pf2;
    o[217] = (uint8_t)e->bus.bank; o[218] = (uint8_t)e->bus.scheme;
}

/* Inverse of orc_atari_pack; rebuilds the flat array (RAM mirrors and the
 * bank window) the way HeadlessBus keeps it. */
void orc_atari_unpack(orc_atari *e, const uint8_t *o)
{
    orc_cpu *c = &e->cpu;
    orc_tia *t = &e->bus.tia;
    orc_pia *p = &e->bus.pia;
    c->a = o[0]; c->x = o[1]; c->y = o[2]; c->sp = o[3];
    c->pc = get16(o, 4); c->p = o[6];
    memcpy(p->ram, o + 8, 128);
    p->pa_out = o[136]; p->pa_ddr = o[137]; p->pa_in = o[138];
    p->pb_out = o[139]; p->pb_ddr = o[140]; p->pb_in = o[141];
    p->tval = get16(o, 142); p->tps = get16(o, 144); p->tsel = get16(o, 146);
    p->tuf = o[148];
    t->clock = get16(o, 152); t->scanline = get16(o, 154);
    t->wsync = o[156]; t->frame_complete = o[157]; t->vsync = o[158]; t->vblank = o[159];
    t->cxm0p = o[160]; t->cxm1p = o[161]; t->cxp0fb = o[162]; t->cxp1fb = o[163];
    t->cxm0fb = o[164]; t->cxm1fb = o[165]; t->cxblpf = o[166]; t->cxppmm = o[167];
    t->inpt4 = o[168]; t->inpt5 = o[169]; t->paddle0 = o[170]; t->paddle1 = o[171];
    t->paddle_counter = get16(o, 172); t->paddle_dumped = o[174];
    t->pos_p0 = get16(o, 176); t->pos_p1 = get16(o, 178); t->pos_m0 = get16(o, 180);
    t->pos_m1 = get16(o, 182); t->pos_bl = get16(o, 184);
    t->hmp0 = o[186]; t->hmp1 = o[187]; t->hmm0 = o[188]; t->hmm1 = o[189]; t->hmbl = o[190];
    t->grp0 = o[192]; t->grp1 = o[193]; t->grp0_old = o[194]; t->grp1_old = o[195];
    t->enam0 = o[196]; t->enam1 = o[197]; t->enabl = o[198]; t->enabl_old = o[199];
    t->vdelp0 = o[200]; t->vdelp1 = o[201]; t->vdelbl = o[202];
    t->resmp0 = o[203]; t->resmp1 = o[204]; t->nusiz0 = o[205]; t->nusiz1 = o[206];
    t->colup0 = o[207]; t->colup1 = o[208]; t->colupf = o[209]; t->colubk = o[210];
    t->ctrlpf = o[211]; t->refp0 = o[212]; t->refp1 = o[213];
    t->pf0 = o[214]; t->pf1 = o[215]; t->pf2 = o[216];
    e->bus.bank = o[217];
    orc_load_bank(&e->bus, e->bus.bank);
    orc_sync_ram(&e->bus);
}

/* ---------------------------------------------------------------- handle API */
typedef struct {
    orc_atari e;
    uint8_t *rom;
} orc_handle;

void *orc_new(const uint8_t *rom, uint32_t rom_len)
{
    orc_handle *h = (orc_handle *)calloc(1, sizeof(orc_handle));
    h->rom = (uint8_t *)malloc(rom_len);
    memcpy(h->rom, rom, rom_len);
    orc_atari_init(&h->e, h->rom, rom_len);
    return h;
}
void orc_free(void *hp)
{
    orc_handle *h = (orc_handle *)hp;
    if (!h) return;
    free(h->rom);
    free(h);
}
void orc_reset_frame(void *hp)
{
    orc_handle *h = (orc_handle *)hp;
    orc_atari_init(&h->e, h->rom, h->e.bus.rom_len);
    orc_atari_run_frame(&h->e);
}
uint64_t orc_step(void *hp, uint32_t action, uint32_t frames)
{
    orc_handle *h = (orc_handle *)hp;
    orc_atari_set_action(&h->e, action);
    uint64_t tot = 0;
    for (uint32_t f = 0; f < frames; f++) tot += orc_atari_run_frame(&h->e);
    return tot;
}
void orc_get_state(void *hp, uint8_t *out) { orc_atari_pack(&((orc_handle *)hp)->e, out); }
void orc_set_state(void *hp, const uint8_t *in) { orc_atari_unpack(&((orc_handle *)hp)->e, in); }
void orc_get_counters(void *hp, uint64_t out[4])
{
    orc_atari *e = &((orc_handle *)hp)->e;
    out[0] = e->cpu.cycles;
    out[1] = e->last_frame_cycles;
    out[2] = e->instructions;
    out[3] = (uint64_t)e->error;
}
void orc_set_cycles(void *hp, uint64_t cyc) { ((orc_handle *)hp)->e.cpu.cycles = cyc; }
void orc_get_ram(void *hp, uint8_t out[128]) { memcpy(out, &((orc_handle *)hp)->e.bus.mem[0x80], 128); }
void orc_get_mem(void *hp, uint8_t out[8192]) { memcpy(out, ((orc_handle *)hp)->e.bus.mem, 8192); }
uint8_t orc_h_bus_read(void *hp, uint16_t a) { return orc_bus_read(&((orc_handle *)hp)->e.bus, a); }
void orc_h_bus_write(void *hp, uint16_t a, uint8_t v) { orc_bus_write(&((orc_handle *)hp)->e.bus, a, v); }

/* ---------------------------------------------------------------- batch API */
typedef struct {
    uint8_t *rom;
    uint32_t rom_len;
    uint32_t n;
    orc_atari *envs;
} orc_batch;

typedef struct {
    orc_batch *b;
    uint32_t lo, hi;
    int kind; /* 0 reset, 1 run, 2 run_fs */
    const uint8_t *actions;
    const uint8_t *fs;
    uint32_t frames;
} orc_job;

static void *orc_worker(void *arg)
{
    orc_job *j = (orc_job *)arg;
    orc_batch *b = j->b;
    for (uint32_t i = j->lo; i < j->hi; i++) {
        orc_atari *e = &b->envs[i];
        if (j->kind == 0) {
            orc_atari_init(e, b->rom, b->rom_len);
            orc_atari_run_frame(e);
        } else if (j->kind == 1) {
            for (uint32_t f = 0; f < j->frames; f++) {
                orc_atari_set_action(e, j->actions[(size_t)f * b->n + i]);
                orc_atari_run_frame(e);
            }
        } else {
            orc_atari_set_action(e, j->actions[i]);
            for (uint32_t f = 0; f < j->fs[i]; f++) orc_atari_run_frame(e);
        }
    }
    return NULL;
}

static void orc_parallel(orc_batch *b, int kind, const uint8_t *actions, const uint8_t *fs,
                         uint32_t frames, uint32_t nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > b->n) nthreads = b->n ? b->n : 1;
    pthread_t th[256];
    orc_job jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (uint32_t t = 0; t < nthreads; t++) {
        jobs[t].b = b;
        jobs[t].lo = (uint32_t)((uint64_t)b->n * t / nthreads);
        jobs[t].hi = (uint32_t)((uint64_t)b->n * (t + 1) / nthreads);
        jobs[t].kind = kind;
        jobs[t].actions = actions;
        jobs[t].fs = fs;
        jobs[t].frames = frames;
    }
    if (nthreads == 1) { orc_worker(&jobs[0]); return; }
    for (uint32_t t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, orc_worker, &jobs[t]);
    for (uint32_t t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

void *orc_batch_new(const uint8_t *rom, uint32_t rom_len, uint32_t n)
{
    orc_batch *b = (orc_batch *)calloc(1, sizeof(orc_batch));
    b->rom = (uint8_t *)malloc(rom_len);
    memcpy(b->rom, rom, rom_len);
    b->rom_len = rom_len;
    b->n = n;
    b->envs = (orc_atari *)calloc(n ? n : 1, sizeof(orc_atari));
    for (uint32_t i = 0; i < n; i++) orc_atari_init(&b->envs[i], b->rom, rom_len);
    return b;
}
void orc_batch_free(void *bp)
{
    orc_batch *b = (orc_batch *)bp;
    if (!b) return;
    free(b->envs);
    free(b->rom);
    free(b);
}
void orc_batch_reset(void *bp, uint32_t nthreads) { orc_parallel((orc_batch *)bp, 0, NULL, NULL, 0, nthreads); }
void orc_batch_run(void *bp, const uint8_t *actions, uint32_t frames, uint32_t nthreads)
{
    orc_parallel((orc_batch *)bp, 1, actions, NULL, frames, nthreads);
}
void orc_batch_run_fs(void *bp, const uint8_t *actions, const uint8_t *fs, uint32_t nthreads)
{
    orc_parallel((orc_batch *)bp, 2, actions, fs, 0, nthreads);
}
void orc_batch_states(void *bp, uint8_t *out)
{
    orc_batch *b = (orc_batch *)bp;
    for (uint32_t i = 0; i < b->n; i++) orc_atari_pack(&b->envs[i], out + (size_t)i * ORC_STATE_BYTES);
}
void orc_batch_set_states(void *bp, const uint8_t *in)
{
    orc_batch *b = (orc_batch *)bp;
    for (uint32_t i = 0; i < b->n; i++) orc_atari_unpack(&b->envs[i], in + (size_t)i * ORC_STATE_BYTES);
}
void orc_batch_counters(void *bp, uint64_t *out)
{
    orc_batch *b = (orc_batch *)bp;
    for (uint32_t i = 0; i < b->n; i++) {
        out[4 * i + 0] = b->envs[i].cpu.cycles;
        out[4 * i + 1] = b->envs[i].last_frame_cycles;
        out[4 * i + 2] = b->envs[i].instructions;
        out[4 * i + 3] = (uint64_t)b->envs[i].error;
    }
}
/* Restores every counter (the inverse of orc_batch_counters). */
void orc_batch_set_counters(void *bp, const uint64_t *in)
{
    orc_batch *b = (orc_batch *)bp;
    for (uint32_t i = 0; i < b->n; i++) {
        b->envs[i].cpu.cycles = in[4 * i + 0];
        b->envs[i].last_frame_cycles = in[4 * i + 1];
        b->envs[i].instructions = in[4 * i + 2];
        b->envs[i].error = (int)in[4 * i + 3];
    }
}
void orc_batch_set_cycles(void *bp, const uint64_t *cyc)
{
    orc_batch *b = (orc_batch *)bp;
    for (uint32_t i = 0; i < b->n; i++) b->envs[i].cpu.cycles = cyc[i];
}
void orc_batch_obs(void *bp, uint8_t *out)
{
    orc_batch *b = (orc_batch *)bp;
    for (uint32_t i = 0; i < b->n; i++) memcpy(out + (size_t)i * 128, &b->envs[i].bus.mem[0x80], 128);
}
void *orc_batch_env(void *bp, uint32_t i) { return &((orc_batch *)bp)->envs[i]; }

double orc_batch_bench(void *bp, const uint8_t *actions, uint32_t frames, uint32_t nthreads)
{
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    orc_batch_run(bp, actions, frames, nthreads);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ---------------------------------------------------------------- FlatMemory */
void *orc_flat_new(void) { return calloc(1, sizeof(orc_flat)); }
void orc_flat_free(void *f) { free(f); }
void orc_flat_load(void *fp, uint16_t base, const uint8_t *d, uint32_t len)
{
    orc_flat *f = (orc_flat *)fp;
    for (uint32_t i = 0; i < len; i++) f->ram[(uint16_t)(base + i)] = d[i];
}
void orc_flat_poke(void *fp, uint16_t a, uint8_t v) { ((orc_flat *)fp)->ram[a] = v; }
uint8_t orc_flat_peek(void *fp, uint16_t a) { return ((orc_flat *)fp)->ram[a]; }

static void regs_from(orc_cpu *c, const uint32_t r[6])
{
    c->pc = (uint16_t)r[0]; c->sp = (uint8_t)r[1]; c->a = (uint8_t)r[2];
    c->x = (uint8_t)r[3]; c->y = (uint8_t)r[4]; c->p = (uint8_t)r[5];
}
static void regs_to(const orc_cpu *c, uint32_t r[6])
{
    r[0] = c->pc; r[1] = c->sp; r[2] = c->a; r[3] = c->x; r[4] = c->y; r[5] = c->p;
}

/* Cpu::new + Cpu::reset on FlatMemory (src/cpu.rs:107-134) */
void orc_flat_reset(void *fp, uint32_t regs[6])
{
    orc_flat *f = (orc_flat *)fp;
    orc_cpu c;
    memset(&c, 0, sizeof(c));
    c.pc = (uint16_t)(f->ram[0xFFFC] | (f->ram[0xFFFD] << 8));
    c.sp = 0xFD;
    regs_to(&c, regs);
}

int orc_flat_step(void *fp, uint32_t regs[6], uint32_t *trace_out, uint32_t *ntrace)
{
    orc_flat *f = (orc_flat *)fp;
    orc_cpu c;
    memset(&c, 0, sizeof(c));
    regs_from(&c, regs);
    f->ntrace = 0;
    int cy = orc_step_flat(&c, f);
    regs_to(&c, regs);
    if (trace_out) memcpy(trace_out, f->trace, f->ntrace * sizeof(uint32_t));
    if (ntrace) *ntrace = f->ntrace;
    return cy;
}

void orc_alu_sweep(uint8_t op, uint16_t *out)
{
    orc_flat *f = (orc_flat *)calloc(1, sizeof(orc_flat));
    for (uint32_t i = 0; i < 262144; i++) {
        uint32_t a = i & 0xFF, v = (i >> 8) & 0xFF, cin = (i >> 16) & 1, d = (i >> 17) & 1;
        orc_cpu c;
        memset(&c, 0, sizeof(c));
        c.pc = 0x8000; c.sp = 0xFD; c.a = (uint8_t)a;
        c.p = (uint8_t)((cin ? ORC_C : 0) | (d ? ORC_D : 0));
        f->ram[0x8000] = op;
        f->ram[0x8001] = (uint8_t)v;
        f->ntrace = 0;
        orc_step_flat(&c, f);
        out[i] = (uint16_t)((c.p << 8) | c.a);
    }
    free(f);
}

/* ---------------------------------------------------------------- component hooks */
/* Applies a list of TIA / PIA operations to a console unpacked from a
 * 222-byte state (tests drive the components through the packed state).
 * op = (kind << 24) | (addr << 8) | val; results[i] receives read values,
 * skipped-cycle counts or 0.
 *   0 tia.tick_n(val)   1 tia.skip_to_scanline_end   2 tia.write(addr, val)
 *   3 tia.read(addr)    4 pia.tick_n(val)            5 pia.read(addr)
 *   6 pia.write(addr, val)                           7 pia.tick() x val   */
void orc_component_run(const uint8_t *rom, uint32_t rom_len, uint8_t *state,
                       const uint32_t *ops, uint32_t nops, uint32_t *results)
{
    orc_atari *e = (orc_atari *)calloc(1, sizeof(orc_atari));
    orc_atari_init(e, rom, rom_len);
    orc_atari_unpack(e, state);
    for (uint32_t i = 0; i < nops; i++) {
        uint32_t k = ops[i] >> 24, a = (ops[i] >> 8) & 0xFFFF, v = ops[i] & 0xFF;
        uint32_t r = 0;
        switch (k) {
        case 0: orc_tia_tick_n(&e->bus.tia, (uint8_t)v); break;
        case 1: r = orc_tia_skip_to_scanline_end(&e->bus.tia); break;
        case 2: orc_tia_write(&e->bus.tia, (uint16_t)a, (uint8_t)v); break;
        case 3: r = orc_tia_read(&e->bus.tia, (uint16_t)a); break;
        case 4: orc_pia_tick_n(&e->bus.pia, (uint16_t)v); break;
        case 5: r = orc_pia_read(&e->bus.pia, (uint16_t)a); break;
        case 6: orc_pia_write(&e->bus.pia, (uint16_t)a, (uint8_t)v); break;
        case 7: for (uint32_t j = 0; j < v; j++) orc_pia_tick(&e->bus.pia); break;
        default: break;
        }
        if (results) results[i] = r;
    }
    orc_atari_pack(e, state);
    free(e);
}

/* Timer sweep: for each of k states (tval, tps, tsel, tuf) and step n, runs
 * tick_n(n) and, separately, n single tick() calls.  out_n / out_1 receive
 * (tval, tps, tsel, tuf); uf_seen[i] = 1 if the per-cycle run underflowed. */
void orc_pia_sweep(const uint16_t *in, const uint16_t *n, uint32_t k,
                   uint16_t *out_n, uint16_t *out_1, uint8_t *uf_seen)
{
    for (uint32_t i = 0; i < k; i++) {
        orc_pia a, b;
        orc_pia_default(&a);
        a.tval = in[4 * i]; a.tps = in[4 * i + 1]; a.tsel = in[4 * i + 2]; a.tuf = (uint8_t)in[4 * i + 3];
        b = a;
        orc_pia_tick_n(&a, n[i]);
        uint8_t seen = 0;
        for (uint16_t j = 0; j < n[i]; j++) {
            uint8_t before = b.tuf;
            orc_pia_tick(&b);
            if (b.tuf && !before) seen = 1;
        }
        out_n[4 * i] = a.tval; out_n[4 * i + 1] = a.tps; out_n[4 * i + 2] = a.tsel; out_n[4 * i + 3] = a.tuf;
        out_1[4 * i] = b.tval; out_1[4 * i + 1] = b.tps; out_1[4 * i + 2] = b.tsel; out_1[4 * i + 3] = b.tuf;
        uf_seen[i] = seen;
    }
}

/* ---------------------------------------------------------------- rendering console */
#include "orc_render.inc"
