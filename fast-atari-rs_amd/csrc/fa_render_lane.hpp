// fa_render_lane.hpp -- one lane of the rendering console: a cycle-stepped
// 6502 on the rendering bus and the pixel-rendering TIA.
//
// Reference: Atari::{new, set_action, run_frame} (src/atari.rs:226-305),
// the rendering Bus (src/atari.rs:107-224), Tia (src/tia.rs:98-675), Pia
// (src/pia.rs:80-228) and the CPU's bus-access order (src/cpu.rs:245-1146).
// Included by fa_render.hip, which declares the workgroup's LDS as
// `r_lds` before including this file.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/fastatari.h"
#include "fa_render.hpp"

namespace fa {
namespace rnd {

constexpr uint32_t kW = FA_FRAME_WIDTH, kH = FA_FRAME_HEIGHT, kFrameBytes = kW * kH;
constexpr uint32_t kChunks = kFrameBytes / 16u;   // 16-byte chunks of a frame
constexpr uint32_t kRState = FA_RENDER_STATE_BYTES;
constexpr uint32_t kLoopBCap = 100000u;           // run_frame's loop B cap (src/atari.rs:296)
constexpr uint32_t kLoopACap = 1000000u;          // loop A is unbounded in the reference (FA_ENV_VSYNC_HELD here)
constexpr uint32_t kNoGroup = 0xFFFFFFFFu;
constexpr uint32_t kNoWord = 0xFFFFFFFFu;

struct RArgs {
    uint8_t *states;              // kRState bytes per env
    uint64_t *cpu_cycles, *frame_cycles, *instr;
    uint32_t *status;
    const uint8_t *actions;       // [f * stride + i]; NULL: no set_action
    uint8_t *fb;                  // kFrameBytes per env
    const uint32_t *tables;       // kTableBytes
    const uint8_t *rom;           // image, rom_bytes (a multiple of 4096)
    uint32_t rom_bytes, hot_lo, hot_n;
    uint32_t n, k, stride;
};

__device__ __forceinline__ uint32_t rd8(uint32_t o) { return r_lds[o]; }
__device__ __forceinline__ void wr8(uint32_t o, uint32_t v) { r_lds[o] = (uint8_t)v; }
__device__ __forceinline__ uint32_t rd32(uint32_t o) { return *reinterpret_cast<const uint32_t *>(r_lds + o); }
__device__ __forceinline__ void wr32(uint32_t o, uint32_t v) { *reinterpret_cast<uint32_t *>(r_lds + o) = v; }
__device__ __forceinline__ uint32_t rev8(uint32_t v) { return __builtin_bitreverse32(v) >> 24; }
// A select of two values.  Written as a call so that a choice between two
// members of the lane's state is a choice of values, never of addresses,
// and the state stays in registers.
__device__ __forceinline__ uint32_t pick(bool c, uint32_t t, uint32_t f) { return c ? t : f; }
// All ones when bit k of w is set, else zero.
__device__ __forceinline__ uint32_t lbm(uint32_t w, uint32_t k) { return (uint32_t)((int32_t)(w << (31u - k)) >> 31); }

// RESMx / RESBL: (clock - 68 + 5) mod 160 (calc_resp_position_simple, src/tia.rs:548-550)
__device__ __forceinline__ uint32_t resp_simple(uint32_t clock) {
    uint32_t t = clock + 97u;
    t = t >= 160u ? t - 160u : t;
    return t >= 160u ? t - 160u : t;
}
// RESP0 / RESP1 (calc_resp_position, src/tia.rs:536-545)
__device__ __forceinline__ uint32_t resp_player(uint32_t clock) {
    const uint32_t dc = (clock + (clock < 68u ? 3u : 4u) + 156u) % 228u;
    return dc >= 68u ? dc - 68u : dc + 92u;
}
// HMOVE: position minus the signed upper nibble, mod 160 (apply_hmove_offset, src/tia.rs:552-555)
__device__ __forceinline__ uint32_t hmove_pos(uint32_t pos, uint32_t hm) {
    const int32_t q = ((int32_t)pos - (((int32_t)(int8_t)hm) >> 4)) % 160;
    return (uint32_t)(q < 0 ? q + 160 : q);
}

// Copies after the first, as a mask over the offsets {16, 32, 64}, by NUSIZ
// mode 0..7 (src/tia.rs:470-480 and 507-514: modes 5 and 7 draw one copy).
__device__ __forceinline__ uint32_t copy_mask(uint32_t nusiz) { return (0x06043210u >> ((nusiz & 7u) * 4u)) & 7u; }
// Each bit of b8 spread over 2 / 4 bits (bit i -> bits 2i..2i+1 / 4i..4i+3).
__device__ __forceinline__ uint32_t spread2(uint32_t b8) {
    uint32_t x = b8 & 0xFFu;
    x = (x | (x << 4)) & 0x0F0Fu;
    x = (x | (x << 2)) & 0x3333u;
    x = (x | (x << 1)) & 0x5555u;
    return x * 3u;
}
__device__ __forceinline__ uint32_t spread4(uint32_t b8) {
    uint32_t x = b8 & 0xFFu;
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x * 0xFu;
}
// ORs a pattern (bit j = pixel s + j, at most 32 pixels) into a 160-pixel
// line mask (m[w] bit i = pixel 32w + i), wrapping at pixel 160 as object
// positions do (src/tia.rs:462-466).
__device__ __forceinline__ void place(uint32_t (&m)[kMaskWords], uint32_t pat, uint32_t s) {
    const uint32_t w = s >> 5, r = s & 31u;
    const uint32_t lo = pat << r, hi = r ? (pat >> (32u - r)) : 0u;
    const uint32_t w1 = w == kMaskWords - 1u ? 0u : w + 1u;
#pragma unroll
    for (uint32_t k = 0; k < kMaskWords; ++k) m[k] |= (k == w ? lo : 0u) | (k == w1 ? hi : 0u);
}
__device__ __forceinline__ uint32_t add160(uint32_t pos, uint32_t off) {
    const uint32_t q = pos + off;
    return q >= 160u ? q - 160u : q;
}
// Line mask of a player (get_player_pixel, src/tia.rs:451-494): byte g
// (leftmost pixel in bit 7, mirrored when refp), 1 / 2 / 4 pixels a bit by
// NUSIZ mode, copies 16 / 32 / 64 pixels right of the first.
__device__ __forceinline__ void player_mask(uint32_t (&m)[kMaskWords], uint32_t pos, uint32_t g, uint32_t refp,
                                            uint32_t nusiz) {
    const uint32_t mode = nusiz & 7u;
    const uint32_t b8 = refp ? (g & 0xFFu) : rev8(g);
    const uint32_t pat = mode == 5u ? spread2(b8) : (mode == 7u ? spread4(b8) : b8);
    const uint32_t cm = copy_mask(nusiz);
    place(m, pat, pos);
    if (cm & 1u) place(m, pat, add160(pos, 16u));
    if (cm & 2u) place(m, pat, add160(pos, 32u));
    if (cm & 4u) place(m, pat, add160(pos, 64u));
}
// Line mask of a missile (get_missile_pixel, src/tia.rs:497-522): NUSIZ bits
// 4-5 give 1 / 2 / 4 / 8 pixels, copies as the player's.
__device__ __forceinline__ void missile_mask(uint32_t (&m)[kMaskWords], uint32_t pos, uint32_t on, uint32_t nusiz) {
    const uint32_t pat = on ? (1u << (1u << ((nusiz >> 4) & 3u))) - 1u : 0u;
    const uint32_t cm = copy_mask(nusiz);
    place(m, pat, pos);
    if (cm & 1u) place(m, pat, add160(pos, 16u));
    if (cm & 2u) place(m, pat, add160(pos, 32u));
    if (cm & 4u) place(m, pat, add160(pos, 64u));
}

// One lane's console.
struct RLane {
    // CPU and the instruction in progress
    uint32_t a, x, y, sp, p, pc;
    uint32_t busy, rx, d0, dg, dh, t, cyc, opnd, lo, hi, dat, res;
    // run_frame's loop: in loop A (VSYNC held at the frame's start), cycles so far
    uint32_t in_a, fcyc;
    // TIA
    uint32_t clock, line, wsync, fcomp, vsync, vblank;
    uint32_t colup0, colup1, colupf, colubk, ctrlpf, pf0, pf1, pf2, pf20;
    uint32_t grp0, grp1, grp0o, grp1o, refp0, refp1, nusiz0, nusiz1, vdelp0, vdelp1, vdelbl;
    uint32_t enam0, enam1, enabl, enablo, resmp0, resmp1;
    uint32_t pp0, pp1, pm0, pm1, pbl, hp0, hp1, hmm0, hmm1, hbl;
    uint32_t cxa, cxb;            // CXM0P, CXM1P, CXP0FB, CXP1FB / CXM0FB, CXM1FB, CXBLPF, CXPPMM, a byte each
    uint64_t seen;                // object sets drawn since the latches were last brought up to date
    uint32_t inpt4, inpt5, pad0, pad1, pcount, pdump;
    uint32_t hmove, vtop, vbtop, ptop, pframes, l4, l5, len;
    // RIOT and cartridge
    uint32_t pa_out, pa_ddr, pa_in, pb_out, pb_ddr, pb_in, tval, tps, tsel, tuf, bank;
    // frame output: the open pixel group (row << 4 | group), its drawn-pixel
    // mask, and the zero fill's next chunk (kChunks: none pending)
    uint32_t gkey, gmask, zq;
    // addressing
    uint32_t ram0, grp0ff, msk0, stride4;   // LDS: this lane's RAM, pixel group and line masks
    uint32_t hot_lo, hot_n, rom0;
    uint8_t *frame;                   // this env's frame in HBM
    // The six objects' line masks (P0, P1, M0, M1, BL, PF; LDS, rebuilt when
    // a register they depend on is written) and, in registers, the words
    // holding pixels 32 cw .. 32 cw + 31 of each (cw = kNoWord: not loaded).
    uint32_t cw, c_p0, c_p1, c_m0, c_m1, c_bl, c_pf;

    __device__ __forceinline__ uint32_t ram_off(uint32_t i) const { return ram0 + (i >> 2) * stride4 + (i & 3u); }
    __device__ __forceinline__ uint32_t grp_off(uint32_t i) const { return grp0ff + (i >> 2) * stride4 + (i & 3u); }
    __device__ __forceinline__ uint32_t msk_off(uint32_t obj, uint32_t w) const {
        return msk0 + (obj * kMaskWords + w) * stride4;
    }

    // the 20 playfield bits in display order: bit i lights pixels 4i..4i+3 of the left half
    __device__ __forceinline__ void update_pf20() {
        pf20 = ((pf0 >> 4) & 0xFu) | (rev8(pf1) << 4) | ((pf2 & 0xFFu) << 12);
    }

    // ---------------------------------------------------------- frame output
    __device__ __forceinline__ void flush_group() {
        if (gmask == 0u) return;
        uint32_t w[4];
        for (uint32_t k = 0; k < 4u; ++k) w[k] = rd32(grp0ff + k * stride4);
        uint4 *dst = reinterpret_cast<uint4 *>(frame + (size_t)(gkey >> 4) * kW + (gkey & 15u) * 16u);
        if (gmask != 0xFFFFu) {
            // pixels the group did not draw keep the frame's bytes
            const uint4 old = *dst;
            const uint32_t ow[4] = {old.x, old.y, old.z, old.w};
            for (uint32_t k = 0; k < 4u; ++k) {
                const uint32_t m4 = (gmask >> (4u * k)) & 0xFu;
                const uint32_t bm = ((m4 * 0x00204081u) & 0x01010101u) * 0xFFu;   // bit i -> byte i
                w[k] = (w[k] & bm) | (ow[k] & ~bm);
            }
        }
        *dst = make_uint4(w[0], w[1], w[2], w[3]);
        gmask = 0;
    }
    // one chunk of a pending frame clear
    __device__ __forceinline__ void zero_step() {
        if (zq < kChunks) {
            reinterpret_cast<uint4 *>(frame)[zq] = make_uint4(0, 0, 0, 0);
            zq += 1u;
        }
    }

    // ---------------------------------------------------------- line masks
    __device__ __forceinline__ void store_mask(uint32_t obj, const uint32_t (&m)[kMaskWords]) {
#pragma unroll
        for (uint32_t k = 0; k < kMaskWords; ++k) wr32(msk_off(obj, k), m[k]);
    }
    // Rebuilds the masks of the objects in `objs` (bit i: object i).
    __device__ void rebuild(uint32_t objs) {
        if (objs & 1u) {
            uint32_t m[kMaskWords] = {0, 0, 0, 0, 0};
            player_mask(m, pp0, pick(vdelp0, grp0o, grp0), refp0, nusiz0);
            store_mask(0, m);
        }
        if (objs & 2u) {
            uint32_t m[kMaskWords] = {0, 0, 0, 0, 0};
            player_mask(m, pp1, pick(vdelp1, grp1o, grp1), refp1, nusiz1);
            store_mask(1, m);
        }
        if (objs & 4u) {
            uint32_t m[kMaskWords] = {0, 0, 0, 0, 0};
            missile_mask(m, pm0, enam0 && !resmp0, nusiz0);
            store_mask(2, m);
        }
        if (objs & 8u) {
            uint32_t m[kMaskWords] = {0, 0, 0, 0, 0};
            missile_mask(m, pm1, enam1 && !resmp1, nusiz1);
            store_mask(3, m);
        }
        if (objs & 16u) {
            // ball (get_ball_pixel, src/tia.rs:525-532): CTRLPF bits 4-5 give 1 / 2 / 4 / 8 pixels
            uint32_t m[kMaskWords] = {0, 0, 0, 0, 0};
            const uint32_t on = pick(vdelbl, enablo, enabl);
            place(m, on ? (1u << (1u << ((ctrlpf >> 4) & 3u))) - 1u : 0u, pbl);
            store_mask(4, m);
        }
        if (objs & 32u) {
            // playfield (get_playfield_bit, src/tia.rs:432-448): each of the 20
            // bits lights 4 pixels of the left half; the right half repeats
            // it, or mirrors it when CTRLPF bit 0 is set
            const uint32_t l0 = spread4(pf20), l1 = spread4(pf20 >> 8), l2 = spread4((pf20 >> 16) & 0xFu);
            uint32_t r2, r3, r4;
            if (ctrlpf & 1u) {
                r2 = __builtin_bitreverse32(l2);
                r3 = __builtin_bitreverse32(l1);
                r4 = __builtin_bitreverse32(l0);
            } else {
                r2 = spread4(pf20 & 0xFu) << 16;
                r3 = spread4(pf20 >> 4);
                r4 = spread4(pf20 >> 12);
            }
            const uint32_t m[kMaskWords] = {l0, l1, (l2 & 0xFFFFu) | (r2 & 0xFFFF0000u), r3, r4};
            store_mask(5, m);
        }
        cw = kNoWord;
    }
    __device__ __forceinline__ void load_words(uint32_t w) {
        c_p0 = rd32(msk_off(0, w));
        c_p1 = rd32(msk_off(1, w));
        c_m0 = rd32(msk_off(2, w));
        c_m1 = rd32(msk_off(3, w));
        c_bl = rd32(msk_off(4, w));
        c_pf = rd32(msk_off(5, w));
        cw = w;
    }

    // ---------------------------------------------------------- TIA
    // Tia::resize_commit, src/tia.rs:269-292
    __device__ __forceinline__ void resize_commit() {
        if (vbtop > 262u) return;
        const uint32_t cl = vbtop < 23u ? 23u : (vbtop > 32u ? 32u : vbtop);
        const uint32_t cand = pick(cl < vtop, cl, vtop);
        if (cand != ptop) {
            ptop = cand;
            pframes = 3;
        } else if (pframes > 0u) {
            pframes -= 1u;
            if (pframes == 0u) vtop = ptop;
        }
        vbtop = vtop;
    }

    // end of a scanline (src/tia.rs:371-397)
    __device__ __forceinline__ void line_end() {
        clock = 0;
        line += 1u;
        wsync = 0;
        hmove = 0;
        if (line == vtop) {       // framebuffer.fill(0): the open group is dropped, the zero fill restarts
            gmask = 0;
            gkey = kNoGroup;
            zq = 0;
        }
        if (!pdump && pcount < 256u) pcount += 1u;
        if (line >= 262u) {
            line = 0;
            fcomp = 1;
        }
    }

    // One TIA clock (Tia::tick, src/tia.rs:296-399).  A playfield write takes
    // effect at once here: the reference queues it for one clock, and the
    // clock that applies it follows every write before any pixel is drawn.
    __device__ __forceinline__ void tick() {
        const bool blank = (vblank & 2u) != 0u;
        if (len && !inpt4) l4 = 0;
        if (len && !inpt5) l5 = 0;
        if (!blank && line >= 23u && line < vbtop) vbtop = line;
        const uint32_t vis = (line - vtop) & 0xFFFFu;
        const uint32_t px = (clock - 68u) & 0xFFFFu;
        if (clock >= 68u && px < kW && !blank && vis < kH) {
            if ((px >> 5) != cw) load_words(px >> 5);
            const uint32_t sh = px & 31u;
            const uint32_t p0 = (c_p0 >> sh) & 1u, p1 = (c_p1 >> sh) & 1u, m0 = (c_m0 >> sh) & 1u;
            const uint32_t m1 = (c_m1 >> sh) & 1u, bl = (c_bl >> sh) & 1u, pf = (c_pf >> sh) & 1u;
            const uint32_t pfc = pick(ctrlpf & 2u, pick(px < 80u, colup0, colup1), colupf);
            const uint32_t s0 = p0 | m0, s1 = p1 | m1, sf = pf | bl;
            uint32_t color;
            if (ctrlpf & 4u) color = pick(sf, pfc, pick(s0, colup0, pick(s1, colup1, colubk)));
            else color = pick(s0, colup0, pick(s1, colup1, pick(sf, pfc, colubk)));
            color = pick(hmove && px < 8u, colubk, color);
            const uint32_t key = (vis << 4) | (px >> 4);
            if (key != gkey) {
                flush_group();
                gkey = key;
            }
            wr8(grp_off(px & 15u), color);
            gmask |= 1u << (px & 15u);
            seen |= 1ull << (p0 | (p1 << 1) | (m0 << 2) | (m1 << 3) | (bl << 4) | (pf << 5));
        }
        clock += 1u;
        if (clock >= 228u) line_end();
    }

    // The collision latches (src/tia.rs:414-428) from the object sets drawn
    // since the last fold: a latch bit is set when a drawn set holds both of
    // its objects.
    __device__ __forceinline__ void fold_collisions() {
        constexpr uint32_t P0 = 0, P1 = 1, M0 = 2, M1 = 3, BL = 4, PF = 5;
        const uint64_t sn = seen;
        auto hit = [sn](uint64_t sets, uint32_t bit) { return (sn & sets) ? (1u << bit) : 0u; };
        cxa |= hit(overlap_sets(M0, P1), 7) | hit(overlap_sets(M0, P0), 6) | hit(overlap_sets(M1, P0), 15) |
               hit(overlap_sets(M1, P1), 14) | hit(overlap_sets(P0, PF), 23) | hit(overlap_sets(P0, BL), 22) |
               hit(overlap_sets(P1, PF), 31) | hit(overlap_sets(P1, BL), 30);
        cxb |= hit(overlap_sets(M0, PF), 7) | hit(overlap_sets(M0, BL), 6) | hit(overlap_sets(M1, PF), 15) |
               hit(overlap_sets(M1, BL), 14) | hit(overlap_sets(BL, PF), 23) | hit(overlap_sets(P0, P1), 31) |
               hit(overlap_sets(M0, M1), 30);
        seen = 0;
    }

    // Tia::read, src/tia.rs:557-590
    __device__ __forceinline__ uint32_t tia_read(uint32_t a13) {
        const uint32_t r = a13 & 0x0Fu;
        uint32_t v = 0;
        if (r < 8u && seen) fold_collisions();
        if (r < 8u) v = (pick(r < 4u, cxa, cxb) >> ((r & 3u) * 8u)) & 0xFFu;
        else if (r == 8u) v = pcount >= pad0 ? 0x80u : 0u;
        else if (r == 9u) v = pcount >= pad1 ? 0x80u : 0u;
        else if (r < 12u) v = 0x80u;
        else if (r == 12u) v = len ? l4 : (inpt4 ? 0x80u : 0u);
        else if (r == 13u) v = len ? l5 : (inpt5 ? 0x80u : 0u);
        return v;
    }

    // Tia::write, src/tia.rs:593-673
    __device__ __forceinline__ void tia_write(uint32_t a13, uint32_t v) {
        uint32_t objs = 0;               // line masks to rebuild
        switch (a13 & 0x3Fu) {
        case 0x00:
            if (!(vsync & 2u) && (v & 2u)) {
                resize_commit();
                line = 0;
                fcomp = 1;
            }
            vsync = v;
            break;
        case 0x01:
            if (v & 0x40u) {
                len = 1;
                l4 = 0x80u;
                l5 = 0x80u;
            } else {
                len = 0;
            }
            if (v & 0x80u) {
                pcount = 0;
                pdump = 1;
            } else {
                pdump = 0;
            }
            vblank = v;
            break;
        case 0x02: wsync = 1; break;
        case 0x04: nusiz0 = v; objs = 1u | 4u; break;
        case 0x05: nusiz1 = v; objs = 2u | 8u; break;
        case 0x06: colup0 = v; break;
        case 0x07: colup1 = v; break;
        case 0x08: colupf = v; break;
        case 0x09: colubk = v; break;
        case 0x0A: ctrlpf = v; objs = 16u | 32u; break;
        case 0x0B: refp0 = (v >> 3) & 1u; objs = 1u; break;
        case 0x0C: refp1 = (v >> 3) & 1u; objs = 2u; break;
        case 0x0D: pf0 = v; update_pf20(); objs = 32u; break;
        case 0x0E: pf1 = v; update_pf20(); objs = 32u; break;
        case 0x0F: pf2 = v; update_pf20(); objs = 32u; break;
        case 0x10: pp0 = resp_player(clock); objs = 1u; break;
        case 0x11: pp1 = resp_player(clock); objs = 2u; break;
        case 0x12: pm0 = resp_simple(clock); objs = 4u; break;
        case 0x13: pm1 = resp_simple(clock); objs = 8u; break;
        case 0x14: pbl = resp_simple(clock); objs = 16u; break;
        case 0x1B: grp0o = grp0; grp0 = v; objs = 1u; break;
        case 0x1C: grp1o = grp1; grp1 = v; enablo = enabl; objs = 2u | 16u; break;
        case 0x1D: enam0 = (v >> 1) & 1u; objs = 4u; break;
        case 0x1E: enam1 = (v >> 1) & 1u; objs = 8u; break;
        case 0x1F: enabl = (v >> 1) & 1u; objs = 16u; break;
        case 0x20: hp0 = v; break;
        case 0x21: hp1 = v; break;
        case 0x22: hmm0 = v; break;
        case 0x23: hmm1 = v; break;
        case 0x24: hbl = v; break;
        case 0x25: vdelp0 = v & 1u; objs = 1u; break;
        case 0x26: vdelp1 = v & 1u; objs = 2u; break;
        case 0x27: vdelbl = v & 1u; objs = 16u; break;
        case 0x28: resmp0 = (v >> 1) & 1u; objs = 4u; break;
        case 0x29: resmp1 = (v >> 1) & 1u; objs = 8u; break;
        case 0x2A:
            pp0 = hmove_pos(pp0, hp0);
            pp1 = hmove_pos(pp1, hp1);
            pm0 = hmove_pos(pm0, hmm0);
            pm1 = hmove_pos(pm1, hmm1);
            pbl = hmove_pos(pbl, hbl);
            hmove = 1;
            objs = 31u;
            break;
        case 0x2B: hp0 = hp1 = hmm0 = hmm1 = hbl = 0; break;
        case 0x2C: cxa = cxb = 0; seen = 0; break;
        default: break;           // RSYNC, audio, unused
        }
        if (objs) rebuild(objs);
    }

    // ---------------------------------------------------------- RIOT
    // Pia::tick, src/pia.rs:80-103
    __device__ __forceinline__ void pia_tick() {
        if (tval == 0u && tuf) {
            tval = 0xFFu;
            return;
        }
        tps = (tps - 1u) & 0xFFFFu;
        if (tps == 0u) {
            tps = tuf ? 1u : tsel;
            if (tval == 0u) {
                tuf = 1;
                tval = 0xFFu;
            } else {
                tval -= 1u;
            }
        }
    }
    // Pia::read / write on the I/O and timer registers, src/pia.rs:169-228
    __device__ __forceinline__ uint32_t pia_read_io(uint32_t a13) {
        switch (a13 & 7u) {
        case 0: return (pa_in & ~pa_ddr & 0xFFu) | (pa_out & pa_ddr);
        case 1: return pa_ddr;
        case 2: return (pb_in & ~pb_ddr & 0xFFu) | (pb_out & pb_ddr);
        case 3: return pb_ddr;
        case 4: tuf = 0; return tval & 0xFFu;
        case 5: return tuf ? 0x80u : 0u;
        default: return 0u;
        }
    }
    __device__ __forceinline__ void pia_write_io(uint32_t a13, uint32_t v) {
        if (a13 & 0x10u) {        // TIM1T / TIM8T / TIM64T / T1024T
            tval = v;
            tuf = 0;
            tsel = 1u << ((0x0A060300u >> ((a13 & 3u) * 8u)) & 0xFFu);
            tps = tsel;
            return;
        }
        switch (a13 & 3u) {
        case 0: pa_out = v; break;
        case 1: pa_ddr = v; break;
        case 2: pb_out = v; break;
        default: pb_ddr = v; break;
        }
    }

    // ---------------------------------------------------------- bus
    // check_bankswitch: hotspot lo12 - hot_lo < hot_n in the cartridge window (src/atari.rs:152-193)
    __device__ __forceinline__ void bankswitch(uint32_t a13) {
        const uint32_t hd = (a13 & 0xFFFu) - hot_lo;
        if ((a13 & 0x1000u) && hd < hot_n) bank = hd;
    }
    // Bus::read's decode (src/atari.rs:206-218): TIA, RAM, RIOT, cartridge
    __device__ __forceinline__ uint32_t bus_read(uint32_t addr) {
        const uint32_t a13 = addr & 0x1FFFu;
        bankswitch(a13);
        uint32_t v;
        if ((a13 & 0x1080u) == 0u) v = tia_read(a13);
        else if ((a13 & 0x1280u) == 0x0080u) v = rd8(ram_off(a13 & 0x7Fu));
        else if ((a13 & 0x1280u) == 0x0280u) v = pia_read_io(a13);
        else v = rd8(rom0 + (bank << 12) + (a13 & 0xFFFu));
        return v;
    }
    __device__ __forceinline__ void bus_write(uint32_t addr, uint32_t v) {
        const uint32_t a13 = addr & 0x1FFFu;
        bankswitch(a13);
        if ((a13 & 0x1080u) == 0u) tia_write(a13, v);
        else if ((a13 & 0x1280u) == 0x0080u) wr8(ram_off(a13 & 0x7Fu), v);
        else if ((a13 & 0x1280u) == 0x0280u) pia_write_io(a13, v);
    }

    // ---------------------------------------------------------- CPU
    // Decimal ADC / SBC (src/cpu.rs:155-166, 180-191) after the binary
    // result and flags are in place: ADC replaces A, N, V and C; SBC only A.
    __device__ __forceinline__ void decimal(bool sub, uint32_t a0, uint32_t v, uint32_t c) {
        if (!sub) {
            uint32_t l = (a0 & 0xFu) + (v & 0xFu) + c;
            l += l > 9u ? 6u : 0u;
            uint32_t h = (a0 >> 4) + (v >> 4) + (l > 0xFu ? 1u : 0u);
            const uint32_t inter = ((h << 4) | (l & 0xFu)) & 0xFFu;
            uint32_t np = p & ~0xC1u;
            np |= (h & 8u) ? 0x80u : 0u;
            np |= (~(a0 ^ v) & (a0 ^ inter) & 0x80u) ? 0x40u : 0u;
            h += h > 9u ? 6u : 0u;
            np |= h > 0xFu ? 1u : 0u;
            p = np;
            a = ((h & 0xFu) << 4) | (l & 0xFu);
        } else {
            int32_t l = (int32_t)(a0 & 0xFu) - (int32_t)(v & 0xFu) - (int32_t)(c ^ 1u);
            const int32_t lb = l < 0 ? 1 : 0;
            l -= lb ? 6 : 0;
            int32_t h = (int32_t)(a0 >> 4) - (int32_t)(v >> 4) - lb;
            h -= h < 0 ? 6 : 0;
            a = (((uint32_t)h & 0xFu) << 4) | ((uint32_t)l & 0xFu);
        }
    }
    __device__ __forceinline__ uint32_t reg(uint32_t r) const {
        return pick(r == R_A, a, pick(r == R_X, x, pick(r == R_Y, y, sp)));
    }
    __device__ __forceinline__ void set_reg(uint32_t r, uint32_t v) {
        if (r == R_A) a = v;
        else if (r == R_X) x = v;
        else if (r == R_Y) y = v;
        else if (r == R_SP) sp = v;
    }
    __device__ __forceinline__ void set_zn(uint32_t v) { p = (p & ~0x82u) | (v & 0x80u) | (v == 0u ? 0x02u : 0u); }

    // The register / flag effect of the instruction (src/cpu.rs:245-1146) on
    // its operand: the data byte read (memory or immediate) or the source
    // register.  Straight-line code steered by the opcode's one-hot words
    // (g: result and register selection, h: flag masks), so lanes at
    // different opcodes share it; a memory read-modify-write's result goes to
    // `res` for the final write.
    __device__ __forceinline__ void alu() {
        const uint32_t spec = (rx >> RX_SPEC_SHIFT) & 7u;
        if (spec != SX_NONE) {           // PLA / PLP
            if (spec == SX_PLA) {
                a = dat;
                set_zn(dat);
            } else {
                p = dat & ~0x30u;
            }
            return;
        }
        const uint32_t g = dg, h = dh;
        const uint32_t s = (a & lbm(g, G_SRC)) | (x & lbm(g, G_SRC + 1u)) | (y & lbm(g, G_SRC + 2u)) |
                           (sp & lbm(g, G_SRC + 3u));
        const uint32_t v = pick((d0 & (D_IMM | D_RD)) != 0u, dat, s);
        const uint32_t p0 = p, a0 = a;
        const uint32_t lg = (a & v & lbm(h, H_M_AND)) | ((a ^ v) & lbm(h, H_M_XOR)) | (v & lbm(h, H_M_LD));
        const uint32_t vv = v ^ (0xFFu & lbm(d0, 25u));               // D_INV
        const uint32_t cin = p & 1u;
        const uint32_t sum = s + vv + (cin | ((d0 >> 26) & 1u));      // D_CIN1
        const uint32_t ar = sum & 0xFFu;
        const uint32_t shl = ((v << 1) | (cin & lbm(g, G_CIN_SHL))) & 0xFFu;
        const uint32_t shr = (v >> 1) | ((cin << 7) & lbm(g, G_CIN_SHR));
        const uint32_t inc = (v + (g >> G_DELTA)) & 0xFFu;
        const uint32_t r = (lg & lbm(g, G_R_LOG)) | (ar & lbm(g, G_R_ARI)) | (shl & lbm(g, G_R_SHL)) |
                           (shr & lbm(g, G_R_SHR)) | (inc & lbm(g, G_R_INC)) | (s & lbm(g, G_R_SRC));
        // N, V, Z, C candidates, then the opcode's clear / take / set masks
        const uint32_t mbit = lbm(g, G_BIT);
        const uint32_t nv = (v & mbit) | (r & ~mbit);
        const uint32_t vf = (((~(s ^ vv) & (s ^ ar)) >> 1) & 0x40u & ~mbit) | (v & 0x40u & mbit);
        const uint32_t cf = (((sum >> 8) & lbm(g, G_C_SUM)) | ((v >> 7) & lbm(g, G_C_B7)) | (v & lbm(g, G_C_B0))) & 1u;
        const uint32_t cand = (nv & 0x80u) | vf | (r == 0u ? 0x02u : 0u) | cf;
        p = (p & ~(h & 0xFFu)) | (cand & (h >> H_TAKE)) | ((h >> H_SET) & 0xFFu);
        const uint32_t ma = lbm(g, G_DST), mx = lbm(g, G_DST + 1u), my = lbm(g, G_DST + 2u), ms = lbm(g, G_DST + 3u);
        a = (r & ma) | (a & ~ma);
        x = (r & mx) | (x & ~mx);
        y = (r & my) | (y & ~my);
        sp = (r & ms) | (sp & ~ms);
        res = r;
        if ((d0 & D_DEC) && (p0 & 0x08u)) decimal((d0 & D_INV) != 0u, a0, v, cin);
    }

    __device__ __forceinline__ uint32_t idx() const { return pick(rx & RX_IDXY, y, x); }
    __device__ __forceinline__ uint32_t base() const { return (hi << 8) | lo; }
    __device__ __forceinline__ uint32_t ea() const { return (base() + idx()) & 0xFFFFu; }
    __device__ __forceinline__ bool cross() const { return ((base() ^ ea()) & 0xFF00u) != 0u; }
    __device__ __forceinline__ bool taken() const {
        const uint32_t fm = (dh >> D1_FMASK_SHIFT) & 0xFFu;
        return ((p & fm) != 0u) == ((dh >> D1_FVAL_SHIFT) != 0u);
    }
    // branch target from the PC after the operand (src/cpu.rs:925-940)
    __device__ __forceinline__ uint32_t target() const { return (pc + opnd - ((opnd & 0x80u) << 1)) & 0xFFFFu; }

    __device__ __forceinline__ uint32_t uaddr(uint32_t u) const {
        switch (u & 15u) {
        case AS_PC: return pc;
        case AS_OPND: return opnd;
        case AS_ZPI: return (opnd + idx()) & 0xFFu;
        case AS_PTRX0: return (opnd + x) & 0xFFu;
        case AS_PTRX1: return (opnd + x + 1u) & 0xFFu;
        case AS_PTR1: return (opnd + 1u) & 0xFFu;
        case AS_ABS: return base();
        case AS_WRONG: return (hi << 8) | ((lo + idx()) & 0xFFu);
        case AS_EA: return ea();
        case AS_STACK: return 0x100u | sp;
        case AS_VEC0: return 0xFFFEu;
        case AS_VEC1: return 0xFFFFu;
        case AS_JIND1: return (hi << 8) | ((lo + 1u) & 0xFFu);
        default: return (pc & 0xFF00u) | (target() & 0xFFu);   // AS_BRWRONG
        }
    }
    __device__ __forceinline__ uint32_t uwdata(uint32_t u) const {
        switch ((u >> 5) & 7u) {
        case WS_STORE: return reg((rx >> RX_SREG_SHIFT) & 3u);
        case WS_DAT: return dat;
        case WS_RES: return res;
        case WS_PCH: return pc >> 8;
        case WS_PCL: return pc & 0xFFu;
        default: return p | 0x30u;   // WS_PPUSH: B and U set (src/cpu.rs:1066, 1096)
        }
    }
};

// Atari::set_action (src/atari.rs:249-268) on action index 0..17 (Action,
// src/atari.rs:9-81): up / down / left / right pull SWCHA bits 4 / 5 / 6 / 7
// low, fire pulls INPT4 low, and paddle 0 reads 200 for left, 50 for right.
__device__ __forceinline__ void set_action(RLane &e, uint32_t act) {
    // bit 0 up, 1 down, 2 left, 3 right, 4 fire (src/atari.rs:60-79)
    constexpr uint32_t kDec[18] = {0x00, 0x10, 0x01, 0x08, 0x04, 0x02, 0x09, 0x05, 0x0A,
                                   0x06, 0x11, 0x18, 0x14, 0x12, 0x19, 0x15, 0x1A, 0x16};
    const uint32_t d = act < 18u ? kDec[act] : 0u;
    e.pa_in = 0xFFu & ~((d & 0xFu) << 4);
    e.inpt4 = (d & 0x10u) ? 0u : 1u;
    e.pad0 = (d & 4u) ? 200u : ((d & 8u) ? 50u : 128u);
}

// ---------------------------------------------------------------- state I/O
// The render state: AtariStateGpu offsets (src/cuda_env.rs:12-38) with
// tia_clock in TIA clocks, then visible_top @224, vblank_top @226,
// pending_top @228 (u16), pending_frames @230, hmove_pending @231,
// inpt4_latch @232, inpt5_latch @233, input_latch_enabled @234.
__device__ __forceinline__ uint32_t sb(const uint32_t *w, uint32_t o) { return (w[o >> 2] >> ((o & 3u) * 8u)) & 0xFFu; }
__device__ __forceinline__ uint32_t sh(const uint32_t *w, uint32_t o) { return sb(w, o) | (sb(w, o + 1u) << 8); }

__device__ __forceinline__ void load_lane(RLane &e, const uint8_t *src) {
    uint32_t w[64];
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    for (int i = 0; i < 16; ++i) {
        const uint4 v = s4[i];
        w[4 * i] = v.x;
        w[4 * i + 1] = v.y;
        w[4 * i + 2] = v.z;
        w[4 * i + 3] = v.w;
    }
    e.a = sb(w, 0); e.x = sb(w, 1); e.y = sb(w, 2); e.sp = sb(w, 3); e.pc = sh(w, 4); e.p = sb(w, 6);
    _Pragma("unroll") for (uint32_t q = 0; q < 32u; ++q) wr32(e.ram0 + q * e.stride4, w[2 + q]);   // RAM, bytes 8..135
    e.pa_out = sb(w, 136); e.pa_ddr = sb(w, 137); e.pa_in = sb(w, 138);
    e.pb_out = sb(w, 139); e.pb_ddr = sb(w, 140); e.pb_in = sb(w, 141);
    e.tval = sh(w, 142); e.tps = sh(w, 144); e.tsel = sh(w, 146); e.tuf = sb(w, 148);
    e.clock = sh(w, 152); e.line = sh(w, 154); e.wsync = sb(w, 156); e.fcomp = sb(w, 157);
    e.vsync = sb(w, 158); e.vblank = sb(w, 159);
    e.cxa = w[40];                                                                // bytes 160..163
    e.cxb = w[41];                                                                // bytes 164..167
    e.inpt4 = sb(w, 168); e.inpt5 = sb(w, 169); e.pad0 = sb(w, 170); e.pad1 = sb(w, 171);
    e.pcount = sh(w, 172); e.pdump = sb(w, 174);
    e.pp0 = sh(w, 176); e.pp1 = sh(w, 178); e.pm0 = sh(w, 180); e.pm1 = sh(w, 182); e.pbl = sh(w, 184);
    e.hp0 = sb(w, 186); e.hp1 = sb(w, 187); e.hmm0 = sb(w, 188); e.hmm1 = sb(w, 189); e.hbl = sb(w, 190);
    e.grp0 = sb(w, 192); e.grp1 = sb(w, 193); e.grp0o = sb(w, 194); e.grp1o = sb(w, 195);
    e.enam0 = sb(w, 196); e.enam1 = sb(w, 197); e.enabl = sb(w, 198); e.enablo = sb(w, 199);
    e.vdelp0 = sb(w, 200); e.vdelp1 = sb(w, 201); e.vdelbl = sb(w, 202); e.resmp0 = sb(w, 203);
    e.resmp1 = sb(w, 204); e.nusiz0 = sb(w, 205); e.nusiz1 = sb(w, 206);
    e.colup0 = sb(w, 207); e.colup1 = sb(w, 208); e.colupf = sb(w, 209); e.colubk = sb(w, 210);
    e.ctrlpf = sb(w, 211); e.refp0 = sb(w, 212); e.refp1 = sb(w, 213);
    e.pf0 = sb(w, 214); e.pf1 = sb(w, 215); e.pf2 = sb(w, 216); e.bank = sb(w, 217);
    e.vtop = sh(w, 224); e.vbtop = sh(w, 226); e.ptop = sh(w, 228); e.pframes = sb(w, 230);
    e.hmove = sb(w, 231); e.l4 = sb(w, 232); e.l5 = sb(w, 233); e.len = sb(w, 234);
    e.update_pf20();
}

__device__ __forceinline__ void put8(uint32_t *w, uint32_t o, uint32_t v) {
    const uint32_t s = (o & 3u) * 8u;
    w[o >> 2] = (w[o >> 2] & ~(0xFFu << s)) | ((v & 0xFFu) << s);
}
__device__ __forceinline__ void put16(uint32_t *w, uint32_t o, uint32_t v) {
    put8(w, o, v);
    put8(w, o + 1u, v >> 8);
}

__device__ __forceinline__ void store_lane(const RLane &e, uint8_t *dst, uint32_t scheme) {
    uint32_t w[64];
    for (int i = 0; i < 64; ++i) w[i] = 0;
    put8(w, 0, e.a); put8(w, 1, e.x); put8(w, 2, e.y); put8(w, 3, e.sp); put16(w, 4, e.pc); put8(w, 6, e.p);
    _Pragma("unroll") for (uint32_t q = 0; q < 32u; ++q) w[2 + q] = rd32(e.ram0 + q * e.stride4);
    put8(w, 136, e.pa_out); put8(w, 137, e.pa_ddr); put8(w, 138, e.pa_in);
    put8(w, 139, e.pb_out); put8(w, 140, e.pb_ddr); put8(w, 141, e.pb_in);
    put16(w, 142, e.tval); put16(w, 144, e.tps); put16(w, 146, e.tsel); put8(w, 148, e.tuf);
    put16(w, 152, e.clock); put16(w, 154, e.line); put8(w, 156, e.wsync); put8(w, 157, e.fcomp);
    put8(w, 158, e.vsync); put8(w, 159, e.vblank);
    w[40] = e.cxa;
    w[41] = e.cxb;
    put8(w, 168, e.inpt4); put8(w, 169, e.inpt5); put8(w, 170, e.pad0); put8(w, 171, e.pad1);
    put16(w, 172, e.pcount); put8(w, 174, e.pdump);
    put16(w, 176, e.pp0); put16(w, 178, e.pp1); put16(w, 180, e.pm0); put16(w, 182, e.pm1); put16(w, 184, e.pbl);
    put8(w, 186, e.hp0); put8(w, 187, e.hp1); put8(w, 188, e.hmm0); put8(w, 189, e.hmm1); put8(w, 190, e.hbl);
    put8(w, 192, e.grp0); put8(w, 193, e.grp1); put8(w, 194, e.grp0o); put8(w, 195, e.grp1o);
    put8(w, 196, e.enam0); put8(w, 197, e.enam1); put8(w, 198, e.enabl); put8(w, 199, e.enablo);
    put8(w, 200, e.vdelp0); put8(w, 201, e.vdelp1); put8(w, 202, e.vdelbl); put8(w, 203, e.resmp0);
    put8(w, 204, e.resmp1); put8(w, 205, e.nusiz0); put8(w, 206, e.nusiz1);
    put8(w, 207, e.colup0); put8(w, 208, e.colup1); put8(w, 209, e.colupf); put8(w, 210, e.colubk);
    put8(w, 211, e.ctrlpf); put8(w, 212, e.refp0); put8(w, 213, e.refp1);
    put8(w, 214, e.pf0); put8(w, 215, e.pf1); put8(w, 216, e.pf2); put8(w, 217, e.bank); put8(w, 218, scheme);
    put16(w, 224, e.vtop); put16(w, 226, e.vbtop); put16(w, 228, e.ptop); put8(w, 230, e.pframes);
    put8(w, 231, e.hmove); put8(w, 232, e.l4); put8(w, 233, e.l5); put8(w, 234, e.len);
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    for (int i = 0; i < 16; ++i) d4[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// ---------------------------------------------------------------- one lane
// k frames of env idx: for each, set_action and run_frame (src/atari.rs:249-305).
// Every loop iteration is one CPU cycle: a bus access from the opcode fetch
// or the micro-program (or a WSYNC stall), three TIA clocks with a write
// landing before the third, and one RIOT clock.  bs and tid place the lane's
// RAM and pixel group in the workgroup's LDS (tables at 0, the ROM image at
// kTableBytes, then bs * 128 bytes of RAM and bs * 16 bytes of pixel groups).
__device__ __forceinline__ void render_lane(const RArgs &g, uint32_t bs, uint32_t tid, uint32_t idx) {
    const uint32_t rom0 = kTableBytes;
    const uint32_t ram_base = rom0 + g.rom_bytes;
    const uint32_t grp_base = ram_base + 128u * bs;
    const uint32_t msk_base = grp_base + 16u * bs;
    RLane e;
    e.stride4 = 4u * bs;
    e.ram0 = ram_base + 4u * tid;
    e.grp0ff = grp_base + 4u * tid;
    e.msk0 = msk_base + 4u * tid;
    e.hot_lo = g.hot_lo;
    e.hot_n = g.hot_n;
    e.rom0 = rom0;
    e.frame = g.fb + (size_t)idx * kFrameBytes;
    uint8_t *st_ptr = g.states + (size_t)idx * kRState;
    load_lane(e, st_ptr);
    e.rebuild(63u);
    const uint32_t scheme = st_ptr[218];
    e.busy = 0;
    e.rx = e.d0 = e.dg = e.dh = e.t = e.cyc = e.opnd = e.lo = e.hi = e.dat = e.res = 0;
    e.seen = 0;
    e.gkey = kNoGroup;
    e.gmask = 0;
    e.zq = kChunks;
    uint64_t cpu_cyc = g.cpu_cycles[idx], instr = g.instr[idx], fcyc_last = g.frame_cycles[idx];
    uint32_t st = g.status[idx];
    uint32_t f = 0;                      // frame index within this launch
    bool live = g.k > 0u && st == FA_ENV_OK;
    if (live) {
        if (g.actions) set_action(e, g.actions[idx]);
        e.fcomp = 0;
        e.fcyc = 0;
        e.in_a = (e.vsync & 2u) ? 1u : 0u;
    }
    while (live) {
        e.zero_step();
        // ---- this cycle's bus access: at an instruction boundary, run_frame's
        // loop test, then a WSYNC stall or the opcode fetch; else the next
        // micro-op of the instruction in progress
        bool frame_end = false;
        uint32_t acc = 0;                // 0: stall, 1: read, 2: write
        uint32_t addr = 0, wval = 0, u = 0;
        if (!e.busy) {
            if (e.in_a && !(e.vsync & 2u)) e.in_a = 0;
            if (!e.in_a && (e.vsync & 2u)) frame_end = true;
            else if (!e.wsync) {
                acc = 1;
                addr = e.pc;
            }
        } else {
            const uint32_t prog = e.rx & 31u;
            u = rd32((prog * kProgLen + e.t) * 4u);
            if ((u & U_IFCROSS) && !e.cross()) {
                e.t += 1u;
                u = rd32((prog * kProgLen + e.t) * 4u);
            }
            if (u & U_SPINC) e.sp = (e.sp + 1u) & 0xFFu;
            addr = e.uaddr(u);
            acc = (u & U_W) ? 2u : 1u;
            wval = e.uwdata(u);
        }
        if (!frame_end) {
            // ---- the bus cycle (src/atari.rs:196-236)
            uint32_t v = 0;
            if (acc == 1u) v = e.bus_read(addr);
#pragma nounroll
            for (uint32_t c = 0; c < 3u; ++c) {
                if (c == 2u && acc == 2u) e.bus_write(addr, wval);
                e.tick();
            }
            e.pia_tick();
            // ---- the cycle's effect
            bool done = false;           // one run_one_cycle of the frame loop ended
            if (!e.busy) {
                if (acc == 0u) {
                    e.fcyc += 1u;        // WSYNC stall: one cycle (src/atari.rs:309-312)
                    done = true;
                } else {
                    e.pc = (e.pc + 1u) & 0xFFFFu;
                    const uint4 dsc = *reinterpret_cast<const uint4 *>(r_lds + kProgBytes + v * 16u);
                    e.rx = dsc.x;
                    e.d0 = dsc.y;
                    e.dg = dsc.z;
                    e.dh = dsc.w;
                    if ((e.rx & 31u) == P_ILLEGAL) {
                        st = FA_ENV_ILLEGAL_OPCODE;   // the reference panics (src/cpu.rs:1139-1141)
                        frame_end = true;
                    } else {
                        e.busy = 1;
                        e.t = 0;
                        e.cyc = 1;
                    }
                }
            } else {
                e.cyc += 1u;
                if (acc == 1u) {
                    switch ((u >> 5) & 7u) {
                    case DS_OPND: e.opnd = v; break;
                    case DS_LO: e.lo = v; break;
                    case DS_HI: e.hi = v; break;
                    case DS_DAT: e.dat = v; break;
                    case DS_P: e.p = v & ~0x30u; break;
                    default: break;
                    }
                }
                if (u & U_PCINC) e.pc = (e.pc + 1u) & 0xFFFFu;
                if (u & U_SPDEC) e.sp = (e.sp - 1u) & 0xFFu;
                if (u & U_SETI) e.p |= 0x04u;
                if (u & U_ALU) e.alu();
                const bool last = (u & U_LAST) || ((u & U_LAST_NT) && !e.taken()) ||
                                  ((u & U_LAST_NC) && ((e.pc ^ e.target()) & 0xFF00u) == 0u);
                if (last) {
                    switch ((u >> 17) & 7u) {
                    case F_JMP: e.pc = e.base(); break;
                    case F_RTS: e.pc = (e.base() + 1u) & 0xFFFFu; break;
                    case F_BRANCH: e.pc = e.target(); break;
                    case F_JMPI: e.pc = (e.hi << 8) | e.dat; break;
                    default: break;
                    }
                    e.busy = 0;
                    e.fcyc += e.cyc;
                    cpu_cyc += e.cyc;
                    instr += 1u;
                    done = true;
                } else {
                    e.t += 1u;
                }
            }
            // ---- run_frame's caps after each run_one_cycle (src/atari.rs:285-300)
            if (done) {
                if (e.in_a) {
                    if (e.fcyc > kLoopACap) {
                        st = FA_ENV_VSYNC_HELD;
                        frame_end = true;
                    }
                } else if (e.fcyc > kLoopBCap) {
                    frame_end = true;
                }
            }
        }
        if (frame_end) {
            fcyc_last = e.fcyc;
            f += 1u;
            if (f < g.k && st == FA_ENV_OK) {
                if (g.actions) set_action(e, g.actions[(size_t)f * g.stride + idx]);
                e.fcomp = 0;
                e.fcyc = 0;
                e.in_a = (e.vsync & 2u) ? 1u : 0u;
            } else {
                live = false;
            }
        }
    }
    // A stopped env still takes each frame's action, and its frames run no
    // cycles (the handle keeps stepping it, frozen).
    if (st != FA_ENV_OK && f < g.k) {
        fcyc_last = 0;
        if (g.actions) set_action(e, g.actions[(size_t)(g.k - 1u) * g.stride + idx]);
    }
    while (e.zq < kChunks) e.zero_step();
    e.fold_collisions();
    e.flush_group();
    store_lane(e, st_ptr, scheme);
    g.cpu_cycles[idx] = cpu_cyc;
    g.instr[idx] = instr;
    g.frame_cycles[idx] = fcyc_last;
    g.status[idx] = st;
}

}  // namespace rnd
}  // namespace fa
