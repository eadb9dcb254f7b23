// fa_frame.hpp -- the frame-step kernel (template) shared by the two
// translation units that instantiate it: fa_kernels.hip (the plain path) and
// fa_kernels_sorted.hip (wave64 opcode grouping).  They are compiled with
// different code-generation options; see fastatari/build.py.
#pragma once

#include <hip/hip_runtime.h>

#include "fa_emu.hpp"
#include "fa_internal.hpp"
#include "fa_uniform.hpp"

namespace fa {

// HeadlessAtari::run_frame (src/atari.rs:584-614).  Returns the frame's
// cycle count; sets `st` on an illegal opcode or a held VSYNC.
//
// Whenever the active lanes agree on the control state (the frame's cycle count
// aside, see u_enter) the wave runs the
// loop in the wave-uniform mode (fa_uniform.hpp) until the frame ends or an
// instruction would make the lanes disagree; the per-lane loop below runs that
// instruction and everything else.  Before each per-lane instruction a cheap
// test (PC and TIA clock) decides whether the full test is worth making; a
// failed full test backs off for a growing number of instructions.
//
// The reference's loop tests, before each instruction, whether VSYNC still
// agrees with the loop (loop A while it is held, loop B until it rises) and
// whether WSYNC is set, and after it the cycle cap.  Here one test covers
// both rare conditions before the instruction and one the cap and an illegal
// opcode after it: neither VSYNC nor WSYNC changes but by a TIA write, so the
// common step runs no exec-mask branching of its own.  A WSYNC skip is
// followed by the next instruction in the same iteration, which is the
// reference's next step once its caps are checked (the skip leaves VSYNC as
// it was and WSYNC clear).
//
// With SORTED (opcode grouping), a step whose wave holds three or more
// distinct opcodes backs the grouping test off like the uniform-mode test
// (1, 2, 4 .. 64 steps), and the steps in between run the plain kernel's
// per-lane step itself, so a divergent phase pays neither the group test nor
// the grouped dispatch's merge with the general path.
template <bool SORTED>
__device__ __forceinline__ uint32_t run_frame(Emu &e, uint32_t &st, uint32_t &cpu_cyc, uint32_t &instr) {
    e.fcomp = 0;
    uint32_t cycles = 0;
    bool in_a = (e.vsync & 2u) != 0u;
    uint32_t cap = in_a ? LOOPA_CAP : LOOPB_CAP;
    uint32_t ina2 = in_a ? 2u : 0u;      // in_a as VSYNC's bit 1
    uint32_t hold = 0, backoff = 0;      // wave-uniform
    uint32_t ghold = 0, gback = 0;       // wave-uniform: grouping attempts (SORTED)
    uint32_t op, b1, b2;                 // ROM bytes at PC, read ahead
    e.fetch_ahead(op, b1, b2);
    while (true) {
        if (hold == 0u) {
            if (u_maybe(e)) {
                UCtl u;
                if (u_enter(e, cycles, in_a, u)) {
                    const uint32_t r = u_run(u, e, st);
                    const bool moved = u.cycles != u.c0;
                    u_leave(u, e, cycles, in_a, cpu_cyc, instr);
                    if (r == U_DONE) break;
                    cap = in_a ? LOOPA_CAP : LOOPB_CAP;
                    ina2 = in_a ? 2u : 0u;
                    e.fetch_ahead(op, b1, b2);
                    // declined at once: back off as after a failed test
                    backoff = moved ? 0u : (backoff ? (backoff < 64u ? 2u * backoff : 64u) : 1u);
                    hold = backoff;
                } else {
                    backoff = backoff ? (backoff < 64u ? 2u * backoff : 64u) : 1u;
                    hold = backoff;
                }
            }
        } else {
            --hold;
        }
        // one test for two rare events: WSYNC, and VSYNC disagreeing with
        // the loop (loop A runs while VSYNC is held, loop B until it rises)
        const uint32_t v2 = e.vsync & 2u;
        if (__builtin_expect((e.wsync | (v2 ^ ina2)) != 0u, 0)) {
            if (!in_a & (v2 != 0u)) break;             // VSYNC rose in loop B: the frame ends
            if (in_a & (v2 == 0u)) {                    // VSYNC dropped: loop B from here
                in_a = false;
                ina2 = 0u;
                cap = LOOPB_CAP;
            }
            if (e.wsync != 0u) {
                // the skip, then the caps; VSYNC is unchanged and WSYNC clear,
                // so the loop's next step is the instruction below
                const uint32_t s = e.tia_skip();
                if (s == 0u) { st = ENV_VSYNC_HELD; break; }   // only a clock >= 228 could stall here
                e.pia_tick(s);
                cycles += s;
                if (cycles > cap) {
                    st = in_a ? ENV_VSYNC_HELD : st;
                    break;
                }
            }
        }
        uint32_t c;
        if constexpr (SORTED) {
            if (ghold == 0u) {
                bool grouped = false;
                c = e.exec<true>(op, b1, b2, &grouped);
                gback = grouped ? 0u : (gback ? (gback < 64u ? 2u * gback : 64u) : 1u);
                ghold = gback;
            } else {
                --ghold;
                c = e.exec<false>(op, b1, b2);
            }
        } else {
            c = e.exec<false>(op, b1, b2);
        }
        e.fetch_ahead(op, b1, b2);
        cycles += c;
        cpu_cyc += c;
        instr += 1u;
        e.tia_tick(c);
        e.pia_tick(c);
        if (__builtin_expect((c == 0u) | (cycles > cap), 0)) {
            // an illegal opcode (0 cycles) changes nothing above and ends the
            // frame; loop A's cap leaves VSYNC held
            instr -= c == 0u ? 1u : 0u;
            st = c == 0u ? ENV_ILLEGAL : st;
            st = in_a & (cycles > cap) ? ENV_VSYNC_HELD : st;
            break;
        }
    }
    return cycles;
}

// RAM bytes of an LDS image (2-byte aligned) to 128 bytes of global memory.
__device__ __forceinline__ void put_obs(uint8_t *dst, uint32_t im) {
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    for (uint32_t part = 0; part < 8u; ++part) {
        const uint32_t off = im + S_RAM + part * 16u;
        uint32_t w[4];
        for (int j = 0; j < 4; ++j) w[j] = ld16(off + 4u * j) | (ld16(off + 4u * j + 2u) << 16);
        d[part] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

template <bool SORTED>
__global__ void __launch_bounds__(256) fa_frame_kernel(StepArgs g) {
    const uint32_t bs = blockDim.x;
    const uint32_t tid = threadIdx.x;
    uint32_t first, count;
    const uint32_t *pk_in = nullptr;   // packed inputs of this workgroup's envs
    uint8_t *obs_out = nullptr;        // obs of this workgroup's envs (last step)
    if (g.htasks) {
        const HostTask t = g.htasks[blockIdx.x];
        first = t.first;
        count = t.count;
        pk_in = t.in;
        obs_out = t.obs;
    } else if (g.tasks) {
        first = g.tasks[2u * blockIdx.x];
        count = g.tasks[2u * blockIdx.x + 1u];
    } else {
        first = blockIdx.x * bs;
        count = g.n > first ? (g.n - first < bs ? g.n - first : bs) : 0u;
    }
    if (!g.htasks) {
        if (g.packed) pk_in = g.packed + first;
        if (g.obs && !g.obs_every) obs_out = g.obs + (size_t)first * 128u;
    }
    // The AoS block is staged in whole 16-byte granules: 8 envs = 1776 bytes.
    const uint32_t staged = (count + 7u) & ~7u;
    const uint32_t img0 = ROM0 + g.rom_bytes;
    const uint32_t img_bytes = staged * STATE_BYTES;
    const uint32_t junk0 = img0 + bs * STATE_BYTES;

    // Stage the descriptors, the ROM and this workgroup's state block into
    // LDS (16-byte loads).
    {
        const uint4 *dsrc = reinterpret_cast<const uint4 *>(g.desc);
        uint4 *ddst = reinterpret_cast<uint4 *>(fa_lds + DESC0);
        for (uint32_t i = tid; i < kDescBytes / 16u; i += bs) ddst[i] = dsrc[i];
        const uint4 *src = reinterpret_cast<const uint4 *>(g.rom);
        uint4 *dst = reinterpret_cast<uint4 *>(fa_lds + ROM0);
        for (uint32_t i = tid; i < g.rom_bytes / 16u; i += bs) dst[i] = src[i];
        const uint4 *ssrc = reinterpret_cast<const uint4 *>(g.states + (size_t)first * STATE_BYTES);
        uint4 *sdst = reinterpret_cast<uint4 *>(fa_lds + img0);
        for (uint32_t i = tid; i < img_bytes / 16u; i += bs) sdst[i] = ssrc[i];
    }
    __syncthreads();

    if (tid < count) {
        const uint32_t idx = first + tid;
        const uint32_t im = img0 + tid * STATE_BYTES;
        uint64_t cpu0, fcyc0, instr0;
        uint32_t st;
        const uint32_t pk = pk_in ? pk_in[tid] : 0u;
        const bool restart = pk_in ? ((pk >> 16) & 0xFFu) != 0u : (g.reset && g.reset[idx]);
        if (restart) {
            // restart: the ROM's reset snapshot and its counters
            const uint16_t *src = reinterpret_cast<const uint16_t *>(g.reset_state);
            for (uint32_t j = 0; j < STATE_BYTES / 2u; ++j) st16(im + 2u * j, src[j]);
            cpu0 = g.reset_counters[0];
            fcyc0 = g.reset_counters[1];
            instr0 = g.reset_counters[2];
            st = ENV_OK;
        } else {
            cpu0 = g.cpu_cycles[idx];
            fcyc0 = g.frame_cycles[idx];
            instr0 = g.instr[idx];
            st = g.status[idx];
        }
        Emu e;
        e.load(im, junk0 + tid, g.hot_lo, g.hot_n);
        uint32_t cpu_cyc = 0, instr = 0;
        uint64_t fcyc = fcyc0;
        for (uint32_t s = 0; s < g.k; ++s) {
            const size_t at = (size_t)s * g.stride + idx;
            uint32_t nfr;
            if (pk_in) {
                if (!(pk & (1u << 24))) e.set_action(pk & 0xFFu);
                nfr = (pk >> 8) & 0xFFu;
            } else {
                if (g.actions) e.set_action(g.actions[at]);
                nfr = g.fs ? (uint32_t)g.fs[at] : 1u;
            }
            for (uint32_t f = 0; f < nfr; ++f) {
                if (st != ENV_OK) { fcyc = 0; continue; }
                fcyc = run_frame<SORTED>(e, st, cpu_cyc, instr);
            }
            if (g.obs && g.obs_every) {
                e.store();
                put_obs(g.obs + at * 128u, im);
            }
        }
        e.store();
        g.cpu_cycles[idx] = cpu0 + cpu_cyc;
        g.frame_cycles[idx] = fcyc;
        g.instr[idx] = instr0 + instr;
        g.status[idx] = st;
    }
    __syncthreads();

    {
        uint4 *sdst = reinterpret_cast<uint4 *>(g.states + (size_t)first * STATE_BYTES);
        const uint4 *ssrc = reinterpret_cast<const uint4 *>(fa_lds + img0);
        for (uint32_t i = tid; i < img_bytes / 16u; i += bs) sdst[i] = ssrc[i];
    }
    if (obs_out) {
        // obs of env j = RAM bytes 8..135 of its image; 8 lanes per env
        for (uint32_t c = tid; c < count * 8u; c += bs) {
            const uint32_t env = c >> 3, part = c & 7u;
            const uint32_t off = img0 + env * STATE_BYTES + S_RAM + part * 16u;
            uint32_t w[4];
            for (int j = 0; j < 4; ++j) w[j] = ld16(off + 4u * j) | (ld16(off + 4u * j + 2u) << 16);
            reinterpret_cast<uint4 *>(obs_out + (size_t)env * 128u)[part] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

}  // namespace fa
