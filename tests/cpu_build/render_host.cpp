// Host build of the rendering console's lane code
// (fast-atari-rs_amd/csrc/fa_render_lane.hpp), for the CPU tests only: the
// lanes of each 64-lane workgroup run one after another over a host array
// standing in for the workgroup's LDS, so the kernel's per-lane logic (its
// micro-programs, TIA, RIOT, LDS layout and frame-group writes) is checked
// against the oracle without a GPU.  Not part of the library.
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define __device__
#define __forceinline__ inline
struct uint4 {
    uint32_t x, y, z, w;
};
static inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return uint4{a, b, c, d}; }

namespace fa {
namespace rnd {
alignas(16) static uint8_t r_lds[160 * 1024];
}  // namespace rnd
}  // namespace fa

#include "../../fast-atari-rs_amd/csrc/fa_render_lane.hpp"

using namespace fa::rnd;

extern "C" uint32_t rh_table_bytes(void) { return kTableBytes; }

// One launch: k frames of n envs, workgroups of bs lanes.  Returns 0, or -1
// when the workgroup's LDS would not fit.
extern "C" int rh_launch(const uint8_t *rom_img, uint32_t rom_bytes, uint32_t hot_lo, uint32_t hot_n, uint8_t *states,
                         uint64_t *cpu, uint64_t *fcyc, uint64_t *instr, uint32_t *status, const uint8_t *actions,
                         uint8_t *fb, uint32_t n, uint32_t k, uint32_t bs) {
    if (kTableBytes + rom_bytes + (size_t)bs * kLaneBytes > sizeof(r_lds)) return -1;
    static uint32_t tables[kTableBytes / 4];
    build_tables(tables);
    RArgs g{};
    g.states = states;
    g.cpu_cycles = cpu;
    g.frame_cycles = fcyc;
    g.instr = instr;
    g.status = status;
    g.actions = actions;
    g.fb = fb;
    g.tables = tables;
    g.rom = rom_img;
    g.rom_bytes = rom_bytes;
    g.hot_lo = hot_lo;
    g.hot_n = hot_n;
    g.n = n;
    g.k = k;
    g.stride = n;
    for (uint32_t first = 0; first < n; first += bs) {
        memset(r_lds, 0xA5, sizeof(r_lds));      // stale LDS: nothing may depend on it
        memcpy(r_lds, tables, kTableBytes);
        memcpy(r_lds + kTableBytes, rom_img, rom_bytes);
        const uint32_t count = n - first < bs ? n - first : bs;
        for (uint32_t tid = 0; tid < count; ++tid) render_lane(g, bs, tid, first + tid);
    }
    return 0;
}
