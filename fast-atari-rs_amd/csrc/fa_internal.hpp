// fa_internal.hpp -- host-side interfaces shared by the library's translation
// units: ROM classes, launch descriptions, the kernel launchers defined in
// fa_kernels.hip, and error reporting for the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fastatari.h"
#include "fa_decode.hpp"

namespace fa {

constexpr uint32_t kStateBytes = 222;     // AtariStateGpu
constexpr uint32_t kLdsRom0 = kDescBytes; // LDS offset of the ROM image (after the descriptor table)
constexpr uint32_t kMaxBlock = 256;

// One workgroup's share of a served launch: a client's slot range and its
// packed inputs / obs in pinned host memory, which the kernel reads and writes
// directly (no copies on the serving path).
struct HostTask {
    uint32_t first, count;        // slots (first % 8 == 0, count <= block)
    const uint32_t *in;           // count packed input words
    uint8_t *obs;                 // count * 128 bytes
};

// Arguments of one frame-step launch.  Envs are addressed by slot index into
// the AoS state array and the per-env counter arrays.
struct StepArgs {
    uint8_t *states;              // AtariStateGpu[], 222-byte stride
    uint64_t *cpu_cycles;         // Cpu.cycles (src/cpu.rs:104)
    uint64_t *frame_cycles;       // return value of the last run_frame
    uint64_t *instr;              // executed instructions
    uint32_t *status;             // FA_ENV_*
    const uint8_t *actions;       // [s * stride + slot]; NULL: no set_action
    const uint8_t *fs;            // frames per step [s * stride + slot]; NULL: 1
    uint8_t *obs;                 // RAM after the last step [slot * 128], or after every step
    uint32_t obs_every;           //   [(s * stride + slot) * 128] when non-zero; NULL: none
    const uint8_t *reset;         // [slot] != 0: restart from reset_state before stepping; NULL: none
    const uint8_t *reset_state;   // 222-byte snapshot (HeadlessAtari::new + one run_frame)
    const uint64_t *reset_counters;   // its cpu_cycles, frame_cycles, instr
    const uint32_t *packed;       // [slot]: action | frames << 8 | restart << 16 | no-action << 24
                                  //   (one step; replaces actions / fs / reset); NULL: none
    const uint32_t *tasks;        // NULL: workgroup b covers slots [b * bs, b * bs + bs) of [0, n);
                                  // else (first, count) per workgroup, first % 8 == 0, count <= bs
    const HostTask *htasks;       // one per workgroup; replaces tasks, packed and obs when set
    const uint32_t *desc;         // kDescBytes: opcode descriptors
    const uint8_t *rom;           // ROM image, rom_bytes (multiple of 4096)
    uint32_t rom_bytes;
    uint32_t hot_lo, hot_n;       // hotspot window (lo12 - hot_lo < hot_n)
    uint32_t n, k, stride;
};

// One cartridge: the image staged into LDS and its power-on state.
struct RomClass {
    std::vector<uint8_t> image;   // Fixed: the mirrored 4 KB window; else every 4 KB bank
    uint32_t rom_bytes = 0;
    int scheme = 0;               // 0 Fixed, 1 F8, 2 F6, 3 F4 (src/atari.rs:97-104)
    uint32_t hot_lo = 0xFFF, hot_n = 0, first_bank = 0;
    AtariStateGpu power_on{};
};

int make_rom_class(const uint8_t *rom, size_t rom_len, RomClass &out);
int check_state(const RomClass &rc, const AtariStateGpu &s, size_t i);

inline size_t lds_bytes(uint32_t rom_bytes, uint32_t bs) {
    return kLdsRom0 + rom_bytes + (size_t)bs * (kStateBytes + 1u);
}
uint32_t default_block(uint32_t rom_bytes, size_t max_lds);

// Kernel launchers (fa_kernels.hip).
hipError_t prepare_kernels(size_t max_lds);
hipError_t launch_step(const StepArgs &g, bool sorted, uint32_t bs, uint32_t grid, hipStream_t s);
// The opcode-grouping kernel lives in its own translation unit (fa_kernels_sorted.hip).
hipError_t prepare_sorted_kernel(size_t max_lds);
hipError_t launch_sorted(const StepArgs &g, uint32_t bs, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_fill(const AtariStateGpu &tmpl, uint8_t *states, uint64_t *cpu, uint64_t *fcyc, uint64_t *instr,
                       uint32_t *status, uint32_t n, uint32_t n_pad, hipStream_t s);
void descriptor_table(uint32_t out[kDescBytes / 4]);

// fa_lib.hip: the batch handle's stream and device obs buffer (n * 128 B; it
// holds the obs of device-side steps and of resets that return no host obs)
hipStream_t env_stream(fa_env *e);
uint8_t *env_obs_device(fa_env *e);
// Host memcpy split over up to 8 threads (parts of at least 1 MB): one core
// copies many megabytes at a fraction of the host's memory bandwidth.
void copy_host(void *dst, const void *src, size_t bytes);

// Errors: the message of the calling thread's last failure.
int fail(int code, const std::string &msg);
std::string &last_error();

}  // namespace fa

#define FA_HIP_TRY(expr)                                                                           \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) return fa::fail(FA_EHIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)
