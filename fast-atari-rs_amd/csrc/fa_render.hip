// fa_render.hip -- the batched rendering console: 160 x 192 frames of NTSC
// palette indices from the pixel-rendering TIA, one wave64 lane per env.
//
// Reference: Atari::{new, set_action, run_frame, framebuffer} and the Env
// impl's reset / step (src/atari.rs:226-378) over the rendering Bus
// (src/atari.rs:107-224) and Tia (src/tia.rs:98-675), with the CPU of
// src/cpu.rs:245-1146.  The micro-programs and the per-opcode decode are in
// fa_render.hpp.
//
// Per env in HBM: the 256-byte render state (the AtariStateGpu layout with
// the TIA clock in TIA clocks, then the rendering TIA's own fields), the
// counters, and the 30,720-byte frame.  Per workgroup in LDS: the
// micro-program / opcode tables, the ROM image, and per lane 128 bytes of RAM
// and a 16-byte pixel group (both lane-interleaved: byte i of lane l at
// ((i >> 2) * block + l) * 4 + (i & 3), so lanes of a wave touching the same
// per-lane offset hit 64 different banks).  Pixels of one row and 16-pixel
// group collect in LDS and go to HBM as one 16-byte store when the beam
// leaves the group.  The TIA's frame clear at the first visible line
// (src/tia.rs:380-382) becomes a zero fill of one 16-byte chunk per bus
// cycle, which stays ahead of the beam (10 chunks a row against 76 cycles a
// scanline).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fa_internal.hpp"
#include "fa_render.hpp"

namespace fa {
namespace rnd {

extern __shared__ __attribute__((aligned(16))) uint8_t r_lds[];

}  // namespace rnd
}  // namespace fa

#include "fa_render_lane.hpp"

namespace fa {
namespace rnd {

// One lane per env; the workgroup stages the tables and the ROM in LDS.
__global__ void __launch_bounds__(256) fa_render_kernel(RArgs g) {
    const uint32_t bs = blockDim.x, tid = threadIdx.x;
    const uint32_t first = blockIdx.x * bs;
    const uint32_t count = g.n > first ? (g.n - first < bs ? g.n - first : bs) : 0u;
    {
        const uint4 *ts = reinterpret_cast<const uint4 *>(g.tables);
        uint4 *td = reinterpret_cast<uint4 *>(r_lds);
        for (uint32_t i = tid; i < kTableBytes / 16u; i += bs) td[i] = ts[i];
        const uint4 *rs = reinterpret_cast<const uint4 *>(g.rom);
        uint4 *rdst = reinterpret_cast<uint4 *>(r_lds + kTableBytes);
        for (uint32_t i = tid; i < g.rom_bytes / 16u; i += bs) rdst[i] = rs[i];
    }
    __syncthreads();
    if (tid >= count) return;
    render_lane(g, bs, tid, first + tid);
}

// Broadcast of the power-on state (16 x 16 bytes) with zero counters and status.
__global__ void fa_render_fill_kernel(const uint4 *tmpl, uint8_t *states, uint64_t *cpu, uint64_t *fcyc,
                                      uint64_t *instr, uint32_t *status, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 *d = reinterpret_cast<uint4 *>(states + (size_t)i * kRState);
    for (int q = 0; q < 16; ++q) d[q] = tmpl[q];
    cpu[i] = 0;
    fcyc[i] = 0;
    instr[i] = 0;
    status[i] = 0;
}

// Zero fill of n16 16-byte chunks (the frames of a fresh console).
__global__ void fa_render_zero_kernel(uint4 *p, size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

}  // namespace rnd
}  // namespace fa

// ======================================================================== host

struct fa_render {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t n = 0;
    fa::RomClass rom;
    uint32_t block = 0;
    size_t max_lds = 65536;
    uint8_t tmpl[FA_RENDER_STATE_BYTES] = {};

    uint32_t *d_tables = nullptr;
    uint8_t *d_rom = nullptr, *d_states = nullptr, *d_frames = nullptr, *d_actions = nullptr, *d_tmpl = nullptr;
    uint64_t *d_cpu = nullptr, *d_fcyc = nullptr, *d_instr = nullptr;
    uint32_t *d_status = nullptr;
    size_t actions_cap = 0;

    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t events_used = 0;
    double ms_done = 0.0;
    uint64_t launches_done = 0;
};

namespace {

using fa::fail;
namespace rnd = fa::rnd;

size_t render_lds(uint32_t rom_bytes, uint32_t bs) { return rnd::kTableBytes + rom_bytes + (size_t)bs * rnd::kLaneBytes; }

uint32_t render_block(const fa_render *r) { return r->block ? r->block : 64u; }

void put16(uint8_t *o, size_t off, uint32_t v) {
    o[off] = (uint8_t)v;
    o[off + 1] = (uint8_t)(v >> 8);
}
uint32_t get16(const uint8_t *o, size_t off) { return (uint32_t)o[off] | ((uint32_t)o[off + 1] << 8); }

// Atari::new (src/atari.rs:238-246): Tia::default with paddles at 128, the
// ideal visible top and open input latches (src/tia.rs:191-240),
// Pia::default (src/pia.rs:46-62), the cartridge's first bank, and
// Cpu::reset (src/cpu.rs:125-134), whose two reset-vector reads go through
// the rendering bus: 6 TIA clocks and 2 RIOT clocks, which leave the timer at
// $FE with the underflow flag set.
void render_power_on(const fa::RomClass &rc, uint8_t *s) {
    std::memset(s, 0, FA_RENDER_STATE_BYTES);
    s[3] = 0xFD;                              // SP
    put16(s, 4, rc.power_on.pc);              // the reset vector of the first bank
    s[138] = 0xFF;                            // port A input
    s[141] = 0x3F;                            // port B input
    put16(s, 142, 0xFE);                      // timer value
    put16(s, 144, 1);                         // prescaler
    put16(s, 146, 1);                         // prescaler select
    s[148] = 1;                               // underflow
    put16(s, 152, 6);                         // TIA clock
    s[168] = 1;                               // INPT4, INPT5 released
    s[169] = 1;
    s[170] = 128;                             // paddles centred
    s[171] = 128;
    s[217] = (uint8_t)rc.first_bank;
    s[218] = (uint8_t)rc.scheme;
    put16(s, 224, 32);                        // visible_top
    put16(s, 226, 262);                       // vblank_top
    put16(s, 228, 32);                        // pending_top
    s[232] = 0x80;                            // input latches
    s[233] = 0x80;
}

int check_render_state(const fa_render *r, const uint8_t *s, size_t i) {
    auto bad = [&](const char *what) { return fail(FA_ESTATE, "render state " + std::to_string(i) + ": " + what); };
    const uint32_t sel = get16(s, 146);
    if (sel != 1 && sel != 8 && sel != 64 && sel != 1024) return bad("timer select must be 1, 8, 64 or 1024");
    if (get16(s, 144) > sel) return bad("timer prescaler above select");
    if (get16(s, 142) > 255) return bad("timer value above 255");
    if (get16(s, 152) >= 228) return bad("TIA clock must be below 228");
    if (get16(s, 154) >= 262) return bad("scanline must be below 262");
    if (s[218] != r->rom.scheme) return bad("scheme differs from the ROM's");
    if (s[217] >= r->rom.rom_bytes / 4096u) return bad("bank out of range for the ROM");
    for (size_t o = 176; o < 186; o += 2)
        if (get16(s, o) >= 160) return bad("object position above 159");
    const uint32_t vt = get16(s, 224), pt = get16(s, 228);
    if (vt < 23 || vt > 32 || pt < 23 || pt > 32) return bad("visible / pending top outside 23..32");
    if (get16(s, 226) > 262) return bad("vblank_top above 262");
    if (s[230] > 3) return bad("pending_frames above 3");
    return FA_OK;
}

int ensure_actions(fa_render *r, size_t bytes) {
    if (bytes <= r->actions_cap) return FA_OK;
    if (r->d_actions) FA_HIP_TRY(hipFree(r->d_actions));
    r->d_actions = nullptr;
    r->actions_cap = 0;
    if (hipMalloc(&r->d_actions, bytes) != hipSuccess) return fail(FA_ENOMEM, "device allocation failed");
    r->actions_cap = bytes;
    return FA_OK;
}

int launch_render(fa_render *r, const uint8_t *d_actions, uint32_t k) {
    if (r->n == 0) return FA_OK;
    rnd::RArgs g{};
    g.states = r->d_states;
    g.cpu_cycles = r->d_cpu;
    g.frame_cycles = r->d_fcyc;
    g.instr = r->d_instr;
    g.status = r->d_status;
    g.actions = d_actions;
    g.fb = r->d_frames;
    g.tables = r->d_tables;
    g.rom = r->d_rom;
    g.rom_bytes = r->rom.rom_bytes;
    g.hot_lo = r->rom.hot_lo;
    g.hot_n = r->rom.hot_n;
    g.n = (uint32_t)r->n;
    g.k = k;
    g.stride = (uint32_t)r->n;
    const uint32_t bs = render_block(r);
    const size_t lds = render_lds(r->rom.rom_bytes, bs);
    if (lds > r->max_lds) return fail(FA_EINVAL, "LDS budget exceeded for this block size");
    const uint32_t grid = (uint32_t)((r->n + bs - 1) / bs);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (r->timing) {
        if (r->events_used == r->events.size()) {
            hipEvent_t a, b;
            FA_HIP_TRY(hipEventCreate(&a));
            FA_HIP_TRY(hipEventCreate(&b));
            r->events.emplace_back(a, b);
        }
        ev0 = r->events[r->events_used].first;
        ev1 = r->events[r->events_used].second;
        r->events_used++;
        FA_HIP_TRY(hipEventRecord(ev0, r->stream));
    }
    hipLaunchKernelGGL(rnd::fa_render_kernel, dim3(grid), dim3(bs), lds, r->stream, g);
    FA_HIP_TRY(hipGetLastError());
    if (r->timing) FA_HIP_TRY(hipEventRecord(ev1, r->stream));
    return FA_OK;
}

int fill_power_on(fa_render *r) {
    if (r->n == 0) return FA_OK;
    const uint32_t n = (uint32_t)r->n;
    hipLaunchKernelGGL(rnd::fa_render_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, r->stream,
                       reinterpret_cast<const uint4 *>(r->d_tmpl), r->d_states, r->d_cpu, r->d_fcyc, r->d_instr,
                       r->d_status, n);
    FA_HIP_TRY(hipGetLastError());
    const size_t n16 = r->n * (size_t)rnd::kChunks;
    hipLaunchKernelGGL(rnd::fa_render_zero_kernel, dim3(2048), dim3(256), 0, r->stream,
                       reinterpret_cast<uint4 *>(r->d_frames), n16);
    FA_HIP_TRY(hipGetLastError());
    return FA_OK;
}

}  // namespace

extern "C" {

int fa_render_create(const uint8_t *rom, size_t rom_len, size_t n, fa_render **out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    *out = nullptr;
    if (n > FA_RENDER_MAX_ENVS) return fail(FA_EINVAL, "n above FA_RENDER_MAX_ENVS");
    fa::RomClass rc;
    int rc_ = fa::make_rom_class(rom, rom_len, rc);
    if (rc_) return rc_;
    auto *r = new (std::nothrow) fa_render();
    if (!r) return fail(FA_ENOMEM, "out of host memory");
    r->n = n;
    r->rom = std::move(rc);
    render_power_on(r->rom, r->tmpl);
    auto cleanup = [&](int code) {
        fa_render_destroy(r);
        return code;
    };
    if (hipGetDevice(&r->device) != hipSuccess) return cleanup(fail(FA_EHIP, "no HIP device"));
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, r->device) == hipSuccess && lds > 0)
        r->max_lds = (size_t)lds;
    if (render_lds(r->rom.rom_bytes, 64) > r->max_lds) return cleanup(fail(FA_EINVAL, "ROM too large for LDS"));
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&rnd::fa_render_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)r->max_lds) != hipSuccess)
        return cleanup(fail(FA_EHIP, "kernel attribute setup failed"));
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FA_EHIP, "hipStreamCreate failed"));
    const size_t m = n ? n : 1;
    if (hipMalloc(&r->d_tables, rnd::kTableBytes) != hipSuccess || hipMalloc(&r->d_rom, r->rom.rom_bytes) != hipSuccess ||
        hipMalloc(&r->d_tmpl, FA_RENDER_STATE_BYTES) != hipSuccess ||
        hipMalloc(&r->d_states, m * FA_RENDER_STATE_BYTES) != hipSuccess ||
        hipMalloc(&r->d_frames, m * (size_t)rnd::kFrameBytes) != hipSuccess ||
        hipMalloc(&r->d_cpu, m * 8) != hipSuccess || hipMalloc(&r->d_fcyc, m * 8) != hipSuccess ||
        hipMalloc(&r->d_instr, m * 8) != hipSuccess || hipMalloc(&r->d_status, m * 4) != hipSuccess)
        return cleanup(fail(FA_ENOMEM, "device allocation failed"));
    std::vector<uint32_t> tables(rnd::kTableBytes / 4);
    rnd::build_tables(tables.data());
    if (hipMemcpy(r->d_rom, r->rom.image.data(), r->rom.rom_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->d_tables, tables.data(), rnd::kTableBytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->d_tmpl, r->tmpl, FA_RENDER_STATE_BYTES, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FA_EHIP, "table / ROM upload failed"));
    if (fill_power_on(r) || hipStreamSynchronize(r->stream) != hipSuccess)
        return cleanup(fail(FA_EHIP, "power-on fill failed"));
    *out = r;
    return FA_OK;
}

void fa_render_destroy(fa_render *r) {
    if (!r) return;
    if (r->stream) hipStreamSynchronize(r->stream);
    for (auto &p : r->events) {
        hipEventDestroy(p.first);
        hipEventDestroy(p.second);
    }
    for (void *p : {(void *)r->d_tables, (void *)r->d_rom, (void *)r->d_tmpl, (void *)r->d_states,
                    (void *)r->d_frames, (void *)r->d_actions, (void *)r->d_cpu, (void *)r->d_fcyc,
                    (void *)r->d_instr, (void *)r->d_status})
        if (p) hipFree(p);
    if (r->stream) hipStreamDestroy(r->stream);
    delete r;
}

size_t fa_render_num_envs(const fa_render *r) { return r ? r->n : 0; }

int fa_render_reset(fa_render *r, uint8_t *frames_out) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    int rc = fill_power_on(r);
    if (rc) return rc;
    rc = launch_render(r, nullptr, 1);
    if (rc) return rc;
    if (frames_out && r->n)
        FA_HIP_TRY(hipMemcpyAsync(frames_out, r->d_frames, r->n * (size_t)rnd::kFrameBytes, hipMemcpyDeviceToHost,
                                  r->stream));
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_step(fa_render *r, const uint8_t *actions, size_t n_actions, uint32_t k, uint8_t *frames_out) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    if (n_actions != (size_t)k * r->n) return fail(FA_EINVAL, "actions length must be k * num_envs");
    if (!actions && n_actions) return fail(FA_EINVAL, "actions is NULL");
    int rc = ensure_actions(r, n_actions);
    if (rc) return rc;
    if (n_actions) FA_HIP_TRY(hipMemcpyAsync(r->d_actions, actions, n_actions, hipMemcpyHostToDevice, r->stream));
    rc = launch_render(r, r->d_actions, k);
    if (rc) return rc;
    if (frames_out && r->n)
        FA_HIP_TRY(hipMemcpyAsync(frames_out, r->d_frames, r->n * (size_t)rnd::kFrameBytes, hipMemcpyDeviceToHost,
                                  r->stream));
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_step_device(fa_render *r, const uint8_t *d_actions, uint32_t k) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    return launch_render(r, d_actions, k);
}

int fa_render_sync(fa_render *r) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_frames_device(fa_render *r, uint8_t **d_frames) {
    if (!r || !d_frames) return fail(FA_EINVAL, "NULL argument");
    *d_frames = r->d_frames;
    return FA_OK;
}

int fa_render_download_frames(fa_render *r, uint8_t *out, size_t n) {
    if (!r || !out) return fail(FA_EINVAL, "NULL argument");
    if (n != r->n) return fail(FA_EINVAL, "n must be num_envs");
    FA_HIP_TRY(hipMemcpyAsync(out, r->d_frames, n * (size_t)rnd::kFrameBytes, hipMemcpyDeviceToHost, r->stream));
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_upload_frames(fa_render *r, const uint8_t *in, size_t n) {
    if (!r || !in) return fail(FA_EINVAL, "NULL argument");
    if (n != r->n) return fail(FA_EINVAL, "n must be num_envs");
    FA_HIP_TRY(hipMemcpyAsync(r->d_frames, in, n * (size_t)rnd::kFrameBytes, hipMemcpyHostToDevice, r->stream));
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_download_states(fa_render *r, uint8_t *out, size_t n) {
    if (!r || !out) return fail(FA_EINVAL, "NULL argument");
    if (n != r->n) return fail(FA_EINVAL, "n must be num_envs");
    FA_HIP_TRY(hipMemcpyAsync(out, r->d_states, n * (size_t)FA_RENDER_STATE_BYTES, hipMemcpyDeviceToHost, r->stream));
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_upload_states(fa_render *r, const uint8_t *in, size_t n) {
    if (!r || !in) return fail(FA_EINVAL, "NULL argument");
    if (n != r->n) return fail(FA_EINVAL, "n must be num_envs");
    for (size_t i = 0; i < n; ++i) {
        int rc = check_render_state(r, in + i * FA_RENDER_STATE_BYTES, i);
        if (rc) return rc;
    }
    std::vector<uint32_t> zero(n, 0u);
    FA_HIP_TRY(hipMemcpyAsync(r->d_states, in, n * (size_t)FA_RENDER_STATE_BYTES, hipMemcpyHostToDevice, r->stream));
    FA_HIP_TRY(hipMemcpyAsync(r->d_status, zero.data(), n * 4, hipMemcpyHostToDevice, r->stream));
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_download_counters(fa_render *r, uint64_t *cpu_cycles, uint64_t *frame_cycles, uint64_t *instructions,
                                uint32_t *status) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    const size_t n = r->n;
    if (cpu_cycles) FA_HIP_TRY(hipMemcpyAsync(cpu_cycles, r->d_cpu, n * 8, hipMemcpyDeviceToHost, r->stream));
    if (frame_cycles) FA_HIP_TRY(hipMemcpyAsync(frame_cycles, r->d_fcyc, n * 8, hipMemcpyDeviceToHost, r->stream));
    if (instructions) FA_HIP_TRY(hipMemcpyAsync(instructions, r->d_instr, n * 8, hipMemcpyDeviceToHost, r->stream));
    if (status) FA_HIP_TRY(hipMemcpyAsync(status, r->d_status, n * 4, hipMemcpyDeviceToHost, r->stream));
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    return FA_OK;
}

int fa_render_power_on_state(const uint8_t *rom, size_t rom_len, uint8_t *out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    fa::RomClass rc;
    int rc_ = fa::make_rom_class(rom, rom_len, rc);
    if (rc_) return rc_;
    render_power_on(rc, out);
    return FA_OK;
}

int fa_render_set_block_size(fa_render *r, uint32_t block) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    if (block != 0 && (block % 64 != 0 || block > 256)) return fail(FA_EINVAL, "block must be 0, 64, 128, 192 or 256");
    if (block && render_lds(r->rom.rom_bytes, block) > r->max_lds)
        return fail(FA_EINVAL, "block too large for the LDS budget with this ROM");
    r->block = block;
    return FA_OK;
}

int fa_render_get_block_size(const fa_render *r, uint32_t *block) {
    if (!r || !block) return fail(FA_EINVAL, "NULL argument");
    *block = render_block(r);
    return FA_OK;
}

int fa_render_set_timing(fa_render *r, int enabled) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    r->timing = enabled != 0;
    return FA_OK;
}

int fa_render_kernel_time(fa_render *r, double *total_ms, uint64_t *launches, int reset) {
    if (!r) return fail(FA_EINVAL, "NULL handle");
    FA_HIP_TRY(hipStreamSynchronize(r->stream));
    for (size_t i = 0; i < r->events_used; ++i) {
        float ms = 0.f;
        FA_HIP_TRY(hipEventElapsedTime(&ms, r->events[i].first, r->events[i].second));
        r->ms_done += ms;
        r->launches_done++;
    }
    r->events_used = 0;
    if (total_ms) *total_ms = r->ms_done;
    if (launches) *launches = r->launches_done;
    if (reset) {
        r->ms_done = 0.0;
        r->launches_done = 0;
    }
    return FA_OK;
}

}  // extern "C"
