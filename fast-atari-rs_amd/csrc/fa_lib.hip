// fa_lib.hip -- frame-step kernels and the C ABI (include/fastatari.h).
//
// Kernels (one lane per env, workgroups of `block` lanes):
//   fa_frame_kernel<false>  the plain drop-in path: convergent waves take
//                           the opcode-specialised fast path, divergent
//                           waves the general per-lane path
//   fa_frame_kernel<true>   wave64 opcode grouping (set_sorted): lanes are
//                           served one distinct opcode at a time (64-bit
//                           ballot + readlane), every group on the
//                           opcode-specialised fast path
//   fa_fill_kernel          power-on broadcast used by reset
// The same frame kernel serves step (k = 1), step_multi (k frames with state
// kept in VGPRs/LDS across frames) and frameskip (per-env frame counts).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastatari.h"
#include "fa_emu.hpp"

static_assert(sizeof(AtariStateGpu) == 222, "AtariStateGpu must be 222 bytes");
static_assert(offsetof(AtariStateGpu, ram) == fa::S_RAM, "ram offset");
static_assert(offsetof(AtariStateGpu, timer_value) == fa::S_TVAL, "timer offset");
static_assert(offsetof(AtariStateGpu, tia_clock) == fa::S_CLOCK, "clock offset");
static_assert(offsetof(AtariStateGpu, pos_p0) == fa::S_POS, "pos offset");
static_assert(offsetof(AtariStateGpu, grp0) == fa::S_GRP0, "grp0 offset");
static_assert(offsetof(AtariStateGpu, colup0) == fa::S_COLUP0, "colup0 offset");
static_assert(offsetof(AtariStateGpu, bank) == fa::S_BANK, "bank offset");

namespace fa {


struct StepArgs {
    uint8_t *states;            // n_pad x 222, AtariStateGpu
    uint64_t *cpu_cycles;       // n
    uint64_t *frame_cycles;     // n
    uint64_t *instr;            // n
    uint32_t *status;           // n
    const uint8_t *actions;     // [f * stride + i], or [i] with fs; NULL: no set_action
    const uint8_t *fs;          // per-env frame counts (frameskip), or NULL
    uint8_t *obs;               // n x 128, or NULL
    const uint32_t *desc;       // 512 words: opcode descriptors (fa_decode.hpp)
    const uint8_t *rom;         // ROM image (rom_bytes, multiple of 4096)
    uint32_t rom_bytes;
    uint32_t hot_lo, hot_n;     // hotspot window (lo12 - hot_lo < hot_n)
    uint32_t n, k, stride;
};

// One CPU instruction for the active lanes.  Returns cycles (0 = illegal).
template <bool SORTED>
__device__ __forceinline__ uint32_t cpu_step(Emu &e) {
    return e.exec<SORTED>();
}

// HeadlessAtari::run_frame (src/atari.rs:584-614).  Returns the frame's
// cycle count; sets `st` on an illegal opcode or a held VSYNC.
template <bool SORTED>
__device__ __forceinline__ uint32_t run_frame(Emu &e, uint32_t &st, uint32_t &cpu_cyc, uint32_t &instr) {
    e.fcomp = 0;
    uint32_t cycles = 0;
    bool in_a = (e.vsync & 2u) != 0u;
    while (true) {
        if (in_a && !(e.vsync & 2u)) in_a = false;
        if (!in_a && (e.vsync & 2u)) break;
        if (e.wsync) {
            uint32_t s = e.tia_skip();
            if (s == 0u) { st = ENV_VSYNC_HELD; break; }   // only a clock >= 228 could stall here
            e.pia_tick(s);
            cycles += s;
        } else {
            uint32_t c = cpu_step<SORTED>(e);
            if (c == 0u) { st = ENV_ILLEGAL; break; }
            cycles += c;
            cpu_cyc += c;
            instr += 1u;
            e.tia_tick(c);
            e.pia_tick(c);
        }
        if (in_a) {
            if (cycles > LOOPA_CAP) { st = ENV_VSYNC_HELD; break; }
        } else if (cycles > LOOPB_CAP) {
            break;
        }
    }
    return cycles;
}

template <bool SORTED>
__global__ void __launch_bounds__(256) fa_frame_kernel(StepArgs g) {
    const uint32_t bs = blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t first = blockIdx.x * bs;
    const uint32_t img0 = ROM0 + g.rom_bytes;
    const uint32_t img_bytes = bs * STATE_BYTES;

    // Stage the descriptors, the ROM and this workgroup's state block into
    // LDS (16-byte loads).
    {
        const uint4 *dsrc = reinterpret_cast<const uint4 *>(g.desc);
        uint4 *ddst = reinterpret_cast<uint4 *>(fa_lds + DESC0);
        for (uint32_t i = tid; i < 2048u / 16u; i += bs) ddst[i] = dsrc[i];
        const uint4 *src = reinterpret_cast<const uint4 *>(g.rom);
        uint4 *dst = reinterpret_cast<uint4 *>(fa_lds + ROM0);
        for (uint32_t i = tid; i < g.rom_bytes / 16u; i += bs) dst[i] = src[i];
        const uint4 *ssrc = reinterpret_cast<const uint4 *>(g.states + (size_t)first * STATE_BYTES);
        uint4 *sdst = reinterpret_cast<uint4 *>(fa_lds + img0);
        for (uint32_t i = tid; i < img_bytes / 16u; i += bs) sdst[i] = ssrc[i];
    }
    __syncthreads();

    const uint32_t idx = first + tid;
    if (idx < g.n) {
        Emu e;
        e.load(img0 + tid * STATE_BYTES, img0 + img_bytes + tid, g.hot_lo, g.hot_n);
        uint32_t st = g.status[idx];
        uint32_t nfr = g.fs ? (uint32_t)g.fs[idx] : g.k;
        uint32_t cpu_cyc = 0, instr = 0, fcyc = 0;
        for (uint32_t f = 0; f < nfr; ++f) {
            if (g.actions) e.set_action(g.fs ? g.actions[idx] : g.actions[(size_t)f * g.stride + idx]);
            if (st != ENV_OK) { fcyc = 0; continue; }
            fcyc = run_frame<SORTED>(e, st, cpu_cyc, instr);
        }
        e.store();
        g.cpu_cycles[idx] += cpu_cyc;
        g.frame_cycles[idx] = fcyc;
        g.instr[idx] += instr;
        g.status[idx] = st;
    }
    __syncthreads();

    {
        uint4 *sdst = reinterpret_cast<uint4 *>(g.states + (size_t)first * STATE_BYTES);
        const uint4 *ssrc = reinterpret_cast<const uint4 *>(fa_lds + img0);
        for (uint32_t i = tid; i < img_bytes / 16u; i += bs) sdst[i] = ssrc[i];
    }
    if (g.obs) {
        // obs[env] = RAM bytes 8..135 of the env's image (2-byte aligned in LDS)
        const uint32_t nv = g.n > first ? (g.n - first < bs ? g.n - first : bs) : 0u;
        for (uint32_t c = tid; c < nv * 8u; c += bs) {
            const uint32_t env = c >> 3, part = c & 7u;
            const uint32_t off = img0 + env * STATE_BYTES + S_RAM + part * 16u;
            uint32_t w[4];
            for (int j = 0; j < 4; ++j) w[j] = ld16(off + 4u * j) | (ld16(off + 4u * j + 2u) << 16);
            reinterpret_cast<uint4 *>(g.obs + (size_t)(first + env) * 128u)[part] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

struct FillArgs {
    uint8_t tmpl[224];
    uint8_t *states;
    uint64_t *cpu_cycles, *frame_cycles, *instr;
    uint32_t *status;
    uint32_t n, n_pad;
};

// Power-on broadcast: every env gets the HeadlessAtari::new state.
__global__ void fa_fill_kernel(FillArgs g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n_pad) return;
    uint8_t *s = g.states + (size_t)i * STATE_BYTES;
    for (uint32_t j = 0; j < STATE_BYTES; ++j) s[j] = g.tmpl[j];
    if (i < g.n) {
        g.cpu_cycles[i] = 0;
        g.frame_cycles[i] = 0;
        g.instr[i] = 0;
        g.status[i] = 0;
    }
}

}  // namespace fa

// ============================================================================
// host side
// ============================================================================
namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) return fail(FA_EHIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

constexpr size_t kPadEnvs = 1024;

}  // namespace

struct fa_env {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t n = 0, n_pad = 0;
    std::vector<uint8_t> rom_image;    // staged into LDS
    uint32_t rom_bytes = 0;
    int scheme = 0;
    uint32_t hot_lo = 0xFFFu, hot_n = 0;
    AtariStateGpu power_on{};
    bool sorted = false;
    uint32_t block = 0;
    size_t max_lds = 65536;

    uint32_t *d_desc = nullptr;
    uint8_t *d_rom = nullptr, *d_states = nullptr, *d_actions = nullptr, *d_fs = nullptr, *d_obs = nullptr;
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

size_t lds_bytes(const fa_env *e, uint32_t bs) {
    return fa::ROM0 + e->rom_bytes + (size_t)bs * (fa::STATE_BYTES + 1u);
}

uint32_t pick_block(const fa_env *e) {
    if (e->block) return e->block;
    for (uint32_t b : {256u, 128u, 64u})
        if (lds_bytes(e, b) <= e->max_lds) return b;
    return 64u;
}

int ensure_actions(fa_env *e, size_t bytes) {
    if (bytes <= e->actions_cap) return FA_OK;
    if (e->d_actions) HIP_TRY(hipFree(e->d_actions));
    e->d_actions = nullptr;
    size_t cap = bytes < 4096 ? 4096 : bytes;
    HIP_TRY(hipMalloc(&e->d_actions, cap));
    e->actions_cap = cap;
    return FA_OK;
}

int launch_frames(fa_env *e, const uint8_t *d_actions, const uint8_t *d_fs, uint32_t k, uint8_t *d_obs) {
    fa::StepArgs g{};
    g.states = e->d_states;
    g.cpu_cycles = e->d_cpu;
    g.frame_cycles = e->d_fcyc;
    g.instr = e->d_instr;
    g.status = e->d_status;
    g.actions = d_actions;
    g.fs = d_fs;
    g.obs = d_obs;
    g.desc = e->d_desc;
    g.rom = e->d_rom;
    g.rom_bytes = e->rom_bytes;
    g.hot_lo = e->hot_lo;
    g.hot_n = e->hot_n;
    g.n = (uint32_t)e->n;
    g.k = k;
    g.stride = (uint32_t)e->n;
    const uint32_t bs = pick_block(e);
    const size_t lds = lds_bytes(e, bs);
    if (lds > e->max_lds) return fail(FA_EINVAL, "LDS budget exceeded for this block size");
    const uint32_t grid = (uint32_t)((e->n + bs - 1) / bs);
    if (grid == 0) return FA_OK;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e->timing) {
        if (e->events_used == e->events.size()) {
            hipEvent_t a, b;
            HIP_TRY(hipEventCreate(&a));
            HIP_TRY(hipEventCreate(&b));
            e->events.emplace_back(a, b);
        }
        ev0 = e->events[e->events_used].first;
        ev1 = e->events[e->events_used].second;
        e->events_used++;
        HIP_TRY(hipEventRecord(ev0, e->stream));
    }
    if (e->sorted)
        hipLaunchKernelGGL(fa::fa_frame_kernel<true>, dim3(grid), dim3(bs), lds, e->stream, g);
    else
        hipLaunchKernelGGL(fa::fa_frame_kernel<false>, dim3(grid), dim3(bs), lds, e->stream, g);
    HIP_TRY(hipGetLastError());
    if (e->timing) HIP_TRY(hipEventRecord(ev1, e->stream));
    return FA_OK;
}

// Power-on state (HeadlessAtari::new: src/atari.rs:410-430, 562-569;
// Cpu::reset src/cpu.rs:125-134; HeadlessTia::default src/headless_tia.rs:90-118;
// Pia::default src/pia.rs:46-62).
AtariStateGpu power_on_state(const fa_env *e, uint32_t first_bank) {
    AtariStateGpu s;
    std::memset(&s, 0, sizeof(s));
    s.sp = 0xFD;
    // the reset vector is read through the bus from the initial bank window
    const uint8_t *win = e->rom_image.data() + (size_t)first_bank * 4096u;
    s.pc = (uint16_t)(win[0xFFC] | (win[0xFFD] << 8));
    s.port_a_input = 0xFF;
    s.port_b_input = 0x3F;
    s.timer_prescaler = 1;
    s.timer_prescaler_select = 1;
    s.inpt4 = 1;
    s.inpt5 = 1;
    s.bank = (uint8_t)first_bank;
    s.scheme = (uint8_t)e->scheme;
    return s;
}

int check_state(const fa_env *e, const AtariStateGpu &s, size_t i) {
    auto bad = [&](const char *what) {
        return fail(FA_ESTATE, "state " + std::to_string(i) + ": " + what);
    };
    if (s.tia_clock % 3 != 0 || s.tia_clock >= 228) return bad("tia_clock must be a multiple of 3 below 228");
    uint16_t sel = s.timer_prescaler_select;
    if (sel != 1 && sel != 8 && sel != 64 && sel != 1024) return bad("timer select must be 1, 8, 64 or 1024");
    if (s.timer_prescaler > sel) return bad("timer prescaler above select");
    if (s.timer_value > 255) return bad("timer value above 255");
    if (s.scheme != e->scheme) return bad("scheme differs from the ROM's");
    if (s.bank >= e->rom_bytes / 4096u) return bad("bank out of range for the ROM");
    const uint16_t pos[5] = {s.pos_p0, s.pos_p1, s.pos_m0, s.pos_m1, s.pos_bl};
    for (uint16_t p : pos)
        if (p >= 160) return bad("object position above 159");
    return FA_OK;
}

}  // namespace

extern "C" {

const char *fa_last_error(void) { return g_err.c_str(); }

int fa_create(const uint8_t *rom, size_t rom_len, size_t n, fa_env **out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    *out = nullptr;
    if (!rom || rom_len == 0) return fail(FA_EINVAL, "empty ROM");
    int scheme;
    if (rom_len <= 4096) scheme = 0;
    else if (rom_len == 8192) scheme = 1;
    else if (rom_len == 16384) scheme = 2;
    else if (rom_len >= 32768) scheme = 3;
    else return fail(FA_EINVAL, "ROM size " + std::to_string(rom_len) +
                                    " is not 1..4096, 8192, 16384 or >= 32768 bytes");
    auto *e = new fa_env();
    e->n = n;
    e->n_pad = ((n + kPadEnvs - 1) / kPadEnvs) * kPadEnvs;
    if (e->n_pad == 0) e->n_pad = kPadEnvs;
    e->scheme = scheme;
    static const uint32_t first_bank[4] = {0, 1, 3, 7};
    static const uint32_t hot_lo[4] = {0xFFFu, 0xFF8u, 0xFF6u, 0xFF4u};
    static const uint32_t hot_n[4] = {0, 2, 4, 8};
    e->hot_lo = hot_lo[scheme];
    e->hot_n = hot_n[scheme];
    if (scheme == 0) {
        // HeadlessBus::load_bank for Fixed: mirror rom[i % len] into 4 KB
        e->rom_image.resize(4096);
        for (size_t i = 0; i < 4096; ++i) e->rom_image[i] = rom[i % rom_len];
    } else {
        size_t bytes = scheme == 1 ? 8192 : scheme == 2 ? 16384 : 32768;
        e->rom_image.assign(rom, rom + bytes);
    }
    e->rom_bytes = (uint32_t)e->rom_image.size();
    e->power_on = power_on_state(e, first_bank[scheme]);

    auto cleanup = [&](int rc) {
        fa_destroy(e);
        return rc;
    };
    int rc;
    if (hipGetDevice(&e->device) != hipSuccess) return cleanup(fail(FA_EHIP, "no HIP device"));
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) == hipSuccess && lds > 0)
        e->max_lds = (size_t)lds;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FA_EHIP, "hipStreamCreate failed"));
    hipFuncSetAttribute((const void *)fa::fa_frame_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)e->max_lds);
    hipFuncSetAttribute((const void *)fa::fa_frame_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)e->max_lds);
    if (hipMalloc(&e->d_desc, 2048) != hipSuccess || hipMalloc(&e->d_rom, e->rom_bytes) != hipSuccess ||
        hipMalloc(&e->d_states, e->n_pad * fa::STATE_BYTES) != hipSuccess ||
        hipMalloc(&e->d_cpu, e->n_pad * 8) != hipSuccess || hipMalloc(&e->d_fcyc, e->n_pad * 8) != hipSuccess ||
        hipMalloc(&e->d_instr, e->n_pad * 8) != hipSuccess || hipMalloc(&e->d_status, e->n_pad * 4) != hipSuccess ||
        hipMalloc(&e->d_obs, e->n_pad * 128) != hipSuccess || hipMalloc(&e->d_fs, e->n_pad) != hipSuccess)
        return cleanup(fail(FA_ENOMEM, "device allocation failed"));
    uint32_t desc[512];
    fa::build_descriptors(desc);
    if (hipMemcpy(e->d_rom, e->rom_image.data(), e->rom_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->d_desc, desc, sizeof(desc), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FA_EHIP, "ROM upload failed"));
    (void)rc;
    // Fresh handles hold the power-on state until reset() is called.
    fa::FillArgs fg{};
    std::memcpy(fg.tmpl, &e->power_on, sizeof(AtariStateGpu));
    fg.states = e->d_states;
    fg.cpu_cycles = e->d_cpu;
    fg.frame_cycles = e->d_fcyc;
    fg.instr = e->d_instr;
    fg.status = e->d_status;
    fg.n = (uint32_t)e->n;
    fg.n_pad = (uint32_t)e->n_pad;
    hipLaunchKernelGGL(fa::fa_fill_kernel, dim3((uint32_t)((e->n_pad + 255) / 256)), dim3(256), 0, e->stream, fg);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(e->stream) != hipSuccess)
        return cleanup(fail(FA_EHIP, "power-on fill failed"));
    *out = e;
    return FA_OK;
}

void fa_destroy(fa_env *e) {
    if (!e) return;
    if (e->stream) hipStreamSynchronize(e->stream);
    for (auto &p : e->events) {
        hipEventDestroy(p.first);
        hipEventDestroy(p.second);
    }
    for (void *p : {(void *)e->d_desc, (void *)e->d_rom, (void *)e->d_states, (void *)e->d_actions, (void *)e->d_fs, (void *)e->d_obs,
                    (void *)e->d_cpu, (void *)e->d_fcyc, (void *)e->d_instr, (void *)e->d_status})
        if (p) hipFree(p);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

int fa_set_sorted(fa_env *e, int sorted) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    e->sorted = sorted != 0;
    return FA_OK;
}

size_t fa_num_envs(const fa_env *e) { return e ? e->n : 0; }

int fa_reset(fa_env *e, uint8_t *obs_out) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    fa::FillArgs fg{};
    std::memcpy(fg.tmpl, &e->power_on, sizeof(AtariStateGpu));
    fg.states = e->d_states;
    fg.cpu_cycles = e->d_cpu;
    fg.frame_cycles = e->d_fcyc;
    fg.instr = e->d_instr;
    fg.status = e->d_status;
    fg.n = (uint32_t)e->n;
    fg.n_pad = (uint32_t)e->n_pad;
    hipLaunchKernelGGL(fa::fa_fill_kernel, dim3((uint32_t)((e->n_pad + 255) / 256)), dim3(256), 0, e->stream, fg);
    HIP_TRY(hipGetLastError());
    int rc = launch_frames(e, nullptr, nullptr, 1, e->d_obs);
    if (rc) return rc;
    if (obs_out) HIP_TRY(hipMemcpyAsync(obs_out, e->d_obs, e->n * 128, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_step_multi(fa_env *e, const uint8_t *actions, size_t n_actions, uint32_t k, uint8_t *obs_out) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    if (n_actions != (size_t)k * e->n) return fail(FA_EINVAL, "actions length must be k * num_envs");
    if (!actions && n_actions) return fail(FA_EINVAL, "actions is NULL");
    int rc = ensure_actions(e, n_actions);
    if (rc) return rc;
    if (n_actions) HIP_TRY(hipMemcpyAsync(e->d_actions, actions, n_actions, hipMemcpyHostToDevice, e->stream));
    rc = launch_frames(e, e->d_actions, nullptr, k, e->d_obs);
    if (rc) return rc;
    if (obs_out) HIP_TRY(hipMemcpyAsync(obs_out, e->d_obs, e->n * 128, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_step(fa_env *e, const uint8_t *actions, size_t n_actions, uint8_t *obs_out) {
    return fa_step_multi(e, actions, n_actions, 1, obs_out);
}

int fa_step_frameskip(fa_env *e, const uint8_t *actions, const uint8_t *fs, size_t n, uint8_t *obs_out) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (n != e->n) return fail(FA_EINVAL, "actions / frameskip length must be num_envs");
    if (n && (!actions || !fs)) return fail(FA_EINVAL, "NULL actions or frameskip");
    int rc = ensure_actions(e, n);
    if (rc) return rc;
    if (n) {
        HIP_TRY(hipMemcpyAsync(e->d_actions, actions, n, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->d_fs, fs, n, hipMemcpyHostToDevice, e->stream));
    }
    rc = launch_frames(e, e->d_actions, e->d_fs, 0, e->d_obs);
    if (rc) return rc;
    if (obs_out) HIP_TRY(hipMemcpyAsync(obs_out, e->d_obs, e->n * 128, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_step_device(fa_env *e, const uint8_t *d_actions, uint32_t k, uint8_t *d_obs) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    return launch_frames(e, d_actions, nullptr, k, d_obs);
}

int fa_step_frameskip_device(fa_env *e, const uint8_t *d_actions, const uint8_t *d_fs, uint8_t *d_obs) {
    if (!e || !d_fs) return fail(FA_EINVAL, "NULL handle or frameskip");
    return launch_frames(e, d_actions, d_fs, 0, d_obs);
}

int fa_sync(fa_env *e) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_download_states(fa_env *e, AtariStateGpu *out, size_t n) {
    if (!e || !out) return fail(FA_EINVAL, "NULL argument");
    if (n != e->n) return fail(FA_EINVAL, "n must be num_envs");
    HIP_TRY(hipMemcpyAsync(out, e->d_states, n * fa::STATE_BYTES, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_upload_states(fa_env *e, const AtariStateGpu *in, size_t n) {
    if (!e || !in) return fail(FA_EINVAL, "NULL argument");
    if (n != e->n) return fail(FA_EINVAL, "n must be num_envs");
    for (size_t i = 0; i < n; ++i) {
        int rc = check_state(e, in[i], i);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(e->d_states, in, n * fa::STATE_BYTES, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(e->d_status, 0, n * 4, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_download_counters(fa_env *e, uint64_t *cpu_cycles, uint64_t *frame_cycles, uint64_t *instructions,
                         uint32_t *status) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (cpu_cycles) HIP_TRY(hipMemcpyAsync(cpu_cycles, e->d_cpu, e->n * 8, hipMemcpyDeviceToHost, e->stream));
    if (frame_cycles) HIP_TRY(hipMemcpyAsync(frame_cycles, e->d_fcyc, e->n * 8, hipMemcpyDeviceToHost, e->stream));
    if (instructions) HIP_TRY(hipMemcpyAsync(instructions, e->d_instr, e->n * 8, hipMemcpyDeviceToHost, e->stream));
    if (status) HIP_TRY(hipMemcpyAsync(status, e->d_status, e->n * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_upload_cpu_cycles(fa_env *e, const uint64_t *cpu_cycles, size_t n) {
    if (!e || !cpu_cycles) return fail(FA_EINVAL, "NULL argument");
    if (n != e->n) return fail(FA_EINVAL, "n must be num_envs");
    HIP_TRY(hipMemcpyAsync(e->d_cpu, cpu_cycles, n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_device_alloc(size_t bytes, void **d_ptr) {
    if (!d_ptr) return fail(FA_EINVAL, "NULL out pointer");
    HIP_TRY(hipMalloc(d_ptr, bytes ? bytes : 1));
    return FA_OK;
}
int fa_device_free(void *d_ptr) {
    if (d_ptr) HIP_TRY(hipFree(d_ptr));
    return FA_OK;
}
int fa_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes) {
    HIP_TRY(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    return FA_OK;
}
int fa_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes) {
    HIP_TRY(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    return FA_OK;
}

int fa_set_timing(fa_env *e, int enabled) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    e->timing = enabled != 0;
    return FA_OK;
}

int fa_kernel_time(fa_env *e, double *total_ms, uint64_t *launches, int reset) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < e->events_used; ++i) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e->events[i].first, e->events[i].second));
        e->ms_done += ms;
        e->launches_done++;
    }
    e->events_used = 0;
    if (total_ms) *total_ms = e->ms_done;
    if (launches) *launches = e->launches_done;
    if (reset) {
        e->ms_done = 0.0;
        e->launches_done = 0;
    }
    return FA_OK;
}

int fa_set_block_size(fa_env *e, uint32_t block) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (block != 0 && (block % 64 != 0 || block > 256 || (kPadEnvs % block) != 0))
        return fail(FA_EINVAL, "block must be 0, 64, 128 or 256");
    if (block && lds_bytes(e, block) > e->max_lds)
        return fail(FA_EINVAL, "block too large for the LDS budget with this ROM");
    e->block = block;
    return FA_OK;
}

int fa_get_max_lds(const fa_env *e, size_t *bytes) {
    if (!e || !bytes) return fail(FA_EINVAL, "NULL argument");
    *bytes = e->max_lds;
    return FA_OK;
}

int fa_get_block_size(const fa_env *e, uint32_t *block) {
    if (!e || !block) return fail(FA_EINVAL, "NULL argument");
    *block = pick_block(e);
    return FA_OK;
}

}  // extern "C"
