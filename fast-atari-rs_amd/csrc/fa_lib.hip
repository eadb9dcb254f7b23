// fa_lib.hip -- the single-ROM C ABI (include/fastatari.h): the drop-in for
// BatchAtariGpu (src/cuda_env.rs:99-191), plus ROM classes, state validation
// and error reporting shared with the batch and server translation units.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "fa_internal.hpp"

namespace fa {

namespace {
thread_local std::string g_err;
}

std::string &last_error() { return g_err; }

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// BankScheme::detect (src/atari.rs:97-104) and HeadlessBus::new / load_bank
// (src/atari.rs:410-449): the staged image and the power-on state
// (HeadlessAtari::new src/atari.rs:562-569, Cpu::reset src/cpu.rs:125-134,
// HeadlessTia::default src/headless_tia.rs:90-118, Pia::default
// src/pia.rs:46-62).
int make_rom_class(const uint8_t *rom, size_t rom_len, RomClass &rc) {
    if (!rom || rom_len == 0) return fail(FA_EINVAL, "empty ROM");
    int scheme;
    if (rom_len <= 4096) scheme = 0;
    else if (rom_len == 8192) scheme = 1;
    else if (rom_len == 16384) scheme = 2;
    else if (rom_len >= 32768) scheme = 3;
    else return fail(FA_EINVAL, "ROM size " + std::to_string(rom_len) +
                                    " is not 1..4096, 8192, 16384 or >= 32768 bytes");
    static const uint32_t first_bank[4] = {0, 1, 3, 7};
    static const uint32_t hot_lo[4] = {0xFFFu, 0xFF8u, 0xFF6u, 0xFF4u};
    static const uint32_t hot_n[4] = {0, 2, 4, 8};
    rc.scheme = scheme;
    rc.hot_lo = hot_lo[scheme];
    rc.hot_n = hot_n[scheme];
    rc.first_bank = first_bank[scheme];
    if (scheme == 0) {
        // Fixed: mirror rom[i % len] into the 4 KB window
        rc.image.resize(4096);
        for (size_t i = 0; i < 4096; ++i) rc.image[i] = rom[i % rom_len];
    } else {
        size_t bytes = scheme == 1 ? 8192 : scheme == 2 ? 16384 : 32768;
        rc.image.assign(rom, rom + bytes);
    }
    rc.rom_bytes = (uint32_t)rc.image.size();
    AtariStateGpu &s = rc.power_on;
    std::memset(&s, 0, sizeof(s));
    s.sp = 0xFD;
    // the reset vector is read through the bus from the initial bank window
    const uint8_t *win = rc.image.data() + (size_t)rc.first_bank * 4096u;
    s.pc = (uint16_t)(win[0xFFC] | (win[0xFFD] << 8));
    s.port_a_input = 0xFF;
    s.port_b_input = 0x3F;
    s.timer_prescaler = 1;
    s.timer_prescaler_select = 1;
    s.inpt4 = 1;
    s.inpt5 = 1;
    s.bank = (uint8_t)rc.first_bank;
    s.scheme = (uint8_t)scheme;
    return FA_OK;
}

int check_state(const RomClass &rc, const AtariStateGpu &s, size_t i) {
    auto bad = [&](const char *what) { return fail(FA_ESTATE, "state " + std::to_string(i) + ": " + what); };
    if (s.tia_clock % 3 != 0 || s.tia_clock >= 228) return bad("tia_clock must be a multiple of 3 below 228");
    uint16_t sel = s.timer_prescaler_select;
    if (sel != 1 && sel != 8 && sel != 64 && sel != 1024) return bad("timer select must be 1, 8, 64 or 1024");
    if (s.timer_prescaler > sel) return bad("timer prescaler above select");
    if (s.timer_value > 255) return bad("timer value above 255");
    if (s.scheme != rc.scheme) return bad("scheme differs from the ROM's");
    if (s.bank >= rc.rom_bytes / 4096u) return bad("bank out of range for the ROM");
    const uint16_t pos[5] = {s.pos_p0, s.pos_p1, s.pos_m0, s.pos_m1, s.pos_bl};
    for (uint16_t p : pos)
        if (p >= 160) return bad("object position above 159");
    return FA_OK;
}

uint32_t default_block(uint32_t rom_bytes, size_t max_lds) {
    for (uint32_t b : {256u, 128u, 64u})
        if (lds_bytes(rom_bytes, b) <= max_lds) return b;
    return 64u;
}

}  // namespace fa

namespace {

constexpr size_t kPadEnvs = 1024;

}  // namespace

struct fa_env {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t n = 0, n_pad = 0;
    fa::RomClass rom;
    bool sorted = false;
    uint32_t block = 0;
    size_t max_lds = 65536;

    uint32_t *d_desc = nullptr;
    uint8_t *d_rom = nullptr, *d_states = nullptr, *d_actions = nullptr, *d_fs = nullptr, *d_obs = nullptr;
    uint64_t *d_cpu = nullptr, *d_fcyc = nullptr, *d_instr = nullptr;
    uint32_t *d_status = nullptr;
    size_t actions_cap = 0;
    uint8_t *h_obs = nullptr, *hv_obs = nullptr;   // pinned obs for host-side steps, and its device address

    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t events_used = 0;
    double ms_done = 0.0;
    uint64_t launches_done = 0;
};

namespace {

using fa::fail;

uint32_t pick_block(const fa_env *e) { return e->block ? e->block : fa::default_block(e->rom.rom_bytes, e->max_lds); }

int ensure_actions(fa_env *e, size_t bytes) {
    if (bytes <= e->actions_cap) return FA_OK;
    if (e->d_actions) FA_HIP_TRY(hipFree(e->d_actions));
    e->d_actions = nullptr;
    size_t cap = bytes < 4096 ? 4096 : bytes;
    FA_HIP_TRY(hipMalloc(&e->d_actions, cap));
    e->actions_cap = cap;
    return FA_OK;
}

// Host-side steps: the kernel writes the obs straight into pinned host memory
// (no separate device-to-host copy), and a few threads copy them into the
// caller's buffer (one core copies 8 MB at a fraction of the host's memory
// bandwidth).
int host_obs(fa_env *e, uint8_t **dev) {
    if (!e->h_obs) {
        FA_HIP_TRY(hipHostMalloc(&e->h_obs, std::max<size_t>(e->n, 1) * 128, hipHostMallocDefault));
        FA_HIP_TRY(hipHostGetDevicePointer((void **)&e->hv_obs, e->h_obs, 0));
    }
    *dev = e->hv_obs;
    return FA_OK;
}

// k steps; each holds its action for fs[s * n + i] frames (1 without fs).
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
    g.rom_bytes = e->rom.rom_bytes;
    g.hot_lo = e->rom.hot_lo;
    g.hot_n = e->rom.hot_n;
    g.n = (uint32_t)e->n;
    g.k = k;
    g.stride = (uint32_t)e->n;
    const uint32_t bs = pick_block(e);
    if (fa::lds_bytes(e->rom.rom_bytes, bs) > e->max_lds) return fail(FA_EINVAL, "LDS budget exceeded for this block size");
    const uint32_t grid = (uint32_t)((e->n + bs - 1) / bs);
    if (grid == 0) return FA_OK;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e->timing) {
        if (e->events_used == e->events.size()) {
            hipEvent_t a, b;
            FA_HIP_TRY(hipEventCreate(&a));
            FA_HIP_TRY(hipEventCreate(&b));
            e->events.emplace_back(a, b);
        }
        ev0 = e->events[e->events_used].first;
        ev1 = e->events[e->events_used].second;
        e->events_used++;
        FA_HIP_TRY(hipEventRecord(ev0, e->stream));
    }
    FA_HIP_TRY(fa::launch_step(g, e->sorted, bs, grid, e->stream));
    if (e->timing) FA_HIP_TRY(hipEventRecord(ev1, e->stream));
    return FA_OK;
}

int fill_power_on(fa_env *e) {
    FA_HIP_TRY(fa::launch_fill(e->rom.power_on, e->d_states, e->d_cpu, e->d_fcyc, e->d_instr, e->d_status,
                               (uint32_t)e->n, (uint32_t)e->n_pad, e->stream));
    return FA_OK;
}

}  // namespace

namespace fa {
hipStream_t env_stream(fa_env *e) { return e->stream; }

// Persistent copy workers: started once for the process (and never stopped:
// they sleep on a condition variable between copies), so a host step pays no
// thread start-up.  One copy uses the pool at a time; a concurrent caller
// copies alone.
namespace {
class CopyPool {
  public:
    static constexpr size_t kWorkers = 7;
    static CopyPool *get() {
        static CopyPool *pool = new CopyPool();     // left running until the process ends
        return pool;
    }
    // Copies `parts` slices of `per` bytes (the last one shorter); false when
    // the pool is in use or has no workers.
    bool copy(uint8_t *d, const uint8_t *s, size_t bytes, size_t per, size_t parts) {
        std::unique_lock<std::mutex> use(use_mu_, std::try_to_lock);
        if (!use.owns_lock() || nworkers_ == 0) return false;
        {
            std::lock_guard<std::mutex> lk(mu_);
            d_ = d;
            s_ = s;
            bytes_ = bytes;
            per_ = per;
            parts_ = parts;
            next_.store(0);
            active_ = nworkers_;
            ++gen_;
        }
        cv_work_.notify_all();
        run_parts();
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return active_ == 0; });
        return true;
    }

  private:
    CopyPool() {
        for (size_t i = 0; i < kWorkers; ++i) {
            try {
                std::thread(&CopyPool::worker, this).detach();
                ++nworkers_;
            } catch (...) {
                break;
            }
        }
    }
    void run_parts() {
        for (size_t p = next_.fetch_add(1); p < parts_; p = next_.fetch_add(1)) {
            const size_t off = p * per_;
            if (off < bytes_) std::memcpy(d_ + off, s_ + off, std::min(per_, bytes_ - off));
        }
    }
    void worker() {
        uint64_t seen = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_work_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            run_parts();
            std::lock_guard<std::mutex> lk(mu_);
            if (--active_ == 0) cv_done_.notify_all();
        }
    }
    std::mutex use_mu_, mu_;
    std::condition_variable cv_work_, cv_done_;
    uint64_t gen_ = 0;
    size_t nworkers_ = 0, active_ = 0;
    uint8_t *d_ = nullptr;
    const uint8_t *s_ = nullptr;
    size_t bytes_ = 0, per_ = 0, parts_ = 0;
    std::atomic<size_t> next_{0};
};
}  // namespace

void copy_host(void *dst, const void *src, size_t bytes) {
    constexpr size_t kPart = 1u << 20;
    const size_t parts = std::min<size_t>(CopyPool::kWorkers + 1, (bytes + kPart - 1) / kPart);
    auto *d = static_cast<uint8_t *>(dst);
    auto *s = static_cast<const uint8_t *>(src);
    if (parts <= 1 || !CopyPool::get()->copy(d, s, bytes, (bytes + parts - 1) / parts, parts))
        std::memcpy(d, s, bytes);
}
uint8_t *env_obs_device(fa_env *e) { return e->d_obs; }
}  // namespace fa

extern "C" {

const char *fa_last_error(void) { return fa::last_error().c_str(); }

int fa_create(const uint8_t *rom, size_t rom_len, size_t n, fa_env **out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    *out = nullptr;
    fa::RomClass rc;
    int rc_ = fa::make_rom_class(rom, rom_len, rc);
    if (rc_) return rc_;
    auto *e = new (std::nothrow) fa_env();
    if (!e) return fail(FA_ENOMEM, "out of host memory");
    e->n = n;
    e->n_pad = ((n + kPadEnvs - 1) / kPadEnvs) * kPadEnvs;
    if (e->n_pad == 0) e->n_pad = kPadEnvs;
    e->rom = std::move(rc);

    auto cleanup = [&](int code) {
        fa_destroy(e);
        return code;
    };
    if (hipGetDevice(&e->device) != hipSuccess) return cleanup(fail(FA_EHIP, "no HIP device"));
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) == hipSuccess && lds > 0)
        e->max_lds = (size_t)lds;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FA_EHIP, "hipStreamCreate failed"));
    if (fa::prepare_kernels(e->max_lds) != hipSuccess) return cleanup(fail(FA_EHIP, "kernel attribute setup failed"));
    if (hipMalloc(&e->d_desc, fa::kDescBytes) != hipSuccess || hipMalloc(&e->d_rom, e->rom.rom_bytes) != hipSuccess ||
        hipMalloc(&e->d_states, e->n_pad * fa::kStateBytes) != hipSuccess ||
        hipMalloc(&e->d_cpu, e->n_pad * 8) != hipSuccess || hipMalloc(&e->d_fcyc, e->n_pad * 8) != hipSuccess ||
        hipMalloc(&e->d_instr, e->n_pad * 8) != hipSuccess || hipMalloc(&e->d_status, e->n_pad * 4) != hipSuccess ||
        hipMalloc(&e->d_obs, e->n_pad * 128) != hipSuccess || hipMalloc(&e->d_fs, e->n_pad) != hipSuccess)
        return cleanup(fail(FA_ENOMEM, "device allocation failed"));
    uint32_t desc[fa::kDescBytes / 4];
    fa::descriptor_table(desc);
    if (hipMemcpy(e->d_rom, e->rom.image.data(), e->rom.rom_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->d_desc, desc, sizeof(desc), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FA_EHIP, "ROM upload failed"));
    // Fresh handles hold the power-on state until reset() is called.
    if (fill_power_on(e) || hipStreamSynchronize(e->stream) != hipSuccess)
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
    for (void *p : {(void *)e->d_desc, (void *)e->d_rom, (void *)e->d_states, (void *)e->d_actions, (void *)e->d_fs,
                    (void *)e->d_obs, (void *)e->d_cpu, (void *)e->d_fcyc, (void *)e->d_instr, (void *)e->d_status})
        if (p) hipFree(p);
    if (e->h_obs) hipHostFree(e->h_obs);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

int fa_set_sorted(fa_env *e, int sorted) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    e->sorted = sorted != 0;
    return FA_OK;
}

size_t fa_num_envs(const fa_env *e) { return e ? e->n : 0; }

int fa_descriptor_table(uint32_t *out, size_t n_words) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    if (n_words != fa::kDescBytes / 4) return fail(FA_EINVAL, "n_words must be 2048");
    fa::descriptor_table(out);
    return FA_OK;
}

int fa_reset(fa_env *e, uint8_t *obs_out) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    int rc = fill_power_on(e);
    if (rc) return rc;
    uint8_t *obs = e->d_obs;
    if (obs_out && (rc = host_obs(e, &obs))) return rc;
    rc = launch_frames(e, nullptr, nullptr, 1, obs);
    if (rc) return rc;
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    if (obs_out) fa::copy_host(obs_out, e->h_obs, e->n * 128);
    return FA_OK;
}

int fa_step_multi(fa_env *e, const uint8_t *actions, size_t n_actions, uint32_t k, uint8_t *obs_out) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    if (n_actions != (size_t)k * e->n) return fail(FA_EINVAL, "actions length must be k * num_envs");
    if (!actions && n_actions) return fail(FA_EINVAL, "actions is NULL");
    int rc = ensure_actions(e, n_actions);
    if (rc) return rc;
    if (n_actions) FA_HIP_TRY(hipMemcpyAsync(e->d_actions, actions, n_actions, hipMemcpyHostToDevice, e->stream));
    uint8_t *obs = e->d_obs;
    if (obs_out && (rc = host_obs(e, &obs))) return rc;
    rc = launch_frames(e, e->d_actions, nullptr, k, obs);
    if (rc) return rc;
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    if (obs_out) fa::copy_host(obs_out, e->h_obs, e->n * 128);
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
        FA_HIP_TRY(hipMemcpyAsync(e->d_actions, actions, n, hipMemcpyHostToDevice, e->stream));
        FA_HIP_TRY(hipMemcpyAsync(e->d_fs, fs, n, hipMemcpyHostToDevice, e->stream));
    }
    uint8_t *obs = e->d_obs;
    if (obs_out && (rc = host_obs(e, &obs))) return rc;
    rc = launch_frames(e, e->d_actions, e->d_fs, 1, obs);
    if (rc) return rc;
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    if (obs_out) fa::copy_host(obs_out, e->h_obs, e->n * 128);
    return FA_OK;
}

int fa_step_device(fa_env *e, const uint8_t *d_actions, uint32_t k, uint8_t *d_obs) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    return launch_frames(e, d_actions, nullptr, k, d_obs);
}

int fa_step_frameskip_device(fa_env *e, const uint8_t *d_actions, const uint8_t *d_fs, uint8_t *d_obs) {
    if (!e || !d_fs) return fail(FA_EINVAL, "NULL handle or frameskip");
    return launch_frames(e, d_actions, d_fs, 1, d_obs);
}

int fa_sync(fa_env *e) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_download_states(fa_env *e, AtariStateGpu *out, size_t n) {
    if (!e || !out) return fail(FA_EINVAL, "NULL argument");
    if (n != e->n) return fail(FA_EINVAL, "n must be num_envs");
    FA_HIP_TRY(hipMemcpyAsync(out, e->d_states, n * fa::kStateBytes, hipMemcpyDeviceToHost, e->stream));
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_upload_states(fa_env *e, const AtariStateGpu *in, size_t n) {
    if (!e || !in) return fail(FA_EINVAL, "NULL argument");
    if (n != e->n) return fail(FA_EINVAL, "n must be num_envs");
    for (size_t i = 0; i < n; ++i) {
        int rc = fa::check_state(e->rom, in[i], i);
        if (rc) return rc;
    }
    FA_HIP_TRY(hipMemcpyAsync(e->d_states, in, n * fa::kStateBytes, hipMemcpyHostToDevice, e->stream));
    FA_HIP_TRY(hipMemsetAsync(e->d_status, 0, n * 4, e->stream));
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_download_counters(fa_env *e, uint64_t *cpu_cycles, uint64_t *frame_cycles, uint64_t *instructions,
                         uint32_t *status) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    if (cpu_cycles) FA_HIP_TRY(hipMemcpyAsync(cpu_cycles, e->d_cpu, e->n * 8, hipMemcpyDeviceToHost, e->stream));
    if (frame_cycles) FA_HIP_TRY(hipMemcpyAsync(frame_cycles, e->d_fcyc, e->n * 8, hipMemcpyDeviceToHost, e->stream));
    if (instructions) FA_HIP_TRY(hipMemcpyAsync(instructions, e->d_instr, e->n * 8, hipMemcpyDeviceToHost, e->stream));
    if (status) FA_HIP_TRY(hipMemcpyAsync(status, e->d_status, e->n * 4, hipMemcpyDeviceToHost, e->stream));
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_upload_cpu_cycles(fa_env *e, const uint64_t *cpu_cycles, size_t n) {
    if (!e || !cpu_cycles) return fail(FA_EINVAL, "NULL argument");
    if (n != e->n) return fail(FA_EINVAL, "n must be num_envs");
    FA_HIP_TRY(hipMemcpyAsync(e->d_cpu, cpu_cycles, n * 8, hipMemcpyHostToDevice, e->stream));
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    return FA_OK;
}

int fa_device_alloc(size_t bytes, void **d_ptr) {
    if (!d_ptr) return fail(FA_EINVAL, "NULL out pointer");
    FA_HIP_TRY(hipMalloc(d_ptr, bytes ? bytes : 1));
    return FA_OK;
}
int fa_device_free(void *d_ptr) {
    if (d_ptr) FA_HIP_TRY(hipFree(d_ptr));
    return FA_OK;
}
int fa_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes) {
    FA_HIP_TRY(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    return FA_OK;
}
int fa_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes) {
    FA_HIP_TRY(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    return FA_OK;
}

int fa_set_timing(fa_env *e, int enabled) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    e->timing = enabled != 0;
    return FA_OK;
}

int fa_kernel_time(fa_env *e, double *total_ms, uint64_t *launches, int reset) {
    if (!e) return fail(FA_EINVAL, "NULL handle");
    FA_HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < e->events_used; ++i) {
        float ms = 0.f;
        FA_HIP_TRY(hipEventElapsedTime(&ms, e->events[i].first, e->events[i].second));
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
    if (block && fa::lds_bytes(e->rom.rom_bytes, block) > e->max_lds)
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
