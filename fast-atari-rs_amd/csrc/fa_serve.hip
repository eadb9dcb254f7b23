// fa_serve.hip -- the env-step server: client envs stay resident in HBM,
// clients are admitted by memory footprint, and step requests that arrive
// while others are on the GPU are coalesced into the next launch of their ROM
// class (continuous batching).
//
// Memory: per ROM class, chunks of kChunkSlots env slots (AtariStateGpu AoS,
// counters, one packed 4-byte input word and 128 bytes of obs per slot),
// allocated on demand against the byte budget.  A client owns a contiguous,
// 64-aligned slot range inside one chunk; freed ranges are reused.
//
// Requests: submit() copies the client's inputs into its pinned staging block
// and queues it.  The dispatcher thread takes every queued request of the
// chunk with the oldest waiting request, builds a task list (one workgroup
// per <= 256 envs of a client range) and enqueues on a free launch context of
// that class: the frame kernel alone, reading each client's packed inputs
// from and writing its obs to the client's pinned host buffers (zero-copy:
// no copy-engine work, which would serialise across classes), then an event.  Each class has one stream (the device exposes few hardware
// queues per process, and kernels of streams that share a queue serialise, so
// one stream per class keeps a slow class from delaying a fast one) and
// kCtxPerClass launch contexts on it; classes overlap on the device.  A launch
// ends with an event; one completer thread per class waits for its launches
// in order and completes their requests, so a slow class never delays the
// completions of a fast one.  With one context per class, the next launch of
// a class takes every request that arrived while the previous one ran.
// A client has at most one request in flight, so concurrent launches touch
// disjoint slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fa_internal.hpp"
#include "fa_serve_wait.hpp"

namespace {

using fa::fail;
using Clock = std::chrono::steady_clock;

constexpr uint32_t kChunkSlots = 65536;
constexpr uint32_t kUnit = 64;                 // slot allocation granule
constexpr uint32_t kUnits = kChunkSlots / kUnit;
constexpr uint32_t kCtxPerClass = 1;
constexpr uint32_t kTaskEnvs = 256;            // envs per workgroup of the admin launches
constexpr double kLaunchLimit = 60.0;          // seconds a completer waits for one launch
// A class's workgroup size: the largest of 256 / 128 / 64 envs whose LDS
// footprint stays within kShareLds, so that workgroups of two classes fit on
// one CU together (160 KB) -- a slow class's long-running workgroups then do
// not keep another class's launches waiting for a free CU.  (Smaller
// workgroups than needed spread a slow class's waves over more CUs, where they
// share SIMDs with the other classes' waves.)
constexpr size_t kShareLds = 80u << 10;
constexpr uint32_t kMaxClientEnvs = kChunkSlots;
constexpr size_t kSlotBytes = fa::kStateBytes + 3 * 8 + 4 + 4 + 128;

struct Chunk {
    uint8_t *d_states = nullptr;
    uint64_t *d_cpu = nullptr, *d_fcyc = nullptr, *d_instr = nullptr;
    uint32_t *d_status = nullptr, *d_in = nullptr;
    uint8_t *d_obs = nullptr;
    std::vector<uint8_t> used;   // per kUnit slots
    uint32_t used_units = 0;
};

struct Client {
    int cls = 0, chunk = 0;
    uint32_t first = 0, n = 0, units = 0;
    uint32_t *h_in = nullptr;    // pinned: packed inputs, n words
    uint8_t *h_obs = nullptr;    // pinned: n * 128
    uint32_t *dv_in = nullptr;   // device addresses of the two
    uint8_t *dv_obs = nullptr;
    bool open = false;
    int state = 0;               // 0 idle, 1 queued, 2 in flight, 3 done
    uint32_t frames = 0;
    Clock::time_point t_submit;
    double latency_ms = 0.0;
};

struct ServerClass;

struct LaunchCtx {
    fa_server *srv = nullptr;
    int cls = 0;
    hipStream_t stream = nullptr;      // the class's stream (not owned)
    hipEvent_t done = nullptr;         // recorded after the launch's obs copies
    fa::HostTask *h_tasks = nullptr;   // pinned, read by the kernel
    fa::HostTask *dv_tasks = nullptr;  // its device address
    size_t task_cap = 0;
    std::vector<int> clients;
    std::vector<fa::HostTask> parts;   // each client's slot range and buffers, copied under the lock
    bool busy = false;
    Clock::time_point t_launch{};      // when the launch was issued
};

struct ServerClass {
    fa::RomClass rom;
    uint8_t *d_rom = nullptr;
    uint8_t *d_snap = nullptr;         // reset snapshot (222 B) + 3 counters, 16-aligned
    std::vector<std::unique_ptr<Chunk>> chunks;
    std::vector<std::unique_ptr<LaunchCtx>> ctx;
    std::deque<int> queue;             // clients with a queued request
    std::deque<LaunchCtx *> inflight;  // launches on the stream, oldest first
    hipStream_t stream = nullptr;
    uint32_t open_clients = 0;
    Clock::time_point hold_until{};    // coalescing window after a completion
    uint32_t block = 256;
    int sorted = 0;
};

struct Sample {
    double ms;
    int cls;
    uint32_t envs, frames;
};

}  // namespace

struct fa_server {
    int device = 0;
    size_t max_lds = 65536;
    size_t budget = 0, used_bytes = 0;
    uint32_t coalesce_us = 300;
    uint32_t *d_desc = nullptr;
    std::deque<ServerClass> cls;       // deque: ServerClass is move-only
    std::deque<Client> clients;        // indexed under mu only; element references stay valid as it grows
    std::vector<int> free_ids;
    hipStream_t admin = nullptr;       // client set-up and downloads
    uint32_t *d_admin_tasks = nullptr;
    std::mutex admin_mu;

    std::mutex mu;
    std::condition_variable cv_dispatch, cv_done;
    std::thread dispatcher;
    std::vector<std::thread> completers;   // one per class
    std::condition_variable cv_inflight;
    bool stop = false, stop_completers = false;
    int error = FA_OK;
    std::string error_msg;
    std::vector<Sample> samples;
    uint64_t launches = 0, env_frames = 0, requests = 0;
    std::vector<int> completions;      // finished requests, drained by the load driver
    bool track_completions = false;
};

namespace {

int free_chunk_bytes(Chunk &c) {
    for (void *p : {(void *)c.d_states, (void *)c.d_cpu, (void *)c.d_fcyc, (void *)c.d_instr, (void *)c.d_status,
                    (void *)c.d_in, (void *)c.d_obs})
        if (p) hipFree(p);
    return FA_OK;
}

int alloc_chunk(fa_server *s, ServerClass &sc) {
    const size_t bytes = (size_t)kChunkSlots * kSlotBytes;
    if (s->used_bytes + bytes > s->budget) return fail(FA_EBUSY, "state budget exhausted");
    auto c = std::make_unique<Chunk>();
    const size_t n = kChunkSlots;
    if (hipMalloc(&c->d_states, n * fa::kStateBytes) != hipSuccess || hipMalloc(&c->d_cpu, n * 8) != hipSuccess ||
        hipMalloc(&c->d_fcyc, n * 8) != hipSuccess || hipMalloc(&c->d_instr, n * 8) != hipSuccess ||
        hipMalloc(&c->d_status, n * 4) != hipSuccess || hipMalloc(&c->d_in, n * 4) != hipSuccess ||
        hipMalloc(&c->d_obs, n * 128) != hipSuccess) {
        free_chunk_bytes(*c);
        return fail(FA_ENOMEM, "device allocation failed");
    }
    c->used.assign(kUnits, 0);
    s->used_bytes += bytes;
    sc.chunks.push_back(std::move(c));
    return FA_OK;
}

// First fit of `units` consecutive free units; -1 when none.
int find_run(const Chunk &c, uint32_t units) {
    uint32_t run = 0;
    for (uint32_t u = 0; u < kUnits; ++u) {
        run = c.used[u] ? 0 : run + 1;
        if (run == units) return (int)(u + 1 - units);
    }
    return -1;
}

fa::StepArgs chunk_args(const fa_server *s, const ServerClass &sc, const Chunk &c) {
    fa::StepArgs g{};
    g.states = c.d_states;
    g.cpu_cycles = c.d_cpu;
    g.frame_cycles = c.d_fcyc;
    g.instr = c.d_instr;
    g.status = c.d_status;
    g.obs = c.d_obs;
    g.packed = c.d_in;
    g.reset_state = sc.d_snap;
    g.reset_counters = reinterpret_cast<const uint64_t *>(sc.d_snap + 224);
    g.desc = s->d_desc;
    g.rom = sc.d_rom;
    g.rom_bytes = sc.rom.rom_bytes;
    g.hot_lo = sc.rom.hot_lo;
    g.hot_n = sc.rom.hot_n;
    g.n = kChunkSlots;
    g.k = 1;
    g.stride = kChunkSlots;
    return g;
}

uint32_t serve_block(uint32_t rom_bytes, size_t max_lds) {
    for (uint32_t b : {256u, 128u}) {
        const size_t l = fa::lds_bytes(rom_bytes, b);
        if (l <= kShareLds && l <= max_lds) return b;
    }
    return 64u;
}

uint32_t fill_tasks(uint32_t *tasks, uint32_t first, uint32_t n) {
    uint32_t t = 0;
    for (uint32_t off = 0; off < n; off += kTaskEnvs) {
        tasks[2 * t] = first + off;
        tasks[2 * t + 1] = std::min(kTaskEnvs, n - off);
        ++t;
    }
    return t;
}

void finish_launch(LaunchCtx *ctx);

// Dispatcher: one launch per iteration; waits for queued work and a free
// launch context of the same class.  After a class's launch completes, its
// clients resubmit one by one: the class is held for up to coalesce_us or a
// tenth of its last launch's time, whichever is longer (or until every open
// client of the class has a request queued) so that the next launch takes
// them together instead of running a near-empty launch first.
void dispatch_loop(fa_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    size_t rr = 0;
    while (true) {
        int pick = -1;
        LaunchCtx *ctx = nullptr;
        while (!s->stop) {
            // the eligible class whose head request waited longest
            Clock::time_point best{}, wake{};
            bool held = false;
            const auto now = Clock::now();
            for (size_t i = 0; i < s->cls.size(); ++i) {
                const size_t ci = (rr + i) % s->cls.size();
                auto &sc = s->cls[ci];
                if (sc.queue.empty()) continue;
                LaunchCtx *free_ctx = nullptr;
                for (auto &c : sc.ctx)
                    if (!c->busy) {
                        free_ctx = c.get();
                        break;
                    }
                if (!free_ctx) continue;
                if (now < sc.hold_until && sc.queue.size() < sc.open_clients) {
                    if (!held || sc.hold_until < wake) wake = sc.hold_until;
                    held = true;
                    continue;
                }
                const auto t = s->clients[sc.queue.front()].t_submit;
                if (pick < 0 || t < best) {
                    best = t;
                    pick = (int)ci;
                    ctx = free_ctx;
                }
            }
            if (pick >= 0) break;
            if (held)
                s->cv_dispatch.wait_until(lk, wake);
            else
                s->cv_dispatch.wait(lk);
        }
        if (s->stop) return;
        rr = (size_t)pick + 1;
        ServerClass &sc = s->cls[pick];
        // every queued request on the head request's chunk
        const int chunk = s->clients[sc.queue.front()].chunk;
        const uint32_t bs = sc.block;
        const int sorted = sc.sorted;
        ctx->clients.clear();
        std::deque<int> rest;
        uint32_t ntasks = 0;
        for (int id : sc.queue) {
            Client &cl = s->clients[id];
            if (cl.chunk == chunk) {
                ctx->clients.push_back(id);
                ntasks += (cl.n + bs - 1) / bs;
                cl.state = 2;
            } else {
                rest.push_back(id);
            }
        }
        sc.queue.swap(rest);
        ctx->busy = true;
        Chunk &ch = *sc.chunks[chunk];
        // the launch's task list is built from a copy taken under the lock:
        // fa_server_open may grow `clients` meanwhile
        ctx->parts.clear();
        for (int id : ctx->clients) {
            const Client &cl = s->clients[id];
            ctx->parts.push_back(fa::HostTask{cl.first, cl.n, cl.dv_in, cl.dv_obs});
        }
        lk.unlock();

        int rc = FA_OK;
        std::string msg;
        auto hip_ok = [&](hipError_t e, const char *what) {
            if (e != hipSuccess && rc == FA_OK) {
                rc = FA_EHIP;
                msg = std::string(what) + ": " + hipGetErrorString(e);
            }
            return e == hipSuccess;
        };
        if (ntasks > ctx->task_cap) {
            // grows rarely: the first launches of a context (idle: its last launch completed)
            if (ctx->h_tasks) hipHostFree(ctx->h_tasks);
            ctx->h_tasks = ctx->dv_tasks = nullptr;
            ctx->task_cap = 0;
            const size_t cap = std::max<size_t>(ntasks * 2, 512);
            if (hip_ok(hipHostMalloc(&ctx->h_tasks, cap * sizeof(fa::HostTask), hipHostMallocDefault), "hipHostMalloc") &&
                hip_ok(hipHostGetDevicePointer((void **)&ctx->dv_tasks, ctx->h_tasks, 0), "hipHostGetDevicePointer"))
                ctx->task_cap = cap;
        }
        if (rc == FA_OK) {
            // one workgroup per <= bs envs of each client; inputs and obs stay in the client's pinned buffers
            uint32_t t = 0;
            for (const fa::HostTask &cl : ctx->parts) {
                for (uint32_t off = 0; off < cl.count; off += bs)
                    ctx->h_tasks[t++] = fa::HostTask{cl.first + off, std::min(bs, cl.count - off), cl.in + off,
                                                     cl.obs + (size_t)off * 128u};
            }
            fa::StepArgs g = chunk_args(s, sc, ch);
            g.packed = nullptr;
            g.obs = nullptr;
            g.htasks = ctx->dv_tasks;
            ctx->t_launch = Clock::now();
            hip_ok(fa::launch_step(g, sorted != 0, bs, t, ctx->stream), "frame kernel");
            hip_ok(hipEventRecord(ctx->done, ctx->stream), "hipEventRecord");
        }
        lk.lock();
        if (rc != FA_OK) {
            // the launch did not go out: fail its requests, free the context
            s->error = rc;
            s->error_msg = msg;
            for (int id : ctx->clients) s->clients[id].state = 3;
            ctx->busy = false;
            s->cv_done.notify_all();
        } else {
            s->launches++;
            sc.inflight.push_back(ctx);
            s->cv_inflight.notify_all();
        }
    }
}

// Completer of one class: waits for the class's launches in stream order,
// each for at most kLaunchLimit seconds.
void complete_loop(fa_server *s, size_t ci) {
    ServerClass &sc = s->cls[ci];
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_inflight.wait(lk, [&] { return s->stop_completers || !sc.inflight.empty(); });
        if (sc.inflight.empty()) return;          // stopping, nothing left
        LaunchCtx *ctx = sc.inflight.front();
        lk.unlock();
        hipError_t e = hipSuccess;
        const fa::WaitEnd w = fa::wait_launch(
            [&] {
                e = hipEventQuery(ctx->done);
                return e == hipSuccess ? 0 : (e == hipErrorNotReady ? 1 : 2);
            },
            [] { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); },
            [](uint64_t n) {
                if (n < 2000) std::this_thread::yield();
                else std::this_thread::sleep_for(std::chrono::microseconds(50));
            },
            kLaunchLimit);
        if (w == fa::WaitEnd::TimedOut) {
            // the launch's kernel has not finished: fail this launch's and
            // every queued request of the class, keep its context busy (the
            // kernel may still write its buffers) and stop waiting on the class
            std::lock_guard<std::mutex> g(s->mu);
            s->error = FA_EHIP;
            s->error_msg = "class " + std::to_string(ci) + ": a launch has not completed after " +
                           std::to_string((int)kLaunchLimit) + " s";
            for (int id : ctx->clients) s->clients[id].state = 3;
            for (int id : sc.queue) s->clients[id].state = 3;
            sc.queue.clear();
            s->cv_done.notify_all();
            return;
        }
        if (w == fa::WaitEnd::Failed) {
            std::lock_guard<std::mutex> g(s->mu);
            s->error = FA_EHIP;
            s->error_msg = std::string("launch failed: ") + hipGetErrorString(e);
        }
        finish_launch(ctx);
        lk.lock();
        sc.inflight.pop_front();
    }
}

// A launch's obs are in host memory: complete its requests.
void finish_launch(LaunchCtx *ctx) {
    fa_server *s = ctx->srv;
    const auto now = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int id : ctx->clients) {
            Client &cl = s->clients[id];
            cl.latency_ms = std::chrono::duration<double, std::milli>(now - cl.t_submit).count();
            cl.state = 3;
            s->samples.push_back({cl.latency_ms, cl.cls, cl.n, cl.frames});
            s->env_frames += (uint64_t)cl.n * cl.frames;
            s->requests++;
            if (s->track_completions) s->completions.push_back(id);
        }
        ctx->busy = false;
        // the window grows with the class's launch time: the clients of a
        // class whose frames take tens of ms resubmit over a few ms, and a
        // launch that leaves some behind costs them a whole further launch
        const auto hold = std::max<Clock::duration>(std::chrono::microseconds(s->coalesce_us),
                                                    (now - ctx->t_launch) / 10);
        s->cls[ctx->cls].hold_until = now + hold;
    }
    s->cv_done.notify_all();
    s->cv_dispatch.notify_all();
}

int check_client(fa_server *s, int id) {
    if (!s) return fail(FA_EINVAL, "NULL server");
    bool ok;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        ok = id >= 0 && (size_t)id < s->clients.size() && s->clients[id].open;
    }
    if (!ok) return fail(FA_EINVAL, "no open client " + std::to_string(id));
    return FA_OK;
}

// Restart n envs of a fresh client range from the class's reset snapshot
// (zero frames, restart flag set) on the admin stream; obs = their RAM.
int init_range(fa_server *s, ServerClass &sc, Chunk &ch, Client &cl) {
    std::lock_guard<std::mutex> admin(s->admin_mu);
    for (uint32_t i = 0; i < cl.n; ++i) cl.h_in[i] = (1u << 16) | (1u << 24);   // restart, no action, 0 frames
    FA_HIP_TRY(hipMemcpyAsync(ch.d_in + cl.first, cl.h_in, (size_t)cl.n * 4, hipMemcpyHostToDevice, s->admin));
    std::vector<uint32_t> tasks(2 * ((cl.n + kTaskEnvs - 1) / kTaskEnvs));
    const uint32_t t = fill_tasks(tasks.data(), cl.first, cl.n);
    int rc = FA_OK;
    if (hipMemcpyAsync(s->d_admin_tasks, tasks.data(), (size_t)t * 8, hipMemcpyHostToDevice, s->admin) != hipSuccess)
        rc = fail(FA_EHIP, "task upload failed");
    fa::StepArgs g = chunk_args(s, sc, ch);
    g.tasks = s->d_admin_tasks;
    if (rc == FA_OK && fa::launch_step(g, false, kTaskEnvs, t, s->admin) != hipSuccess)
        rc = fail(FA_EHIP, "init launch failed");
    if (rc == FA_OK &&
        hipMemcpyAsync(cl.h_obs, ch.d_obs + (size_t)cl.first * 128, (size_t)cl.n * 128, hipMemcpyDeviceToHost,
                       s->admin) != hipSuccess)
        rc = fail(FA_EHIP, "obs download failed");
    if (hipStreamSynchronize(s->admin) != hipSuccess && rc == FA_OK) rc = fail(FA_EHIP, "init failed");
    return rc;
}

// HeadlessAtari::new + one run_frame for one env: the class's reset snapshot.
int make_snapshot(fa_server *s, ServerClass &sc) {
    uint8_t *d_states = nullptr;
    uint64_t *d_cnt = nullptr;
    uint32_t *d_status = nullptr;
    FA_HIP_TRY(hipMalloc(&d_states, 8 * fa::kStateBytes));
    FA_HIP_TRY(hipMalloc(&d_cnt, 3 * 8 * 8));
    FA_HIP_TRY(hipMalloc(&d_status, 8 * 4));
    int rc = FA_OK;
    fa::StepArgs g{};
    g.states = d_states;
    g.cpu_cycles = d_cnt;
    g.frame_cycles = d_cnt + 8;
    g.instr = d_cnt + 16;
    g.status = d_status;
    g.desc = s->d_desc;
    g.rom = sc.d_rom;
    g.rom_bytes = sc.rom.rom_bytes;
    g.hot_lo = sc.rom.hot_lo;
    g.hot_n = sc.rom.hot_n;
    g.n = 1;
    g.k = 1;
    g.stride = 1;
    if (fa::launch_fill(sc.rom.power_on, d_states, g.cpu_cycles, g.frame_cycles, g.instr, d_status, 1, 8, s->admin) !=
            hipSuccess ||
        fa::launch_step(g, false, 64, 1, s->admin) != hipSuccess)
        rc = fail(FA_EHIP, "reset snapshot launch failed");
    if (rc == FA_OK && (hipMemcpyAsync(sc.d_snap, d_states, fa::kStateBytes, hipMemcpyDeviceToDevice, s->admin) != hipSuccess ||
                        hipMemcpyAsync(sc.d_snap + 224, d_cnt, 8, hipMemcpyDeviceToDevice, s->admin) != hipSuccess ||
                        hipMemcpyAsync(sc.d_snap + 232, d_cnt + 8, 8, hipMemcpyDeviceToDevice, s->admin) != hipSuccess ||
                        hipMemcpyAsync(sc.d_snap + 240, d_cnt + 16, 8, hipMemcpyDeviceToDevice, s->admin) != hipSuccess))
        rc = fail(FA_EHIP, "reset snapshot copy failed");
    if (hipStreamSynchronize(s->admin) != hipSuccess && rc == FA_OK) rc = fail(FA_EHIP, "reset snapshot failed");
    hipFree(d_states);
    hipFree(d_cnt);
    hipFree(d_status);
    return rc;
}

}  // namespace

extern "C" {

int fa_server_create(const uint8_t *const *roms, const size_t *rom_lens, size_t nclass, size_t budget_bytes,
                     fa_server **out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    *out = nullptr;
    if (!roms || !rom_lens || nclass == 0) return fail(FA_EINVAL, "need at least one ROM class");
    auto *s = new (std::nothrow) fa_server();
    if (!s) return fail(FA_ENOMEM, "out of host memory");
    auto cleanup = [&](int code) {
        fa_server_destroy(s);
        return code;
    };
    s->budget = budget_bytes;
    s->cls.resize(nclass);
    for (size_t c = 0; c < nclass; ++c) {
        int rc = fa::make_rom_class(roms[c], rom_lens[c], s->cls[c].rom);
        if (rc) return cleanup(fail(rc, "ROM class " + std::to_string(c) + ": " + fa::last_error()));
    }
    if (hipGetDevice(&s->device) != hipSuccess) return cleanup(fail(FA_EHIP, "no HIP device"));
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) == hipSuccess && lds > 0)
        s->max_lds = (size_t)lds;
    if (fa::prepare_kernels(s->max_lds) != hipSuccess) return cleanup(fail(FA_EHIP, "kernel attribute setup failed"));
    if (hipStreamCreateWithFlags(&s->admin, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FA_EHIP, "stream creation failed"));
    if (hipMalloc(&s->d_desc, fa::kDescBytes) != hipSuccess ||
        hipMalloc(&s->d_admin_tasks, 2 * (kMaxClientEnvs / kTaskEnvs) * 4) != hipSuccess)
        return cleanup(fail(FA_ENOMEM, "device allocation failed"));
    uint32_t desc[fa::kDescBytes / 4];
    fa::descriptor_table(desc);
    if (hipMemcpy(s->d_desc, desc, sizeof(desc), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FA_EHIP, "descriptor upload failed"));
    for (size_t c = 0; c < nclass; ++c) {
        ServerClass &sc = s->cls[c];
        if (fa::lds_bytes(sc.rom.rom_bytes, kTaskEnvs) > s->max_lds)
            return cleanup(fail(FA_EINVAL, "ROM too large for 256-env workgroups on this device"));
        if (hipMalloc(&sc.d_rom, sc.rom.rom_bytes) != hipSuccess || hipMalloc(&sc.d_snap, 256) != hipSuccess)
            return cleanup(fail(FA_ENOMEM, "device allocation failed"));
        if (hipMemcpy(sc.d_rom, sc.rom.image.data(), sc.rom.rom_bytes, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(FA_EHIP, "ROM upload failed"));
        int rc = make_snapshot(s, sc);
        if (rc) return cleanup(rc);
        if (hipStreamCreateWithFlags(&sc.stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(FA_EHIP, "stream creation failed"));
        for (uint32_t i = 0; i < kCtxPerClass; ++i) {
            auto ctx = std::make_unique<LaunchCtx>();
            ctx->srv = s;
            ctx->cls = (int)c;
            ctx->stream = sc.stream;
            if (hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming | hipEventBlockingSync) != hipSuccess) {
                sc.ctx.push_back(std::move(ctx));
                return cleanup(fail(FA_EHIP, "event creation failed"));
            }
            sc.ctx.push_back(std::move(ctx));
        }
        sc.block = serve_block(sc.rom.rom_bytes, s->max_lds);
    }
    s->dispatcher = std::thread(dispatch_loop, s);
    for (size_t c = 0; c < nclass; ++c) s->completers.emplace_back(complete_loop, s, c);
    *out = s;
    return FA_OK;
}

void fa_server_destroy(fa_server *s) {
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_dispatch.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->stop_completers = true;
    }
    s->cv_inflight.notify_all();
    for (auto &t : s->completers)
        if (t.joinable()) t.join();
    for (auto &sc : s->cls)
        if (sc.stream) hipStreamSynchronize(sc.stream);
    for (auto &sc : s->cls) {
        for (auto &c : sc.ctx) {
            if (c->h_tasks) hipHostFree(c->h_tasks);
            if (c->done) hipEventDestroy(c->done);
        }
        if (sc.stream) hipStreamDestroy(sc.stream);
        for (auto &ch : sc.chunks) free_chunk_bytes(*ch);
        if (sc.d_rom) hipFree(sc.d_rom);
        if (sc.d_snap) hipFree(sc.d_snap);
    }
    for (auto &cl : s->clients) {
        if (cl.h_in) hipHostFree(cl.h_in);
        if (cl.h_obs) hipHostFree(cl.h_obs);
    }
    if (s->d_desc) hipFree(s->d_desc);
    if (s->d_admin_tasks) hipFree(s->d_admin_tasks);
    if (s->admin) hipStreamDestroy(s->admin);
    delete s;
}

int fa_server_set_class_config(fa_server *s, size_t c, uint32_t block, int sorted) {
    if (!s) return fail(FA_EINVAL, "NULL server");
    if (c >= s->cls.size()) return fail(FA_EINVAL, "class index out of range");
    if (block != 0 && block != 64 && block != 128 && block != 256)
        return fail(FA_EINVAL, "block must be 0 (the default), 64, 128 or 256");
    const uint32_t bs = block ? block : serve_block(s->cls[c].rom.rom_bytes, s->max_lds);
    if (fa::lds_bytes(s->cls[c].rom.rom_bytes, bs) > s->max_lds)
        return fail(FA_EINVAL, "block too large for the LDS budget with this ROM");
    std::lock_guard<std::mutex> lk(s->mu);
    s->cls[c].block = bs;
    s->cls[c].sorted = sorted != 0;
    return FA_OK;
}

// Admission: FA_EBUSY when the client's slots do not fit the budget now.
int fa_server_open(fa_server *s, size_t cls, size_t n_envs, uint8_t *obs_out, int *client_id) {
    if (!s || !client_id) return fail(FA_EINVAL, "NULL argument");
    if (cls >= s->cls.size()) return fail(FA_EINVAL, "class index out of range");
    if (n_envs == 0 || n_envs > kMaxClientEnvs) return fail(FA_EINVAL, "client env count must be 1..65536");
    const uint32_t units = (uint32_t)((n_envs + kUnit - 1) / kUnit);
    std::unique_lock<std::mutex> lk(s->mu);
    ServerClass &sc = s->cls[cls];
    int chunk = -1, at = -1;
    for (size_t i = 0; i < sc.chunks.size() && chunk < 0; ++i) {
        at = find_run(*sc.chunks[i], units);
        if (at >= 0) chunk = (int)i;
    }
    if (chunk < 0) {
        int rc = alloc_chunk(s, sc);
        if (rc) return rc;
        chunk = (int)sc.chunks.size() - 1;
        at = 0;
    }
    Chunk &ch = *sc.chunks[chunk];
    for (uint32_t u = 0; u < units; ++u) ch.used[at + u] = 1;
    ch.used_units += units;
    int id;
    if (!s->free_ids.empty()) {
        id = s->free_ids.back();
        s->free_ids.pop_back();
    } else {
        id = (int)s->clients.size();
        s->clients.emplace_back();
    }
    Client &cl = s->clients[id];
    cl = Client{};
    cl.cls = (int)cls;
    cl.chunk = chunk;
    cl.first = (uint32_t)at * kUnit;
    cl.n = (uint32_t)n_envs;
    cl.units = units;
    cl.open = true;
    sc.open_clients++;
    lk.unlock();
    auto undo = [&](int code) {
        std::lock_guard<std::mutex> g(s->mu);
        for (uint32_t u = 0; u < units; ++u) ch.used[at + u] = 0;
        ch.used_units -= units;
        sc.open_clients--;
        if (cl.h_in) hipHostFree(cl.h_in);
        if (cl.h_obs) hipHostFree(cl.h_obs);
        cl = Client{};
        s->free_ids.push_back(id);
        return code;
    };
    if (hipHostMalloc(&cl.h_in, (size_t)n_envs * 4, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&cl.h_obs, (size_t)n_envs * 128, hipHostMallocDefault) != hipSuccess)
        return undo(fail(FA_ENOMEM, "pinned staging allocation failed"));
    if (hipHostGetDevicePointer((void **)&cl.dv_in, cl.h_in, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&cl.dv_obs, cl.h_obs, 0) != hipSuccess)
        return undo(fail(FA_EHIP, "pinned staging is not device-accessible"));
    int rc = init_range(s, sc, ch, cl);
    if (rc) return undo(rc);
    if (obs_out) std::memcpy(obs_out, cl.h_obs, (size_t)n_envs * 128);
    *client_id = id;
    return FA_OK;
}

int fa_server_close(fa_server *s, int id) {
    int rc = check_client(s, id);
    if (rc) return rc;
    std::unique_lock<std::mutex> lk(s->mu);
    Client &cl = s->clients[id];
    s->cv_done.wait(lk, [&] { return cl.state == 0 || cl.state == 3; });
    Chunk &ch = *s->cls[cl.cls].chunks[cl.chunk];
    s->cls[cl.cls].open_clients--;
    const uint32_t u0 = cl.first / kUnit;
    for (uint32_t u = 0; u < cl.units; ++u) ch.used[u0 + u] = 0;
    ch.used_units -= cl.units;
    hipHostFree(cl.h_in);
    hipHostFree(cl.h_obs);
    cl = Client{};
    s->free_ids.push_back(id);
    return FA_OK;
}

// One step for every env of the client: actions (n bytes), each held for
// `frames` frames; restart (n bytes or NULL) restarts an env from the ROM's
// reset snapshot before its step.  Returns at once; FA_EBUSY while the
// client's previous request is unfinished.
int fa_server_submit(fa_server *s, int id, const uint8_t *actions, const uint8_t *restart, uint32_t frames) {
    int rc = check_client(s, id);
    if (rc) return rc;
    if (!actions) return fail(FA_EINVAL, "actions is NULL");
    if (frames > 255) return fail(FA_EINVAL, "frames must be 0..255");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->error) return fail(s->error, s->error_msg);
    Client &cl = s->clients[id];
    if (cl.state == 1 || cl.state == 2) return fail(FA_EBUSY, "client has a request in flight");
    for (uint32_t i = 0; i < cl.n; ++i)
        cl.h_in[i] = (uint32_t)actions[i] | (frames << 8) | ((restart && restart[i]) ? 1u << 16 : 0u);
    cl.frames = frames;
    cl.state = 1;
    cl.t_submit = Clock::now();
    s->cls[cl.cls].queue.push_back(id);
    lk.unlock();
    s->cv_dispatch.notify_all();
    return FA_OK;
}

// 1 when the client's request has completed, 0 while it is queued or running.
int fa_server_poll(fa_server *s, int id) {
    int rc = check_client(s, id);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    return s->clients[id].state == 3 ? 1 : 0;
}

// Blocks until the client's request completes; copies its obs (n * 128) and
// the request's latency (submit to obs in host memory).
int fa_server_wait(fa_server *s, int id, uint8_t *obs_out, double *latency_ms) {
    int rc = check_client(s, id);
    if (rc) return rc;
    std::unique_lock<std::mutex> lk(s->mu);
    Client &cl = s->clients[id];
    if (cl.state == 0) return fail(FA_EINVAL, "client has no request");
    s->cv_done.wait(lk, [&] { return cl.state == 3; });
    if (s->error) return fail(s->error, s->error_msg);
    if (obs_out) std::memcpy(obs_out, cl.h_obs, (size_t)cl.n * 128);
    if (latency_ms) *latency_ms = cl.latency_ms;
    cl.state = 0;
    return FA_OK;
}

// States and counters of an idle client's envs (no request in flight).
int fa_server_download(fa_server *s, int id, AtariStateGpu *states, uint64_t *cpu_cycles, uint64_t *frame_cycles,
                       uint32_t *status) {
    int rc = check_client(s, id);
    if (rc) return rc;
    std::unique_lock<std::mutex> lk(s->mu);
    Client &cl = s->clients[id];
    if (cl.state == 1 || cl.state == 2) return fail(FA_EBUSY, "client has a request in flight");
    Chunk &ch = *s->cls[cl.cls].chunks[cl.chunk];
    const size_t f = cl.first, n = cl.n;
    lk.unlock();
    std::lock_guard<std::mutex> admin(s->admin_mu);
    if (states) FA_HIP_TRY(hipMemcpyAsync(states, ch.d_states + f * fa::kStateBytes, n * fa::kStateBytes,
                                          hipMemcpyDeviceToHost, s->admin));
    if (cpu_cycles) FA_HIP_TRY(hipMemcpyAsync(cpu_cycles, ch.d_cpu + f, n * 8, hipMemcpyDeviceToHost, s->admin));
    if (frame_cycles) FA_HIP_TRY(hipMemcpyAsync(frame_cycles, ch.d_fcyc + f, n * 8, hipMemcpyDeviceToHost, s->admin));
    if (status) FA_HIP_TRY(hipMemcpyAsync(status, ch.d_status + f, n * 4, hipMemcpyDeviceToHost, s->admin));
    FA_HIP_TRY(hipStreamSynchronize(s->admin));
    return FA_OK;
}

int fa_server_client_info(fa_server *s, int id, size_t *cls, size_t *n_envs) {
    int rc = check_client(s, id);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    if (cls) *cls = (size_t)s->clients[id].cls;
    if (n_envs) *n_envs = s->clients[id].n;
    return FA_OK;
}

int fa_server_set_coalesce_us(fa_server *s, uint32_t us) {
    if (!s) return fail(FA_EINVAL, "NULL server");
    if (us > 1000000) return fail(FA_EINVAL, "coalescing window above 1 s");
    std::lock_guard<std::mutex> lk(s->mu);
    s->coalesce_us = us;
    return FA_OK;
}

int fa_server_memory(fa_server *s, size_t *used_bytes, size_t *budget_bytes, size_t *bytes_per_slot) {
    if (!s) return fail(FA_EINVAL, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (used_bytes) *used_bytes = s->used_bytes;
    if (budget_bytes) *budget_bytes = s->budget;
    if (bytes_per_slot) *bytes_per_slot = kSlotBytes;
    return FA_OK;
}

// Completed-request samples since the last call with reset != 0: latency
// (ms), class, envs and frames per env.  Copies min(cap, count) of them.
int fa_server_samples(fa_server *s, double *ms, uint32_t *cls, uint32_t *envs, uint32_t *frames, size_t cap,
                      size_t *count, uint64_t *launches, int reset) {
    if (!s) return fail(FA_EINVAL, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    const size_t n = std::min(cap, s->samples.size());
    for (size_t i = 0; i < n; ++i) {
        if (ms) ms[i] = s->samples[i].ms;
        if (cls) cls[i] = (uint32_t)s->samples[i].cls;
        if (envs) envs[i] = s->samples[i].envs;
        if (frames) frames[i] = s->samples[i].frames;
    }
    if (count) *count = s->samples.size();
    if (launches) *launches = s->launches;
    if (reset) {
        s->samples.clear();
        s->launches = 0;
    }
    return FA_OK;
}


// ---------------------------------------------------------------------------
// Load driver: stands in for the clients.  Per client, actions and restart
// masks come from per-env SplitMix64 streams (the host package's
// ActionStream: state_i = mix(seed ^ i * GOLDEN + GOLDEN), step s uses
// mix(state_i + (s + 1) * GOLDEN); action = (u64 * 18) >> 64; restart when the
// restart stream's u64 < restart_p * 2^64), so any step of any env can be
// regenerated from the client's seed and step count.  Closed loop (sat ==
// NULL): every client resubmits as soon as its request completes.  Open loop:
// Poisson arrivals at rate_mult[i] * sat[i] per second; arrivals wait in the
// client's backlog while its previous request is in flight.
static inline uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
// How long the load driver waits for in-flight requests after its deadline.
static constexpr std::chrono::seconds kDrainLimit{30};
static constexpr uint64_t kRestartSalt = 0x5E5E7000ull;

int fa_loadgen_run(fa_server *s, const int *client_ids, const fa_client_spec *specs, size_t nclients,
                   double duration_s, const double *sat, uint32_t frames, double restart_p, uint64_t arrival_seed,
                   uint64_t max_steps, uint64_t *steps, double *service_ms, double *e2e_ms, uint32_t *sample_cls,
                   size_t cap, size_t *nsamples, double *wall_s) {
    if (!s || !client_ids || !specs || !steps) return fail(FA_EINVAL, "NULL argument");
    if (frames > 255 || restart_p < 0.0 || restart_p > 1.0) return fail(FA_EINVAL, "bad frames or restart_p");
    std::unordered_map<int, size_t> index;
    for (size_t i = 0; i < nclients; ++i) {
        int rc = check_client(s, client_ids[i]);
        if (rc) return rc;
        index[client_ids[i]] = i;
    }
    const uint64_t thresh =
        restart_p >= 1.0 ? ~0ull : (uint64_t)(restart_p * 18446744073709551616.0);
    std::mt19937_64 rng(arrival_seed);
    std::vector<std::exponential_distribution<double>> gap;
    for (size_t i = 0; i < nclients; ++i) {
        double lam = sat ? std::max(1e-9, specs[i].rate_mult * sat[i]) : 1.0;
        gap.emplace_back(lam);
    }
    using dsec = std::chrono::duration<double>;
    const auto t0 = Clock::now();
    const auto t_end = t0 + std::chrono::duration_cast<Clock::duration>(dsec(duration_s));
    std::vector<Clock::time_point> next_arr(nclients, t0);
    if (sat)
        for (size_t i = 0; i < nclients; ++i)
            next_arr[i] = t0 + std::chrono::duration_cast<Clock::duration>(dsec(gap[i](rng)));
    std::vector<std::deque<Clock::time_point>> backlog(nclients);
    std::vector<Clock::time_point> arrival(nclients);
    std::vector<char> inflight(nclients, 0);
    size_t n_inflight = 0, ns = 0;
    std::vector<uint8_t> act, rst;
    std::vector<int> done;
    // per env: the first SplitMix64 output of its action and restart streams
    // (fixed for the run; a step's values are one more mix of these)
    std::vector<std::vector<uint64_t>> a0s(nclients), r0s(nclients);
    for (size_t i = 0; i < nclients; ++i) {
        const fa_client_spec &sp = specs[i];
        a0s[i].resize(sp.n);
        r0s[i].resize(sp.n);
        for (uint32_t e = 0; e < sp.n; ++e) {
            a0s[i][e] = sm_mix((sp.seed ^ ((uint64_t)e * kGolden)) + kGolden);
            r0s[i][e] = sm_mix(((sp.seed ^ kRestartSalt) ^ ((uint64_t)e * kGolden)) + kGolden);
        }
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->completions.clear();
        s->track_completions = true;
    }
    int result = FA_OK;
    while (true) {
        auto now = Clock::now();
        const bool open = now < t_end;
        if (open) {
            for (size_t i = 0; i < nclients; ++i) {
                if (sat) {
                    while (next_arr[i] <= now) {
                        backlog[i].push_back(next_arr[i]);
                        next_arr[i] += std::chrono::duration_cast<Clock::duration>(dsec(gap[i](rng)));
                    }
                } else if (!inflight[i] && backlog[i].empty()) {
                    backlog[i].push_back(now);
                }
            }
        } else if (n_inflight == 0) {
            break;
        } else if (now > t_end + kDrainLimit) {
            // requests still in flight long after the deadline: name the
            // oldest launch that has not completed and return (its kernel is
            // left to finish on its own)
            std::lock_guard<std::mutex> lk(s->mu);
            std::string what = "no launch in flight";
            double age = -1.0;
            for (size_t c = 0; c < s->cls.size(); ++c) {
                if (s->cls[c].inflight.empty()) continue;
                const double a =
                    std::chrono::duration<double, std::milli>(now - s->cls[c].inflight.front()->t_launch).count();
                if (a > age) {
                    age = a;
                    what = "class " + std::to_string(c) + ": launch issued " + std::to_string((long long)a) +
                           " ms ago has not completed";
                }
            }
            result = fail(FA_EHIP, "load driver drain timed out (" + std::to_string(n_inflight) +
                                       " requests in flight); " + what);
            break;
        }
        // submit for every idle client with a waiting arrival
        for (size_t i = 0; open && i < nclients; ++i) {
            if (inflight[i] || backlog[i].empty()) continue;
            if (max_steps && steps[i] >= max_steps) {
                backlog[i].clear();
                continue;
            }
            const fa_client_spec &sp = specs[i];
            act.resize(sp.n);
            rst.resize(sp.n);
            const uint64_t st = steps[i] + 1;
            const uint64_t *a0 = a0s[i].data(), *r0 = r0s[i].data();
            for (uint32_t e = 0; e < sp.n; ++e) {
                const uint64_t za = sm_mix(a0[e] + st * kGolden);
                const uint64_t zr = sm_mix(r0[e] + st * kGolden);
                act[e] = (uint8_t)(((unsigned __int128)za * 18u) >> 64);
                rst[e] = zr < thresh ? 1 : 0;
            }
            int rc = fa_server_submit(s, client_ids[i], act.data(), rst.data(), frames);
            if (rc) {
                result = rc;
                break;
            }
            arrival[i] = backlog[i].front();
            backlog[i].pop_front();
            inflight[i] = 1;
            ++n_inflight;
        }
        if (result) break;
        // completions
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->completions.empty()) {
                auto until = t_end;
                if (now < t_end && sat)
                    for (size_t i = 0; i < nclients; ++i) until = std::min(until, next_arr[i]);
                if (now >= t_end) until = now + std::chrono::milliseconds(5);
                s->cv_done.wait_until(lk, until, [&] { return !s->completions.empty() || s->error; });
            }
            done.swap(s->completions);
            s->completions.clear();
            if (s->error) {
                result = s->error;
                fail(s->error, s->error_msg);
            }
            const auto t_done = Clock::now();
            for (int id : done) {
                auto it = index.find(id);
                if (it == index.end()) continue;
                const size_t i = it->second;
                Client &cl = s->clients[id];
                cl.state = 0;
                if (ns < cap) {
                    if (service_ms) service_ms[ns] = cl.latency_ms;
                    if (e2e_ms) e2e_ms[ns] = std::chrono::duration<double, std::milli>(t_done - arrival[i]).count();
                    if (sample_cls) sample_cls[ns] = (uint32_t)cl.cls;
                }
                ++ns;
                inflight[i] = 0;
                --n_inflight;
                ++steps[i];
            }
            done.clear();
        }
        if (result) break;
        if (max_steps && n_inflight == 0) {
            bool all = true;
            for (size_t i = 0; i < nclients && all; ++i) all = steps[i] >= max_steps;
            if (all) break;
        }
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->track_completions = false;
        s->completions.clear();
    }
    if (nsamples) *nsamples = ns;
    if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    return result;
}

}  // extern "C"
