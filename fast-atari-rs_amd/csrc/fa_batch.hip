// fa_batch.hip -- heterogeneous batches: several ROM classes (cartridge +
// mapper) with their own env counts, stepped together by one hipGraph launch.
//
// Device layout: one AoS state array with every class's envs in its own slot
// range (slot ranges start at multiples of 256 so a workgroup's staged block
// never reaches into the next class), per-env counters / inputs / outputs
// indexed by the caller's global env index, one ROM image per class.  A step
// of K frames-held-for-fs is a graph of per-class kernel chains that fork from
// the batch stream, run concurrently on per-class streams and join back;
// each class has its own launch shape (block size, envs per workgroup, opcode
// grouping, steps per launch), chosen by fa_batch_tune from timed trials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "fa_internal.hpp"

namespace {

using fa::fail;

struct BatchClass {
    fa::RomClass rom;
    size_t first = 0, count = 0;   // global env index range
    size_t slot = 0;               // first state slot
    uint8_t *d_rom = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t join = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    uint32_t block = 0;            // 0: largest that fits the LDS
    uint32_t lanes = 0;            // envs per workgroup below the block (8, 16, 32), 0: the block
    int sorted = 0;
    uint32_t k_split = 0;          // steps per kernel launch, 0: all
    uint32_t *d_tasks[3] = {};     // (first, count) per workgroup for 32, 16 and 8 envs a workgroup
};

// Fewer envs per workgroup than its lanes spread a class over more waves:
// a divergent wave's frame takes longer the more lanes (and so the more
// different instruction kinds) it holds, so a class that leaves SIMDs idle
// can trade lanes for waves.
bool lanes_ok(uint32_t lanes) { return lanes == 0 || lanes == 8 || lanes == 16 || lanes == 32; }
int lanes_index(uint32_t lanes) { return lanes == 32 ? 0 : (lanes == 16 ? 1 : 2); }

constexpr size_t kSlotAlign = 256;

}  // namespace

struct fa_batch {
    int device = 0;
    size_t max_lds = 65536;
    std::vector<BatchClass> cls;
    size_t n = 0, slots = 0;
    uint32_t *d_desc = nullptr;
    uint8_t *d_states = nullptr;
    uint64_t *d_cpu = nullptr, *d_fcyc = nullptr, *d_instr = nullptr;
    uint32_t *d_status = nullptr;
    uint8_t *d_actions = nullptr, *d_fs = nullptr, *d_obs = nullptr;
    uint32_t k_cap = 0;            // steps the input buffers hold
    uint32_t obs_steps_cap = 0;    // steps the obs buffer holds
    // scratch for tuning trials: one class's states and counters
    uint8_t *d_save = nullptr;
    size_t save_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    std::map<uint64_t, hipGraphExec_t> graphs;
    bool timed_launch = false;
    double last_ms = 0.0;
};

namespace {

uint32_t class_block(const fa_batch *b, const BatchClass &c) {
    return c.block ? c.block : fa::default_block(c.rom.rom_bytes, b->max_lds);
}

fa::StepArgs class_args(const fa_batch *b, const BatchClass &c) {
    fa::StepArgs g{};
    g.states = b->d_states + c.slot * fa::kStateBytes;
    g.cpu_cycles = b->d_cpu + c.first;
    g.frame_cycles = b->d_fcyc + c.first;
    g.instr = b->d_instr + c.first;
    g.status = b->d_status + c.first;
    g.desc = b->d_desc;
    g.rom = c.d_rom;
    g.rom_bytes = c.rom.rom_bytes;
    g.hot_lo = c.rom.hot_lo;
    g.hot_n = c.rom.hot_n;
    g.n = (uint32_t)c.count;
    g.stride = (uint32_t)b->n;
    return g;
}

// The kernel chain of one class for k steps, on stream s.
int enqueue_class(fa_batch *b, const BatchClass &c, uint32_t k, bool use_fs, bool use_actions, int obs_mode,
                  uint32_t block, uint32_t lanes, int sorted, uint32_t k_split, hipStream_t s) {
    if (c.count == 0) return FA_OK;
    const uint32_t bs = block;
    if (fa::lds_bytes(c.rom.rom_bytes, bs) > b->max_lds) return fail(FA_EINVAL, "block too large for the LDS budget");
    const bool spread = lanes != 0 && lanes < bs;
    const uint32_t per = spread ? lanes : bs;
    const uint32_t grid = (uint32_t)((c.count + per - 1) / per);
    const uint32_t ks = k_split ? std::min(k_split, k) : k;
    for (uint32_t s0 = 0; s0 < k; s0 += ks) {
        fa::StepArgs g = class_args(b, c);
        const size_t step0 = (size_t)s0 * b->n;
        g.k = std::min(ks, k - s0);
        if (spread) g.tasks = c.d_tasks[lanes_index(lanes)];
        g.actions = use_actions ? b->d_actions + step0 + c.first : nullptr;
        g.fs = use_fs ? b->d_fs + step0 + c.first : nullptr;
        if (obs_mode == 1)
            g.obs = b->d_obs + c.first * 128u;
        else if (obs_mode == 2) {
            g.obs = b->d_obs + (step0 + c.first) * 128u;
            g.obs_every = 1;
        }
        FA_HIP_TRY(fa::launch_step(g, sorted != 0, bs, grid, s));
    }
    return FA_OK;
}

// Forks the per-class chains off b->stream and joins them back.
int enqueue_all(fa_batch *b, uint32_t k, bool use_fs, bool use_actions, int obs_mode) {
    FA_HIP_TRY(hipEventRecord(b->fork, b->stream));
    for (auto &c : b->cls) {
        if (c.count == 0) continue;
        FA_HIP_TRY(hipStreamWaitEvent(c.stream, b->fork, 0));
        int rc = enqueue_class(b, c, k, use_fs, use_actions, obs_mode, class_block(b, c), c.lanes, c.sorted, c.k_split,
                               c.stream);
        if (rc) return rc;
        FA_HIP_TRY(hipEventRecord(c.join, c.stream));
        FA_HIP_TRY(hipStreamWaitEvent(b->stream, c.join, 0));
    }
    return FA_OK;
}

void drop_graphs(fa_batch *b) {
    for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

int ensure_inputs(fa_batch *b, uint32_t k, int obs_mode) {
    if (k > b->k_cap) {
        if (b->d_actions) FA_HIP_TRY(hipFree(b->d_actions));
        if (b->d_fs) FA_HIP_TRY(hipFree(b->d_fs));
        b->d_actions = b->d_fs = nullptr;
        b->k_cap = 0;
        const size_t bytes = (size_t)k * std::max<size_t>(b->n, 1);
        FA_HIP_TRY(hipMalloc(&b->d_actions, bytes));
        FA_HIP_TRY(hipMalloc(&b->d_fs, bytes));
        b->k_cap = k;
        drop_graphs(b);   // captured graphs hold the old addresses
    }
    const uint32_t steps = obs_mode == 2 ? k : 1u;
    if (steps > b->obs_steps_cap) {
        if (b->d_obs) FA_HIP_TRY(hipFree(b->d_obs));
        b->d_obs = nullptr;
        b->obs_steps_cap = 0;
        FA_HIP_TRY(hipMalloc(&b->d_obs, (size_t)steps * std::max<size_t>(b->n, 1) * 128u));
        b->obs_steps_cap = steps;
        drop_graphs(b);
    }
    return FA_OK;
}

int graph_for(fa_batch *b, uint32_t k, bool use_fs, int obs_mode, hipGraphExec_t *out) {
    const uint64_t key = ((uint64_t)k << 8) | ((uint64_t)use_fs << 4) | (uint64_t)obs_mode;
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return FA_OK;
    }
    FA_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_all(b, k, use_fs, true, obs_mode);
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(FA_EHIP, std::string("graph capture failed: ") + hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ec != hipSuccess) return fail(FA_EHIP, std::string("graph instantiate failed: ") + hipGetErrorString(ec));
    b->graphs[key] = exec;
    *out = exec;
    return FA_OK;
}

int check_batch(const fa_batch *b) { return b ? FA_OK : fail(FA_EINVAL, "NULL batch"); }

}  // namespace

extern "C" {

int fa_batch_create(const uint8_t *const *roms, const size_t *rom_lens, const size_t *counts, size_t nclass,
                    fa_batch **out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    *out = nullptr;
    if (!roms || !rom_lens || !counts || nclass == 0) return fail(FA_EINVAL, "need at least one ROM class");
    auto *b = new (std::nothrow) fa_batch();
    if (!b) return fail(FA_ENOMEM, "out of host memory");
    auto cleanup = [&](int code) {
        fa_batch_destroy(b);
        return code;
    };
    b->cls.resize(nclass);
    for (size_t c = 0; c < nclass; ++c) {
        BatchClass &bc = b->cls[c];
        int rc = fa::make_rom_class(roms[c], rom_lens[c], bc.rom);
        if (rc) return cleanup(fail(rc, "ROM class " + std::to_string(c) + ": " + fa::last_error()));
        bc.first = b->n;
        bc.count = counts[c];
        bc.slot = b->slots;
        b->n += counts[c];
        b->slots += (counts[c] + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
    }
    if (b->n >= (1ull << 31)) return cleanup(fail(FA_EINVAL, "too many envs"));
    if (hipGetDevice(&b->device) != hipSuccess) return cleanup(fail(FA_EHIP, "no HIP device"));
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device) == hipSuccess && lds > 0)
        b->max_lds = (size_t)lds;
    if (fa::prepare_kernels(b->max_lds) != hipSuccess) return cleanup(fail(FA_EHIP, "kernel attribute setup failed"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess)
        return cleanup(fail(FA_EHIP, "stream / event creation failed"));
    const size_t n1 = std::max<size_t>(b->n, 1), s1 = std::max<size_t>(b->slots, 1);
    if (hipMalloc(&b->d_desc, fa::kDescBytes) != hipSuccess || hipMalloc(&b->d_states, s1 * fa::kStateBytes) != hipSuccess ||
        hipMalloc(&b->d_cpu, n1 * 8) != hipSuccess || hipMalloc(&b->d_fcyc, n1 * 8) != hipSuccess ||
        hipMalloc(&b->d_instr, n1 * 8) != hipSuccess || hipMalloc(&b->d_status, n1 * 4) != hipSuccess)
        return cleanup(fail(FA_ENOMEM, "device allocation failed"));
    uint32_t desc[fa::kDescBytes / 4];
    fa::descriptor_table(desc);
    if (hipMemcpy(b->d_desc, desc, sizeof(desc), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FA_EHIP, "descriptor upload failed"));
    for (auto &c : b->cls) {
        if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c.join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreate(&c.t0) != hipSuccess || hipEventCreate(&c.t1) != hipSuccess)
            return cleanup(fail(FA_EHIP, "stream / event creation failed"));
        if (hipMalloc(&c.d_rom, c.rom.rom_bytes) != hipSuccess) return cleanup(fail(FA_ENOMEM, "ROM allocation failed"));
        if (hipMemcpy(c.d_rom, c.rom.image.data(), c.rom.rom_bytes, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(FA_EHIP, "ROM upload failed"));
        // task lists for the spread launch shapes (env ranges start at multiples of 8)
        for (int li = 0; li < 3 && c.count; ++li) {
            const uint32_t per = 32u >> li;
            std::vector<uint32_t> t;
            for (size_t off = 0; off < c.count; off += per) {
                t.push_back((uint32_t)off);
                t.push_back((uint32_t)std::min<size_t>(per, c.count - off));
            }
            if (hipMalloc(&c.d_tasks[li], t.size() * 4) != hipSuccess) return cleanup(fail(FA_ENOMEM, "task list allocation failed"));
            if (hipMemcpy(c.d_tasks[li], t.data(), t.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
                return cleanup(fail(FA_EHIP, "task list upload failed"));
        }
    }
    if (ensure_inputs(b, 1, 1)) return cleanup(FA_ENOMEM);
    // power-on state until reset
    for (auto &c : b->cls) {
        const uint32_t pad = (uint32_t)((c.count + kSlotAlign - 1) / kSlotAlign * kSlotAlign);
        if (fa::launch_fill(c.rom.power_on, b->d_states + c.slot * fa::kStateBytes, b->d_cpu + c.first,
                            b->d_fcyc + c.first, b->d_instr + c.first, b->d_status + c.first, (uint32_t)c.count, pad,
                            b->stream) != hipSuccess)
            return cleanup(fail(FA_EHIP, "power-on fill failed"));
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return cleanup(fail(FA_EHIP, "power-on fill failed"));
    *out = b;
    return FA_OK;
}

void fa_batch_destroy(fa_batch *b) {
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (auto &c : b->cls) {
        if (c.stream) hipStreamSynchronize(c.stream);
        if (c.d_rom) hipFree(c.d_rom);
        for (uint32_t *t : c.d_tasks)
            if (t) hipFree(t);
        if (c.join) hipEventDestroy(c.join);
        if (c.t0) hipEventDestroy(c.t0);
        if (c.t1) hipEventDestroy(c.t1);
        if (c.stream) hipStreamDestroy(c.stream);
    }
    for (void *p : {(void *)b->d_desc, (void *)b->d_states, (void *)b->d_cpu, (void *)b->d_fcyc, (void *)b->d_instr,
                    (void *)b->d_status, (void *)b->d_actions, (void *)b->d_fs, (void *)b->d_obs, (void *)b->d_save})
        if (p) hipFree(p);
    for (hipEvent_t ev : {b->fork, b->t0, b->t1})
        if (ev) hipEventDestroy(ev);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

size_t fa_batch_num_envs(const fa_batch *b) { return b ? b->n : 0; }
size_t fa_batch_num_classes(const fa_batch *b) { return b ? b->cls.size() : 0; }

int fa_batch_class_range(const fa_batch *b, size_t c, size_t *first, size_t *count) {
    if (check_batch(b)) return FA_EINVAL;
    if (c >= b->cls.size()) return fail(FA_EINVAL, "class index out of range");
    if (first) *first = b->cls[c].first;
    if (count) *count = b->cls[c].count;
    return FA_OK;
}

int fa_batch_reset(fa_batch *b, uint8_t *obs_out) {
    if (check_batch(b)) return FA_EINVAL;
    for (auto &c : b->cls) {
        const uint32_t pad = (uint32_t)((c.count + kSlotAlign - 1) / kSlotAlign * kSlotAlign);
        FA_HIP_TRY(fa::launch_fill(c.rom.power_on, b->d_states + c.slot * fa::kStateBytes, b->d_cpu + c.first,
                                   b->d_fcyc + c.first, b->d_instr + c.first, b->d_status + c.first,
                                   (uint32_t)c.count, pad, b->stream));
    }
    int rc = enqueue_all(b, 1, false, false, 1);
    if (rc) return rc;
    if (obs_out && b->n) FA_HIP_TRY(hipMemcpyAsync(obs_out, b->d_obs, b->n * 128, hipMemcpyDeviceToHost, b->stream));
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    return FA_OK;
}

int fa_batch_buffers(fa_batch *b, uint32_t k, int obs_every, uint8_t **d_actions, uint8_t **d_fs, uint8_t **d_obs) {
    if (check_batch(b)) return FA_EINVAL;
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    int rc = ensure_inputs(b, k, obs_every ? 2 : 1);
    if (rc) return rc;
    if (d_actions) *d_actions = b->d_actions;
    if (d_fs) *d_fs = b->d_fs;
    if (d_obs) *d_obs = b->d_obs;
    return FA_OK;
}

int fa_batch_launch(fa_batch *b, uint32_t k, int use_fs, int obs_every) {
    if (check_batch(b)) return FA_EINVAL;
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    const int obs_mode = obs_every ? 2 : 1;
    int rc = ensure_inputs(b, k, obs_mode);
    if (rc) return rc;
    hipGraphExec_t exec = nullptr;
    rc = graph_for(b, k, use_fs != 0, obs_mode, &exec);
    if (rc) return rc;
    FA_HIP_TRY(hipEventRecord(b->t0, b->stream));
    FA_HIP_TRY(hipGraphLaunch(exec, b->stream));
    FA_HIP_TRY(hipEventRecord(b->t1, b->stream));
    b->timed_launch = true;
    return FA_OK;
}

int fa_batch_sync(fa_batch *b, double *launch_ms) {
    if (check_batch(b)) return FA_EINVAL;
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->timed_launch) {
        float ms = 0.f;
        FA_HIP_TRY(hipEventElapsedTime(&ms, b->t0, b->t1));
        b->last_ms = ms;
        b->timed_launch = false;
    }
    if (launch_ms) *launch_ms = b->last_ms;
    return FA_OK;
}

int fa_batch_step(fa_batch *b, const uint8_t *actions, const uint8_t *fs, uint32_t k, int obs_every,
                  uint8_t *obs_out) {
    if (check_batch(b)) return FA_EINVAL;
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    if (!actions && b->n) return fail(FA_EINVAL, "actions is NULL");
    int rc = ensure_inputs(b, k, obs_every ? 2 : 1);
    if (rc) return rc;
    const size_t bytes = (size_t)k * b->n;
    if (bytes) FA_HIP_TRY(hipMemcpyAsync(b->d_actions, actions, bytes, hipMemcpyHostToDevice, b->stream));
    if (fs && bytes) FA_HIP_TRY(hipMemcpyAsync(b->d_fs, fs, bytes, hipMemcpyHostToDevice, b->stream));
    rc = fa_batch_launch(b, k, fs != nullptr, obs_every);
    if (rc) return rc;
    const size_t obs_bytes = (obs_every ? (size_t)k : 1u) * b->n * 128u;
    if (obs_out && obs_bytes) FA_HIP_TRY(hipMemcpyAsync(obs_out, b->d_obs, obs_bytes, hipMemcpyDeviceToHost, b->stream));
    return fa_batch_sync(b, nullptr);
}

int fa_batch_download_states(fa_batch *b, AtariStateGpu *out, size_t n) {
    if (check_batch(b) || !out) return fail(FA_EINVAL, "NULL argument");
    if (n != b->n) return fail(FA_EINVAL, "n must be the batch's env count");
    for (auto &c : b->cls)
        if (c.count)
            FA_HIP_TRY(hipMemcpyAsync(out + c.first, b->d_states + c.slot * fa::kStateBytes, c.count * fa::kStateBytes,
                                      hipMemcpyDeviceToHost, b->stream));
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    return FA_OK;
}

int fa_batch_upload_states(fa_batch *b, const AtariStateGpu *in, size_t n) {
    if (check_batch(b) || !in) return fail(FA_EINVAL, "NULL argument");
    if (n != b->n) return fail(FA_EINVAL, "n must be the batch's env count");
    for (auto &c : b->cls)
        for (size_t i = 0; i < c.count; ++i) {
            int rc = fa::check_state(c.rom, in[c.first + i], c.first + i);
            if (rc) return rc;
        }
    for (auto &c : b->cls)
        if (c.count)
            FA_HIP_TRY(hipMemcpyAsync(b->d_states + c.slot * fa::kStateBytes, in + c.first, c.count * fa::kStateBytes,
                                      hipMemcpyHostToDevice, b->stream));
    if (b->n) FA_HIP_TRY(hipMemsetAsync(b->d_status, 0, b->n * 4, b->stream));
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    return FA_OK;
}

int fa_batch_download_counters(fa_batch *b, uint64_t *cpu_cycles, uint64_t *frame_cycles, uint64_t *instructions,
                               uint32_t *status) {
    if (check_batch(b)) return FA_EINVAL;
    const size_t n = b->n;
    if (n == 0) return FA_OK;
    if (cpu_cycles) FA_HIP_TRY(hipMemcpyAsync(cpu_cycles, b->d_cpu, n * 8, hipMemcpyDeviceToHost, b->stream));
    if (frame_cycles) FA_HIP_TRY(hipMemcpyAsync(frame_cycles, b->d_fcyc, n * 8, hipMemcpyDeviceToHost, b->stream));
    if (instructions) FA_HIP_TRY(hipMemcpyAsync(instructions, b->d_instr, n * 8, hipMemcpyDeviceToHost, b->stream));
    if (status) FA_HIP_TRY(hipMemcpyAsync(status, b->d_status, n * 4, hipMemcpyDeviceToHost, b->stream));
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    return FA_OK;
}

int fa_batch_set_class_config(fa_batch *b, size_t c, uint32_t block, int sorted, uint32_t k_split) {
    if (check_batch(b)) return FA_EINVAL;
    if (c >= b->cls.size()) return fail(FA_EINVAL, "class index out of range");
    if (block != 0 && (block % 64 != 0 || block > fa::kMaxBlock))
        return fail(FA_EINVAL, "block must be 0, 64, 128, 192 or 256");
    if (block && fa::lds_bytes(b->cls[c].rom.rom_bytes, block) > b->max_lds)
        return fail(FA_EINVAL, "block too large for the LDS budget with this ROM");
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    b->cls[c].block = block;
    b->cls[c].sorted = sorted != 0;
    b->cls[c].k_split = k_split;
    drop_graphs(b);
    return FA_OK;
}

int fa_batch_set_class_lanes(fa_batch *b, size_t c, uint32_t lanes) {
    if (check_batch(b)) return FA_EINVAL;
    if (c >= b->cls.size()) return fail(FA_EINVAL, "class index out of range");
    if (!lanes_ok(lanes)) return fail(FA_EINVAL, "lanes must be 0, 8, 16 or 32");
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    b->cls[c].lanes = lanes;
    drop_graphs(b);
    return FA_OK;
}

int fa_batch_get_class_lanes(const fa_batch *b, size_t c, uint32_t *lanes) {
    if (check_batch(b)) return FA_EINVAL;
    if (c >= b->cls.size()) return fail(FA_EINVAL, "class index out of range");
    if (!lanes) return fail(FA_EINVAL, "lanes is NULL");
    *lanes = b->cls[c].lanes;
    return FA_OK;
}

int fa_batch_get_class_config(const fa_batch *b, size_t c, uint32_t *block, int *sorted, uint32_t *k_split) {
    if (check_batch(b)) return FA_EINVAL;
    if (c >= b->cls.size()) return fail(FA_EINVAL, "class index out of range");
    if (block) *block = class_block(b, b->cls[c]);
    if (sorted) *sorted = b->cls[c].sorted;
    if (k_split) *k_split = b->cls[c].k_split;
    return FA_OK;
}

// Autotuner: for each class, times trial launches of that class alone on the
// inputs currently in the batch's buffers (k steps; frame counts from the fs
// buffer when use_fs), restoring the class's states and counters after every
// trial.  Stage 1 picks opcode grouping and block size on a one-step trial,
// and for 64-env workgroups also 32, 16 or 8 envs per workgroup (more, less
// full waves); stage 2 picks steps per launch (a divisor chain of k) on the
// full k steps.
// The winners become the class configs; trial times land in `trials`
// (per class: [ms of the chosen config, ms of the default config]).
int fa_batch_tune(fa_batch *b, uint32_t k, int use_fs, double *trials) {
    if (check_batch(b)) return FA_EINVAL;
    if (k == 0 || k > b->k_cap) return fail(FA_EINVAL, "k must be >= 1 and fit the input buffers (fa_batch_buffers)");
    int rc = ensure_inputs(b, k, 1);
    if (rc) return rc;
    FA_HIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t ci = 0; ci < b->cls.size(); ++ci) {
        BatchClass &c = b->cls[ci];
        if (c.count == 0) continue;
        const size_t sbytes = c.count * fa::kStateBytes, cbytes = c.count * 8;
        const size_t need = sbytes + 3 * cbytes + c.count * 4 + 64;
        if (need > b->save_cap) {
            if (b->d_save) FA_HIP_TRY(hipFree(b->d_save));
            b->d_save = nullptr;
            b->save_cap = 0;
            FA_HIP_TRY(hipMalloc(&b->d_save, need));
            b->save_cap = need;
        }
        uint8_t *st = b->d_states + c.slot * fa::kStateBytes;
        uint8_t *sv = b->d_save;
        auto copy_all = [&](bool save) -> int {
            struct {
                void *dev;
                size_t bytes;
            } parts[5] = {{st, sbytes},
                          {b->d_cpu + c.first, cbytes},
                          {b->d_fcyc + c.first, cbytes},
                          {b->d_instr + c.first, cbytes},
                          {b->d_status + c.first, c.count * 4}};
            size_t off = 0;
            for (auto &p : parts) {
                if (save)
                    FA_HIP_TRY(hipMemcpyAsync(sv + off, p.dev, p.bytes, hipMemcpyDeviceToDevice, c.stream));
                else
                    FA_HIP_TRY(hipMemcpyAsync(p.dev, sv + off, p.bytes, hipMemcpyDeviceToDevice, c.stream));
                off += (p.bytes + 15) / 16 * 16;
            }
            return FA_OK;
        };
        auto trial = [&](uint32_t steps, uint32_t block, uint32_t lanes, int sorted, uint32_t ks, double *ms) -> int {
            int r = copy_all(false);
            if (r) return r;
            FA_HIP_TRY(hipEventRecord(c.t0, c.stream));
            r = enqueue_class(b, c, steps, use_fs != 0, true, 0, block, lanes, sorted, ks, c.stream);
            if (r) return r;
            FA_HIP_TRY(hipEventRecord(c.t1, c.stream));
            FA_HIP_TRY(hipStreamSynchronize(c.stream));
            float f = 0.f;
            FA_HIP_TRY(hipEventElapsedTime(&f, c.t0, c.t1));
            *ms = f;
            return FA_OK;
        };
        rc = copy_all(true);
        if (rc) return rc;
        // stage 1: block x grouping on one step
        uint32_t best_block = fa::default_block(c.rom.rom_bytes, b->max_lds), best_lanes = 0;
        int best_sorted = 0;
        double best = 1e30;
        for (int sorted = 0; sorted < 2; ++sorted)
            for (uint32_t block : {64u, 128u, 256u}) {
                if (fa::lds_bytes(c.rom.rom_bytes, block) > b->max_lds) continue;
                for (uint32_t lanes : {0u, 32u, 16u, 8u}) {
                    if (lanes && block != 64u) break;
                    double ms = 0;
                    rc = trial(1, block, lanes, sorted, 0, &ms);
                    if (rc) return rc;
                    if (ms < best) {
                        best = ms;
                        best_block = block;
                        best_lanes = lanes;
                        best_sorted = sorted;
                    }
                }
            }
        // stage 2: steps per launch on the full k
        double def_ms = 0, best_k = 1e30;
        uint32_t best_split = 0;
        rc = trial(k, fa::default_block(c.rom.rom_bytes, b->max_lds), 0, 0, 0, &def_ms);
        if (rc) return rc;
        for (uint32_t ks = k;; ks = (ks + 1) / 2) {
            double ms = 0;
            rc = trial(k, best_block, best_lanes, best_sorted, ks, &ms);
            if (rc) return rc;
            if (ms < best_k) {
                best_k = ms;
                best_split = ks;
            }
            if (ks == 1) break;
        }
        rc = copy_all(false);
        if (rc) return rc;
        FA_HIP_TRY(hipStreamSynchronize(c.stream));
        c.block = best_block;
        c.lanes = best_lanes;
        c.sorted = best_sorted;
        c.k_split = best_split == k ? 0 : best_split;
        if (trials) {
            trials[2 * ci] = best_k;
            trials[2 * ci + 1] = def_ms;
        }
    }
    drop_graphs(b);
    return FA_OK;
}

}  // extern "C"
