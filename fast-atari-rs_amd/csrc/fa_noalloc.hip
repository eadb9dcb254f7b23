// fa_noalloc.hip -- fa_step_noalloc: the k-frame step of fa_step_device for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The caller passes the states (AtariStateGpu[n], device memory), the
// actions, the obs buffer and optionally a status array.  Everything else the
// frame kernel needs -- the opcode descriptors, the ROM image, the padded
// state block it stages in whole 8-env granules, and the per-env counters --
// lives in one file-scope device array, `scratch`, of fixed size (room for
// FA_NOALLOC_MAX_ENVS envs).  Each GPU holds its own copy of the array until
// the process ends; its address on the current device is taken once with
// hipGetSymbolAddress and used for every kernel argument and copy.  Calls on
// one GPU share the array: calls that overlap in time on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fa_internal.hpp"

namespace {

using fa::fail;

constexpr size_t align256(size_t x) { return (x + 255u) & ~(size_t)255u; }

constexpr size_t kMaxEnvs = FA_NOALLOC_MAX_ENVS;
constexpr size_t kMaxRom = 32768;                       // F4: eight 4 KB banks
constexpr size_t kOffDesc = 0;
constexpr size_t kOffRom = align256(kOffDesc + fa::kDescBytes);
constexpr size_t kOffStates = align256(kOffRom + kMaxRom + 16);
constexpr size_t kOffCpu = align256(kOffStates + kMaxEnvs * fa::kStateBytes);   // kMaxEnvs is a multiple of 8
constexpr size_t kOffFcyc = kOffCpu + 8 * kMaxEnvs;
constexpr size_t kOffInstr = kOffFcyc + 8 * kMaxEnvs;
constexpr size_t kOffStatus = kOffInstr + 8 * kMaxEnvs;
constexpr size_t kScratchBytes = align256(kOffStatus + 4 * kMaxEnvs);
static_assert(kMaxEnvs % 8 == 0, "state staging reads whole 8-env granules");
static_assert(kScratchBytes == FA_NOALLOC_SCRATCH_BYTES, "header states the scratch size");

}  // namespace

// The scratch array.  In host code this name is only a stand-in for the
// symbol: its device address comes from hipGetSymbolAddress (scratch_on_device).
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// Fixed ROMs of fewer than 4096 bytes: mirror img[i % len] over the window
// (bytes [0, len) are already in place; only [len, 4096) is written).
__global__ void fa_mirror_kernel(uint8_t *img, uint32_t len) {
    const uint32_t i = len + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4096u) img[i] = img[i % len];
}

// Zero the first n entries of the three counter arrays, and of the status
// array when `status` is set (a kernel rather than hipMemsetAsync, so that the
// call stays capturable into a graph).
__global__ void fa_zero_kernel(uint64_t *cpu, uint64_t *fcyc, uint64_t *instr, uint32_t *status, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cpu[i] = 0;
    fcyc[i] = 0;
    instr[i] = 0;
    if (status) status[i] = 0u;
}

struct DevScratch {
    uint8_t *ptr = nullptr;     // device address of `scratch` on this device
    size_t max_lds = 65536;
    bool ready = false;
};
constexpr int kMaxDevices = 64;
std::mutex g_mu;
DevScratch g_dev[kMaxDevices];

// Once per device, with the device current: the array's address, the LDS
// limit, the frame kernel's LDS attribute and the descriptor table.
int scratch_on_device(DevScratch **out) {
    int dev = 0;
    FA_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(FA_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    DevScratch &d = g_dev[dev];
    if (!d.ready) {
        void *p = nullptr;
        FA_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(FA_EHIP, "no device address for the scratch array");
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
            d.max_lds = (size_t)lds;
        FA_HIP_TRY(fa::prepare_kernels(d.max_lds));
        uint32_t desc[fa::kDescBytes / 4];
        fa::descriptor_table(desc);
        FA_HIP_TRY(hipMemcpy(static_cast<uint8_t *>(p) + kOffDesc, desc, sizeof(desc), hipMemcpyHostToDevice));
        d.ptr = static_cast<uint8_t *>(p);
        d.ready = true;
    }
    *out = &d;
    return FA_OK;
}

}  // namespace

extern "C" {

int fa_step_noalloc(const uint8_t *rom, size_t rom_len, AtariStateGpu *d_states, uint32_t *d_status,
                    const uint8_t *d_actions, size_t n, uint32_t k, uint8_t *d_obs, void *stream) {
    if (!rom || rom_len == 0) return fail(FA_EINVAL, "empty ROM");
    if (k == 0) return fail(FA_EINVAL, "k must be >= 1");
    if (n > kMaxEnvs)
        return fail(FA_EINVAL, "n = " + std::to_string(n) + " exceeds FA_NOALLOC_MAX_ENVS (" +
                                   std::to_string(kMaxEnvs) + ")");
    if (n && !d_states) return fail(FA_EINVAL, "d_states is NULL");
    // BankScheme::detect and the staged window (src/atari.rs:97-104, 433-449),
    // as make_rom_class, without building a host image
    int scheme;
    if (rom_len <= 4096) scheme = 0;
    else if (rom_len == 8192) scheme = 1;
    else if (rom_len == 16384) scheme = 2;
    else if (rom_len >= 32768) scheme = 3;
    else return fail(FA_EINVAL, "ROM size " + std::to_string(rom_len) + " is not 1..4096, 8192, 16384 or >= 32768 bytes");
    static const uint32_t kHotLo[4] = {0xFFFu, 0xFF8u, 0xFF6u, 0xFF4u};
    static const uint32_t kHotN[4] = {0, 2, 4, 8};
    static const uint32_t kImage[4] = {4096, 8192, 16384, 32768};
    if (n == 0) return FA_OK;
    DevScratch *d = nullptr;
    int rc = scratch_on_device(&d);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t *base = d->ptr;
    const uint32_t rom_bytes = kImage[scheme];
    const uint32_t bs = fa::default_block(rom_bytes, d->max_lds);
    if (fa::lds_bytes(rom_bytes, bs) > d->max_lds) return fail(FA_EINVAL, "ROM too large for this device's LDS");

    // ROM image (the mirrored 4 KB window of a Fixed ROM, else every bank)
    const size_t copy = scheme == 0 ? rom_len : rom_bytes;
    FA_HIP_TRY(hipMemcpyAsync(base + kOffRom, rom, copy, hipMemcpyHostToDevice, s));
    if (scheme == 0 && rom_len < 4096) {
        const uint32_t rest = 4096u - (uint32_t)rom_len;
        hipLaunchKernelGGL(fa_mirror_kernel, dim3((rest + 255u) / 256u), dim3(256), 0, s, base + kOffRom,
                           (uint32_t)rom_len);
        FA_HIP_TRY(hipGetLastError());
    }
    // the caller's states into the padded block; the counters start at zero,
    // the status at the caller's values (zero without d_status)
    FA_HIP_TRY(hipMemcpyAsync(base + kOffStates, d_states, n * fa::kStateBytes, hipMemcpyDeviceToDevice, s));
    uint64_t *cpu = reinterpret_cast<uint64_t *>(base + kOffCpu);
    uint64_t *fcyc = reinterpret_cast<uint64_t *>(base + kOffFcyc);
    uint64_t *instr = reinterpret_cast<uint64_t *>(base + kOffInstr);
    uint32_t *status = reinterpret_cast<uint32_t *>(base + kOffStatus);
    if (d_status) FA_HIP_TRY(hipMemcpyAsync(status, d_status, n * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(fa_zero_kernel, dim3(((uint32_t)n + 255u) / 256u), dim3(256), 0, s, cpu, fcyc, instr,
                       d_status ? nullptr : status, (uint32_t)n);
    FA_HIP_TRY(hipGetLastError());

    fa::StepArgs g{};
    g.states = base + kOffStates;
    g.cpu_cycles = cpu;
    g.frame_cycles = fcyc;
    g.instr = instr;
    g.status = status;
    g.actions = d_actions;
    g.obs = d_obs;
    g.desc = reinterpret_cast<const uint32_t *>(base + kOffDesc);
    g.rom = base + kOffRom;
    g.rom_bytes = rom_bytes;
    g.hot_lo = kHotLo[scheme];
    g.hot_n = kHotN[scheme];
    g.n = (uint32_t)n;
    g.k = k;
    g.stride = (uint32_t)n;
    FA_HIP_TRY(fa::launch_step(g, false, bs, (uint32_t)((n + bs - 1) / bs), s));

    FA_HIP_TRY(hipMemcpyAsync(d_states, base + kOffStates, n * fa::kStateBytes, hipMemcpyDeviceToDevice, s));
    if (d_status) FA_HIP_TRY(hipMemcpyAsync(d_status, status, n * 4, hipMemcpyDeviceToDevice, s));
    return FA_OK;
}

}  // extern "C"
