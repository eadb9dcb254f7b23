// fa_kernels.hip -- the frame-step kernel and its launchers.
//
//   fa_frame_kernel<false>  the plain path: convergent waves take the
//                           opcode-specialised fast path, divergent waves the
//                           general per-lane path
//   fa_frame_kernel<true>   wave64 opcode grouping (set_sorted): lanes are
//                           served one distinct opcode at a time (64-bit
//                           ballot + readlane), every group on the
//                           opcode-specialised fast path
//   fa_fill_kernel          power-on broadcast used by reset
//
// One frame kernel serves every entry point: step (k = 1), step_multi (k
// frames, state kept in VGPRs / LDS across frames), frameskip (k steps of
// per-env frame counts), heterogeneous batches (one launch per ROM class) and
// the server (task lists of client slot ranges, per-env restart from the
// ROM's reset snapshot).
#include <hip/hip_runtime.h>

#include <cstring>

#include "fa_frame.hpp"

static_assert(sizeof(AtariStateGpu) == 222, "AtariStateGpu must be 222 bytes");
static_assert(offsetof(AtariStateGpu, ram) == fa::S_RAM, "ram offset");
static_assert(offsetof(AtariStateGpu, timer_value) == fa::S_TVAL, "timer offset");
static_assert(offsetof(AtariStateGpu, tia_clock) == fa::S_CLOCK, "clock offset");
static_assert(offsetof(AtariStateGpu, pos_p0) == fa::S_POS, "pos offset");
static_assert(offsetof(AtariStateGpu, grp0) == fa::S_GRP0, "grp0 offset");
static_assert(offsetof(AtariStateGpu, colup0) == fa::S_COLUP0, "colup0 offset");
static_assert(offsetof(AtariStateGpu, bank) == fa::S_BANK, "bank offset");
static_assert(fa::ROM0 == fa::kLdsRom0, "LDS layout");
static_assert(fa::STATE_BYTES == fa::kStateBytes, "state size");

namespace fa {

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

hipError_t prepare_kernels(size_t max_lds) {
    hipError_t e = hipFuncSetAttribute((const void *)fa_frame_kernel<false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    if (e != hipSuccess) return e;
    return prepare_sorted_kernel(max_lds);
}

hipError_t launch_step(const StepArgs &g, bool sorted, uint32_t bs, uint32_t grid, hipStream_t s) {
    const size_t lds = lds_bytes(g.rom_bytes, bs);
    if (sorted) return launch_sorted(g, bs, grid, lds, s);
    hipLaunchKernelGGL(fa_frame_kernel<false>, dim3(grid), dim3(bs), lds, s, g);
    return hipGetLastError();
}

hipError_t launch_fill(const AtariStateGpu &tmpl, uint8_t *states, uint64_t *cpu, uint64_t *fcyc, uint64_t *instr,
                       uint32_t *status, uint32_t n, uint32_t n_pad, hipStream_t s) {
    FillArgs fg{};
    std::memcpy(fg.tmpl, &tmpl, sizeof(AtariStateGpu));
    fg.states = states;
    fg.cpu_cycles = cpu;
    fg.frame_cycles = fcyc;
    fg.instr = instr;
    fg.status = status;
    fg.n = n;
    fg.n_pad = n_pad;
    hipLaunchKernelGGL(fa_fill_kernel, dim3((n_pad + 255u) / 256u), dim3(256), 0, s, fg);
    return hipGetLastError();
}

void descriptor_table(uint32_t out[kDescBytes / 4]) { build_descriptors(out); }

}  // namespace fa
