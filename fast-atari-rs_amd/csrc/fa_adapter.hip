// fa_adapter.hip -- the env adapter BatchAtariGpuEnv (src/rl4burn_env.rs:268-380),
// the caller of BatchAtariGpu, over the batch handle of fa_lib.hip.
//
// A step is one launch: every env's action (one broadcast value, or one per
// env) is held for `frameskip` frames through the batch's frameskip path,
// which runs the same frames as the reference's step_multi with the action
// repeated (src/rl4burn_env.rs:339-353: set_action writes only input
// registers the ROM cannot change, so setting it once or every frame is the
// same).  The last frame's RAM becomes f32 observations on the device
// (byte / 255, src/rl4burn_env.rs:314-316), written by the kernel straight
// into pinned host memory and copied from there into the caller's buffer.  An
// optional RAM reader turns each env's RAM into a reward (score change) and a
// terminated flag (lives field 0), also on the device; without one, reward
// and terminated are the reference's constants 0 and false.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "fa_internal.hpp"

struct fa_vecenv {
    fa_env *env = nullptr;
    size_t n = 0;
    uint64_t max_steps = FA_DEFAULT_MAX_STEPS;
    uint32_t frameskip = FA_DEFAULT_FRAMESKIP;
    uint64_t steps = 0;
    uint8_t *d_actions = nullptr;      // n: the step's actions
    uint8_t *d_fs = nullptr;           // n: frameskip of every env
    uint32_t fs_held = 0;              // the value d_fs holds (0: not written yet)
    float *h_obs = nullptr;            // n * 128 f32 observations for host copies, in pinned host memory
    float *hv_obs = nullptr;           // its device address, which the f32 kernel writes
    // RAM reader (reward / termination)
    bool reading = false;
    fa_ram_reader reader{};
    int64_t *d_score = nullptr;        // n: score after the last reset / step
    float *d_reward = nullptr;         // n
    uint8_t *d_term = nullptr;         // n
    std::vector<float> h_reward;
    std::vector<uint8_t> h_term;
};

namespace {

using fa::fail;

// byte / 255 in IEEE single precision, round to nearest (the reference's
// `b as f32 / 255.0`); one byte per lane, grid-stride
__global__ void __launch_bounds__(256) fa_obs_f32_kernel(const uint8_t *in, float *out, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = __fdiv_rn((float)in[i], 255.0f);
}

// One env per lane: score from the RAM bytes (BCD or binary, most significant
// first), reward = score - previous score (0 when `first`), terminated when
// the masked lives field is 0.
__global__ void __launch_bounds__(256) fa_reader_kernel(const uint8_t *obs8, uint32_t n, fa_ram_reader r,
                                                        int first, int64_t *score, float *reward, uint8_t *term) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t *ram = obs8 + (size_t)i * FA_OBS_BYTES;
    int64_t s = 0;
    for (uint32_t k = 0; k < r.n_score; ++k) {
        const uint32_t b = ram[r.score_addr[k] & 0x7Fu];
        s = r.score_bcd ? s * 100 + (int64_t)((b >> 4) * 10u + (b & 15u)) : s * 256 + (int64_t)b;
    }
    reward[i] = first ? 0.0f : (float)(s - score[i]);
    score[i] = s;
    const uint32_t lives = ((uint32_t)ram[r.lives_addr & 0x7Fu] >> (r.lives_shift & 7u)) & r.lives_mask;
    term[i] = (r.lives_mask != 0 && lives == 0) ? 1 : 0;
}

hipError_t launch_reader(fa_vecenv *v, int first) {
    hipLaunchKernelGGL(fa_reader_kernel, dim3((unsigned)((v->n + 255) / 256)), dim3(256), 0, fa::env_stream(v->env),
                       (const uint8_t *)fa::env_obs_device(v->env), (uint32_t)v->n, v->reader, first, v->d_score,
                       v->d_reward, v->d_term);
    return hipGetLastError();
}

hipError_t launch_obs_f32(const uint8_t *in, float *out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t blocks = (n + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL(fa_obs_f32_kernel, dim3(grid), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

// frameskip path with the given device actions; obs as f32 into d_out
int launch_step(fa_vecenv *v, const uint8_t *d_act, float *d_out) {
    hipStream_t s = fa::env_stream(v->env);
    if (v->fs_held != v->frameskip) {
        FA_HIP_TRY(hipMemsetAsync(v->d_fs, (int)v->frameskip, v->n, s));
        v->fs_held = v->frameskip;
    }
    uint8_t *obs8 = fa::env_obs_device(v->env);
    int rc = fa_step_frameskip_device(v->env, d_act, v->d_fs, obs8);
    if (rc) return rc;
    FA_HIP_TRY(launch_obs_f32(obs8, d_out, v->n * FA_OBS_BYTES, s));
    if (v->reading) FA_HIP_TRY(launch_reader(v, 0));
    return FA_OK;
}

// Step count and truncation; reward / terminated: the reference's 0 / false,
// or with a reader the sum of the per-env rewards and whether every env is
// terminated (from the host copies finish_host made).
void count_step(fa_vecenv *v, float *reward, int *terminated, int *truncated) {
    v->steps += 1;
    double sum = 0.0;
    bool all = v->reading;
    if (v->reading && (reward || terminated)) {
        for (size_t i = 0; i < v->n; ++i) {
            sum += v->h_reward[i];
            all = all && v->h_term[i];
        }
    }
    if (reward) *reward = (float)sum;        // src/rl4burn_env.rs:347-352 without a reader
    if (terminated) *terminated = all ? 1 : 0;
    if (truncated) *truncated = (v->max_steps > 0 && v->steps >= v->max_steps) ? 1 : 0;
}

int finish_host(fa_vecenv *v, float *obs_out) {
    hipStream_t s = fa::env_stream(v->env);
    if (v->reading) {
        FA_HIP_TRY(hipMemcpyAsync(v->h_reward.data(), v->d_reward, v->n * sizeof(float), hipMemcpyDeviceToHost, s));
        FA_HIP_TRY(hipMemcpyAsync(v->h_term.data(), v->d_term, v->n, hipMemcpyDeviceToHost, s));
    }
    FA_HIP_TRY(hipStreamSynchronize(s));
    if (obs_out) fa::copy_host(obs_out, v->h_obs, v->n * FA_OBS_BYTES * sizeof(float));
    return FA_OK;
}

}  // namespace

extern "C" {

int fa_vecenv_create(const uint8_t *rom, size_t rom_len, size_t n, fa_vecenv **out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    *out = nullptr;
    if (n == 0) return fail(FA_EINVAL, "n must be >= 1");
    auto *v = new (std::nothrow) fa_vecenv();
    if (!v) return fail(FA_ENOMEM, "out of host memory");
    v->n = n;
    int rc = fa_create(rom, rom_len, n, &v->env);
    if (rc) {
        delete v;
        return rc;
    }
    if (hipMalloc(&v->d_actions, n) != hipSuccess || hipMalloc(&v->d_fs, n) != hipSuccess ||
        hipHostMalloc(&v->h_obs, n * FA_OBS_BYTES * sizeof(float), hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer((void **)&v->hv_obs, v->h_obs, 0) != hipSuccess) {
        fa_vecenv_destroy(v);
        return fail(FA_ENOMEM, "device allocation failed");
    }
    *out = v;
    return FA_OK;
}

void fa_vecenv_destroy(fa_vecenv *v) {
    if (!v) return;
    if (v->env) hipStreamSynchronize(fa::env_stream(v->env));
    for (void *p : {(void *)v->d_actions, (void *)v->d_fs, (void *)v->d_score, (void *)v->d_reward,
                    (void *)v->d_term})
        if (p) hipFree(p);
    if (v->h_obs) hipHostFree(v->h_obs);
    fa_destroy(v->env);
    delete v;
}

int fa_vecenv_set_max_steps(fa_vecenv *v, uint64_t max_steps) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    v->max_steps = max_steps;
    return FA_OK;
}

int fa_vecenv_set_frameskip(fa_vecenv *v, uint32_t frameskip) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    if (frameskip < 1 || frameskip > 255) return fail(FA_EINVAL, "frameskip must be 1..255");
    v->frameskip = frameskip;
    return FA_OK;
}

int fa_vecenv_set_sorted(fa_vecenv *v, int sorted) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    return fa_set_sorted(v->env, sorted);
}

size_t fa_vecenv_num_envs(const fa_vecenv *v) { return v ? v->n : 0; }

uint64_t fa_vecenv_step_count(const fa_vecenv *v) { return v ? v->steps : 0; }

fa_env *fa_vecenv_batch(fa_vecenv *v) { return v ? v->env : nullptr; }

int fa_vecenv_reset(fa_vecenv *v, float *obs_out) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    v->steps = 0;
    int rc = fa_reset(v->env, nullptr);
    if (rc) return rc;
    FA_HIP_TRY(launch_obs_f32(fa::env_obs_device(v->env), v->hv_obs, v->n * FA_OBS_BYTES, fa::env_stream(v->env)));
    if (v->reading) FA_HIP_TRY(launch_reader(v, 1));
    return finish_host(v, obs_out);
}

int fa_vecenv_step(fa_vecenv *v, uint32_t action, float *obs_out, float *reward, int *terminated, int *truncated) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    FA_HIP_TRY(hipMemsetAsync(v->d_actions, (int)(uint8_t)action, v->n, fa::env_stream(v->env)));
    int rc = launch_step(v, v->d_actions, v->hv_obs);
    if (rc) return rc;
    rc = finish_host(v, obs_out);
    if (rc) return rc;
    count_step(v, reward, terminated, truncated);
    return FA_OK;
}

int fa_vecenv_step_each(fa_vecenv *v, const uint8_t *actions, float *obs_out, float *reward, int *terminated,
                        int *truncated) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    if (!actions) return fail(FA_EINVAL, "actions is NULL");
    FA_HIP_TRY(hipMemcpyAsync(v->d_actions, actions, v->n, hipMemcpyHostToDevice, fa::env_stream(v->env)));
    int rc = launch_step(v, v->d_actions, v->hv_obs);
    if (rc) return rc;
    rc = finish_host(v, obs_out);
    if (rc) return rc;
    count_step(v, reward, terminated, truncated);
    return FA_OK;
}

int fa_vecenv_step_device(fa_vecenv *v, const uint8_t *d_actions, uint32_t action, float *d_obs, int *truncated) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    if (!d_obs) return fail(FA_EINVAL, "d_obs is NULL");
    if (!d_actions) {
        FA_HIP_TRY(hipMemsetAsync(v->d_actions, (int)(uint8_t)action, v->n, fa::env_stream(v->env)));
        d_actions = v->d_actions;
    }
    int rc = launch_step(v, d_actions, d_obs);
    if (rc) return rc;
    count_step(v, nullptr, nullptr, truncated);
    return FA_OK;
}

int fa_vecenv_sync(fa_vecenv *v) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    return fa_sync(v->env);
}

int fa_vecenv_set_reader(fa_vecenv *v, const fa_ram_reader *r) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    if (!r) {
        v->reading = false;
        return FA_OK;
    }
    if (r->n_score > 4) return fail(FA_EINVAL, "n_score must be 0..4");
    if (r->lives_shift > 7) return fail(FA_EINVAL, "lives_shift must be 0..7");
    for (uint32_t k = 0; k < r->n_score; ++k)
        if (r->score_addr[k] > 127) return fail(FA_EINVAL, "score_addr must be 0..127 (RAM offsets)");
    if (r->lives_addr > 127) return fail(FA_EINVAL, "lives_addr must be 0..127 (RAM offset)");
    if (!v->d_score) {
        if (hipMalloc(&v->d_score, v->n * sizeof(int64_t)) != hipSuccess ||
            hipMalloc(&v->d_reward, v->n * sizeof(float)) != hipSuccess || hipMalloc(&v->d_term, v->n) != hipSuccess)
            return fail(FA_ENOMEM, "device allocation failed");
        v->h_reward.assign(v->n, 0.0f);
        v->h_term.assign(v->n, 0);
    }
    v->reader = *r;
    v->reading = true;
    FA_HIP_TRY(launch_reader(v, 1));
    return finish_host(v, nullptr);
}

int fa_vecenv_rewards(fa_vecenv *v, float *rewards, uint8_t *terminated) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    hipStream_t s = fa::env_stream(v->env);
    if (!v->reading) {
        if (rewards) std::fill(rewards, rewards + v->n, 0.0f);
        if (terminated) std::fill(terminated, terminated + v->n, (uint8_t)0);
        return FA_OK;
    }
    if (rewards) FA_HIP_TRY(hipMemcpyAsync(rewards, v->d_reward, v->n * sizeof(float), hipMemcpyDeviceToHost, s));
    if (terminated) FA_HIP_TRY(hipMemcpyAsync(terminated, v->d_term, v->n, hipMemcpyDeviceToHost, s));
    FA_HIP_TRY(hipStreamSynchronize(s));
    return FA_OK;
}

int fa_obs_to_f32(const uint8_t *d_in, float *d_out, size_t n_bytes) {
    if (n_bytes && (!d_in || !d_out)) return fail(FA_EINVAL, "NULL buffer");
    FA_HIP_TRY(launch_obs_f32(d_in, d_out, n_bytes, nullptr));
    FA_HIP_TRY(hipStreamSynchronize(nullptr));
    return FA_OK;
}

}  // extern "C"
