// fa_adapter.hip -- the env adapter BatchAtariGpuEnv (src/rl4burn_env.rs:268-380),
// the caller of BatchAtariGpu, over the batch handle of fa_lib.hip.
//
// A step is one launch: every env's action (one broadcast value, or one per
// env) is held for `frameskip` frames through the batch's frameskip path,
// which runs the same frames as the reference's step_multi with the action
// repeated (src/rl4burn_env.rs:339-353: set_action writes only input
// registers the ROM cannot change, so setting it once or every frame is the
// same).  The last frame's RAM becomes f32 observations on the device
// (byte / 255, src/rl4burn_env.rs:314-316) before one copy to the host.
#include <hip/hip_runtime.h>

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
    float *d_obs = nullptr;            // n * 128 f32 observations for host copies
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
    return FA_OK;
}

void count_step(fa_vecenv *v, float *reward, int *terminated, int *truncated) {
    v->steps += 1;
    if (reward) *reward = 0.0f;              // src/rl4burn_env.rs:347-352
    if (terminated) *terminated = 0;
    if (truncated) *truncated = (v->max_steps > 0 && v->steps >= v->max_steps) ? 1 : 0;
}

int finish_host(fa_vecenv *v, float *obs_out) {
    hipStream_t s = fa::env_stream(v->env);
    if (obs_out)
        FA_HIP_TRY(hipMemcpyAsync(obs_out, v->d_obs, v->n * FA_OBS_BYTES * sizeof(float), hipMemcpyDeviceToHost, s));
    FA_HIP_TRY(hipStreamSynchronize(s));
    return FA_OK;
}

}  // namespace

extern "C" {

int fa_vecenv_create(const uint8_t *rom, size_t rom_len, size_t n, fa_vecenv **out) {
    if (!out) return fail(FA_EINVAL, "out is NULL");
    *out = nullptr;
    if (n == 0) return fail(FA_EINVAL, "n must be >= 1");
    auto *v = new fa_vecenv();
    v->n = n;
    int rc = fa_create(rom, rom_len, n, &v->env);
    if (rc) {
        delete v;
        return rc;
    }
    if (hipMalloc(&v->d_actions, n) != hipSuccess || hipMalloc(&v->d_fs, n) != hipSuccess ||
        hipMalloc(&v->d_obs, n * FA_OBS_BYTES * sizeof(float)) != hipSuccess) {
        fa_vecenv_destroy(v);
        return fail(FA_ENOMEM, "device allocation failed");
    }
    *out = v;
    return FA_OK;
}

void fa_vecenv_destroy(fa_vecenv *v) {
    if (!v) return;
    if (v->env) hipStreamSynchronize(fa::env_stream(v->env));
    for (void *p : {(void *)v->d_actions, (void *)v->d_fs, (void *)v->d_obs})
        if (p) hipFree(p);
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
    FA_HIP_TRY(launch_obs_f32(fa::env_obs_device(v->env), v->d_obs, v->n * FA_OBS_BYTES, fa::env_stream(v->env)));
    return finish_host(v, obs_out);
}

int fa_vecenv_step(fa_vecenv *v, uint32_t action, float *obs_out, float *reward, int *terminated, int *truncated) {
    if (!v) return fail(FA_EINVAL, "NULL handle");
    FA_HIP_TRY(hipMemsetAsync(v->d_actions, (int)(uint8_t)action, v->n, fa::env_stream(v->env)));
    int rc = launch_step(v, v->d_actions, v->d_obs);
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
    int rc = launch_step(v, v->d_actions, v->d_obs);
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

int fa_obs_to_f32(const uint8_t *d_in, float *d_out, size_t n_bytes) {
    if (n_bytes && (!d_in || !d_out)) return fail(FA_EINVAL, "NULL buffer");
    FA_HIP_TRY(launch_obs_f32(d_in, d_out, n_bytes, nullptr));
    FA_HIP_TRY(hipStreamSynchronize(nullptr));
    return FA_OK;
}

}  // extern "C"
