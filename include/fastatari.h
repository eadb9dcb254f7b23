/*
 * fastatari.h -- C ABI of the MI355X batched Atari 2600 frame step.
 *
 * Drop-in boundary for the reference's GPU path (RPP1011/fast-atari-rs):
 * every entry point below names the reference interface it replaces.  The
 * handle owns all device memory; host buffers are plain pointers and sizes.
 * Every function returns FA_OK (0) or a negative FA_E* code; fa_last_error()
 * gives a message for the calling thread's last failure.  A handle is not
 * thread-safe (as the reference's `&mut self` methods), calls on one handle
 * are serialised on the handle's own HIP stream.
 */
#ifndef FASTATARI_H
#define FASTATARI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FA_OK 0
#define FA_EINVAL -1     /* bad argument (wrong length, NULL pointer, bad ROM size) */
#define FA_EHIP -2       /* HIP runtime error */
#define FA_ESTATE -3     /* an uploaded state is not one the emulator can reach */
#define FA_ENOMEM -4     /* device allocation failed */
#define FA_EBUSY -5      /* server: over the memory budget, or a request already in flight */

/* ---- per-env emulation status (fa_download_counters) ---- */
#define FA_ENV_OK 0
#define FA_ENV_ILLEGAL_OPCODE 1   /* reference panics here (src/cpu.rs:1139-1141) */
#define FA_ENV_VSYNC_HELD 2       /* VSYNC held > 1,000,000 cycles (reference loops forever) */

/*
 * AtariStateGpu -- the reference's packed per-env state, byte for byte
 * (src/cuda_env.rs:12-38, cuda/src/state_layout.cuh:16-88): 222 bytes, align 2.
 */
typedef struct AtariStateGpu {
    uint8_t a, x, y, sp;
    uint16_t pc;
    uint8_t status, _pad0;
    uint8_t ram[128];
    uint8_t port_a_output, port_a_ddr, port_a_input;
    uint8_t port_b_output, port_b_ddr, port_b_input;
    uint16_t timer_value, timer_prescaler, timer_prescaler_select;
    uint8_t timer_underflow, _pad1, _pad2, _pad3;
    uint16_t tia_clock, tia_scanline;
    uint8_t tia_wsync, tia_frame_complete, tia_vsync, tia_vblank;
    uint8_t cxm0p, cxm1p, cxp0fb, cxp1fb, cxm0fb, cxm1fb, cxblpf, cxppmm;
    uint8_t inpt4, inpt5, paddle0, paddle1;
    uint16_t paddle_counter;
    uint8_t paddle_dumped, _pad4;
    uint16_t pos_p0, pos_p1, pos_m0, pos_m1, pos_bl;
    uint8_t hmp0, hmp1, hmm0, hmm1, hmbl, _pad5;
    uint8_t grp0, grp1, grp0_old, grp1_old;
    uint8_t enam0, enam1, enabl, enabl_old;
    uint8_t vdelp0, vdelp1, vdelbl;
    uint8_t resmp0, resmp1, nusiz0, nusiz1;
    uint8_t colup0, colup1, colupf, colubk;
    uint8_t ctrlpf, refp0, refp1;
    uint8_t pf0, pf1, pf2;
    uint8_t bank, scheme;
    uint8_t _pad_end[2];
} AtariStateGpu;

#define FA_STATE_BYTES 222
#define FA_OBS_BYTES 128

typedef struct fa_env fa_env;

/* BatchAtariGpu::new(rom, n) -- src/cuda_env.rs:112-126.
 * rom_len: 1..4096 (Fixed, mirrored), 8192 (F8), 16384 (F6), >= 32768 (F4;
 * the first 32 KB are used) -- the sizes BankScheme::detect + load_bank accept
 * without panicking (src/atari.rs:97-104, 433-449). */
int fa_create(const uint8_t *rom, size_t rom_len, size_t n, fa_env **out);

/* Drop: releases device memory (src/cuda_env.rs:99-109 owns the same buffers). */
void fa_destroy(fa_env *env);

/* BatchAtariGpu::set_sorted -- src/cuda_env.rs:129.  Non-zero enables the
 * wave64 opcode-grouping interpreter. */
int fa_set_sorted(fa_env *env, int sorted);

/* BatchAtariGpu::reset -- src/cuda_env.rs:131-142: every env is powered on
 * (HeadlessAtari::new) and runs one frame without an action.  obs_out
 * (n * 128 bytes, may be NULL) receives each env's RAM. */
int fa_reset(fa_env *env, uint8_t *obs_out);

/* BatchAtariGpu::step -- src/cuda_env.rs:144-161.  actions: n bytes (action
 * index 0..17, others = NOOP); obs_out: n * 128 bytes.  n_actions != n is
 * FA_EINVAL (the reference asserts). */
int fa_step(fa_env *env, const uint8_t *actions, size_t n_actions, uint8_t *obs_out);

/* BatchAtariGpu::step_multi -- src/cuda_env.rs:166-184.  actions: k * n
 * bytes laid out [f * n + i]; obs_out: RAM after the last frame, n * 128. */
int fa_step_multi(fa_env *env, const uint8_t *actions, size_t n_actions, uint32_t k, uint8_t *obs_out);

/* BatchAtariGpu::download_states -- src/cuda_env.rs:186-188. */
int fa_download_states(fa_env *env, AtariStateGpu *out, size_t n);

/* BatchAtariGpu::num_envs -- src/cuda_env.rs:190. */
size_t fa_num_envs(const fa_env *env);

/* Message for the calling thread's last non-OK return. */
const char *fa_last_error(void);

/* ---- extensions (no reference counterpart; they ride alongside) ---- */

/* Replace every env's state (checkpoint restore, seeding).  States must be
 * reachable ones: clock % 3 == 0 and < 228, select in {1, 8, 64, 1024},
 * prescaler <= select, bank valid for the ROM, scheme equal to the ROM's. */
int fa_upload_states(fa_env *env, const AtariStateGpu *in, size_t n);

/* Per-env counters: cumulative Cpu.cycles (src/cpu.rs:104), cycles of the
 * last frame (HeadlessAtari::run_frame's return value, src/atari.rs:584-598),
 * cumulative executed instructions, and FA_ENV_* status.  Any pointer may be
 * NULL. */
int fa_download_counters(fa_env *env, uint64_t *cpu_cycles, uint64_t *frame_cycles,
                         uint64_t *instructions, uint32_t *status);
int fa_upload_cpu_cycles(fa_env *env, const uint64_t *cpu_cycles, size_t n);

/* Hold each env's action for fs[i] frames (set_action + fs[i] x run_frame);
 * obs after each env's last frame. */
int fa_step_frameskip(fa_env *env, const uint8_t *actions, const uint8_t *fs, size_t n, uint8_t *obs_out);

/* Device-resident variants: d_actions (k * n bytes, or n with d_fs) and
 * d_obs (n * 128, may be NULL) are device pointers; nothing is copied and the
 * call returns after enqueueing on the handle's stream. */
int fa_step_device(fa_env *env, const uint8_t *d_actions, uint32_t k, uint8_t *d_obs);
int fa_step_frameskip_device(fa_env *env, const uint8_t *d_actions, const uint8_t *d_fs, uint8_t *d_obs);
int fa_sync(fa_env *env);

/* Device memory helpers for callers without a GPU framework. */
int fa_device_alloc(size_t bytes, void **d_ptr);
int fa_device_free(void *d_ptr);
int fa_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
int fa_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);

/* Kernel timing: when enabled, every frame-step launch is bracketed by HIP
 * events on the handle's stream; fa_kernel_time returns the summed kernel
 * milliseconds and the launch count since the last reset. */
int fa_set_timing(fa_env *env, int enabled);
int fa_kernel_time(fa_env *env, double *total_ms, uint64_t *launches, int reset);

/* Launch shape: workgroup size (multiple of 64, <= 1024; 0 = automatic). */
int fa_set_block_size(fa_env *env, uint32_t block);
int fa_get_block_size(const fa_env *env, uint32_t *block);
/* LDS bytes available to one workgroup on the handle's device. */
int fa_get_max_lds(const fa_env *env, size_t *bytes);
/* The per-opcode descriptor table the kernels stage into LDS: 256 opcodes x 8
 * words (the decode d0 / d1, the same decode as one-hot masks d2 / d3 and the
 * per-lane register selectors d4 .. d7, see csrc/fa_decode.hpp); n_words must
 * be 2048.  Host only (no GPU needed). */
int fa_descriptor_table(uint32_t *out, size_t n_words);

/* ---- heterogeneous batches (several ROM classes stepped together) ----
 * No reference counterpart: the reference binds one ROM per BatchAtariGpu
 * (src/cuda_env.rs:112-126).  Envs of class c occupy the global indices
 * [first_c, first_c + count_c) in class order; per-env inputs and outputs use
 * those indices.  A step launch is one hipGraph (captured per k / flags) of
 * per-class kernel chains running concurrently. */
typedef struct fa_batch fa_batch;

int fa_batch_create(const uint8_t *const *roms, const size_t *rom_lens, const size_t *counts, size_t nclass,
                    fa_batch **out);
void fa_batch_destroy(fa_batch *b);
size_t fa_batch_num_envs(const fa_batch *b);
size_t fa_batch_num_classes(const fa_batch *b);
int fa_batch_class_range(const fa_batch *b, size_t c, size_t *first, size_t *count);
/* BatchAtariGpu::reset semantics for every env of every class. */
int fa_batch_reset(fa_batch *b, uint8_t *obs_out);
/* k steps: actions [s * N + i]; fs [s * N + i] frames each action is held
 * (NULL = 1); obs_out: N * 128 after the last step, or k * N * 128 after
 * every step when obs_every.  Synchronous. */
int fa_batch_step(fa_batch *b, const uint8_t *actions, const uint8_t *fs, uint32_t k, int obs_every,
                  uint8_t *obs_out);
/* Device-resident stepping: the batch's own input / output buffers (sized
 * for k steps) -- write actions / fs there, then fa_batch_launch replays the
 * captured graph on the batch stream and fa_batch_sync waits for it and
 * reports the launch's device time. */
int fa_batch_buffers(fa_batch *b, uint32_t k, int obs_every, uint8_t **d_actions, uint8_t **d_fs, uint8_t **d_obs);
int fa_batch_launch(fa_batch *b, uint32_t k, int use_fs, int obs_every);
int fa_batch_sync(fa_batch *b, double *launch_ms);
int fa_batch_download_states(fa_batch *b, AtariStateGpu *out, size_t n);
int fa_batch_upload_states(fa_batch *b, const AtariStateGpu *in, size_t n);
int fa_batch_download_counters(fa_batch *b, uint64_t *cpu_cycles, uint64_t *frame_cycles, uint64_t *instructions,
                               uint32_t *status);
/* Launch shape per class: block (0 = largest that fits), opcode grouping,
 * steps per kernel launch (0 = all k in one launch). */
int fa_batch_set_class_config(fa_batch *b, size_t c, uint32_t block, int sorted, uint32_t k_split);
int fa_batch_get_class_config(const fa_batch *b, size_t c, uint32_t *block, int *sorted, uint32_t *k_split);
/* Envs per workgroup below the block: 8, 16 or 32 spread the class over more,
 * less full waves (a divergent wave's frame is shorter with fewer lanes); 0 =
 * the block.  No counterpart in the reference, whose kernels run one fixed
 * launch shape (cuda/src/atari_kernel.cu:104-113). */
int fa_batch_set_class_lanes(fa_batch *b, size_t c, uint32_t lanes);
int fa_batch_get_class_lanes(const fa_batch *b, size_t c, uint32_t *lanes);
/* Autotuner: times trial launches of each class alone on the inputs in the
 * batch buffers (states restored after every trial) and adopts the fastest
 * block / envs per workgroup / grouping / steps-per-launch.  trials (2 * nclass, may be NULL):
 * per class, ms of the chosen and of the default configuration. */
int fa_batch_tune(fa_batch *b, uint32_t k, int use_fs, double *trials);

/* ---- env-step server (continuous batching) ----
 * No reference counterpart.  Client envs stay resident in HBM; clients are
 * admitted against budget_bytes of device memory (FA_EBUSY when they do not
 * fit); step requests that arrive while others run are coalesced into the
 * next launch of their ROM class.  A client has at most one request in
 * flight.  Functions may be called from any thread. */
typedef struct fa_server fa_server;

int fa_server_create(const uint8_t *const *roms, const size_t *rom_lens, size_t nclass, size_t budget_bytes,
                     fa_server **out);
void fa_server_destroy(fa_server *s);
/* New client of n_envs (1..65536) envs on ROM class cls, each started from
 * the ROM's reset state (power-on + one frame); obs_out (n * 128, may be
 * NULL) receives their RAM. */
int fa_server_open(fa_server *s, size_t cls, size_t n_envs, uint8_t *obs_out, int *client_id);
int fa_server_close(fa_server *s, int client_id);
/* One step of every env of the client: actions[i] held for `frames` frames;
 * restart (n bytes, may be NULL) != 0 restarts env i from the reset state
 * first.  Returns immediately. */
int fa_server_submit(fa_server *s, int client_id, const uint8_t *actions, const uint8_t *restart, uint32_t frames);
/* 1 when the client's request is done, 0 otherwise (negative on error). */
int fa_server_poll(fa_server *s, int client_id);
/* Waits for the client's request; obs_out (n * 128) and the request's
 * latency from submit to obs in host memory. */
int fa_server_wait(fa_server *s, int client_id, uint8_t *obs_out, double *latency_ms);
/* States / counters of an idle client's envs. */
int fa_server_download(fa_server *s, int client_id, AtariStateGpu *states, uint64_t *cpu_cycles,
                       uint64_t *frame_cycles, uint32_t *status);
int fa_server_client_info(fa_server *s, int client_id, size_t *cls, size_t *n_envs);
/* Launch shape of one ROM class: workgroup size (64, 128 or 256 envs; 0 picks
 * the default, the largest whose LDS footprint leaves room on a CU for
 * another class's workgroup) and opcode grouping. */
int fa_server_set_class_config(fa_server *s, size_t cls, uint32_t block, int sorted);
/* Coalescing window (default 300 us): after a launch of a class completes,
 * its next launch waits up to this long, or a tenth of the completed launch's
 * time if that is longer, for the class's clients to resubmit (it starts at
 * once when all of them have). */
int fa_server_set_coalesce_us(fa_server *s, uint32_t us);
int fa_server_memory(fa_server *s, size_t *used_bytes, size_t *budget_bytes, size_t *bytes_per_slot);
/* Per completed request since the last reset: latency ms, class, envs,
 * frames; *count = samples held; *launches = kernel launches. */
int fa_server_samples(fa_server *s, double *ms, uint32_t *cls, uint32_t *envs, uint32_t *frames, size_t cap,
                      size_t *count, uint64_t *launches, int reset);

/* Load driver standing in for the clients (benchmarks, tests).  specs[i]
 * describes client_ids[i]: its envs' actions and restarts come from per-env
 * SplitMix64 streams of specs[i].seed, step steps[i] onwards (steps is
 * updated).  sat == NULL: closed loop; else Poisson arrivals at
 * rate_mult * sat[i] per second.  Per completed request: service latency
 * (submit -> obs in host memory), end-to-end latency (arrival -> obs) and
 * class, the first `cap` of them; *nsamples = completed requests. */
typedef struct fa_client_spec {
    uint32_t cls, n;
    double rate_mult;
    uint64_t seed;
} fa_client_spec;
int fa_loadgen_run(fa_server *s, const int *client_ids, const fa_client_spec *specs, size_t nclients,
                   double duration_s, const double *sat, uint32_t frames, double restart_p, uint64_t arrival_seed,
                   uint64_t max_steps, uint64_t *steps, double *service_ms, double *e2e_ms, uint32_t *sample_cls,
                   size_t cap, size_t *nsamples, double *wall_s);

/* ---- env adapter: BatchAtariGpuEnv (src/rl4burn_env.rs:268-380), the caller
 * of BatchAtariGpu.  One action broadcast to every env and held for
 * `frameskip` frames per step, observations = RAM / 255 as f32 (n * 128,
 * normalised on the device), reward 0 and never terminated (the reference's
 * constants), truncated once the step count reaches max_steps. */
typedef struct fa_vecenv fa_vecenv;
#define FA_NUM_ACTIONS 18
#define FA_DEFAULT_MAX_STEPS 108000
#define FA_DEFAULT_FRAMESKIP 4

/* BatchAtariGpuEnv::new (src/rl4burn_env.rs:282-295): max_steps 108000,
 * frameskip 4, unsorted. */
int fa_vecenv_create(const uint8_t *rom, size_t rom_len, size_t n, fa_vecenv **out);
void fa_vecenv_destroy(fa_vecenv *v);
/* with_max_steps (0: never truncated), with_frameskip (>= 1), with_sorted
 * (src/rl4burn_env.rs:297-313). */
int fa_vecenv_set_max_steps(fa_vecenv *v, uint64_t max_steps);
int fa_vecenv_set_frameskip(fa_vecenv *v, uint32_t frameskip);
int fa_vecenv_set_sorted(fa_vecenv *v, int sorted);
size_t fa_vecenv_num_envs(const fa_vecenv *v);
/* The step count since the last reset. */
uint64_t fa_vecenv_step_count(const fa_vecenv *v);
/* The underlying batch (download_states, counters, ...). */
fa_env *fa_vecenv_batch(fa_vecenv *v);
/* Env::reset (src/rl4burn_env.rs:318-325): step count 0, every env to the
 * ROM's reset state (BatchAtariGpu::reset); obs_out: n * 128 f32 (host). */
int fa_vecenv_reset(fa_vecenv *v, float *obs_out);
/* Env::step (src/rl4burn_env.rs:327-358): `action` (as u8; indices above 17
 * are no-ops) for every env, held for frameskip frames in one launch; obs of
 * the last frame (n * 128 f32, host; NULL: not copied).  reward / terminated /
 * truncated may be NULL. */
int fa_vecenv_step(fa_vecenv *v, uint32_t action, float *obs_out, float *reward, int *terminated, int *truncated);
/* The same step with a per-env action (actions: n bytes, host). */
int fa_vecenv_step_each(fa_vecenv *v, const uint8_t *actions, float *obs_out, float *reward, int *terminated,
                        int *truncated);
/* Device form: actions from device memory (d_actions: n bytes; NULL: `action`
 * for every env), f32 obs into device memory d_obs (n * 128 floats); queued on
 * the batch's stream, complete after fa_vecenv_sync.  Counts the step and
 * reports truncation at once. */
int fa_vecenv_step_device(fa_vecenv *v, const uint8_t *d_actions, uint32_t action, float *d_obs, int *truncated);
int fa_vecenv_sync(fa_vecenv *v);
/* u8 observations to f32 in [0, 1] (byte / 255, IEEE round-to-nearest), device
 * memory to device memory, n_bytes values; synchronous. */
int fa_obs_to_f32(const uint8_t *d_in, float *d_out, size_t n_bytes);

/* Reward / termination readers over RAM.  The reference leaves both to
 * game-specific wrappers and returns 0 / false (src/rl4burn_env.rs:347-352,
 * src/atari.rs:347-352).  Score: the RAM bytes score_addr[0 .. n_score)
 * (offsets 0..127 into the 128 RAM bytes, most significant first), packed BCD
 * (two decimal digits a byte) when score_bcd, else binary; an env's reward
 * for a step is its score after the step minus its score before, as f32.
 * Lives: (ram[lives_addr] >> lives_shift) & lives_mask; an env is terminated
 * while that is 0 (lives_mask 0: never).  Evaluated on the GPU after reset and
 * after every step; with a reader set, fa_vecenv_step's reward is the sum over
 * envs and terminated is 1 when every env is terminated. */
typedef struct fa_ram_reader {
    uint8_t score_addr[4];
    uint8_t n_score;                   /* 0..4; 0: every reward 0 */
    uint8_t score_bcd;
    uint8_t lives_addr, lives_shift, lives_mask;
    uint8_t _pad[3];
} fa_ram_reader;
/* reader NULL: none (reward 0, never terminated).  Scores start from the
 * current RAM (reset sets them again). */
int fa_vecenv_set_reader(fa_vecenv *v, const fa_ram_reader *reader);
/* Per-env reward and terminated flag of the last reset / step (host, n each). */
int fa_vecenv_rewards(fa_vecenv *v, float *rewards, uint8_t *terminated);

/* ---- measurement ----
 * No reference counterpart.  The device's VALU issue peak, measured: `reps`
 * launches (after a warm-up) of a kernel of independent v_add_u32 chains with
 * 8 waves per SIMD and `iters` loop iterations per wave; returns the mean
 * launch time and the wave64 adds per launch. */
int fa_probe_valu_peak(uint32_t iters, uint32_t reps, double *ms_per_launch, uint64_t *adds_per_launch);

/* ---- stepping without allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fa_step_device's k-frame step on n envs whose states
 * (AtariStateGpu[n]) live in the caller's device memory.  d_actions: k * n
 * bytes laid out [f * n + i] (NULL: no set_action); d_obs: n * 128 bytes, RAM
 * after the last frame (may be NULL); d_status: n FA_ENV_* words, read and
 * written back (NULL: every env starts at FA_ENV_OK; an env with a non-zero
 * status does not run).  ROM as fa_create's.  The descriptors, the ROM image,
 * the staged state block and the counters live in a fixed device array inside
 * the library (FA_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends):
 * n above FA_NOALLOC_MAX_ENVS is FA_EINVAL, and the call makes no device
 * allocation (no hipMalloc / hipFree or allocating library call).  Enqueued on `stream` (a hipStream_t; NULL = the null
 * stream), returns without waiting.  Calls that overlap in time on one GPU
 * share the array and race; the first call on a device sets it up. */
#define FA_NOALLOC_MAX_ENVS 65536
#define FA_NOALLOC_SCRATCH_BYTES 16425216
int fa_step_noalloc(const uint8_t *rom, size_t rom_len, AtariStateGpu *d_states, uint32_t *d_status,
                    const uint8_t *d_actions, size_t n, uint32_t k, uint8_t *d_obs, void *stream);

/* ---- rendering console: pixel frames ----
 * A batch of the reference's rendering `Atari` (src/atari.rs:226-378): the
 * per-clock TIA (src/tia.rs:98-675) drawing 160 x 192 frames of NTSC palette
 * indices, on the rendering bus whose every CPU access clocks the TIA three
 * times and the RIOT once (src/atari.rs:107-224).  A frame is one run_frame
 * (src/atari.rs:277-305); each env's frame is Atari::framebuffer
 * (src/atari.rs:320-323) after its last frame.  The render state is
 * FA_RENDER_STATE_BYTES per env: the AtariStateGpu layout (tia_clock counts
 * TIA clocks 0..227 here), then visible_top (u16 @224), vblank_top (u16 @226),
 * pending_top (u16 @228), pending_frames @230, hmove_pending @231,
 * inpt4_latch @232, inpt5_latch @233, input_latch_enabled @234, zeros to 256. */
typedef struct fa_render fa_render;
#define FA_FRAME_WIDTH 160
#define FA_FRAME_HEIGHT 192
#define FA_FRAME_BYTES (FA_FRAME_WIDTH * FA_FRAME_HEIGHT)
#define FA_RENDER_STATE_BYTES 256
#define FA_RENDER_MAX_ENVS (1u << 20)

/* n consoles of one ROM (sizes as fa_create's), each at Atari::new
 * (src/atari.rs:238-246) with an all-zero frame. */
int fa_render_create(const uint8_t *rom, size_t rom_len, size_t n, fa_render **out);
void fa_render_destroy(fa_render *r);
size_t fa_render_num_envs(const fa_render *r);
/* Env::reset (src/atari.rs:330-340) on every env: Atari::new, then one
 * run_frame with no action set.  frames_out: n * FA_FRAME_BYTES (host; NULL:
 * not copied). */
int fa_render_reset(fa_render *r, uint8_t *frames_out);
/* Env::step (src/atari.rs:342-357) k times: frame f sets actions[f * n + i]
 * (indices above 17 are no-ops, as Action::from_index's default) then runs a
 * frame.  actions: k * n bytes (host); frames_out as fa_render_reset's. */
int fa_render_step(fa_render *r, const uint8_t *actions, size_t n_actions, uint32_t k, uint8_t *frames_out);
/* Device form: actions in device memory (k * n bytes, NULL: none set); queued
 * on the handle's stream, complete after fa_render_sync. */
int fa_render_step_device(fa_render *r, const uint8_t *d_actions, uint32_t k);
int fa_render_sync(fa_render *r);
/* The handle's device frame array (n * FA_FRAME_BYTES, env i at i * FA_FRAME_BYTES). */
int fa_render_frames_device(fa_render *r, uint8_t **d_frames);
int fa_render_download_frames(fa_render *r, uint8_t *out, size_t n);
int fa_render_upload_frames(fa_render *r, const uint8_t *in, size_t n);
/* Render states (n * FA_RENDER_STATE_BYTES).  An upload is checked for
 * reachable values (FA_ESTATE) and clears every env's status. */
int fa_render_download_states(fa_render *r, uint8_t *out, size_t n);
int fa_render_upload_states(fa_render *r, const uint8_t *in, size_t n);
/* Cpu.cycles, the last run_frame's cycles, instructions, FA_ENV_* (any may be NULL). */
int fa_render_download_counters(fa_render *r, uint64_t *cpu_cycles, uint64_t *frame_cycles, uint64_t *instructions,
                                uint32_t *status);
/* Atari::new's state for a ROM (FA_RENDER_STATE_BYTES into out). */
int fa_render_power_on_state(const uint8_t *rom, size_t rom_len, uint8_t *out);
/* Workgroup size (0: default 64; else 64, 128, 192 or 256 within the LDS budget). */
int fa_render_set_block_size(fa_render *r, uint32_t block);
int fa_render_get_block_size(const fa_render *r, uint32_t *block);
/* Per-launch HIP event timing on the handle's stream, as fa_set_timing / fa_kernel_time. */
int fa_render_set_timing(fa_render *r, int enabled);
int fa_render_kernel_time(fa_render *r, double *total_ms, uint64_t *launches, int reset);

#ifdef __cplusplus
}
#endif
#endif
