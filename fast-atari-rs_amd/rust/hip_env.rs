//! HIP-accelerated batch Atari environment -- drop-in for `src/cuda_env.rs`.
//!
//! Same public surface as the reference's `BatchAtariGpu`
//! (src/cuda_env.rs:99-191): `new`, `set_sorted`, `reset`, `step`,
//! `step_multi`, `download_states`, `num_envs`, plus `AtariStateGpu` with the
//! identical 222-byte `#[repr(C)]` layout.  All work is done by
//! libfastatari_hip.so through the C ABI declared in include/fastatari.h.
//!
//! Wiring (Cargo.toml): `hip = []` feature; `src/lib.rs`:
//! `#[cfg(feature = "hip")] pub mod hip_env;`.  `build.rs` builds and links
//! the library (see build.rs next to this file).  No Rust toolchain exists in
//! the build container, so this file is shipped as source.

use std::ffi::CStr;
use std::os::raw::{c_char, c_int};

/// Packed per-env state; byte-identical to `AtariStateGpu` of
/// src/cuda_env.rs:12-38 and to the C struct in include/fastatari.h, with the
/// reference's field names, so callers that name fields compile unchanged.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct AtariStateGpu {
    pub a: u8, pub x: u8, pub y: u8, pub sp: u8,
    pub pc: u16, pub status: u8, pub _pad0: u8,
    pub ram: [u8; 128],
    pub port_a_output: u8, pub port_a_ddr: u8, pub port_a_input: u8,
    pub port_b_output: u8, pub port_b_ddr: u8, pub port_b_input: u8,
    pub timer_value: u16, pub timer_prescaler: u16, pub timer_prescaler_select: u16,
    pub timer_underflow: u8, pub _pad1: u8, pub _pad2: u8, pub _pad3: u8,
    pub tia_clock: u16, pub tia_scanline: u16,
    pub tia_wsync: u8, pub tia_frame_complete: u8, pub tia_vsync: u8, pub tia_vblank: u8,
    pub cxm0p: u8, pub cxm1p: u8, pub cxp0fb: u8, pub cxp1fb: u8,
    pub cxm0fb: u8, pub cxm1fb: u8, pub cxblpf: u8, pub cxppmm: u8,
    pub inpt4: u8, pub inpt5: u8,
    pub paddle0: u8, pub paddle1: u8, pub paddle_counter: u16,
    pub paddle_dumped: u8, pub _pad4: u8,
    pub pos_p0: u16, pub pos_p1: u16, pub pos_m0: u16, pub pos_m1: u16, pub pos_bl: u16,
    pub hmp0: u8, pub hmp1: u8, pub hmm0: u8, pub hmm1: u8, pub hmbl: u8, pub _pad5: u8,
    pub grp0: u8, pub grp1: u8, pub grp0_old: u8, pub grp1_old: u8,
    pub enam0: u8, pub enam1: u8, pub enabl: u8, pub enabl_old: u8,
    pub vdelp0: u8, pub vdelp1: u8, pub vdelbl: u8,
    pub resmp0: u8, pub resmp1: u8, pub nusiz0: u8, pub nusiz1: u8,
    pub colup0: u8, pub colup1: u8, pub colupf: u8, pub colubk: u8,
    pub ctrlpf: u8, pub refp0: u8, pub refp1: u8,
    pub pf0: u8, pub pf1: u8, pub pf2: u8,
    pub bank: u8, pub scheme: u8,
    pub _pad_end: [u8; 2],
}

const _: () = assert!(std::mem::size_of::<AtariStateGpu>() == 222);

use crate::atari::{BankScheme, HeadlessAtari};

impl AtariStateGpu {
    /// Packs a CPU emulator into the GPU layout (src/cuda_env.rs:43-94): the
    /// RIOT RAM goes to `ram`, booleans become 0 / 1, the mapper becomes
    /// `scheme` 0..3 (Fixed, F8, F6, F4) and every pad byte is 0.  Callers
    /// that build states on the host upload them with
    /// `BatchAtariGpu::upload_states`.
    pub fn from_headless(atari: &HeadlessAtari) -> Self {
        let (cpu, bus) = (&atari.cpu, &atari.bus);
        let (t, p) = (&bus.tia, &bus.pia);
        let scheme = match bus.scheme {
            BankScheme::Fixed => 0,
            BankScheme::F8 => 1,
            BankScheme::F6 => 2,
            BankScheme::F4 => 3,
        };
        AtariStateGpu {
            a: cpu.a, x: cpu.x, y: cpu.y, sp: cpu.sp, pc: cpu.pc, status: cpu.status.0, _pad0: 0,
            ram: p.ram,
            port_a_output: p.port_a_output, port_a_ddr: p.port_a_ddr, port_a_input: p.port_a_input,
            port_b_output: p.port_b_output, port_b_ddr: p.port_b_ddr, port_b_input: p.port_b_input,
            timer_value: p.timer_value,
            timer_prescaler: p.timer_prescaler,
            timer_prescaler_select: p.timer_prescaler_select,
            timer_underflow: u8::from(p.timer_underflow), _pad1: 0, _pad2: 0, _pad3: 0,
            tia_clock: t.clock, tia_scanline: t.scanline,
            tia_wsync: u8::from(t.wsync), tia_frame_complete: u8::from(t.frame_complete),
            tia_vsync: t.vsync, tia_vblank: t.vblank,
            cxm0p: t.cxm0p, cxm1p: t.cxm1p, cxp0fb: t.cxp0fb, cxp1fb: t.cxp1fb,
            cxm0fb: t.cxm0fb, cxm1fb: t.cxm1fb, cxblpf: t.cxblpf, cxppmm: t.cxppmm,
            inpt4: u8::from(t.inpt4), inpt5: u8::from(t.inpt5),
            paddle0: t.paddle0, paddle1: t.paddle1, paddle_counter: t.paddle_counter,
            paddle_dumped: u8::from(t.paddle_dumped), _pad4: 0,
            pos_p0: t.pos_p0, pos_p1: t.pos_p1, pos_m0: t.pos_m0, pos_m1: t.pos_m1, pos_bl: t.pos_bl,
            hmp0: t.hmp0, hmp1: t.hmp1, hmm0: t.hmm0, hmm1: t.hmm1, hmbl: t.hmbl, _pad5: 0,
            grp0: t.grp0, grp1: t.grp1, grp0_old: t.grp0_old, grp1_old: t.grp1_old,
            enam0: u8::from(t.enam0), enam1: u8::from(t.enam1),
            enabl: u8::from(t.enabl), enabl_old: u8::from(t.enabl_old),
            vdelp0: u8::from(t.vdelp0), vdelp1: u8::from(t.vdelp1), vdelbl: u8::from(t.vdelbl),
            resmp0: u8::from(t.resmp0), resmp1: u8::from(t.resmp1),
            nusiz0: t.nusiz0, nusiz1: t.nusiz1,
            colup0: t.colup0, colup1: t.colup1, colupf: t.colupf, colubk: t.colubk,
            ctrlpf: t.ctrlpf, refp0: u8::from(t.refp0), refp1: u8::from(t.refp1),
            pf0: t.pf0, pf1: t.pf1, pf2: t.pf2,
            bank: bus.bank as u8,
            scheme,
            _pad_end: [0; 2],
        }
    }
}

#[repr(C)]
pub struct FaEnv {
    _private: [u8; 0],
}

extern "C" {
    fn fa_create(rom: *const u8, rom_len: usize, n: usize, out: *mut *mut FaEnv) -> c_int;
    fn fa_destroy(env: *mut FaEnv);
    fn fa_set_sorted(env: *mut FaEnv, sorted: c_int) -> c_int;
    fn fa_reset(env: *mut FaEnv, obs_out: *mut u8) -> c_int;
    fn fa_step(env: *mut FaEnv, actions: *const u8, n_actions: usize, obs_out: *mut u8) -> c_int;
    fn fa_step_multi(env: *mut FaEnv, actions: *const u8, n_actions: usize, k: u32, obs_out: *mut u8) -> c_int;
    fn fa_download_states(env: *mut FaEnv, out: *mut AtariStateGpu, n: usize) -> c_int;
    fn fa_num_envs(env: *const FaEnv) -> usize;
    fn fa_last_error() -> *const c_char;
    fn fa_upload_states(env: *mut FaEnv, states: *const AtariStateGpu, n: usize) -> c_int;
    fn fa_download_counters(env: *mut FaEnv, cpu_cycles: *mut u64, frame_cycles: *mut u64,
                            instructions: *mut u64, status: *mut u32) -> c_int;
    fn fa_step_frameskip(env: *mut FaEnv, actions: *const u8, fs: *const u8, n: usize, obs_out: *mut u8) -> c_int;
    fn fa_batch_create(roms: *const *const u8, rom_lens: *const usize, counts: *const usize, nclass: usize,
                       out: *mut *mut FaBatch) -> c_int;
    fn fa_batch_destroy(b: *mut FaBatch);
    fn fa_batch_num_envs(b: *const FaBatch) -> usize;
    fn fa_batch_reset(b: *mut FaBatch, obs_out: *mut u8) -> c_int;
    fn fa_batch_step(b: *mut FaBatch, actions: *const u8, fs: *const u8, k: u32, obs_every: c_int,
                     obs_out: *mut u8) -> c_int;
    fn fa_batch_tune(b: *mut FaBatch, k: u32, use_fs: c_int, trials: *mut f64) -> c_int;
    fn fa_batch_buffers(b: *mut FaBatch, k: u32, obs_every: c_int, d_actions: *mut *mut u8, d_fs: *mut *mut u8,
                        d_obs: *mut *mut u8) -> c_int;
    fn fa_memcpy_h2d(d_dst: *mut u8, h_src: *const u8, bytes: usize) -> c_int;
}

#[repr(C)]
pub struct FaBatch {
    _private: [u8; 0],
}

/// Error of a failed library call (the reference returns cudarc's `DriverError`).
#[derive(Debug, Clone)]
pub struct HipError {
    pub code: i32,
    pub message: String,
}

fn check(rc: c_int) -> Result<(), HipError> {
    if rc == 0 {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(fa_last_error()) }.to_string_lossy().into_owned();
    Err(HipError { code: rc, message })
}

pub struct BatchAtariGpu {
    env: *mut FaEnv,
    n: usize,
}

unsafe impl Send for BatchAtariGpu {}

impl BatchAtariGpu {
    pub fn new(rom: Vec<u8>, n: usize) -> Result<Self, HipError> {
        let mut env = std::ptr::null_mut();
        check(unsafe { fa_create(rom.as_ptr(), rom.len(), n, &mut env) })?;
        Ok(Self { env, n })
    }

    pub fn set_sorted(&mut self, sorted: bool) {
        unsafe { fa_set_sorted(self.env, sorted as c_int) };
    }

    pub fn reset(&mut self) -> Result<Vec<Vec<u8>>, HipError> {
        let mut obs = vec![0u8; self.n * 128];
        check(unsafe { fa_reset(self.env, obs.as_mut_ptr()) })?;
        Ok(obs.chunks(128).map(|c| c.to_vec()).collect())
    }

    pub fn step(&mut self, actions: &[u8]) -> Result<Vec<u8>, HipError> {
        assert_eq!(actions.len(), self.n);
        let mut obs = vec![0u8; self.n * 128];
        check(unsafe { fa_step(self.env, actions.as_ptr(), actions.len(), obs.as_mut_ptr()) })?;
        Ok(obs)
    }

    pub fn step_multi(&mut self, actions: &[u8], k: usize) -> Result<Vec<u8>, HipError> {
        assert_eq!(actions.len(), k * self.n);
        let mut obs = vec![0u8; self.n * 128];
        check(unsafe { fa_step_multi(self.env, actions.as_ptr(), actions.len(), k as u32, obs.as_mut_ptr()) })?;
        Ok(obs)
    }

    pub fn download_states(&self) -> Result<Vec<AtariStateGpu>, HipError> {
        let mut out: Vec<AtariStateGpu> = Vec::with_capacity(self.n);
        check(unsafe { fa_download_states(self.env, out.as_mut_ptr(), self.n) })?;
        unsafe { out.set_len(self.n) };
        Ok(out)
    }

    pub fn num_envs(&self) -> usize {
        unsafe { fa_num_envs(self.env) }
    }
}

// Extensions (no counterpart in src/cuda_env.rs).
impl BatchAtariGpu {
    /// Replace every env's state (checkpoint restore).
    pub fn upload_states(&mut self, states: &[AtariStateGpu]) -> Result<(), HipError> {
        assert_eq!(states.len(), self.n);
        check(unsafe { fa_upload_states(self.env, states.as_ptr(), states.len()) })
    }

    /// (cpu cycles, last-frame cycles, instructions, status) per env.
    pub fn counters(&mut self) -> Result<(Vec<u64>, Vec<u64>, Vec<u64>, Vec<u32>), HipError> {
        let (mut c, mut f, mut i, mut s) = (vec![0u64; self.n], vec![0u64; self.n], vec![0u64; self.n], vec![0u32; self.n]);
        check(unsafe { fa_download_counters(self.env, c.as_mut_ptr(), f.as_mut_ptr(), i.as_mut_ptr(), s.as_mut_ptr()) })?;
        Ok((c, f, i, s))
    }

    /// Env i holds actions[i] for fs[i] frames; obs after each env's last frame.
    pub fn step_frameskip(&mut self, actions: &[u8], fs: &[u8]) -> Result<Vec<u8>, HipError> {
        assert_eq!(actions.len(), self.n);
        assert_eq!(fs.len(), self.n);
        let mut obs = vec![0u8; self.n * 128];
        check(unsafe { fa_step_frameskip(self.env, actions.as_ptr(), fs.as_ptr(), actions.len(), obs.as_mut_ptr()) })?;
        Ok(obs)
    }
}

impl Drop for BatchAtariGpu {
    fn drop(&mut self) {
        unsafe { fa_destroy(self.env) };
    }
}

/// Several ROM classes stepped together by one hipGraph launch.
pub struct BatchMixGpu {
    b: *mut FaBatch,
    n: usize,
}

unsafe impl Send for BatchMixGpu {}

impl BatchMixGpu {
    pub fn new(roms: &[Vec<u8>], counts: &[usize]) -> Result<Self, HipError> {
        assert_eq!(counts.len(), roms.len());
        let ptrs: Vec<*const u8> = roms.iter().map(|r| r.as_ptr()).collect();
        let lens: Vec<usize> = roms.iter().map(|r| r.len()).collect();
        let mut b = std::ptr::null_mut();
        check(unsafe { fa_batch_create(ptrs.as_ptr(), lens.as_ptr(), counts.as_ptr(), roms.len(), &mut b) })?;
        let n = unsafe { fa_batch_num_envs(b) };
        Ok(Self { b, n })
    }

    pub fn reset(&mut self) -> Result<Vec<u8>, HipError> {
        let mut obs = vec![0u8; self.n * 128];
        check(unsafe { fa_batch_reset(self.b, obs.as_mut_ptr()) })?;
        Ok(obs)
    }

    /// k steps: actions / fs laid out [s * N + i]; obs after the last step.
    pub fn step(&mut self, actions: &[u8], fs: Option<&[u8]>, k: u32) -> Result<Vec<u8>, HipError> {
        assert_eq!(actions.len(), k as usize * self.n);
        if let Some(f) = fs {
            assert_eq!(f.len(), k as usize * self.n);
        }
        let mut obs = vec![0u8; self.n * 128];
        let fsp = fs.map_or(std::ptr::null(), |f| f.as_ptr());
        check(unsafe { fa_batch_step(self.b, actions.as_ptr(), fsp, k, 0, obs.as_mut_ptr()) })?;
        Ok(obs)
    }

    /// Autotune block size / grouping / steps per launch per ROM class.
    pub fn tune(&mut self, actions: &[u8], fs: Option<&[u8]>, k: u32) -> Result<(), HipError> {
        assert_eq!(actions.len(), k as usize * self.n);
        if let Some(f) = fs {
            assert_eq!(f.len(), k as usize * self.n);
        }
        let (mut da, mut df, mut d_obs) = (std::ptr::null_mut(), std::ptr::null_mut(), std::ptr::null_mut());
        check(unsafe { fa_batch_buffers(self.b, k, 0, &mut da, &mut df, &mut d_obs) })?;
        check(unsafe { fa_memcpy_h2d(da, actions.as_ptr(), actions.len()) })?;
        if let Some(f) = fs {
            check(unsafe { fa_memcpy_h2d(df, f.as_ptr(), f.len()) })?;
        }
        check(unsafe { fa_batch_tune(self.b, k, fs.is_some() as c_int, std::ptr::null_mut()) })
    }
}

impl Drop for BatchMixGpu {
    fn drop(&mut self) {
        unsafe { fa_batch_destroy(self.b) };
    }
}

// ---- env adapter: `BatchAtariGpuEnv` of src/rl4burn_env.rs:268-380 over the
// fa_vecenv_* C ABI (f32 observations normalised on the GPU).  A maintainer
// puts this beside the `cuda` variant under `#[cfg(feature = "hip")]`.

use rl4burn_core::env::space::Space;
use rl4burn_core::env::{Env, Step};

#[repr(C)]
pub struct FaVecEnv {
    _private: [u8; 0],
}

extern "C" {
    fn fa_vecenv_create(rom: *const u8, rom_len: usize, n: usize, out: *mut *mut FaVecEnv) -> c_int;
    fn fa_vecenv_destroy(v: *mut FaVecEnv);
    fn fa_vecenv_set_max_steps(v: *mut FaVecEnv, max_steps: u64) -> c_int;
    fn fa_vecenv_set_frameskip(v: *mut FaVecEnv, frameskip: u32) -> c_int;
    fn fa_vecenv_set_sorted(v: *mut FaVecEnv, sorted: c_int) -> c_int;
    fn fa_vecenv_reset(v: *mut FaVecEnv, obs_out: *mut f32) -> c_int;
    fn fa_vecenv_step(v: *mut FaVecEnv, action: u32, obs_out: *mut f32, reward: *mut f32, terminated: *mut c_int,
                      truncated: *mut c_int) -> c_int;
    fn fa_vecenv_set_reader(v: *mut FaVecEnv, reader: *const FaRamReader) -> c_int;
    fn fa_vecenv_rewards(v: *mut FaVecEnv, rewards: *mut f32, terminated: *mut u8) -> c_int;
}

/// Reward / termination reader over RAM (fa_ram_reader): score bytes (RAM
/// offsets, most significant first; packed BCD or binary) and a lives field.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct FaRamReader {
    pub score_addr: [u8; 4],
    pub n_score: u8,
    pub score_bcd: u8,
    pub lives_addr: u8,
    pub lives_shift: u8,
    pub lives_mask: u8,
    pub _pad: [u8; 3],
}

/// N instances behind one `Env`: the action is broadcast to every instance
/// and held for `frameskip` frames; observations are the 128 RAM bytes of
/// each instance as f32 in [0, 1], flattened (`obs.len() == N * 128`).
pub struct BatchAtariGpuEnv {
    v: *mut FaVecEnv,
    n: usize,
}

unsafe impl Send for BatchAtariGpuEnv {}

impl BatchAtariGpuEnv {
    /// Panics if HIP initialisation fails (as the CUDA variant does).
    pub fn new(rom: Vec<u8>, n: usize) -> Self {
        let mut v = std::ptr::null_mut();
        check(unsafe { fa_vecenv_create(rom.as_ptr(), rom.len(), n, &mut v) })
            .expect("failed to initialise HIP BatchAtariGpuEnv");
        Self { v, n }
    }

    pub fn with_max_steps(self, max_steps: usize) -> Self {
        check(unsafe { fa_vecenv_set_max_steps(self.v, max_steps as u64) }).expect("with_max_steps");
        self
    }

    pub fn with_frameskip(self, frameskip: usize) -> Self {
        check(unsafe { fa_vecenv_set_frameskip(self.v, frameskip as u32) }).expect("frameskip must be 1..255");
        self
    }

    pub fn with_sorted(self, sorted: bool) -> Self {
        check(unsafe { fa_vecenv_set_sorted(self.v, sorted as c_int) }).expect("with_sorted");
        self
    }

    pub fn num_envs(&self) -> usize {
        self.n
    }

    /// Per-env rewards and termination read from RAM on the GPU (the
    /// reference leaves both to game-specific wrappers); `step` then returns
    /// their sum and whether every instance is terminated.
    pub fn with_reader(self, reader: FaRamReader) -> Self {
        check(unsafe { fa_vecenv_set_reader(self.v, &reader) }).expect("invalid RAM reader");
        self
    }

    /// Per-instance reward and terminated flag of the last reset / step.
    pub fn rewards(&mut self) -> (Vec<f32>, Vec<bool>) {
        let (mut r, mut t) = (vec![0f32; self.n], vec![0u8; self.n]);
        check(unsafe { fa_vecenv_rewards(self.v, r.as_mut_ptr(), t.as_mut_ptr()) }).expect("rewards");
        (r, t.into_iter().map(|x| x != 0).collect())
    }
}

impl Env for BatchAtariGpuEnv {
    type Observation = Vec<f32>;
    type Action = usize;

    fn reset(&mut self) -> Vec<f32> {
        let mut obs = vec![0f32; self.n * 128];
        check(unsafe { fa_vecenv_reset(self.v, obs.as_mut_ptr()) }).expect("HIP reset failed");
        obs
    }

    fn step(&mut self, action: usize) -> Step<Vec<f32>> {
        let mut obs = vec![0f32; self.n * 128];
        let (mut reward, mut terminated, mut truncated) = (0f32, 0 as c_int, 0 as c_int);
        check(unsafe {
            fa_vecenv_step(self.v, action as u8 as u32, obs.as_mut_ptr(), &mut reward, &mut terminated, &mut truncated)
        })
        .expect("HIP step failed");
        Step { observation: obs, reward, terminated: terminated != 0, truncated: truncated != 0 }
    }

    fn observation_space(&self) -> Space {
        let size = self.n * 128;
        Space::Box { low: vec![0.0; size], high: vec![1.0; size] }
    }

    fn action_space(&self) -> Space {
        Space::Discrete(18)
    }
}

impl Drop for BatchAtariGpuEnv {
    fn drop(&mut self) {
        unsafe { fa_vecenv_destroy(self.v) };
    }
}

// ---------------------------------------------------------------------------
// Rendering console: pixel frames (Atari, src/atari.rs:226-378)
// ---------------------------------------------------------------------------

pub const FRAME_WIDTH: usize = 160;
pub const FRAME_HEIGHT: usize = 192;
pub const RENDER_STATE_BYTES: usize = 256;

#[repr(C)]
pub struct FaRender {
    _private: [u8; 0],
}

extern "C" {
    fn fa_render_create(rom: *const u8, rom_len: usize, n: usize, out: *mut *mut FaRender) -> c_int;
    fn fa_render_destroy(r: *mut FaRender);
    fn fa_render_reset(r: *mut FaRender, frames_out: *mut u8) -> c_int;
    fn fa_render_step(r: *mut FaRender, actions: *const u8, n_actions: usize, k: u32, frames_out: *mut u8) -> c_int;
    fn fa_render_download_states(r: *mut FaRender, out: *mut u8, n: usize) -> c_int;
    fn fa_render_num_envs(r: *const FaRender) -> usize;
}

/// n rendering consoles of one ROM on the GPU: `Env::reset` / `Env::step`
/// (src/atari.rs:330-357) for every env, frames as 160 x 192 NTSC palette
/// indices per env (`Atari::framebuffer`, src/atari.rs:320-323).
pub struct BatchAtariRenderGpu {
    r: *mut FaRender,
    n: usize,
}

unsafe impl Send for BatchAtariRenderGpu {}

impl BatchAtariRenderGpu {
    pub fn new(rom: Vec<u8>, n: usize) -> Result<Self, HipError> {
        let mut r = std::ptr::null_mut();
        check(unsafe { fa_render_create(rom.as_ptr(), rom.len(), n, &mut r) })?;
        Ok(Self { r, n })
    }

    /// Atari::new + one run_frame per env; returns n frames.
    pub fn reset(&mut self) -> Result<Vec<u8>, HipError> {
        let mut frames = vec![0u8; self.n * FRAME_WIDTH * FRAME_HEIGHT];
        check(unsafe { fa_render_reset(self.r, frames.as_mut_ptr()) })?;
        Ok(frames)
    }

    /// k frames; actions laid out [f * n + i]; returns the frames after the last.
    pub fn step_multi(&mut self, actions: &[u8], k: usize) -> Result<Vec<u8>, HipError> {
        assert_eq!(actions.len(), k * self.n);
        let mut frames = vec![0u8; self.n * FRAME_WIDTH * FRAME_HEIGHT];
        check(unsafe { fa_render_step(self.r, actions.as_ptr(), actions.len(), k as u32, frames.as_mut_ptr()) })?;
        Ok(frames)
    }

    pub fn step(&mut self, actions: &[u8]) -> Result<Vec<u8>, HipError> {
        self.step_multi(actions, 1)
    }

    /// The 256-byte render states (AtariStateGpu layout, then the rendering TIA's fields).
    pub fn download_states(&self) -> Result<Vec<u8>, HipError> {
        let mut out = vec![0u8; self.n * RENDER_STATE_BYTES];
        check(unsafe { fa_render_download_states(self.r, out.as_mut_ptr(), self.n) })?;
        Ok(out)
    }

    pub fn num_envs(&self) -> usize {
        unsafe { fa_render_num_envs(self.r) }
    }
}

impl Drop for BatchAtariRenderGpu {
    fn drop(&mut self) {
        unsafe { fa_render_destroy(self.r) };
    }
}
