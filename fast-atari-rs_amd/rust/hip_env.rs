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
/// src/cuda_env.rs:12-38 and to the C struct in include/fastatari.h.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct AtariStateGpu {
    pub a: u8, pub x: u8, pub y: u8, pub sp: u8,
    pub pc: u16, pub status: u8, pub _pad0: u8,
    pub ram: [u8; 128],
    pub port_a_output: u8, pub port_a_ddr: u8, pub port_a_input: u8,
    pub port_b_output: u8, pub port_b_ddr: u8, pub port_b_input: u8,
    pub timer_value: u16, pub timer_prescaler: u16, pub timer_prescaler_select: u16,
    pub timer_underflow: u8, pub _pad1: [u8; 3],
    pub tia_clock: u16, pub tia_scanline: u16,
    pub tia_wsync: u8, pub tia_frame_complete: u8, pub tia_vsync: u8, pub tia_vblank: u8,
    pub collisions: [u8; 8],
    pub inpt4: u8, pub inpt5: u8, pub paddle0: u8, pub paddle1: u8,
    pub paddle_counter: u16, pub paddle_dumped: u8, pub _pad4: u8,
    pub pos: [u16; 5],
    pub hm: [u8; 5], pub _pad5: u8,
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

impl Drop for BatchAtariGpu {
    fn drop(&mut self) {
        unsafe { fa_destroy(self.env) };
    }
}
