"""BatchAtariGpuEnv: the env adapter over BatchAtariGpu, mirroring the
reference's `BatchAtariGpuEnv` (src/rl4burn_env.rs:268-380) method by method
over the fa_vecenv_* C ABI (include/fastatari.h).

`step(action)` broadcasts one action to every env and holds it for
`frameskip` frames; observations are the 128 RAM bytes of every env as f32
in [0, 1] (byte / 255, computed on the GPU), flattened to n * 128 values;
reward is 0 and `terminated` False (the reference's constants); `truncated`
once the step count reaches max_steps (0: never).  `step_each` takes one
action per env; `step_device` keeps actions and f32 observations in device
memory.  `with_reader` adds per-env rewards and termination read from RAM on
the GPU (the reference leaves both to game-specific wrappers).
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import numpy as np

from .env import BatchAtariGpu, _check, _p8, _u8, load_library

NUM_ACTIONS = 18                 # src/rl4burn_env.rs:26
DEFAULT_MAX_STEPS = 108_000      # src/rl4burn_env.rs:29
DEFAULT_FRAMESKIP = 4            # src/rl4burn_env.rs:32


class RamReader(C.Structure):
    """fa_ram_reader: score bytes (RAM offsets, most significant first; BCD or
    binary) and a lives field; see include/fastatari.h."""
    _fields_ = [("score_addr", C.c_uint8 * 4), ("n_score", C.c_uint8), ("score_bcd", C.c_uint8),
                ("lives_addr", C.c_uint8), ("lives_shift", C.c_uint8), ("lives_mask", C.c_uint8),
                ("_pad", C.c_uint8 * 3)]


class Step(NamedTuple):
    """rl4burn's Step (observation, reward, terminated, truncated)."""
    observation: np.ndarray
    reward: float
    terminated: bool
    truncated: bool


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class BatchAtariGpuEnv:
    def __init__(self, rom, n: int):
        L = load_library()
        self._rom = _u8(rom)
        self.n = int(n)
        h = C.c_void_p()
        _check(L.fa_vecenv_create(_p8(self._rom), len(self._rom), self.n, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().fa_vecenv_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- builder methods (src/rl4burn_env.rs:297-313)
    def with_max_steps(self, max_steps: int) -> "BatchAtariGpuEnv":
        _check(load_library().fa_vecenv_set_max_steps(self._h, int(max_steps)))
        return self

    def with_frameskip(self, frameskip: int) -> "BatchAtariGpuEnv":
        _check(load_library().fa_vecenv_set_frameskip(self._h, int(frameskip)))
        return self

    def with_sorted(self, sorted_: bool) -> "BatchAtariGpuEnv":
        _check(load_library().fa_vecenv_set_sorted(self._h, int(bool(sorted_))))
        return self

    def with_reader(self, score_addrs=(), bcd: bool = False, lives_addr: int = 0, lives_shift: int = 0,
                    lives_mask: int = 0) -> "BatchAtariGpuEnv":
        """Reward = change of the score held in RAM bytes `score_addrs` (offsets
        0..127 into the RAM observation, most significant first; packed BCD
        when `bcd`); terminated while (ram[lives_addr] >> lives_shift) &
        lives_mask is 0 (mask 0: never).  Evaluated on the GPU."""
        r = RamReader()
        addrs = list(score_addrs)
        if len(addrs) > 4:
            raise ValueError("at most 4 score bytes")
        for k, a in enumerate(addrs):
            r.score_addr[k] = int(a)
        r.n_score, r.score_bcd = len(addrs), int(bool(bcd))
        r.lives_addr, r.lives_shift, r.lives_mask = int(lives_addr), int(lives_shift), int(lives_mask)
        _check(load_library().fa_vecenv_set_reader(self._h, C.byref(r)))
        return self

    def without_reader(self) -> "BatchAtariGpuEnv":
        _check(load_library().fa_vecenv_set_reader(self._h, None))
        return self

    def rewards(self) -> tuple[np.ndarray, np.ndarray]:
        """Per-env reward (f32) and terminated flag of the last reset / step."""
        rew = np.empty(self.n, np.float32)
        term = np.empty(self.n, np.uint8)
        _check(load_library().fa_vecenv_rewards(self._h, _fp(rew), _p8(term)))
        return rew, term.astype(bool)

    def num_envs(self) -> int:
        return int(load_library().fa_vecenv_num_envs(self._h))

    def step_count(self) -> int:
        return int(load_library().fa_vecenv_step_count(self._h))

    # ---- Env (src/rl4burn_env.rs:318-380)
    def reset(self) -> np.ndarray:
        obs = np.empty(self.n * 128, np.float32)
        _check(load_library().fa_vecenv_reset(self._h, _fp(obs)))
        return obs

    def _finish(self, fn, *args) -> Step:
        obs = np.empty(self.n * 128, np.float32)
        rew, term, trunc = C.c_float(), C.c_int(), C.c_int()
        _check(fn(self._h, *args, _fp(obs), C.byref(rew), C.byref(term), C.byref(trunc)))
        return Step(obs, float(rew.value), bool(term.value), bool(trunc.value))

    def step(self, action: int) -> Step:
        return self._finish(load_library().fa_vecenv_step, int(action) & 0xFFFFFFFF)

    def step_each(self, actions) -> Step:
        a = _u8(actions).reshape(-1)
        if a.size != self.n:
            raise ValueError("actions length must be num_envs")
        return self._finish(load_library().fa_vecenv_step_each, _p8(a))

    def step_device(self, d_obs: int, action: int = 0, d_actions: int | None = None) -> bool:
        """Queue one step: actions from device memory d_actions (n bytes) or
        `action` for every env; f32 obs into device memory d_obs (n * 128
        floats).  Returns `truncated`; call sync() before reading d_obs."""
        trunc = C.c_int()
        _check(load_library().fa_vecenv_step_device(self._h, C.c_void_p(d_actions) if d_actions else None,
                                                    int(action) & 0xFFFFFFFF, C.c_void_p(d_obs), C.byref(trunc)))
        return bool(trunc.value)

    def sync(self):
        _check(load_library().fa_vecenv_sync(self._h))

    def observation_space(self) -> dict:
        return {"kind": "box", "low": 0.0, "high": 1.0, "shape": (self.n * 128,)}

    def action_space(self) -> dict:
        return {"kind": "discrete", "n": NUM_ACTIONS}

    def batch(self) -> BatchAtariGpu:
        """The underlying BatchAtariGpu (states, counters), owned by this env."""
        b = BatchAtariGpu.__new__(BatchAtariGpu)
        b._rom = self._rom
        b.n = self.n
        b._h = C.c_void_p(load_library().fa_vecenv_batch(self._h))
        b._obs = np.empty(self.n * 128, np.uint8)
        b.close = lambda: None            # not owned
        return b


def obs_to_f32(d_in: int, d_out: int, n_bytes: int):
    """u8 observations in device memory to f32 (byte / 255) in device memory."""
    _check(load_library().fa_obs_to_f32(C.c_void_p(d_in), C.c_void_p(d_out), int(n_bytes)))
