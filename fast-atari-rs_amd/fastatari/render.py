"""Batched rendering console: pixel frames from the GPU rendering TIA.

A batch of the reference's rendering `Atari` (src/atari.rs:226-378) behind
its Env interface (src/env.rs, implemented at src/atari.rs:326-378):

    BatchAtariRender(rom, n)        Atari::new                   (:238-246)
    .reset() -> [n, 192, 160] u8    Env::reset                   (:330-340)
    .step(actions) -> frames        Env::step                    (:342-357)
    .step_multi(actions, k)         k x (set_action + run_frame) (:249-305)
    .framebuffer()                  Atari::framebuffer           (:320-323)
    .download_states()              the 256-byte render states
    .num_envs()

Frames are NTSC palette indices, 160 x 192 per env (observation_shape
(192, 160, 1), src/atari.rs:369-371).  A wrong action length raises
ValueError; HIP failures raise FastAtariError.  The kernel is
fa_render_kernel in csrc/fa_render.hip; there is no host fallback."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .env import FastAtariError, _check, _p8, _u8, load_library

FRAME_WIDTH = 160
FRAME_HEIGHT = 192
FRAME_BYTES = FRAME_WIDTH * FRAME_HEIGHT
RENDER_STATE_BYTES = 256


def _lib():
    """The library with every signature bound (env.load_library binds the
    fa_render_* entry points with the rest)."""
    return load_library()


def power_on_state(rom) -> np.ndarray:
    """Atari::new's render state for a ROM (256 bytes)."""
    r = _u8(rom)
    out = np.zeros(RENDER_STATE_BYTES, np.uint8)
    _check(_lib().fa_render_power_on_state(_p8(r), r.size, _p8(out)))
    return out


class BatchAtariRender:
    """n rendering consoles of one ROM stepped together on the GPU."""

    def __init__(self, rom, n: int):
        L = _lib()
        r = _u8(rom)
        h = C.c_void_p()
        _check(L.fa_render_create(_p8(r), r.size, int(n), C.byref(h)))
        self._h = h
        self._n = int(n)

    def close(self):
        """Frees the handle's device memory now (also done on collection)."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib().fa_render_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num_envs(self) -> int:
        return self._n

    def _frames_buf(self) -> np.ndarray:
        return np.empty((self._n, FRAME_HEIGHT, FRAME_WIDTH), np.uint8)

    def reset(self) -> np.ndarray:
        out = self._frames_buf()
        _check(_lib().fa_render_reset(self._h, _p8(out)))
        return out

    def step_multi(self, actions, k: int, frames: bool = True):
        """k frames; actions: k * n bytes laid out [f * n + i].  Returns the
        frames after the last one (or None with frames=False)."""
        a = _u8(actions)
        if a.size != k * self._n:
            raise ValueError(f"actions length {a.size} != k * num_envs = {k * self._n}")
        out = self._frames_buf() if frames else None
        _check(_lib().fa_render_step(self._h, _p8(a), a.size, int(k), _p8(out) if frames else None))
        return out

    def step(self, actions, frames: bool = True):
        return self.step_multi(actions, 1, frames)

    def step_device(self, d_actions: int | None, k: int):
        """Queued k-frame step with actions in device memory (an int pointer)."""
        _check(_lib().fa_render_step_device(self._h, C.c_void_p(d_actions) if d_actions else None, int(k)))

    def sync(self):
        _check(_lib().fa_render_sync(self._h))

    def framebuffer(self) -> np.ndarray:
        out = self._frames_buf()
        _check(_lib().fa_render_download_frames(self._h, _p8(out), self._n))
        return out

    def upload_frames(self, frames: np.ndarray):
        f = np.ascontiguousarray(frames, dtype=np.uint8)
        if f.size != self._n * FRAME_BYTES:
            raise ValueError("frames must hold num_envs * 30720 bytes")
        _check(_lib().fa_render_upload_frames(self._h, _p8(f), self._n))

    def frames_device(self) -> int:
        p = C.c_void_p()
        _check(_lib().fa_render_frames_device(self._h, C.byref(p)))
        return p.value

    def download_states(self) -> np.ndarray:
        out = np.empty((self._n, RENDER_STATE_BYTES), np.uint8)
        _check(_lib().fa_render_download_states(self._h, _p8(out), self._n))
        return out

    def upload_states(self, states: np.ndarray):
        s = np.ascontiguousarray(states, dtype=np.uint8)
        if s.size != self._n * RENDER_STATE_BYTES:
            raise ValueError("states must hold num_envs * 256 bytes")
        _check(_lib().fa_render_upload_states(self._h, _p8(s), self._n))

    def counters(self) -> dict:
        n = self._n
        cpu = np.empty(n, np.uint64)
        fc = np.empty(n, np.uint64)
        ins = np.empty(n, np.uint64)
        st = np.empty(n, np.uint32)
        u64 = C.POINTER(C.c_uint64)
        _check(_lib().fa_render_download_counters(self._h, cpu.ctypes.data_as(u64), fc.ctypes.data_as(u64),
                                                  ins.ctypes.data_as(u64),
                                                  st.ctypes.data_as(C.POINTER(C.c_uint32))))
        return {"cpu_cycles": cpu, "frame_cycles": fc, "instructions": ins, "status": st}

    def set_block_size(self, block: int):
        _check(_lib().fa_render_set_block_size(self._h, int(block)))

    def block_size(self) -> int:
        b = C.c_uint32()
        _check(_lib().fa_render_get_block_size(self._h, C.byref(b)))
        return b.value

    def set_timing(self, enabled: bool):
        _check(_lib().fa_render_set_timing(self._h, 1 if enabled else 0))

    def kernel_time(self, reset: bool = True):
        ms = C.c_double()
        n = C.c_uint64()
        _check(_lib().fa_render_kernel_time(self._h, C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value


__all__ = ["BatchAtariRender", "FRAME_WIDTH", "FRAME_HEIGHT", "FRAME_BYTES", "RENDER_STATE_BYTES",
           "power_on_state", "FastAtariError"]
