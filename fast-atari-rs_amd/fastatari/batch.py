"""Heterogeneous batches: several ROM classes (cartridge + mapper), each with
its own env count, stepped together by one hipGraph launch.

    BatchMixGpu([rom_a, rom_b, ...], [n_a, n_b, ...])
    .reset() -> [N, 128]              every env: power-on + one frame
    .step(actions[k, N], fs[k, N])    k steps, action i held fs frames
    .tune(k)                          per-class block / grouping / steps-per-launch
    .download_states() / .counters()

Env i of class c has the global index class_range(c)[0] + i.  There is no
reference counterpart (the reference binds one ROM per BatchAtariGpu,
src/cuda_env.rs:112-126); the per-env semantics are those of
HeadlessAtari::set_action + fs x run_frame (src/atari.rs:571-598).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .env import STATE_DTYPE, _check, _p8, _u8, load_library


class BatchMixGpu:
    def __init__(self, roms, counts):
        L = load_library()
        if len(roms) != len(counts) or not roms:
            raise ValueError("need one env count per ROM class")
        self._roms = [_u8(r) for r in roms]
        n = len(self._roms)
        ptrs = (C.POINTER(C.c_uint8) * n)(*[_p8(r) for r in self._roms])
        lens = (C.c_size_t * n)(*[len(r) for r in self._roms])
        cnts = (C.c_size_t * n)(*[int(c) for c in counts])
        h = C.c_void_p()
        _check(L.fa_batch_create(ptrs, lens, cnts, n, C.byref(h)))
        self._h = h
        self.n = int(L.fa_batch_num_envs(h))
        self.nclass = n

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().fa_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num_envs(self) -> int:
        return self.n

    def class_range(self, c: int) -> tuple[int, int]:
        f, n = C.c_size_t(), C.c_size_t()
        _check(load_library().fa_batch_class_range(self._h, int(c), C.byref(f), C.byref(n)))
        return f.value, n.value

    def reset(self) -> np.ndarray:
        obs = np.empty((self.n, 128), np.uint8)
        _check(load_library().fa_batch_reset(self._h, _p8(obs)))
        return obs

    def step(self, actions, fs=None, obs_every: bool = False) -> np.ndarray:
        """actions: [k, N] (or [N] for one step); fs: same shape or None."""
        a = _u8(actions)
        k = a.size // max(self.n, 1) if self.n else 1
        if a.size != k * self.n:
            raise ValueError("actions must hold k * num_envs bytes")
        f = None
        if fs is not None:
            f = _u8(fs)
            if f.size != a.size:
                raise ValueError("fs must have the shape of actions")
        obs = np.empty(((k if obs_every else 1) * self.n, 128), np.uint8)
        _check(load_library().fa_batch_step(self._h, _p8(a), _p8(f) if f is not None else None, k,
                                            int(bool(obs_every)), _p8(obs)))
        return obs.reshape(k, self.n, 128) if obs_every else obs

    # ---- device-resident stepping (graph replay on the batch's own buffers)
    def buffers(self, k: int, obs_every: bool = False) -> tuple[int, int, int]:
        da, df, do = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(load_library().fa_batch_buffers(self._h, int(k), int(bool(obs_every)), C.byref(da), C.byref(df),
                                               C.byref(do)))
        return da.value, df.value, do.value

    def launch(self, k: int, use_fs: bool = True, obs_every: bool = False):
        _check(load_library().fa_batch_launch(self._h, int(k), int(bool(use_fs)), int(bool(obs_every))))

    def sync(self) -> float:
        ms = C.c_double()
        _check(load_library().fa_batch_sync(self._h, C.byref(ms)))
        return ms.value

    def write_inputs(self, actions, fs=None):
        """Copies [k, N] actions (and fs) into the batch's device buffers."""
        a = _u8(actions)
        k = a.size // self.n
        da, df, _ = self.buffers(k)
        L = load_library()
        _check(L.fa_memcpy_h2d(C.c_void_p(da), a.ctypes.data_as(C.c_void_p), a.size))
        if fs is not None:
            f = _u8(fs)
            _check(L.fa_memcpy_h2d(C.c_void_p(df), f.ctypes.data_as(C.c_void_p), f.size))
        return k

    def read_obs(self, k: int = 1, obs_every: bool = False) -> np.ndarray:
        _, _, do = self.buffers(k, obs_every)
        rows = (k if obs_every else 1) * self.n
        out = np.empty((rows, 128), np.uint8)
        _check(load_library().fa_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(do), out.nbytes))
        return out

    # ---- state
    def download_states(self) -> np.ndarray:
        out = np.zeros(self.n, STATE_DTYPE)
        _check(load_library().fa_batch_download_states(self._h, out.ctypes.data_as(C.c_void_p), self.n))
        return out

    def download_states_bytes(self) -> np.ndarray:
        return self.download_states().view(np.uint8).reshape(self.n, 222)

    def upload_states(self, states):
        s = np.ascontiguousarray(states)
        if s.dtype != STATE_DTYPE:
            s = np.ascontiguousarray(s, dtype=np.uint8).reshape(self.n, 222)
        _check(load_library().fa_batch_upload_states(self._h, s.ctypes.data_as(C.c_void_p), self.n))

    def counters(self) -> dict:
        n = self.n
        cyc, fc, ins = (np.zeros(n, np.uint64) for _ in range(3))
        st = np.zeros(n, np.uint32)
        p64 = C.POINTER(C.c_uint64)
        _check(load_library().fa_batch_download_counters(
            self._h, cyc.ctypes.data_as(p64), fc.ctypes.data_as(p64), ins.ctypes.data_as(p64),
            st.ctypes.data_as(C.POINTER(C.c_uint32))))
        return {"cpu_cycles": cyc, "frame_cycles": fc, "instructions": ins, "status": st}

    # ---- launch shapes
    def set_class_config(self, c: int, block: int = 0, sorted_: bool = False, k_split: int = 0):
        _check(load_library().fa_batch_set_class_config(self._h, int(c), int(block), int(bool(sorted_)),
                                                        int(k_split)))

    def set_class_lanes(self, c: int, lanes: int):
        """Envs per workgroup below the block (8, 16, 32; 0: the block)."""
        _check(load_library().fa_batch_set_class_lanes(self._h, int(c), int(lanes)))

    def class_config(self, c: int) -> dict:
        b, s, k, ln = C.c_uint32(), C.c_int(), C.c_uint32(), C.c_uint32()
        _check(load_library().fa_batch_get_class_config(self._h, int(c), C.byref(b), C.byref(s), C.byref(k)))
        _check(load_library().fa_batch_get_class_lanes(self._h, int(c), C.byref(ln)))
        return {"block": b.value, "lanes": ln.value, "sorted": bool(s.value), "k_split": k.value}

    def tune(self, actions, fs=None) -> list[dict]:
        """Autotunes every class on these inputs ([k, N]); env states are
        left as they were.  Returns per class the chosen config and the trial
        times (ms) of the chosen and of the default config."""
        k = self.write_inputs(actions, fs)
        trials = (C.c_double * (2 * self.nclass))()
        _check(load_library().fa_batch_tune(self._h, k, int(fs is not None), trials))
        out = []
        for c in range(self.nclass):
            d = self.class_config(c)
            d.update(tuned_ms=trials[2 * c], default_ms=trials[2 * c + 1])
            out.append(d)
        return out
