"""ctypes binding of the C ABI and the BatchAtariGpu host class.

Method names, argument meaning and error behaviour follow the reference's
`BatchAtariGpu` (src/cuda_env.rs:99-191):

    BatchAtariGpu(rom, n)          BatchAtariGpu::new          (:112-126)
    .set_sorted(flag)              set_sorted                  (:129)
    .reset() -> [n, 128] uint8     reset                       (:131-142)
    .step(actions) -> n*128 uint8  step                        (:144-161)
    .step_multi(actions, k)        step_multi                  (:166-184)
    .download_states()             download_states             (:186-188)
    .num_envs()                    num_envs                    (:190)

A wrong action length raises ValueError (the reference's assert_eq! panics);
HIP failures raise FastAtariError (the reference's DriverError).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))

FA_OK = 0
FA_EINVAL = -1
FA_EHIP = -2
FA_ESTATE = -3
FA_ENOMEM = -4
FA_EBUSY = -5

# AtariStateGpu (include/fastatari.h; src/cuda_env.rs:12-38), 222 bytes.
_FIELDS = [
    ("a", "u1", 0), ("x", "u1", 1), ("y", "u1", 2), ("sp", "u1", 3), ("pc", "<u2", 4), ("status", "u1", 6),
    ("ram", ("u1", 128), 8),
    ("port_a_output", "u1", 136), ("port_a_ddr", "u1", 137), ("port_a_input", "u1", 138),
    ("port_b_output", "u1", 139), ("port_b_ddr", "u1", 140), ("port_b_input", "u1", 141),
    ("timer_value", "<u2", 142), ("timer_prescaler", "<u2", 144), ("timer_prescaler_select", "<u2", 146),
    ("timer_underflow", "u1", 148),
    ("tia_clock", "<u2", 152), ("tia_scanline", "<u2", 154), ("tia_wsync", "u1", 156),
    ("tia_frame_complete", "u1", 157), ("tia_vsync", "u1", 158), ("tia_vblank", "u1", 159),
    ("cxm0p", "u1", 160), ("cxm1p", "u1", 161), ("cxp0fb", "u1", 162), ("cxp1fb", "u1", 163),
    ("cxm0fb", "u1", 164), ("cxm1fb", "u1", 165), ("cxblpf", "u1", 166), ("cxppmm", "u1", 167),
    ("inpt4", "u1", 168), ("inpt5", "u1", 169), ("paddle0", "u1", 170), ("paddle1", "u1", 171),
    ("paddle_counter", "<u2", 172), ("paddle_dumped", "u1", 174),
    ("pos_p0", "<u2", 176), ("pos_p1", "<u2", 178), ("pos_m0", "<u2", 180), ("pos_m1", "<u2", 182),
    ("pos_bl", "<u2", 184),
    ("hmp0", "u1", 186), ("hmp1", "u1", 187), ("hmm0", "u1", 188), ("hmm1", "u1", 189), ("hmbl", "u1", 190),
    ("grp0", "u1", 192), ("grp1", "u1", 193), ("grp0_old", "u1", 194), ("grp1_old", "u1", 195),
    ("enam0", "u1", 196), ("enam1", "u1", 197), ("enabl", "u1", 198), ("enabl_old", "u1", 199),
    ("vdelp0", "u1", 200), ("vdelp1", "u1", 201), ("vdelbl", "u1", 202),
    ("resmp0", "u1", 203), ("resmp1", "u1", 204), ("nusiz0", "u1", 205), ("nusiz1", "u1", 206),
    ("colup0", "u1", 207), ("colup1", "u1", 208), ("colupf", "u1", 209), ("colubk", "u1", 210),
    ("ctrlpf", "u1", 211), ("refp0", "u1", 212), ("refp1", "u1", 213),
    ("pf0", "u1", 214), ("pf1", "u1", 215), ("pf2", "u1", 216), ("bank", "u1", 217), ("scheme", "u1", 218),
]
STATE_DTYPE = np.dtype({"names": [f[0] for f in _FIELDS], "formats": [f[1] for f in _FIELDS],
                        "offsets": [f[2] for f in _FIELDS], "itemsize": 222})


class FastAtariError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fastatari error {code}: {msg}")
        self.code = code


class Busy(FastAtariError):
    """FA_EBUSY: over the server's memory budget, or a request in flight."""


_lib = None


def lib_path() -> str:
    return os.path.join(PKG, "libfastatari_hip.so")


def load_library():
    """Load libfastatari_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise FastAtariError(FA_EHIP, f"{path} is not built (run fastatari.build.build())")
    L = C.CDLL(path)
    vp, u8p, u32p, u64p, sz = (C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint64), C.c_size_t)
    sigs = {
        "fa_create": (C.c_int, [u8p, sz, sz, C.POINTER(vp)]),
        "fa_destroy": (None, [vp]),
        "fa_set_sorted": (C.c_int, [vp, C.c_int]),
        "fa_reset": (C.c_int, [vp, u8p]),
        "fa_step": (C.c_int, [vp, u8p, sz, u8p]),
        "fa_step_multi": (C.c_int, [vp, u8p, sz, C.c_uint32, u8p]),
        "fa_download_states": (C.c_int, [vp, vp, sz]),
        "fa_num_envs": (sz, [vp]),
        "fa_last_error": (C.c_char_p, []),
        "fa_upload_states": (C.c_int, [vp, vp, sz]),
        "fa_download_counters": (C.c_int, [vp, u64p, u64p, u64p, u32p]),
        "fa_upload_cpu_cycles": (C.c_int, [vp, u64p, sz]),
        "fa_step_frameskip": (C.c_int, [vp, u8p, u8p, sz, u8p]),
        "fa_step_device": (C.c_int, [vp, vp, C.c_uint32, vp]),
        "fa_step_frameskip_device": (C.c_int, [vp, vp, vp, vp]),
        "fa_sync": (C.c_int, [vp]),
        "fa_device_alloc": (C.c_int, [sz, C.POINTER(vp)]),
        "fa_device_free": (C.c_int, [vp]),
        "fa_memcpy_h2d": (C.c_int, [vp, vp, sz]),
        "fa_memcpy_d2h": (C.c_int, [vp, vp, sz]),
        "fa_set_timing": (C.c_int, [vp, C.c_int]),
        "fa_kernel_time": (C.c_int, [vp, C.POINTER(C.c_double), u64p, C.c_int]),
        "fa_set_block_size": (C.c_int, [vp, C.c_uint32]),
        "fa_get_block_size": (C.c_int, [vp, u32p]),
        "fa_get_max_lds": (C.c_int, [vp, C.POINTER(C.c_size_t)]),
        # heterogeneous batches
        "fa_batch_create": (C.c_int, [C.POINTER(u8p), C.POINTER(sz), C.POINTER(sz), sz, C.POINTER(vp)]),
        "fa_batch_destroy": (None, [vp]),
        "fa_batch_num_envs": (sz, [vp]),
        "fa_batch_num_classes": (sz, [vp]),
        "fa_batch_class_range": (C.c_int, [vp, sz, C.POINTER(sz), C.POINTER(sz)]),
        "fa_batch_reset": (C.c_int, [vp, u8p]),
        "fa_batch_step": (C.c_int, [vp, u8p, u8p, C.c_uint32, C.c_int, u8p]),
        "fa_batch_buffers": (C.c_int, [vp, C.c_uint32, C.c_int, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
        "fa_batch_launch": (C.c_int, [vp, C.c_uint32, C.c_int, C.c_int]),
        "fa_batch_sync": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "fa_batch_download_states": (C.c_int, [vp, vp, sz]),
        "fa_batch_upload_states": (C.c_int, [vp, vp, sz]),
        "fa_batch_download_counters": (C.c_int, [vp, u64p, u64p, u64p, u32p]),
        "fa_batch_set_class_config": (C.c_int, [vp, sz, C.c_uint32, C.c_int, C.c_uint32]),
        "fa_batch_get_class_config": (C.c_int, [vp, sz, u32p, C.POINTER(C.c_int), u32p]),
        "fa_batch_set_class_lanes": (C.c_int, [vp, sz, C.c_uint32]),
        "fa_batch_get_class_lanes": (C.c_int, [vp, sz, u32p]),
        "fa_batch_tune": (C.c_int, [vp, C.c_uint32, C.c_int, C.POINTER(C.c_double)]),
        # server
        "fa_server_create": (C.c_int, [C.POINTER(u8p), C.POINTER(sz), sz, sz, C.POINTER(vp)]),
        "fa_server_destroy": (None, [vp]),
        "fa_server_open": (C.c_int, [vp, sz, sz, u8p, C.POINTER(C.c_int)]),
        "fa_server_close": (C.c_int, [vp, C.c_int]),
        "fa_server_submit": (C.c_int, [vp, C.c_int, u8p, u8p, C.c_uint32]),
        "fa_server_poll": (C.c_int, [vp, C.c_int]),
        "fa_server_wait": (C.c_int, [vp, C.c_int, u8p, C.POINTER(C.c_double)]),
        "fa_server_download": (C.c_int, [vp, C.c_int, vp, u64p, u64p, u32p]),
        "fa_server_client_info": (C.c_int, [vp, C.c_int, C.POINTER(sz), C.POINTER(sz)]),
        "fa_server_set_class_config": (C.c_int, [vp, sz, C.c_uint32, C.c_int]),
        "fa_descriptor_table": (C.c_int, [u32p, sz]),
        "fa_vecenv_create": (C.c_int, [u8p, sz, sz, C.POINTER(vp)]),
        "fa_vecenv_destroy": (None, [vp]),
        "fa_vecenv_set_max_steps": (C.c_int, [vp, C.c_uint64]),
        "fa_vecenv_set_frameskip": (C.c_int, [vp, C.c_uint32]),
        "fa_vecenv_set_sorted": (C.c_int, [vp, C.c_int]),
        "fa_vecenv_num_envs": (sz, [vp]),
        "fa_vecenv_step_count": (C.c_uint64, [vp]),
        "fa_vecenv_batch": (vp, [vp]),
        "fa_vecenv_reset": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "fa_vecenv_step": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "fa_vecenv_step_each": (C.c_int, [vp, u8p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "fa_vecenv_step_device": (C.c_int, [vp, vp, C.c_uint32, vp, C.POINTER(C.c_int)]),
        "fa_vecenv_sync": (C.c_int, [vp]),
        "fa_obs_to_f32": (C.c_int, [vp, vp, sz]),
        "fa_step_noalloc": (C.c_int, [u8p, sz, vp, vp, vp, sz, C.c_uint32, vp, vp]),
        "fa_probe_valu_peak": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(C.c_double), u64p]),
        "fa_vecenv_set_reader": (C.c_int, [vp, vp]),
        "fa_vecenv_rewards": (C.c_int, [vp, C.POINTER(C.c_float), u8p]),
        "fa_server_memory": (C.c_int, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
        "fa_server_set_coalesce_us": (C.c_int, [vp, C.c_uint32]),
        "fa_server_samples": (C.c_int, [vp, C.POINTER(C.c_double), u32p, u32p, u32p, sz, C.POINTER(sz), u64p,
                                        C.c_int]),
        # rendering console
        "fa_render_create": (C.c_int, [u8p, sz, sz, C.POINTER(vp)]),
        "fa_render_destroy": (None, [vp]),
        "fa_render_num_envs": (sz, [vp]),
        "fa_render_reset": (C.c_int, [vp, u8p]),
        "fa_render_step": (C.c_int, [vp, u8p, sz, C.c_uint32, u8p]),
        "fa_render_step_device": (C.c_int, [vp, vp, C.c_uint32]),
        "fa_render_sync": (C.c_int, [vp]),
        "fa_render_frames_device": (C.c_int, [vp, C.POINTER(vp)]),
        "fa_render_download_frames": (C.c_int, [vp, u8p, sz]),
        "fa_render_upload_frames": (C.c_int, [vp, u8p, sz]),
        "fa_render_download_states": (C.c_int, [vp, u8p, sz]),
        "fa_render_upload_states": (C.c_int, [vp, u8p, sz]),
        "fa_render_download_counters": (C.c_int, [vp, u64p, u64p, u64p, u32p]),
        "fa_render_power_on_state": (C.c_int, [u8p, sz, u8p]),
        "fa_render_set_block_size": (C.c_int, [vp, C.c_uint32]),
        "fa_render_get_block_size": (C.c_int, [vp, u32p]),
        "fa_render_set_timing": (C.c_int, [vp, C.c_int]),
        "fa_render_kernel_time": (C.c_int, [vp, C.POINTER(C.c_double), u64p, C.c_int]),
        "fa_loadgen_run": (C.c_int, [vp, C.POINTER(C.c_int), vp, sz, C.c_double, C.POINTER(C.c_double),
                                     C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, u64p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), u32p, sz, C.POINTER(sz),
                                     C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != FA_OK:
        msg = load_library().fa_last_error().decode(errors="replace")
        if rc == FA_EINVAL:
            raise ValueError(msg)
        if rc == FA_EBUSY:
            raise Busy(rc, msg)
        raise FastAtariError(rc, msg)


def _u8(a) -> np.ndarray:
    if isinstance(a, np.ndarray):
        return np.ascontiguousarray(a, dtype=np.uint8)
    return np.frombuffer(bytes(a), dtype=np.uint8).copy()


def _p8(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class DeviceBuffer:
    """A raw device allocation owned by Python (actions / obs kept in HBM)."""

    def __init__(self, nbytes: int):
        L = load_library()
        self.nbytes = nbytes
        p = C.c_void_p()
        _check(L.fa_device_alloc(nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, host: np.ndarray, offset: int = 0):
        h = np.ascontiguousarray(host)
        assert offset + h.nbytes <= self.nbytes
        _check(load_library().fa_memcpy_h2d(C.c_void_p(self.ptr + offset), h.ctypes.data_as(C.c_void_p), h.nbytes))

    def download(self, nbytes: int | None = None, offset: int = 0) -> np.ndarray:
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        _check(load_library().fa_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), n))
        return out

    def at(self, offset: int) -> int:
        return self.ptr + offset

    def free(self):
        if getattr(self, "ptr", None):
            load_library().fa_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


NOALLOC_MAX_ENVS = 65536


def step_noalloc(rom, d_states: int, n: int, k: int = 1, d_actions: int | None = None, d_obs: int | None = None,
                 d_status: int | None = None, stream: int | None = None):
    """fa_step_noalloc: k frames on n envs whose states are in caller device
    memory (device pointers as ints), with no allocation by the library;
    enqueued on `stream` (NULL stream by default), returns without waiting."""
    r = _u8(rom)
    _check(load_library().fa_step_noalloc(_p8(r), r.size, C.c_void_p(d_states),
                                          C.c_void_p(d_status) if d_status else None,
                                          C.c_void_p(d_actions) if d_actions else None, int(n), int(k),
                                          C.c_void_p(d_obs) if d_obs else None,
                                          C.c_void_p(stream) if stream else None))


def probe_valu_peak(iters: int = 4096, reps: int = 5) -> dict:
    """Measured VALU issue peak (fa_probe_valu_peak): wave64 v_add_u32 per
    second over the chip."""
    ms = C.c_double()
    adds = C.c_uint64()
    _check(load_library().fa_probe_valu_peak(int(iters), int(reps), C.byref(ms), C.byref(adds)))
    return {"ms_per_launch": ms.value, "adds_per_launch": adds.value,
            "g_wave_instr_per_s": adds.value / (ms.value * 1e-3) / 1e9}


class BatchAtariGpu:
    """N Atari 2600 consoles stepped on the GPU (one wave64 lane each)."""

    def __init__(self, rom, n: int):
        L = load_library()
        self._rom = _u8(rom)
        self.n = int(n)
        h = C.c_void_p()
        _check(L.fa_create(_p8(self._rom), len(self._rom), self.n, C.byref(h)))
        self._h = h
        self._obs = np.empty(self.n * 128, np.uint8)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().fa_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the reference's method set
    def set_sorted(self, sorted_: bool):
        _check(load_library().fa_set_sorted(self._h, int(bool(sorted_))))

    def reset(self) -> np.ndarray:
        obs = np.empty((self.n, 128), np.uint8)
        _check(load_library().fa_reset(self._h, _p8(obs)))
        return obs

    def step(self, actions) -> np.ndarray:
        a = _u8(actions)
        obs = np.empty(self.n * 128, np.uint8)
        _check(load_library().fa_step(self._h, _p8(a), a.size, _p8(obs)))
        return obs

    def step_multi(self, actions, k: int) -> np.ndarray:
        a = _u8(actions).reshape(-1)
        obs = np.empty(self.n * 128, np.uint8)
        _check(load_library().fa_step_multi(self._h, _p8(a), a.size, int(k), _p8(obs)))
        return obs

    def download_states(self) -> np.ndarray:
        out = np.zeros(self.n, STATE_DTYPE)
        _check(load_library().fa_download_states(self._h, out.ctypes.data_as(C.c_void_p), self.n))
        return out

    def num_envs(self) -> int:
        return int(load_library().fa_num_envs(self._h))

    # ---- extensions
    def download_states_bytes(self) -> np.ndarray:
        return self.download_states().view(np.uint8).reshape(self.n, 222)

    def upload_states(self, states):
        s = np.ascontiguousarray(states)
        if s.dtype != STATE_DTYPE:
            s = np.ascontiguousarray(s, dtype=np.uint8)
            if s.size != self.n * 222:
                raise ValueError(f"states: expected {self.n} x 222 bytes, got {s.size} bytes")
            s = s.reshape(self.n, 222)
        elif s.size != self.n:
            raise ValueError(f"states: expected {self.n} states, got {s.size}")
        _check(load_library().fa_upload_states(self._h, s.ctypes.data_as(C.c_void_p), s.shape[0]))

    def counters(self) -> dict:
        n = self.n
        cyc = np.zeros(n, np.uint64)
        fc = np.zeros(n, np.uint64)
        ins = np.zeros(n, np.uint64)
        st = np.zeros(n, np.uint32)
        _check(load_library().fa_download_counters(
            self._h, cyc.ctypes.data_as(C.POINTER(C.c_uint64)), fc.ctypes.data_as(C.POINTER(C.c_uint64)),
            ins.ctypes.data_as(C.POINTER(C.c_uint64)), st.ctypes.data_as(C.POINTER(C.c_uint32))))
        return {"cpu_cycles": cyc, "frame_cycles": fc, "instructions": ins, "status": st}

    def upload_cpu_cycles(self, cycles):
        c = np.ascontiguousarray(cycles, dtype=np.uint64)
        _check(load_library().fa_upload_cpu_cycles(self._h, c.ctypes.data_as(C.POINTER(C.c_uint64)), c.size))

    def step_frameskip(self, actions, fs) -> np.ndarray:
        a = _u8(actions)
        f = _u8(fs)
        if a.size != f.size:
            raise ValueError(f"actions ({a.size}) and frameskips ({f.size}) differ in length")
        obs = np.empty(self.n * 128, np.uint8)
        _check(load_library().fa_step_frameskip(self._h, _p8(a), _p8(f), a.size, _p8(obs)))
        return obs

    def step_device(self, d_actions: int, k: int = 1, d_obs: int | None = None):
        _check(load_library().fa_step_device(self._h, C.c_void_p(d_actions), int(k),
                                             C.c_void_p(d_obs) if d_obs else None))

    def step_frameskip_device(self, d_actions: int, d_fs: int, d_obs: int | None = None):
        _check(load_library().fa_step_frameskip_device(self._h, C.c_void_p(d_actions), C.c_void_p(d_fs),
                                                       C.c_void_p(d_obs) if d_obs else None))

    def sync(self):
        _check(load_library().fa_sync(self._h))

    def set_timing(self, on: bool):
        _check(load_library().fa_set_timing(self._h, int(bool(on))))

    def kernel_time(self, reset: bool = False):
        ms = C.c_double()
        cnt = C.c_uint64()
        _check(load_library().fa_kernel_time(self._h, C.byref(ms), C.byref(cnt), int(reset)))
        return ms.value, cnt.value

    def set_block_size(self, block: int):
        _check(load_library().fa_set_block_size(self._h, int(block)))

    def max_lds(self) -> int:
        b = C.c_size_t()
        _check(load_library().fa_get_max_lds(self._h, C.byref(b)))
        return b.value

    def block_size(self) -> int:
        b = C.c_uint32()
        _check(load_library().fa_get_block_size(self._h, C.byref(b)))
        return b.value
