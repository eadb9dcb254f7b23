"""Env-step server: client envs resident in HBM, admission by memory
footprint, continuous batching of step requests (native dispatcher in
csrc/fa_serve.hip).

    srv = EnvServer([rom_a, rom_b, ...], budget_bytes=64 << 30)
    cid, obs0 = srv.open(cls, n_envs)        # Busy when over budget
    srv.submit(cid, actions, restart=None, frames=1)
    obs, latency_ms = srv.wait(cid)
    srv.close(cid)

Per env, a step is HeadlessAtari::set_action + frames x run_frame
(src/atari.rs:571-598); a restart first puts the env back to the ROM's reset
state (HeadlessAtari::new + one run_frame, as BatchAtariGpu::reset,
src/cuda_env.rs:131-142).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .env import STATE_DTYPE, _check, _p8, _u8, load_library

# fa_client_spec (include/fastatari.h)
SPEC_DTYPE = np.dtype({"names": ["cls", "n", "rate_mult", "seed"], "formats": ["<u4", "<u4", "<f8", "<u8"],
                       "offsets": [0, 4, 8, 16], "itemsize": 24})


class EnvServer:
    def __init__(self, roms, budget_bytes: int = 64 << 30):
        L = load_library()
        self._roms = [_u8(r) for r in roms]
        n = len(self._roms)
        ptrs = (C.POINTER(C.c_uint8) * n)(*[_p8(r) for r in self._roms])
        lens = (C.c_size_t * n)(*[len(r) for r in self._roms])
        h = C.c_void_p()
        _check(L.fa_server_create(ptrs, lens, n, int(budget_bytes), C.byref(h)))
        self._h = h
        self.nclass = n
        self._n = {}

    def close_server(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().fa_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close_server()
        except Exception:
            pass

    def open(self, cls: int, n_envs: int) -> tuple[int, np.ndarray]:
        obs = np.empty((int(n_envs), 128), np.uint8)
        cid = C.c_int()
        _check(load_library().fa_server_open(self._h, int(cls), int(n_envs), _p8(obs), C.byref(cid)))
        self._n[cid.value] = int(n_envs)
        return cid.value, obs

    def close(self, cid: int):
        _check(load_library().fa_server_close(self._h, int(cid)))
        self._n.pop(cid, None)

    def submit(self, cid: int, actions, restart=None, frames: int = 1):
        a = _u8(actions)
        if a.size != self._n[cid]:
            raise ValueError("actions must hold one byte per client env")
        r = None
        if restart is not None:
            r = _u8(restart)
            if r.size != a.size:
                raise ValueError("restart must hold one byte per client env")
        _check(load_library().fa_server_submit(self._h, int(cid), _p8(a), _p8(r) if r is not None else None,
                                               int(frames)))

    def poll(self, cid: int) -> bool:
        rc = load_library().fa_server_poll(self._h, int(cid))
        if rc < 0:
            _check(rc)
        return rc == 1

    def wait(self, cid: int, want_obs: bool = True) -> tuple[np.ndarray | None, float]:
        obs = np.empty((self._n[cid], 128), np.uint8) if want_obs else None
        ms = C.c_double()
        _check(load_library().fa_server_wait(self._h, int(cid), _p8(obs) if obs is not None else None,
                                             C.byref(ms)))
        return obs, ms.value

    def download(self, cid: int) -> dict:
        n = self._n[cid]
        st = np.zeros(n, STATE_DTYPE)
        cyc = np.zeros(n, np.uint64)
        fc = np.zeros(n, np.uint64)
        status = np.zeros(n, np.uint32)
        p64 = C.POINTER(C.c_uint64)
        _check(load_library().fa_server_download(self._h, int(cid), st.ctypes.data_as(C.c_void_p),
                                                 cyc.ctypes.data_as(p64), fc.ctypes.data_as(p64),
                                                 status.ctypes.data_as(C.POINTER(C.c_uint32))))
        return {"states": st.view(np.uint8).reshape(n, 222), "cpu_cycles": cyc, "frame_cycles": fc,
                "status": status}

    def set_class_config(self, cls: int, sorted_: bool = False, block: int = 0):
        """Opcode grouping and workgroup size (64 / 128 / 256 envs, 0: the
        default) of one ROM class's launches."""
        _check(load_library().fa_server_set_class_config(self._h, int(cls), int(block), int(bool(sorted_))))

    def set_coalesce_us(self, us: int):
        _check(load_library().fa_server_set_coalesce_us(self._h, int(us)))

    def memory(self) -> dict:
        u, b, s = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _check(load_library().fa_server_memory(self._h, C.byref(u), C.byref(b), C.byref(s)))
        return {"used_bytes": u.value, "budget_bytes": b.value, "bytes_per_slot": s.value}

    def run_load(self, cids, specs, duration_s: float, sat=None, frames: int = 1, restart_p: float = 0.01,
                 arrival_seed: int = 1, max_steps: int = 0, steps=None, cap: int = 1 << 20) -> dict:
        """Native load driver (fa_loadgen_run): specs are loadgen.ClientSpec;
        steps (uint64 per client, updated in place) counts each client's
        completed steps, which fixes its next actions / restarts."""
        nc = len(cids)
        spec_arr = np.zeros(nc, SPEC_DTYPE)
        for i, sp in enumerate(specs):
            spec_arr[i] = (sp.cls, sp.n, sp.rate_mult, sp.seed & ((1 << 64) - 1))
        ids = (C.c_int * nc)(*[int(c) for c in cids])
        if steps is None:
            steps = np.zeros(nc, np.uint64)
        sat_arr = None if sat is None else np.ascontiguousarray(sat, dtype=np.float64)
        svc = np.zeros(cap, np.float64)
        e2e = np.zeros(cap, np.float64)
        cls = np.zeros(cap, np.uint32)
        ns = C.c_size_t()
        wall = C.c_double()
        pd = C.POINTER(C.c_double)
        _check(load_library().fa_loadgen_run(
            self._h, ids, spec_arr.ctypes.data_as(C.c_void_p), nc, float(duration_s),
            sat_arr.ctypes.data_as(pd) if sat_arr is not None else None, int(frames), float(restart_p),
            int(arrival_seed), int(max_steps), steps.ctypes.data_as(C.POINTER(C.c_uint64)),
            svc.ctypes.data_as(pd), e2e.ctypes.data_as(pd), cls.ctypes.data_as(C.POINTER(C.c_uint32)), cap,
            C.byref(ns), C.byref(wall)))
        m = min(ns.value, cap)
        return {"service_ms": svc[:m], "e2e_ms": e2e[:m], "cls": cls[:m], "requests": ns.value,
                "wall_s": wall.value, "steps": steps}

    def samples(self, reset: bool = False) -> dict:
        cnt = C.c_size_t()
        launches = C.c_uint64()
        _check(load_library().fa_server_samples(self._h, None, None, None, None, 0, C.byref(cnt),
                                                C.byref(launches), 0))
        n = cnt.value
        ms = np.zeros(n, np.float64)
        cls = np.zeros(n, np.uint32)
        envs = np.zeros(n, np.uint32)
        frames = np.zeros(n, np.uint32)
        p32 = C.POINTER(C.c_uint32)
        _check(load_library().fa_server_samples(self._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                                cls.ctypes.data_as(p32), envs.ctypes.data_as(p32),
                                                frames.ctypes.data_as(p32), n, C.byref(cnt), C.byref(launches),
                                                int(bool(reset))))
        return {"ms": ms, "cls": cls, "envs": envs, "frames": frames, "launches": launches.value}
