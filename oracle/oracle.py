"""ORACLE -- test infrastructure only.

ctypes loader for ``oracle/liborc.so``, the C restatement of the reference
CPU emulator (src/cpu.rs, src/headless_tia.rs, src/pia.rs, src/atari.rs of
RPP1011/fast-atari-rs).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module, always as the checker or
the CPU baseline -- never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
STATE_BYTES = 222

_lib = None


def build() -> str:
    """Compile oracle/liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def build_native(out_path: str) -> str:
    """Compile the oracle with -O3 -march=native for this host into out_path
    (the CPU baseline's build; not used by the tests)."""
    subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={out_path}"], check=True)
    return out_path


def use_library(path: str):
    """Load the oracle from `path` (e.g. the native build) for every later
    call; must precede the first use of lib()."""
    global _lib
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    lib(path)


def lib(path: str | None = None):
    global _lib
    if _lib is None:
        path = path or LIB_PATH
        if path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(path)
        vp, u8p, u16p, u32p, u64p = (C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint16),
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_uint64))
        sig = {
            "orc_new": (vp, [u8p, C.c_uint32]),
            "orc_free": (None, [vp]),
            "orc_reset_frame": (None, [vp]),
            "orc_step": (C.c_uint64, [vp, C.c_uint32, C.c_uint32]),
            "orc_get_state": (None, [vp, u8p]),
            "orc_set_state": (None, [vp, u8p]),
            "orc_get_counters": (None, [vp, u64p]),
            "orc_set_cycles": (None, [vp, C.c_uint64]),
            "orc_get_ram": (None, [vp, u8p]),
            "orc_get_mem": (None, [vp, u8p]),
            "orc_h_bus_read": (C.c_uint8, [vp, C.c_uint16]),
            "orc_h_bus_write": (None, [vp, C.c_uint16, C.c_uint8]),
            "orc_batch_new": (vp, [u8p, C.c_uint32, C.c_uint32]),
            "orc_batch_free": (None, [vp]),
            "orc_batch_reset": (None, [vp, C.c_uint32]),
            "orc_batch_run": (None, [vp, u8p, C.c_uint32, C.c_uint32]),
            "orc_batch_run_fs": (None, [vp, u8p, u8p, C.c_uint32]),
            "orc_batch_states": (None, [vp, u8p]),
            "orc_batch_set_states": (None, [vp, u8p]),
            "orc_batch_counters": (None, [vp, u64p]),
            "orc_batch_set_cycles": (None, [vp, u64p]),
            "orc_batch_set_counters": (None, [vp, u64p]),
            "orc_batch_obs": (None, [vp, u8p]),
            "orc_batch_bench": (C.c_double, [vp, u8p, C.c_uint32, C.c_uint32]),
            "orc_flat_new": (vp, []),
            "orc_flat_free": (None, [vp]),
            "orc_flat_load": (None, [vp, C.c_uint16, u8p, C.c_uint32]),
            "orc_flat_poke": (None, [vp, C.c_uint16, C.c_uint8]),
            "orc_flat_peek": (C.c_uint8, [vp, C.c_uint16]),
            "orc_flat_reset": (None, [vp, u32p]),
            "orc_flat_step": (C.c_int, [vp, u32p, u32p, u32p]),
            "orc_alu_sweep": (None, [C.c_uint8, u16p]),
            "orc_component_run": (None, [u8p, C.c_uint32, u8p, u32p, C.c_uint32, u32p]),
            "orc_pia_sweep": (None, [u16p, u16p, C.c_uint32, u16p, u16p, u8p]),
            "orc_r_new": (vp, [u8p, C.c_uint32]),
            "orc_r_free": (None, [vp]),
            "orc_r_step": (C.c_uint64, [vp, C.c_uint32, C.c_uint32]),
            "orc_r_get_state": (None, [vp, u8p]),
            "orc_r_set_state": (None, [vp, u8p]),
            "orc_r_get_frame": (None, [vp, u8p]),
            "orc_r_set_frame": (None, [vp, u8p]),
            "orc_r_get_counters": (None, [vp, u64p]),
            "orc_r_bus_read": (C.c_uint8, [vp, C.c_uint16]),
            "orc_r_bus_write": (None, [vp, C.c_uint16, C.c_uint8]),
            "orc_r_tia_write": (None, [vp, C.c_uint16, C.c_uint8]),
            "orc_r_tia_ticks": (None, [vp, C.c_uint32]),
            "orc_rbatch_new": (vp, [u8p, C.c_uint32, C.c_uint32]),
            "orc_rbatch_free": (None, [vp]),
            "orc_rbatch_reset": (None, [vp, C.c_uint32]),
            "orc_rbatch_run": (None, [vp, u8p, C.c_uint32, C.c_uint32]),
            "orc_rbatch_states": (None, [vp, u8p]),
            "orc_rbatch_set_states": (None, [vp, u8p]),
            "orc_rbatch_frames": (None, [vp, u8p]),
            "orc_rbatch_counters": (None, [vp, u64p]),
            "orc_rbatch_bench": (C.c_double, [vp, u8p, C.c_uint32, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def _u8(buf) -> np.ndarray:
    return np.ascontiguousarray(np.frombuffer(bytes(buf), dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf,
                                dtype=np.uint8)


class Console:
    """One HeadlessAtari (src/atari.rs:555-620)."""

    def __init__(self, rom):
        self._rom = _u8(rom)
        self.h = lib().orc_new(_p(self._rom, C.c_uint8), len(self._rom))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def reset_frame(self):
        lib().orc_reset_frame(self.h)

    def step(self, action: int, frames: int = 1) -> int:
        return int(lib().orc_step(self.h, action, frames))

    def state(self) -> np.ndarray:
        out = np.zeros(STATE_BYTES, np.uint8)
        lib().orc_get_state(self.h, _p(out, C.c_uint8))
        return out

    def set_state(self, s):
        s = _u8(s)
        lib().orc_set_state(self.h, _p(s, C.c_uint8))

    def counters(self):
        out = np.zeros(4, np.uint64)
        lib().orc_get_counters(self.h, _p(out, C.c_uint64))
        return [int(x) for x in out]

    def set_cycles(self, c: int):
        lib().orc_set_cycles(self.h, c)

    def ram(self) -> np.ndarray:
        out = np.zeros(128, np.uint8)
        lib().orc_get_ram(self.h, _p(out, C.c_uint8))
        return out

    def mem(self) -> np.ndarray:
        out = np.zeros(8192, np.uint8)
        lib().orc_get_mem(self.h, _p(out, C.c_uint8))
        return out

    def bus_read(self, addr: int) -> int:
        return int(lib().orc_h_bus_read(self.h, addr))

    def bus_write(self, addr: int, v: int):
        lib().orc_h_bus_write(self.h, addr, v)


class Batch:
    """N independent consoles; mirrors BatchAtariGpu's semantics on the CPU."""

    def __init__(self, rom, n: int, nthreads: int = 8):
        self._rom = _u8(rom)
        self.n = n
        self.nthreads = nthreads
        self.h = lib().orc_batch_new(_p(self._rom, C.c_uint8), len(self._rom), n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_batch_free(self.h)
            self.h = None

    def reset(self):
        """BatchAtariGpu::reset (src/cuda_env.rs:131-142): new + one frame."""
        lib().orc_batch_reset(self.h, self.nthreads)

    def run(self, actions: np.ndarray):
        """actions: [frames, n] uint8; one set_action + run_frame per entry."""
        a = np.ascontiguousarray(actions, dtype=np.uint8)
        frames = a.shape[0] if a.ndim == 2 else 1
        lib().orc_batch_run(self.h, _p(a, C.c_uint8), frames, self.nthreads)

    def run_frameskip(self, actions: np.ndarray, fs: np.ndarray):
        a = np.ascontiguousarray(actions, dtype=np.uint8)
        f = np.ascontiguousarray(fs, dtype=np.uint8)
        lib().orc_batch_run_fs(self.h, _p(a, C.c_uint8), _p(f, C.c_uint8), self.nthreads)

    def states(self) -> np.ndarray:
        out = np.zeros((self.n, STATE_BYTES), np.uint8)
        lib().orc_batch_states(self.h, _p(out, C.c_uint8))
        return out

    def set_states(self, s: np.ndarray):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        lib().orc_batch_set_states(self.h, _p(s, C.c_uint8))

    def counters(self) -> np.ndarray:
        """[n, 4]: cpu.cycles, last frame cycles, instructions, error."""
        out = np.zeros((self.n, 4), np.uint64)
        lib().orc_batch_counters(self.h, _p(out, C.c_uint64))
        return out

    def set_cycles(self, c: np.ndarray):
        c = np.ascontiguousarray(c, dtype=np.uint64)
        lib().orc_batch_set_cycles(self.h, _p(c, C.c_uint64))

    def set_counters(self, c: np.ndarray):
        """[n, 4] as counters()."""
        c = np.ascontiguousarray(c, dtype=np.uint64)
        lib().orc_batch_set_counters(self.h, _p(c, C.c_uint64))

    def restart(self, mask, snap_state: np.ndarray, snap_counters: np.ndarray):
        """Envs where mask is set go back to a reset snapshot (state and
        counters), as a restarted env of the server."""
        m = np.asarray(mask, dtype=bool)
        if not m.any():
            return
        st = self.states()
        st[m] = snap_state
        self.set_states(st)
        cn = self.counters()
        cn[m] = snap_counters
        self.set_counters(cn)

    def obs(self) -> np.ndarray:
        out = np.zeros((self.n, 128), np.uint8)
        lib().orc_batch_obs(self.h, _p(out, C.c_uint8))
        return out

    def bench(self, actions: np.ndarray, nthreads: int) -> float:
        a = np.ascontiguousarray(actions, dtype=np.uint8)
        return float(lib().orc_batch_bench(self.h, _p(a, C.c_uint8), a.shape[0], nthreads))


RSTATE_BYTES = 256
FRAME_BYTES = 160 * 192


class RConsole:
    """One rendering Atari (src/atari.rs:226-318 over src/tia.rs)."""

    def __init__(self, rom):
        self._rom = _u8(rom)
        self.h = lib().orc_r_new(_p(self._rom, C.c_uint8), len(self._rom))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_r_free(self.h)
            self.h = None

    def step(self, action: int, frames: int = 1) -> int:
        return int(lib().orc_r_step(self.h, action, frames))

    def state(self) -> np.ndarray:
        out = np.zeros(RSTATE_BYTES, np.uint8)
        lib().orc_r_get_state(self.h, _p(out, C.c_uint8))
        return out

    def set_state(self, s):
        s = _u8(s)
        lib().orc_r_set_state(self.h, _p(s, C.c_uint8))

    def frame(self) -> np.ndarray:
        out = np.zeros(FRAME_BYTES, np.uint8)
        lib().orc_r_get_frame(self.h, _p(out, C.c_uint8))
        return out.reshape(192, 160)

    def set_frame(self, f):
        f = _u8(np.ascontiguousarray(f).reshape(-1))
        lib().orc_r_set_frame(self.h, _p(f, C.c_uint8))

    def counters(self):
        out = np.zeros(4, np.uint64)
        lib().orc_r_get_counters(self.h, _p(out, C.c_uint64))
        return [int(x) for x in out]

    def bus_read(self, addr: int) -> int:
        return int(lib().orc_r_bus_read(self.h, addr))

    def bus_write(self, addr: int, v: int):
        lib().orc_r_bus_write(self.h, addr, v)

    def tia_write(self, addr: int, v: int):
        """A TIA register write with no bus cycle around it."""
        lib().orc_r_tia_write(self.h, addr, v)

    def tia_ticks(self, n: int):
        """n TIA clocks with no CPU."""
        lib().orc_r_tia_ticks(self.h, n)


class RBatch:
    """N rendering consoles (frames as 160 x 192 palette indices)."""

    def __init__(self, rom, n: int, nthreads: int = 8):
        self._rom = _u8(rom)
        self.n = n
        self.nthreads = nthreads
        self.h = lib().orc_rbatch_new(_p(self._rom, C.c_uint8), len(self._rom), n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_rbatch_free(self.h)
            self.h = None

    def reset(self):
        """Atari::new + one run_frame per env."""
        lib().orc_rbatch_reset(self.h, self.nthreads)

    def run(self, actions: np.ndarray):
        a = np.ascontiguousarray(actions, dtype=np.uint8)
        frames = a.shape[0] if a.ndim == 2 else 1
        lib().orc_rbatch_run(self.h, _p(a, C.c_uint8), frames, self.nthreads)

    def states(self) -> np.ndarray:
        out = np.zeros((self.n, RSTATE_BYTES), np.uint8)
        lib().orc_rbatch_states(self.h, _p(out, C.c_uint8))
        return out

    def set_states(self, s: np.ndarray):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        lib().orc_rbatch_set_states(self.h, _p(s, C.c_uint8))

    def frames(self) -> np.ndarray:
        out = np.zeros((self.n, 192, 160), np.uint8)
        lib().orc_rbatch_frames(self.h, _p(out, C.c_uint8))
        return out

    def counters(self) -> np.ndarray:
        out = np.zeros((self.n, 4), np.uint64)
        lib().orc_rbatch_counters(self.h, _p(out, C.c_uint64))
        return out

    def bench(self, actions: np.ndarray, nthreads: int) -> float:
        a = np.ascontiguousarray(actions, dtype=np.uint8)
        return float(lib().orc_rbatch_bench(self.h, _p(a, C.c_uint8), a.shape[0], nthreads))


class Flat:
    """FlatMemory test bus (src/cpu.rs:69-90) with a bus-access trace."""

    def __init__(self):
        self.h = lib().orc_flat_new()
        self.regs = np.zeros(6, np.uint32)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_flat_free(self.h)
            self.h = None

    def load(self, base: int, data):
        d = _u8(data)
        lib().orc_flat_load(self.h, base, _p(d, C.c_uint8), len(d))

    def poke(self, addr: int, v: int):
        lib().orc_flat_poke(self.h, addr, v)

    def peek(self, addr: int) -> int:
        return int(lib().orc_flat_peek(self.h, addr))

    def reset(self):
        lib().orc_flat_reset(self.h, _p(self.regs, C.c_uint32))

    def step(self):
        """Returns (cycles, trace) with trace = [(kind 'R'|'W', addr, value)]."""
        tr = np.zeros(64, np.uint32)
        nt = np.zeros(1, np.uint32)
        cy = lib().orc_flat_step(self.h, _p(self.regs, C.c_uint32), _p(tr, C.c_uint32), _p(nt, C.c_uint32))
        trace = [("W" if (t >> 24) else "R", int((t >> 8) & 0xFFFF), int(t & 0xFF)) for t in tr[: int(nt[0])]]
        return cy, trace

    @property
    def pc(self):
        return int(self.regs[0])

    @property
    def a(self):
        return int(self.regs[2])

    def set(self, pc=None, sp=None, a=None, x=None, y=None, p=None):
        for i, v in enumerate((pc, sp, a, x, y, p)):
            if v is not None:
                self.regs[i] = v


def alu_sweep(op: int) -> np.ndarray:
    """Run opcode `op` (immediate mode) over every (a, v, C, D); returns
    (p << 8) | a, indexed by a | v<<8 | C<<16 | D<<17."""
    out = np.zeros(262144, np.uint16)
    lib().orc_alu_sweep(op, _p(out, C.c_uint16))
    return out


def component_run(state, ops, rom=None):
    """Apply TIA/PIA operations (see orc_component_run) to a packed state.
    Returns (new_state, results)."""
    rom = _u8(rom if rom is not None else bytes(4096))
    st = np.array(_u8(state), dtype=np.uint8, copy=True)
    o = np.ascontiguousarray(ops, dtype=np.uint32)
    res = np.zeros(len(o), np.uint32)
    lib().orc_component_run(_p(rom, C.c_uint8), len(rom), _p(st, C.c_uint8), _p(o, C.c_uint32), len(o),
                            _p(res, C.c_uint32))
    return st, res


def pia_sweep(states: np.ndarray, n: np.ndarray):
    """states: [k, 4] (tval, tps, tsel, tuf); returns (tick_n result, per-cycle
    tick() result, per-cycle run underflowed?)."""
    s = np.ascontiguousarray(states, dtype=np.uint16)
    nn = np.ascontiguousarray(n, dtype=np.uint16)
    k = len(nn)
    a = np.zeros((k, 4), np.uint16)
    b = np.zeros((k, 4), np.uint16)
    u = np.zeros(k, np.uint8)
    lib().orc_pia_sweep(_p(s, C.c_uint16), _p(nn, C.c_uint16), k, _p(a, C.c_uint16), _p(b, C.c_uint16),
                        _p(u, C.c_uint8))
    return a, b, u.astype(bool)
