"""Host build of the rendering console's lane code for the CPU tests (see
tests/cpu_build/render_host.cpp): the same per-lane code the GPU kernel runs,
compiled with clang++ for the host, driven like BatchAtariRender."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpu_build", "render_host.cpp")
CLANG = "/opt/rocm/llvm/bin/clang++"

STATE_BYTES = 256
FRAME_BYTES = 160 * 192


def build(out_dir: str) -> C.CDLL:
    out = os.path.join(out_dir, "render_host.so")
    subprocess.run([CLANG, "-std=c++17", "-O2", "-shared", "-fPIC", "-o", out, SRC], check=True)
    L = C.CDLL(out)
    vp = C.c_void_p
    L.rh_launch.restype = C.c_int
    L.rh_launch.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, vp, vp, vp, vp, vp, C.c_uint32,
                            C.c_uint32, C.c_uint32]
    return L


def rom_image(rom: bytes):
    """The staged image and hotspot window of a ROM (BankScheme::detect,
    src/atari.rs:97-104; the cartridge window of src/atari.rs:141-193)."""
    n = len(rom)
    if n <= 4096:
        return np.frombuffer(bytes(rom[i % n] for i in range(4096)), np.uint8).copy(), 0xFFF, 0
    if n == 8192:
        return np.frombuffer(rom, np.uint8).copy(), 0xFF8, 2
    if n == 16384:
        return np.frombuffer(rom, np.uint8).copy(), 0xFF6, 4
    return np.frombuffer(rom[:32768], np.uint8).copy(), 0xFF4, 8


class HostRender:
    def __init__(self, L, rom: bytes, n: int, power_on: np.ndarray, block: int = 64):
        self.L, self.n, self.block = L, n, block
        self.image, self.hot_lo, self.hot_n = rom_image(rom)
        self.power_on = power_on
        self.states = np.tile(power_on, (n, 1))
        self.frames = np.zeros((n, 192, 160), np.uint8)
        self.cpu = np.zeros(n, np.uint64)
        self.fcyc = np.zeros(n, np.uint64)
        self.instr = np.zeros(n, np.uint64)
        self.status = np.zeros(n, np.uint32)

    def _launch(self, actions, k: int):
        a = None if actions is None else np.ascontiguousarray(actions, dtype=np.uint8)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        rc = self.L.rh_launch(p(self.image), self.image.size, self.hot_lo, self.hot_n, p(self.states), p(self.cpu),
                              p(self.fcyc), p(self.instr), p(self.status), None if a is None else p(a),
                              p(self.frames), self.n, k, self.block)
        assert rc == 0

    def reset(self):
        self.states[:] = self.power_on
        self.frames[:] = 0
        self.cpu[:] = 0
        self.fcyc[:] = 0
        self.instr[:] = 0
        self.status[:] = 0
        self._launch(None, 1)

    def step_multi(self, actions: np.ndarray, k: int):
        self._launch(actions, k)

    def counters(self) -> np.ndarray:
        return np.stack([self.cpu, self.fcyc, self.instr, self.status.astype(np.uint64)], axis=1)
