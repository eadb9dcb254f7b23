"""The rendering console without a GPU: the library's power-on state against
the oracle's Atari::new, and the kernel's per-lane code (csrc/fa_render_lane.hpp,
built for the host by tests/cpu_build/render_host.cpp) against the rendering
oracle -- states, 160 x 192 frames and counters after every launch, on every
ROM class, a graphics kernel, frames cut by run_frame's cycle cap, an illegal
opcode and a held VSYNC.  The GPU tests run the same code on the GPU."""
import numpy as np
import pytest

import render_host as rh

RH = {}


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    if "L" not in RH:
        RH["L"] = rh.build(str(tmp_path_factory.mktemp("render_host")))
    return RH["L"]


def _roms():
    from fastatari import workloads as W
    mix = W.mix_roms(W.config_seed(4))
    return {
        "config1_fixed4k": W.rom_for(1),
        "config2_f8": W.rom_for(2),
        "config3_f6_branchy": W.rom_for(3),
        "mix_fixed2k": mix[0],
        "mix_f4": mix[3],
        "graphics": W.gen_render_rom(100),
        "graphics_no_vsync": W.gen_render_rom(7, vsync=False),
    }


def _illegal_rom():
    """Runs three frames of a WSYNC kernel, then executes opcode $02."""
    from fastatari.asm6502 import Asm
    a = Asm(0xF000, 4096)
    a.op("SEI")
    a.op("CLD")
    a.op("LDX", "imm", 0xFF)
    a.op("TXS")
    a.op("LDA", "imm", 0)
    a.op("STA", "zp", 0x80)
    a.label("frame")
    a.op("LDA", "imm", 2)
    a.op("STA", "zp", 0x00)
    a.op("STA", "zp", 0x02)
    a.op("LDA", "imm", 0)
    a.op("STA", "zp", 0x00)
    a.op("LDX", "imm", 200)
    a.label("line")
    a.op("STX", "zp", 0x09)                  # COLUBK = line counter
    a.op("STA", "zp", 0x02)
    a.op("DEX")
    a.op("BNE", "rel", "line")
    a.op("INC", "zp", 0x80)
    a.op("LDA", "zp", 0x80)
    a.op("CMP", "imm", 3)
    a.op("BNE", "rel", "frame")
    a.raw(0x02)
    rom = bytearray(a.link())
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)


def _vsync_held_rom():
    """Sets VSYNC and never clears it: loop A of run_frame does not end."""
    from fastatari.asm6502 import Asm
    a = Asm(0xF000, 4096)
    a.op("LDA", "imm", 2)
    a.op("STA", "zp", 0x00)
    a.label("spin")
    a.op("JMP", "abs", "spin")
    rom = bytearray(a.link())
    rom[0xFFC], rom[0xFFD] = 0x00, 0xF0
    return bytes(rom)


def compare(orc, L, rom, n, chunks, block=64, seed=1):
    from fastatari.render import power_on_state
    h = rh.HostRender(L, rom, n, power_on_state(rom), block)
    r = orc.RBatch(rom, n)
    h.reset()
    r.reset()
    rng = np.random.default_rng(seed)

    def check(tag):
        assert np.array_equal(h.states, r.states()), f"{tag}: states differ"
        assert np.array_equal(h.counters(), r.counters()), f"{tag}: counters differ"
        f1, f2 = h.frames, r.frames()
        if not np.array_equal(f1, f2):
            i = int(np.nonzero((f1 != f2).reshape(n, -1).any(1))[0][0])
            ys, xs = np.nonzero(f1[i] != f2[i])
            pytest.fail(f"{tag}: env {i} frame differs at {len(ys)} pixels, first row {ys[0]} x {xs[0]}")

    check("reset")
    for j, k in enumerate(chunks):
        a = rng.integers(0, 18, size=(k, n), dtype=np.uint8)
        h.step_multi(a, k)
        r.run(a)
        check(f"launch {j} (k={k})")
    return h


@pytest.mark.parametrize("name", list(_roms()))
def test_power_on_state_matches_the_oracle(orc, name):
    from fastatari.render import power_on_state
    rom = _roms()[name]
    assert np.array_equal(power_on_state(rom), orc.RConsole(rom).state())


@pytest.mark.parametrize("name", list(_roms()))
def test_lane_code_matches_the_oracle(orc, host, name):
    rom = _roms()[name]
    n = 32 if name.startswith("config3") else 48
    compare(orc, host, rom, n, [1, 2, 5])


def test_graphics_frames_are_not_trivial(orc, host):
    """The graphics kernel's frames hold many colours and its objects latch
    collisions (so the comparisons exercise the pixel path)."""
    from fastatari import workloads as W
    h = compare(orc, host, W.gen_render_rom(101), 32, [3])        # an odd frame count: latches kept
    assert len(np.unique(h.frames)) > 20
    cx = h.states[:, 160:168]
    assert (cx != 0).any(0).sum() >= 4


@pytest.mark.parametrize("block", [128, 256])
def test_lane_code_block_sizes(orc, host, block):
    from fastatari import workloads as W
    compare(orc, host, W.gen_render_rom(102), 40, [2, 3], block=block)


def test_illegal_opcode_stops_the_env(orc, host):
    h = compare(orc, host, _illegal_rom(), 8, [2, 2, 3])
    assert (h.status == 1).all()
    assert (h.fcyc == 0).all()


def test_vsync_held_is_capped(orc, host):
    h = compare(orc, host, _vsync_held_rom(), 2, [1])
    assert (h.status == 2).all()
