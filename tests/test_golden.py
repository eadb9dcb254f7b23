"""Committed fixtures (tests/golden/golden.json): the seeded ROM generators
are deterministic, and the oracle reproduces its recorded per-frame state
hashes.  The GPU variant checks the HIP path against the same hashes."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

with open(os.path.join(HERE, "golden", "golden.json")) as f:
    GOLD = json.load(f)

CASES = make_golden.cases()


@pytest.mark.parametrize("name", sorted(GOLD["cases"]))
def test_generator_is_deterministic(name):
    import hashlib
    rom, seed = CASES[name]
    assert hashlib.sha256(rom).hexdigest() == GOLD["cases"][name]["rom_sha256"]
    assert seed == GOLD["cases"][name]["seed"]


@pytest.mark.parametrize("name", sorted(GOLD["cases"]))
def test_oracle_matches_golden(orc, name):
    rom, seed = CASES[name]
    assert make_golden.oracle_trace(rom, seed) == GOLD["cases"][name]["state_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLD["cases"]))
def test_gpu_matches_golden(name):
    import fastatari
    from fastatari import workloads as W
    rom, seed = CASES[name]
    n = GOLD["n_envs"]
    gpu = fastatari.BatchAtariGpu(rom, n)
    gpu.reset()

    def dig():
        c = gpu.counters()
        cnt = np.stack([c["cpu_cycles"], c["frame_cycles"], c["instructions"],
                        c["status"].astype(np.uint64)], axis=1)
        return make_golden.digest(gpu.download_states_bytes(), cnt)

    got = [dig()]
    acts = W.ActionStream(seed, n)
    for _ in range(GOLD["frames"]):
        gpu.step(acts.next())
        got.append(dig())
    want = GOLD["cases"][name]["state_sha256"]
    bad = [i for i, (g, w) in enumerate(zip(got, want)) if g != w]
    assert not bad, f"{name}: first differing frame {bad[0]} (0 = after reset)"
