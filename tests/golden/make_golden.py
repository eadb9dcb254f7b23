"""Writes tests/golden/golden.json: committed fixtures for the seeded workloads.

For every ROM class it records the SHA-256 of the generated ROM (pins the
generator) and, for 8 envs stepped with their SplitMix64 action streams, the
SHA-256 of the oracle's packed states + counters after reset and after each
of the first frames.  The oracle is the C restatement of the reference CPU
emulator; the GPU tests compare the HIP path against these hashes too.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "fast-atari-rs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

N_ENVS = 8
FRAMES = 6


def cases():
    from fastatari import workloads as W
    return {
        "config1_fixed4k": (W.rom_for(1), W.config_seed(1)),
        "config2_f8": (W.rom_for(2), W.config_seed(2)),
        "config3_f6_branchy": (W.rom_for(3), W.config_seed(3)),
        "mix_fixed2k": (W.rom_for(4, 0), W.config_seed(4)),
        "mix_f4": (W.rom_for(4, 3), W.config_seed(4) + 3),
    }


def digest(states, counters) -> str:
    h = hashlib.sha256()
    h.update(states.tobytes())
    h.update(counters.astype("<u8").tobytes())
    return h.hexdigest()


def oracle_trace(rom, seed, n=N_ENVS, frames=FRAMES):
    from oracle import oracle
    from fastatari import workloads as W
    b = oracle.Batch(rom, n, nthreads=1)
    b.reset()
    out = [digest(b.states(), b.counters())]
    acts = W.ActionStream(seed, n)
    for _ in range(frames):
        b.run(acts.next()[None, :])
        out.append(digest(b.states(), b.counters()))
    return out


def main():
    data = {"n_envs": N_ENVS, "frames": FRAMES, "cases": {}}
    for name, (rom, seed) in cases().items():
        data["cases"][name] = {
            "rom_sha256": hashlib.sha256(rom).hexdigest(),
            "rom_bytes": len(rom),
            "seed": seed,
            "state_sha256": oracle_trace(rom, seed),
        }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
