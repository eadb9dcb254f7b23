"""Helpers shared by the GPU parity tests: run the HIP path and the oracle on
the same seeded inputs and compare everything bit-exactly."""
import numpy as np

FIELD_NAMES = None


def describe_state_diff(g: np.ndarray, r: np.ndarray, env_ids=None) -> str:
    from fastatari import STATE_DTYPE
    bad = np.nonzero((g != r).any(axis=1))[0]
    if len(bad) == 0:
        return "equal"
    i = bad[0]
    gs = g[i].view(STATE_DTYPE)[0]
    rs = r[i].view(STATE_DTYPE)[0]
    diffs = []
    for name in STATE_DTYPE.names:
        if not np.array_equal(gs[name], rs[name]):
            gv, rv = gs[name], rs[name]
            if name == "ram":
                k = int(np.nonzero(gv != rv)[0][0])
                diffs.append(f"ram[{k:#x}] gpu={int(gv[k]):#x} ref={int(rv[k]):#x}")
            else:
                diffs.append(f"{name} gpu={int(gv):#x} ref={int(rv):#x}")
    env = env_ids[i] if env_ids is not None else i
    return f"{len(bad)} envs differ; env {env}: " + ", ".join(diffs[:8])


def assert_same(gpu, ref, where: str, env_ids=None, gpu_rows=None):
    """Compare full packed state, cycle counters, instruction counts and
    status of the GPU handle (rows `gpu_rows`, default all) with the oracle."""
    g = gpu.download_states_bytes()
    c = gpu.counters()
    if gpu_rows is not None:
        g = g[gpu_rows]
        c = {k: v[gpu_rows] for k, v in c.items()}
    r = ref.states()
    rc = ref.counters()
    assert np.array_equal(g, r), f"{where}: state mismatch: {describe_state_diff(g, r, env_ids)}"
    for col, key in ((0, "cpu_cycles"), (1, "frame_cycles"), (2, "instructions"), (3, "status")):
        a = c[key].astype(np.uint64)
        b = rc[:, col]
        if not np.array_equal(a, b):
            i = int(np.nonzero(a != b)[0][0])
            raise AssertionError(f"{where}: {key} mismatch at env {i}: gpu={int(a[i])} ref={int(b[i])}")
