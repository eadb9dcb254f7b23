"""Serving workload: 256 clients on the four mixed ROMs, each owning
floor(64 * 32**u) envs (u ~ U[0, 1)), stepping with Poisson arrivals at
r x its saturation rate (r ~ U[0.5, 2]) and restarting ~1 % of envs per
step.  run_load drives an EnvServer from Python (tests at small scale);
EnvServer.run_load is the native driver with the same streams, used at full
scale.  Both record per-request service latency (submit -> obs in host
memory, measured natively) and end-to-end latency (arrival -> obs, including
the client's own backlog).

Every client's results depend only on its own request sequence (actions,
restarts, frames), which is drawn from the client's seed, so a client run
alone and the same client run in the mix see identical results.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .workloads import GOLDEN, MASK64, ActionStream, Rng, _mulhi, splitmix64


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


RESTART_P = 0.01


@dataclass
class ClientSpec:
    cls: int            # ROM class
    n: int              # envs
    rate_mult: float    # arrival rate / saturation rate
    seed: int


def serve_clients(seed: int, n_clients: int = 256, n_classes: int = 4, lo: int = 64, ratio: float = 32.0):
    rng = Rng(seed)
    out = []
    for c in range(n_clients):
        cls = rng.below(n_classes)
        u = rng.u64() / 2.0 ** 64
        n = int(lo * ratio ** u)
        r = 0.5 + 1.5 * (rng.u64() / 2.0 ** 64)
        out.append(ClientSpec(cls, n, r, splitmix64((seed ^ ((c + 1) * GOLDEN)) & MASK64)))
    return out


class ClientStream:
    """The request sequence of one client: per step, actions (n bytes) and a
    restart mask (each env with probability RESTART_P)."""

    def __init__(self, spec: ClientSpec, frames: int = 1, record=None):
        self.spec = spec
        self.frames = frames
        self.actions = ActionStream(spec.seed, spec.n)
        self.restarts = ActionStream(spec.seed ^ 0x5E5E7000, spec.n)
        self.threshold = np.uint64(int(RESTART_P * 2 ** 64))
        self.record = record        # env indices whose inputs are kept
        self.history = []

    def inputs_at(self, steps: int, envs):
        """Actions and restart masks of steps 0..steps-1 for the given envs
        ([steps, len(envs)] each), by jumping the SplitMix64 streams ahead:
        step s of env i draws mix(state_i + (s + 1) * GOLDEN)."""
        envs = np.asarray(envs, dtype=np.int64)
        fresh_a = ActionStream(self.spec.seed, self.spec.n)
        fresh_r = ActionStream(self.spec.seed ^ 0x5E5E7000, self.spec.n)
        k = np.arange(1, steps + 1, dtype=np.uint64)[:, None]
        with np.errstate(over="ignore"):
            za = _mix(fresh_a.state[envs][None, :] + k * np.uint64(GOLDEN))
            zr = _mix(fresh_r.state[envs][None, :] + k * np.uint64(GOLDEN))
        return _mulhi(za, 18).astype(np.uint8), (zr < self.threshold).astype(np.uint8)

    def next(self):
        a = self.actions.next()
        r = (self.restarts.next_u64() < self.threshold).astype(np.uint8)
        if self.record is not None:
            self.history.append((a[self.record].copy(), r[self.record].copy()))
        return a, r


@dataclass
class LoadResult:
    service_ms: list = field(default_factory=list)
    e2e_ms: list = field(default_factory=list)
    cls: list = field(default_factory=list)
    env_frames: int = 0
    requests: int = 0
    wall_s: float = 0.0
    steps: dict = field(default_factory=dict)      # client index -> completed steps


def run_load(srv, specs, cids, streams, duration_s: float, sat_rate=None, seed: int = 1, max_steps=None):
    """Drives the server for duration_s.  sat_rate None: closed loop (every
    client resubmits as soon as its request completes); else per-client
    saturation rates (steps / s) and Poisson arrivals at rate_mult x that.
    max_steps caps the steps of every client (then the run may end early)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nc = len(specs)
    t0 = time.perf_counter()
    next_arr = np.full(nc, t0)
    if sat_rate is not None:
        lam = np.array([s.rate_mult for s in specs]) * np.asarray(sat_rate, dtype=np.float64)
        next_arr = t0 + rng.exponential(1.0 / lam)
    backlog = [[] for _ in range(nc)]     # arrival times not yet submitted
    inflight = {}                         # client index -> arrival time
    res = LoadResult(steps={i: 0 for i in range(nc)})
    submitted = np.zeros(nc, np.int64)
    t_end = t0 + duration_s
    while True:
        now = time.perf_counter()
        open_loop = sat_rate is not None
        if now < t_end:
            if open_loop:
                due = np.nonzero(next_arr <= now)[0]
                for i in due:
                    backlog[i].append(next_arr[i])
                    next_arr[i] += rng.exponential(1.0 / lam[i])
            else:
                for i in range(nc):
                    if i not in inflight and not backlog[i]:
                        backlog[i].append(now)
        elif not inflight:
            break
        acted = False
        for i in range(nc):
            if i in inflight or not backlog[i] or now >= t_end:
                continue
            if max_steps is not None and submitted[i] >= max_steps:
                backlog[i].clear()
                continue
            a, r = streams[i].next()
            srv.submit(cids[i], a, r, streams[i].frames)
            inflight[i] = backlog[i].pop(0)
            submitted[i] += 1
            acted = True
        for i in list(inflight):
            if srv.poll(cids[i]):
                _, ms = srv.wait(cids[i], want_obs=False)
                done = time.perf_counter()
                res.service_ms.append(ms)
                res.e2e_ms.append((done - inflight.pop(i)) * 1e3)
                res.cls.append(specs[i].cls)
                res.env_frames += specs[i].n * streams[i].frames
                res.requests += 1
                res.steps[i] += 1
                acted = True
        if max_steps is not None and not inflight and all(submitted[i] >= max_steps for i in range(nc)):
            break
        if not acted:
            time.sleep(0.0001)
    res.wall_s = time.perf_counter() - t0
    return res


def summarize(res: LoadResult, n_classes: int = 4) -> dict:
    def pct(v, q):
        return float(np.percentile(v, q)) if len(v) else None
    svc = np.asarray(res.service_ms)
    e2e = np.asarray(res.e2e_ms)
    cls = np.asarray(res.cls)
    out = {
        "requests": res.requests,
        "env_frames": res.env_frames,
        "wall_s": round(res.wall_s, 3),
        "fps": res.env_frames / res.wall_s if res.wall_s > 0 else 0.0,
        "p50_ms": pct(svc, 50), "p99_ms": pct(svc, 99),
        "e2e_p50_ms": pct(e2e, 50), "e2e_p99_ms": pct(e2e, 99),
        "per_class": {},
    }
    for c in range(n_classes):
        m = cls == c
        if m.any():
            out["per_class"][c] = {"requests": int(m.sum()), "p50_ms": pct(svc[m], 50), "p99_ms": pct(svc[m], 99)}
    return out
