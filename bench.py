"""Benchmark: aggregate frames/s of the batched Atari 2600 frame step on one GPU.

Workload (one "step" = one frame of every env): the seeded 8 KB F8
bankswitched WSYNC ROM (bank switch every 16 scanlines), 65,536 envs,
per-env SplitMix64 action streams, actions resident in HBM before timing.
`value` is the plain per-frame path (fa_step_device, k = 1) over exactly
--steps frames; the batched path (step_multi, K = 16 frames per launch) is
timed separately beside it.  `configs` adds config 1 (the Fixed 4 KB ROM at
1,000 and 65,536 envs) and config 3 (the branchy 16 KB F6 ROM at 65,536 envs,
plain and with opcode grouping).  `valu_peak` is the device's VALU issue
peak measured in the same run (fa_probe_valu_peak); the VALU-issue roofline
divides the kernel's SQ_INSTS_VALU per launch (rocprofv3 --pmc, recorded in
profiles/ for the current kernel sources only) by the live kernel time.  The
CPU baseline is the oracle (the C restatement of the reference emulator),
built -O3 -march=native on the measuring host, at 1, 2, 4, ... threads up to
the CPU share of one GPU, on the headline ROM and on the branchy ROM.

The same line carries the heterogeneous batch ("mix": four ROM classes in
one autotuned hipGraph launch of K = 8 steps with per-env frameskip) and the
served mode ("served": 256 clients with continuous batching; p50 / p99 step
latency beside aggregate frames/s; its bit-exactness, 1 % of envs against the
oracle, is checked in tests/test_gpu_serve.py) and the env adapter
("env_adapter": BatchAtariGpuEnv steps with f32 observations), and the
rendering console ("render": 160 x 192 palette-index frames drawn by the
per-clock TIA, with an HBM roofline on the frame bytes and the rendering
oracle's CPU rate beside it).  "state_io"
is the HBM rate of the frame-boundary state traffic alone (launches that run
zero frames per env: state and counters in, state, counters and obs out).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

# The served mode keeps one stream per ROM class plus an admin stream; the
# HIP runtime maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless
# set) and serialises the kernels of streams that share one.  Set before
# anything initialises HIP.
os.environ["GPU_MAX_HW_QUEUES"] = "8"

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "fast-atari-rs_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per env-frame of the per-frame step: read state 222 +
# action 1 + counters 28, write state 222 + obs 128 + counters 28 (counters:
# cpu cycles, frame cycles, instructions as u64, status as u32).
BYTES_PER_ENV_FRAME = (222 + 1 + 28) + (222 + 128 + 28)
N_ENVS = 65536
# Nominal VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz / 4 clocks per wave64
# VALU op; bench.py measures the device's own figure (fa_probe_valu_peak).
VALU_PEAK_G = 256 * 4 * 2.4 / 4
# HBM bytes and VALU wave-instructions per launch measured with rocprofv3
# --pmc for the default workload (profiles/*.md give the commands).
PROFILE_FIGURES = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


WORKLOADS = {"fixed4k": 1, "f8": 2, "branchy16k": 3}
WORKLOAD_NAMES = {1: "Fixed 4KB WSYNC ROM, per-frame step",
                  2: "F8 8KB WSYNC ROM (bank switch every 16 lines), per-frame step",
                  3: "F6 16KB no-WSYNC branchy ROM (all 151 opcodes), per-frame step"}


def calibrate_copy(mib):
    """Copy kernel of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
    scale the counters read for the frame kernel."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.arange(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "dispatches": 3}), flush=True)


def bench_mix(args, fastatari, W):
    """Heterogeneous launch: four ROM classes, 1K / 40K / 12K / 3K envs, K = 8
    steps of per-env frameskip 1-4 per hipGraph launch, autotuned shapes."""
    seed = W.config_seed(4)
    roms = W.mix_roms(seed)
    counts = list(W.MIX_COUNTS)
    mix = fastatari.BatchMixGpu(roms, counts)
    mix.reset()
    n, k = mix.n, 8
    fs = np.broadcast_to(W.frameskips(seed, n), (k, n)).copy()
    stream = W.ActionStream(seed, n)
    t_tune = time.perf_counter()
    tuned = mix.tune(stream.frames(k), fs)
    t_tune = time.perf_counter() - t_tune
    launches = args.mix_launches
    inputs = [stream.frames(k) for _ in range(launches + 1)]
    mix.write_inputs(inputs[0], fs)
    mix.launch(k, use_fs=True)
    mix.sync()                                     # warm-up (graph instantiation)
    dev_ms = 0.0
    t0 = time.perf_counter()
    for L in range(launches):
        mix.write_inputs(inputs[L + 1], fs)
        mix.launch(k, use_fs=True)
        dev_ms += mix.sync()
    wall = time.perf_counter() - t0
    frames = int(fs.astype(np.int64).sum()) * launches
    errors = int((mix.counters()["status"] != 0).sum())
    names = ["2KB Fixed WSYNC", "8KB F8 WSYNC", "16KB F6 no-WSYNC", "32KB F4 WSYNC"]
    return {
        "workload": "4 ROM classes (1K/40K/12K/3K envs), K=8 steps x per-env frameskip 1-4, one hipGraph launch",
        "fps": round(frames / wall, 1),
        "device_fps": round(frames / (dev_ms * 1e-3), 1),
        "ms_per_launch": round(wall / launches * 1e3, 3),
        "env_frames_per_launch": frames // launches,
        "launches": launches,
        "tune_s": round(t_tune, 2),
        "env_errors": errors,
        "classes": [dict(rom=nm, envs=c, **t) for nm, c, t in zip(names, counts, tuned)],
    }


def _pct(v, q):
    return round(float(np.percentile(v, q)), 3) if len(v) else None


def bench_adapter(fastatari, rom, n, steps=16):
    """The env adapter BatchAtariGpuEnv (src/rl4burn_env.rs:268-380) on the
    same ROM and env count: one broadcast action per step held for the
    default frameskip of 4 frames, f32 observations (RAM / 255) computed on
    the GPU; API-level (obs copied to host memory every step) and with the
    observations left in device memory.  frames/s = envs x 4 x steps / wall."""
    env = fastatari.BatchAtariGpuEnv(rom, n)
    env.reset()
    for s in range(2):
        env.step(s)
    t0 = time.perf_counter()
    for s in range(steps):
        env.step(s % 18)
    host_s = time.perf_counter() - t0
    d_obs = fastatari.DeviceBuffer(n * 128 * 4)
    env.step_device(d_obs.ptr, action=0)
    env.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        env.step_device(d_obs.ptr, action=s % 18)
    env.sync()
    dev_s = time.perf_counter() - t0
    frames = n * 4 * steps
    env.close()
    d_obs.free()
    return {"frameskip": 4, "steps": steps, "obs": "f32 RAM / 255 (n x 128)",
            "api_fps": round(frames / host_s, 1), "device_obs_fps": round(frames / dev_s, 1),
            "ms_per_step_api": round(host_s / steps * 1e3, 3), "ms_per_step_device": round(dev_s / steps * 1e3, 3)}


def bench_serve(args, fastatari, W):
    """Serving: 256 clients on the four mixed ROMs (native load driver).  A
    closed-loop phase finds each client's saturation rate; then Poisson
    arrivals at 0.5-2x that rate; ~1 % env restarts per step.  Reports the
    service latency (submit -> obs in host memory) and end-to-end latency
    (arrival -> obs) percentiles and aggregate frames/s, beside the same env
    population stepped as one plain batch (one frame per step)."""
    from fastatari import loadgen as LG
    roms = W.mix_roms(W.config_seed(4))
    specs = LG.serve_clients(W.config_seed(5))
    counts = [sum(sp.n for sp in specs if sp.cls == c) for c in range(4)]
    n_envs = sum(counts)

    # the same population as one plain batch: every env one frame per step
    mix = fastatari.BatchMixGpu(roms, counts)
    mix.reset()
    acts = W.ActionStream(W.config_seed(5), n_envs)
    mix.write_inputs(acts.next())
    mix.launch(1, use_fs=False)
    mix.sync()
    reps = 8
    t0 = time.perf_counter()
    for _ in range(reps):
        mix.launch(1, use_fs=False)
        mix.sync()
    plain_fps = n_envs * reps / (time.perf_counter() - t0)
    mix.close()

    srv = fastatari.EnvServer(roms, budget_bytes=64 << 30)
    cids = [srv.open(sp.cls, sp.n)[0] for sp in specs]
    steps = np.zeros(len(specs), np.uint64)
    closed = srv.run_load(cids, specs, args.serve_seconds / 2, steps=steps, arrival_seed=1)
    closed_steps = closed["steps"].copy()
    sat = np.maximum(closed_steps, 1) / closed["wall_s"]
    before = steps.copy()
    opened = srv.run_load(cids, specs, args.serve_seconds, sat=sat, steps=steps, arrival_seed=5)
    mem = srv.memory()
    srv.close_server()
    n_by = np.array([sp.n for sp in specs], np.float64)
    cls = np.array([sp.cls for sp in specs])

    def line(res, prev):
        done = res["steps"].astype(np.float64) - prev
        out = {"fps": round(float((done * n_by).sum()) / res["wall_s"], 1), "requests": int(res["requests"]),
               "p50_ms": _pct(res["service_ms"], 50), "p99_ms": _pct(res["service_ms"], 99),
               "e2e_p50_ms": _pct(res["e2e_ms"], 50), "e2e_p99_ms": _pct(res["e2e_ms"], 99), "per_class": {}}
        for c in range(4):
            m = res["cls"] == c
            if m.any():
                out["per_class"][str(c)] = {"p50_ms": _pct(res["service_ms"][m], 50),
                                            "p99_ms": _pct(res["service_ms"][m], 99),
                                            "fps": round(float((done * n_by)[cls == c].sum()) / res["wall_s"], 1)}
        return out

    return {
        "workload": f"256 clients, {n_envs} envs resident (64-2047 per client, log-uniform) on the 4 mixed ROMs; "
                    "Poisson arrivals at 0.5-2x saturation; 1% restarts per step; native load driver",
        "plain_batch_fps": round(plain_fps, 1),
        # run_load returns `steps` itself, which the open-loop phase advanced
        # further: the closed phase's counts are the copy taken after it
        "closed_loop": line(dict(closed, steps=closed_steps), np.zeros(len(specs))),
        "open_loop": line(opened, before.astype(np.float64)),
        "memory": mem,
    }


def kernel_source_hash(prefixes=("fa_emu", "fa_uniform", "fa_frame", "fa_decode", "fa_kernels", "fa_internal"),
                       units=("fa_kernels.hip", "fa_kernels_sorted.hip")):
    """Hash of everything that shapes a kernel's code: the sources whose names
    start with `prefixes` and the compile options of its translation `units`.
    Profile counters recorded for another hash are not used."""
    import hashlib
    from fastatari import build as fb
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "fast-atari-rs_amd", "csrc")
    for name in sorted(os.listdir(csrc)):
        if name.startswith(prefixes):
            with open(os.path.join(csrc, name), "rb") as f:
                h.update(name.encode() + b"\0" + f.read())
    h.update(repr(fb.FLAGS).encode())
    for src, extra in fb.SOURCES:
        if src in units:
            h.update(repr((src, extra)).encode())
    return h.hexdigest()[:16]


def render_source_hash():
    """kernel_source_hash of the rendering console's kernel."""
    return kernel_source_hash(("fa_render", "fa_decode"), ("fa_render.hip",))


def profile_figures(workload, n, source_hash=None):
    """rocprofv3 counter figures of a workload at n envs, if they were
    recorded for the current kernel sources."""
    if not os.path.exists(PROFILE_FIGURES):
        return {}
    with open(PROFILE_FIGURES) as f:
        m = json.load(f).get(workload, {})
    if m.get("envs") != n or m.get("source_hash") != (source_hash or kernel_source_hash()):
        return {}
    return m


def time_plain(gpu, d_act, n, first, frames, d_obs):
    """`frames` per-frame launches (k = 1) on actions resident in HBM: wall
    frames/s and the mean kernel time (HIP events on the handle's stream)."""
    gpu.sync()
    gpu.kernel_time(reset=True)
    t0 = time.perf_counter()
    for f in range(frames):
        gpu.step_device(d_act.at((first + f) * n), 1, d_obs.ptr)
    gpu.sync()
    wall = time.perf_counter() - t0
    ms, launches = gpu.kernel_time(reset=True)
    return n * frames / wall, wall, ms / max(launches, 1)


def bench_config(fastatari, W, cfg, n, warm, frames, sorted_=False):
    """Per-frame step of a config's ROM: frames/s, kernel time per frame step,
    emulated instructions per env-frame."""
    rom = W.rom_for(cfg)
    acts = W.ActionStream(W.config_seed(cfg), n).frames(warm + frames)
    d_act = fastatari.DeviceBuffer(acts.nbytes)
    d_act.upload(acts)
    d_obs = fastatari.DeviceBuffer(n * 128)
    gpu = fastatari.BatchAtariGpu(rom, n)
    gpu.set_sorted(sorted_)
    gpu.reset()
    gpu.set_timing(True)
    for f in range(warm):
        gpu.step_device(d_act.at(f * n), 1, d_obs.ptr)
    ins0 = int(gpu.counters()["instructions"].sum())
    fps, wall, kms = time_plain(gpu, d_act, n, warm, frames, d_obs)
    c = gpu.counters()
    out = {"envs": n, "frames": frames, "sorted": bool(sorted_), "fps": round(fps, 1),
           "ms_per_frame_step": round(wall / frames * 1e3, 4), "kernel_ms_per_frame_step": round(kms, 4),
           "instructions_per_frame": round((int(c["instructions"].sum()) - ins0) / (n * frames), 1),
           "env_errors": int((c["status"] != 0).sum())}
    gpu.close()
    d_act.free()
    d_obs.free()
    return out


def bench_render(fastatari, W, key, name, rom, n, warm, k, launches, valu_peak=None):
    """The rendering console (fa_render_kernel): k frames per launch of n
    envs drawing 160 x 192 palette-index frames, actions resident in HBM.
    HBM roofline on the algorithmic bytes: each env-frame writes its frame
    (30,720 B) and reads its action; each launch reads and writes the
    256-byte render state and the counters."""
    from fastatari.render import FRAME_BYTES, RENDER_STATE_BYTES, BatchAtariRender
    acts = W.ActionStream(W.config_seed(1), n).frames(warm + k * launches)
    d_act = fastatari.DeviceBuffer(acts.nbytes)
    d_act.upload(acts)
    g = BatchAtariRender(rom, n)
    g.reset()
    g.set_timing(True)
    if warm:
        g.step_device(d_act.ptr, warm)
    g.sync()
    g.kernel_time(reset=True)
    ins0 = int(g.counters()["instructions"].sum())
    t0 = time.perf_counter()
    for L in range(launches):
        g.step_device(d_act.at((warm + L * k) * n), k)
    g.sync()
    wall = time.perf_counter() - t0
    ms, nl = g.kernel_time(reset=True)
    kms = ms / max(nl, 1)
    c = g.counters()
    frames = n * k * launches
    alg = n * (k * (FRAME_BYTES + 1) + 2 * (RENDER_STATE_BYTES + 28))
    achieved = alg / (kms * 1e-3) / 1e9
    measured = profile_figures(f"render_{key}", n, render_source_hash()) if key else {}
    if measured.get("k") != k:
        measured = {}
    out = {"workload": name, "envs": n, "k": k, "launches": launches, "fps": round(frames / wall, 1),
           "kernel_ms_per_launch": round(kms, 3), "kernel_ms_per_frame_step": round(kms / k, 3),
           "instructions_per_frame": round((int(c["instructions"].sum()) - ins0) / frames, 1),
           "env_errors": int((c["status"] != 0).sum()), "block": g.block_size(),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 6), "bytes_per_launch": alg,
                        "traffic": measured.get("traffic_bytes"),
                        "note": "algorithmic bytes: frame + action per env-frame, render state + counters in "
                                "and out per launch; traffic: FETCH_SIZE + WRITE_SIZE per launch (rocprofv3, "
                                "profiles/)"},
           "source_hash": render_source_hash()}
    if "valu_wave_instr" in measured and valu_peak:
        valu = measured["valu_wave_instr"] / (kms * 1e-3) / 1e9
        iss = measured["wave_instr"] / (kms * 1e-3) / 1e9
        out["valu_issue"] = {"bound": "valu", "achieved": round(valu, 2), "peak": round(valu_peak, 1),
                             "unit": "G wave64-instr/s", "frac": round(valu / valu_peak, 4),
                             "issue_frac": round(iss / valu_peak, 4),
                             "note": "SQ_INSTS_VALU (issue_frac: all SQ_INSTS_*) per launch, rocprofv3 --pmc for "
                                     "the same kernel sources, / live kernel time, against the measured peak"}
    g.close()
    d_act.free()
    return out


def render_cpu_baseline(W, threads):
    """The rendering oracle (native build when loaded) on the config-1 ROM:
    frames/s at each thread count, each thread 96 envs x 16 frames (about
    0.7 s a point)."""
    from oracle import oracle
    rom = W.rom_for(1)
    pts = []
    for t in threads:
        n = 96 * t
        b = oracle.RBatch(rom, n, nthreads=t)
        b.reset()
        acts = W.ActionStream(W.config_seed(1), n).frames(16)
        secs = b.bench(acts, t)
        pts.append({"threads": t, "fps": round(n * 16 / secs, 1), "seconds": round(secs, 2),
                    "sample": f"{n} envs x 16 frames"})
    return pts


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, os.cpu_count() or 1, avail


def cpu_baseline_curve(W, threads_max, rom_cfgs):
    """The oracle, built -O3 -march=native on this host, timed on the same
    workloads at 1, 2, 4, ... threads (each thread owns a contiguous slice of
    envs; harness shape of examples/benchmark.rs:23-94).  Each point runs a
    fixed per-thread share of envs and frames (about 1-2 s)."""
    import tempfile
    from oracle import oracle
    build = "-O3 -march=native (built on this host)"
    try:
        path = oracle.build_native(os.path.join(tempfile.mkdtemp(prefix="orc_"), "liborc_native.so"))
        oracle.use_library(path)
    except Exception as e:                                   # noqa: BLE001 - fall back to the portable build
        build = f"-O3 portable (native build failed: {type(e).__name__})"
    out = {}
    for name, cfg, envs_per_thread, frames in rom_cfgs:
        rom = W.rom_for(cfg)
        curve = []
        t = 1
        while True:
            n = envs_per_thread * t
            b = oracle.Batch(rom, n, nthreads=t)
            b.reset()
            acts = W.ActionStream(W.config_seed(cfg), n).frames(frames)
            secs = b.bench(acts, t)
            curve.append({"threads": t, "fps": round(n * frames / secs, 1), "seconds": round(secs, 2),
                          "sample": f"{n} envs x {frames} frames"})
            if t >= threads_max:
                break
            t = min(2 * t, threads_max)
        out[name] = curve
    return out, build


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=320, help="timed frame steps of the headline (plain) path")
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--k", type=int, default=16, help="frames per launch of the batched path")
    ap.add_argument("--envs", type=int, default=N_ENVS)
    ap.add_argument("--sorted", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="workgroup size (0: automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="most host threads for the CPU baseline (0: the CPU share of one GPU, at most 16)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="f8",
                    help="f8: the headline 8 KB F8 WSYNC ROM; fixed4k / branchy16k: the other ROM classes")
    ap.add_argument("--no-configs", action="store_true", help="skip the config 1 / config 3 lines")
    ap.add_argument("--no-mix", action="store_true", help="skip the heterogeneous-batch line")
    ap.add_argument("--mix-launches", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving line")
    ap.add_argument("--no-adapter", action="store_true", help="skip the env-adapter line")
    ap.add_argument("--no-render", action="store_true", help="skip the rendering-console line")
    ap.add_argument("--serve-seconds", type=float, default=6.0, help="open-loop serving time")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(envs are independent; sharding is a per-device split of the env range).", file=sys.stderr)
        sys.exit(2)

    import fastatari
    from fastatari import env as fenv
    from fastatari import workloads as W

    cfg = WORKLOADS[args.workload]
    seed = W.config_seed(cfg)
    rom = W.rom_for(cfg)
    n = args.envs
    steps, warm, k = max(1, args.steps), args.warmup, args.k
    multi_launches = max(4, steps // k)

    log(f"bench: {WORKLOAD_NAMES[cfg]}, {n} envs, warmup {warm}, timed {steps} frames (plain) + "
        f"{multi_launches} launches of K={k}")
    acts = W.ActionStream(seed, n).frames(warm + steps + multi_launches * k)      # [frames, n]
    d_act = fastatari.DeviceBuffer(acts.nbytes)
    d_act.upload(acts)
    d_obs = fastatari.DeviceBuffer(n * 128)

    gpu = fastatari.BatchAtariGpu(rom, n)
    gpu.set_sorted(args.sorted)
    if args.block:
        gpu.set_block_size(args.block)
    gpu.reset()
    gpu.set_timing(True)

    # ---- plain per-frame path (the drop-in for atari_frame_kernel): exactly `steps` frames
    for f in range(warm):
        gpu.step_device(d_act.at(f * n), 1, d_obs.ptr)
    plain_fps, plain_wall, kernel_ms = time_plain(gpu, d_act, n, warm, steps, d_obs)

    # ---- batched path: K frames per launch, state kept on-chip across frames
    t0 = time.perf_counter()
    for L in range(multi_launches):
        gpu.step_device(d_act.at((warm + steps + L * k) * n), k, d_obs.ptr)
    gpu.sync()
    multi_wall = time.perf_counter() - t0
    multi_ms_total, multi_n = gpu.kernel_time(reset=True)
    multi_fps = n * k * multi_launches / multi_wall

    counters = gpu.counters()
    instr_total = int(counters["instructions"].sum())
    frames_total = n * (1 + warm + steps + multi_launches * k)
    status_bad = int((counters["status"] != 0).sum())

    # ---- API-level rate (host actions in, host obs out each frame, as the reference measures)
    api_frames = 16
    t0 = time.perf_counter()
    for f in range(api_frames):
        gpu.step(acts[f])
    api_fps = n * api_frames / (time.perf_counter() - t0)
    gpu.kernel_time(reset=True)

    # ---- the frame-boundary state traffic alone: zero frames per env, so a
    # launch stages state and counters in and writes state, counters and obs
    # back without emulating
    d_fs0 = fastatari.DeviceBuffer(n)
    d_fs0.upload(np.zeros(n, np.uint8))
    for _ in range(4):
        gpu.step_frameskip_device(d_act.ptr, d_fs0.ptr, d_obs.ptr)
    gpu.sync()
    gpu.kernel_time(reset=True)
    io_launches = 32
    for _ in range(io_launches):
        gpu.step_frameskip_device(d_act.ptr, d_fs0.ptr, d_obs.ptr)
    gpu.sync()
    io_ms_total, io_n = gpu.kernel_time(reset=True)
    io_ms = io_ms_total / max(io_n, 1)
    io_gbs = (BYTES_PER_ENV_FRAME + 1) * n / (io_ms * 1e-3) / 1e9     # + the frameskip byte
    d_fs0.free()
    gpu.close()
    d_act.free()
    d_obs.free()

    # ---- the VALU issue peak, measured (independent v_add_u32 chains, 8 waves per SIMD)
    peak = fenv.probe_valu_peak()
    valu_peak = peak["g_wave_instr_per_s"]

    achieved_gbs = BYTES_PER_ENV_FRAME * n / (kernel_ms * 1e-3) / 1e9
    measured = profile_figures(args.workload, n)
    out = {
        "metric": "aggregate frames/sec at 64K envs (plain vs batched); p50/p99 served step latency",
        "value": round(plain_fps, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(plain_wall / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": f"synthetic (seeded {len(rom) // 1024} KB ROM, SplitMix64 per-env actions)",
        "config": {"workload": WORKLOAD_NAMES[cfg],
                   "envs": n, "frames": steps, "sorted": bool(args.sorted), "block": args.block or "auto"},
        "batched": {"k": k, "launches": multi_launches, "frames": multi_launches * k, "fps": round(multi_fps, 1),
                    "ms_per_launch": round(multi_wall / multi_launches * 1e3, 4),
                    "kernel_ms_per_launch": round(multi_ms_total / max(multi_n, 1), 4)},
        "api_level_fps": round(api_fps, 1),
        "kernel_ms_per_frame_step": round(kernel_ms, 4),
        "emulated_instr_per_s": round(instr_total / frames_total * plain_fps, 1),
        "instructions_per_frame": round(instr_total / frames_total, 1),
        "env_errors": status_bad,
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 6), "traffic": measured.get("traffic_bytes")},
        "valu_peak": {"measured": round(valu_peak, 1), "nominal": VALU_PEAK_G, "unit": "G wave64-instr/s",
                      "ms_per_launch": round(peak["ms_per_launch"], 3),
                      "note": "fa_probe_valu_peak: independent v_add_u32 chains, 8 waves per SIMD"},
        "state_io": {"achieved": round(io_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(io_gbs / HBM_PEAK_GBS, 4), "ms_per_launch": round(io_ms, 4),
                     "bytes_per_launch": (BYTES_PER_ENV_FRAME + 1) * n,
                     "note": "zero-frame launches: state + counters in, state + counters + obs out"},
        "kernel_source_hash": kernel_source_hash(),
    }
    if "valu_wave_instr" in measured:
        valu = measured["valu_wave_instr"] / (kernel_ms * 1e-3) / 1e9
        out["valu_issue"] = {"bound": "valu", "achieved": round(valu, 2), "peak": round(valu_peak, 1),
                             "unit": "G wave64-instr/s", "frac": round(valu / valu_peak, 4),
                             "note": "SQ_INSTS_VALU per launch (rocprofv3 --pmc, profiles/, same kernel sources) "
                                     "/ live kernel time, against the measured peak"}
    if "wave_instr" in measured:
        # every wave-instruction (VALU, SALU, branch, LDS, SMEM) against one
        # issue per SIMD every 4 clocks: the bound of a SIMD's single wave
        iss = measured["wave_instr"] / (kernel_ms * 1e-3) / 1e9
        out["issue"] = {"achieved": round(iss, 2), "peak": round(valu_peak, 1), "unit": "G wave-instr/s",
                        "frac": round(iss / valu_peak, 4),
                        "note": "SQ_INSTS_{VALU,SALU,BRANCH,LDS,SMEM} per launch (rocprofv3, profiles/) / live "
                                "kernel time"}
    if not measured:
        out["valu_issue"] = None
        out["valu_issue_note"] = "no rocprofv3 counters recorded for these kernel sources (hash differs)"

    if not args.no_configs:
        log("bench: config 1 (Fixed 4 KB, 1,000 envs and 65,536 envs) and config 3 (branchy F6, plain / sorted)")
        c1 = bench_config(fastatari, W, 1, 1000, 8, 64)
        c1_64k = bench_config(fastatari, W, 1, 65536, 8, 32)
        # config 3 as defined: 65,536 envs x 300 frames from the reset state,
        # every frame timed (divergence grows over the first frames)
        c3 = bench_config(fastatari, W, 3, 65536, 0, 300)
        c3s = bench_config(fastatari, W, 3, 65536, 0, 300, sorted_=True)
        c3_fig = profile_figures("branchy16k", 65536)
        occ = {}
        if c3_fig:
            for mode in ("plain", "sorted"):
                f = c3_fig.get(mode, {})
                if "exec_occupancy" in f:
                    occ[mode] = f["exec_occupancy"]
        out["configs"] = {
            "1": {"workload": WORKLOAD_NAMES[1], "1000_envs": c1, "65536_envs": c1_64k},
            "3": {"workload": WORKLOAD_NAMES[3], "plain": c3, "sorted": c3s,
                  "sorted_vs_plain": round(c3s["fps"] / c3["fps"], 3),
                  "exec_mask_occupancy": occ or None},
        }
    if not args.no_render:
        log("bench: rendering console (pixel TIA frames, 65,536 and 1,000 envs)")
        c1name = "config-1 ROM (seeded Fixed 4 KB WSYNC kernel)"
        gname = "graphics kernel (seeded 4 KB Fixed, sprites / playfield / HMOVE / mid-line RESPx)"
        out["render"] = {
            "config1_65536": bench_render(fastatari, W, "config1_65536", c1name, W.rom_for(1), 65536, 1, 4, 2,
                                          valu_peak),
            "graphics_65536": bench_render(fastatari, W, "graphics_65536", gname, W.gen_render_rom(100), 65536,
                                           1, 4, 2, valu_peak),
            "config1_1000": bench_render(fastatari, W, "config1_1000", c1name, W.rom_for(1), 1000, 1, 16, 2,
                                         valu_peak),
        }
    if not args.no_adapter:
        log("bench: env adapter (broadcast action, frameskip 4, f32 obs)")
        out["env_adapter"] = bench_adapter(fastatari, rom, n)
    if not args.no_mix:
        log("bench: heterogeneous batch (4 ROM classes, K=8, autotuned)")
        out["mix"] = bench_mix(args, fastatari, W)
    if not args.no_serve:
        log("bench: serving (256 clients, continuous batching)")
        out["served"] = bench_serve(args, fastatari, W)
        ol = out["served"]["open_loop"]
        out["served_p50_ms"], out["served_p99_ms"] = ol["p50_ms"], ol["p99_ms"]
    if not args.no_cpu:
        model, nproc, avail = host_info()
        # the CPU share of one GPU on a shared host: at most 16 threads
        tmax = args.cpu_threads or min(16, avail)
        log(f"bench: CPU baseline (oracle, 1..{tmax} threads of {nproc})")
        curves, build = cpu_baseline_curve(W, tmax, [("f8", cfg if args.workload == "f8" else 2, 256, 96),
                                                     ("branchy16k", 3, 64, 16)])
        head = curves["f8"]
        out["cpu_baseline"] = {"value": head[-1]["fps"], "unit": "frames/s", "cores": head[-1]["threads"],
                               "kind": "port", "sample": head[-1]["sample"] + " of the headline ROM/actions",
                               "single_thread_fps": head[0]["fps"], "build": build,
                               "host": {"cpu_model": model, "nproc": nproc, "affinity": avail,
                                        "threads_used": head[-1]["threads"]},
                               "scaling": curves["f8"],
                               "branchy16k": {"value": curves["branchy16k"][-1]["fps"],
                                              "single_thread_fps": curves["branchy16k"][0]["fps"],
                                              "scaling": curves["branchy16k"]}}
        if not args.no_render:
            rc = render_cpu_baseline(W, sorted({1, tmax}))
            out["cpu_baseline"]["render_config1"] = {"value": rc[-1]["fps"], "single_thread_fps": rc[0]["fps"],
                                                     "kind": "port", "scaling": rc}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
