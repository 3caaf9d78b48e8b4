"""Benchmark of the FTL hair-strand step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload single|batched|skinned|served|all]

Prints ONE JSON line.  `value` is control-point updates per second of the
single large groom stepped by the plain kernel (262,144 strands x 32 points,
wind, rotating emitter, two colliders, dt = 1/60), inputs resident in HBM.
The other workloads are reported in extra keys of the same line.

Timing: W untimed steps, then exactly K steps bracketed by device
synchronisation; HIP events recorded on the stream the kernels run on give
the per-launch duration used by the roofline figure.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
BYTES_PER_CP = 64      # read + write of float4 position and float4 velocity
DT = np.float32(1.0 / 60.0)
METRIC = "FTL control-point updates/s & % HBM peak; p99 step latency, batched vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def c2_frame(t):
    """Per-step inputs of the single-groom workload at time t: emitter rotation
    about y at 1 rad/s, sphere r=1 at the origin and r=0.3 orbiting the y axis
    at radius 1.2, 1 rad/s."""
    c, s = np.cos(t), np.sin(t)
    m = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)
    sph = np.array([[0, 0, 0, 1.0], [1.2 * np.cos(t), 0.0, 1.2 * np.sin(t), 0.3]], np.float32)
    return m, sph


def c2_params(ba):
    return ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0, iterations=8,
                          wind_amplitude=2.0, wind_frequency=2.0, wind_scroll=(0.5, 0.0, 0.25), wind_seed=2)


def bench_single(ba, hipm, steps, warmup, S=262144, N=32):
    pos, vel, roots = ba.sphere_groom(S, N, radius=1.0, length=0.2, jitter=0.05, seed=2024)
    ctx = ba.Context(0)
    g = ba.Groom(ctx, pos, vel, rest_roots=roots)
    g.set_params(c2_params(ba))
    g.set_root_mode(ba.FTL_ROOT_REST)
    stream = hipm.Stream()
    e0, e1 = hipm.Event(), hipm.Event()

    def step(k):
        t = float(k + 1) * float(DT)
        m, sph = c2_frame(t)
        g.set_root_transform(m)
        g.set_colliders(sph)
        g.set_time(t)
        g.step(DT, stream.h)

    for k in range(warmup):
        step(k)
    hipm.device_synchronize()
    w0 = time.perf_counter()
    e0.record(stream.h)
    for k in range(warmup, warmup + steps):
        step(k)
    e1.record(stream.h)
    hipm.device_synchronize()
    wall = time.perf_counter() - w0
    gpu_ms = e0.elapsed_ms(e1)
    p, v = g.download()
    assert np.isfinite(p).all() and np.isfinite(v).all(), "non-finite state after the benchmark"
    lanes = ctx.lanes_for(N)
    g.close()
    ctx.close()
    return dict(S=S, N=N, wall_s=wall, gpu_ms=gpu_ms, lanes=lanes)


def cpu_baseline(ba, seconds_target=12.0):
    """The FP32 oracle (a plain C++ transliteration of the reference shader)
    on the host cores, on a bounded sample of the single-groom workload."""
    from oracle import ftl_oracle as orc
    from tests.helpers import oracle_params_for

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    S, N = 32768, 32  # 1/8 of the strands of the workload
    pos, vel, roots = ba.sphere_groom(S, N, radius=1.0, length=0.2, jitter=0.05, seed=2024)
    params = c2_params(ba)
    wind = ba.wind_table(params)
    done, k, t0 = 0, 0, time.perf_counter()
    while True:
        t = float(k + 1) * float(DT)
        m, sph = c2_frame(t)
        op = oracle_params_for(params, N, DT, spheres=[tuple(x) for x in sph], model=m, root_mode=1, time=t,
                               wind=wind)
        orc.step_f32(op, pos, vel, roots, nthreads=threads)
        done += S * N
        k += 1
        el = time.perf_counter() - t0
        if el >= seconds_target or k >= 2000:
            break
    return dict(value=done / el, unit="CP updates/s", cores=threads, kind="port",
                sample="%d strands x %d points (1/8 of the single-groom workload), %d steps, %.1f s, "
                       "OpenMP over strands" % (S, N, k, el))


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (strands and grooms shard with no "
              "exchange; see the design notes).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import barbu_amd as ba
    from barbu_amd import _hip as hipm

    log("single groom: 262144 x 32, %d warmup + %d timed steps" % (args.warmup, args.steps))
    r = bench_single(ba, hipm, args.steps, args.warmup)
    cps = r["S"] * r["N"]
    ms_per_step = 1000.0 * r["wall_s"] / args.steps
    value = cps * args.steps / r["wall_s"]
    launch_s = r["gpu_ms"] / 1000.0 / args.steps
    achieved = cps * BYTES_PER_CP / launch_s / 1e9
    traffic = load_traffic("single")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "CP updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single groom, plain kernel: 262,144 strands x 32 points, wind (value noise, "
                               "2 m/s^2), emitter rotating 1 rad/s about y, spheres r=1 at 0 and r=0.3 orbiting "
                               "at 1.2, dt=1/60, 8 iterations, damping 0.9",
                   "strands": r["S"], "points": r["N"], "lanes_per_strand": r["lanes"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_us": launch_s * 1e6, "bytes_per_launch": cps * BYTES_PER_CP},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)")
        out["cpu_baseline"] = cpu_baseline(ba, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
