"""Benchmark of the FTL hair-strand step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload single|batched|skinned|served|all]

Prints ONE JSON line.  `value` is control-point updates per second of the
single large groom stepped by the plain kernel (262,144 strands x 32 points,
wind, rotating emitter, two colliders, dt = 1/60), inputs resident in HBM.
The other workloads are reported in extra keys of the same line.

Timing: W untimed steps, then exactly K steps bracketed by device
synchronisation.  `roofline.kernel_us` is the mean duration of the step
kernel itself: each timed launch is bracketed by two HIP events that take the
kernel's own start and end timestamps (ftl_groom_step_timed), on the stream the
kernel runs on.  The window of the whole timed region on that stream is
reported beside it (`stream_window_us`).
"""
import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from barbu_amd import workloads as wl  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
BYTES_PER_CP = 64      # read + write of float4 position and float4 velocity
DT = wl.DT
METRIC = "FTL control-point updates/s & % HBM peak; p99 step latency, batched vs plain"
TUNED = os.path.join(ROOT, "barbu_amd", "tuned_gfx950.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def context(ba):
    """A context with the launch-shape tables measured on MI355X."""
    ctx = ba.Context(0)
    ctx.autotune(TUNED)  # loads the tables (no timing) when they match the device
    return ctx


def set_tuning(path):
    """Use another launch-shape table (a JSON written by ftl_autotune)."""
    global TUNED
    TUNED = os.path.abspath(path)


def bench_single(ba, hipm, steps, warmup, S=262144, N=32, lanes=None):
    pos, vel, roots = ba.sphere_groom(S, N, radius=1.0, length=0.2, jitter=0.05, seed=2024)
    ctx = context(ba)
    g = ba.Groom(ctx, pos, vel, rest_roots=roots)
    g.set_params(wl.c2_params(ba))
    g.set_root_mode(ba.FTL_ROOT_REST)
    if lanes:
        g.set_lanes(lanes)
    del pos, vel, roots
    stream = hipm.Stream()
    e0, e1 = hipm.Event(), hipm.Event()
    # per-step inputs, evaluated before the timed region
    total = warmup + steps
    times = [float(k + 1) * float(DT) for k in range(total)]
    cms, sphs = [], []
    for t in times:
        m, sph = wl.c2_frame(t)
        cms.append(np.ascontiguousarray(m.T))
        sphs.append(np.ascontiguousarray(sph))
    ks = [(hipm.Event(), hipm.Event()) for _ in range(steps)]
    L = ba.lib()
    h = g.handle
    dt = float(DT)

    def inputs(k):
        ba.check(L.ftl_groom_set_root_transform(h, cms[k].ctypes.data), "ftl_groom_set_root_transform")
        ba.check(L.ftl_groom_set_colliders(h, sphs[k].ctypes.data, 2), "ftl_groom_set_colliders")
        ba.check(L.ftl_groom_set_time(h, times[k]), "ftl_groom_set_time")

    for k in range(warmup):
        inputs(k)
        ba.check(L.ftl_groom_step(h, dt, stream.h), "ftl_groom_step")
    hipm.device_synchronize()
    w0 = time.perf_counter()
    e0.record(stream.h)
    for i, k in enumerate(range(warmup, total)):
        inputs(k)
        ba.check(L.ftl_groom_step_timed(h, dt, stream.h, ks[i][0].h, ks[i][1].h), "ftl_groom_step_timed")
    e1.record(stream.h)
    hipm.device_synchronize()
    wall = time.perf_counter() - w0
    gpu_ms = e0.elapsed_ms(e1)
    kern_ms = np.array([a.elapsed_ms(b) for a, b in ks])
    p, v = g.download()
    assert np.isfinite(p).all() and np.isfinite(v).all(), "non-finite state after the benchmark"
    lanes = lanes or ctx.lanes_for(N, wind=True)
    g.close()
    copy_gbs = copy_reference(ba, hipm, ctx, S * N * 32)
    ctx.close()
    return dict(S=S, N=N, wall_s=wall, gpu_ms=gpu_ms, lanes=lanes, copy_gbs=copy_gbs,
                kernel_us_mean=float(kern_ms.mean() * 1000.0), kernel_us_median=float(np.median(kern_ms) * 1000.0),
                kernel_us_min=float(kern_ms.min() * 1000.0), kernel_us_max=float(kern_ms.max() * 1000.0))


def copy_reference(ba, hipm, ctx, nbytes, reps=20):
    """GB/s of a plain streaming copy of nbytes (the step's read + write
    volume is 2 x nbytes): the practical bandwidth ceiling on this device."""
    L = ba.lib()
    a, b = C.c_void_p(), C.c_void_p()
    ba.check(L.ftl_diag_alloc(ctx.handle, nbytes, C.byref(a)), "ftl_diag_alloc")
    ba.check(L.ftl_diag_alloc(ctx.handle, nbytes, C.byref(b)), "ftl_diag_alloc")
    stream = hipm.Stream()
    e0, e1 = hipm.Event(), hipm.Event()
    for _ in range(3):
        ba.check(L.ftl_diag_copy(ctx.handle, a, b, nbytes, stream.h), "ftl_diag_copy")
    e0.record(stream.h)
    for _ in range(reps):
        ba.check(L.ftl_diag_copy(ctx.handle, a, b, nbytes, stream.h), "ftl_diag_copy")
    e1.record(stream.h)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    L.ftl_diag_free(ctx.handle, a, nbytes)
    L.ftl_diag_free(ctx.handle, b, nbytes)
    return 2.0 * nbytes / (ms / 1000.0) / 1e9


def host_info():
    """The host the CPU baseline ran on: CPU model, nproc, the cores this
    process may run on, and the thread count the environment grants."""
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc", "--all"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        nproc = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return dict(cpu_model=model, nproc=nproc, affinity_cores=affinity, omp_num_threads=omp or None)


def cpu_threads(info):
    """Every core the process may run on, within the thread count the
    environment grants (OMP_NUM_THREADS: a GPU box's host share)."""
    n = info["affinity_cores"]
    if info["omp_num_threads"]:
        try:
            n = min(n, int(info["omp_num_threads"]))
        except ValueError:
            pass
    return max(1, n)


def cpu_baseline(ba, seconds_target=12.0):
    """The FP32 oracle (a plain C++ transliteration of the reference shader)
    on the host cores, on a bounded sample of the single-groom workload."""
    from oracle import ftl_oracle as orc
    from tests.helpers import oracle_params_for

    info = host_info()
    threads = cpu_threads(info)
    S, N = 32768, 32  # 1/8 of the strands of the workload
    pos, vel, roots = ba.sphere_groom(S, N, radius=1.0, length=0.2, jitter=0.05, seed=2024)
    params = wl.c2_params(ba)
    wind = ba.wind_table(params)

    def run(nthreads, seconds, k0):
        done, k, t0 = 0, 0, time.perf_counter()
        while True:
            t = float(k0 + k + 1) * float(DT)
            m, sph = wl.c2_frame(t)
            op = oracle_params_for(params, N, DT, spheres=[tuple(x) for x in sph], model=m, root_mode=1, time=t,
                                   wind=wind)
            orc.step_f32(op, pos, vel, roots, nthreads=nthreads)
            done += S * N
            k += 1
            el = time.perf_counter() - t0
            if el >= seconds or k >= 2000:
                return done / el, k, el

    v_all, k_all, el_all = run(threads, seconds_target, 0)
    v_one, _, _ = run(1, seconds_target / 6.0, k_all)
    # the CPU configuration: 4,096 x 16, one sphere, 100 steps
    c1 = {}
    for nt in sorted({1, threads}):
        p1, v1, _ = ba.sphere_groom(4096, 16, radius=1.0, length=0.2, jitter=0.05, seed=1, coherent=False)
        op = oracle_params_for(ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0), 16, DT,
                               spheres=[(0.0, 0.0, 0.0, 1.0)])
        c0 = time.perf_counter()
        for _ in range(100):
            orc.step_f32(op, p1, v1, nthreads=nt)
        c1["threads_%d_s" % nt] = time.perf_counter() - c0
        c1["threads_%d_cps" % nt] = 4096 * 16 * 100 / c1["threads_%d_s" % nt]
    return dict(value=v_all, unit="CP updates/s", cores=threads, kind="port",
                sample="%d strands x %d points (1/8 of the single-groom workload), %d steps, %.1f s, "
                       "OpenMP over strands on %d threads" % (S, N, k_all, el_all, threads),
                single_thread_value=v_one, c1_4096x16_100_steps=c1, host=info)


def bench_batched(ba, hipm, frames, warmup, comp=None):
    comp = comp or wl.c3_grooms()
    ctx = context(ba)
    grooms = wl.make_c3(ba, ctx, comp)
    cps = sum(c["S"] * c["N"] for c in comp)
    stream = hipm.Stream()
    e0, e1 = hipm.Event(), hipm.Event()
    batch = ba.Batch(ctx, grooms)
    ntiles, grid = batch.info()
    dts = np.full(len(grooms), DT, np.float32)
    res = {}
    # Time runs on through both legs: the plain leg continues the motion where
    # the batched one stopped (a jump back in time would teleport every root).
    inputs = wl.c3_inputs_ahead(comp, (np.arange(2 * (warmup + frames)) + 1) * float(DT))
    frame0 = 0
    for mode in ("batched", "plain"):
        for k in range(frame0, frame0 + warmup):
            wl.c3_set_frame(ba, grooms, inputs, k)
            if mode == "batched":
                batch.launch(dts, stream.h)
            else:
                for g in grooms:
                    g.step(DT, stream.h)
        hipm.device_synchronize()
        w0 = time.perf_counter()
        e0.record(stream.h)
        for k in range(frame0 + warmup, frame0 + warmup + frames):
            wl.c3_set_frame(ba, grooms, inputs, k)
            if mode == "batched":
                batch.launch(dts, stream.h)
            else:
                for g in grooms:
                    g.step(DT, stream.h)
        e1.record(stream.h)
        hipm.device_synchronize()
        wall = time.perf_counter() - w0
        res[mode] = dict(wall_s=wall, gpu_ms=e0.elapsed_ms(e1))
        frame0 += warmup + frames
    for g in grooms:
        p, v = g.download()
        assert np.isfinite(p).all() and np.isfinite(v).all(), "non-finite state after the batched benchmark"
    batch.close()
    for g in grooms:
        g.close()
    ctx.close()
    return dict(cps=cps, ngrooms=len(comp), ntiles=ntiles, grid=grid, **res)


def bench_served(ba, hipm, seconds, warm, ncreatures=256, seed=11, strand_scale=1.0, coalesce_us=0,
                 target_bytes=0, inflight=2):
    """Served request loop: resident creatures, Poisson step requests submitted
    as they fall due, continuous batching (8 GB per-batch cap), latency p50/p99
    from submit to completion."""
    comp = wl.served_composition(ncreatures, seed, strand_scale)
    ctx = context(ba)
    srv = ba.Server(ctx, batch_cap_bytes=8 << 30, max_inflight=inflight, coalesce_us=coalesce_us,
                    batch_target_bytes=target_bytes)
    params = wl.c3_params(ba)
    sid = []
    for c in comp:
        pos, vel, roots = wl.c3_state(ba, c)
        cid, resident = srv.add(pos, vel, rest_roots=roots, params=params)
        if not resident:
            raise RuntimeError("served workload does not fit the HBM budget")
        sid.append(cid)
        del pos, vel, roots
    cps = sum(c["S"] * c["N"] for c in comp)
    times, idx, reqs, rates = wl.served_schedule(ba, comp, warm + seconds, seed)
    ids = np.asarray(sid, np.int32)[idx]
    n = times.shape[0]
    t0 = time.perf_counter()
    i = 0
    reset = False
    while i < n:
        now = time.perf_counter() - t0
        if not reset and now >= warm:
            srv.reset_stats()
            reset = True
        j = int(np.searchsorted(times, now, side="right"))
        if j > i:
            srv.submit(ids[i:j], reqs[i:j])
            i = j
        else:
            time.sleep(min(2e-4, max(0.0, times[i] - now)))
    srv.drain(120000)
    st = srv.stats()
    # A strand whose consecutive points meet (two points pushed onto the same
    # spot of a collider) divides by a zero length, as normalize(0) in the
    # shader does; the oracle reproduces such strands bit for bit.  The server
    # quarantines such a creature; counted, not hidden.
    bad = int(sum(srv.status(c) for c in sid))
    srv.close()
    ctx.close()
    span = st["span_s"]
    return dict(creatures=len(comp), control_points=cps, requests=int(st["completed"]),
                quarantined_creatures=bad, dropped_requests=int(st["dropped"]),
                rate_hz_mean=float(np.mean(rates)), seconds=seconds, warmup_s=warm,
                coalesce_us=coalesce_us, batch_target_bytes=target_bytes, max_inflight=inflight,
                p50_ms=st["p50_us"] / 1000.0, p99_ms=st["p99_us"] / 1000.0, max_ms=st["max_us"] / 1000.0,
                mean_ms=st["mean_us"] / 1000.0, batches=int(st["batches"]),
                mean_batch_requests=st["mean_batch_requests"],
                value=st["cps_stepped"] / span if span > 0 else 0.0, unit="CP updates/s",
                value_finite=st["cps_finite"] / span if span > 0 else 0.0,
                per_busy_second=st["cps_stepped"] / st["busy_s"] if st["busy_s"] > 0 else 0.0,
                demand=float(np.sum(rates * np.array([c["S"] * c["N"] for c in comp]))),
                device_busy_frac=st["busy_s"] / span if span > 0 else 0.0)


def bench_skinned(ba, hipm, steps, warmup, S=4_194_304, N=32):
    """Skinned roots: 4,194,304 strands x 32 points bound to a 100k-triangle
    mesh skinned by 64 bones (4 influences, dual quaternions, smooth random
    rotations up to 30 degrees), 4 colliders; one skinning pass + one step per
    frame, 120 frames."""
    sc = ba.skinned_scene(S, N, ntris=100_000, nbones=64, radius=1.0, length=0.2, jitter=0.05, seed=5)
    ctx = context(ba)
    skin = ba.Skin(ctx, sc["rest"], sc["joints"], sc["weights"], sc["tris"], sc["nbones"])
    g = ba.Groom(ctx, sc["pos"], sc["vel"])
    g.bind_skin(skin, sc["tri"], sc["bary"])
    g.set_params(wl.c5_params(ba))
    nverts = sc["rest"].shape[0]
    del sc["pos"], sc["vel"]
    stream = hipm.Stream()
    e0, e1 = hipm.Event(), hipm.Event()
    bones = [ba.bones(sc, (k + 1) * float(DT)) for k in range(warmup + steps)]
    sphs = [wl.c5_spheres((k + 1) * float(DT)) for k in range(warmup + steps)]
    ks = [(hipm.Event(), hipm.Event()) for _ in range(steps)]

    def frame(k, ev=None):
        skin.set_bones(bones[k])
        skin.update(stream.h)
        g.set_colliders(sphs[k])
        if ev is None:
            g.step(DT, stream.h)
        else:
            g.step_timed(DT, stream.h, ev[0].h, ev[1].h)

    for k in range(warmup):
        frame(k)
    hipm.device_synchronize()
    w0 = time.perf_counter()
    e0.record(stream.h)
    for i, k in enumerate(range(warmup, warmup + steps)):
        frame(k, ks[i])
    e1.record(stream.h)
    hipm.device_synchronize()
    wall = time.perf_counter() - w0
    gpu_ms = e0.elapsed_ms(e1)
    kern_ms = np.array([a.elapsed_ms(b) for a, b in ks])
    nonfinite = g.status()
    lanes = ctx.lanes_for(N)
    g.close()
    skin.close()
    ctx.close()
    return dict(S=S, N=N, nverts=nverts, wall_s=wall, gpu_ms=gpu_ms, lanes=lanes, nonfinite=bool(nonfinite),
                kernel_us_mean=float(kern_ms.mean() * 1000.0))


def bench_expand(ba, hipm, reps, warmup, ntris=100_000, npoints=4, seed=5):
    """Render-side strand expansion (the stream pass of the reference renderer):
    one guide strand per vertex of a 100k-triangle scalp (the lat/long sphere
    mesh of the skinned scene), initialised as Hair::init_simulation does with
    the reference's 4 control points, central-difference tangents, a seeded
    random table and the reference tessellation settings (3 instances, 2
    lines, 16 subsegments)."""
    sc = ba.skinned_scene(1, npoints, ntris=ntris, nbones=4, seed=seed)
    roots = sc["rest"][:, :3].astype(np.float64)
    normals = roots / np.linalg.norm(roots, axis=1, keepdims=True)
    pos, vel = ba.reference_groom(roots, normals, npoints=npoints, maxlength=0.2, seed=seed)
    tan = ba.strand_tangents(pos)
    rnd = ba.random_table(seed, 4096)
    ctx = context(ba)
    g = ba.Groom(ctx, pos, vel)
    e = ba.Expand(ctx, pos.shape[0], npoints, sc["tris"], tan, rnd)
    p = ba.expand_params()
    n = e.vertex_count(p)
    buf = ba.DeviceBuffer(ctx, n * 16)
    ptr, _ = g.device_buffers()
    stream = hipm.Stream()
    e0, e1 = hipm.Event(), hipm.Event()
    for _ in range(warmup):
        e.run(ptr, buf.ptr, n, p, stream.h)
    hipm.device_synchronize()
    e0.record(stream.h)
    for _ in range(reps):
        e.run(ptr, buf.ptr, n, p, stream.h)
    e1.record(stream.h)
    hipm.device_synchronize()
    ms = e0.elapsed_ms(e1) / reps
    head = buf.download(np.float32)[:8 * 4]
    finite = bool(np.isfinite(head).all())
    ntri = sc["tris"].shape[0]
    written = n * 16
    read = pos.shape[0] * npoints * 32 + ntri * 16 + rnd.nbytes
    buf.close()
    e.close()
    g.close()
    ctx.close()
    return dict(ntris=ntri, strands=pos.shape[0], npoints=npoints, vertices=n, ms=ms, written=written, read=read,
                finite=finite)


def bench_marschner(ba, hipm, reps=200, res=128):
    """Marschner::generate (the two shading lookup tables) on the device."""
    ctx = context(ba)
    p = ba.marschner_params()
    nb = res * res * 16
    m, n = ba.DeviceBuffer(ctx, nb), ba.DeviceBuffer(ctx, nb)
    stream = hipm.Stream()
    e0, e1 = hipm.Event(), hipm.Event()
    from barbu_amd import _ffi
    import ctypes as C
    run = lambda: _ffi.check(_ffi.lib().ftl_marschner_luts(ctx.handle, C.byref(p), res, m.ptr, n.ptr, stream.h),
                             "ftl_marschner_luts")
    for _ in range(5):
        run()
    hipm.device_synchronize()
    e0.record(stream.h)
    for _ in range(reps):
        run()
    e1.record(stream.h)
    hipm.device_synchronize()
    us = e0.elapsed_ms(e1) * 1000.0 / reps
    m.close()
    n.close()
    ctx.close()
    return dict(res=res, us=us)


def load_traffic(workload):
    """HBM bytes per launch of the workload's kernel, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/traffic.json (None if
    not profiled)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
        return None if rec is None else float(rec["bytes"])
    except (OSError, ValueError, KeyError, TypeError):
        return None




def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--workload", default="all",
                    choices=["single", "batched", "served", "skinned", "expand", "marschner", "all"])
    ap.add_argument("--served-seconds", type=float, default=4.0)
    ap.add_argument("--served-coalesce-us", type=int, default=2000,
                    help="coalescing window of the served loop (0: launch a batch as soon as a slot frees)")
    ap.add_argument("--served-target-mb", type=int, default=2048,
                    help="traffic at which a coalescing batch launches at once, MB")
    ap.add_argument("--skinned-steps", type=int, default=120)
    ap.add_argument("--batched-frames", type=int, default=60)
    ap.add_argument("--tuning", default=None, help="launch-shape table to load instead of the shipped one")
    ap.add_argument("--lanes", type=int, default=0, choices=[0, 1, 2, 4],
                    help="lanes per strand of the single-groom workload (0: the tuned table)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (strands and grooms would shard with no "
              "exchange between GPUs).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import barbu_amd as ba
    from barbu_amd import _hip as hipm

    if args.tuning:
        set_tuning(args.tuning)

    log("single groom: 262144 x 32, %d warmup + %d timed steps" % (args.warmup, args.steps))
    r = bench_single(ba, hipm, args.steps, args.warmup, lanes=args.lanes or None)
    cps = r["S"] * r["N"]
    ms_per_step = 1000.0 * r["wall_s"] / args.steps
    value = cps * args.steps / r["wall_s"]
    kernel_s = r["kernel_us_mean"] * 1e-6
    achieved = cps * BYTES_PER_CP / kernel_s / 1e9
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("single"),
                     "kernel": "ftl_plain_wind_kernel",
                     "kernel_us": r["kernel_us_mean"], "kernel_us_median": r["kernel_us_median"],
                     "kernel_us_min": r["kernel_us_min"], "kernel_us_max": r["kernel_us_max"],
                     "stream_window_us": r["gpu_ms"] * 1000.0 / args.steps,
                     "bytes_per_launch": cps * BYTES_PER_CP,
                     "copy_reference_gbs": r["copy_gbs"]},
    }
    if args.workload in ("batched", "all"):
        log("batched: 64 uneven grooms, %d frames (graph) and %d frames of per-groom launches"
            % (args.batched_frames, args.batched_frames))
        b = bench_batched(ba, hipm, args.batched_frames, 3)
        fb = b["batched"]
        fp = b["plain"]
        v_b = b["cps"] * args.batched_frames / fb["wall_s"]
        v_p = b["cps"] * args.batched_frames / fp["wall_s"]
        frame_s = fb["gpu_ms"] / 1000.0 / args.batched_frames
        gb = b["cps"] * BYTES_PER_CP / frame_s / 1e9
        out["batched"] = {
            "workload": "64 grooms, strands log-uniform [1024, 262144], points {8,16,32,64}, lengths U(0.05,0.4), "
                        "random rigid motion, 1-3 spheres; one graph launch per frame",
            "control_points": b["cps"], "tiles": b["ntiles"], "grid": b["grid"], "frames": args.batched_frames,
            "value": v_b, "unit": "CP updates/s", "ms_per_frame": 1000.0 * fb["wall_s"] / args.batched_frames,
            "gpu_ms_per_frame": fb["gpu_ms"] / args.batched_frames,
            "plain_value": v_p, "plain_ms_per_frame": 1000.0 * fp["wall_s"] / args.batched_frames,
            "plain_gpu_ms_per_frame": fp["gpu_ms"] / args.batched_frames,
            "speedup_vs_plain": v_b / v_p,
            "roofline": {"bound": "hbm", "achieved": gb, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gb / HBM_PEAK_GBS, "traffic": load_traffic("batched"),
                         "frame_us": frame_s * 1e6, "bytes_per_frame": b["cps"] * BYTES_PER_CP,
                         "timing": "HIP events around the timed graph launches on their stream; a frame is "
                                   "the descriptor upload plus one tile-list kernel per launch class, the "
                                   "classes running side by side"},
        }
    if args.workload in ("served", "all"):
        target = args.served_target_mb << 20
        log("served: 256 creatures (sizes as the batched workload), Poisson requests at 60 Hz +-30%%, %.1f s, "
            "coalescing %d us" % (args.served_seconds, args.served_coalesce_us))
        sv = bench_served(ba, hipm, args.served_seconds, 1.0, coalesce_us=args.served_coalesce_us,
                          target_bytes=target)
        sv["workload"] = ("256 resident creatures, sizes drawn as the batched workload (strands log-uniform "
                          "[1024, 262144], points {8,16,32,64}); requests Poisson per creature at 60 Hz x "
                          "(1 + U(-0.3, 0.3)); 8 GB per-batch cap; 2 batches in flight")
        out["served"] = sv
        log("served at the stated total (about 20M control points: strand counts scaled by 1/18)")
        sv20 = bench_served(ba, hipm, args.served_seconds, 1.0, strand_scale=1.0 / 18.0,
                            coalesce_us=args.served_coalesce_us, target_bytes=target)
        sv20["workload"] = "as served, strand counts scaled by 1/18 (about 20M control points)"
        out["served_20m"] = sv20
    if args.workload in ("skinned", "all"):
        log("skinned: 4,194,304 x 32, 100k-triangle mesh, 64 bones, 4 colliders, %d frames" % args.skinned_steps)
        k = bench_skinned(ba, hipm, args.skinned_steps, 5)
        kcps = k["S"] * k["N"]
        kgb = kcps * BYTES_PER_CP / (k["kernel_us_mean"] * 1e-6) / 1e9
        out["skinned"] = {
            "workload": "4,194,304 strands x 32 points, roots bound to a 100k-triangle mesh skinned by 64 bones "
                        "(4 influences, dual quaternions, rotations <= 30 deg), 4 colliders, dt=1/60",
            "control_points": kcps, "mesh_vertices": k["nverts"], "frames": args.skinned_steps,
            "value": kcps * args.skinned_steps / k["wall_s"], "unit": "CP updates/s",
            "ms_per_frame": 1000.0 * k["wall_s"] / args.skinned_steps,
            "gpu_ms_per_frame": k["gpu_ms"] / args.skinned_steps, "lanes_per_strand": k["lanes"],
            "nonfinite": k["nonfinite"],
            "roofline": {"bound": "hbm", "achieved": kgb, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": kgb / HBM_PEAK_GBS, "traffic": load_traffic("skinned"),
                         "kernel": "ftl_plain_kernel", "kernel_us": k["kernel_us_mean"]},
        }
    if args.workload in ("expand", "all"):
        log("render-side expansion: 100k-triangle scalp, 4-point guides, 3 x 2 x 16 tessellation")
        x = bench_expand(ba, hipm, 50, 3)
        xgb = (x["written"] + x["read"]) / (x["ms"] / 1000.0) / 1e9
        out["expand"] = {
            "workload": "render-side strand expansion (the reference renderer's tessellation stream pass): "
                        "%d-triangle scalp, %d guide strands x %d points, 3 instances x 2 lines x 16 "
                        "subsegments" % (x["ntris"], x["strands"], x["npoints"]),
            "vertices": x["vertices"], "bytes_written": x["written"], "bytes_read": x["read"],
            "ms_per_pass": x["ms"], "achieved_gbs": xgb, "frac_hbm_peak": xgb / HBM_PEAK_GBS,
            "finite": x["finite"],
        }
    if args.workload in ("marschner", "all"):
        log("Marschner lookup tables: 2 x 128 x 128")
        mr = bench_marschner(ba, hipm)
        out["marschner"] = {"workload": "Marschner::generate: M and N lookup tables, 128 x 128 float4 each, "
                                        "reference shading parameters",
                            "us_per_generate": mr["us"]}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)")
        out["cpu_baseline"] = cpu_baseline(ba, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
