"""Synthetic workloads of the FTL step, shared by the benchmark and the tests.

Each function builds the inputs of one configuration with a fixed seed:

* single groom (262,144 strands x 32 points): rotating emitter, value-noise
  wind, a head sphere and an orbiting sphere (`c2_*`);
* batched uneven grooms (64 grooms, strands log-uniform in [1024, 262144],
  points from {8, 16, 32, 64}): `c3_*`, `rigid*`, `world_spheres`;
* served creatures (256 of the batched distribution, Poisson requests at
  60 Hz +-30%): `served_schedule`;
* skinned roots (4,194,304 x 32 on a 100k-triangle mesh, 64 bones):
  `c5_spheres`.

The reference has no workloads of its own (its demo steps one 448-strand
groom per frame, src/Application.cc:38-39); these are host-side generators
only and hold no compute of the step.
"""
import numpy as np

DT = np.float32(1.0 / 60.0)


# --- single groom -----------------------------------------------------------

def c2_frame(t):
    """Per-step inputs of the single-groom workload at time t: emitter rotation
    about y at 1 rad/s (row-major 4x4), sphere r=1 at the origin and r=0.3
    orbiting the y axis at radius 1.2, 1 rad/s."""
    c, s = np.cos(t), np.sin(t)
    m = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)
    sph = np.array([[0, 0, 0, 1.0], [1.2 * np.cos(t), 0.0, 1.2 * np.sin(t), 0.3]], np.float32)
    return m, sph


def c2_params(ba):
    return ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0, iterations=8,
                          wind_amplitude=2.0, wind_frequency=2.0, wind_scroll=(0.5, 0.0, 0.25), wind_seed=2)


# --- batched uneven grooms --------------------------------------------------

def c3_grooms(seed=3, ngrooms=64):
    """Composition of the batched workload: 64 grooms, strand counts
    log-uniform in [1024, 262144], points per strand from {8,16,32,64},
    lengths U(0.05, 0.4), a random rigid motion each (rotation about a random
    axis at 0.5-2 rad/s plus a swaying translation), 1-3 sphere colliders that
    move with the groom (a head sphere r=1 at its centre, 0-2 more)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(ngrooms):
        S = int(round(np.exp(rng.uniform(np.log(1024), np.log(262144)))))
        N = int(rng.choice([8, 16, 32, 64]))
        L = float(rng.uniform(0.05, 0.4))
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        w = float(rng.uniform(0.5, 2.0))
        sway = rng.uniform(0.0, 0.2, size=3)
        ncol = int(rng.integers(1, 4))
        extra = []
        for _ in range(ncol - 1):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            extra.append((*(d * rng.uniform(1.0, 1.5)), rng.uniform(0.2, 0.5)))
        center = np.array([(i % 8) * 3.0, 0.0, (i // 8) * 3.0])
        out.append(dict(S=S, N=N, L=L, axis=axis, w=w, sway=sway, spheres=[(0, 0, 0, 1.0)] + extra,
                        center=center, seed=1000 + i))
    return out


def c3_params(ba):
    return ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0, iterations=8)


def rigid(c, t):
    """Row-major 4x4 world transform of groom c at time t."""
    return rigid_many(c, [t])[0]


def rigid_many(c, t):
    """rigid() for an array of times: (len(t), 4, 4) row-major."""
    t = np.asarray(t, np.float64)
    a = c["w"] * t
    k = c["axis"]
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    K2 = K @ K
    m = np.zeros((t.shape[0], 4, 4))
    m[:, :3, :3] = np.eye(3) + np.sin(a)[:, None, None] * K + (1 - np.cos(a))[:, None, None] * K2
    m[:, :3, 3] = c["center"] + c["sway"][None, :] * np.sin(2.0 * t)[:, None]
    m[:, 3, 3] = 1.0
    return m.astype(np.float32)


def world_spheres(c, m):
    """The groom's colliders under its row-major transform m: (k, 4)."""
    sp = np.array(c["spheres"], np.float32)
    out = sp.copy()
    out[:, :3] = sp[:, :3] @ m[:3, :3].T + m[:3, 3]
    return out


def c3_state(ba, c):
    """Initial state of groom c at its world placement at t = 0 (roots stay in
    object space: the step applies the transform): (pos, vel, rest_roots)."""
    pos, vel, roots = ba.sphere_groom(c["S"], c["N"], radius=1.0, length=c["L"], jitter=0.05, seed=c["seed"])
    m0 = rigid(c, 0.0)
    pos[..., :3] = pos[..., :3] @ m0[:3, :3].T + m0[:3, 3]
    return pos, vel, roots


def make_c3(ba, ctx, comp, keep=None):
    """The grooms of composition comp, on the device.  keep(i, pos, vel,
    roots), if given, sees each groom's initial state before it is freed."""
    grooms = []
    params = c3_params(ba)
    for i, c in enumerate(comp):
        pos, vel, roots = c3_state(ba, c)
        if keep is not None:
            keep(i, pos, vel, roots)
        g = ba.Groom(ctx, pos, vel, rest_roots=roots)
        g.set_params(params)
        g.set_root_mode(ba.FTL_ROOT_REST)
        grooms.append(g)
        del pos, vel, roots
    return grooms


def c3_inputs_ahead(comp, times):
    """The per-frame inputs of the batched workload for every frame, computed
    ahead: per groom, the column-major root transforms (frames, 16) and the
    world-space colliders (frames, k, 4)."""
    out = []
    for c in comp:
        ms = rigid_many(c, times)
        cm = np.ascontiguousarray(ms.transpose(0, 2, 1)).reshape(len(times), 16)
        sp = np.array(c["spheres"], np.float32)
        ws = np.repeat(sp[None], len(times), axis=0)
        ws[:, :, :3] = np.einsum("kj,nij->nki", sp[:, :3], ms[:, :3, :3]) + ms[:, None, :3, 3]
        out.append((np.ascontiguousarray(cm, np.float32), np.ascontiguousarray(ws, np.float32)))
    return out


def c3_set_frame(ba, grooms, inputs, f):
    """Hands frame f of c3_inputs_ahead to the grooms (two library calls per
    groom)."""
    L = ba.lib()
    for g, (cm, ws) in zip(grooms, inputs):
        ba.check(L.ftl_groom_set_root_transform(g.handle, cm[f].ctypes.data), "ftl_groom_set_root_transform")
        ba.check(L.ftl_groom_set_colliders(g.handle, ws[f].ctypes.data, ws.shape[1]), "ftl_groom_set_colliders")


# --- served creatures -------------------------------------------------------

def served_composition(ncreatures=256, seed=11, strand_scale=1.0):
    comp = c3_grooms(seed=seed, ngrooms=ncreatures)
    for c in comp:
        c["S"] = max(64, int(round(c["S"] * strand_scale)))
    return comp


def served_requests(ba, c, nreq, first=0):
    """Requests first .. first + nreq - 1 of creature c: its k-th request steps
    dt = 1/60 at simulated time (k + 1)/60 under its rigid motion."""
    tsim = (np.arange(first, first + nreq) + 1) * float(DT)
    ms = rigid_many(c, tsim)
    r = ba.make_requests(nreq, DT)
    r["time"] = tsim
    r["has_transform"] = 1
    r["model"] = np.transpose(ms, (0, 2, 1)).reshape(-1, 16)
    sp = np.array(c["spheres"], np.float32)
    wsp = np.einsum("nij,kj->nki", ms[:, :3, :3], sp[:, :3]) + ms[:, None, :3, 3]
    r["ncolliders"] = sp.shape[0]
    sph = np.zeros((nreq, 4, 4), np.float32)
    sph[:, : sp.shape[0], :3] = wsp
    sph[:, : sp.shape[0], 3] = sp[:, 3]
    r["spheres"] = sph.reshape(-1, 16)
    return r


def served_schedule(ba, comp, seconds, seed, rate=60.0, jitter=0.3):
    """Arrival times and inputs of every step request: creature i asks at a
    Poisson rate 60 Hz x (1 + U(-0.3, 0.3))."""
    rng = np.random.default_rng(seed)
    rates = rate * (1.0 + rng.uniform(-jitter, jitter, size=len(comp)))
    times, ids, reqs = [], [], []
    for i, c in enumerate(comp):
        n = rng.poisson(rates[i] * seconds * 1.2) + 8
        arr = np.cumsum(rng.exponential(1.0 / rates[i], size=n))
        arr = arr[arr < seconds]
        if arr.size == 0:
            continue
        times.append(arr)
        ids.append(np.full(arr.size, i, np.int32))
        reqs.append(served_requests(ba, c, arr.size))
    times = np.concatenate(times)
    order = np.argsort(times, kind="stable")
    return times[order], np.concatenate(ids)[order], np.concatenate(reqs)[order], rates


# --- skinned roots ----------------------------------------------------------

def c5_spheres(t):
    """Four colliders of the skinned workload at time t: the head (r=1 at the
    origin) and three spheres orbiting it."""
    return np.array([[0.0, 0.0, 0.0, 1.0],
                     [1.3 * np.cos(t), 0.2, 1.3 * np.sin(t), 0.35],
                     [-1.2, -0.5 + 0.2 * np.sin(2 * t), 0.3, 0.4],
                     [0.3 * np.sin(t), 1.25, -1.1 * np.cos(t), 0.3]], np.float32)


def c5_params(ba):
    return ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0, iterations=8)
