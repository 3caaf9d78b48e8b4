"""Writes the regression fixtures of the FTL step: small inputs, their
parameters, and the FP32 and FP64 oracle outputs (oracle/ftl_oracle.cpp).

    python tests/golden/make_fixtures.py      (rewrites ftl_cases.json / .npz)

The cases follow the known-answer checks of the oracle tests (K1 hanging
strand, K6 three-point strand, K7 tiny random strands) plus small instances of
each feature on the path (reference preset, rest roots under a moving model
with two colliders, DFTL clamp, wind).  They lock the oracle and the GPU path
against regressions; the closed forms in tests/test_oracle.py pin the oracle
itself."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import barbu_amd as ba  # noqa: E402
from oracle import ftl_oracle as orc  # noqa: E402
from tests.helpers import oracle_params_for  # noqa: E402

DT = 1.0 / 60.0


def hanging(N, seg=0.05):
    pos = np.zeros((1, N, 4), np.float32)
    pos[0, :, 1] = -seg * np.arange(N)
    pos[0, 1:, 3] = seg
    return pos, np.zeros_like(pos)


def random_strands(S, N, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((S, N, 4), np.float32)
    pos[:, 0, :3] = rng.normal(size=(S, 3)) * 0.6
    for i in range(1, N):
        d = rng.normal(size=(S, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        ln = rng.uniform(0.05, 0.2, size=S)
        pos[:, i, :3] = pos[:, i - 1, :3] + d * ln[:, None]
        pos[:, i, 3] = ln
    vel = np.zeros_like(pos)
    vel[:, 1:, :3] = rng.normal(scale=0.3, size=(S, N - 1, 3))
    return pos, vel


def rot_y(a, t=(0.0, 0.0, 0.0)):
    c, s = np.cos(a), np.sin(a)
    return [[c, 0, s, t[0]], [0, 1, 0, t[1]], [-s, 0, c, t[2]], [0, 0, 0, 1]]


def cases():
    out = []
    pos, vel = hanging(9)
    out.append(dict(name="k1_hanging_n9", pos=pos, vel=vel, steps=1, spheres=[],
                    params=dict(force_scale=20.0, damping=0.8, length_scale=1.0)))
    pos = np.array([[[0, 0, 0, 0], [0, -1, 0, 1], [0, -2, 0, 1]]], np.float32)
    vel = np.zeros_like(pos)
    vel[0, 2, :3] = (1, 0, 0)
    out.append(dict(name="k6_three_point", pos=pos, vel=vel, steps=1, spheres=[], dt=1.0,
                    params=dict(force_scale=0.0, damping=0.8, length_scale=1.0, iterations=1)))
    for N in (2, 3, 4):
        pos, vel = random_strands(4, N, 10 + N)
        out.append(dict(name="k7_tiny_n%d" % N, pos=pos, vel=vel, steps=1, spheres=[[0, 0, 0, 1.0]],
                        params=dict(force_scale=20.0)))
    pos, vel, _ = ba.sphere_groom(16, 16, radius=1.0, length=0.2, jitter=0.05, seed=3)
    out.append(dict(name="c1_slice_16x16", pos=pos, vel=vel, steps=5, spheres=[[0, 0, 0, 1.0]],
                    params=dict(force_scale=1.0, damping=0.9, length_scale=1.0)))
    rng = np.random.default_rng(7)
    n = rng.normal(size=(8, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    pos, vel = ba.reference_groom((0.25 * n).astype(np.float32), n.astype(np.float32), npoints=4, maxlength=0.5,
                                  seed=3)
    out.append(dict(name="reference_preset_8x4", pos=pos, vel=vel, steps=3, spheres=[[0, 0, 0, 0.25]],
                    params=dict()))
    pos, vel, roots = ba.sphere_groom(8, 8, radius=1.0, length=0.3, jitter=0.05, seed=5)
    out.append(dict(name="rest_roots_model_two_spheres_8x8", pos=pos, vel=vel, roots=roots, steps=2,
                    spheres=[[0, 0, 0, 1.0], [0.9, 0.3, 0.2, 0.4]], root_mode=ba.FTL_ROOT_REST,
                    model=rot_y(0.3, (0.02, -0.01, 0.0)),
                    params=dict(force_scale=1.0, damping=0.9, length_scale=1.0)))
    pos, vel = random_strands(6, 8, 99)
    vel *= 40.0
    out.append(dict(name="dftl_clamp_6x8", pos=pos, vel=vel, steps=2, spheres=[[0, 0, 0, 0.5]],
                    params=dict(velocity_mode=ba.FTL_VELOCITY_DFTL, max_velocity=2.0, force_scale=1.0)))
    pos, vel, _ = ba.sphere_groom(8, 8, radius=1.0, length=0.2, jitter=0.05, seed=6)
    out.append(dict(name="wind_8x8", pos=pos, vel=vel, steps=2, spheres=[[0, 0, 0, 1.0]], time=1.25,
                    params=dict(force_scale=1.0, damping=0.9, length_scale=1.0, wind_amplitude=2.0,
                                wind_frequency=2.0, wind_scroll=(0.5, 0.0, 0.25), wind_seed=2)))
    return out


def run_oracle(case, double):
    """Steps a fixture case through the oracle; returns (pos, vel)."""
    params = ba.make_params(**case["params"])
    N = case["pos"].shape[1]
    dt = np.float32(case.get("dt", DT))
    wind = ba.wind_table(params) if params.wind_amplitude != 0.0 else None
    model = case.get("model")
    op = oracle_params_for(params, N, dt, spheres=[tuple(s) for s in case["spheres"]],
                           model=None if model is None else np.asarray(model, np.float32),
                           root_mode=case.get("root_mode", 0), time=case.get("time", 0.0), wind=wind)
    if double:
        p, v = case["pos"].astype(np.float64), case["vel"].astype(np.float64)
    else:
        p, v = case["pos"].astype(np.float32).copy(), case["vel"].astype(np.float32).copy()
    roots = case.get("roots")
    for _ in range(case["steps"]):
        (orc.step_f64 if double else orc.step_f32)(op, p, v, roots)
    return p, v


def main():
    meta, arrays = [], {}
    for c in cases():
        p32, v32 = run_oracle(c, False)
        p64, v64 = run_oracle(c, True)
        name = c["name"]
        arrays[name + ".pos"] = c["pos"]
        arrays[name + ".vel"] = c["vel"]
        if c.get("roots") is not None:
            arrays[name + ".roots"] = c["roots"]
        arrays[name + ".pos32"] = p32
        arrays[name + ".vel32"] = v32
        arrays[name + ".pos64"] = p64
        arrays[name + ".vel64"] = v64
        m = {k: c[k] for k in ("name", "steps", "spheres", "params") }
        for k in ("dt", "model", "root_mode", "time"):
            if k in c:
                m[k] = c[k]
        m["params"] = {k: (list(v) if isinstance(v, tuple) else v) for k, v in m["params"].items()}
        m["has_roots"] = c.get("roots") is not None
        meta.append(m)
    with open(os.path.join(HERE, "ftl_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ftl_cases.npz"), **arrays)
    print("wrote %d cases" % len(meta))


if __name__ == "__main__":
    main()
