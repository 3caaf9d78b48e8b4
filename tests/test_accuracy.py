"""FP32 error against FP64 of the FTL step, after 1 step and after 100 steps,
on the 4,096 x 16 workload (gravity, damping 0.9, one sphere r = 1 at the
origin, dt = 1/60).  The FP32 figures hold for the GPU path too: it is
bit-identical to the FP32 oracle (tests/test_gpu_parity.py).

Two measures, both max |FP32 - FP64| over every position and velocity
component, relative to the strand length (0.2):
  * trajectory: both precisions run 100 steps from the same start;
  * per step: at every step the FP64 step starts from the FP32 state, so the
    error is that of one FP32 step (the stated per-step bar is 1e-5).

    python tests/test_accuracy.py      prints the table
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import ftl_oracle as orc  # noqa: E402

F = np.float32
L = 0.2
STEPS = 100


def c1_state(nstrands=4096, npoints=16, seed=1):
    """The 4,096 x 16 start state: roots uniform on the unit sphere, straight
    strands along the normal + N(0, 0.05) jitter, 15 equal segments of 0.2/15."""
    rng = np.random.default_rng(seed)
    n = rng.normal(size=(nstrands, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = n + rng.normal(scale=0.05, size=n.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    seg = L / (npoints - 1)
    i = np.arange(npoints)
    pos = np.zeros((nstrands, npoints, 4), np.float32)
    pos[..., :3] = n[:, None, :] + (i * seg)[None, :, None] * d[:, None, :]
    pos[:, 1:, 3] = seg
    return pos, np.zeros_like(pos)


def params(npoints):
    return orc.make_params(npoints, F(1 / 60), iterations=8, length_scale=1.0, damping=0.9, force_scale=1.0,
                           spheres=[(0.0, 0.0, 0.0, 1.0)])


def err(a32, a64):
    return float(np.max(np.abs(a32[..., :3].astype(np.float64) - a64[..., :3])))


def strand_errs(a32, a64):
    """Per-strand max |FP32 - FP64|."""
    return np.max(np.abs(a32[..., :3].astype(np.float64) - a64[..., :3]), axis=(1, 2))


def fp32_vs_fp64(steps=STEPS):
    pos, vel = c1_state()
    P = params(pos.shape[1])
    p32, v32 = pos.copy(), vel.copy()
    p64, v64 = pos.astype(np.float64), vel.astype(np.float64)
    rows = {}
    worst_step = (0.0, 0.0)
    for k in range(steps):
        # one FP64 step from the current FP32 state
        q64, w64 = p32.astype(np.float64), v32.astype(np.float64)
        orc.step_f64(P, q64, w64)
        orc.step_f32(P, p32, v32, nthreads=orc.max_threads())
        orc.step_f64(P, p64, v64)
        one = (err(p32, q64), err(v32, w64))
        worst_step = (max(worst_step[0], one[0]), max(worst_step[1], one[1]))
        if k + 1 in (1, 10, steps):
            per = strand_errs(p32, p64) / L
            rows[k + 1] = dict(traj_pos=err(p32, p64) / L, traj_vel=err(v32, v64) / L, step_pos=one[0] / L,
                               step_vel=one[1] / L, worst_step_pos=worst_step[0] / L,
                               worst_step_vel=worst_step[1] / L, traj_median=float(np.median(per)),
                               traj_p99=float(np.percentile(per, 99)), diverged=int(np.sum(per > 1e-4)))
    return rows


def test_fp32_error_against_fp64():
    rows = fp32_vs_fp64()
    r1, r100 = rows[1], rows[STEPS]
    # one step: inside the 1e-5 x strand length bar (positions near 1.0 carry
    # FP32 rounding of a few ulp, ~1e-7 absolute each)
    assert r1["traj_pos"] < 1e-5 and r1["traj_vel"] < 1e-5
    # every single step of the 100, from the FP32 state
    assert r100["worst_step_pos"] < 1e-5 and r100["worst_step_vel"] < 1e-5
    # Trajectories: most strands stay within FP32 rounding of FP64; strands
    # sliding in contact with the sphere diverge (the contact is the only
    # non-smooth part of the step), which the max shows.
    assert r100["traj_median"] < 1e-5
    assert r100["diverged"] < 0.15 * 4096


if __name__ == "__main__":
    rows = fp32_vs_fp64()
    print("| after step | trajectory pos max | median strand | p99 strand | strands > 1e-4 | trajectory vel max "
          "| that step pos | that step vel | worst step so far pos | vel |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for k, r in sorted(rows.items()):
        print("| %d | %.2e | %.2e | %.2e | %d | %.2e | %.2e | %.2e | %.2e | %.2e |" % (
            k, r["traj_pos"], r["traj_median"], r["traj_p99"], r["diverged"], r["traj_vel"], r["step_pos"],
            r["step_vel"], r["worst_step_pos"], r["worst_step_vel"]))
    print("(max |FP32 - FP64| over all components / strand length %.1f)" % L)
