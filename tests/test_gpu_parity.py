"""GPU parity: the HIP step (libftl.so, through the C ABI) against the FP32
oracle (oracle/ftl_oracle.cpp) on the same inputs, step by step.

Tolerance (the stated bar): per step, |GPU - oracle| <= 1e-5 x strand length on
every position and velocity component.  The kernels perform the same IEEE
operations as the FP32 oracle, so results are also expected to be bit-identical,
which is asserted separately.
"""
import numpy as np
import pytest

import barbu_amd as ba
from oracle import ftl_oracle as orc
from tests.helpers import assert_bitexact, max_abs, oracle_params_for

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)
TOL_REL = 1e-5  # of strand length


def strand_length(pos):
    return float(np.sum(pos[0, 1:, 3])) if pos.shape[1] > 1 else 1.0


def c1_params():
    # 4,096 strands x 16 points, gravity (0,-9.81,0), damping 0.9, 1 sphere r=1 at 0
    return ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)


def run_compare(ctx, pos, vel, params, spheres, steps, *, lanes=0, roots=None, root_mode=0, models=None,
                times=None, bitexact=True, label=""):
    """Steps a groom on the GPU and the oracle in lockstep; checks every step."""
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    g = ba.Groom(ctx, pos, vel, rest_roots=roots)
    g.set_params(params)
    g.set_colliders(np.asarray(spheres, np.float32).reshape(-1, 4))
    g.set_lanes(lanes)
    if root_mode:
        g.set_root_mode(root_mode)
    wind = ba.wind_table(params) if params.wind_amplitude != 0.0 else None
    P, V = pos.copy(), vel.copy()
    L = strand_length(pos)
    worst = 0.0
    for k in range(steps):
        model = models(k) if models else None
        t = times(k) if times else 0.0
        if model is not None:
            g.set_root_transform(model)
        g.set_time(t)
        g.step(DT)
        op = oracle_params_for(params, pos.shape[1], DT, spheres=[tuple(s) for s in np.reshape(spheres, (-1, 4))],
                               model=model, root_mode=root_mode, time=t, wind=wind)
        orc.step_f32(op, P, V, roots)
        gp, gv = g.download()
        d = max(max_abs(gp, P), max_abs(gv, V))
        worst = max(worst, d)
        assert d <= TOL_REL * L, "%s step %d: max |gpu - oracle| = %g > %g" % (label, k, d, TOL_REL * L)
        if bitexact:
            assert_bitexact(gp, P, "%s step %d positions" % (label, k))
            assert_bitexact(gv, V, "%s step %d velocities" % (label, k))
    g.close()
    return P, V, worst


def test_c1_workload_100_steps(ctx):
    """4,096 x 16, C1 parameters, 100 steps; FP32 error vs FP64 reported."""
    pos, vel, _ = ba.sphere_groom(4096, 16, radius=1.0, length=0.2, jitter=0.05, seed=1, coherent=False)
    params = c1_params()
    P, V, worst = run_compare(ctx, pos, vel, params, [(0, 0, 0, 1)], 100, label="c1")
    # FP32 (oracle == GPU) vs FP64 after 1 and after 100 steps
    P32, V32 = pos.copy(), vel.copy()
    P64, V64 = pos.astype(np.float64), vel.astype(np.float64)
    op = oracle_params_for(params, 16, DT, spheres=[(0, 0, 0, 1)])
    errs = {}
    for k in range(100):
        orc.step_f32(op, P32, V32)
        orc.step_f64(op, P64, V64)
        if k in (0, 99):
            errs[k + 1] = (max_abs(P32, P64), max_abs(V32, V64))
    print("\nFP32 vs FP64 after 1 step: pos %.3g vel %.3g; after 100 steps: pos %.3g vel %.3g (strand length 0.2)"
          % (errs[1][0], errs[1][1], errs[100][0], errs[100][1]))
    assert np.array_equal(P32.view(np.uint32), P.view(np.uint32))
    assert errs[1][0] < 1e-5


def test_reference_preset(ctx):
    """Shipped shader constants (sigma 1.45, delta 0.80, force 20 g, 4 points, one
    sphere) on a groom built like Hair::init_simulation."""
    rng = np.random.default_rng(7)
    n = rng.normal(size=(448, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    roots = (0.25 * n).astype(np.float32)
    pos, vel = ba.reference_groom(roots, n.astype(np.float32), npoints=4, maxlength=0.5, seed=3)
    run_compare(ctx, pos, vel, ba.reference_params(), [(0.0, 0.0, 0.0, 0.25)], 12, label="reference")


CLASSES = [(4, 1), (8, 1), (8, 2), (16, 1), (16, 2), (16, 4), (32, 1), (32, 2), (32, 4), (64, 2), (64, 4), (64, 1)]
GUARDED = [2, 3, 5, 7, 9, 12, 15, 17, 20, 31, 33, 40, 48, 49, 55, 63]


@pytest.mark.parametrize("npoints,lanes", CLASSES)
def test_launch_classes(ctx, npoints, lanes):
    pos, vel, _ = ba.sphere_groom(1000, npoints, radius=1.0, length=0.3, seed=npoints * 10 + lanes)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.45)
    spheres = [(0, 0, 0, 1), (0.3, 1.1, 0.0, 0.35)]
    run_compare(ctx, pos, vel, params, spheres, 3, lanes=lanes, label="N=%d G=%d" % (npoints, lanes))


@pytest.mark.parametrize("lanes", [0, 1])
@pytest.mark.parametrize("npoints", GUARDED)
def test_guarded_point_counts(ctx, npoints, lanes):
    """Ragged point counts: by default the two-lane guarded classes (the last
    lane of a strand partly filled) above 8 points; with one lane asked for,
    the one-lane guarded classes."""
    pos, vel, _ = ba.sphere_groom(333, npoints, radius=1.0, length=0.25, seed=npoints)
    params = ba.make_params(force_scale=1.0, damping=0.85, length_scale=1.2)
    run_compare(ctx, pos, vel, params, [(0, 0, 0, 1)], 3, lanes=lanes, label="N=%d lanes=%d" % (npoints, lanes))


@pytest.mark.parametrize("npoints", [3, 5, 8, 20, 32, 48])
def test_tips_inside_a_collider(ctx, npoints):
    """A sphere that contains the strand tips: every constraint iteration
    pushes the last points, which must be collided exactly once per
    iteration (a second push moves a point on the surface by rounding)."""
    pos, vel, _ = ba.sphere_groom(400, npoints, radius=1.0, length=0.25, seed=50 + npoints)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)
    run_compare(ctx, pos, vel, params, [(0, 0, 0, 1.15)], 3, label="tips N=%d" % npoints)


def test_dftl_mode_with_clamp(ctx):
    pos, vel, _ = ba.sphere_groom(700, 16, radius=1.0, length=0.2, seed=11)
    vel[:, 1:, :3] = np.random.default_rng(1).normal(scale=3.0, size=vel[:, 1:, :3].shape)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0,
                            velocity_mode=ba.FTL_VELOCITY_DFTL, max_velocity=2.0)
    run_compare(ctx, pos, vel, params, [(0, 0, 0, 1), (0.5, 0.8, 0.0, 0.4)], 5, label="dftl")


@pytest.mark.parametrize("npoints", [12, 20, 40, 63])
def test_dftl_mode_ragged(ctx, npoints):
    """DFTL velocities (they re-read the step's input positions) on ragged
    strands: the two-lane guarded classes read only the strand's own points,
    the last one included."""
    pos, vel, _ = ba.sphere_groom(3000, npoints, radius=1.0, length=0.3, seed=60 + npoints)
    vel[:, 1:, :3] = np.random.default_rng(npoints).normal(scale=3.0, size=vel[:, 1:, :3].shape)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.2,
                            velocity_mode=ba.FTL_VELOCITY_DFTL, max_velocity=2.0)
    run_compare(ctx, pos, vel, params, [(0, 0, 0, 1), (0.5, 0.8, 0.0, 0.4)], 3, label="dftl N=%d" % npoints)


def test_wind_and_rotating_emitter(ctx):
    """C2 features at small size: value-noise wind, emitter rotating about y at
    1 rad/s (rest roots), two colliders (one orbiting)."""
    pos, vel, roots = ba.sphere_groom(2000, 32, radius=1.0, length=0.2, seed=5)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0, wind_amplitude=2.0,
                            wind_frequency=2.0, wind_scroll=(0.5, 0.1, -0.3), wind_seed=99)

    def model(k):
        a = float(k + 1) * float(DT)
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)

    spheres = [(0, 0, 0, 1), (1.2, 0.0, 0.0, 0.3)]
    run_compare(ctx, pos, vel, params, spheres, 6, roots=roots, root_mode=ba.FTL_ROOT_REST, models=model,
                times=lambda k: (k + 1) * float(DT), label="wind")


@pytest.mark.parametrize("nstrands,npoints,lanes",
                         [(150100, 32, 2), (140000, 20, 1), (150100, 16, 2), (140000, 20, 0), (100100, 40, 0)])
def test_wind_lds_workgroups(ctx, nstrands, npoints, lanes):
    """Plain launches with wind run workgroups of four tiles that share an
    LDS copy of the half lattice (the 2-lane 32- and 16-point classes, the
    one-lane guarded 32-point class and the two-lane guarded 16- and 32-point
    classes).  Hundreds of workgroups, the last one partial and its last tile
    partial."""
    pos, vel, _ = ba.sphere_groom(nstrands, npoints, radius=1.0, length=0.3, seed=5)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.45, wind_amplitude=2.0,
                            wind_frequency=2.0, wind_scroll=(0.5, 0.0, 0.25), wind_seed=7)
    spheres = [(0, 0, 0, 1), (0.3, 1.1, 0.0, 0.35)]
    run_compare(ctx, pos, vel, params, spheres, 2, lanes=lanes, times=lambda k: (k + 1) * float(DT),
                label="wind lds n=%d N=%d G=%d" % (nstrands, npoints, lanes))


@pytest.mark.parametrize("ncol", [0, 3, 4])
def test_collider_counts_and_iterations(ctx, ncol):
    pos, vel, _ = ba.sphere_groom(500, 8, radius=1.0, length=0.4, seed=ncol)
    spheres = [(0, 0, 0, 1), (0.2, 1.0, 0.1, 0.3), (-0.4, 0.9, 0.0, 0.25), (0.0, -1.2, 0.0, 0.5)][:ncol]
    for iters in (1, 3):
        params = ba.make_params(force_scale=20.0, damping=0.8, length_scale=1.45, iterations=iters)
        run_compare(ctx, pos, vel, params, spheres if ncol else np.zeros((0, 4)), 2,
                    label="ncol=%d K=%d" % (ncol, iters))


def test_model_transform_root(ctx):
    """FTL_ROOT_MODEL with a non-identity uModel applied to the current roots."""
    pos, vel, _ = ba.sphere_groom(300, 16, seed=21)
    m = np.array([[1, 0, 0, 0.01], [0, 1, 0, -0.02], [0, 0, 1, 0.005], [0, 0, 0, 1]], np.float32)
    run_compare(ctx, pos, vel, c1_params(), [(0, 0, 0, 1)], 3, models=lambda k: m, label="model")


def test_partial_tiles_and_single_strand(ctx):
    for S in (1, 63, 65, 129):
        pos, vel, _ = ba.sphere_groom(S, 32, seed=S)
        for lanes in (1, 2, 4):
            run_compare(ctx, pos, vel, c1_params(), [(0, 0, 0, 1)], 2, lanes=lanes, label="S=%d G=%d" % (S, lanes))


def test_nonfinite_flag(ctx):
    pos, vel, _ = ba.sphere_groom(200, 16, seed=4)
    g = ba.Groom(ctx, pos, vel)
    g.step(DT)
    assert g.status() is False
    vel[17, 5, 0] = np.nan
    g.upload(pos, vel)
    g.step(DT)
    assert g.status() is True
    p, v = g.download()
    assert np.isnan(p[17, -1]).any()
    assert np.isfinite(p[16]).all() and np.isfinite(p[18]).all()
    assert g.status() is False  # cleared by the read


def _mixed_grooms(ctx, seed=0):
    rng = np.random.default_rng(seed)
    grooms, states = [], []
    # exact classes and ragged point counts (guarded classes)
    for i, (S, N) in enumerate([(1000, 8), (130, 16), (777, 32), (64, 64), (2049, 16), (5, 32), (900, 64),
                                (300, 5), (200, 20), (100, 48), (50, 2)]):
        pos, vel, roots = ba.sphere_groom(S, N, center=tuple(rng.normal(size=3) * 2), length=float(rng.uniform(0.05, 0.4)),
                                          seed=100 + i)
        g = ba.Groom(ctx, pos, vel, rest_roots=roots)
        params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)
        g.set_params(params)
        c = roots[0, :3] - np.asarray([0, 0, 0])
        spheres = np.array([[*(pos[0, 0, :3] * 0 + np.mean(roots[:, :3], axis=0)), 1.0],
                            [*(np.mean(roots[:, :3], axis=0) + [0.0, 1.1, 0.0]), 0.3]], np.float32)[: 1 + i % 2]
        g.set_colliders(spheres)
        g.set_root_mode(ba.FTL_ROOT_REST)
        grooms.append(g)
        states.append(dict(pos=pos.copy(), vel=vel.copy(), roots=roots, params=params, spheres=spheres, N=N))
        del c
    return grooms, states


def _oracle_frame(states, dts, models):
    for s, dt, m in zip(states, dts, models):
        op = oracle_params_for(s["params"], s["N"], dt, spheres=[tuple(x) for x in s["spheres"]], model=m,
                               root_mode=ba.FTL_ROOT_REST)
        orc.step_f32(op, s["pos"], s["vel"], s["roots"])


def _rigid(k, i):
    a = 0.3 * (k + 1) + i
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0, 0.01 * k], [s, c, 0, 0.0], [0, 0, 1, -0.02 * i], [0, 0, 0, 1]], np.float32)


def test_batched_step_matches_oracle(ctx):
    grooms, states = _mixed_grooms(ctx)
    for k in range(3):
        models = [_rigid(k, i) for i in range(len(grooms))]
        for g, m in zip(grooms, models):
            g.set_root_transform(m)
        dts = [np.float32(1 / 60), np.float32(1 / 90)] * 8
        dts = dts[: len(grooms)]
        ba.batch_step(ctx, grooms, dts)
        _oracle_frame(states, dts, models)
        for i, (g, s) in enumerate(zip(grooms, states)):
            p, v = g.download()
            assert_bitexact(p, s["pos"], "batch groom %d step %d pos" % (i, k))
            assert_bitexact(v, s["vel"], "batch groom %d step %d vel" % (i, k))


def test_graph_replays_match_oracle(ctx):
    """Captured frames replayed several times, with eager work and
    synchronisation between replays; every replay checked."""
    grooms, states = _mixed_grooms(ctx, seed=3)
    b = ba.Batch(ctx, grooms)
    ntiles, grid = b.info()
    assert ntiles > 0 and grid > 0
    for k in range(4):
        models = [_rigid(k, i) for i in range(len(grooms))]
        for g, m in zip(grooms, models):
            g.set_root_transform(m)
        dts = [np.float32(1 / 60)] * len(grooms)
        b.launch(dts)
        _oracle_frame(states, dts, models)
        # eager work between replays
        grooms[0].download()
        for i, (g, s) in enumerate(zip(grooms, states)):
            p, v = g.download()
            assert_bitexact(p, s["pos"], "graph groom %d replay %d pos" % (i, k))
            assert_bitexact(v, s["vel"], "graph groom %d replay %d vel" % (i, k))
    b.close()


def test_batch_refuses_a_destroyed_groom(ctx):
    """A captured batch whose groom has been destroyed refuses its next launch
    (rather than stepping freed memory); the batch and the other grooms are
    then destroyed cleanly, in either order."""
    pos, vel, _ = ba.sphere_groom(300, 16, seed=8)
    grooms = [ba.Groom(ctx, pos, vel) for _ in range(3)]
    b = ba.Batch(ctx, grooms)
    b.launch([DT] * 3)
    grooms[1].close()
    with pytest.raises(ba.FtlError) as e:
        b.launch([DT] * 3)
    assert e.value.status == ba.FTL_ERR_INVALID_ARGUMENT
    b2 = ba.Batch(ctx, [grooms[0], grooms[2]])
    b2.launch([DT] * 2)
    b.close()
    grooms[0].close()
    b2.close()
    grooms[2].close()


def test_alone_equals_batched(ctx):
    """Each groom's result is the same stepped alone and in the mix."""
    grooms_a, _ = _mixed_grooms(ctx, seed=9)
    grooms_b, _ = _mixed_grooms(ctx, seed=9)
    for k in range(2):
        for i, (ga, gb) in enumerate(zip(grooms_a, grooms_b)):
            m = _rigid(k, i)
            ga.set_root_transform(m)
            gb.set_root_transform(m)
            ga.step(DT)
        ba.batch_step(ctx, grooms_b, [DT] * len(grooms_b))
    for ga, gb in zip(grooms_a, grooms_b):
        pa, va = ga.download()
        pb, vb = gb.download()
        assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))


def test_spans_many_grooms_one_class(ctx):
    """40 grooms of one launch class with 1 to 40 tiles each (some with a
    partial last tile), in shuffled order: each workgroup of the class kernel
    finds its groom by a binary search over the class's spans.  Every groom
    equals its own plain step bit for bit, through ftl_batch_step and through
    a captured batch."""
    rng = np.random.default_rng(21)
    spt = 64 // ctx.lanes_for(16)
    shapes = []
    for t in rng.permutation(np.arange(1, 41)):
        S = int(spt * t - (int(rng.integers(1, spt)) if t % 3 == 0 else 0))
        shapes.append(S)
    states = [ba.sphere_groom(S, 16, radius=1.0, length=0.25, seed=300 + i) for i, S in enumerate(shapes)]
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)
    sph = np.array([[0, 0, 0, 1.0], [0.2, 1.1, 0.0, 0.3]], np.float32)

    def make():
        out = []
        for pos, vel, _ in states:
            g = ba.Groom(ctx, pos, vel)
            g.set_params(params)
            g.set_colliders(sph)
            out.append(g)
        return out

    alone, eager, graph = make(), make(), make()
    b = ba.Batch(ctx, graph)
    ntiles, _ = b.info()
    assert ntiles == sum(-(-S // spt) for S in shapes)
    for k in range(2):
        for g in alone:
            g.step(DT)
        ba.batch_step(ctx, eager, [DT] * len(eager))
        b.launch([DT] * len(graph))
    for i, (ga, ge, gg) in enumerate(zip(alone, eager, graph)):
        pa, va = ga.download()
        for gx, label in ((ge, "batch_step"), (gg, "graph")):
            px, vx = gx.download()
            assert_bitexact(px, pa, "%s groom %d (%d strands) positions" % (label, i, shapes[i]))
            assert_bitexact(vx, va, "%s groom %d (%d strands) velocities" % (label, i, shapes[i]))
    b.close()
    for g in alone + eager + graph:
        g.close()


def test_skinned_roots(ctx):
    sc = ba.skinned_scene(3000, 16, ntris=2000, nbones=16, seed=2)
    skin = ba.Skin(ctx, sc["rest"], sc["joints"], sc["weights"], sc["tris"], sc["nbones"])
    g = ba.Groom(ctx, sc["pos"], sc["vel"])
    g.bind_skin(skin, sc["tri"], sc["bary"])
    params = c1_params()
    g.set_params(params)
    g.set_colliders(np.array([[0, 0, 0, 1.0]], np.float32))
    P, V = sc["pos"].copy(), sc["vel"].copy()
    for k in range(3):
        dq = ba.bones(sc, (k + 1) * float(DT) * 10)
        skin.set_bones(dq)
        skin.update()
        g.step(DT)
        verts = orc.skin_f32(dq, sc["joints"], sc["weights"], sc["rest"])
        assert_bitexact(skin.download(), verts, "skinned vertices step %d" % k)
        roots = orc.bary_roots_f32(sc["tris"], sc["tri"], sc["bary"], verts)
        op = oracle_params_for(params, 16, DT, spheres=[(0, 0, 0, 1.0)], root_mode=ba.FTL_ROOT_REST)
        orc.step_f32(op, P, V, roots)
        gp, gv = g.download()
        assert_bitexact(gp, P, "skinned step %d pos" % k)
        assert_bitexact(gv, V, "skinned step %d vel" % k)
    g.close()
    skin.close()


@pytest.mark.parametrize("S", [262144, 524288])
def test_c2_full_size_properties(ctx, S):
    """262,144 x 32 (the single-groom workload), and twice as many strands,
    whose 537 MB of state take the non-temporal-store form of the wind kernel:
    a few steps checked against the oracle on a strand sample, plus
    size-independent invariants on all strands: finite state, roots on the
    rotating emitter, segment lengths where no collider touches, and
    non-penetration of the last collider."""
    N = 32
    print("c2 properties S=%d" % S, flush=True)
    pos, vel, roots = ba.sphere_groom(S, N, radius=1.0, length=0.2, seed=2024)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0, wind_amplitude=2.0,
                            wind_frequency=2.0, wind_scroll=(0.4, 0.0, 0.2), wind_seed=7)
    g = ba.Groom(ctx, pos, vel, rest_roots=roots)
    g.set_params(params)
    g.set_root_mode(ba.FTL_ROOT_REST)
    wind = ba.wind_table(params)
    sample = np.arange(0, S, 97)
    P, V, R = pos[sample].copy(), vel[sample].copy(), roots[sample].copy()
    for k in range(4):
        t = (k + 1) * float(DT)
        c, s = np.cos(t), np.sin(t)
        m = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)
        sph = np.array([[0, 0, 0, 1], [1.2 * np.cos(t), 0, 1.2 * np.sin(t), 0.3]], np.float32)
        g.set_root_transform(m)
        g.set_colliders(sph)
        g.set_time(t)
        g.step(DT)
        op = oracle_params_for(params, N, DT, spheres=[tuple(x) for x in sph], model=m, root_mode=ba.FTL_ROOT_REST,
                               time=t, wind=wind)
        orc.step_f32(op, P, V, R)
    gp, gv = g.download()
    assert_bitexact(gp[sample], P, "c2 sample pos")
    assert_bitexact(gv[sample], V, "c2 sample vel")
    assert np.isfinite(gp).all() and np.isfinite(gv).all()
    # roots follow the emitter
    r = roots[:, :3] @ m[:3, :3].T
    assert np.abs(gp[:, 0, :3] - r).max() < 1e-5
    # non-penetration of the last collider
    d1 = np.linalg.norm(gp[:, 1:, :3] - sph[1, :3], axis=-1)
    assert d1.min() >= 0.3 * (1 - 1e-5)
    # inextensibility for strands no collider reaches
    far0 = np.linalg.norm(gp[:, 1:, :3], axis=-1).min(axis=1) > 1.0 + 1e-3
    far1 = np.linalg.norm(gp[:, 1:, :3] - sph[1, :3], axis=-1).min(axis=1) > 0.3 + 1e-3
    free = far0 & far1
    seg = np.linalg.norm(np.diff(gp[free, :, :3], axis=1), axis=-1)
    assert free.sum() > 0
    assert np.abs(seg / gp[free, 1:, 3] - 1.0).max() < 1e-4
    g.close()


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_collider_reach_bound_edges(ctx, lanes):
    """Colliders placed around the edge of what a strand can reach in one step
    (root distance ~ radius + strand length), and chains where one collider's
    push carries points toward the next: skipping unreachable colliders must
    not change any result."""
    rng = np.random.default_rng(17 + lanes)
    S, N, L = 640, 32, 0.3
    pos, vel, _ = ba.sphere_groom(S, N, radius=1.0, length=L, seed=40 + lanes)
    vel[:, 1:, :3] = rng.normal(scale=2.0, size=vel[:, 1:, :3].shape)
    params = ba.make_params(force_scale=20.0, damping=0.85, length_scale=1.0)
    base = pos[rng.integers(0, S), 0, :3].astype(np.float64)
    d = base / np.linalg.norm(base)
    spheres = [(0, 0, 0, 1.0),
               (*(base + d * (0.25 + L * 0.999)), 0.25),    # just inside the chain reach
               (*(base + d * (0.25 + L * 1.001)), 0.25),    # just outside it
               (*(base + d * 0.9 + rng.normal(scale=0.05, size=3)), 0.45)]
    run_compare(ctx, pos, vel, params, spheres, 3, lanes=lanes, label="reach G=%d" % lanes)


def test_tuned_lane_tables(tmp_path):
    """The shipped tuning file loads both lanes tables and both class tables
    (plain steps with wind take their own), also rewritten compactly (no
    blanks, other key order), and a ragged point count does not depend on
    them."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "barbu_amd",
                        "tuned_gfx950.json")
    with open(path) as f:
        table = json.load(f)
    compact = tmp_path / "compact.json"
    compact.write_text(json.dumps(dict(reversed(list(table.items()))), separators=(",", ":")))
    for p in (path, str(compact)):
        c = ba.Context(0)
        try:
            c.autotune(p)  # matches the device: loads, no timing
            for n in (4, 8, 16, 32, 64):
                assert c.lanes_for(n) == table["lanes"][str(n)]
                assert c.lanes_for(n, wind=True) == table["lanes_wind"][str(n)]
                assert c.class_for(n) == table["class"][str(n)]
                assert c.class_for(n, True) == table["class_wind"][str(n)]
            assert c.lanes_for(20) == 2 and c.lanes_for(63) == 4  # pipelined guarded classes
        finally:
            c.close()


@pytest.mark.parametrize("npoints,lanes", [(32, 1), (32, 2), (64, 4), (20, 0), (20, 1), (40, 0), (63, 0)])
def test_negative_length_scale_colliders(ctx, npoints, lanes):
    """A negative length scale places each point |sigma * l| from its leader
    on the other side: the per-wave collider cull must bound the chain by the
    absolute lengths.  A sphere within the strands' reach, exact and guarded
    classes, 1, 2 and 4 lanes per strand."""
    pos, vel, _ = ba.sphere_groom(700, npoints, radius=1.0, length=0.3, seed=70 + npoints)
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=-1.0)
    base = pos[5, 0, :3].astype(np.float64)
    d = base / np.linalg.norm(base)
    # inside the chain's reach from the roots, on both sides of the surface
    spheres = [(0, 0, 0, 1.0), (*(base - d * 0.2), 0.15), (*(base + d * 0.15), 0.1)]
    run_compare(ctx, pos, vel, params, spheres, 3, lanes=lanes, label="sigma=-1 N=%d G=%d" % (npoints, lanes))


def test_batch_step_alternating_streams(ctx):
    """Two batches of differently shaped grooms stepped alternately on two
    streams without host synchronisation between them: each batch's
    descriptors must stay intact until its kernels have read them."""
    from barbu_amd import _hip as hipm
    sa, sb = hipm.Stream(), hipm.Stream()
    ga, sta = _mixed_grooms(ctx, seed=21)
    shapes_b = [(3000, 64), (5000, 8), (77, 33), (4096, 32)]
    gb, stb = [], []
    for i, (S, N) in enumerate(shapes_b):
        pos, vel, roots = ba.sphere_groom(S, N, length=0.2, seed=300 + i)
        g = ba.Groom(ctx, pos, vel, rest_roots=roots)
        params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)
        g.set_params(params)
        spheres = np.array([[0, 0, 0, 1.0]], np.float32)
        g.set_colliders(spheres)
        g.set_root_mode(ba.FTL_ROOT_REST)
        gb.append(g)
        stb.append(dict(pos=pos.copy(), vel=vel.copy(), roots=roots, params=params, spheres=spheres, N=N))
    for k in range(3):
        ba.batch_step(ctx, ga, [DT] * len(ga), sa.h)
        ba.batch_step(ctx, gb, [DT] * len(gb), sb.h)
        _oracle_frame(sta, [DT] * len(ga), [None] * len(ga))
        _oracle_frame(stb, [DT] * len(gb), [None] * len(gb))
    sa.synchronize()
    sb.synchronize()
    for name, grooms, states in (("a", ga, sta), ("b", gb, stb)):
        for i, (g, s) in enumerate(zip(grooms, states)):
            p, v = g.download()
            assert_bitexact(p, s["pos"], "stream %s groom %d pos" % (name, i))
            assert_bitexact(v, s["vel"], "stream %s groom %d vel" % (name, i))


def test_tuned_tile_shape_variants(tmp_path):
    """A tuning table that picks tile-shape variants (another load/store chunk
    or occupancy target) is loaded, the picked classes run plain and batched
    steps, and the results stay bit-identical with the oracle."""
    import json
    import os
    shipped = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "barbu_amd",
                           "tuned_gfx950.json")
    table = json.load(open(shipped))
    variants = {}
    for cls in range(64):
        try:
            s = ba.class_shape(cls)
        except ba.FtlError:
            break
        n = s["points_per_lane"] * s["lanes"]
        if not s["guarded"] and n in (8, 16, 32, 64) and s["points_per_lane"] <= 32:
            variants.setdefault(n, []).append(cls)
    pick = {n: max(c) for n, c in variants.items()}  # the last class of each count: a variant where one exists
    table["class"] = {str(n): c for n, c in pick.items()}
    table["class_wind"] = {str(n): c for n, c in pick.items()}
    path = tmp_path / "tuned.json"
    path.write_text(json.dumps(table))
    c = ba.Context(0)
    try:
        c.autotune(str(path))
        for n, cls in pick.items():
            assert c.class_for(n) == cls and c.class_for(n, wind=True) == cls
            assert c.lanes_for(n) == ba.class_shape(cls)["lanes"]
        params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.2)
        spheres = [(0, 0, 0, 1), (0.3, 1.1, 0.0, 0.35)]
        for n in sorted(pick):
            pos, vel, _ = ba.sphere_groom(700, n, radius=1.0, length=0.3, seed=90 + n)
            run_compare(c, pos, vel, params, spheres, 2, label="variant class %d N=%d" % (pick[n], n))
        # batched through the picked classes
        grooms, states = [], []
        for i, n in enumerate(sorted(pick)):
            pos, vel, roots = ba.sphere_groom(900 + 37 * i, n, length=0.2, seed=120 + i)
            g = ba.Groom(c, pos, vel, rest_roots=roots)
            p = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)
            g.set_params(p)
            sph = np.array([[0, 0, 0, 1.0]], np.float32)
            g.set_colliders(sph)
            g.set_root_mode(ba.FTL_ROOT_REST)
            grooms.append(g)
            states.append(dict(pos=pos.copy(), vel=vel.copy(), roots=roots, params=p, spheres=sph, N=n))
        b = ba.Batch(c, grooms)
        for k in range(2):
            b.launch([DT] * len(grooms))
            _oracle_frame(states, [DT] * len(grooms), [None] * len(grooms))
        for i, (g, s) in enumerate(zip(grooms, states)):
            p_, v_ = g.download()
            assert_bitexact(p_, s["pos"], "variant batch groom %d pos" % i)
            assert_bitexact(v_, s["vel"], "variant batch groom %d vel" % i)
        b.close()
        for g in grooms:
            g.close()
    finally:
        c.close()
