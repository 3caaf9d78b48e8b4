"""Serving loop (include/ftl_server.h) on the GPU: results of a creature are
the same alone and in the mix (and equal the FP32 oracle), admission by HBM
budget, per-batch traffic cap, in-flight slots, request errors.

Tolerance: 1e-5 x strand length per component (asserted), bit-identity
expected and asserted: every batch runs the kernels of the plain step, and a
step's inputs travel with its request."""
import numpy as np
import pytest

import barbu_amd as ba
from oracle import ftl_oracle as orc
from tests.helpers import assert_bitexact, max_abs, oracle_params_for

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def params():
    return ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0, iterations=8)


def creature(S, N, seed):
    pos, vel, roots = ba.sphere_groom(S, N, radius=1.0, length=0.25, jitter=0.05, seed=seed)
    return pos, vel, roots


def motion(seed, k):
    """Per-request inputs of creature `seed` at request k: a rotation about y,
    a translation, two spheres that move with it."""
    t = (k + 1) * float(DT)
    a = (0.5 + 0.1 * seed) * t
    c, s = np.cos(a), np.sin(a)
    m = np.array([[c, 0, s, 0.05 * k], [0, 1, 0, 0.0], [-s, 0, c, -0.02 * k], [0, 0, 0, 1]], np.float32)
    sph = np.array([[0.05 * k, 0, -0.02 * k, 1.0], [1.1 * np.cos(a), -0.4, 1.1 * np.sin(a), 0.35]], np.float32)
    return t, m, sph


def plain_reference(ctx, pos, vel, roots, steps, seed):
    g = ba.Groom(ctx, pos, vel, rest_roots=roots)
    g.set_params(params())
    g.set_root_mode(ba.FTL_ROOT_REST)
    for k in range(steps):
        t, m, sph = motion(seed, k)
        g.set_root_transform(m)
        g.set_colliders(sph)
        g.set_time(t)
        g.step(DT)
    out = g.download()
    g.close()
    return out


def requests(seed, steps):
    r = ba.make_requests(steps, DT)
    for k in range(steps):
        t, m, sph = motion(seed, k)
        ba.set_request(r[k], time=t, model=m, spheres=sph)
    return r


SHAPES = [(700, 8, 11), (513, 16, 12), (300, 32, 13), (130, 64, 14)]
STEPS = 4


def test_server_alone_equals_mix_and_oracle(ctx):
    data = {seed: creature(S, N, seed) for S, N, seed in SHAPES}
    ref = {seed: plain_reference(ctx, *data[seed], STEPS, seed) for _, _, seed in SHAPES}

    # the plain path itself against the oracle (small creature)
    S, N, seed = SHAPES[0]
    pos, vel, roots = data[seed]
    P, V = pos.copy(), vel.copy()
    for k in range(STEPS):
        t, m, sph = motion(seed, k)
        op = oracle_params_for(params(), N, DT, spheres=[tuple(x) for x in sph], model=m,
                               root_mode=ba.FTL_ROOT_REST, time=t)
        orc.step_f32(op, P, V, roots)
    assert max_abs(ref[seed][0], P) <= 1e-5 * 0.25 and max_abs(ref[seed][1], V) <= 1e-5 * 0.25
    assert_bitexact(ref[seed][0], P, "plain vs oracle positions")
    assert_bitexact(ref[seed][1], V, "plain vs oracle velocities")

    # each creature alone in a server
    for S, N, seed in SHAPES:
        srv = ba.Server(ctx)
        cid, resident = srv.add(*data[seed][:2], rest_roots=data[seed][2], params=params())
        assert resident
        srv.submit(np.full(STEPS, cid), requests(seed, STEPS))
        srv.drain(60000)
        p, v = srv.download(cid)
        assert_bitexact(p, ref[seed][0], "alone N=%d positions" % N)
        assert_bitexact(v, ref[seed][1], "alone N=%d velocities" % N)
        st = srv.stats()
        assert st["completed"] == STEPS and st["submitted"] == STEPS
        srv.close()

    # all of them in one server, requests interleaved in a random order that
    # keeps each creature's own order
    srv = ba.Server(ctx)
    ids = {}
    for S, N, seed in SHAPES:
        ids[seed], resident = srv.add(*data[seed][:2], rest_roots=data[seed][2], params=params())
        assert resident
    rng = np.random.default_rng(5)
    order = np.repeat([seed for _, _, seed in SHAPES], STEPS)
    rng.shuffle(order)
    nxt = {seed: 0 for _, _, seed in SHAPES}
    reqs = {seed: requests(seed, STEPS) for _, _, seed in SHAPES}
    tickets = []
    for seed in order:
        k = nxt[seed]
        nxt[seed] += 1
        tickets.append(srv.submit([ids[seed]], reqs[seed][k:k + 1])[0])
    assert list(tickets) == sorted(tickets)
    srv.wait(tickets[-1], 60000)
    srv.drain(60000)
    for S, N, seed in SHAPES:
        p, v = srv.download(ids[seed])
        assert_bitexact(p, ref[seed][0], "mix N=%d positions" % N)
        assert_bitexact(v, ref[seed][1], "mix N=%d velocities" % N)
        assert not srv.status(ids[seed])
    st = srv.stats()
    assert st["completed"] == len(order) and st["batches"] >= STEPS
    assert 0 < st["p50_us"] <= st["p99_us"] <= st["max_us"]
    assert st["cps_stepped"] == sum(S * N for S, N, _ in SHAPES) * STEPS
    srv.close()


def test_server_admission_by_budget(ctx):
    shapes = [(600, 16, 21), (600, 16, 22), (600, 16, 23)]
    data = {seed: creature(S, N, seed) for S, N, seed in shapes}
    per = 600 * 16 * 32 + 600 * 16 + 256  # pos + vel + rest roots + status
    srv = ba.Server(ctx, hbm_budget_bytes=2 * per + 1000)
    ids = []
    for S, N, seed in shapes:
        cid, resident = srv.add(*data[seed][:2], rest_roots=data[seed][2], params=params())
        ids.append((cid, resident))
    assert [r for _, r in ids] == [True, True, False]
    st = srv.stats()
    assert st["resident"] == 2 and st["waiting"] == 1
    reqs = {seed: requests(seed, 2) for _, _, seed in shapes}
    t_wait = srv.submit([ids[2][0]] * 2, reqs[23])
    srv.submit([ids[0][0]] * 2, reqs[21])
    srv.drain(60000)
    with pytest.raises(ba.FtlError) as e:
        srv.wait(t_wait[-1], 200)  # still queued behind the budget
    assert e.value.status == ba.FTL_ERR_BUSY
    assert not srv.resident(ids[2][0])
    srv.remove(ids[1][0])  # frees memory: the waiting creature is admitted
    assert srv.resident(ids[2][0])
    srv.wait(t_wait[-1], 60000)
    p, v = srv.download(ids[2][0])
    ref = plain_reference(ctx, *data[23], 2, 23)
    assert_bitexact(p, ref[0], "admitted creature positions")
    assert_bitexact(v, ref[1], "admitted creature velocities")
    with pytest.raises(ba.FtlError) as e:
        srv.add(*creature(5000, 64, 3)[:2], params=params())  # larger than the whole budget
    assert e.value.status == ba.FTL_ERR_BUDGET
    srv.close()


@pytest.mark.parametrize("cap,inflight", [(64, 1), (64, 3), (0, 4)])
def test_server_batch_cap_and_slots(ctx, cap, inflight):
    shapes = SHAPES[:3]
    data = {seed: creature(S, N, seed) for S, N, seed in shapes}
    srv = ba.Server(ctx, batch_cap_bytes=cap, max_inflight=inflight)
    ids = {seed: srv.add(*data[seed][:2], rest_roots=data[seed][2], params=params())[0] for _, _, seed in shapes}
    for k in range(3):
        for _, _, seed in shapes:
            srv.submit([ids[seed]], requests(seed, 3)[k:k + 1])
    srv.drain(60000)
    for S, N, seed in shapes:
        p, v = srv.download(ids[seed])
        ref = plain_reference(ctx, *data[seed], 3, seed)
        assert_bitexact(p, ref[0], "cap=%d N=%d" % (cap, N))
        assert_bitexact(v, ref[1], "cap=%d N=%d" % (cap, N))
    st = srv.stats()
    assert st["completed"] == 9
    if cap == 64:  # one creature per batch
        assert st["batches"] == 9 and st["mean_batch_requests"] == 1.0
    srv.close()


def test_server_request_errors(ctx):
    srv = ba.Server(ctx)
    cid, _ = srv.add(*creature(64, 8, 1)[:2], params=params())
    bad = ba.make_requests(1, 0.0)
    with pytest.raises(ba.FtlError):
        srv.submit([cid], bad)
    with pytest.raises(ba.FtlError):
        srv.submit([cid + 100], ba.make_requests(1))
    r = ba.make_requests(1)
    r["has_transform"] = 1
    r["model"] = np.eye(4, dtype=np.float32).reshape(16)
    r["model"][0, 3] = 1.0  # last row not (0, 0, 0, 1)
    with pytest.raises(ba.FtlError):
        srv.submit([cid], r)
    srv.remove(cid)
    with pytest.raises(ba.FtlError):
        srv.download(cid)
    srv.close()


def test_server_removal_cancels_pending_requests(ctx):
    """Requests still queued when their creature is removed complete without a
    step, and ftl_server_wait says so (FTL_ERR_CANCELLED); a creature waiting
    for memory keeps its requests queued until it is removed the same way.  A
    coalescing window of a second holds the requests of the resident creature
    in the queue."""
    with pytest.raises(ba.FtlError) as e:
        ba.Server(ctx, coalesce_us=2_000_000)
    assert e.value.status == ba.FTL_ERR_INVALID_ARGUMENT
    per = 500 * 16 * 32 + 256
    srv = ba.Server(ctx, hbm_budget_bytes=per + 1000, coalesce_us=1_000_000, batch_target_bytes=1 << 40)
    a, res_a = srv.add(*creature(500, 16, 51)[:2], params=params())
    b, res_b = srv.add(*creature(500, 16, 52)[:2], params=params())
    assert res_a and not res_b
    ta = srv.submit([a, a], requests(51, 2))
    tb = srv.submit([b], requests(52, 1))
    srv.remove(a)  # admits b; b's request stays queued in the window
    for t in ta:
        with pytest.raises(ba.FtlError) as e:
            srv.wait(t, 60000)
        assert e.value.status == ba.FTL_ERR_CANCELLED
    assert srv.resident(b)
    srv.remove(b)
    with pytest.raises(ba.FtlError) as e:
        srv.wait(tb[0], 60000)
    assert e.value.status == ba.FTL_ERR_CANCELLED
    st = srv.stats()
    assert st["dropped"] == 3 and st["completed"] == 0 and st["cps_stepped"] == 0
    srv.close()


def test_server_rejects_invalid_params_before_queueing(ctx):
    """A creature with invalid parameters is refused at ftl_server_add even
    when it would wait for memory, so it cannot block the creatures queued
    behind it."""
    shapes = [(600, 16, 31), (600, 16, 32)]
    data = {seed: creature(S, N, seed) for S, N, seed in shapes}
    per = 600 * 16 * 32 + 600 * 16 + 256
    srv = ba.Server(ctx, hbm_budget_bytes=per + 1000)
    a, res_a = srv.add(*data[31][:2], rest_roots=data[31][2], params=params())
    assert res_a
    bad = params()
    bad.iterations = 0
    with pytest.raises(ba.FtlError) as e:
        srv.add(*data[32][:2], rest_roots=data[32][2], params=bad)  # would have to wait for memory
    assert e.value.status == ba.FTL_ERR_INVALID_ARGUMENT
    nan = params()
    nan.damping = float("nan")
    with pytest.raises(ba.FtlError):
        srv.add(*data[32][:2], rest_roots=data[32][2], params=nan)
    b, res_b = srv.add(*data[32][:2], rest_roots=data[32][2], params=params())
    assert not res_b and srv.stats()["waiting"] == 1
    t = srv.submit([b, b], requests(32, 2))
    srv.remove(a)  # frees memory: the valid creature behind is admitted
    assert srv.resident(b)
    srv.wait(t[-1], 60000)
    p, v = srv.download(b)
    ref = plain_reference(ctx, *data[32], 2, 32)
    assert_bitexact(p, ref[0], "admitted creature positions")
    assert_bitexact(v, ref[1], "admitted creature velocities")
    assert srv.stats()["failed_admissions"] == 0
    srv.close()


def test_server_quarantines_nonfinite_creatures(ctx):
    """A creature whose state holds a NaN is quarantined after its first step:
    its later requests are not stepped (FTL_ERR_NONFINITE), the other
    creatures are unaffected, finite-only throughput leaves it out, and an
    upload of clean state lifts the quarantine."""
    shapes = [(400, 16, 41), (300, 32, 42)]
    data = {seed: creature(S, N, seed) for S, N, seed in shapes}
    srv = ba.Server(ctx)
    ids = {seed: srv.add(*data[seed][:2], rest_roots=data[seed][2], params=params())[0] for _, _, seed in shapes}
    pos, vel, roots = data[41]
    bad_vel = vel.copy()
    bad_vel[17, 5, 0] = np.nan
    srv.upload(ids[41], pos, bad_vel)
    t0 = srv.submit([ids[41], ids[42]], np.concatenate([requests(41, 1), requests(42, 1)]))
    srv.wait(t0[0], 60000)  # the step itself ran
    srv.wait(t0[1], 60000)
    assert srv.status(ids[41]) and not srv.status(ids[42])
    st = srv.stats()
    assert st["quarantined"] == 1
    assert st["cps_stepped"] - st["cps_finite"] == 400 * 16
    # later requests of the quarantined creature complete without a step
    p_before, v_before = srv.download(ids[41])
    t1 = srv.submit([ids[41]], requests(41, 2)[1:2])
    with pytest.raises(ba.FtlError) as e:
        srv.wait(t1[0], 60000)
    assert e.value.status == ba.FTL_ERR_NONFINITE
    p_after, v_after = srv.download(ids[41])
    assert np.array_equal(p_before.view(np.uint32), p_after.view(np.uint32))
    assert srv.stats()["dropped"] == 1
    # the neighbour steps on, bit-identical to its own reference
    t2 = srv.submit([ids[42]], requests(42, 2)[1:2])
    srv.wait(t2[0], 60000)
    p, v = srv.download(ids[42])
    ref = plain_reference(ctx, *data[42], 2, 42)
    assert_bitexact(p, ref[0], "neighbour positions")
    # clean state lifts the quarantine
    srv.upload(ids[41], pos, vel)
    assert not srv.status(ids[41])
    t3 = srv.submit([ids[41]], requests(41, 1))
    srv.wait(t3[0], 60000)
    p, v = srv.download(ids[41])
    ref = plain_reference(ctx, *data[41], 1, 41)
    assert_bitexact(p, ref[0], "re-uploaded creature positions")
    assert_bitexact(v, ref[1], "re-uploaded creature velocities")
    assert srv.stats()["quarantined"] == 0
    srv.close()
