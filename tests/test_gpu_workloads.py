"""GPU parity at the full sizes of the batched, served and skinned workloads.

Each test builds the workload exactly as the benchmark does
(barbu_amd/workloads.py) and steps it through the C ABI:

* batched: 64 uneven grooms (58 M control points), two graph-captured frames;
* served: 256 creatures at the stated total of about 20 M control points
  (strand counts of the batched distribution scaled by 1/18), interleaved
  requests through the serving loop with a coalescing window;
* skinned: 4,194,304 x 32 on a 100k-triangle mesh skinned by 64 bones, four
  colliders, two frames.

The FP32 oracle checks, bit for bit, a strided sample of strands of every
groom (every creature in full for the served workload) and the skinned mesh
vertices; size-independent invariants are checked on every strand: finite
state, roots on their source, non-penetration of the last collider, and
segment lengths |sigma * l| for strands no collider reaches.  Tolerance (the
stated bar, 1e-5 x strand length) is implied by bit-identity.
"""
import numpy as np
import pytest

import barbu_amd as ba
from barbu_amd import workloads as wl
from oracle import ftl_oracle as orc
from tests.helpers import assert_bitexact, oracle_params_for

pytestmark = pytest.mark.gpu

DT = wl.DT
THREADS = max(1, min(16, orc.max_threads()))


def transform_rows(m, x):
    """Row-major affine m applied to points x[..., 3] in float64."""
    return x.astype(np.float64) @ m[:3, :3].T.astype(np.float64) + m[:3, 3].astype(np.float64)


def check_invariants(gp, gv, root_src, spheres, sigma, label, chunk=1 << 18):
    """Size-independent checks on every strand of a groom after a step.
    root_src: the expected root positions [S, 3] (float64); spheres: the
    step's colliders (k, 4), applied in order."""
    S = gp.shape[0]
    sp = np.asarray(spheres, np.float64).reshape(-1, 4)
    nfree = 0
    for a in range(0, S, chunk):
        p = gp[a:a + chunk]
        v = gv[a:a + chunk]
        assert np.isfinite(p).all() and np.isfinite(v).all(), "%s: non-finite state" % label
        assert np.all(v[..., 3] == 0.0), "%s: velocity w must be 0" % label
        # roots on their source (the transform is evaluated in float32 by the step)
        d0 = np.abs(p[:, 0, :3].astype(np.float64) - root_src[a:a + chunk]).max()
        assert d0 < 1e-5, "%s: root off its source by %g" % (label, d0)
        x = p[:, 1:, :3].astype(np.float64)
        # float32 rounding of the coordinates themselves: a few ulp of the
        # largest coordinate of the chunk
        ulps = 8.0 * 2.0 ** -23 * max(1.0, float(np.abs(x).max()))
        if sp.shape[0]:
            c, r = sp[-1, :3], sp[-1, 3]
            dl = np.sqrt(((x - c) ** 2).sum(-1)).min()
            assert dl >= r * (1.0 - 1e-5) - ulps, "%s: a point is inside the last collider (%g < %g)" % (
                label, dl, r)
        free = np.ones(p.shape[0], bool)
        for k in range(sp.shape[0]):
            dk = np.sqrt(((x - sp[k, :3]) ** 2).sum(-1)).min(axis=1)
            free &= dk > sp[k, 3] + 1e-3
        if free.any():
            seg = np.sqrt((np.diff(p[free, :, :3].astype(np.float64), axis=1) ** 2).sum(-1))
            want = np.abs(sigma * p[free, 1:, 3].astype(np.float64))
            err = np.abs(seg - want) - (1e-4 * want + ulps)
            assert err.max() <= 0.0, "%s: segment length off by %g beyond 1e-4 relative" % (label, err.max())
            nfree += int(free.sum())
    return nfree


def test_batched_workload_64_grooms(ctx):
    """The batched workload at full size: 64 grooms, strands log-uniform in
    [1024, 262144], points from {8, 16, 32, 64}, rigid motions, 1-3 spheres;
    two frames of the captured graph."""
    comp = wl.c3_grooms()
    assert len(comp) == 64
    samples = {}

    def keep(i, pos, vel, roots):
        sel = np.arange(0, pos.shape[0], 37)
        samples[i] = (sel, pos[sel].copy(), vel[sel].copy(), roots.copy())

    grooms = wl.make_c3(ba, ctx, comp, keep)
    total = sum(c["S"] * c["N"] for c in comp)
    assert total > 50_000_000
    b = ba.Batch(ctx, grooms)
    params = wl.c3_params(ba)
    frames = 2
    inputs = wl.c3_inputs_ahead(comp, (np.arange(frames) + 1) * float(DT))
    print("\nbatched workload: %d grooms, %d control points, %d tiles" % (len(comp), total, b.info()[0]), flush=True)
    for f in range(frames):
        wl.c3_set_frame(ba, grooms, inputs, f)
        b.launch(np.full(len(grooms), DT, np.float32))
        for i, c in enumerate(comp):
            sel, P, V, R = samples[i]
            cm, ws = inputs[i]
            m = cm[f].reshape(4, 4).T
            op = oracle_params_for(params, c["N"], DT, spheres=[tuple(x) for x in ws[f]], model=m,
                                   root_mode=ba.FTL_ROOT_REST)
            orc.step_f32(op, P, V, R[sel], nthreads=THREADS)
    nfree = 0
    for i, (g, c) in enumerate(zip(grooms, comp)):
        gp, gv = g.download()
        sel, P, V, R = samples[i]
        label = "groom %d (%d x %d)" % (i, c["S"], c["N"])
        assert_bitexact(gp[sel], P, label + " sampled positions")
        assert_bitexact(gv[sel], V, label + " sampled velocities")
        cm, ws = inputs[i]
        m = cm[frames - 1].reshape(4, 4).T
        nfree += check_invariants(gp, gv, transform_rows(m, R[:, :3]), ws[frames - 1], params.length_scale, label)
        del gp, gv
    assert nfree > 0
    b.close()
    for g in grooms:
        g.close()


def test_served_workload_256_creatures(ctx):
    """256 creatures of the batched distribution at about 20 M control points
    (strand counts / 18), three requests each submitted interleaved through the
    serving loop (coalescing window 300 us, 2 slots): every creature equals the
    oracle bit for bit, and a sample stepped alone equals the mix."""
    comp = wl.served_composition(256, seed=11, strand_scale=1.0 / 18.0)
    total = sum(c["S"] * c["N"] for c in comp)
    assert 15_000_000 < total < 30_000_000
    params = wl.c3_params(ba)
    steps = 3
    srv = ba.Server(ctx, coalesce_us=300, max_inflight=2)
    states = []
    for c in comp:
        pos, vel, roots = wl.c3_state(ba, c)
        cid, resident = srv.add(pos, vel, rest_roots=roots, params=params)
        assert resident
        states.append((cid, pos, vel, roots))
    reqs = [wl.served_requests(ba, c, steps) for c in comp]
    rng = np.random.default_rng(4)
    order = rng.permutation(np.repeat(np.arange(len(comp)), steps))
    nxt = np.zeros(len(comp), int)
    print("\nserved workload: %d creatures, %d control points, %d requests" % (len(comp), total, order.size),
          flush=True)
    for a in range(0, order.size, 24):
        ids, rr = [], []
        for i in order[a:a + 24]:
            ids.append(states[i][0])
            rr.append(reqs[i][nxt[i]:nxt[i] + 1])
            nxt[i] += 1
        srv.submit(np.array(ids, np.int32), np.concatenate(rr))
    srv.drain(120000)
    st = srv.stats()
    assert st["completed"] == order.size and st["dropped"] == 0 and st["quarantined"] == 0
    mixed = {}
    for i, (c, (cid, pos, vel, roots)) in enumerate(zip(comp, states)):
        P, V = pos.copy(), vel.copy()
        for k in range(steps):
            r = reqs[i][k]
            m = r["model"].reshape(4, 4).T
            sph = r["spheres"].reshape(4, 4)[: r["ncolliders"]]
            op = oracle_params_for(params, c["N"], r["dt"], spheres=[tuple(x) for x in sph], model=m,
                                   root_mode=ba.FTL_ROOT_REST)
            orc.step_f32(op, P, V, roots, nthreads=THREADS)
        gp, gv = srv.download(cid)
        assert_bitexact(gp, P, "creature %d (%d x %d) positions" % (i, c["S"], c["N"]))
        assert_bitexact(gv, V, "creature %d (%d x %d) velocities" % (i, c["S"], c["N"]))
        assert not srv.status(cid)
        if i % 43 == 0:
            mixed[i] = (gp, gv)
    srv.close()
    # the same creatures alone
    for i, (gp, gv) in mixed.items():
        cid0, pos, vel, roots = states[i]
        alone = ba.Server(ctx)
        cid, resident = alone.add(pos, vel, rest_roots=roots, params=params)
        assert resident
        alone.submit(np.full(steps, cid, np.int32), reqs[i])
        alone.drain(60000)
        ap, av = alone.download(cid)
        assert np.array_equal(ap.view(np.uint32), gp.view(np.uint32)), "creature %d alone != mix" % i
        assert np.array_equal(av.view(np.uint32), gv.view(np.uint32)), "creature %d alone != mix" % i
        alone.close()


def test_skinned_workload_4m_strands(ctx):
    """4,194,304 x 32 strands bound to a 100k-triangle mesh skinned by 64 bones
    (dual quaternions), four colliders, two frames: every skinned vertex
    bit-exact, sampled strands bit-exact, invariants on all strands."""
    S, N = 4_194_304, 32
    sc = ba.skinned_scene(S, N, ntris=100_000, nbones=64, radius=1.0, length=0.2, jitter=0.05, seed=5)
    print("\nskinned workload: %d x %d, %d vertices, %d triangles" % (S, N, sc["rest"].shape[0],
                                                                       sc["tris"].shape[0]), flush=True)
    skin = ba.Skin(ctx, sc["rest"], sc["joints"], sc["weights"], sc["tris"], sc["nbones"])
    g = ba.Groom(ctx, sc["pos"], sc["vel"])
    g.bind_skin(skin, sc["tri"], sc["bary"])
    params = wl.c5_params(ba)
    g.set_params(params)
    sample = np.arange(0, S, 97)
    P, V = sc["pos"][sample].copy(), sc["vel"][sample].copy()
    del sc["pos"], sc["vel"]
    for k in range(2):
        t = (k + 1) * float(DT)
        dq = ba.bones(sc, t)
        sph = wl.c5_spheres(t)
        skin.set_bones(dq)
        skin.update()
        g.set_colliders(sph)
        g.step(DT)
        verts = orc.skin_f32(dq, sc["joints"], sc["weights"], sc["rest"])
        assert_bitexact(skin.download(), verts, "skinned vertices frame %d" % k)
        roots = orc.bary_roots_f32(sc["tris"], sc["tri"][sample], sc["bary"][sample], verts)
        op = oracle_params_for(params, N, DT, spheres=[tuple(x) for x in sph], root_mode=ba.FTL_ROOT_REST)
        orc.step_f32(op, P, V, roots, nthreads=THREADS)
    gp, gv = g.download()
    assert_bitexact(gp[sample], P, "skinned sampled positions")
    assert_bitexact(gv[sample], V, "skinned sampled velocities")
    all_roots = orc.bary_roots_f32(sc["tris"], sc["tri"], sc["bary"], verts)[:, :3].astype(np.float64)
    nfree = check_invariants(gp, gv, all_roots, sph, params.length_scale, "skinned")
    assert nfree > 0
    assert g.status() is False
    g.close()
    skin.close()
