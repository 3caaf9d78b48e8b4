"""Pins the CPU oracle (oracle/ftl_oracle.cpp) to what the mathematics and the
reference shader fix, independently of its own code:

  K1  hanging strand: closed-form fixed point and first-iteration corrections
  K2  inextensibility after a step without collisions
  K3  non-penetration of the (last) collider
  K4  root: identity keeps it in place with zero velocity; a transform moves it
  K5  single-point collision: closed-form push-out and reflection
  K6  3-point strand: damping 0.80 and the index of the shifted correction
  K7  FP64 vs FP32 on small random states
  K9  wind value noise: lattice values, continuity, bound, periodicity
  K8  dual-quaternion skinning: identity, single bone = rigid motion, antipodal
      quaternions (in test_skinning_oracle below)
  inversesqrt: exhaustive over all binades (two binades + exact scaling)

A plausible transliteration mistake (wrong sign, dropped term, shifted index,
per-index damping, wrong axis) changes at least one of these results.
"""
import numpy as np
import pytest

from oracle import ftl_oracle as orc

F = np.float32


def step(P, pos, vel, roots=None):
    pos = np.array(pos, np.float32, order="C", copy=True)
    vel = np.array(vel, np.float32, order="C", copy=True)
    orc.step_f32(P, pos, vel, roots)
    return pos, vel


def step64(P, pos, vel, roots=None):
    pos = np.array(pos, np.float64, order="C", copy=True)
    vel = np.array(vel, np.float64, order="C", copy=True)
    orc.step_f64(P, pos, vel, None if roots is None else np.ascontiguousarray(roots, np.float64))
    return pos, vel


# --- inversesqrt -------------------------------------------------------------

def test_rsqrt_exhaustive_two_binades():
    """All floats in [1, 4): |rsqrt - 1/sqrt| <= 2 ulp (GLSL's bound for
    inversesqrt).  Scaling x by 4 halves the result exactly (the seed and the
    Newton steps commute with powers of two), so this covers every normal float."""
    lo, hi = np.float32(1.0).view(np.uint32), np.float32(4.0).view(np.uint32)
    x = np.arange(lo, hi, dtype=np.uint32).view(np.float32)
    y = orc.rsqrt_f32(x)
    exact = 1.0 / np.sqrt(x.astype(np.float64))
    ulp = np.spacing(exact.astype(np.float32)).astype(np.float64)
    err = np.abs(y.astype(np.float64) - exact) / ulp
    assert err.max() <= 2.0, err.max()


def test_rsqrt_power_of_four_scaling():
    rng = np.random.default_rng(0)
    x = rng.uniform(1.0, 4.0, 100000).astype(np.float32)
    for k in (-20, -3, -1, 1, 5, 30):
        s = np.float32(4.0) ** k
        a = orc.rsqrt_f32(x * s)
        b = orc.rsqrt_f32(x) / np.float32(2.0) ** k
        assert np.array_equal(a, b.astype(np.float32))


# --- K1 hanging strand ---------------------------------------------------------

def hanging(N, a=0.125, top=2.0):
    pos = np.zeros((1, N, 4), np.float32)
    for i in range(N):
        pos[0, i] = (0.0, top - i * a, 0.0, 0.0 if i == 0 else a)
    return pos, np.zeros_like(pos)


@pytest.mark.parametrize("N", [2, 4, 8, 16, 33])
def test_k1_hanging_strand_fixed_point(N):
    pos, vel = hanging(N)
    dt = F(1 / 60)
    P = orc.make_params(N, dt, length_scale=1.0, damping=0.8, force_scale=20.0, spheres=None)
    p, v = step(P, pos, vel)
    assert np.abs(p - pos).max() < 4e-7  # restored to rounding
    assert np.abs(v).max() < 1e-6        # iterations 2..8 correct nothing


@pytest.mark.parametrize("N", [3, 4, 9])
def test_k1_first_iteration_corrections(N):
    """K = 1: every point is moved back up by h = |F| dt^2; u_i = delta*(0,h,0)
    for 1 <= i <= N-2 and u_{N-1} = (0,h,0).  Pins the force (20 x gravity in
    the shipped shader), the damping and the shift."""
    pos, vel = hanging(N)
    dt = F(1 / 60)
    P = orc.make_params(N, dt, iterations=1, length_scale=1.0, damping=0.8, force_scale=20.0, spheres=None)
    p, v = step(P, pos, vel)
    h = 196.2 * (1 / 60) ** 2
    for i in range(1, N - 1):
        np.testing.assert_allclose(v[0, i, :3], (0.0, 0.8 * h, 0.0), atol=2e-6)
    np.testing.assert_allclose(v[0, N - 1, :3], (0.0, h, 0.0), atol=2e-6)
    np.testing.assert_array_equal(v[0, 0], 0.0)


# --- K6 three-point strand -------------------------------------------------

def test_k6_damping_index_and_value():
    """x0=(0,0,0) root, x1=(0,-1,0), x2=(0,-2,0), rest lengths 1, no force,
    dt=1, v2=(1,0,0): after predict p2=(1,-2,0).  The sweep leaves p1, moves
    p2 to p1 + (1,-1,0)/sqrt2, so c1 = 0 and c2 = (1/sqrt2 - 1, 1 - 1/sqrt2, 0).
    With one iteration u1 = 0.80 c2 (the next point's correction, lane 0's
    damping) and u2 = c2."""
    pos = np.array([[[0, 0, 0, 0], [0, -1, 0, 1], [0, -2, 0, 1]]], np.float32)
    vel = np.zeros_like(pos)
    vel[0, 2, :3] = (1, 0, 0)
    P = orc.make_params(3, 1.0, iterations=1, length_scale=1.0, damping=0.8, force_scale=0.0, spheres=None)
    p, v = step(P, pos, vel)
    r = 1 / np.sqrt(2.0)
    c2 = np.array([r - 1, 1 - r, 0.0])
    np.testing.assert_allclose(p[0, 2, :3], (r, -1 - r, 0.0), atol=3e-7)
    np.testing.assert_allclose(v[0, 1, :3], 0.8 * c2, atol=3e-7)
    np.testing.assert_allclose(v[0, 2, :3], c2, atol=3e-7)
    # the shipped constant is lane 0's value of mix(0.80, 0.94, smoothstep(0,1,t)*t)
    t = 0.0
    assert F(0.80 + (0.94 - 0.80) * (t * t * (3 - 2 * t)) * t) == F(0.80)
    # later iterations find the constraints satisfied
    P8 = orc.make_params(3, 1.0, iterations=8, length_scale=1.0, damping=0.8, force_scale=0.0, spheres=None)
    p8, v8 = step(P8, pos, vel)
    assert np.abs(v8[0, 1:, :3]).max() < 1e-6
    np.testing.assert_allclose(p8[0, 2, :3], (r, -1 - r, 0.0), atol=3e-7)


def test_length_scale_stretches_segments():
    pos, vel = hanging(5)
    P = orc.make_params(5, F(1 / 60), length_scale=1.45, force_scale=0.0, spheres=None)
    p, _ = step(P, pos, vel)
    seg = np.linalg.norm(np.diff(p[0, :, :3].astype(np.float64), axis=0), axis=1)
    np.testing.assert_allclose(seg, 1.45 * 0.125, rtol=1e-6)
    assert np.all(p[0, :, 3] == pos[0, :, 3])  # rest lengths pass through


# --- K5 single-point collision ---------------------------------------------------

@pytest.mark.parametrize("axis", [0, 1, 2])
def test_k5_push_out_and_reflect(axis):
    """Root at 2 e_a, rest length 1, sphere r = 1.5 at the origin: the point at
    1 e_a is inside.  Every iteration the sweep pulls it to 1 e_a and the
    collider pushes it to 1.5 e_a.  Last iteration: correction -0.5 e_a,
    reflected about n = e_a -> +0.5 e_a."""
    e = np.zeros(3, np.float32)
    e[axis] = 1.0
    pos = np.zeros((1, 2, 4), np.float32)
    pos[0, 0, :3] = 2 * e
    pos[0, 1, :3] = 1 * e
    pos[0, 1, 3] = 1.0
    vel = np.zeros_like(pos)
    for K, u in ((1, 0.0), (2, 0.5), (8, 0.5)):
        P = orc.make_params(2, F(1 / 60), iterations=K, length_scale=1.0, force_scale=0.0,
                            spheres=[(0, 0, 0, 1.5)])
        p, v = step(P, pos, vel)
        np.testing.assert_allclose(p[0, 1, :3], 1.5 * e, atol=2e-7)
        np.testing.assert_allclose(v[0, 1, :3], u * e, atol=2e-7)
        np.testing.assert_array_equal(p[0, 0, :3], pos[0, 0, :3])


def test_collision_reflection_closed_form():
    """Two iterations on a 2-point strand against one sphere: iteration 1 ends
    at q (its contact point); iteration 2's sweep moves q to s = root + l (q -
    root)/|q - root| and the collider pushes s to r s/|s| (n = s/|s|).  The
    stored velocity is the reflection of the correction s - q about n:
    u = (s - q) - 2 ((s - q).n) n."""
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(40):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pos = np.zeros((1, 2, 4), np.float32)
        pos[0, 0, :3] = 2.0 * d
        pos[0, 1, :3] = 2.0 * d - 1.0 * (d + 0.3 * rng.normal(size=3))
        pos[0, 1, 3] = 1.0
        vel = np.zeros_like(pos)
        kw = dict(length_scale=1.0, force_scale=0.0, spheres=[(0, 0, 0, 1.5)])
        q, _ = step(orc.make_params(2, F(1 / 60), iterations=1, **kw), pos, vel)
        p, v = step(orc.make_params(2, F(1 / 60), iterations=2, **kw), pos, vel)
        root = pos[0, 0, :3].astype(np.float64)
        q = q[0, 1, :3].astype(np.float64)
        s = root + (q - root) / np.linalg.norm(q - root)
        if np.linalg.norm(s) >= 1.5 - 1e-4:
            continue  # iteration 2 does not collide
        n = s / np.linalg.norm(s)
        c = s - q
        np.testing.assert_allclose(p[0, 1, :3], 1.5 * n, atol=2e-6)
        np.testing.assert_allclose(v[0, 1, :3], c - 2 * np.dot(c, n) * n, atol=2e-6)
        checked += 1
    assert checked >= 10


# --- K2 / K3 / K4 invariants ----------------------------------------------------

def random_strands(S, N, seed, length=0.3):
    rng = np.random.default_rng(seed)
    n = rng.normal(size=(S, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    pos = np.zeros((S, N, 4), np.float32)
    seg = length / (N - 1)
    for i in range(N):
        pos[:, i, :3] = n * (1.0 + i * seg) + rng.normal(scale=0.01, size=(S, 3)) * (i > 0)
        pos[:, i, 3] = 0.0 if i == 0 else seg
    vel = np.zeros_like(pos)
    vel[:, 1:, :3] = rng.normal(scale=0.5, size=(S, N - 1, 3))
    return pos, vel


def test_k2_inextensible_without_colliders():
    pos, vel = random_strands(200, 16, 1)
    P = orc.make_params(16, F(1 / 60), length_scale=1.3, force_scale=1.0, spheres=None)
    p, _ = step(P, pos, vel)
    seg = np.linalg.norm(np.diff(p[:, :, :3].astype(np.float64), axis=1), axis=-1)
    np.testing.assert_allclose(seg, 1.3 * pos[:, 1:, 3].astype(np.float64), rtol=3e-5)


def test_k3_non_penetration():
    pos, vel = random_strands(300, 12, 2)
    c, r = np.array([0.3, 0.9, 0.0]), 0.6
    P = orc.make_params(12, F(1 / 60), force_scale=20.0, spheres=[(0, 0, 0, 1.0), (*c, r)])
    p, _ = step(P, pos, vel)
    d = np.linalg.norm(p[:, 1:, :3].astype(np.float64) - c, axis=-1)
    assert d.min() >= r * (1 - 1e-6)


def test_k4_root_identity_and_transform():
    pos, vel = random_strands(50, 6, 3)
    P = orc.make_params(6, F(1 / 60), force_scale=20.0)
    p, v = step(P, pos, vel)
    np.testing.assert_array_equal(p[:, 0, :3], pos[:, 0, :3])
    np.testing.assert_array_equal(v[:, 0], 0.0)
    # rest roots moved by M: p0 = M r, u0 = p0 - x0
    a = 0.4
    M = np.array([[np.cos(a), 0, np.sin(a), 0.1], [0, 1, 0, -0.2], [-np.sin(a), 0, np.cos(a), 0.3], [0, 0, 0, 1]],
                 np.float32)
    roots = np.ascontiguousarray(pos[:, 0, :].copy())
    P2 = orc.make_params(6, F(1 / 60), force_scale=20.0, root_mode=1, model=M)
    p2, v2 = step(P2, pos, vel, roots)
    want = roots[:, :3].astype(np.float64) @ M[:3, :3].T.astype(np.float64) + M[:3, 3]
    np.testing.assert_allclose(p2[:, 0, :3], want, atol=1e-6)
    np.testing.assert_allclose(v2[:, 0, :3], want - pos[:, 0, :3], atol=1e-6)


def test_dftl_velocity_mode_and_clamp():
    """Eq. 9: v = ((p - x) - u)/dt, then |v| <= max_velocity."""
    pos, vel = random_strands(100, 8, 4)
    P = orc.make_params(8, F(1 / 60), velocity_mode=1, max_velocity=1.5, force_scale=1.0, spheres=None)
    p, v = step(P, pos, vel)
    sp = np.linalg.norm(v[:, :, :3].astype(np.float64), axis=-1)
    assert sp.max() <= 1.5 * (1 + 1e-5)
    assert sp.max() > 1.0  # the clamp is active on this input
    # unclamped: reconstruct u from the shipped mode on the same input
    P0 = orc.make_params(8, F(1 / 60), velocity_mode=0, force_scale=1.0, spheres=None)
    p0, u = step(P0, pos, vel)
    np.testing.assert_array_equal(p0, p)
    Pb = orc.make_params(8, F(1 / 60), velocity_mode=1, max_velocity=1e9, force_scale=1.0, spheres=None)
    _, vb = step(Pb, pos, vel)
    want = ((p0[:, :, :3].astype(np.float64) - pos[:, :, :3]) - u[:, :, :3]) * 60.0
    np.testing.assert_allclose(vb[:, :, :3], want, rtol=2e-4, atol=2e-4)


def test_colliders_applied_in_order():
    """A point inside two overlapping spheres A and B: each collider pushes it
    out in turn, so it ends on the surface of the sphere applied last, for
    either order."""
    pos = np.zeros((1, 2, 4), np.float32)
    pos[0, 0, :3] = (0, 3, 0)
    pos[0, 1, :3] = (0, 2.1, 0)
    pos[0, 1, 3] = 0.9
    vel = np.zeros_like(pos)
    A, B = (0, 2.2, 0, 0.3), (0.2, 1.8, 0, 0.5)
    for order in ([A, B], [B, A]):
        P = orc.make_params(2, F(1 / 60), iterations=1, length_scale=1.0, force_scale=0.0, spheres=order)
        p, _ = step(P, pos, vel)
        last = order[-1]
        assert abs(np.linalg.norm(p[0, 1, :3] - np.array(last[:3])) - last[3]) < 1e-6


# --- K7 FP64 vs FP32 -------------------------------------------------------------------

@pytest.mark.parametrize("N", [2, 3, 4])
def test_k7_fp64_vs_fp32_tiny(N):
    pos, vel = random_strands(4, N, 10 + N)
    P = orc.make_params(N, F(1 / 60), force_scale=20.0, spheres=[(0, 0, 0, 1.0)])
    p32, v32 = step(P, pos, vel)
    p64, v64 = step64(P, pos, vel)
    L = float(pos[0, 1:, 3].sum())
    assert np.abs(p32 - p64).max() <= 1e-5 * L
    assert np.abs(v32 - v64).max() <= 1e-5 * L


# --- K9 wind ---------------------------------------------------------------------------

def wind_params(table, freq=1.0, off=(0, 0, 0)):
    return orc.make_params(4, F(1 / 60), wind_table=table, wind_frequency=freq, wind_offset=off)


def synthetic_table(G=16, seed=0):
    rng = np.random.default_rng(seed)
    t = np.zeros((G ** 3, 4), np.float32)
    t[:, :3] = rng.uniform(-2, 2, size=(G ** 3, 3))
    return t


def test_k9_wind_lattice_values_and_continuity():
    G = 16
    t = synthetic_table(G)
    P = wind_params(t)
    rng = np.random.default_rng(1)
    pts = rng.integers(-40, 40, size=(50, 3)).astype(np.float32)
    w = orc.wind_f32(P, pts)
    idx = np.mod(pts.astype(int), G)
    want = t[(idx[:, 2] * G + idx[:, 1]) * G + idx[:, 0], :3]
    np.testing.assert_array_equal(w, want)
    # continuity across a cell face and boundedness
    x = np.array([[3.0, 5.5, -2.25]], np.float32)
    lo, hi = x.copy(), x.copy()
    lo[0, 0] = np.nextafter(np.float32(3.0), np.float32(0))
    w0, w1, w2 = orc.wind_f32(P, lo), orc.wind_f32(P, x), orc.wind_f32(P, hi)
    assert np.abs(w0 - w1).max() < 1e-5 and np.array_equal(w1, w2)
    many = orc.wind_f32(P, rng.uniform(-10, 10, size=(500, 3)).astype(np.float32))
    assert np.abs(many).max() <= 2.0
    # periodic over G cells
    y = rng.uniform(-4, 4, size=(20, 3)).astype(np.float32)
    np.testing.assert_allclose(orc.wind_f32(P, y + G), orc.wind_f32(P, y), atol=1e-5)


def test_wind_enters_the_force():
    """With a constant wind w (all lattice entries equal) the predict step is
    the no-wind predict with force + w."""
    G = 16
    t = np.zeros((G ** 3, 4), np.float32)
    t[:, :3] = (0.7, -0.2, 0.4)
    pos, vel = random_strands(30, 5, 6)
    Pw = orc.make_params(5, F(1 / 60), iterations=1, force_scale=1.0, gravity=(0, -9.81, 0), spheres=None,
                         wind_table=t, wind_frequency=2.0)
    Pn = orc.make_params(5, F(1 / 60), iterations=1, force_scale=1.0, gravity=(0.7, F(-9.81) + F(-0.2), 0.4),
                         spheres=None)
    a = step(Pw, pos, vel)
    b = step(Pn, pos, vel)
    np.testing.assert_array_equal(a[0], b[0])


# --- K8 dual-quaternion skinning ---------------------------------------------------------

def quat_to_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def dq_from(q, t):
    x, y, z, w = q
    tx, ty, tz = t
    # dual = 0.5 * (t, 0) * q
    d = 0.5 * np.array([w * tx + ty * z - tz * y, w * ty + tz * x - tx * z, w * tz + tx * y - ty * x,
                        -(tx * x + ty * y + tz * z)])
    return np.concatenate([q, d]).astype(np.float32)


def test_k8_single_bone_is_rigid_motion():
    rng = np.random.default_rng(5)
    nb = 4
    qs, ts = [], []
    for _ in range(nb):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        qs.append(q)
        ts.append(rng.normal(size=3))
    dq = np.stack([dq_from(q, t) for q, t in zip(qs, ts)])
    rest = np.zeros((nb, 4), np.float32)
    rest[:, :3] = rng.normal(size=(nb, 3))
    joints = np.array([[b, (b + 1) % nb, (b + 2) % nb, (b + 3) % nb] for b in range(nb)], np.int32)
    weights = np.tile(np.array([1.0, 0.0, 0.0, 0.0], np.float32), (nb, 1))
    out = orc.skin_f32(dq, joints, weights, rest)
    for b in range(nb):
        want = quat_to_mat(qs[b]) @ rest[b, :3].astype(np.float64) + ts[b]
        np.testing.assert_allclose(out[b, :3], want, atol=2e-6)


def test_k8_identity_and_antipodal():
    rng = np.random.default_rng(6)
    ident = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 0], np.float32), (4, 1))
    rest = np.zeros((10, 4), np.float32)
    rest[:, :3] = rng.normal(size=(10, 3))
    joints = np.tile(np.arange(4, dtype=np.int32), (10, 1))
    w = rng.uniform(0.1, 1.0, size=(10, 4))
    w /= w.sum(axis=1, keepdims=True)
    weights = w.astype(np.float32)
    np.testing.assert_allclose(orc.skin_f32(ident, joints, weights, rest)[:, :3], rest[:, :3], atol=1e-6)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    t = rng.normal(size=3)
    a = np.stack([dq_from(q, t)] * 4)
    b = a.copy()
    b[1] = -b[1]  # antipodal copy of the same rigid motion
    b[2] = -b[2]
    oa = orc.skin_f32(a, joints, weights, rest)
    ob = orc.skin_f32(b, joints, weights, rest)
    np.testing.assert_allclose(oa[:, :3], ob[:, :3], atol=2e-6)
    want = rest[:, :3].astype(np.float64) @ quat_to_mat(q).T + t
    np.testing.assert_allclose(oa[:, :3], want, atol=2e-6)
    # a zero first weight leaves the vertex unchanged (apply_skinning early out)
    w0 = weights.copy()
    w0[:, 0] = 0.0
    np.testing.assert_array_equal(orc.skin_f32(a, joints, w0, rest), rest)


def test_bary_roots():
    verts = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]], np.float32)
    tris = np.array([[0, 1, 2]], np.int32)
    bary = np.array([[0, 0], [1, 0], [0, 1], [0.25, 0.5]], np.float32)
    r = orc.bary_roots_f32(tris, np.zeros(4, np.int32), bary, verts)
    np.testing.assert_allclose(r[:, :3], [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.25, 0.5, 0]], atol=1e-7)


def test_zero_length_segment_is_not_finite():
    """A point on top of its leader: normalize(0) has no direction.  GLSL gives
    inversesqrt(0) = +inf and 0 * inf = NaN; the FP32 inversesqrt here also
    yields NaN for 0 (the Newton step computes 0 * inf), so the strand from
    that point on becomes NaN, as in the reference (the step reports it through
    the groom's non-finite flag)."""
    assert np.isnan(orc.rsqrt_f32(np.zeros(1, np.float32))[0])
    pos = np.array([[[0, 0, 0, 0], [0, -0.1, 0, 0.1], [0, -0.1, 0, 0.1], [0, -0.3, 0, 0.1]]], np.float32)
    P = orc.make_params(4, F(1 / 60), iterations=1, length_scale=1.0, force_scale=0.0, spheres=None)
    p, v = step(P, pos, np.zeros_like(pos))
    assert np.isfinite(p[0, :2]).all()
    assert np.isnan(p[0, 2:, :3]).all()
