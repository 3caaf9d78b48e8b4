"""Groom initialisation on the device (Hair::init_simulation, src/fx/hair.cc:236-328).

Pins of the oracle (CPU):
  * its SplitMix64 draw walks the stream step by step; the library's host
    generator (ftl_scn_reference_groom) jumps to draw j directly: the two agree
    bit for bit, and the layout is the reference's (w = rest length, root 0);
  * glm::simplex (GLM is not part of the reference tree; the oracle restates
    the published webgl-noise algorithm): zero on the skewed lattice, range
    within [-1, 1] with mean ~0, bounded slope, and the period the 289-entry
    permutation ring gives (a wrong modulus or skew constant breaks it);
  * tangents: the root gets 0.15 n, the tip 0.2 (-n + c), the inner points
    s_i c_i with s_i linear in i (exact without curl: 0).
  * the curly tangents' sine and cosine (one FP32 definition, shared with the
    device): against float64 within 1.5 ulp, quadrant points, symmetries.
GPU: positions/velocities and tangents, curled or not, bit for bit.
"""
import numpy as np
import pytest

from oracle import ftl_oracle as orc


def scalp(n=300, seed=0):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((n, 3))
    r /= np.linalg.norm(r, axis=1, keepdims=True)
    return r.astype(np.float32), r.astype(np.float32)


def test_oracle_matches_host_generator():
    import barbu_amd as ba
    roots, normals = scalp()
    for N, L, seed in ((4, 0.5, 0), (16, 0.2, 123), (2, 1.0, 2 ** 63 + 5)):
        p1, v1 = orc.reference_groom_f32(roots, normals, N, L, seed)
        p2, v2 = ba.reference_groom(roots, normals, npoints=N, maxlength=L, seed=seed)
        assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32)) and not v1.any() and not v2.any()
        assert (p1[:, 0, 3] == 0).all() and np.array_equal(p1[:, 0, :3], roots)
        # offsets grow linearly by the per-strand factor r_j in [0.9, 1.1]
        seg = p1[:, 1:, 3].astype(np.float64)
        assert (seg > 0).all() and ((seg / (L / N) > 0.9 - 1e-6) & (seg / (L / N) < 1.1 + 1e-6)).all()


def test_trig_against_float64():
    """The FP32 sine and cosine of the curly tangents: within 1.5 ulp of the
    float64 values over the generator's range (ulps of max(|f|, 1e-3)), exact
    at the quadrant points, odd/even, and pi/2-shifted into each other (a
    wrong reduction constant, quadrant or coefficient sign fails these)."""
    xs = np.concatenate([np.linspace(-200, 200, 40001), np.random.default_rng(3).uniform(-60, 60, 20000),
                         np.arange(-64, 65) * (np.pi / 2)]).astype(np.float32)
    for f, ref in ((orc.sin_f32, np.sin), (orc.cos_f32, np.cos)):
        got = np.array([f(x) for x in xs], np.float64)
        want = ref(xs.astype(np.float64))
        ulp = np.exp2(np.floor(np.log2(np.maximum(np.abs(want), 1e-3))) - 23)
        assert (np.abs(got - want) / ulp).max() <= 1.5
    assert orc.sin_f32(0.0) == 0.0 and orc.cos_f32(0.0) == 1.0
    assert orc.sin_f32(np.float32(np.pi / 2)) == 1.0 and orc.cos_f32(np.float32(np.pi)) == -1.0
    for x in np.random.default_rng(4).uniform(-50, 50, 200).astype(np.float32):
        assert orc.sin_f32(-x) == -orc.sin_f32(x) and orc.cos_f32(-x) == orc.cos_f32(x)
        assert abs(orc.sin_f32(x) - orc.cos_f32(np.float32(x - np.float32(np.pi / 2)))) < 1e-6


def test_simplex_properties():
    s = orc.simplex2_f32
    C0 = (3 - np.sqrt(3)) / 6
    assert s(0.0, 0.0) == 0.0
    rng = np.random.default_rng(1)
    for i, j in rng.integers(-40, 40, (50, 2)):
        v = np.array([i, j], np.float64) - (i + j) * C0       # unskewed lattice point
        assert abs(s(*v)) < 1e-4, (i, j, s(*v))
    pts = rng.uniform(-60, 60, (20000, 2))
    vals = np.array([s(x, y) for x, y in pts])
    assert vals.min() >= -1.0 and vals.max() <= 1.0 and abs(vals.mean()) < 0.02 and vals.std() > 0.2
    eps = 1e-3
    d = np.array([abs(s(x + eps, y) - s(x, y)) for x, y in pts[:4000]]) / eps
    assert d.max() < 10.0   # C1 noise: bounded slope (about 6.3 observed)
    # the permutation ring of 289: a shift by 289 lattice steps along one axis
    shift = np.array([289.0, 0.0]) - 289.0 * C0
    for x, y in pts[:300] * 0.1:
        a, b = s(x, y), s(x + shift[0], y + shift[1])
        assert abs(a - b) < 2e-3, (x, y, a, b)
    # and not by 288 (a sanity check that the test can fail)
    shift = np.array([288.0, 0.0]) - 288.0 * C0
    diff = [abs(s(x, y) - s(x + shift[0], y + shift[1])) for x, y in pts[:300] * 0.1]
    assert max(diff) > 0.1


def test_tangent_structure():
    _, normals = scalp(50, seed=2)
    for N in (2, 3, 4, 16):
        flat = orc.reference_tangents_f32(normals, N, 0.5, curly=False)
        assert np.array_equal(flat[:, 0, :3], np.float32(0.15) * normals)
        assert np.array_equal(flat[:, -1, :3], np.float32(0.2) * (-normals))
        assert not flat[:, 1:-1].any()
        cur = orc.reference_tangents_f32(normals, N, 0.5, curly=True)
        assert np.array_equal(cur[:, 0], flat[:, 0]) and (cur[..., 3] == 0).all()
        tip = cur[:, -1, :3] / np.float32(0.2) + normals       # c = (cos 4 pi k, -0.71 k, sin 2.7 pi k)
        assert (np.abs(tip[:, 0]) <= 1 + 1e-5).all() and (np.abs(tip[:, 2]) <= 1 + 1e-5).all()
        assert (np.abs(tip[:, 1]) <= 0.71 * 1.25 + 1e-5).all()
        if N > 2:  # inner: s_i c_i, |c_i| <= 5.8 * |(10.7, 2.3 * 10 * 1, 20.5)|
            s = 0.1 * np.arange(1, N - 1) / (N - 1) * 0.125 * np.sqrt(0.5)
            c = cur[:, 1:-1, :3] / s[None, :, None]
            assert (np.abs(c[..., 0]) <= 5.8 * 10.7 * (1 + 1e-5)).all()
            assert (np.abs(c[..., 2]) <= 5.8 * 20.5 * (1 + 1e-5)).all()


@pytest.mark.gpu
def test_gpu_groom_init_matches_oracle(ctx):
    import barbu_amd as ba
    roots, normals = scalp(5000, seed=3)
    for N, L, seed in ((4, 0.5, 9), (32, 0.2, 77), (64, 0.3, 1)):
        print("device init N=%d" % N, flush=True)
        g = ba.reference_groom_device(ctx, roots, normals, npoints=N, maxlength=L, seed=seed)
        p, v = g.download()
        P, V = orc.reference_groom_f32(roots, normals, N, L, seed)
        assert np.array_equal(p.view(np.uint32), P.view(np.uint32)) and np.array_equal(v, V)
        g.close()


@pytest.mark.gpu
def test_gpu_tangents_match_oracle(ctx):
    import barbu_amd as ba
    roots, normals = scalp(3000, seed=4)
    for N in (2, 4, 16):
        print("device tangents N=%d" % N, flush=True)
        pos, _ = orc.reference_groom_f32(roots, normals, N, 0.5, 0)
        tris = np.arange(30, dtype=np.int32).reshape(10, 3)
        e = ba.Expand(ctx, roots.shape[0], N, tris, np.zeros_like(pos), ba.random_table(0, 64))
        e.reference_tangents(normals, 0.5, curly=False)
        got = e.tangents()
        assert np.array_equal(got.view(np.uint32), orc.reference_tangents_f32(normals, N, 0.5, False).view(np.uint32))
        e.reference_tangents(normals, 0.5, curly=True)
        got = e.tangents()
        want = orc.reference_tangents_f32(normals, N, 0.5, True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), np.abs(got - want).max()
        e.close()
