"""Marschner lookup tables (src/fx/marschner.cc:35-66; cs_marschner_m.glsl,
cs_marschner_n.glsl with inc_marschner_n.glsl, inc_solver.glsl, inc_fresnel.glsl).

Pins of the oracle (CPU):
  * the FP32 restatement agrees with an FP64 evaluation of the same formulas;
  * M depends on the half angle, so it is symmetric under (x, y) -> (y, x),
    bit for bit; its lobes are normalised gaussians in degrees, peaking at the
    shifts (1, -0.5, -1.5) ar with height 1 / (sqrt(2 pi) width);
  * the cubic solver: every root it returns is a root (residual; Cardano's
    branch within the error of the reference's pow(x, 0.333) cube root), and
    the counts follow the discriminant;
  * Np(0) is the Fresnel term at gamma = -phi/2 times |cos gamma| / 4 (the
    linear root of its polynomial and 1 / |2 d/dgamma|);
  * Fresnel: no reflection without an interface (eta = 1), total reflection
    past the critical angle;
  * N terms lie in [0, 1];
  * the FP32 sin, cos, asin, acos, exp, log and pow the tables use (one
    definition, shared with the device): within 2 ulp (pow 3) of float64 on the
    tables' ranges, exact at their special points, with their symmetries.
GPU: the device tables against the FP32 oracle, bit for bit on every texel
(the well- and ill-conditioned alike: both evaluate the same operations).
"""
import numpy as np
import pytest

from oracle import ftl_oracle as orc


def _ulps(got, want):
    return np.abs(got - want) / np.exp2(np.floor(np.log2(np.maximum(np.abs(want), 1e-30))) - 23)


def test_shared_transcendentals_against_float64():
    rng = np.random.default_rng(5)
    f = np.vectorize(orc.math_f32)
    x = np.concatenate([np.linspace(-1, 1, 20001), rng.uniform(-1, 1, 20000)]).astype(np.float32)
    x64 = x.astype(np.float64)
    assert _ulps(f("asin", x), np.arcsin(x64)).max() <= 2.0
    assert _ulps(f("acos", x), np.arccos(x64)).max() <= 2.0
    e = rng.uniform(-100, 88, 20000).astype(np.float32)
    assert _ulps(f("exp", e), np.exp(e.astype(np.float64))).max() <= 1.0
    lg = np.exp2(rng.uniform(-30, 30, 20000)).astype(np.float32)
    want = np.log(lg.astype(np.float64))
    assert (np.abs(f("log", lg) - want) / np.exp2(np.floor(np.log2(np.maximum(np.abs(want), 1e-3))) - 23)).max() <= 2.0
    b = rng.uniform(0, 40, 20000).astype(np.float32)
    assert _ulps(f("pow", b, np.float32(0.333)), b.astype(np.float64) ** np.float64(np.float32(0.333))).max() <= 3.0
    # special points and symmetries
    assert orc.math_f32("asin", 0) == 0 and orc.math_f32("asin", 1) == np.float32(np.pi / 2)
    assert orc.math_f32("acos", 1) == 0 and orc.math_f32("acos", -1) == np.float32(np.pi)
    assert orc.math_f32("acos", 0) == np.float32(np.pi / 2) and orc.math_f32("exp", 0) == 1
    assert orc.math_f32("log", 1) == 0 and orc.math_f32("pow", 0, 0.333) == 0 and orc.math_f32("pow", 1, 0.333) == 1
    assert np.isnan(orc.math_f32("asin", 1.5)) and np.isnan(orc.math_f32("log", -1.0))
    assert orc.math_f32("log", 0.0) == -np.inf and orc.math_f32("exp", -200) == 0 and orc.math_f32("exp", 100) == np.inf
    for v in rng.uniform(0, 1, 200).astype(np.float32):
        assert orc.math_f32("asin", -v) == -orc.math_f32("asin", v)
        assert abs(orc.math_f32("acos", v) + orc.math_f32("asin", v) - np.float32(np.pi / 2)) < 3e-7


def test_fp32_vs_fp64():
    M, N = orc.marschner_luts()
    M64, N64 = orc.marschner_luts(dtype=np.float64)
    assert np.isfinite(M).all() and np.abs(M - M64).max() < 1e-6
    fin = np.isfinite(N) & np.isfinite(N64)
    assert fin.mean() > 0.99
    err = np.abs(N[fin] - N64[fin])
    assert np.quantile(err, 0.999) < 1e-4 and err.max() < 0.05   # a few ill-conditioned texels


def test_m_symmetry_and_lobes():
    for ar, br in ((-5.0, 5.0), (3.0, 8.0)):
        M, _ = orc.marschner_luts(ar=ar, br=br)
        assert np.array_equal(M, np.transpose(M, (1, 0, 2)))
        for k, (shift, width) in enumerate(((1.0, 1.0), (-0.5, 0.5), (-1.5, 2.0))):
            peak = 1.0 / (np.sqrt(2 * np.pi) * abs(width * br))
            assert M[..., k].max() <= peak * (1 + 1e-5)
            # the texel of largest value has a half angle near the shift
            y, x = np.unravel_index(np.argmax(M[..., k]), M.shape[:2])
            th = 0.5 * (np.arcsin(2 * x / 128 - 1) + np.arcsin(2 * y / 128 - 1))
            assert abs(np.degrees(th) - shift * ar) < 1.5
        # the fourth channel is cos(theta_d), theta_d the half difference
        y, x = np.mgrid[0:128, 0:128]
        td = 0.5 * (np.arcsin(2 * x / 128 - 1) - np.arcsin(2 * y / 128 - 1))
        assert np.abs(M[..., 3] - np.cos(td)).max() < 2e-6


def test_cubic_solver():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        A, C, D = rng.uniform(-3, 3, 3)
        if abs(A) < 0.1:   # the normalised form divides by A: keep it well conditioned
            continue
        r = orc.cubic_f64(A, 0.0, C, D)
        # three real roots come from the trigonometric form, whose second and
        # third use the reference's Pi() = 3.141564 (off by 2.9e-5); one from
        # Cardano with the reference's pow(x, 0.333) for the cube root, off by
        # |ln x| (1/3 - 0.333) relative
        for k, g in enumerate(r):
            tol = (1e-9 if k == 0 else 1e-3) if len(r) == 3 else 2e-2
            assert abs(A * g ** 3 + C * g + D) < tol * (1 + abs(D) + abs(A) * abs(g) ** 3), (A, C, D, g)
        disc = (3 * (C / A) / 9) ** 3 + (-27 * (D / A) / 54) ** 2   # Q^3 + R^2 of the normalised cubic
        if abs(disc) > 1e-6 and abs(D / A) > 1e-5:
            assert len(r) == (1 if disc > 0 else 3)
    # |A| below the solver's epsilon: the linear case C g + D = 0
    assert np.allclose(orc.cubic_f64(1e-8, 0.0, -2.0, -0.7), [-0.35])


def test_np0_closed_form():
    eta, absorption = 1.55, 0.2
    for ct in (0.3, 0.7, 0.95):
        ep = np.sqrt(eta * eta - (1 - ct * ct)) / ct
        epa = eta * eta / ep
        c = np.arcsin(1 / ep)
        for phi in np.linspace(0.05, 3.1, 12):
            g = -phi / 2
            want = orc.fresnel_f32(ep, epa, np.cos(g), np.sin(g)) * abs(np.cos(g)) / 4
            assert abs(orc.marschner_np_f32(0, absorption, ep, epa, c, phi) - min(want, 1.0)) < 2e-6


def test_fresnel_limits():
    for a in np.linspace(0.0, 1.5, 7):
        assert orc.fresnel_f32(1.0, 1.0, np.cos(a), np.sin(a)) < 1e-12
    # from inside a denser fibre (eta < 1 seen from the fibre) past the critical angle
    assert orc.fresnel_f32(1 / 1.55, 1 / 1.55, np.cos(1.2), np.sin(1.2)) == 1.0


def test_n_range():
    _, N = orc.marschner_luts()
    assert (N[..., :3] >= 0).all() and (N[..., :3] <= 1).all() and (N[..., 3] == 1).all()
    # the generic branch gives TRT = 0 at the reference settings wherever
    # cos(theta_d) > 0: its cubic's root lies outside [-pi/2, pi/2], where the
    # Fresnel term saturates at 1
    assert not N[65:, :, 2].any()


@pytest.mark.gpu
@pytest.mark.parametrize("res,eta,absorption,ar,br", [(128, 1.55, 0.2, -5.0, 5.0), (64, 1.4, 0.5, 2.0, 7.0),
                                                       (256, 1.8, 0.05, -3.0, 4.0)])
def test_gpu_matches_oracle(ctx, res, eta, absorption, ar, br):
    import barbu_amd as ba
    print("marschner res=%d eta=%g" % (res, eta), flush=True)
    M, N = ba.marschner_luts(ctx, ba.marschner_params(eta=eta, absorption=absorption, ar=ar, br=br), res)
    Mo, No = orc.marschner_luts(res, eta, absorption, ar, br)
    M64, N64 = orc.marschner_luts(res, eta, absorption, ar, br, dtype=np.float64)
    assert np.array_equal(M.view(np.uint32), Mo.view(np.uint32)), np.abs(M - Mo).max()
    # bit for bit, NaN texels (a NaN's sign and payload are the hardware's) at the same places
    nan = np.isnan(No)
    assert (np.isnan(N) == nan).all()
    assert np.array_equal(N[~nan].view(np.uint32), No[~nan].view(np.uint32)), np.nanmax(np.abs(N - No))
    # (the conditioning of the formulas, for the record: FP32 against FP64)
    well = np.isfinite(No) & np.isfinite(N64) & (np.abs(No - N64) < 1e-4)
    assert well.mean() > 0.98, well.mean()
