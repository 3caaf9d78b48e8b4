"""The skinning stage that feeds skinned roots (src/shaders/shared/
inc_skinning.glsl:15-129; skeleton_controller.cc:248-265): dual-quaternion
and linear blending of positions and normals, kNoJoint, joint matrices to dual
quaternions, and each strand's root frame.

CPU tests pin the oracle by what the mathematics fixes (glm, which the
reference uses for the matrix-to-dual-quaternion step, is not in the reference
tree: parity against glm itself is unpinned):
  * one bone, linear blending = the rigid transform (positions and normals);
  * kNoJoint = the identity;
  * two bones, linear blending = sum_k w_k M_k v (FP64 closed form);
  * DQ(M) v = M v for rigid M, in all four branches of the conversion;
  * skinned normals of rigid bones stay unit, and equal R n;
  * antipodal dual quaternions give the same result;
  * root frames: the step's roots and a unit normal.
GPU tests compare the kernels with the oracle bit for bit."""
import numpy as np
import pytest

import barbu_amd as ba
from oracle import ftl_oracle as orc
from tests.helpers import assert_bitexact, oracle_params_for

DT = np.float32(1.0 / 60.0)


def rot(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def rigid_rows(R, t):
    """The 3x4 rows [R | t] as float32 (12 floats)."""
    m = np.zeros((3, 4))
    m[:, :3] = R
    m[:, 3] = t
    return m.astype(np.float32)


def random_rigid(rng, max_angle=np.pi):
    return rigid_rows(rot(rng.normal(size=3), rng.uniform(-max_angle, max_angle)), rng.normal(size=3))


def dq_to_rows(dq):
    """Rows [R | t] of a unit dual quaternion (real xyzw, dual xyzw), float64."""
    x, y, z, w = (float(c) for c in dq[:4])
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    d = np.asarray(dq[4:], np.float64)
    # t = 2 d q*: vector part of 2 (d x q_conj)
    qv = np.array([x, y, z])
    t = 2.0 * (-d[3] * qv + w * d[:3] - np.cross(d[:3], qv))
    return R, t


def mesh(rng, nverts, nbones):
    rest = np.zeros((nverts, 4), np.float32)
    rest[:, :3] = rng.normal(size=(nverts, 3))
    rest[:, 3] = rng.uniform(size=nverts)
    n = rng.normal(size=(nverts, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    normals = np.zeros((nverts, 4), np.float32)
    normals[:, :3] = n
    joints = rng.integers(0, nbones, size=(nverts, 4)).astype(np.int32)
    w = rng.dirichlet(np.ones(4), size=nverts)
    weights = np.zeros((nverts, 4), np.float32)
    weights[:, :3] = w[:, :3]
    return rest, normals, joints, weights


# --- CPU: the oracle pinned ------------------------------------------------------

def test_lbs_one_bone_is_the_rigid_transform():
    rng = np.random.default_rng(1)
    rest, normals, _, _ = mesh(rng, 500, 1)
    M = random_rigid(rng)
    joints = np.zeros((500, 4), np.int32)
    weights = np.zeros((500, 4), np.float32)
    weights[:, 0] = 1.0
    out, nout = orc.skin_full(orc.SKIN_LBS, M, joints, weights, rest, normals)
    Md = M.astype(np.float64)
    want = rest[:, :3].astype(np.float64) @ Md[:, :3].T + Md[:, 3]
    assert np.abs(out[:, :3] - want).max() < 2e-6 * max(1.0, np.abs(want).max())
    assert np.array_equal(out[:, 3], rest[:, 3])
    wn = normals[:, :3].astype(np.float64) @ Md[:, :3].T
    assert np.abs(nout[:, :3] - wn).max() < 2e-6
    assert np.abs(np.linalg.norm(nout[:, :3], axis=1) - 1.0).max() < 2e-6


def test_lbs_no_joint_is_the_identity():
    rng = np.random.default_rng(2)
    rest, normals, _, weights = mesh(rng, 400, 3)
    joints = np.full((400, 4), 0xFF, np.int32)
    bones = np.stack([random_rigid(rng) for _ in range(3)])  # never read
    out, nout = orc.skin_full(orc.SKIN_LBS, bones, joints, weights, rest, normals)
    # sum of the four weights is 1 up to rounding
    assert np.abs(out[:, :3] - rest[:, :3]).max() < 4e-7 * np.abs(rest[:, :3]).max() * 4
    assert np.abs(nout[:, :3] - normals[:, :3]).max() < 2e-6
    # mixed: kNoJoint in one slot blends the identity with the bone
    joints2 = np.zeros((400, 4), np.int32)
    joints2[:, 1] = 0xFF
    out2, _ = orc.skin_full(orc.SKIN_LBS, bones, joints2, weights, rest, normals)
    w = weights.astype(np.float64)
    w3 = 1.0 - w[:, :3].sum(1)
    B0 = bones[0].astype(np.float64)
    v = rest[:, :3].astype(np.float64)
    mv = v @ B0[:, :3].T + B0[:, 3]
    want = (w[:, 0] + w[:, 2] + w3)[:, None] * mv + w[:, 1:2] * v
    assert np.abs(out2[:, :3] - want).max() < 1e-5


def test_lbs_blend_matches_fp64_closed_form():
    rng = np.random.default_rng(3)
    nb = 6
    rest, normals, joints, weights = mesh(rng, 2000, nb)
    bones = np.stack([random_rigid(rng) for _ in range(nb)])
    out, nout = orc.skin_full(orc.SKIN_LBS, bones, joints, weights, rest, normals)
    out64, nout64 = orc.skin_full(orc.SKIN_LBS, bones, joints, weights, rest, normals, dtype=np.float64)
    w = weights.astype(np.float64)
    w4 = np.concatenate([w[:, :3], (1.0 - w[:, :3].sum(1))[:, None]], axis=1)
    B = bones.astype(np.float64)
    v = rest[:, :3].astype(np.float64)
    n = normals[:, :3].astype(np.float64)
    want = np.zeros_like(v)
    wn = np.zeros_like(n)
    for k in range(4):
        Mk = B[joints[:, k]]
        want += w4[:, k:k + 1] * (np.einsum("nij,nj->ni", Mk[:, :, :3], v) + Mk[:, :, 3])
        wn += w4[:, k:k + 1] * np.einsum("nij,nj->ni", Mk[:, :, :3], n)
    assert np.abs(out[:, :3] - want).max() < 1e-5
    assert np.abs(nout[:, :3] - wn).max() < 1e-5
    assert np.abs(out64[:, :3] - want).max() < 1e-6
    # weights below Epsilon leave the vertex and its normal untouched
    weights[:5, 0] = 0.0
    o2, n2 = orc.skin_full(orc.SKIN_LBS, bones, joints, weights, rest, normals)
    assert np.array_equal(o2[:5], rest[:5]) and np.array_equal(n2[:5], normals[:5])


@pytest.mark.parametrize("axis,angle", [((0.3, -0.5, 0.8), 0.7),   # trace > 0
                                        ((1, 0.02, 0.01), 3.1),     # R00 largest
                                        ((0.01, 1, 0.02), 3.0),     # R11 largest
                                        ((0.02, 0.01, 1), 3.12),    # R22 largest
                                        ((0, 0, 1), np.pi)])        # exact half turn
def test_dualquat_from_matrix_reproduces_the_transform(axis, angle):
    """DQ(M) v = M v: the library's conversion, applied by the oracle's
    dual-quaternion blending with one bone, is the rigid transform M."""
    rng = np.random.default_rng(4)
    R = rot(axis, angle)
    t = np.array([0.5, -1.25, 2.0])
    M = rigid_rows(R, t)
    dq = ba.dualquat_from_matrix(M)
    assert abs(np.linalg.norm(dq[:4].astype(np.float64)) - 1.0) < 1e-6
    Rq, tq = dq_to_rows(dq)
    assert np.abs(Rq - M[:, :3]).max() < 2e-6 and np.abs(tq - M[:, 3]).max() < 2e-6
    rest, normals, _, _ = mesh(rng, 300, 1)
    joints = np.zeros((300, 4), np.int32)
    weights = np.zeros((300, 4), np.float32)
    weights[:, 0] = 1.0
    out, nout = orc.skin_full(orc.SKIN_DQ, dq[None], joints, weights, rest, normals)
    Md = M.astype(np.float64)
    want = rest[:, :3].astype(np.float64) @ Md[:, :3].T + Md[:, 3]
    assert np.abs(out[:, :3] - want).max() < 1e-5
    wn = normals[:, :3].astype(np.float64) @ Md[:, :3].T
    assert np.abs(nout[:, :3] - wn).max() < 2e-6
    # rigid bones keep unit normals
    assert np.abs(np.linalg.norm(nout[:, :3], axis=1) - 1.0).max() < 2e-6
    # the antipodal dual quaternion is the same transform
    o2, n2 = orc.skin_full(orc.SKIN_DQ, -dq[None], joints, weights, rest, normals)
    assert np.abs(o2[:, :3] - out[:, :3]).max() < 1e-6 and np.abs(n2[:, :3] - nout[:, :3]).max() < 1e-6


def test_dq_normals_blend_and_identity():
    rng = np.random.default_rng(5)
    rest, normals, joints, weights = mesh(rng, 1000, 4)
    ident = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 0], np.float32), (4, 1))
    out, nout = orc.skin_full(orc.SKIN_DQ, ident, joints, weights, rest, normals)
    assert np.array_equal(out[:, :3], rest[:, :3]) and np.array_equal(nout, normals)
    # the positions of skin_full equal the long-standing positions-only oracle
    bones = np.stack([ba.dualquat_from_matrix(random_rigid(rng, 0.5)) for _ in range(4)])
    out, nout = orc.skin_full(orc.SKIN_DQ, bones, joints, weights, rest, normals)
    assert np.array_equal(out.view(np.uint32), orc.skin_f32(bones, joints, weights, rest).view(np.uint32))
    assert np.abs(np.linalg.norm(nout[:, :3], axis=1) - 1.0).max() < 5e-6
    out64, nout64 = orc.skin_full(orc.SKIN_DQ, bones, joints, weights, rest, normals, dtype=np.float64)
    assert np.abs(nout - nout64).max() < 2e-6


def test_root_frames_oracle():
    rng = np.random.default_rng(6)
    nv, nt, S = 300, 200, 1000
    verts = np.zeros((nv, 4), np.float32)
    verts[:, :3] = rng.normal(size=(nv, 3))
    nrm = np.zeros((nv, 4), np.float32)
    d = rng.normal(size=(nv, 3))
    nrm[:, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    tris = rng.integers(0, nv, size=(nt, 3)).astype(np.int32)
    tri = rng.integers(0, nt, size=S).astype(np.int32)
    b = rng.dirichlet(np.ones(3), size=S)
    bary = b[:, 1:].astype(np.float32)
    roots, normals = orc.root_frames_f32(tris, tri, bary, verts, nrm)
    assert np.array_equal(roots.view(np.uint32), orc.bary_roots_f32(tris, tri, bary, verts).view(np.uint32))
    w1, w2 = bary[:, 0].astype(np.float64), bary[:, 1].astype(np.float64)
    w0 = 1.0 - w1 - w2
    n = (w0[:, None] * nrm[tris[tri, 0], :3] + w1[:, None] * nrm[tris[tri, 1], :3]
         + w2[:, None] * nrm[tris[tri, 2], :3])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.abs(normals[:, :3] - n).max() < 1e-5
    assert np.all(normals[:, 3] == 0.0)


def test_dualquat_from_matrix_rejects_bad_input():
    with pytest.raises(ba.FtlError):
        ba.dualquat_from_matrix(np.full(12, np.nan, np.float32))


# --- GPU: the kernels against the oracle ------------------------------------------

def scene_normals(rest):
    n = rest[:, :3].astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    out = np.zeros_like(rest)
    out[:, :3] = n
    return out


def scene_matrices(sc, t):
    """The scene's bones at time t as 3x4 row matrices (from its dual quaternions)."""
    dq = ba.bones(sc, t)
    rows = np.zeros((dq.shape[0], 3, 4), np.float32)
    for b in range(dq.shape[0]):
        R, tt = dq_to_rows(dq[b])
        rows[b, :, :3] = R
        rows[b, :, 3] = tt
    return rows


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [ba.FTL_SKIN_DQ, ba.FTL_SKIN_LBS])
def test_gpu_skinning_modes_with_normals(ctx, mode):
    sc = ba.skinned_scene(3000, 16, ntris=2000, nbones=16, seed=12)
    normals = scene_normals(sc["rest"])
    joints = sc["joints"].copy()
    if mode == ba.FTL_SKIN_LBS:
        joints[::7, 2] = ba.FTL_NO_JOINT  # kNoJoint: the identity
    skin = ba.Skin(ctx, sc["rest"], joints, sc["weights"], sc["tris"], sc["nbones"])
    skin.set_rest_normals(normals)
    if mode == ba.FTL_SKIN_LBS:
        with pytest.raises(ba.FtlError):
            skin.set_mode(ba.FTL_SKIN_DQ)  # a kNoJoint index has no dual quaternion
    skin.set_mode(mode)
    for k in range(3):
        t = (k + 1) * 0.3
        if mode == ba.FTL_SKIN_DQ:
            bones = ba.bones(sc, t)
            skin.set_bones(bones)
        else:
            bones = scene_matrices(sc, t)
            skin.set_matrices(bones)
        skin.update()
        want, wantn = orc.skin_full(mode, bones, joints, sc["weights"], sc["rest"], normals)
        assert_bitexact(skin.download(), want, "mode %d positions %d" % (mode, k))
        assert_bitexact(skin.download_normals(), wantn, "mode %d normals %d" % (mode, k))
    skin.close()


@pytest.mark.gpu
def test_gpu_joint_matrices_both_modes(ctx):
    """skin_i = global_pose_i * inverse_bind_i, converted to dual quaternions for
    DQ mode and used as rows for LBS mode, as SkeletonController does."""
    rng = np.random.default_rng(13)
    sc = ba.skinned_scene(100, 8, ntris=2000, nbones=8, seed=13)
    normals = scene_normals(sc["rest"])
    G = np.zeros((8, 4, 4), np.float32)
    IB = np.zeros((8, 4, 4), np.float32)
    for b in range(8):
        G[b, :3] = random_rigid(rng, 0.6)
        G[b, 3, 3] = 1.0
        IB[b, :3] = random_rigid(rng, 0.6)
        IB[b, 3, 3] = 1.0
    rows = np.einsum("bij,bjk->bik", G.astype(np.float64), IB.astype(np.float64))[:, :3].astype(np.float32)
    skin = ba.Skin(ctx, sc["rest"], sc["joints"], sc["weights"], sc["tris"], sc["nbones"])
    skin.set_rest_normals(normals)
    skin.set_joint_matrices(G, IB)
    for mode in (ba.FTL_SKIN_LBS, ba.FTL_SKIN_DQ):
        skin.set_mode(mode)
        skin.update()
        bones = rows if mode == ba.FTL_SKIN_LBS else np.stack([ba.dualquat_from_matrix(r) for r in rows])
        want, wantn = orc.skin_full(mode, bones, sc["joints"], sc["weights"], sc["rest"], normals)
        assert_bitexact(skin.download(), want, "joint matrices mode %d positions" % mode)
        assert_bitexact(skin.download_normals(), wantn, "joint matrices mode %d normals" % mode)
    # both modes agree to rounding on rigid bones
    skin.set_mode(ba.FTL_SKIN_LBS)
    skin.update()
    p_lbs = skin.download()
    skin.set_mode(ba.FTL_SKIN_DQ)
    skin.update()
    single = sc["weights"][:, 0] >= 1.0 - 1e-6
    if single.any():
        assert np.abs(skin.download()[single, :3] - p_lbs[single, :3]).max() < 1e-4
    skin.close()


@pytest.mark.gpu
def test_gpu_lbs_skinned_roots_step_and_root_frames(ctx):
    """A groom rooted on a linear-blend skinned mesh: each step bit-exact with
    the oracle; the root frames (roots, unit normals) bit-exact; strands regrown
    on the device from the root frames equal the host generator."""
    sc = ba.skinned_scene(2500, 16, ntris=2000, nbones=16, seed=14)
    normals = scene_normals(sc["rest"])
    skin = ba.Skin(ctx, sc["rest"], sc["joints"], sc["weights"], sc["tris"], sc["nbones"])
    skin.set_rest_normals(normals)
    skin.set_mode(ba.FTL_SKIN_LBS)
    g = ba.Groom(ctx, sc["pos"], sc["vel"])
    g.bind_skin(skin, sc["tri"], sc["bary"])
    params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)
    g.set_params(params)
    g.set_colliders(np.array([[0, 0, 0, 1.0]], np.float32))
    P, V = sc["pos"].copy(), sc["vel"].copy()
    for k in range(3):
        m = scene_matrices(sc, (k + 1) * 0.2)
        skin.set_matrices(m)
        skin.update()
        g.step(DT)
        verts, vn = orc.skin_full(ba.FTL_SKIN_LBS, m, sc["joints"], sc["weights"], sc["rest"], normals)
        roots = orc.bary_roots_f32(sc["tris"], sc["tri"], sc["bary"], verts)
        op = oracle_params_for(params, 16, DT, spheres=[(0, 0, 0, 1.0)], root_mode=ba.FTL_ROOT_REST)
        orc.step_f32(op, P, V, roots)
        gp, gv = g.download()
        assert_bitexact(gp, P, "lbs step %d pos" % k)
        assert_bitexact(gv, V, "lbs step %d vel" % k)
    r, n = g.root_frames()
    wr, wn = orc.root_frames_f32(sc["tris"], sc["tri"], sc["bary"], verts, vn)
    assert_bitexact(r, wr, "root frames roots")
    assert_bitexact(n, wn, "root frames normals")
    regrown = ba.reference_groom_device(ctx, r[:, :3], n[:, :3], npoints=16, maxlength=0.2, seed=3)
    hp, hv = ba.reference_groom(r[:, :3], n[:, :3], npoints=16, maxlength=0.2, seed=3)
    dp, dv = regrown.download()
    assert np.array_equal(dp.view(np.uint32), hp.view(np.uint32))
    regrown.close()
    g.close()
    skin.close()


@pytest.mark.gpu
def test_gpu_skin_destroyed_before_its_grooms(ctx):
    """Destroying a skin unbinds the grooms bound to it: a step with skinned
    roots is then refused (no bound skin) rather than reading the freed mesh,
    another root mode steps on, and a groom bound to a second skin, or
    destroyed first, detaches cleanly."""
    sc = ba.skinned_scene(300, 8, ntris=200, nbones=4, seed=21)
    skin = ba.Skin(ctx, sc["rest"], sc["joints"], sc["weights"], sc["tris"], sc["nbones"])
    other = ba.Skin(ctx, sc["rest"], sc["joints"], sc["weights"], sc["tris"], sc["nbones"])
    g = ba.Groom(ctx, sc["pos"], sc["vel"])
    h = ba.Groom(ctx, sc["pos"], sc["vel"])
    g.bind_skin(skin, sc["tri"], sc["bary"])
    h.bind_skin(skin, sc["tri"], sc["bary"])
    h.bind_skin(other, sc["tri"], sc["bary"])  # moves h to the second skin
    skin.update()
    g.step(DT)
    skin.close()
    with pytest.raises(ba.FtlError) as e:
        g.step(DT)
    assert e.value.status == ba.FTL_ERR_INVALID_ARGUMENT
    with pytest.raises(ba.FtlError):
        g.set_root_mode(ba.FTL_ROOT_SKINNED)
    g.set_root_mode(ba.FTL_ROOT_MODEL)
    g.step(DT)
    other.update()
    h.step(DT)  # still bound to the skin that lives
    h.close()  # detaches from `other` before it goes
    other.close()
    g.close()
