"""ctypes wrapper of the CPU oracle (oracle/ftl_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, as the checker.  The product (barbu_amd,
libftl.so) never imports or calls it.

Build with `make -C oracle` (also done by __graft_entry__.build()).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libftl_oracle.so")


class OracleParams(C.Structure):
    _fields_ = [
        ("npoints", C.c_int),
        ("iterations", C.c_int),
        ("velocity_mode", C.c_int),
        ("root_mode", C.c_int),
        ("dt", C.c_float),
        ("inv_dt", C.c_float),
        ("length_scale", C.c_float),
        ("damping", C.c_float),
        ("max_velocity", C.c_float),
        ("force", C.c_float * 3),
        ("ncolliders", C.c_int),
        ("spheres", (C.c_float * 4) * 4),
        ("model", C.c_float * 12),
        ("model_is_identity", C.c_int),
        ("wind_grid", C.c_int),
        ("wind_frequency", C.c_float),
        ("wind_offset", C.c_float * 3),
        ("wind_table", C.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.oracle_rsqrt_f32.restype = C.c_float
        L.oracle_rsqrt_f32.argtypes = [C.c_float]
        L.oracle_rsqrt_f32_array.argtypes = [vp, vp, C.c_int64]
        for n in ("oracle_step_f32", "oracle_step_f64"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [C.POINTER(OracleParams), C.c_int, vp, vp, vp]
        L.oracle_step_f32_mt.restype = C.c_int
        L.oracle_step_f32_mt.argtypes = [C.POINTER(OracleParams), C.c_int, vp, vp, vp, C.c_int]
        L.oracle_wind_f32.argtypes = [C.POINTER(OracleParams), vp, vp]
        L.oracle_skin_f32.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        L.oracle_skin_f64.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        L.oracle_bary_roots_f32.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        for n in ("oracle_skin_full_f32", "oracle_skin_full_f64"):
            getattr(L, n).argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_root_frames_f32.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_max_threads.restype = C.c_int
        for n in ("oracle_expand_f32", "oracle_expand_f64"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [vp, vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_float, vp]
        L.oracle_reference_groom_f32.argtypes = [C.c_int, vp, vp, C.c_int, C.c_float, C.c_uint64, vp, vp]
        L.oracle_simplex2_f32.restype = C.c_float
        L.oracle_simplex2_f32.argtypes = [C.c_float, C.c_float]
        for fn in (L.oracle_sin_f32, L.oracle_cos_f32, L.oracle_asin_f32, L.oracle_acos_f32, L.oracle_exp_f32,
                   L.oracle_log_f32):
            fn.restype = C.c_float
            fn.argtypes = [C.c_float]
        L.oracle_pow_f32.restype = C.c_float
        L.oracle_pow_f32.argtypes = [C.c_float, C.c_float]
        L.oracle_reference_tangents_f32.argtypes = [C.c_int, C.c_int, vp, C.c_float, C.c_int, vp]
        L.oracle_marschner_f32.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, vp, vp]
        L.oracle_marschner_f64.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, vp, vp]
        L.oracle_marschner_np_f32.restype = C.c_float
        L.oracle_marschner_np_f32.argtypes = [C.c_int] + [C.c_float] * 5
        L.oracle_cubic_f64.restype = C.c_int
        L.oracle_cubic_f64.argtypes = [C.c_double] * 4 + [vp]
        L.oracle_fresnel_f32.restype = C.c_float
        L.oracle_fresnel_f32.argtypes = [C.c_float] * 4
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def f32(x):
    return np.float32(x)


def make_params(npoints, dt, *, iterations=8, velocity_mode=0, root_mode=0, length_scale=1.45,
                damping=0.80, max_velocity=100.0, gravity=(0.0, -9.81, 0.0), force_scale=20.0,
                spheres=((0.0, 0.0, 0.0, 1.0),), model=None, wind_table=None, wind_frequency=2.0,
                wind_offset=(0.0, 0.0, 0.0)):
    """Oracle parameters.  Derived values are formed in float32 exactly as a GLSL
    host would hand them over: force = force_scale * gravity, inv_dt = 1 / dt.
    model: 4x4 row-major affine matrix or None (identity)."""
    P = OracleParams()
    P.npoints = int(npoints)
    P.iterations = int(iterations)
    P.velocity_mode = int(velocity_mode)
    P.root_mode = int(root_mode)
    P.dt = f32(dt)
    P.inv_dt = f32(1.0) / f32(dt)
    P.length_scale = f32(length_scale)
    P.damping = f32(damping)
    P.max_velocity = f32(max_velocity)
    for c in range(3):
        P.force[c] = f32(force_scale) * f32(gravity[c])
    spheres = [] if spheres is None else list(spheres)
    P.ncolliders = len(spheres)
    for k, s in enumerate(spheres):
        for c in range(4):
            P.spheres[k][c] = f32(s[c])
    if model is None:
        m = np.eye(4, dtype=np.float32)
        P.model_is_identity = 1
    else:
        m = np.asarray(model, np.float32).reshape(4, 4)
        P.model_is_identity = int(np.array_equal(m, np.eye(4, dtype=np.float32)))
    for r in range(3):
        for c in range(4):
            P.model[4 * r + c] = m[r, c]
    P._keep = None
    if wind_table is not None:
        wt = np.ascontiguousarray(wind_table, np.float32)
        P._keep = wt
        P.wind_grid = int(round(wt.shape[0] ** (1.0 / 3.0)))
        P.wind_frequency = f32(wind_frequency)
        for c in range(3):
            P.wind_offset[c] = f32(wind_offset[c])
        P.wind_table = wt.ctypes.data
    return P


def step_f32(P, pos, vel, roots=None, nthreads=1):
    """One step in place on float32 arrays [S, N, 4]."""
    assert pos.dtype == np.float32 and vel.dtype == np.float32 and pos.flags.c_contiguous
    if roots is not None:
        roots = np.ascontiguousarray(roots, np.float32)
    S = pos.shape[0]
    if nthreads > 1:
        r = lib().oracle_step_f32_mt(C.byref(P), S, _p(pos), _p(vel), _p(roots), int(nthreads))
    else:
        r = lib().oracle_step_f32(C.byref(P), S, _p(pos), _p(vel), _p(roots))
    if r != 0:
        raise ValueError("oracle_step_f32 rejected its arguments")


def step_f64(P, pos, vel, roots=None):
    """One step in place on float64 arrays [S, N, 4] (same formulas in double)."""
    assert pos.dtype == np.float64 and vel.dtype == np.float64 and pos.flags.c_contiguous
    if roots is not None:
        roots = np.ascontiguousarray(roots, np.float64)  # the FP64 instantiation reads double roots
    r = lib().oracle_step_f64(C.byref(P), pos.shape[0], _p(pos), _p(vel), _p(roots))
    if r != 0:
        raise ValueError("oracle_step_f64 rejected its arguments")


def rsqrt_f32(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().oracle_rsqrt_f32_array(_p(x), _p(out), x.size)
    return out


def wind_f32(P, x):
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    out = np.empty_like(x)
    for i in range(x.shape[0]):
        lib().oracle_wind_f32(C.byref(P), _p(x[i]), _p(out[i]))
    return out


def skin_f32(dq, joints, weights, rest):
    dq = np.ascontiguousarray(dq, np.float32)
    joints = np.ascontiguousarray(joints, np.int32)
    weights = np.ascontiguousarray(weights, np.float32)
    rest = np.ascontiguousarray(rest, np.float32)
    out = np.empty_like(rest)
    lib().oracle_skin_f32(rest.shape[0], _p(dq), _p(joints), _p(weights), _p(rest), _p(out))
    return out


def skin_f64(dq, joints, weights, rest):
    dq = np.ascontiguousarray(dq, np.float32)
    joints = np.ascontiguousarray(joints, np.int32)
    weights = np.ascontiguousarray(weights, np.float32)
    rest = np.ascontiguousarray(rest, np.float32)
    out = np.empty_like(rest)
    lib().oracle_skin_f64(rest.shape[0], _p(dq), _p(joints), _p(weights), _p(rest), _p(out))
    return out


SKIN_DQ, SKIN_LBS = 0, 1


def skin_full(mode, bones, joints, weights, rest, normals, dtype=np.float32):
    """Skinned positions and normals (float32 [nverts, 4] each): mode 0 dual
    quaternions (8 floats per bone), mode 1 linear blending (12 floats per
    bone, the rows of its 3x4 matrix; joint 0xff = identity).  dtype
    float64 evaluates the same formulas in double."""
    bones = np.ascontiguousarray(bones, np.float32)
    joints = np.ascontiguousarray(joints, np.int32)
    weights = np.ascontiguousarray(weights, np.float32)
    rest = np.ascontiguousarray(rest, np.float32)
    normals = np.ascontiguousarray(normals, np.float32)
    out = np.empty_like(rest)
    nout = np.empty_like(normals)
    fn = lib().oracle_skin_full_f64 if dtype == np.float64 else lib().oracle_skin_full_f32
    fn(int(mode), rest.shape[0], _p(bones), _p(joints), _p(weights), _p(rest), _p(normals), _p(out), _p(nout))
    return out, nout


def root_frames_f32(tris, tri, bary, verts, normals):
    """Each strand's root and unit normal on its bound triangle: (roots, normals), float32 [S, 4]."""
    tris = np.ascontiguousarray(tris, np.int32)
    tri = np.ascontiguousarray(tri, np.int32)
    bary = np.ascontiguousarray(bary, np.float32)
    verts = np.ascontiguousarray(verts, np.float32)
    normals = np.ascontiguousarray(normals, np.float32)
    r = np.empty((tri.shape[0], 4), np.float32)
    n = np.empty((tri.shape[0], 4), np.float32)
    lib().oracle_root_frames_f32(tri.shape[0], _p(tris), _p(tri), _p(bary), _p(verts), _p(normals), _p(r), _p(n))
    return r, n


def bary_roots_f32(tris, tri, bary, verts):
    tris = np.ascontiguousarray(tris, np.int32)
    tri = np.ascontiguousarray(tri, np.int32)
    bary = np.ascontiguousarray(bary, np.float32)
    verts = np.ascontiguousarray(verts, np.float32)
    out = np.empty((tri.shape[0], 4), np.float32)
    lib().oracle_bary_roots_f32(tri.shape[0], _p(tris), _p(tri), _p(bary), _p(verts), _p(out))
    return out


def expand_vertex_count(ntris, npoints, ninst=3, nlines=2, nsub=16):
    return 2 * ninst * ntris * (npoints - 1) * nlines * nsub


def _expand(fn, dtype, pos, tan, tris, rnd, ninst, nlines, nsub, scale):
    pos = np.ascontiguousarray(pos, np.float32)
    tan = np.ascontiguousarray(tan, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    rnd = np.ascontiguousarray(rnd, np.float32)
    S, N = pos.shape[0], pos.shape[1]
    T = tris.shape[0]
    out = np.empty((expand_vertex_count(T, N, ninst, nlines, nsub), 4), dtype)
    rc = fn(_p(pos), _p(tan), S, N, _p(tris), T, _p(rnd), rnd.shape[0], ninst, nlines, nsub, f32(scale), _p(out))
    if rc != 0:
        raise ValueError("expansion: a triangle or random index is out of range")
    return out


def expand_f32(pos, tan, tris, rnd, ninst=3, nlines=2, nsub=16, scale=1.0):
    """Render-side strand expansion, FP32 (the kernel's evaluation order).
    pos, tan: [S, N, 4]; tris: [T, 3]; rnd: [R, 4]; returns [V, 4] float32."""
    return _expand(lib().oracle_expand_f32, np.float32, pos, tan, tris, rnd, ninst, nlines, nsub, scale)


def expand_f64(pos, tan, tris, rnd, ninst=3, nlines=2, nsub=16, scale=1.0):
    """The same expansion in FP64, evaluated as the GLSL writes it."""
    return _expand(lib().oracle_expand_f64, np.float64, pos, tan, tris, rnd, ninst, nlines, nsub, scale)


def reference_groom_f32(roots, normals, npoints, maxlength, seed):
    """Hair::init_simulation's positions/velocities (seeded SplitMix64 draws)."""
    roots = np.ascontiguousarray(roots, np.float32).reshape(-1, 3)
    normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    n = roots.shape[0]
    pos = np.empty((n, npoints, 4), np.float32)
    vel = np.empty_like(pos)
    lib().oracle_reference_groom_f32(n, _p(roots), _p(normals), npoints, f32(maxlength), int(seed), _p(pos), _p(vel))
    return pos, vel


def simplex2_f32(x, y):
    return float(lib().oracle_simplex2_f32(f32(x), f32(y)))


def sin_f32(x):
    """The FP32 sine of the groom generator's curly tangents."""
    return float(lib().oracle_sin_f32(f32(x)))


def cos_f32(x):
    return float(lib().oracle_cos_f32(f32(x)))


def math_f32(name, *args):
    """One of the shared FP32 transcendentals of the Marschner tables:
    'asin', 'acos', 'exp', 'log' or 'pow' (two arguments)."""
    return float(getattr(lib(), "oracle_%s_f32" % name)(*[f32(a) for a in args]))


def reference_tangents_f32(normals, npoints, maxlength, curly=True):
    normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    n = normals.shape[0]
    out = np.empty((n, npoints, 4), np.float32)
    lib().oracle_reference_tangents_f32(n, npoints, _p(normals), f32(maxlength), 1 if curly else 0, _p(out))
    return out


def marschner_luts(res=128, eta=1.55, absorption=0.2, ar=-5.0, br=5.0, dtype=np.float32):
    """Marschner M and N tables [res, res, 4] (float32 restatement, or float64)."""
    M = np.empty((res, res, 4), dtype)
    N = np.empty_like(M)
    fn = lib().oracle_marschner_f32 if dtype == np.float32 else lib().oracle_marschner_f64
    fn(res, f32(eta), f32(absorption), f32(ar), f32(br), _p(M), _p(N))
    return M, N


def marschner_np_f32(p, absorption, eperp, epar, c, phi):
    return float(lib().oracle_marschner_np_f32(p, *(f32(v) for v in (absorption, eperp, epar, c, phi))))


def cubic_f64(A, B, C, D):
    r = np.zeros(3)
    n = lib().oracle_cubic_f64(A, B, C, D, _p(r))
    return r[:n]


def fresnel_f32(eperp, epar, c, s):
    return float(lib().oracle_fresnel_f32(f32(eperp), f32(epar), f32(c), f32(s)))


def max_threads():
    return lib().oracle_max_threads()
