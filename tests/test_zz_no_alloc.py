"""ftl_batch_step_noalloc (include/ftl.h): the batched step whose descriptors
and span list live in a fixed device array instead of buffers the library
allocates.  Its results must equal ftl_batch_step's bit for bit (the bar of
the batched path, which the other GPU tests tie to the FP32 oracle), on two
calls in a row, the second with a smaller batch (the array keeps the first
call's contents; nothing of them may leak into the second); a batch larger
than the array is refused with a clear error and steps nothing."""
import ctypes as C

import numpy as np
import pytest

import barbu_amd as ba
from barbu_amd import _ffi
from oracle import ftl_oracle as orc
from tests.helpers import assert_bitexact, oracle_params_for

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)
SHAPES = [(1000, 8), (130, 16), (777, 32), (64, 64), (2049, 16), (5, 32), (900, 64), (300, 5), (200, 20),
          (100, 48), (50, 2), (4096, 64)]


def make_grooms(ctx, shapes, seed):
    rng = np.random.default_rng(seed)
    grooms, states = [], []
    for i, (S, N) in enumerate(shapes):
        pos, vel, roots = ba.sphere_groom(S, N, center=tuple(rng.normal(size=3)), length=float(rng.uniform(0.05, 0.4)),
                                          seed=500 + i)
        g = ba.Groom(ctx, pos, vel, rest_roots=roots)
        params = ba.make_params(force_scale=1.0, damping=0.9, length_scale=1.0)
        g.set_params(params)
        sph = np.array([[*np.mean(roots[:, :3], axis=0), 1.0], [*(np.mean(roots[:, :3], axis=0) + [0, 1.1, 0]), 0.3]],
                       np.float32)[: 1 + i % 2]
        g.set_colliders(sph)
        g.set_root_mode(ba.FTL_ROOT_REST)
        grooms.append(g)
        states.append(dict(pos=pos.copy(), vel=vel.copy(), roots=roots, params=params, spheres=sph, N=N))
    return grooms, states


def step_noalloc(ctx, grooms, dts, stream=None):
    arr = (C.c_void_p * len(grooms))(*[g.handle.value for g in grooms])
    dt = np.ascontiguousarray(dts, np.float32)
    return ba.lib().ftl_batch_step_noalloc(ctx.handle, arr, len(grooms), dt.ctypes.data, stream)


def test_noalloc_matches_batch_step_two_calls(ctx):
    ref, states = make_grooms(ctx, SHAPES, 1)
    out, _ = make_grooms(ctx, SHAPES, 1)
    dts = np.array([1 / 60, 1 / 90] * 6, np.float32)[: len(SHAPES)]
    # call 1: every groom
    ba.batch_step(ctx, ref, dts)
    assert step_noalloc(ctx, out, dts) == _ffi.FTL_OK, ba.lib().ftl_last_error()
    for i, (a, b) in enumerate(zip(ref, out)):
        pa, va = a.download()
        pb, vb = b.download()
        assert_bitexact(pb, pa, "call 1 groom %d positions" % i)
        assert_bitexact(vb, va, "call 1 groom %d velocities" % i)
    # the batched step itself against the oracle on a small groom
    s = states[5]
    op = oracle_params_for(s["params"], s["N"], dts[5], spheres=[tuple(x) for x in s["spheres"]],
                           root_mode=ba.FTL_ROOT_REST)
    orc.step_f32(op, s["pos"], s["vel"], s["roots"])
    assert_bitexact(out[5].download()[0], s["pos"], "call 1 groom 5 against the oracle")
    # call 2: a smaller batch (fewer grooms and tiles) on a stream
    from barbu_amd import _hip as hipm
    st = hipm.Stream()
    sub = [3, 5, 7, 10]
    ba.batch_step(ctx, [ref[i] for i in sub], dts[sub])
    assert step_noalloc(ctx, [out[i] for i in sub], dts[sub], st.h) == _ffi.FTL_OK, ba.lib().ftl_last_error()
    st.synchronize()
    for i in range(len(SHAPES)):
        pa, va = ref[i].download()
        pb, vb = out[i].download()
        assert_bitexact(pb, pa, "call 2 groom %d positions" % i)
        assert_bitexact(vb, va, "call 2 groom %d velocities" % i)
    for g in ref + out:
        g.close()


def test_noalloc_refuses_a_batch_beyond_its_scratch(ctx):
    """The array holds 288 B of descriptor and a 16 B span per groom (about
    13,800 grooms): a batch of more grooms is refused before anything is
    planned or stepped."""
    pos, vel, _ = ba.sphere_groom(64, 8, radius=1.0, length=0.2, seed=4)
    g = ba.Groom(ctx, pos, vel)
    before = g.download()
    n = (4 << 20) // (288 + 16) + 1
    st = step_noalloc(ctx, [g] * n, np.full(n, DT, np.float32))
    assert st == _ffi.FTL_ERR_INVALID_ARGUMENT
    assert b"no-allocation scratch" in ba.lib().ftl_last_error()
    after = g.download()
    assert np.array_equal(before[0].view(np.uint32), after[0].view(np.uint32))
    g.close()


def test_noalloc_groom_of_many_tiles(ctx):
    """The scratch holds one record per groom, whatever its tile count: a
    groom of 16.8M two-point strands (262,145 tiles of 64 strands) steps
    bit-identically to ftl_batch_step."""
    S = (262144 + 1) * 64
    rng = np.random.default_rng(3)
    roots = rng.normal(size=(S, 3)).astype(np.float32)
    normals = roots / np.linalg.norm(roots, axis=1, keepdims=True)
    a = ba.reference_groom_device(ctx, roots, normals, npoints=2, maxlength=0.1, seed=1)
    b = ba.reference_groom_device(ctx, roots, normals, npoints=2, maxlength=0.1, seed=1)
    dt = np.array([DT], np.float32)
    ba.batch_step(ctx, [a], dt)
    assert step_noalloc(ctx, [b], dt) == 0
    pa, va = a.download()
    pb, vb = b.download()
    assert_bitexact(pb, pa, "positions")
    assert_bitexact(vb, va, "velocities")
    a.close()
    b.close()
