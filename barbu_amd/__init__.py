"""barbu_amd: the FTL hair-strand step of tcoppex/barbu on MI355X (gfx950).

The compute path is libftl.so (HIP kernels + C ABI, include/ftl.h).  This
package is a thin ctypes layer over it for Python callers and tests; it has no
compute of its own and fails loudly when the library is missing.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import (FTL_ROOT_MODEL, FTL_ROOT_REST, FTL_ROOT_SKINNED, FTL_SKIN_DQ, FTL_SKIN_LBS, FTL_NO_JOINT,
                   FTL_VELOCITY_DFTL,
                   FTL_VELOCITY_SHIPPED, FTL_MAX_COLLIDERS, FTL_MAX_POINTS, FTL_WIND_GRID,
                   FTL_OK, FTL_ERR_INVALID_ARGUMENT, FTL_ERR_DEVICE, FTL_ERR_OUT_OF_MEMORY, FTL_ERR_BUDGET,
                   FTL_ERR_NOT_INITIALIZED, FTL_ERR_NONFINITE, FTL_ERR_UNSUPPORTED, FTL_ERR_BUSY, FTL_ERR_CANCELLED,
                   FtlError, Params, check, lib)

__all__ = ["Context", "Groom", "Skin", "Batch", "Server", "make_requests", "set_request", "Params", "reference_params", "FtlError",
           "sphere_groom", "reference_groom", "skinned_scene", "bones", "wind_table",
           "Expand", "expand_params", "strand_tangents", "random_table",
           "FTL_ROOT_MODEL", "FTL_ROOT_REST", "FTL_ROOT_SKINNED", "FTL_VELOCITY_SHIPPED",
           "FTL_VELOCITY_DFTL", "Hair"]


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def reference_params():
    """The shipped shader's constants (ftl_params_reference)."""
    p = Params()
    lib().ftl_params_reference(C.byref(p))
    return p


def make_params(**kw):
    """Reference parameters with fields overridden by keyword."""
    p = reference_params()
    names = {n for n, _ in Params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError("unknown parameter %r" % k)
        if k in ("gravity", "wind_scroll"):
            for i in range(3):
                getattr(p, k)[i] = float(v[i])
        else:
            setattr(p, k, v)
    return p


def class_shape(cls):
    """Shape of launch class cls: dict(points_per_lane, lanes, guarded, waves, chunk)."""
    v = [C.c_int(0) for _ in range(5)]
    check(lib().ftl_class_shape(int(cls), *[C.byref(x) for x in v]), "ftl_class_shape")
    return dict(zip(("points_per_lane", "lanes", "guarded", "waves", "chunk"), (x.value for x in v)))


def dualquat_from_matrix(m3x4):
    """The dual quaternion (8 floats: real xyzw, dual xyzw) of a rigid 3x4
    transform given as its three rows (ftl_dualquat_from_matrix; host only)."""
    m = _f32(m3x4).reshape(12)
    out = np.empty(8, np.float32)
    check(lib().ftl_dualquat_from_matrix(_ptr(m), _ptr(out)), "ftl_dualquat_from_matrix")
    return out


def wind_table(params):
    """The wind lattice selected by params: float32 [G^3, 4]."""
    G = FTL_WIND_GRID
    out = np.zeros((G * G * G, 4), np.float32)
    g = C.c_int(0)
    check(lib().ftl_wind_table(C.byref(params), _ptr(out), C.byref(g)), "ftl_wind_table")
    return out


class Context:
    """A HIP device bound to the library (ftl_init)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        check(lib().ftl_init(int(device), C.byref(self._h)), "ftl_init")

    @property
    def handle(self):
        return self._h

    def memory(self):
        v = C.c_uint64(0)
        check(lib().ftl_ctx_memory(self._h, C.byref(v)), "ftl_ctx_memory")
        return v.value

    def autotune(self, cache_path=None, force=False):
        path = cache_path.encode() if cache_path else None
        check(lib().ftl_autotune(self._h, path, 1 if force else 0), "ftl_autotune")

    def lanes_for(self, npoints, wind=False):
        """Lanes per strand the library picks for npoints (plain steps with
        wind take their own table)."""
        return (lib().ftl_lanes_for_wind if wind else lib().ftl_lanes_for)(self._h, int(npoints))

    def class_for(self, npoints, wind=False):
        """The launch class (tile shape) the library picks for npoints."""
        return int(lib().ftl_class_for_points(self._h, int(npoints), 1 if wind else 0))

    def close(self):
        if self._h:
            check(lib().ftl_shutdown(self._h), "ftl_shutdown")
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Groom:
    """One groom's device state (ftl_groom_*)."""

    def __init__(self, ctx, pos, vel=None, rest_roots=None, _handle=None):
        if _handle is not None:   # an existing handle (device-side initialisation)
            self.ctx, self._h = ctx, _handle
            self.nstrands, self.npoints = int(pos[0]), int(pos[1])
            return
        pos = _f32(pos)
        n = pos.shape[0]
        self.nstrands, self.npoints = int(pos.shape[0]), int(pos.shape[1])
        pos = pos.reshape(self.nstrands, self.npoints, 4)
        vel = None if vel is None else _f32(vel, pos.shape)
        roots = None if rest_roots is None else _f32(rest_roots, (n, 4))
        d = _ffi.GroomDesc(self.nstrands, self.npoints, pos.ctypes.data,
                           None if vel is None else vel.ctypes.data,
                           None if roots is None else roots.ctypes.data)
        self._h = C.c_void_p()
        self.ctx = ctx
        check(lib().ftl_groom_create(ctx.handle, C.byref(d), C.byref(self._h)), "ftl_groom_create")

    @property
    def handle(self):
        return self._h

    def set_params(self, params):
        check(lib().ftl_groom_set_params(self._h, C.byref(params)), "ftl_groom_set_params")

    def get_params(self):
        p = Params()
        check(lib().ftl_groom_get_params(self._h, C.byref(p)), "ftl_groom_get_params")
        return p

    def set_colliders(self, spheres):
        s = _f32(spheres).reshape(-1, 4)
        check(lib().ftl_groom_set_colliders(self._h, _ptr(s), int(s.shape[0])), "ftl_groom_set_colliders")

    def set_root_transform(self, m):
        """m: 4x4 matrix (row-major numpy, math convention); stored column-major."""
        m = np.asarray(m, np.float32).reshape(4, 4)
        cm = np.ascontiguousarray(m.T)
        check(lib().ftl_groom_set_root_transform(self._h, _ptr(cm)), "ftl_groom_set_root_transform")

    def set_root_mode(self, mode):
        check(lib().ftl_groom_set_root_mode(self._h, int(mode)), "ftl_groom_set_root_mode")

    def set_rest_roots(self, roots):
        r = _f32(roots, (self.nstrands, 4))
        check(lib().ftl_groom_set_rest_roots(self._h, _ptr(r)), "ftl_groom_set_rest_roots")

    def set_time(self, t):
        check(lib().ftl_groom_set_time(self._h, float(t)), "ftl_groom_set_time")

    def set_lanes(self, lanes):
        check(lib().ftl_groom_set_lanes(self._h, int(lanes)), "ftl_groom_set_lanes")

    def step(self, dt, stream=None):
        check(lib().ftl_groom_step(self._h, float(dt), stream), "ftl_groom_step")

    def step_timed(self, dt, stream, start_event, stop_event):
        """step() with its kernel bracketed by two hipEvent_t handles (ftl_groom_step_timed)."""
        check(lib().ftl_groom_step_timed(self._h, float(dt), stream, start_event, stop_event),
              "ftl_groom_step_timed")

    def download(self):
        pos = np.empty((self.nstrands, self.npoints, 4), np.float32)
        vel = np.empty_like(pos)
        check(lib().ftl_groom_download(self._h, _ptr(pos), _ptr(vel)), "ftl_groom_download")
        return pos, vel

    def upload(self, pos=None, vel=None):
        shape = (self.nstrands, self.npoints, 4)
        pos = None if pos is None else _f32(pos, shape)
        vel = None if vel is None else _f32(vel, shape)
        check(lib().ftl_groom_upload(self._h, _ptr(pos), _ptr(vel)), "ftl_groom_upload")

    def device_buffers(self):
        p, v = C.c_void_p(), C.c_void_p()
        check(lib().ftl_groom_device_buffers(self._h, C.byref(p), C.byref(v)), "ftl_groom_device_buffers")
        return p.value, v.value

    def status(self):
        nf = C.c_int(0)
        check(lib().ftl_groom_status(self._h, C.byref(nf)), "ftl_groom_status")
        return bool(nf.value)

    def bind_skin(self, skin, tri, bary):
        tri = np.ascontiguousarray(tri, np.int32)
        bary = _f32(bary, (self.nstrands, 2))
        check(lib().ftl_groom_bind_skin(self._h, skin.handle, _ptr(tri), _ptr(bary)), "ftl_groom_bind_skin")

    def root_frames(self, normals=True):
        """Skinned roots and (if normals) unit root normals, float32 [nstrands, 4] each."""
        r = np.empty((self.nstrands, 4), np.float32)
        n = np.empty((self.nstrands, 4), np.float32) if normals else None
        check(lib().ftl_groom_root_frames(self._h, _ptr(r), _ptr(n)), "ftl_groom_root_frames")
        return (r, n) if normals else r

    def close(self):
        if self._h:
            lib().ftl_groom_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Skin:
    """Dual-quaternion skinned mesh (ftl_skin_*)."""

    def __init__(self, ctx, rest, joints, weights, tris, nbones):
        self.rest = _f32(rest).reshape(-1, 4)
        self.joints = np.ascontiguousarray(joints, np.int32).reshape(-1, 4)
        self.weights = _f32(weights).reshape(-1, 4)
        self.tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        self.nbones = int(nbones)
        d = _ffi.SkinDesc(self.rest.shape[0], self.tris.shape[0], self.nbones, self.rest.ctypes.data,
                          self.joints.ctypes.data, self.weights.ctypes.data, self.tris.ctypes.data)
        self._h = C.c_void_p()
        check(lib().ftl_skin_create(ctx.handle, C.byref(d), C.byref(self._h)), "ftl_skin_create")

    @property
    def handle(self):
        return self._h

    def set_mode(self, mode):
        """FTL_SKIN_DQ (dual quaternions) or FTL_SKIN_LBS (linear blending)."""
        check(lib().ftl_skin_set_mode(self._h, int(mode)), "ftl_skin_set_mode")

    def set_bones(self, dq):
        dq = _f32(dq).reshape(-1, 8)
        check(lib().ftl_skin_set_bones(self._h, _ptr(dq), int(dq.shape[0])), "ftl_skin_set_bones")

    def set_matrices(self, m3x4):
        """Skinning matrices, (nbones, 3, 4): the rows of each bone's affine matrix."""
        m = _f32(m3x4).reshape(-1, 12)
        check(lib().ftl_skin_set_matrices(self._h, _ptr(m), int(m.shape[0])), "ftl_skin_set_matrices")

    def set_joint_matrices(self, global_pose, inverse_bind=None):
        """Row-major (nbones, 4, 4) global poses and inverse bind matrices; the
        library forms skin_i = global_i @ inverse_bind_i (column-major at the C ABI)."""
        g = np.ascontiguousarray(np.transpose(_f32(global_pose).reshape(-1, 4, 4), (0, 2, 1)))
        ib = None if inverse_bind is None else np.ascontiguousarray(
            np.transpose(_f32(inverse_bind).reshape(-1, 4, 4), (0, 2, 1)))
        check(lib().ftl_skin_set_joint_matrices(self._h, _ptr(g), _ptr(ib), int(g.shape[0])),
              "ftl_skin_set_joint_matrices")

    def set_rest_normals(self, normals):
        n = _f32(normals).reshape(self.rest.shape[0], 4)
        check(lib().ftl_skin_set_rest_normals(self._h, _ptr(n)), "ftl_skin_set_rest_normals")

    def update(self, stream=None):
        check(lib().ftl_skin_update(self._h, stream), "ftl_skin_update")

    def download(self):
        out = np.empty((self.rest.shape[0], 4), np.float32)
        check(lib().ftl_skin_download(self._h, _ptr(out)), "ftl_skin_download")
        return out

    def download_normals(self):
        out = np.empty((self.rest.shape[0], 4), np.float32)
        check(lib().ftl_skin_download_normals(self._h, _ptr(out)), "ftl_skin_download_normals")
        return out

    def close(self):
        if self._h:
            lib().ftl_skin_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A fixed set of grooms stepped as one graph-captured frame (ftl_batch_*)."""

    def __init__(self, ctx, grooms):
        self.grooms = list(grooms)
        arr = (C.c_void_p * len(self.grooms))(*[g.handle.value for g in self.grooms])
        self._h = C.c_void_p()
        check(lib().ftl_batch_create(ctx.handle, arr, len(self.grooms), C.byref(self._h)), "ftl_batch_create")

    def launch(self, dt, stream=None):
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (len(self.grooms),)))
        check(lib().ftl_batch_launch(self._h, _ptr(dt), stream), "ftl_batch_launch")

    def info(self):
        nt, gr = C.c_int(0), C.c_int(0)
        check(lib().ftl_batch_info(self._h, C.byref(nt), C.byref(gr)), "ftl_batch_info")
        return nt.value, gr.value

    def close(self):
        if self._h:
            lib().ftl_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_step(ctx, grooms, dt, stream=None):
    """One batched step over grooms (ftl_batch_step): a tile-list kernel per launch class."""
    arr = (C.c_void_p * len(grooms))(*[g.handle.value for g in grooms])
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (len(grooms),)))
    check(lib().ftl_batch_step(ctx.handle, arr, len(grooms), _ptr(dt), stream), "ftl_batch_step")


# --- serving loop ---------------------------------------------------------------

# numpy layout of ftl_step_request (144 bytes, no padding)
REQUEST_DTYPE = np.dtype([("dt", "<f4"), ("time", "<f4"), ("has_transform", "<i4"), ("model", "<f4", (16,)),
                          ("ncolliders", "<i4"), ("spheres", "<f4", (16,))])
assert REQUEST_DTYPE.itemsize == C.sizeof(_ffi.StepRequest)


def make_requests(n, dt=1.0 / 60.0, time=0.0):
    """n step requests that keep every creature's transform and colliders."""
    r = np.zeros(n, REQUEST_DTYPE)
    r["dt"] = dt
    r["time"] = time
    r["ncolliders"] = -1
    return r


def set_request(r, dt=None, time=None, model=None, spheres=None):
    """Fills one request (a REQUEST_DTYPE record); model is a row-major 4x4 (as
    Groom.set_root_transform), spheres an (k, 4) array."""
    if dt is not None:
        r["dt"] = dt
    if time is not None:
        r["time"] = time
    if model is not None:
        r["has_transform"] = 1
        r["model"] = np.asarray(model, np.float32).reshape(4, 4).T.reshape(16)
    if spheres is not None:
        sp = np.asarray(spheres, np.float32).reshape(-1, 4)
        r["ncolliders"] = sp.shape[0]
        r["spheres"] = 0.0
        r["spheres"][: sp.size] = sp.reshape(-1)
    return r


class Server:
    """Resident creatures stepped by continuous batching (ftl_server_*)."""

    def __init__(self, ctx, hbm_budget_bytes=0, batch_cap_bytes=0, max_inflight=0, coalesce_us=0,
                 batch_target_bytes=0):
        cfg = _ffi.ServerConfig(int(hbm_budget_bytes), int(batch_cap_bytes), int(max_inflight), int(coalesce_us),
                                int(batch_target_bytes))
        self._h = C.c_void_p()
        self.ctx = ctx
        self.shapes = {}
        check(lib().ftl_server_create(ctx.handle, C.byref(cfg), C.byref(self._h)), "ftl_server_create")

    def add(self, pos, vel=None, rest_roots=None, params=None):
        """Returns (id, resident)."""
        pos = _f32(pos)
        S, N = int(pos.shape[0]), int(pos.shape[1])
        pos = pos.reshape(S, N, 4)
        vel = None if vel is None else _f32(vel, pos.shape)
        roots = None if rest_roots is None else _f32(rest_roots, (S, 4))
        d = _ffi.GroomDesc(S, N, pos.ctypes.data, None if vel is None else vel.ctypes.data,
                           None if roots is None else roots.ctypes.data)
        p = params if params is not None else reference_params()
        cid, res = C.c_int(0), C.c_int(0)
        check(lib().ftl_server_add(self._h, C.byref(d), C.byref(p), C.byref(cid), C.byref(res)), "ftl_server_add")
        self.shapes[cid.value] = (S, N)
        return cid.value, bool(res.value)

    def remove(self, cid):
        check(lib().ftl_server_remove(self._h, int(cid)), "ftl_server_remove")
        self.shapes.pop(cid, None)

    def resident(self, cid):
        r = C.c_int(0)
        check(lib().ftl_server_resident(self._h, int(cid), C.byref(r)), "ftl_server_resident")
        return bool(r.value)

    def submit(self, ids, reqs):
        """Queues requests (REQUEST_DTYPE records) for creatures ids; returns tickets."""
        ids = np.ascontiguousarray(np.atleast_1d(ids), np.int32)
        reqs = np.ascontiguousarray(np.atleast_1d(reqs), REQUEST_DTYPE)
        if ids.shape[0] != reqs.shape[0]:
            raise ValueError("ids and requests differ in length")
        tickets = np.zeros(ids.shape[0], np.uint64)
        check(lib().ftl_server_submit(self._h, int(ids.shape[0]), ids.ctypes.data, reqs.ctypes.data,
                                      tickets.ctypes.data), "ftl_server_submit")
        return tickets

    def step(self, cid, dt, time=0.0, model=None, spheres=None):
        r = make_requests(1, dt, time)
        set_request(r[0], model=model, spheres=spheres)
        return int(self.submit([cid], r)[0])

    def wait(self, ticket, timeout_ms=-1):
        check(lib().ftl_server_wait(self._h, int(ticket), int(timeout_ms)), "ftl_server_wait")

    def drain(self, timeout_ms=-1):
        check(lib().ftl_server_drain(self._h, int(timeout_ms)), "ftl_server_drain")

    def download(self, cid):
        if cid not in self.shapes:  # let the library report the unknown creature
            check(lib().ftl_server_download(self._h, int(cid), None, None), "ftl_server_download")
        S, N = self.shapes[cid]
        pos = np.empty((S, N, 4), np.float32)
        vel = np.empty_like(pos)
        check(lib().ftl_server_download(self._h, int(cid), _ptr(pos), _ptr(vel)), "ftl_server_download")
        return pos, vel

    def status(self, cid):
        """True when the creature is quarantined (a step left non-finite values)."""
        nf = C.c_int(0)
        check(lib().ftl_server_status(self._h, int(cid), C.byref(nf)), "ftl_server_status")
        return bool(nf.value)

    def upload(self, cid, pos, vel):
        """Replaces the creature's state and lifts its quarantine (ftl_server_upload)."""
        S, N = self.shapes[cid]
        pos = _f32(pos, (S, N, 4))
        vel = _f32(vel, (S, N, 4))
        check(lib().ftl_server_upload(self._h, int(cid), _ptr(pos), _ptr(vel)), "ftl_server_upload")

    def stats(self):
        st = _ffi.ServerStats()
        check(lib().ftl_server_get_stats(self._h, C.byref(st)), "ftl_server_get_stats")
        return {name: getattr(st, name) for name, _ in st._fields_}

    def reset_stats(self):
        check(lib().ftl_server_reset_stats(self._h), "ftl_server_reset_stats")

    def close(self):
        if self._h:
            lib().ftl_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory owned by the context (ftl_diag_alloc), for outputs such as
    the expansion's vertices."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, int(nbytes)
        self.ptr = C.c_void_p()
        check(lib().ftl_diag_alloc(ctx.handle, self.nbytes, C.byref(self.ptr)), "ftl_diag_alloc")

    def download(self, dtype=np.float32, shape=None):
        from . import _hip
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        _hip.copy_to_host(out, self.ptr)
        return out if shape is None else out.reshape(shape)

    def close(self):
        if self.ptr:
            lib().ftl_diag_free(self.ctx.handle, self.ptr, self.nbytes)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def expand_params(ninstances=None, nlines=None, nsubsegments=None, tangent_scale=None):
    """Tessellation settings of the stream pass; unset fields keep the reference defaults."""
    p = _ffi.ExpandParams()
    lib().ftl_expand_params_reference(C.byref(p))
    if ninstances is not None:
        p.ninstances = int(ninstances)
    if nlines is not None:
        p.nlines = int(nlines)
    if nsubsegments is not None:
        p.nsubsegments = int(nsubsegments)
    if tangent_scale is not None:
        p.tangent_scale = float(tangent_scale)
    return p


class Expand:
    """Render-side strand expansion (ftl_expand_*, include/ftl_render.h): the
    tessellation stream pass of the reference renderer on the device."""

    def __init__(self, ctx, nstrands, npoints, tris, tangents, random):
        self.ctx = ctx
        self.nstrands, self.npoints = int(nstrands), int(npoints)
        self.tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        tangents = _f32(tangents).reshape(self.nstrands * self.npoints, 4)
        self.random = _f32(random).reshape(-1, 4)
        self._h = C.c_void_p()
        check(lib().ftl_expand_create(ctx.handle, self.nstrands, self.npoints, _ptr(self.tris), self.tris.shape[0],
                                      _ptr(tangents), _ptr(self.random), self.random.shape[0], C.byref(self._h)),
              "ftl_expand_create")

    @property
    def handle(self):
        return self._h

    def set_tangents(self, tangents):
        tangents = _f32(tangents).reshape(self.nstrands * self.npoints, 4)
        check(lib().ftl_expand_set_tangents(self._h, _ptr(tangents)), "ftl_expand_set_tangents")

    def reference_tangents(self, normals, maxlength, curly=True):
        """The reference's initial tangents, computed on the device (ftl_expand_reference_tangents)."""
        normals = _f32(normals).reshape(self.nstrands, 3)
        check(lib().ftl_expand_reference_tangents(self._h, _ptr(normals), float(maxlength), 1 if curly else 0),
              "ftl_expand_reference_tangents")

    def tangents(self):
        """The current tangents, float32 [nstrands, npoints, 4] (ftl_expand_get_tangents)."""
        out = np.empty((self.nstrands, self.npoints, 4), np.float32)
        check(lib().ftl_expand_get_tangents(self._h, _ptr(out)), "ftl_expand_get_tangents")
        return out

    def vertex_count(self, params=None):
        p = params or expand_params()
        return int(lib().ftl_expand_vertex_count(self._h, C.byref(p)))

    def run(self, positions, out, capacity, params=None, stream=None):
        """positions, out: device pointers; capacity in vertices."""
        p = params or expand_params()
        check(lib().ftl_expand_run(self._h, C.byref(p), positions, out, int(capacity), stream), "ftl_expand_run")

    def run_groom(self, groom, params=None, stream=None):
        """Expands a groom's current positions; returns float32 [vertices, 4]."""
        from . import _hip
        n = self.vertex_count(params)
        buf = DeviceBuffer(self.ctx, max(n, 1) * 16)
        try:
            pos, _ = groom.device_buffers()
            self.run(pos, buf.ptr, n, params, stream)
            _hip.device_synchronize()
            return buf.download(np.float32, (n, 4))
        finally:
            buf.close()

    def close(self):
        if self._h:
            lib().ftl_expand_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def marschner_params(**kw):
    """Marschner shading parameters (reference defaults, then the given fields)."""
    p = _ffi.MarschnerParams()
    lib().ftl_marschner_params_reference(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, float(v))
    return p


def marschner_luts(ctx, params=None, res=128, stream=None):
    """Marschner::generate on the device (ftl_marschner_luts): (M, N), float32 [res, res, 4] each."""
    from . import _hip
    p = params or marschner_params()
    nb = res * res * 16
    m, n = DeviceBuffer(ctx, nb), DeviceBuffer(ctx, nb)
    try:
        check(lib().ftl_marschner_luts(ctx.handle, C.byref(p), int(res), m.ptr, n.ptr, stream), "ftl_marschner_luts")
        _hip.device_synchronize()
        return m.download(np.float32, (res, res, 4)), n.download(np.float32, (res, res, 4))
    finally:
        m.close()
        n.close()


# --- scenario generators (CPU) ----------------------------------------------

def strand_tangents(pos):
    """Central-difference tangents of strands [S, N, 4] (ftl_scn_strand_tangents)."""
    pos = _f32(pos)
    S, N = pos.shape[0], pos.shape[1]
    out = np.empty_like(pos)
    check(lib().ftl_scn_strand_tangents(_ptr(pos), S, N, _ptr(out)), "ftl_scn_strand_tangents")
    return out


def random_table(seed=0, n=4096):
    """The renderer's random table, seeded (ftl_scn_random_table): float32 [n, 4]."""
    out = np.empty((n, 4), np.float32)
    check(lib().ftl_scn_random_table(int(seed), int(n), _ptr(out)), "ftl_scn_random_table")
    return out


def sphere_groom(nstrands, npoints, center=(0, 0, 0), radius=1.0, length=0.2, jitter=0.05, seed=0,
                 coherent=True):
    """Strands on a sphere (ftl_scn_sphere_groom): (pos, vel, rest_roots)."""
    d = _ffi.ScnSphere(int(nstrands), int(npoints), (C.c_float * 3)(*center), float(radius), float(length),
                       float(jitter), int(seed), 1 if coherent else 0)
    pos = np.empty((nstrands, npoints, 4), np.float32)
    vel = np.empty_like(pos)
    roots = np.empty((nstrands, 4), np.float32)
    check(lib().ftl_scn_sphere_groom(C.byref(d), _ptr(pos), _ptr(vel), _ptr(roots)), "ftl_scn_sphere_groom")
    return pos, vel, roots


def reference_groom_device(ctx, roots, normals, npoints=4, maxlength=0.5, seed=0):
    """Hair::init_simulation's state built on the device (ftl_groom_create_reference): a Groom."""
    roots = _f32(roots).reshape(-1, 3)
    normals = _f32(normals).reshape(-1, 3)
    h = C.c_void_p()
    check(lib().ftl_groom_create_reference(ctx.handle, roots.shape[0], _ptr(roots), _ptr(normals), int(npoints),
                                           float(maxlength), int(seed), C.byref(h)), "ftl_groom_create_reference")
    return Groom(ctx, (roots.shape[0], int(npoints)), _handle=h)


def reference_groom(roots, normals, npoints=4, maxlength=0.5, seed=0):
    """Hair::init_simulation with a seeded stream (ftl_scn_reference_groom)."""
    roots = _f32(roots).reshape(-1, 3)
    normals = _f32(normals).reshape(-1, 3)
    n = roots.shape[0]
    pos = np.empty((n, npoints, 4), np.float32)
    vel = np.empty_like(pos)
    check(lib().ftl_scn_reference_groom(n, _ptr(roots), _ptr(normals), int(npoints), float(maxlength), int(seed),
                                        _ptr(pos), _ptr(vel)), "ftl_scn_reference_groom")
    return pos, vel


def _skinned_desc(nstrands, npoints, ntris, nbones, radius, length, jitter, max_angle, seed):
    return _ffi.ScnSkinned(int(nstrands), int(npoints), int(ntris), int(nbones), float(radius), float(length),
                           float(jitter), float(max_angle), int(seed))


def skinned_scene(nstrands, npoints, ntris=100_000, nbones=64, radius=1.0, length=0.2, jitter=0.05,
                  max_angle=np.deg2rad(30.0), seed=0):
    """Skinned emitter scene: dict with mesh arrays, binding and groom state."""
    d = _skinned_desc(nstrands, npoints, ntris, nbones, radius, length, jitter, max_angle, seed)
    nv, nt = C.c_int(0), C.c_int(0)
    check(lib().ftl_scn_skinned_sizes(C.byref(d), C.byref(nv), C.byref(nt)), "ftl_scn_skinned_sizes")
    rest = np.empty((nv.value, 4), np.float32)
    joints = np.empty((nv.value, 4), np.int32)
    weights = np.empty((nv.value, 4), np.float32)
    tris = np.empty((nt.value, 3), np.int32)
    check(lib().ftl_scn_skinned_mesh(C.byref(d), _ptr(rest), _ptr(joints), _ptr(weights), _ptr(tris)),
          "ftl_scn_skinned_mesh")
    tri = np.empty(nstrands, np.int32)
    bary = np.empty((nstrands, 2), np.float32)
    pos = np.empty((nstrands, npoints, 4), np.float32)
    vel = np.empty_like(pos)
    check(lib().ftl_scn_skinned_groom(C.byref(d), _ptr(rest), _ptr(tris), _ptr(tri), _ptr(bary), _ptr(pos),
                                      _ptr(vel)), "ftl_scn_skinned_groom")
    return dict(desc=d, rest=rest, joints=joints, weights=weights, tris=tris, tri=tri, bary=bary, pos=pos,
                vel=vel, nbones=nbones)


def bones(scene, t):
    """Bone dual quaternions of a skinned scene at time t: float32 [nbones, 8]."""
    d = scene["desc"]
    dq = np.empty((d.nbones, 8), np.float32)
    check(lib().ftl_scn_bones(C.byref(d), float(t), _ptr(dq)), "ftl_scn_bones")
    return dq


from .hair import Hair  # noqa: E402
