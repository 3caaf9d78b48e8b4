"""Shared helpers for the tests: mirror a groom's step into oracle parameters
and compare results."""
import numpy as np

from oracle import ftl_oracle as orc


def oracle_params_for(params, npoints, dt, *, spheres, model=None, root_mode=0, time=0.0, wind=None):
    """Oracle parameters for one library step with ftl params `params` (a
    barbu_amd.Params), colliders `spheres`, 4x4 row-major root transform
    `model`, root mode and wind-field time."""
    kw = dict(iterations=params.iterations, velocity_mode=params.velocity_mode, root_mode=root_mode,
              length_scale=params.length_scale, damping=params.damping, max_velocity=params.max_velocity,
              gravity=tuple(params.gravity), force_scale=params.force_scale, spheres=spheres, model=model)
    if params.wind_amplitude != 0.0:
        off = tuple(np.float32(float(np.float32(s)) * float(time)) for s in params.wind_scroll)
        kw.update(wind_table=wind, wind_frequency=params.wind_frequency, wind_offset=off)
    return orc.make_params(npoints, dt, **kw)


def max_abs(a, b):
    return float(np.max(np.abs(a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)))) if a.size else 0.0


def assert_bitexact(gpu, ref, what):
    g = np.ascontiguousarray(gpu).view(np.uint32)
    r = np.ascontiguousarray(ref).view(np.uint32)
    same = (g == r) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        idx = np.argwhere(~same)
        i = tuple(idx[0])
        raise AssertionError("%s: %d of %d values differ; first at %s: gpu=%r oracle=%r (max |d| = %g)" % (
            what, int((~same).sum()), same.size, i, gpu[i], ref[i],
            np.nanmax(np.abs(gpu.astype(np.float64) - ref.astype(np.float64)))))
