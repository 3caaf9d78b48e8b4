"""Regression fixtures (tests/golden, written by tests/golden/make_fixtures.py
from the oracle): the oracle still produces them bit for bit, FP32 stays
within 1e-5 x strand length of FP64 on every case, and the GPU path (through
the C ABI) reproduces the FP32 outputs bit for bit."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures as mf  # noqa: E402
from tests.helpers import assert_bitexact  # noqa: E402

META = json.load(open(os.path.join(HERE, "golden", "ftl_cases.json")))
ARR = np.load(os.path.join(HERE, "golden", "ftl_cases.npz"))
NAMES = [m["name"] for m in META]


def case(name):
    m = dict(next(x for x in META if x["name"] == name))
    m["pos"] = ARR[name + ".pos"]
    m["vel"] = ARR[name + ".vel"]
    m["roots"] = ARR[name + ".roots"] if m["has_roots"] else None
    if "model" in m:
        m["model"] = np.asarray(m["model"], np.float32)
    return m


def strand_length(c):
    return float(np.max(np.sum(c["pos"][:, 1:, 3], axis=1)))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    c = case(name)
    p32, v32 = mf.run_oracle(c, False)
    p64, v64 = mf.run_oracle(c, True)
    assert_bitexact(p32, ARR[name + ".pos32"], name + " fp32 positions")
    assert_bitexact(v32, ARR[name + ".vel32"], name + " fp32 velocities")
    assert np.array_equal(p64, ARR[name + ".pos64"]) and np.array_equal(v64, ARR[name + ".vel64"])
    tol = 1e-5 * strand_length(c)
    assert np.abs(p32[..., :3] - p64[..., :3]).max() <= tol
    assert np.abs(v32[..., :3] - v64[..., :3]).max() <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_fixture(ctx, name):
    import barbu_amd as ba
    c = case(name)
    g = ba.Groom(ctx, c["pos"], c["vel"], rest_roots=c["roots"])
    g.set_params(ba.make_params(**c["params"]))
    g.set_colliders(np.asarray(c["spheres"], np.float32).reshape(-1, 4))
    if "model" in c:
        g.set_root_transform(c["model"])
    if c.get("root_mode"):
        g.set_root_mode(c["root_mode"])
    g.set_time(c.get("time", 0.0))
    for _ in range(c["steps"]):
        g.step(np.float32(c.get("dt", mf.DT)))
    p, v = g.download()
    g.close()
    assert_bitexact(p, ARR[name + ".pos32"], name + " gpu positions")
    assert_bitexact(v, ARR[name + ".vel32"], name + " gpu velocities")
