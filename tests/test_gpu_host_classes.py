"""The reference-named C++ classes (MPMSimulator & co., host/mpmx_classes.*)
driven the way the reference's viewer drives them (one simulateStick per
frame): their state equals the C ABI called directly, bit for bit, and
tracks the serial FP64 oracle along the same trajectory; with the CFL step
on and the per-frame readback off, the frames equal mpmx_run_cfl."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, max_rel_field, oracle_params

pytestmark = pytest.mark.gpu


def read_partio(path):
    with open(path, "rb") as f:
        assert f.read(8) == b"MPMXPRT1"
        n = int(np.frombuffer(f.read(8), np.uint64)[0])
        data = np.frombuffer(f.read(), np.float64)
    x = data[: 3 * n].reshape(n, 3)
    v = data[3 * n: 6 * n].reshape(n, 3)
    c = data[6 * n: 7 * n]
    a = data[7 * n: 8 * n]
    return x, v, c, a


def demo(tmp_path, *extra):
    out = str(tmp_path / "demo.bin")
    exe = os.path.join(PKG, "lib", "mpmx_demo")
    r = subprocess.run([exe, "3", "30", "1", "0.1", out, *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return read_partio(out)


def test_demo_driver_matches_c_abi_and_oracle(mpmx, oracle, tmp_path):
    xd, vd, cd, ad = demo(tmp_path)
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    ctx = mpmx.Context(p)
    ctx.set_particles(x, v, m)
    ctx.run(30, dt)
    # the same kernels in the same order: the same bits
    assert np.array_equal(xd, ctx.get("x")) and np.array_equal(vd, ctx.get("v"))
    assert np.array_equal(cd, ctx.get("c")) and np.array_equal(ad, ctx.get("alpha"))
    # and the oracle's trajectory of the same FP32 inputs (the 30-substep
    # trajectory bars of test_gpu_parity: x to 1e-4 of its scale)
    sim = oracle.Sim(oracle_params(oracle, p), x.astype(np.float32), v, m.astype(np.float32))
    for _ in range(30):
        sim.substep(dt)
    assert max_rel_field(xd, sim.get("x")) < 1e-4
    assert np.abs(cd - sim.get("c")).max() < 1e-2


def test_demo_cfl_frames_equal_run_cfl(mpmx, tmp_path):
    xd, vd, cd, ad = demo(tmp_path, "cfl")
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    ctx = mpmx.Context(p)
    ctx.set_particles(x, v, m)
    t = ctx.run_cfl(30, dt, 0.3)
    assert 0 < t <= 30 * dt * (1 + 1e-7)
    assert np.array_equal(xd, ctx.get("x")) and np.array_equal(vd, ctx.get("v"))
