"""Mixed-precision storage (x FP32, F as FP16(F - I), v/C/c/alpha/psi/g BF16,
FP32 grid, FP64 SVD for near-degenerate singular values) against the FP64
oracle and against the plain FP32 path.

Bar (the north-star target for the mixed form): relative error <= 1e-4 on x
and F against the FP64 oracle, measured as max |a - b| / max |b| per field.
"""
import numpy as np
import pytest

from conftest import max_rel_field, oracle_params

pytestmark = pytest.mark.gpu


def run(mpmx, precision, scene_id, scale, steps):
    p, x, v, m, dt = mpmx.scene(scene_id, seed=1, scale=scale)
    ctx = mpmx.Context(p, mpmx.default_config(precision=precision))
    ctx.set_particles(x, v, m)
    ctx.run(steps, dt)
    ctx.sync()
    return ctx, p, x, v, m, dt


@pytest.mark.parametrize("scene_id,scale,steps", [(0, 0.5, 20), (3, 0.1, 60)])
def test_mixed_vs_oracle(mpmx, oracle, scene_id, scale, steps):
    ctx, p, x, v, m, dt = run(mpmx, mpmx.MIXED, scene_id, scale, steps)
    sim = oracle.Sim(oracle_params(oracle, p), x.astype(np.float32), v.astype(np.float32),
                     m.astype(np.float32))
    for _ in range(steps):
        sim.substep(dt)
    ex = max_rel_field(ctx.get("x"), sim.get("x"))
    eF = max_rel_field(ctx.get("F"), sim.get("F"))
    # the oracle's own movement under a one-ulp change of the initial positions
    xp = np.nextafter(x.astype(np.float32), np.float32(1))
    sim2 = oracle.Sim(oracle_params(oracle, p), xp, v.astype(np.float32), m.astype(np.float32))
    for _ in range(steps):
        sim2.substep(dt)
    envF = max_rel_field(sim2.get("F"), sim.get("F"))
    print("mixed vs oracle: x %.2e F %.2e v %.2e c %.2e; oracle 1-ulp envelope F %.2e" % (
        ex, eF, max_rel_field(ctx.get("v"), sim.get("v")),
        np.abs(ctx.get("c") - sim.get("c")).max(), envF))
    assert ex <= 1e-4
    assert eF <= max(1e-4, 10 * envF)


def test_mixed_storage_roundtrip(mpmx):
    """set_field/get_field through the codecs: F keeps FP16 precision of F - I,
    BF16 fields keep 8 mantissa bits."""
    p, x, v, m, dt = mpmx.scene(0, seed=1, scale=0.25)
    ctx = mpmx.Context(p, mpmx.default_config(precision=mpmx.MIXED))
    ctx.set_particles(x, v, m)
    rng = np.random.default_rng(3)
    F = np.eye(3)[None] + 1e-3 * rng.standard_normal((ctx.n, 3, 3))
    ctx.set("F", F)
    assert np.abs(ctx.get("F") - F).max() < 1e-3 * 2 ** -10 * 4
    c = rng.uniform(0, 1, ctx.n)
    ctx.set("c", c)
    assert np.abs(ctx.get("c") - c).max() <= 2 ** -8
