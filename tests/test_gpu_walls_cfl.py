"""Sticky walls and the device-side CFL step against the FP64 oracle.

Walls: a batch of six small cubes, each thrown at one of the six faces of
its own domain from just outside the 4-cell wall layer, so that every
substep stops grid nodes at i <= 4 or i > N - 4 along each axis and sign
(collideWithGround, TimeIntegrator.cu:117-165).  After 20 substeps the GPU's
state is loaded into the oracle and both take one substep.

CFL: mpmx_max_velocity_sq and mpmx_compute_dt against the oracle's
computeDt (Simulator.cpp:200-225) on the GPU's own velocities, with and
without the wave-speed term, and mpmx_run_cfl (dt on the device, no host
round trip) against the host loop computeDt -> substep, bit for bit.
"""
import numpy as np
import pytest

from conftest import max_rel_field, oracle_params

pytestmark = pytest.mark.gpu

STATE = ("x", "v", "F", "B", "c", "alpha", "psi", "g")


def wall_batch(mpmx, N=64, speed=4.0):
    """Six cubes of 6^3 cells x 8 ppc; cube a moves at `speed` towards face a
    (x-, x+, y-, y+, z-, z+).  The leading particles' stencils reach node 4
    (low faces, leading face 4.6 cells from the boundary) or node N - 3
    (high faces, 3.6 cells), the first nodes inside each wall layer."""
    rng = np.random.default_rng(11)
    dx = 1.0 / N
    ps, xs, vs, ms = [], [], [], []
    for a in range(6):
        axis, sign = a // 2, (-1 if a % 2 == 0 else 1)
        p = mpmx.default_params()
        p.N = N
        p.V0 = dx ** 3 / 8
        p.E = 2e4
        p.yield_stress = 200.0
        lo = np.full(3, 0.5 - 3 * dx)
        lo[axis] = 4.6 * dx if sign < 0 else 1.0 - 3.6 * dx - 6 * dx
        cells = np.stack(np.meshgrid(*[np.arange(6)] * 3, indexing="ij"), -1).reshape(-1, 3)
        sub = np.stack(np.meshgrid(*[np.array([0.25, 0.75])] * 3, indexing="ij"), -1).reshape(-1, 3)
        pts = (cells[:, None, :] + sub[None]).reshape(-1, 3)
        pts = lo + (pts + rng.uniform(-0.125, 0.125, pts.shape)) * dx
        v = np.zeros_like(pts)
        v[:, axis] = sign * speed
        ps.append(p)
        xs.append(pts)
        vs.append(v)
        ms.append(np.full(len(pts), 1000.0 * p.V0))
    counts = np.array([len(q) for q in xs], np.int64)
    return ps, counts, np.concatenate(xs), np.concatenate(vs), np.concatenate(ms)


def test_walls_one_step(mpmx, oracle):
    ps, counts, x, v, m = wall_batch(mpmx)
    dt = 1e-4
    ctx = mpmx.Context(ps)
    ctx.set_particles_batch(counts, x, v, m)
    ctx.run(20, dt)
    st = {k: ctx.get(k) for k in STATE}
    mass = ctx.get("mass")
    ctx.substep(dt)
    after = {k: ctx.get(k) for k in STATE}
    start = np.concatenate([[0], np.cumsum(counts)])
    bars = {"x": 1e-6, "v": 1e-4, "F": 1e-5, "B": 1e-3}
    for s, p in enumerate(ps):
        sl = slice(int(start[s]), int(start[s + 1]))
        axis, sign = s // 2, (-1 if s % 2 == 0 else 1)
        # the walls have acted: the leading particles have slowed down
        assert (after["v"][sl, axis] * sign).min() < 0.9 * 4.0
        sim = oracle.Sim(oracle_params(oracle, p), st["x"][sl], st["v"][sl], mass[sl])
        for k in ("F", "B", "c", "alpha", "psi", "g"):
            sim.set(k, st[k][sl])
        sim.substep(dt)
        errs = {k: max_rel_field(after[k][sl], sim.get(k)) for k in bars}
        print("wall %d (axis %d, sign %+d):" % (s, axis, sign), {k: "%.1e" % e for k, e in errs.items()})
        for k, e in errs.items():
            assert e < bars[k], (s, k, e)


def test_cfl_against_oracle(mpmx, oracle):
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    ctx = mpmx.Context(p)
    ctx.set_particles(x, v, m)
    ctx.run(10, dt)
    xg, vg, mg = ctx.get("x"), ctx.get("v"), ctx.get("mass")
    sim = oracle.Sim(oracle_params(oracle, p), xg, vg, mg)
    vmax2 = ctx.max_velocity_sq()
    dt_o, sp = sim.cfl_dt(1.0, 0.3)
    assert vmax2 == pytest.approx(sp * sp, rel=1e-6)
    for dflt in (1.0, 1e-6):   # CFL-limited, then the default
        assert ctx.compute_dt(dflt, 0.3) == pytest.approx(sim.cfl_dt(dflt, 0.3)[0], rel=1e-6)
    # with the elastic wave speed: run one CFL substep and read its step back
    dtw_o, spw = sim.cfl_dt(1.0, 0.3, wave=True)
    assert spw > sp
    t = ctx.run_cfl(1, 1.0, 0.3, wave=True)
    assert t == pytest.approx(dtw_o, rel=1e-6)


def test_batch_compute_dt_uses_each_scene_grid(mpmx, oracle):
    ps, counts, x, v, m, dt = mpmx.scene_batch(5, seed=5, num_scenes=12)
    ctx = mpmx.Context(ps)
    ctx.set_particles_batch(counts, x, v, m)
    ctx.run(3, dt)
    vg, xg, mg = ctx.get("v"), ctx.get("x"), ctx.get("mass")
    start = np.concatenate([[0], np.cumsum(counts)])
    expect = 1.0
    for s, p in enumerate(ps):
        sl = slice(int(start[s]), int(start[s + 1]))
        sim = oracle.Sim(oracle_params(oracle, p), xg[sl], vg[sl], mg[sl])
        expect = min(expect, sim.cfl_dt(1.0, 0.3)[0])
    assert ctx.compute_dt(1.0, 0.3) == pytest.approx(expect, rel=1e-6)


def test_run_cfl_equals_host_loop(mpmx):
    """Device-side dt (no host round trip) against computeDt -> substep on the
    host, bit for bit: both hand the kernels the same FP32 step."""
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    a, b = mpmx.Context(p), mpmx.Context(p)
    a.set_particles(x, v, m)
    b.set_particles(x, v, m)
    t = a.run_cfl(12, 2e-4, 0.3)
    total = 0.0
    for _ in range(12):
        d = b.compute_dt(2e-4, 0.3)
        b.substep(d)
        total += float(np.float32(d))
    a.sync()
    b.sync()
    assert t == pytest.approx(total, rel=1e-12)
    for k in ("x", "v", "F", "c"):
        assert np.array_equal(a.get(k), b.get(k)), k
