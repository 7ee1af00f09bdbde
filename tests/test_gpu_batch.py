"""Batched scenes: many independent scenes advanced by one context.

Bars:
  - every scene of a batch equals the same scene run alone, bit for bit, with
    the reference's block-max PCG reductions (dot_mode 0): batching changes
    the page space and the launch shapes, not one operation on a scene's
    data (fixed-point block sums are order-free; the block maxima of the PCG
    do not depend on how the partials are grouped);
  - with true dot products (dot_mode 1) a scene's sums are grouped
    differently: bound 1e-5 of each field's scale;
  - one substep of a batched scene from its own yielded state against the
    oracle, at the single-scene bars;
  - per-scene FP64 mass and momentum = the sums over that scene's particles.
"""
import numpy as np
import pytest

from conftest import max_rel_field, oracle_params

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "B", "c", "alpha", "psi", "g")


def small_batch(mpmx, num_scenes=6, scale=0.004, seed=4, models=None):
    ps, counts, x, v, m, dt = mpmx.scene_batch(4, seed=seed, scale=scale, num_scenes=num_scenes)
    for k, model in (models or {}).items():
        ps[k].model = model
    return ps, counts, x, v, m, dt


def run_batch(mpmx, ps, counts, x, v, m, dt, steps, cfg):
    ctx = mpmx.Context(ps, cfg)
    ctx.set_particles_batch(counts, x, v, m)
    ctx.run(steps, dt)
    ctx.sync()
    return ctx


def run_alone(mpmx, p, x, v, m, dt, steps, cfg):
    ctx = mpmx.Context(p, cfg)
    ctx.set_particles(x, v, m)
    ctx.run(steps, dt)
    ctx.sync()
    return ctx


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_batch_equals_scenes_alone(mpmx, precision):
    """Six scenes (cube, notched bar, sphere; grids 64^3-256^3; one scene with
    the fixed corotated energy, which has no phase field, one Neo-Hookean),
    25 substeps: each scene's state is bit-identical to the scene run alone."""
    ps, counts, x, v, m, dt = small_batch(mpmx, models={1: 0, 2: 1})
    prec = mpmx.MIXED if precision == "mixed" else mpmx.FP32
    cfg = mpmx.default_config(precision=prec)
    steps = 25
    ctx = run_batch(mpmx, ps, counts, x, v, m, dt, steps, cfg)
    print("batch: %d scenes, %d particles, N %s" % (len(ps), ctx.n, [p.N for p in ps]))
    got = {k: ctx.get(k) for k in FIELDS}
    start = np.concatenate([[0], np.cumsum(counts)])
    for s, p in enumerate(ps):
        a, b = start[s], start[s + 1]
        alone = run_alone(mpmx, p, x[a:b], v[a:b], m[a:b], dt, steps, cfg)
        for k in FIELDS:
            assert np.array_equal(got[k][a:b], alone.get(k)), (s, k)
        bs, ss = ctx.scene_stats(s), alone.stats()
        assert bs["cg_iters"] == ss["cg_iters"], s
        assert bs["numPage"] == ss["numPage"], s
        assert bs["mass"] == pytest.approx(ss["mass"], rel=1e-12)
        alone.close()
    assert ctx.stats()["cg_breakdown"] == 0


def test_batch_true_dot_mode(mpmx):
    """dot_mode 1: true inner products; a scene's partial sums are grouped by
    the batch's workgroups instead of the lone scene's, so the PCG scalars
    differ in the last bits: 1e-5 of each field's scale after 25 substeps."""
    ps, counts, x, v, m, dt = small_batch(mpmx, num_scenes=4)
    for p in ps:
        p.dot_mode = 1
    cfg = mpmx.default_config()
    ctx = run_batch(mpmx, ps, counts, x, v, m, dt, 25, cfg)
    got = {k: ctx.get(k) for k in ("x", "v", "F", "c")}
    start = np.concatenate([[0], np.cumsum(counts)])
    for s, p in enumerate(ps):
        a, b = start[s], start[s + 1]
        alone = run_alone(mpmx, p, x[a:b], v[a:b], m[a:b], dt, 25, cfg)
        for k in ("x", "v", "F"):
            assert max_rel_field(got[k][a:b], alone.get(k)) < 1e-5, (s, k)
        assert np.abs(got["c"][a:b] - alone.get("c")).max() < 1e-5
        alone.close()


def test_batch_one_step_vs_oracle(mpmx, oracle):
    """A batched notched bar (gripped: strained, yielding) and sphere (in free
    flight) 30 substeps in: their own state loaded into the oracle, one
    substep each, at the single-scene one-step bars.  The sphere translates
    rigidly, so its B and psi are rounding noise around zero and only x, v, F,
    c and alpha carry a scale to compare against."""
    ps, counts, x, v, m, dt = small_batch(mpmx, num_scenes=3, scale=0.01)
    ctx = mpmx.Context(ps)
    ctx.set_particles_batch(counts, x, v, m)
    ctx.run(30, dt)
    st = {k: ctx.get(k) for k in FIELDS}
    mass = ctx.get("mass")
    ctx.substep(dt)
    after = {k: ctx.get(k) for k in ("x", "v", "F", "B", "psi", "c", "alpha")}
    start = np.concatenate([[0], np.cumsum(counts)])
    bars = {"x": 1e-6, "v": 1e-4, "F": 1e-5, "B": 1e-3, "psi": 1e-4, "c": 1e-4, "alpha": 1e-4}
    compared = {1: ("x", "v", "F", "B", "psi"), 2: ("x", "v", "F")}
    for s in (1, 2):
        a, b = start[s], start[s + 1]
        sim = oracle.Sim(oracle_params(oracle, ps[s]), st["x"][a:b], st["v"][a:b], mass[a:b])
        for k in ("F", "B", "c", "alpha", "psi", "g"):
            sim.set(k, st[k][a:b])
        sim.substep(dt)
        errs = {k: max_rel_field(after[k][a:b], sim.get(k)) for k in compared[s]}
        for k in ("c", "alpha"):
            errs[k] = float(np.abs(after[k][a:b] - sim.get(k)).max())
        print("batched scene %d (N %d) one step:" % (s, ps[s].N), {k: "%.1e" % e for k, e in errs.items()})
        if s == 1:
            assert (st["alpha"][a:b] > 0).any()   # the bar has yielded
        for k, e in errs.items():
            assert e < bars[k], (s, k)


def test_batch_scene_stats(mpmx):
    """Per-scene statistics: counts, contiguous ranges in upload order and
    FP64 mass / momentum of exactly that scene's particles."""
    ps, counts, x, v, m, dt = small_batch(mpmx, num_scenes=5)
    ctx = mpmx.Context(ps)
    ctx.set_particles_batch(counts, x, v, m)
    ctx.run(5, dt)
    ctx.sync()
    vel = ctx.get("v")
    mass = ctx.get("mass")
    start = np.concatenate([[0], np.cumsum(counts)])
    for s in range(len(ps)):
        st = ctx.scene_stats(s)
        a, b = start[s], start[s + 1]
        assert st["num_particles"] == counts[s] and st["first_particle"] == a
        assert st["mass"] == pytest.approx(mass[a:b].astype(np.float32).astype(np.float64).sum(), rel=1e-9)
        mom = (mass[a:b, None] * vel[a:b]).sum(0)
        assert np.allclose(st["momentum"], mom, rtol=1e-6, atol=1e-9 * abs(st["mass"]))
        assert st["numPage"] > 0 and st["numTotalPage"] >= st["numPage"]
    total = ctx.stats()
    assert total["mass"] == pytest.approx(sum(ctx.scene_stats(s)["mass"] for s in range(len(ps))), rel=1e-12)
