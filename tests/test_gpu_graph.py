"""Captured substeps (mpmx_config.use_graph): replays of the two substep
graphs give the same bits as the eager launches, replay after replay, with
eager work and host synchronisation between the replays; for one scene and a
mixed-precision batch, MLS and APIC."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "B", "c", "alpha", "psi", "g")


def pair(mpmx, kind):
    if kind == "batch_mixed":
        ps, counts, x, v, m, dt = mpmx.scene_batch(4, seed=4, scale=0.004, num_scenes=5)
        out = []
        for graph in (1, 0):
            ctx = mpmx.Context(ps, mpmx.default_config(precision=mpmx.MIXED, use_graph=graph))
            ctx.set_particles_batch(counts, x, v, m)
            out.append(ctx)
        return out, dt
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    p.transfer = 1 if kind == "apic" else 0
    out = []
    for graph in (1, 0):
        ctx = mpmx.Context(p, mpmx.default_config(use_graph=graph))
        ctx.set_particles(x, v, m)
        out.append(ctx)
    return out, dt


@pytest.mark.parametrize("kind", ["mls", "apic", "batch_mixed"])
def test_graph_replays_equal_eager(mpmx, kind):
    (g, e), dt = pair(mpmx, kind)
    for rep in range(10):
        g.run(3, dt)
        e.run(3, dt)
        sg, se = g.stats(), e.stats()   # eager work and syncs between replays
        assert sg["cg_iters"] == se["cg_iters"] and sg["numPage"] == se["numPage"], rep
        for k in FIELDS:
            assert np.array_equal(g.get(k), e.get(k)), (rep, k)
    assert g.stats()["substeps"] == 30
