"""The substep's two sort paths (mpmx_config.sort_variant): the LSD radix sort
of every key (1) and the mover sort (2: the particles whose key changed are
sorted and merged with the rest) give the same order and the same state, bit
for bit, substep after substep: eager, in captured graphs, through the public
mpmx_bin_sort, for one scene and a mixed-precision batch.  The velocities are
also scaled up so that most particles change cell each substep (the mover
sort then merges mostly movers)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "B", "c", "alpha", "psi", "g")


def contexts(mpmx, kind, graph=0, vscale=1.0):
    out = []
    if kind == "batch_mixed":
        ps, counts, x, v, m, dt = mpmx.scene_batch(4, seed=4, scale=0.004, num_scenes=5)
        for variant in (1, 2):
            ctx = mpmx.Context(ps, mpmx.default_config(precision=mpmx.MIXED, use_graph=graph,
                                                       sort_variant=variant))
            ctx.set_particles_batch(counts, x, v * vscale, m)
            out.append(ctx)
        return out, dt
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    for variant in (1, 2):
        ctx = mpmx.Context(p, mpmx.default_config(use_graph=graph, sort_variant=variant))
        ctx.set_particles(x, v * vscale, m)
        out.append(ctx)
    return out, dt


def same(a, b, tag):
    oa, ka = a.order()
    ob, kb = b.order()
    assert np.array_equal(oa, ob), tag
    assert np.array_equal(ka, kb), tag
    for k in FIELDS:
        assert np.array_equal(a.get(k), b.get(k)), (tag, k)


@pytest.mark.parametrize("kind", ["scene", "batch_mixed"])
@pytest.mark.parametrize("graph", [0, 1])
def test_mover_sort_equals_radix_sort(mpmx, kind, graph):
    (r, m), dt = contexts(mpmx, kind, graph)
    for rep in range(6):
        r.run(4, dt)
        m.run(4, dt)
        same(r, m, rep)
        # the public sort on its own, then on
        r.bin_sort()
        m.bin_sort()
        same(r, m, ("sorted", rep))
    assert r.stats()["substeps"] == 24


def test_mover_sort_with_most_particles_moving(mpmx):
    (r, m), dt = contexts(mpmx, "scene", vscale=8.0)
    for rep in range(4):
        r.substep(dt)
        m.substep(dt)
        same(r, m, rep)


def test_default_sort_matches(mpmx):
    """sort_variant 0 picks a path per substep from the mover count; the order
    is the radix sort's either way."""
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    ctxs = []
    for variant, vs in ((0, 1.0), (1, 1.0), (0, 8.0), (1, 8.0)):
        ctx = mpmx.Context(p, mpmx.default_config(sort_variant=variant))
        ctx.set_particles(x, v * vs, m)
        ctxs.append(ctx)
    for rep in range(5):
        for ctx in ctxs:
            ctx.run(2, dt)
        same(ctxs[0], ctxs[1], ("slow", rep))
        same(ctxs[2], ctxs[3], ("fast", rep))


def test_mover_sort_without_matching_tables(mpmx):
    """A sort that follows another sort with no table build in between (stage
    calls: sort, G2P, sort) places the movers by searching the previous keys
    instead of through the page tables; the order is still the radix sort's."""
    (r, m), dt = contexts(mpmx, "scene", vscale=2.0)
    for ctx in (r, m):
        ctx.run(3, dt)
    for rep in range(3):
        for ctx in (r, m):
            ctx.bin_sort()
            ctx.g2p(dt)   # moves the particles with the last grid; no table build
        for ctx in (r, m):
            ctx.bin_sort()
        same(r, m, rep)


def test_mover_sort_equals_radix_sort_full_batch(mpmx):
    """The 256-scene mixed batch at full size (12,983,932 particles): the two
    sort paths give the same order after every few substeps, and the same
    positions at the end."""
    ps, counts, x, v, m, dt = mpmx.scene_batch(5, seed=5)
    ctxs = []
    for variant in (1, 2):
        ctx = mpmx.Context(ps, mpmx.default_config(precision=mpmx.MIXED, sort_variant=variant))
        ctx.set_particles_batch(counts, x, v, m)
        ctxs.append(ctx)
    r, mv = ctxs
    for rep in range(4):
        r.run(5, dt)
        mv.run(5, dt)
        r.bin_sort()
        mv.bin_sort()
        orr, kr = r.order()
        om, km = mv.order()
        assert np.array_equal(orr, om), rep
        assert np.array_equal(kr, km), rep
    assert np.array_equal(r.get("x"), mv.get("x"))
    for ctx in ctxs:
        ctx.close()
