"""mpmx_compute_dt_noalloc (static device scratch, no allocation) against
mpmx_compute_dt (context-owned partials), bit for bit, on two calls in a
row, the second with a smaller context, then on a batch of scenes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def noalloc(mpmx, ctx, dt_default, cfl):
    L = mpmx.lib()
    fn = L.mpmx_compute_dt_noalloc
    fn.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double)]
    fn.restype = C.c_int
    out = C.c_double(-1.0)
    st = fn(ctx.h, dt_default, cfl, C.byref(out))
    assert st == mpmx.OK, (st, L.mpmx_last_error(ctx.h))
    return out.value


def test_noalloc_equals_compute_dt(mpmx):
    # first call: the impact (8,192 particles) after 10 substeps
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    big = mpmx.Context(p)
    big.set_particles(x, v, m)
    big.run(10, dt)
    for dflt in (1.0, 1e-7):   # CFL-limited, then the default
        want = big.compute_dt(dflt, 0.3)
        got = noalloc(mpmx, big, dflt, 0.3)
        assert got == want, (dflt, got, want)
    # second call, a smaller context (1,000 particles, another grid)
    p2, x2, v2, m2, dt2 = mpmx.scene(0, seed=1, scale=0.25)
    small = mpmx.Context(p2)
    small.set_particles(x2, v2 + np.array([0.3, -1.2, 0.7]), m2)
    small.run(3, dt2)
    assert len(m2) < len(m)
    want = small.compute_dt(1.0, 0.3)
    got = noalloc(mpmx, small, 1.0, 0.3)
    assert got == want and got < 1.0
    # and the first context again: the shared scratch holds nothing stale
    assert noalloc(mpmx, big, 1.0, 0.3) == big.compute_dt(1.0, 0.3)


def test_noalloc_batch(mpmx):
    ps, counts, x, v, m, dt = mpmx.scene_batch(5, seed=5, num_scenes=12)
    ctx = mpmx.Context(ps)
    ctx.set_particles_batch(counts, x, v, m)
    ctx.run(3, dt)
    got = noalloc(mpmx, ctx, 1.0, 0.3)
    assert got == ctx.compute_dt(1.0, 0.3) and got < 1.0
