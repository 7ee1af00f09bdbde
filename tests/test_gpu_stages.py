"""The full substep against the stage-by-stage API.

mpmx_substep computes the stress and the P2G's per-particle terms in one
pass and builds the block source lists with the page tables; calling the
stages one by one (ExplicitTimeIntegrator's own sequence) must give the same
bits, for both transfers and the storage forms.  The fused pass stores sigma
and FP only with mpmx_config.store_stress; without it the substep's results
are the same and sigma / FP are not readable (test_stress_fields_optional).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "B", "c", "alpha", "psi", "g", "sigma", "FP", "vol")


PREC = {"fp32": 0, "mixed": 1, "mixed_f16": 2}


@pytest.mark.parametrize("precision,transfer", [("fp32", 0), ("mixed", 0), ("mixed_f16", 0), ("fp32", 1)])
def test_substep_equals_stages(mpmx, precision, transfer):
    p, x, v, m, dt = mpmx.scene(3, seed=2, scale=0.1)
    p.transfer = transfer
    cfg = mpmx.default_config(precision=PREC[precision], store_stress=1)
    fused = mpmx.Context(p, cfg)
    fused.set_particles(x, v, m)
    staged = mpmx.Context(p, cfg)
    staged.set_particles(x, v, m)
    for _ in range(40):
        fused.substep(dt)
        staged.bin_sort()
        staged.rebuild()
        staged.stress()
        staged.p2g(dt)
        staged.grid_update(dt)
        staged.phase_field(dt)
        staged.g2p(dt)
    fused.sync()
    staged.sync()
    print("staged vs fused: %d particles, %s, transfer %d" % (fused.n, precision, transfer))
    for k in FIELDS:
        assert np.array_equal(fused.get(k), staged.get(k)), k
    assert fused.stats()["cg_iters"] == staged.stats()["cg_iters"]
    assert (fused.get("alpha") > 0).any()   # the impact has yielded: the return mapping ran
    fused.close()
    staged.close()


def test_stress_fields_optional(mpmx):
    """Without store_stress the fused substep gives the same state and keeps
    sigma / FP unreadable (MPMX_ERR_ARG) until a stage-API stress pass."""
    p, x, v, m, dt = mpmx.scene(3, seed=2, scale=0.1)
    a = mpmx.Context(p, mpmx.default_config(store_stress=1))
    b = mpmx.Context(p, mpmx.default_config())
    for c in (a, b):
        c.set_particles(x, v, m)
        c.run(10, dt)
        c.sync()
    for k in ("x", "v", "F", "B", "c", "alpha", "psi", "g", "vol"):
        assert np.array_equal(a.get(k), b.get(k)), k
    with pytest.raises(mpmx.MpmxError) as e:
        b.get("sigma")
    assert e.value.status == mpmx.ERR_ARG
    b.bin_sort()
    b.rebuild()
    b.stress()
    assert np.isfinite(b.get("sigma")).all() and np.isfinite(b.get("FP")).all()
