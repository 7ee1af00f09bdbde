"""The full substep against the stage-by-stage API.

mpmx_substep computes the stress and the P2G's per-particle terms in one
pass and builds the block source lists with the page tables; calling the
stages one by one (ExplicitTimeIntegrator's own sequence) must give the same
bits, for both transfers and both storage forms.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "B", "c", "alpha", "psi", "g", "sigma", "FP", "vol")


@pytest.mark.parametrize("precision,transfer", [("fp32", 0), ("mixed", 0), ("fp32", 1)])
def test_substep_equals_stages(mpmx, precision, transfer):
    p, x, v, m, dt = mpmx.scene(3, seed=2, scale=0.1)
    p.transfer = transfer
    cfg = mpmx.default_config(precision=mpmx.MIXED if precision == "mixed" else mpmx.FP32)
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
