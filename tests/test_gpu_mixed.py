"""Mixed-precision storage against the FP64 oracle.

MPMX_MIXED keeps x and F in FP32 and stores v, C, c, alpha, psi and g in
BF16 (FP32 grid accumulation); MPMX_MIXED_F16 also stores F as FP16(F - I).

Bars (max |a - b| / max |b| per field, c and alpha absolute):
* one substep from the GPU's own yielded, damaged impact state loaded into
  the oracle: x <= 1e-5 and F <= 1e-4 for MPMX_MIXED (the 1e-4 bar of the
  mixed form); the BF16 fields within their storage rounding (2^-9 per
  store, bars 4e-3 / 8e-3);
* 60-substep trajectories: x <= 1e-4; F for MPMX_MIXED_F16 within 3x what
  the same codecs do to the oracle's own trajectory (FP16 holds |F - I| to
  2^-11 relative, so where plastic flow drives |F - I| to ~0.5 one store
  costs ~2.4e-4 and the impact amplifies it).
"""
import numpy as np
import pytest

from conftest import max_rel_field, oracle_params

pytestmark = pytest.mark.gpu

STATE = ("x", "v", "F", "B", "c", "alpha", "psi", "g")
BARS_MIXED = {"x": 1e-5, "v": 4e-3, "F": 1e-4, "B": 8e-3, "psi": 4e-3, "c": 4e-3, "alpha": 8e-3}


def run(mpmx, precision, scene_id, scale, steps):
    p, x, v, m, dt = mpmx.scene(scene_id, seed=1, scale=scale)
    ctx = mpmx.Context(p, mpmx.default_config(precision=precision))
    ctx.set_particles(x, v, m)
    ctx.run(steps, dt)
    ctx.sync()
    return ctx, p, x, v, m, dt


def bf16(a):
    """float32 -> bfloat16 -> float, round to nearest even (the device codec)."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def round_mixed(sim, f16):
    """Store the oracle's state through the mixed-precision codecs."""
    sim.set("x", sim.get("x").astype(np.float32))
    I = np.eye(3)[None]
    F = sim.get("F")
    sim.set("F", (F - I).astype(np.float32).astype(np.float16).astype(np.float64) + I if f16
            else F.astype(np.float32))
    for k in ("v", "B", "c", "alpha", "psi", "g"):
        sim.set(k, bf16(sim.get(k)))


@pytest.mark.parametrize("transfer", [0, 1])
def test_mixed_one_step_from_gpu_state(mpmx, oracle, transfer):
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    p.transfer = transfer
    ctx = mpmx.Context(p, mpmx.default_config(precision=mpmx.MIXED))
    ctx.set_particles(x, v, m)
    ctx.run(45, dt)
    st = {k: ctx.get(k) for k in STATE}
    assert (st["alpha"] > 0).sum() > 100 and st["c"].min() < 0.99
    sim = oracle.Sim(oracle_params(oracle, p), st["x"], st["v"], ctx.get("mass"))
    for k in ("F", "B", "c", "alpha", "psi", "g"):
        sim.set(k, st[k])
    ctx.substep(dt)
    sim.substep(dt)
    errs = {k: max_rel_field(ctx.get(k), sim.get(k)) for k in ("x", "v", "F", "B", "psi")}
    for k in ("c", "alpha"):
        errs[k] = float(np.abs(ctx.get(k) - sim.get(k)).max())
    print("mixed one step (transfer %d):" % transfer, {k: "%.1e" % e for k, e in errs.items()})
    for k, e in errs.items():
        assert e < BARS_MIXED[k], k


@pytest.mark.parametrize("form,scene_id,scale,steps", [("mixed", 0, 0.5, 20), ("mixed", 3, 0.1, 60),
                                                      ("mixed_f16", 0, 0.5, 20), ("mixed_f16", 3, 0.1, 60)])
def test_mixed_trajectory_vs_oracle(mpmx, oracle, form, scene_id, scale, steps):
    f16 = form == "mixed_f16"
    ctx, p, x, v, m, dt = run(mpmx, mpmx.MIXED_F16 if f16 else mpmx.MIXED, scene_id, scale, steps)
    args = (oracle_params(oracle, p), x.astype(np.float32), v.astype(np.float32), m.astype(np.float32))
    sim, emu = oracle.Sim(*args), oracle.Sim(*args)
    for _ in range(steps):
        sim.substep(dt)
        emu.substep(dt)
        round_mixed(emu, f16)
    ex = max_rel_field(ctx.get("x"), sim.get("x"))
    eF = max_rel_field(ctx.get("F"), sim.get("F"))
    envF = max_rel_field(emu.get("F"), sim.get("F"))
    print("%s vs oracle: x %.2e F %.2e v %.2e c %.2e; codec-emulating oracle vs oracle: x %.2e F %.2e" % (
        form, ex, eF, max_rel_field(ctx.get("v"), sim.get("v")), np.abs(ctx.get("c") - sim.get("c")).max(),
        max_rel_field(emu.get("x"), sim.get("x")), envF))
    assert ex <= 1e-4
    assert eF <= max(1e-4, 3 * envF)


@pytest.mark.parametrize("form", ["mixed", "mixed_f16"])
def test_mixed_storage_roundtrip(mpmx, form):
    """set_field/get_field through the codecs: F exact in FP32 (MPMX_MIXED) or
    to FP16 precision of F - I; BF16 fields keep 8 significant bits."""
    p, x, v, m, dt = mpmx.scene(0, seed=1, scale=0.25)
    prec = mpmx.MIXED_F16 if form == "mixed_f16" else mpmx.MIXED
    ctx = mpmx.Context(p, mpmx.default_config(precision=prec))
    ctx.set_particles(x, v, m)
    rng = np.random.default_rng(3)
    F = np.eye(3)[None] + 1e-3 * rng.standard_normal((ctx.n, 3, 3))
    ctx.set("F", F)
    if form == "mixed":
        assert np.array_equal(ctx.get("F"), F.astype(np.float32).astype(np.float64))
    else:
        assert np.abs(ctx.get("F") - F).max() < 1e-3 * 2 ** -10 * 4
    c = rng.uniform(0, 1, ctx.n)
    ctx.set("c", c)
    assert np.abs(ctx.get("c") - c).max() <= 2 ** -8


def test_fp32_svd_with_fp64_fallback(mpmx, oracle):
    """stress_fp64 = 0: FP32 SVD and return mapping, redone in FP64 where the
    singular-value gap is below svd_gap = 1e-4 (near-degenerate F, which the
    initial F = I and every rotation are).  One substep from a yielded state
    against the oracle, and the fallback count is reported."""
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    cfg = mpmx.default_config(precision=mpmx.MIXED, stress_fp64=0, svd_gap=1e-4)
    ctx = mpmx.Context(p, cfg)
    ctx.set_particles(x, v, m)
    ctx.run(45, dt)
    st = {k: ctx.get(k) for k in STATE}
    sim = oracle.Sim(oracle_params(oracle, p), st["x"], st["v"], ctx.get("mass"))
    for k in ("F", "B", "c", "alpha", "psi", "g"):
        sim.set(k, st[k])
    ctx.substep(dt)
    sim.substep(dt)
    eF = max_rel_field(ctx.get("F"), sim.get("F"))
    ex = max_rel_field(ctx.get("x"), sim.get("x"))
    fb = ctx.svd_fallbacks()
    print("FP32 stress + FP64 fallback: x %.1e F %.1e, fallbacks %d of %d particles x 46 substeps" % (
        ex, eF, fb, ctx.n))
    assert fb > 0
    assert ex < 1e-5 and eF < 1e-4
