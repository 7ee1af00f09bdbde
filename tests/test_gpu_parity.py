"""GPU parity against the serial FP64 oracle, through the C ABI.

Bars (stated per test):
  - integer work (cell keys, canonical sort order, page tables): bit-exact;
  - per-particle constitutive update on identical inputs: FP32 vs FP64,
    relative to the field's scale;
  - P2G grid sums on identical inputs: FP32 atomics vs FP64 serial sums;
  - one substep from a yielded, damaged GPU state loaded into the oracle:
    per-field bounds from FP32 rounding (the strict test of every kernel);
  - multi-substep trajectories: x within 1e-4 of its scale; v and F within
    3x of what FP32 storage alone does to the oracle's trajectory, since the
    model amplifies rounding in yielded material.
"""
import numpy as np
import pytest

from conftest import max_rel_field, oracle_params

pytestmark = pytest.mark.gpu


def make(mpmx, scene_id, scale, seed=1, **over):
    p, x, v, m, dt = mpmx.scene(scene_id, seed=seed, scale=scale)
    for k, val in over.items():
        setattr(p, k, val)
    ctx = mpmx.Context(p)
    ctx.set_particles(x, v, m)
    return ctx, p, x, v, m, dt


def oracle_tables_check(oracle, ctx, p, prev_order, x32):
    """Keys of the GPU's own float32 positions, stably sorted from prev_order,
    must give exactly the GPU's order, keys and tables."""
    keys_prev = oracle.cell_keys(x32[prev_order], p.N)
    perm, sk = oracle.sort_keys(keys_prev)
    order, keys = ctx.order()
    assert np.array_equal(order, prev_order[perm])
    assert np.array_equal(keys, sk)
    ot = oracle.page_tables(sk, p.vpage)
    gt = ctx.tables()
    for k in ("numPage", "numTotalPage", "numVirtualPage", "numCell"):
        assert gt[k] == ot[k], k
    for k in ("page_key", "page2particle", "adj", "vpo", "target", "particle2cell", "cell2particle"):
        assert np.array_equal(np.asarray(gt[k]).astype(np.int64), np.asarray(ot[k]).astype(np.int64)), k


@pytest.mark.parametrize("scene_id,scale", [(0, 0.5), (1, 0.02), (3, 0.1)])
def test_keys_sort_tables_bitexact_initial(mpmx, oracle, scene_id, scale):
    ctx, p, x, v, m, dt = make(mpmx, scene_id, scale)
    ident = np.arange(ctx.n, dtype=np.int32)
    oracle_tables_check(oracle, ctx, p, ident, x.astype(np.float32))


@pytest.mark.parametrize("vpage", [64, 128, 256])
def test_keys_sort_tables_bitexact_after_steps(mpmx, oracle, vpage):
    ctx, p, x, v, m, dt = make(mpmx, 3, 0.1, vpage=vpage)
    for _ in range(3):
        prev, _k = ctx.order()
        ctx.substep(dt)
        ctx.bin_sort()
        ctx.rebuild()
        x32 = ctx.get("x").astype(np.float32)
        oracle_tables_check(oracle, ctx, p, prev, x32)


@pytest.mark.parametrize("model,stress_fp64,bar", [(2, 1, 1e-6), (2, 0, 1e-3), (1, 1, 1e-6), (0, 1, 1e-6)])
def test_stress_parity(mpmx, oracle, model, stress_fp64, bar):
    """Constitutive update on identical FP32 inputs, for the von Mises (2),
    Neo-Hookean (1) and fixed corotated (0) energies.  FP64 arithmetic
    (default): only the FP32 rounding of the outputs remains, bound 1e-6 of
    each field's scale.  FP32 arithmetic: the SVD of near-equal singular values
    and the cancellation in y = |tau| - g R sqrt(2/3) cost up to ~1e-4; bound
    1e-3."""
    p, x, v, m, dt = mpmx.scene(0, seed=1, scale=0.5)
    p.model = model
    ctx = mpmx.Context(p, mpmx.default_config(stress_fp64=stress_fp64, svd_gap=0.0))
    ctx.set_particles(x, v, m)
    n = ctx.n
    rng = np.random.default_rng(7)
    F = np.eye(3)[None] + 0.04 * rng.standard_normal((n, 3, 3))
    F[: n // 2] = np.eye(3)[None] + 0.3 * rng.standard_normal((n // 2, 3, 3))  # plastic half
    F = np.where(np.linalg.det(F)[:, None, None] > 0.2, F, np.eye(3)[None])
    c = rng.uniform(0.3, 1.0, n)
    g = rng.uniform(0.1, 1.0, n)
    alpha = np.where(rng.random(n) < 0.5, 0.0, rng.uniform(0.0, 0.9, n))
    psi = np.where(rng.random(n) < 0.5, 0.0, 1e6 - 1.0)
    F32 = F.astype(np.float32).astype(np.float64)
    ins = {"F": F32, "c": c.astype(np.float32), "g": g.astype(np.float32),
           "alpha": alpha.astype(np.float32), "psi": psi.astype(np.float32)}
    for k, a in ins.items():
        ctx.set(k, a)
    ctx.stress()
    gpu = {k: ctx.get(k) for k in ("F", "sigma", "g", "alpha", "psi", "FP", "vol")}
    op = oracle_params(oracle, p)
    ref = {k: [] for k in gpu}
    for i in range(n):
        args = (float(ins["c"][i]), float(ins["g"][i]), float(ins["alpha"][i]), float(ins["psi"][i]))
        if model == 2:
            r = oracle.von_mises(op, ins["F"][i], *args)
        elif model == 1:
            r = oracle.neo_hookean(op, ins["F"][i], args[0], args[3])
            r.update(F=ins["F"][i], g=args[1], alpha=args[2])
        else:
            r = {"sigma": oracle.fixed_corotated(op, ins["F"][i]), "F": ins["F"][i], "g": args[1],
                 "alpha": args[2], "psi": args[3], "FP": 1.0, "vol": gpu["vol"][i]}
        for k in ref:
            ref[k].append(r[k])
    ref = {k: np.array(a) for k, a in ref.items()}
    errs = {k: max_rel_field(gpu[k], ref[k]) for k in ("F", "sigma", "FP", "vol", "g", "psi")}
    errs["alpha"] = np.abs(gpu["alpha"] - ref["alpha"]).max()
    print("stress parity (model %d, fp64=%d):" % (model, stress_fp64), {k: "%.1e" % e for k, e in errs.items()})
    for k, e in errs.items():
        assert e < bar, k


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_p2g_parity(mpmx, oracle, variant):
    """Raw P2G sums (before the grid update) of the scatter (1), the gather
    (2) and the owner-block (3) transfer on the same plastic, damaged state."""
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    ctx = mpmx.Context(p, mpmx.default_config(p2g_variant=variant))
    ctx.set_particles(x, v, m)
    # a few substeps so that B, sigma, c are non-trivial
    ctx.run(40, dt)
    ctx.bin_sort()
    ctx.enable_p2g_copy()
    ctx.rebuild()
    ctx.stress()
    ctx.p2g(dt)
    ctx.grid_update(dt)
    order, keys = ctx.order()
    tab = ctx.tables()
    f = {k: ctx.get(k)[order] for k in ("x", "v", "B", "sigma", "mass", "c", "vol", "FP")}
    ref = oracle.p2g_mls(oracle_params(oracle, p), keys, f["x"].astype(np.float32).astype(np.float64),
                         f["v"], f["B"], f["sigma"], f["mass"], f["c"], f["vol"], f["FP"], dt,
                         tab["page_key"])
    for ch in range(8):
        gpu = ctx.grid(ch, p2g=True)
        # FP32 sums of up to 8 x 27 contributions against FP64: 1e-5 of the channel scale
        assert max_rel_field(gpu, ref[ch]) < 1e-5, ch


STATE = ("x", "v", "F", "B", "c", "alpha", "psi", "g")


@pytest.mark.parametrize("transfer,variant,model", [(0, 2, 2), (1, 2, 2), (0, 1, 2), (1, 1, 2), (0, 2, 1), (1, 2, 0),
                                                     (0, 3, 2), (1, 3, 2)])
def test_one_step_from_gpu_state(mpmx, oracle, transfer, variant, model):
    """Shadow step: the GPU's own state 45 substeps into the impact (yielded,
    damaged particles) is loaded into the oracle and both advance one substep.
    The model amplifies rounding in yielded material over many substeps, so
    per-substep parity on identical inputs is the bar for every kernel."""
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    p.transfer = transfer
    p.model = model
    ctx = mpmx.Context(p, mpmx.default_config(p2g_variant=variant))
    ctx.set_particles(x, v, m)
    ctx.run(45, dt)
    st = {k: ctx.get(k) for k in STATE}
    mass = ctx.get("mass")
    if model == 2:
        assert (st["alpha"] > 0).sum() > 100 and st["c"].min() < 0.99
    sim = oracle.Sim(oracle_params(oracle, p), st["x"], st["v"], mass)
    for k in ("F", "B", "c", "alpha", "psi", "g"):
        sim.set(k, st[k])
    ctx.substep(dt)
    sim.substep(dt)
    errs = {k: max_rel_field(ctx.get(k), sim.get(k)) for k in ("x", "v", "F", "B", "psi")}
    for k in ("c", "alpha"):
        errs[k] = np.abs(ctx.get(k) - sim.get(k)).max()
    print("one step (transfer %d, variant %d, model %d):" % (transfer, variant, model),
          {k: "%.1e" % e for k, e in errs.items()})
    bars = {"x": 1e-6, "v": 1e-4, "F": 1e-5, "B": 1e-3, "psi": 1e-4, "c": 1e-4, "alpha": 1e-4}
    for k, e in errs.items():
        assert e < bars[k], k


def sensitivity(oracle, p, x, v, m, dt, steps, ref):
    """How far the oracle itself moves when its particle state is rounded to
    float32 after every substep (FP32 storage, FP64 arithmetic): the
    conditioning of the trajectory under the rounding an FP32 build must do."""
    sim = oracle.Sim(oracle_params(oracle, p), x.astype(np.float32), v, m.astype(np.float32))
    for _ in range(steps):
        sim.substep(dt)
        sim.round_f32()
    return {k: max_rel_field(sim.get(k), ref.get(k)) for k in ("x", "v", "F")}


def trajectory(mpmx, oracle, scene_id, scale, steps, **over):
    ctx, p, x, v, m, dt = make(mpmx, scene_id, scale, **over)
    sim = oracle.Sim(oracle_params(oracle, p), x.astype(np.float32), v, m.astype(np.float32))
    for _ in range(steps):
        ctx.substep(dt)
        sim.substep(dt)
    ctx.sync()
    return ctx, sim


def test_trajectory_cube(mpmx, oracle):
    """Free-falling cube, 20 substeps: strains stay ~0, so FP32 tracks FP64 closely."""
    ctx, sim = trajectory(mpmx, oracle, 0, 0.5, 20)
    assert max_rel_field(ctx.get("x"), sim.get("x")) < 1e-6
    assert max_rel_field(ctx.get("v"), sim.get("v")) < 1e-4
    assert max_rel_field(ctx.get("F"), sim.get("F")) < 1e-5
    assert np.abs(ctx.get("c") - sim.get("c")).max() < 1e-4


def trajectory_check(mpmx, oracle, steps, **over):
    """60 substeps of the impact: x within 1e-4 of its scale; v and F, which
    the model amplifies in yielded material, within 3x the distance between
    the oracle and the oracle with FP32 state storage (or 1e-4, whichever is
    larger)."""
    ctx, sim = trajectory(mpmx, oracle, 3, 0.1, steps, **over)
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    for k, val in over.items():
        setattr(p, k, val)
    env = sensitivity(oracle, p, x, v, m, dt, steps, sim)
    errs = {k: max_rel_field(ctx.get(k), sim.get(k)) for k in ("x", "v", "F")}
    dc = np.abs(ctx.get("c") - sim.get("c")).max()
    print("trajectory %s: gpu-oracle %s  oracle fp32-storage envelope %s  dc %.1e" % (
        over, {k: "%.1e" % e for k, e in errs.items()}, {k: "%.1e" % e for k, e in env.items()}, dc))
    assert ctx.stats()["cg_breakdown"] == 0
    assert errs["x"] < 1e-4
    for k in ("v", "F"):
        assert errs[k] < max(1e-4, 3 * env[k]), k
    assert dc < 1e-2


def test_trajectory_impact(mpmx, oracle):
    trajectory_check(mpmx, oracle, 60)


def test_trajectory_impact_apic(mpmx, oracle):
    """The reference's default transfer (P2G_APIC / G2P_APIC): stress as a grid
    force, F updated with sum v grad w^T, no velocity clamp."""
    trajectory_check(mpmx, oracle, 60, transfer=1)


def test_true_dot_mode(mpmx, oracle):
    """dot_mode 1 (true inner products in the PCG) on the impact scene."""
    trajectory_check(mpmx, oracle, 60, dot_mode=1)


def test_zero_weight_node_free_flight(mpmx):
    """A lone particle with two coordinates exactly on cell centres reaches
    nodes with weight 0; they stay inactive and the particle moves as in free
    flight, v = v0 + g dt per substep (closed form, FP32 rounding only)."""
    p = mpmx.default_params()
    p.N = 64
    p.V0 = (1.0 / 64) ** 3 / 8
    x = np.array([[20.5 / 64, 30.5 / 64, 40.25 / 64]])
    v0 = np.array([[0.2, -0.1, 0.05]])
    ctx = mpmx.Context(p)
    ctx.set_particles(x, v0, np.array([1e-3]))
    dt = 1e-4
    ctx.run(3, dt)
    v = ctx.get("v")
    assert np.isfinite(v).all() and np.isfinite(ctx.get("c")).all()
    assert np.abs(v - (v0 + np.array([0, -10.0, 0]) * 3 * dt)).max() < 1e-6
    assert ctx.get("c")[0] == pytest.approx(1.0, abs=1e-6)


def test_error_codes(mpmx):
    """Out-of-domain input is refused on the host; a grid capacity too small
    for the scene is reported as MPMX_ERR_CAPACITY, not a fault."""
    p, x, v, m, dt = mpmx.scene(0, seed=1, scale=0.5)
    ctx = mpmx.Context(p)
    bad = x.copy()
    bad[0, 0] = 1.5 / p.N
    with pytest.raises(mpmx.MpmxError) as e:
        ctx.set_particles(bad, v, m)
    assert e.value.status == mpmx.ERR_OUT_OF_DOMAIN
    ctx2 = mpmx.Context(p, mpmx.default_config(max_pages=8))
    ctx2.set_particles(x, v, m)
    with pytest.raises(mpmx.MpmxError) as e:
        ctx2.run(2, dt)
        ctx2.sync()
    assert e.value.status == mpmx.ERR_CAPACITY


def test_conservation_free_flight(mpmx):
    """No walls touched, gravity only: FP64 momentum changes by m g t exactly."""
    ctx, p, x, v, m, dt = make(mpmx, 0, 0.5)
    s0 = ctx.stats()
    ctx.run(10, dt)
    s1 = ctx.stats()
    assert abs(s1["mass"] - s0["mass"]) <= 1e-12 * s0["mass"]
    expect = s0["momentum"][1] + s0["mass"] * (-10.0) * 10 * dt
    assert abs(s1["momentum"][1] - expect) < 1e-5 * abs(expect)
    assert abs(s1["momentum"][0]) < 1e-6 * abs(expect)
    assert abs(s1["momentum"][2]) < 1e-6 * abs(expect)
