"""Benchmark: explicit MLS-MPM substeps with plastic phase-field fracture.

Headline workload (default, `--workload batch256_mixed`): 256 independent
scenes (cube, notched bar, sphere in turn; 5k-200k particles each, log-uniform,
12,983,932 in all; grids 64^3-256^3; E log-uniform in [1e4, 1e6]) advanced
together, one substep per step, with mixed-precision particle state (x and F
FP32; v, C, c, alpha, psi, g BF16; FP32 grid sums; FP64 conservation sums).
A "step" is one full substep of every scene: sort, page tables, von Mises
stress, P2G, grid update and boundaries, phase-field PCG, G2P.

Timing: `--warmup` substeps, then `--advance` more untimed substeps (into the
stretched regime of the bars), then the autotuner picks launch shapes on a
restored snapshot, then exactly `--steps` substeps are timed with HIP events
on the context stream (inputs resident in HBM; a device synchronisation on
both sides).  `value` = particles x substeps / seconds of that window;
`config.timed_substeps` names the window and `cg_iters_last` the PCG
iterations of its last substep.

`roofline`: the block P2G kernel (`k_p2g_block`, the largest kernel) timed
with HIP events around each launch in an instrumented pass of the same
workload.  achieved = algorithmic bytes / average launch time, with the
algorithmic bytes of an FP32 P2G, 104 B per particle (x, v, C, sigma, c and
the permutation read) + 56 B per active grid node (7 channels read and
written); G = 64 x pages.  `g2p` gives the same for the block G2P with
156 B per particle + 16 B per node.  `traffic` is the P2G kernel's HBM bytes
per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
this workload (null where none is committed).

`conservation`: FP64 sums over the particles before a substep against the
FP64 sums of the grid right after its P2G (mass and momentum are exactly
conserved by the transfer; the fixed-point sums truncate each term below
2^-30 of its block's largest), and the drift of the particles' FP64 mass and
momentum over the timed window.

`cpu_baseline`: the serial FP64 oracle on this host's cores (one thread) on a
bounded sample of the same workload: one substep of 16 seeded scenes (the
largest among them) from the GPU's own state.  The same substep on the GPU
gives `error_vs_oracle`, max |gpu - oracle| / max |oracle| per field (c and
alpha absolute) of the headline form.

`compare` (unless `--no-compare`): the plain FP32 single-scene path on the
notched bar (1,048,576 particles, 256^3), the two-block impact (8,192,000
particles, 512^3) in FP32 and mixed state, the 64-scene FP32 batch, the
256-scene batch in FP32, and the 256 scenes one at a time through the plain
FP32 single-scene path; `speedup` divides the headline by each.

`--copy-probe MIB` runs only a plain device copy (out = 1 * in over MIB MiB
of float32, three times): the calibration kernel of the counter runs.
"""
import argparse
import importlib.util
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "explicit-gpumpm-ductile-fracture_amd")

BASELINE_METRIC = "particle-substeps/s & %HBM roofline, plain FP32 vs mixed/batched at ≤1e-4 err"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "batch256_mixed": ("batch", 5, "mixed",
                       "256 scenes (cube / notched bar / sphere, 5k-200k particles each, grids 64^3-256^3, "
                       "E in [1e4, 1e6]), mixed state (x, F FP32; v, C, c, alpha, psi, g BF16), FP32 grid, "
                       "dt=1e-4 (the batch's CFL-rule step)"),
    "batch256_fp32": ("batch", 5, "fp32", "the same 256 scenes, FP32 state"),
    "batch64_fp32": ("batch", 4, "fp32", "64 scenes (10k-500k particles each), FP32 state"),
    "bar1m": ("scene", 1, "fp32", "notched bar under tension: 1,048,576 particles, 256^3 sparse grid, grips "
                                  "+-0.5 m/s, critical stretch 1.05, phase-field damage, FP32, dt=1e-4"),
    "impact8m": ("scene", 3, "fp32", "two-block impact: 8,192,000 particles, 512^3 sparse grid, +-2.5 m/s, FP32"),
    "impact8m_mixed": ("scene", 3, "mixed", "two-block impact: 8,192,000 particles, 512^3, mixed state"),
    "cube64k": ("scene", 0, "fp32", "cube: 64,000 particles (8 ppc), 128^3 sparse grid, FP32, dt=1e-4"),
}


def load_binding():
    spec = importlib.util.spec_from_file_location("mpmx_binding", os.path.join(PKG, "mpmx.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mpmx_binding"] = mod
    spec.loader.exec_module(mod)
    return mod


def host_cpu():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def make_context(mpmx, name):
    kind, ident, precision, _ = WORKLOADS[name]
    cfg = mpmx.default_config(precision=mpmx.MIXED if precision == "mixed" else mpmx.FP32)
    if kind == "batch":
        ps, counts, x, v, m, dt = mpmx.scene_batch(ident, seed=ident)
        ctx = mpmx.Context(ps, cfg)
        ctx.set_particles_batch(counts, x, v, m)
        return ctx, ps, counts, dt
    p, x, v, m, dt = mpmx.scene(ident, seed=1, scale=1.0)
    ctx = mpmx.Context(p, cfg)
    ctx.set_particles(x, v, m)
    return ctx, [p], np.array([len(m)], np.int64), dt


def timed(mpmx, name, steps, warmup, advance, autotune=True):
    """Warm up, pre-advance, tune, then time exactly `steps` substeps."""
    ctx, ps, counts, dt = make_context(mpmx, name)
    ctx.run(warmup + advance, dt)
    ctx.sync()
    tuning = ctx.autotune(dt, 3) if autotune else ctx.tuning()
    s0 = ctx.stats()
    ms = ctx.time_run(steps, dt)
    s1 = ctx.stats()
    n = int(counts.sum())
    rec = {"workload": WORKLOADS[name][3], "particles": n, "scenes": len(ps),
           "value": n * steps / (ms / 1e3), "unit": "particle-substeps/s", "ms_per_step": ms / steps,
           "steps": steps, "timed_substeps": [warmup + advance + 1, warmup + advance + steps],
           "cg_iters_last": s1["cg_iters"], "cg_breakdown": s1["cg_breakdown"], "tuning": tuning,
           "pages_total": s1["numTotalPage"]}
    return ctx, ps, counts, dt, rec, s0, s1


def kernel_rates(ctx, n, profile_steps, dt):
    """Per-kernel event timing of `profile_steps` substeps; P2G and G2P rates."""
    ctx.timing_begin()
    ctx.run(profile_steps, dt)
    kt = ctx.timing_end()
    st = ctx.stats()
    G = 64 * st["numTotalPage"]
    out = {"stage_ms_per_step": {k: v[1] / max(profile_steps, 1) for k, v in kt.items()}}
    for key, kname, per_p, per_g in (("p2g", "p2g_block", 104.0, 56.0), ("g2p", "g2p_block", 156.0, 16.0)):
        launches, total = kt.get(kname, (0, 0.0))
        avg = total / launches if launches else None
        bytes_ = per_p * n + per_g * G
        out[key] = {"kernel": kname, "avg_launch_ms": avg, "algorithmic_bytes_per_launch": bytes_,
                    "achieved_gbs": bytes_ / (avg * 1e-3) / 1e9 if avg else None, "active_nodes": G}
    return out


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_p2g_block_%s.md")


def measured_traffic(workload):
    """HBM bytes per launch of k_p2g_block from the committed counter summary
    of this workload (line `traffic_bytes_per_launch: B`), or None."""
    path = TRAFFIC_SUMMARY % workload
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def conservation(ctx, dt, s0, s1):
    """P2G transfer conservation of one substep, FP64, and the drift of the
    particle sums over the timed window."""
    ctx.enable_p2g_copy()
    ctx.run(1, dt)                 # the copy starts with this substep's grid update
    before = ctx.stats()
    ctx.run(1, dt)
    grid = [ctx.grid(ch, p2g=True).astype(np.float64).sum() for ch in range(4)]
    pm = np.array(before["momentum"])
    scale = max(np.abs(pm).max(), 1e-30)
    return {"p2g_mass_rel_err": abs(grid[0] - before["mass"]) / before["mass"],
            "p2g_momentum_rel_err": float(np.abs(np.array(grid[1:]) - pm).max() / scale),
            "timed_window_mass_drift": (s1["mass"] - s0["mass"]) / s0["mass"],
            "timed_window_momentum_change": [b - a for a, b in zip(s0["momentum"], s1["momentum"])]}


def cpu_baseline_and_error(mpmx, ctx, ps, counts, dt, sample=16):
    """The serial FP64 oracle, one thread, on one substep of a sample of the
    workload's scenes from the GPU's state; the GPU's own substep against it."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import oracle

    fields = ("x", "v", "F", "B", "c", "alpha", "psi", "g")
    st = {k: ctx.get(k) for k in fields}
    mass = ctx.get("mass")
    ctx.run(1, dt)
    after = {k: ctx.get(k) for k in fields}
    start = np.concatenate([[0], np.cumsum(counts)])
    if len(ps) > sample:
        rng = np.random.default_rng(5)
        pick = sorted(set(rng.choice(len(ps), sample - 1, replace=False).tolist()) | {int(np.argmax(counts))})
    else:
        pick = list(range(len(ps)))
    worst, sec, work = {}, 0.0, 0
    for s in pick:
        sl = slice(int(start[s]), int(start[s + 1]))
        op = oracle.Params()
        for name, _ in oracle.Params._fields_:
            if name == "gravity":
                op.gravity[:] = list(ps[s].gravity)
            else:
                setattr(op, name, getattr(ps[s], name))
        sim = oracle.Sim(op, st["x"][sl], st["v"][sl], mass[sl])
        for k in ("F", "B", "c", "alpha", "psi", "g"):
            sim.set(k, st[k][sl])
        t0 = time.perf_counter()
        sim.substep(dt)
        sec += time.perf_counter() - t0
        work += int(counts[s])
        # fields that vanish in a scene are measured against a floor of their
        # natural scale: B against max|v| dx, psi against mu 1e-8 (a 1e-4 strain)
        vref = sim.get("v")
        mu = ps[s].E / (2 * (1 + ps[s].nu))
        floor = {"B": np.abs(vref).max() / ps[s].N, "psi": 1e-8 * mu}
        for k in ("x", "v", "F", "B", "psi"):
            ref = vref if k == "v" else sim.get(k)
            e = float(np.abs(after[k][sl] - ref).max() / max(np.abs(ref).max(), floor.get(k, 0.0), 1e-30))
            worst[k] = max(worst.get(k, 0.0), e)
        for k in ("c", "alpha"):
            worst[k] = max(worst.get(k, 0.0), float(np.abs(after[k][sl] - sim.get(k)).max()))
    base = {"value": work / sec, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": "one substep of %d of the %d scenes (%d particles, the largest scene among them) from the "
                      "GPU's state on the serial FP64 oracle, 1 thread" % (len(pick), len(ps), work),
            "host": host_cpu()}
    return base, {"fields": worst, "scenes": len(pick), "particles": work,
                  "definition": "one substep from the GPU's state: max|gpu - oracle| / max|oracle| per field "
                                "(c, alpha absolute; B at least against max|v| dx, psi against mu 1e-8)"}


def one_at_a_time(mpmx, config, steps, warmup):
    """The batch's scenes through the plain FP32 single-scene path, one
    context after another: particle-substeps/s over the summed GPU time."""
    ps, counts, x, v, m, dt = mpmx.scene_batch(config, seed=config)
    start = np.concatenate([[0], np.cumsum(counts)])
    total_ms, work = 0.0, 0
    for s, p in enumerate(ps):
        a, b = start[s], start[s + 1]
        ctx = mpmx.Context(p)
        ctx.set_particles(x[a:b], v[a:b], m[a:b])
        ctx.run(warmup, dt)
        ctx.sync()
        total_ms += ctx.time_run(steps, dt)
        work += int(b - a) * steps
        ctx.close()
    return {"workload": "the 256 scenes one after another, plain FP32 single-scene path", "value":
            work / (total_ms / 1e3), "unit": "particle-substeps/s", "scenes": len(ps), "steps_per_scene": steps}


def copy_probe(mib):
    """Plain device copy of `mib` MiB, three launches, for counter calibration."""
    import torch

    n = (mib << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_kib_read": n * 4 // 1024, "copy_probe_kib_written": n * 4 // 1024,
                      "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--advance", type=int, default=200, help="untimed substeps after the warm-up")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="batch256_mixed")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU box here has one "
              "GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    mpmx = load_binding()
    name = args.workload
    t0 = time.perf_counter()
    ctx, ps, counts, dt, rec, s0, s1 = timed(mpmx, name, args.steps, args.warmup, args.advance,
                                              not args.no_autotune)
    wall = time.perf_counter() - t0
    n = rec["particles"]
    rates = kernel_rates(ctx, n, args.profile_steps, dt)
    p2g = rates["p2g"]
    achieved = p2g["achieved_gbs"] or 0.0
    mixed = WORKLOADS[name][2] == "mixed"
    out = {
        "metric": BASELINE_METRIC,
        "value": rec["value"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 storage (x, F f32), f32 compute" if mixed else "f32",
        "stress_arith": "f64",
        "data": "synthetic (seeded SplitMix64 scene generator)",
        "config": {"workload": rec["workload"], "name": name, "particles": n, "scenes": rec["scenes"],
                   "grids": "/".join("%d^3" % g for g in sorted({q.N for q in ps})),
                   "advance": args.advance, "timed_substeps": rec["timed_substeps"],
                   "cg_iters_last": rec["cg_iters_last"], "cg_breakdown": rec["cg_breakdown"],
                   "pages_total": rec["pages_total"], "active_nodes": p2g["active_nodes"],
                   "wall_s_setup_and_timing": wall, "tuning": rec["tuning"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(name),
                     "kernel": p2g["kernel"], "algorithmic_bytes_per_launch": p2g["algorithmic_bytes_per_launch"],
                     "avg_launch_ms": p2g["avg_launch_ms"],
                     "g2p": {"kernel": "g2p_block", "achieved": rates["g2p"]["achieved_gbs"],
                             "frac": (rates["g2p"]["achieved_gbs"] or 0.0) / HBM_PEAK_GBS,
                             "avg_launch_ms": rates["g2p"]["avg_launch_ms"],
                             "algorithmic_bytes_per_launch": rates["g2p"]["algorithmic_bytes_per_launch"]}},
        "stage_ms_per_step": rates["stage_ms_per_step"],
        "conservation": conservation(ctx, dt, s0, s1),
    }
    if not args.no_cpu:
        out["cpu_baseline"], out["error_vs_oracle"] = cpu_baseline_and_error(mpmx, ctx, ps, counts, dt)
    ctx.close()
    if not args.no_compare:
        cmp = {}
        steps = min(args.steps, 50)
        for other, adv in (("bar1m", 500), ("impact8m", 20), ("impact8m_mixed", 20), ("batch64_fp32", 50),
                           ("batch256_fp32", args.advance)):
            if other == name:
                continue
            c2, _, _, _, r, _, _ = timed(mpmx, other, steps, 5, adv)
            c2.close()
            cmp[other] = r
        cmp["batch256_one_at_a_time_fp32"] = one_at_a_time(mpmx, 5, 3, 2)
        out["compare"] = cmp
        out["speedup"] = {k: rec["value"] / v["value"] for k, v in cmp.items()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
