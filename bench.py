"""Benchmark: explicit MLS-MPM substeps with plastic phase-field fracture.

Workload (default): the notched bar under tension — 1,048,576 particles in a
256^3 sparse grid, grips at +-0.5 m/s, psi+ gated at stretch 1.05, FP32,
dt = 1e-4.  A "step" is one full substep (sort, page tables, von Mises stress,
P2G, grid update + boundaries, phase-field PCG, G2P) over the resident scene.

Prints one JSON line.  `value` = particles x substeps / seconds of the timed
region (inputs resident in HBM; HIP events on the context stream bracket it,
with a device synchronisation on both sides).  `roofline` comes from a second,
instrumented pass of the same scene in which every stage is bracketed by HIP
events; achieved = algorithmic bytes of the block P2G kernel (its payload
read once, 80 B per particle for MLS and 112 B for APIC, plus 52 B written per
grid node) / its average launch time.  `cpu_baseline` runs the
serial FP64 oracle on the host for a bounded sample of the same scene.
`roofline.traffic` is the P2G kernel's HBM traffic per launch as measured by
rocprofv3 counter runs (FETCH_SIZE + WRITE_SIZE, each scaled by what the
counter reads for a plain copy of known size), read from the committed counter
summary of this workload; null for other workloads.

`batched` reports the batch forms on the same GPU: the 256-scene batch
(5k-200k particles per scene, ~13M in all) with mixed-precision state and the
64-scene batch (10k-500k, ~9.8M) in FP32, each advanced as one context, and
the 256 scenes run one at a time through the plain FP32 single-scene path
(the speed-up of the batched mixed form over that).  `--no-batch` skips it.

`impact_8M` reports the two-block impact (8,192,000 particles, 512^3) in
FP32 and with mixed-precision state, each after the autotuner.

`--copy-probe MIB` runs only a plain device copy (out = 1 * in over MIB MiB of
float32, three times): the calibration kernel of those counter runs.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "explicit-gpumpm-ductile-fracture_amd")

BASELINE_METRIC = "particle-substeps/s & %HBM roofline, plain FP32 vs mixed/batched at ≤1e-4 err"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    1: "notched bar under tension: 1,048,576 particles, 256^3 sparse grid, grips +-0.5 m/s, "
       "critical stretch 1.05, phase-field damage, FP32, dt=1e-4",
    0: "cube: 64,000 particles (8 ppc), 128^3 sparse grid, FP32, dt=1e-4",
    2: "notched bar 8 ppc: 1,048,576 particles, 256^3 sparse grid, FP32, dt=1e-4",
    3: "two-block impact: 8,192,000 particles, 512^3 sparse grid, +-2.5 m/s, FP32, dt=1e-4",
}


def load_binding():
    spec = importlib.util.spec_from_file_location("mpmx_binding", os.path.join(PKG, "mpmx.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mpmx_binding"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(scene_id, steps):
    """Serial FP64 oracle on this host for `steps` substeps of the same scene."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import oracle

    mpmx = load_binding()
    p, x, v, m, dt = mpmx.scene(scene_id, seed=1, scale=1.0)
    op = oracle.Params()
    for name, _ in oracle.Params._fields_:
        if name == "gravity":
            op.gravity[:] = list(p.gravity)
        else:
            setattr(op, name, getattr(p, name))
    sim = oracle.Sim(op, x, v, m)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.substep(dt)
    sec = time.perf_counter() - t0
    return {"value": len(m) * steps / sec, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": "%d substeps of the same scene (%d particles) on the serial FP64 oracle, 1 thread"
                      % (steps, len(m))}


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_%s_notched_bar_1M.md")


def measured_traffic(kname, scene, precision):
    """HBM bytes per launch of kernel `kname` from its committed counter summary
    (line `traffic_bytes_per_launch: B`), for the default workload only."""
    path = TRAFFIC_SUMMARY % kname
    if scene != 1 or precision != "fp32" or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


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


def batch_run(mpmx, config, precision, steps, warmup):
    """One batch config advanced as one context: particle-substeps/s."""
    ps, counts, x, v, m, dt = mpmx.scene_batch(config, seed=config)
    cfg = mpmx.default_config(precision=mpmx.MIXED if precision == "mixed" else mpmx.FP32)
    ctx = mpmx.Context(ps, cfg)
    ctx.set_particles_batch(counts, x, v, m)
    ctx.run(warmup, dt)
    ctx.sync()
    tuning = ctx.autotune(dt, 3)
    ms = ctx.time_run(steps, dt)
    st = ctx.stats()
    n = int(counts.sum())
    ctx.close()
    return {"config": "%d scenes, %d particles (%d-%d per scene), grids %s, %s" % (
                len(ps), n, counts.min(), counts.max(), "/".join("%d^3" % g for g in sorted({q.N for q in ps})),
                precision),
            "value": n * steps / (ms / 1e3), "unit": "particle-substeps/s", "ms_per_step": ms / steps,
            "steps": steps, "warmup": warmup, "cg_iters_last": st["cg_iters"], "cg_breakdown": st["cg_breakdown"],
            "tuning": tuning}


def scene_run(mpmx, scene_id, precision, steps, warmup):
    """One scene through the single-scene path: particle-substeps/s."""
    p, x, v, m, dt = mpmx.scene(scene_id, seed=1, scale=1.0)
    cfg = mpmx.default_config(precision=mpmx.MIXED if precision == "mixed" else mpmx.FP32)
    ctx = mpmx.Context(p, cfg)
    ctx.set_particles(x, v, m)
    ctx.run(warmup, dt)
    ctx.sync()
    ctx.autotune(dt, 3)
    ms = ctx.time_run(steps, dt)
    st = ctx.stats()
    n = ctx.n
    ctx.close()
    return {"value": n * steps / (ms / 1e3), "unit": "particle-substeps/s", "ms_per_step": ms / steps,
            "particles": n, "precision": precision, "steps": steps, "cg_iters_last": st["cg_iters"]}


def scenes_one_at_a_time(mpmx, config, steps, warmup):
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
    return {"value": work / (total_ms / 1e3), "unit": "particle-substeps/s", "scenes": len(ps),
            "steps_per_scene": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--scene", type=int, default=1)
    ap.add_argument("--precision", choices=["fp32", "mixed"], default="fp32")
    ap.add_argument("--vpage", type=int, default=0)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=30)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU box here has one "
              "GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    mpmx = load_binding()
    p, x, v, m, dt = mpmx.scene(args.scene, seed=1, scale=1.0)
    if args.vpage:
        p.vpage = args.vpage
    cfg = mpmx.default_config(precision=mpmx.MIXED if args.precision == "mixed" else mpmx.FP32)
    ctx = mpmx.Context(p, cfg)
    ctx.set_particles(x, v, m)
    n = ctx.n

    ctx.run(args.warmup, dt)
    ctx.sync()
    # launch shapes from the autotuner (trials on a restored snapshot; the
    # shapes change no result)
    tuning = ctx.tuning() if args.no_autotune else ctx.autotune(dt, 5)
    t0 = time.perf_counter()
    ms = ctx.time_run(args.steps, dt)   # events on the context stream, synchronised on both sides
    wall = time.perf_counter() - t0
    st = ctx.stats()
    value = n * args.steps / (ms / 1e3)

    # instrumented pass: per-stage event timing of the same scene
    ctx.timing_begin()
    ctx.run(args.profile_steps, dt)
    kt = ctx.timing_end()
    st2 = ctx.stats()
    # dominant kernel: the block P2G (8-channel sums, grid update and PCG
    # set-up).  Algorithmic bytes: its per-particle payload read once (MLS 80 B,
    # APIC 112 B) and 32 B of grid channels + 20 B of PCG vectors written per
    # node.  The whole P2G stage (payload kernel + block kernel) is also
    # reported against the 104 B per particle + 56 B per node of a P2G that
    # reads x, v, C, sigma, c and the permutation directly.
    kname = "p2g_block" if "p2g_block" in kt else "p2g"
    p2g_launches, p2g_ms = kt.get(kname, (0, 0.0))
    G = 64 * st2["numTotalPage"]
    pay = 112.0 if p.transfer == 1 else 80.0
    p2g_bytes = (pay if kname == "p2g_block" else 104.0) * n + 52.0 * G
    achieved = p2g_bytes / (p2g_ms / p2g_launches * 1e-3) / 1e9 if p2g_launches else 0.0
    prep_launches, prep_ms = kt.get("p2g_prep", (0, 0.0))
    stage_ms = (p2g_ms + prep_ms) / p2g_launches if p2g_launches else 0.0
    p2g_stage_gbs = (104.0 * n + 56.0 * G) / (stage_ms * 1e-3) / 1e9 if stage_ms else 0.0
    # G2P: reads x, F, c (52 B) and writes x, v, B, F, c, key (104 B) per
    # particle; reads v, c, sum w and x of the PCG (24 B) per node
    gname = "g2p_block" if "g2p_block" in kt else "g2p"
    g2p_launches, g2p_ms = kt.get(gname, (0, 0.0))
    g2p_bytes = 156.0 * n + 24.0 * G
    g2p_gbs = g2p_bytes / (g2p_ms / g2p_launches * 1e-3) / 1e9 if g2p_launches else 0.0

    out = {
        "metric": BASELINE_METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "bf16/fp16 storage, f32 compute",
        "stress_arith": "f64",
        "data": "synthetic (seeded SplitMix64 scene generator)",
        "config": {"workload": WORKLOADS.get(args.scene, "scene %d" % args.scene), "particles": n,
                   "grid": "%d^3" % p.N, "vpage": p.vpage, "precision": args.precision,
                   "pages_total": st["numTotalPage"], "active_nodes": G,
                   "cg_iters_last": st["cg_iters"], "wall_s": wall, "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kname, args.scene, args.precision), "kernel": kname,
                     "algorithmic_bytes_per_launch": p2g_bytes,
                     "avg_launch_ms": p2g_ms / p2g_launches if p2g_launches else None,
                     "p2g_stage_gbs": p2g_stage_gbs, "g2p_gbs": g2p_gbs},
        "stage_ms_per_step": {k: v[1] / max(args.profile_steps, 1) for k, v in kt.items()},
    }
    if not args.no_batch:
        ctx.close()
        c5 = batch_run(mpmx, 5, "mixed", args.batch_steps, 5)
        c4 = batch_run(mpmx, 4, "fp32", args.batch_steps, 5)
        alone = scenes_one_at_a_time(mpmx, 5, 3, 2)
        out["batched"] = {"mixed_256": c5, "fp32_64": c4, "plain_fp32_one_scene_at_a_time_256": alone,
                          "speedup_mixed_batch_vs_one_at_a_time": c5["value"] / alone["value"]}
        out["impact_8M"] = {"workload": WORKLOADS[3], "fp32": scene_run(mpmx, 3, "fp32", 30, 5),
                            "mixed": scene_run(mpmx, 3, "mixed", 30, 5)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.scene, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
