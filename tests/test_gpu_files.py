"""Particle files and state snapshots on the GPU path: checkpoint / resume
continues bit-exactly (one scene and a batch, FP32 and mixed state), the PLY
particle dump holds the context's own state, and the launch-shape autotuner
leaves the state where it was and changes no result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "B", "c", "alpha", "psi", "g")


def state(ctx):
    return {k: ctx.get(k) for k in FIELDS}


@pytest.mark.parametrize("kind", ["single_fp32", "single_mixed", "single_mixed_f16", "batch_fp32"])
def test_checkpoint_resume_is_exact(mpmx, tmp_path, kind):
    """Run 6 substeps, checkpoint, run 6 more; a context loaded from the
    checkpoint runs the same 6 and ends bit-identical."""
    if kind.startswith("single"):
        p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
        prec = {"fp32": mpmx.FP32, "mixed": mpmx.MIXED, "f16": mpmx.MIXED_F16}[kind.rsplit("_", 1)[1]]
        ctx = mpmx.Context(p, mpmx.default_config(precision=prec))
        ctx.set_particles(x, v, m)
    else:
        ps, counts, x, v, m, dt = mpmx.scene_batch(4, seed=4, scale=0.004, num_scenes=5)
        ctx = mpmx.Context(ps)
        ctx.set_particles_batch(counts, x, v, m)
    ctx.run(6, dt)
    path = str(tmp_path / "run.ckpt")
    ctx.save_checkpoint(path)
    ctx.run(6, dt)
    want = state(ctx)
    resumed = mpmx.Context.load_checkpoint(path)
    assert resumed.n == ctx.n and resumed.num_scenes == ctx.num_scenes
    assert resumed.stats()["substeps"] == 6
    resumed.run(6, dt)
    got = state(resumed)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k


def test_particle_dump(mpmx, tmp_path):
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    ctx = mpmx.Context(p)
    ctx.set_particles(x, v, m)
    ctx.run(3, dt)
    path = str(tmp_path / "particles.ply")
    ctx.write_particles(path)
    raw = open(path, "rb").read()
    head, body = raw.split(b"end_header\n", 1)
    assert b"element vertex %d" % ctx.n in head
    rec = np.frombuffer(body, dtype=np.dtype([("x", "<f4", 3), ("v", "<f4", 3), ("c", "<f4"), ("s", "<i4")]))
    assert len(rec) == ctx.n
    assert np.array_equal(rec["x"], ctx.get("x").astype(np.float32))
    assert np.array_equal(rec["v"], ctx.get("v").astype(np.float32))
    assert np.array_equal(rec["c"], ctx.get("c").astype(np.float32))
    assert (rec["s"] == 0).all()


def test_autotune_changes_no_result(mpmx):
    """The autotuner's trials run on a snapshot it restores; every launch
    shape gives the same bits, so a tuned run equals an untuned one."""
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    a = mpmx.Context(p)
    a.set_particles(x, v, m)
    a.run(3, dt)
    before = state(a)
    best = a.autotune(dt, 2)
    assert best["particle_wg"] in (64, 128, 256)
    after = state(a)
    for k in FIELDS:
        assert np.array_equal(before[k], after[k]), k
    assert a.stats()["substeps"] == 3
    a.run(6, dt)
    b = mpmx.Context(p)
    b.set_particles(x, v, m)
    b.set_tuning(particle_wg=64, p2g_pct=200, g2p_pct=50, cg_pct=50)
    b.run(9, dt)
    got, want = state(a), state(b)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k


def test_checkpoint_rejects_inconsistent_files(mpmx, tmp_path):
    """A truncated checkpoint or one whose scene counts do not add up to its
    particle count is refused before anything is allocated from it."""
    import struct

    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=0.1)
    ctx = mpmx.Context(p)
    ctx.set_particles(x, v, m)
    path = tmp_path / "run.ckpt"
    ctx.save_checkpoint(str(path))
    raw = path.read_bytes()
    (tmp_path / "short.ckpt").write_bytes(raw[:-100])
    # header: magic[8], int32 version, num_scenes, precision, sorted, int64 n, ...
    bad_n = raw[:24] + struct.pack("<q", 1 << 29) + raw[32:]
    (tmp_path / "badn.ckpt").write_bytes(bad_n)
    for name in ("short.ckpt", "badn.ckpt"):
        with pytest.raises(mpmx.MpmxError) as e:
            mpmx.Context.load_checkpoint(str(tmp_path / name))
        assert e.value.status == mpmx.ERR_ARG
    assert mpmx.Context.load_checkpoint(str(path)).n == ctx.n
