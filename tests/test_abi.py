"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/mpmx.h declares, fills defaults, generates deterministic scenes,
and refuses to create a context without a device."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mpmx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpmx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(mpmx):
    L = mpmx.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(mpmx.EXPORTS)


def test_defaults_are_reference_constants(mpmx):
    p = mpmx.default_params()
    assert (p.N, p.cg_max_iter, p.vpage, p.wall_cells) == (128, 100, 256, 4)
    assert (p.yield_stress, p.hardening, p.alpha_c, p.dlambda_cap, p.eps_cap) == (800, 3, 1, 100, 3)
    assert (p.Gc0, p.residual_c, p.ps, p.M) == (3, 0.1, 0.01, 50)
    assert p.theta == float(np.float32(1e-4)) and p.vmax == float(np.float32(11.137))
    assert list(p.gravity) == [0, -10, 0]
    c = mpmx.default_config()
    assert c.precision == mpmx.FP32 and c.svd_gap == 1e-4


def test_params_layout_matches_oracle(mpmx, oracle):
    assert [f[0] for f in mpmx.Params._fields_] == [f[0] for f in oracle.Params._fields_]
    assert ctypes.sizeof(mpmx.Params) == ctypes.sizeof(oracle.Params)


def test_scenes_are_seeded_and_in_domain(mpmx):
    for sid, scale, n_expect in ((0, 1.0, 64000), (1, 0.05, 52429), (3, 0.1, 8192)):
        p, x, v, m, dt = mpmx.scene(sid, seed=3, scale=scale)
        assert len(m) == n_expect
        p2, x2, v2, m2, dt2 = mpmx.scene(sid, seed=3, scale=scale)
        assert np.array_equal(x, x2) and np.array_equal(m, m2)
        lo, hi = 2.0 / p.N, 1 - 2.0 / p.N
        assert x.min() >= lo and x.max() <= hi
        assert dt == 1e-4 and p.V0 > 0 and m.min() > 0
    p, x, v, m, dt = mpmx.scene(1, seed=1, scale=1.0)
    assert len(m) == 1048576 and p.N == 256 and p.grip_v == 0.5 and p.crit_stretch == 1.05
    p, x, v, m, dt = mpmx.scene(3, seed=1, scale=1.0)
    assert len(m) == 8192000 and p.N == 512 and set(np.unique(v[:, 2])) == {-2.5, 2.5}


def test_no_device_is_reported(mpmx):
    if mpmx.device_count() > 0:
        return
    try:
        mpmx.Context(mpmx.default_params())
    except mpmx.MpmxError as e:
        assert e.status == mpmx.ERR_NO_DEVICE
    else:
        raise AssertionError("context created without a device")


def test_batch_scenes_are_seeded_and_sized(mpmx):
    """Batch configs: 64 scenes of 10k-500k particles (config 4) and 256 of
    5k-200k (config 5), log-uniform; grids 64^3-256^3; every scene inside its
    own unit cube; deterministic per seed."""
    for cfg, S, lo_n, hi_n in ((4, 64, 1e4, 5e5), (5, 256, 5e3, 2e5)):
        ps, counts, x, v, m, dt = mpmx.scene_batch(cfg, seed=cfg)
        assert len(ps) == S and len(counts) == S and counts.sum() == len(m)
        assert counts.min() >= 0.99 * lo_n and counts.max() <= 1.01 * hi_n
        assert {p.N for p in ps} <= {64, 128, 256} and len({p.N for p in ps}) == 3
        Es = np.array([p.E for p in ps])
        assert Es.min() >= 1e4 and Es.max() <= 1e6
        start = np.concatenate([[0], np.cumsum(counts)])
        for s, p in enumerate(ps):
            xs = x[start[s]:start[s + 1]]
            assert xs.min() >= 2.0 / p.N and xs.max() <= 1 - 2.0 / p.N
            assert p.V0 > 0 and m[start[s]] > 0
        ps2, counts2, x2, *_ = mpmx.scene_batch(cfg, seed=cfg)
        assert np.array_equal(counts, counts2) and np.array_equal(x, x2)
    ps, counts, *_ = mpmx.scene_batch(4, seed=4, scale=0.01, num_scenes=5)
    assert len(ps) == 5 and counts.sum() < 30000


def test_batch_create_validates_arguments(mpmx):
    """Batch-wide settings must agree, and the keys must fit 32 bits; these
    are refused before any device is touched."""
    L = mpmx.lib()
    import ctypes as C

    def create(params, cfg=None):
        arr = (mpmx.Params * len(params))(*params)
        h = C.c_void_p()
        c = cfg if cfg is not None else mpmx.default_config()
        return L.mpmx_create_batch(len(params), arr, C.byref(c), C.byref(h))

    a = mpmx.default_params(N=64)
    b = mpmx.default_params(N=64, transfer=1)
    assert create([a, b]) == mpmx.ERR_ARG
    big = [mpmx.default_params(N=512) for _ in range(64)]   # 6 + 21 + 6 > 32
    assert create(big) == mpmx.ERR_ARG
    assert create([a, a], mpmx.default_config(p2g_variant=1)) == mpmx.ERR_ARG
    for bad in (-1, 4):
        assert create([a], mpmx.default_config(p2g_variant=bad)) == mpmx.ERR_ARG
    for bad in (-1, 3):
        assert create([a], mpmx.default_config(sort_variant=bad)) == mpmx.ERR_ARG
    if mpmx.device_count() == 0:
        assert create([a, a]) == mpmx.ERR_NO_DEVICE
