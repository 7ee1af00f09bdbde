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
