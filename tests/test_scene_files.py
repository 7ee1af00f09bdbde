"""Scene files without a GPU: the point-list OBJ loader (Sample::readObj,
LoadSample.cpp:180-194) and the level-set sampler (Sample::LoadSDF +
GenerateUniformSamples + Initial_data, LoadSample.cpp:3-160), pinned by
brute force on small synthetic files: the sample count from an independent
count of the level set's inside cells, every sample inside by an
independent trilinear evaluation, and the placement inverted in closed form."""
import os

import numpy as np
import pytest

REF_OBJ = "/root/reference/Assets/sample/tube_10.obj"


def test_obj_points(mpmx, tmp_path):
    p = tmp_path / "pts.obj"
    p.write_text("# points\nv 0.5 0.6 0.7\nvn 0 1 0\nv 0.1 0.2 0.3\nf 1 2 3\nv\t0.25 0.5 0.75\n")
    x = mpmx.read_obj_points(str(p))
    assert np.allclose(x, [[0.5, 0.4, 0.7], [0.1, 0.0, 0.3], [0.25, 0.3, 0.75]])
    assert np.allclose(mpmx.read_obj_points(str(p), 0.0)[:, 1], [0.6, 0.2, 0.5])
    with pytest.raises(OSError):
        mpmx.read_obj_points(str(tmp_path / "missing.obj"))


@pytest.mark.skipif(not os.path.exists(REF_OBJ), reason="reference asset not present")
def test_reference_obj_asset(mpmx):
    """The reference's own point list: 100,330 particles, first one as the
    file's first line moved down by 0.2."""
    x = mpmx.read_obj_points(REF_OBJ)
    assert x.shape == (100330, 3)
    first = [float(t) for t in open(REF_OBJ).readline().split()[1:]]
    assert np.allclose(x[0], [first[0], first[1] - 0.2, first[2]])


def sphere_sdf(path, n=20, r=6.0):
    c = (n - 1) / 2.0
    g = np.arange(n, dtype=np.float64)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    phi = np.sqrt((I - c) ** 2 + (J - c) ** 2 + (K - c) ** 2) - r
    with open(path, "w") as f:
        f.write("%d %d %d\n0 0 0\n0.1\n" % (n, n, n))
        # i fastest
        f.write("\n".join("%.17g" % v for v in phi.transpose(2, 1, 0).reshape(-1)))
    return phi


def trilinear(phi, p):
    x0 = np.floor(p).astype(int)
    x1 = np.ceil(p).astype(int)
    f = p - x0
    out = np.zeros(len(p))
    for q in range(8):
        idx = [x1[:, a] if (q >> (2 - a)) & 1 else x0[:, a] for a in range(3)]
        w = np.ones(len(p))
        for a in range(3):
            w *= f[:, a] if (q >> (2 - a)) & 1 else 1 - f[:, a]
        out += w * phi[idx[0], idx[1], idx[2]]
    return out


def test_sdf_sampling(mpmx, tmp_path):
    path = str(tmp_path / "ball.sdf")
    n = 20
    phi = sphere_sdf(path, n)
    N, center, res, spc = 128, (64, 64, 64), (48, 48, 48), 2.0
    x = mpmx.sample_sdf(path, N, center, res, spc, seed=3)
    # count: lattice cells with a corner inside, times spc * scale^3
    inside = phi < 0
    corners = [inside[a:n - 1 + a, b:n - 1 + b, c:n - 1 + c] for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    valid = int(np.logical_or.reduce(corners).sum())
    scale = min(r / (n - 3) for r in res)
    assert len(x) == int(valid * spc * scale ** 3)
    # placement: x = ((s - 1) scale + offset + 4) / N, offset = c - res/2 (y: c + res/4)
    off = np.array([center[0] - res[0] // 2, center[1] + res[1] // 4, center[2] - res[2] // 2], float)
    s = (x * N - off - 4) / scale + 1
    assert s.min() >= 0 and s.max() <= n - 1
    assert (trilinear(phi, s) < 0).all()
    # the samples fill the ball: centroid at the ball's centre
    assert np.allclose(s.mean(0), (n - 1) / 2, atol=0.3)
    x2 = mpmx.sample_sdf(path, N, center, res, spc, seed=3)
    assert np.array_equal(x, x2)
    assert not np.array_equal(x, mpmx.sample_sdf(path, N, center, res, spc, seed=4))
