"""Pins of the serial FP64 oracle (CPU only).

Each oracle function is checked against something it does not share code
with: the reference's own SPGridMask/Materials build (oracle/_ref), LAPACK
singular values (tests/golden/svd_golden.npz), closed forms, invariants and
brute force on small inputs.  The reference ships no tests or golden vectors
of its own for this path.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MASKS = (0x49249249249240C0, 0x2492492492492030, 0x924924924924900C)


def pdep(v, mask):
    out, bit = 0, 0
    for b in range(64):
        if (mask >> b) & 1:
            if (v >> bit) & 1:
                out |= 1 << b
            bit += 1
    return out


def pext(v, mask):
    out, bit = 0, 0
    for b in range(64):
        if (mask >> b) & 1:
            if (v >> b) & 1:
                out |= 1 << bit
            bit += 1
    return out


# ---------------------------------------------------------------- keys
def test_masks_match_reference_build(oracle):
    m = [int(a) for a in oracle.masks()]
    assert tuple(m) == MASKS
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    rm = np.zeros(3, np.uint64)
    R.ref_masks(rm)
    assert [int(a) for a in rm] == m


def test_linear_offset_matches_reference_and_pdep(oracle):
    rng = np.random.default_rng(1)
    R = oracle.ref()
    L = oracle.lib()
    for _ in range(400):
        i, j, k = (int(a) for a in rng.integers(0, 1024, 3))
        off = L.mo_linear_offset(i, j, k)
        assert off == pdep(i, MASKS[0]) | pdep(j, MASKS[1]) | pdep(k, MASKS[2])
        if R is not None:
            assert off == R.ref_linear_offset(i, j, k)
        for a, q in zip(MASKS, (i, j, k)):
            assert L.mo_bit_pack(a, off) == q  # pext inverts pdep
            assert L.mo_bit_spread(a, q) == pdep(q, a)


def test_packed_add_is_coordinate_addition(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    for _ in range(300):
        c = rng.integers(0, 500, 3)
        d = rng.integers(0, 8, 3) * 4
        a = L.mo_linear_offset(*(int(v) for v in c))
        b = L.mo_linear_offset(*(int(v) for v in d))
        assert L.mo_packed_add(a, b) == L.mo_linear_offset(*(int(v) for v in c + d))


def test_cell_key_layout(oracle):
    """page = offset >> 12 is the Morton code (z bit 0, y bit 1, x bit 2 of each
    triple) of (i,j,k) >> 2; bits 2..7 hold the in-page cell (i&3, j&3, k&3)."""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for _ in range(200):
        i, j, k = (int(a) for a in rng.integers(0, 1024, 3))
        off = L.mo_linear_offset(i, j, k)
        pi, pj, pk = i >> 2, j >> 2, k >> 2
        morton = 0
        for b in range(10):
            morton |= ((pk >> b) & 1) << (3 * b) | ((pj >> b) & 1) << (3 * b + 1) | ((pi >> b) & 1) << (3 * b + 2)
        assert off >> 12 == morton
        assert (off & 0xFC) >> 2 == ((i & 3) << 4 | (j & 3) << 2 | (k & 3))
        assert off & 0xF03 == 0


def test_base_node_exact_at_cell_edges(oracle):
    """floor(x N + 1/2) - 1 evaluated exactly for float32 positions, including
    positions one float ulp either side of the cell edge x N + 1/2 = k."""
    N = 128
    for k in (10, 64, 100):
        edge = (k - 0.5) / N
        below = np.nextafter(np.float32(edge), np.float32(0))
        above = np.nextafter(np.float32(edge), np.float32(1))
        x = np.array([[below, above, np.float32(edge)]], np.float32)
        keys = oracle.cell_keys(np.repeat(x, 1, 0), N)
        L = oracle.lib()
        assert L.mo_base_node(float(below), N) == k - 2
        assert L.mo_base_node(float(above), N) == k - 1
        assert L.mo_base_node(float(np.float32(edge)), N) == k - 1
        assert keys[0] == L.mo_linear_offset(k - 2, k - 1, k - 1)


def test_sort_is_stable_key_order(oracle):
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 50, 5000).astype(np.uint64) << np.uint64(12)
    perm, sk = oracle.sort_keys(keys)
    ref = sorted(range(len(keys)), key=lambda i: (int(keys[i]), i))
    assert np.array_equal(perm, np.array(ref, np.int32))
    assert np.array_equal(sk, keys[perm])


def brute_tables(sorted_keys, vpage):
    pages = []
    for k in sorted_keys:
        pk = int(k) >> 12
        if not pages or pages[-1] != pk:
            pages.append(pk)
    pset = set(pages)

    def coords(pk):
        x = y = z = 0
        for b in range(21):
            z |= ((pk >> (3 * b)) & 1) << b
            y |= ((pk >> (3 * b + 1)) & 1) << b
            x |= ((pk >> (3 * b + 2)) & 1) << b
        return x, y, z

    def morton(x, y, z):
        m = 0
        for b in range(21):
            m |= ((z >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((x >> b) & 1) << (3 * b + 2)
        return m

    nbrs = {}
    supp = set()
    for pk in pages:
        x, y, z = coords(pk)
        lst = []
        for q in range(1, 8):
            nk = morton(x + (q >> 2 & 1), y + (q >> 1 & 1), z + (q & 1))
            lst.append(nk)
            if nk not in pset:
                supp.add(nk)
        nbrs[pk] = lst
    allp = pages + sorted(supp)
    ids = {pk: i for i, pk in enumerate(allp)}
    adj = np.array([[ids[nbrs[pk][q]] for pk in pages] for q in range(7)], np.int32)
    counts = [sum(1 for k in sorted_keys if int(k) >> 12 == pk) for pk in pages]
    vp = [(c + vpage - 1) // vpage for c in counts]
    return allp, adj, np.concatenate([[0], np.cumsum(vp)]).astype(np.int32), counts


@pytest.mark.parametrize("vpage", [64, 256])
def test_page_tables_brute_force(oracle, vpage):
    rng = np.random.default_rng(5)
    N = 64
    x = rng.uniform(0.2, 0.6, (3000, 3))
    x[:1000] = rng.uniform(0.45, 0.5, (1000, 3))   # a dense clump: multi-virtual-page pages
    keys = oracle.cell_keys(x, N)
    perm, sk = oracle.sort_keys(keys)
    t = oracle.page_tables(sk, vpage)
    allp, adj, vpo, counts = brute_tables(sk, vpage)
    assert t["numPage"] == len(counts)
    assert list(t["page_key"].astype(np.int64)) == allp
    assert np.array_equal(t["adj"], adj)
    assert np.array_equal(t["vpo"], vpo)
    assert np.array_equal(np.diff(t["page2particle"]), np.array(counts, np.int32))
    tgt = np.repeat(np.arange(1, t["numPage"] + 1), np.diff(vpo))
    assert np.array_equal(t["target"], tgt)
    cells = np.unique(sk)
    assert t["numCell"] == len(cells)
    assert np.array_equal(sk[t["cell2particle"][:-1]], cells)
    assert np.array_equal(t["particle2cell"], np.searchsorted(cells, sk) + 1)


# ---------------------------------------------------------------- SVD
def test_svd_against_lapack_golden(oracle):
    g = np.load(os.path.join(HERE, "golden", "svd_golden.npz"))
    for F, sv, signed in zip(g["F"], g["sv"], g["signed"]):
        U, S, V = oracle.svd3(F)
        scale = max(np.abs(F).max(), 1.0)
        assert np.abs(U @ np.diag(S) @ V.T - F).max() < 1e-12 * scale
        assert np.abs(U.T @ U - np.eye(3)).max() < 1e-12
        assert np.abs(V.T @ V - np.eye(3)).max() < 1e-12
        assert np.abs(np.sort(np.abs(S))[::-1] - sv).max() < 1e-12 * scale
        # conventions: S0 >= S1 >= |S2|, only S2 may be negative
        assert S[0] >= S[1] - 1e-15 and S[1] >= abs(S[2]) - 1e-15 and S[1] >= 0
        if abs(np.linalg.det(F)) > 1e-10:
            assert np.sign(S[2]) == np.sign(signed[2])


def test_svd_special_cases(oracle):
    U, S, V = oracle.svd3(np.eye(3))
    assert np.allclose(S, 1) and np.allclose(U, np.eye(3)) and np.allclose(V, np.eye(3))
    U, S, V = oracle.svd3(np.diag([1.0, 3.0, 2.0]))
    assert np.allclose(S, [3, 2, 1])
    U, S, V = oracle.svd3(np.diag([1.0, 1.0, -1.0]))
    assert np.allclose(S, [1, 1, -1])
    U, S, V = oracle.svd3(np.zeros((3, 3)))
    assert np.allclose(S, 0) and np.allclose(U @ U.T, np.eye(3))


# ---------------------------------------------------------------- constitutive
def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))


def test_lame_matches_reference_build(oracle):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    out = np.zeros(4)
    R.ref_material(1e5, 0.3, out)
    assert np.allclose(out[:3], lame(1e5, 0.3), rtol=1e-15) and out[3] == 50.0


def test_inverse_closed_form(oracle):
    lam, mu, _ = lame(1e5, 0.3)
    C = lam * np.ones((3, 3)) + 2 * mu * np.eye(3)
    D = oracle.inverse3(C)
    closed = np.eye(3) / (2 * mu) - lam / (2 * mu * (2 * mu + 3 * lam)) * np.ones((3, 3))
    assert np.abs(D - closed).max() < 1e-15 * np.abs(closed).max() * 10


def test_von_mises_rest_and_rotation(oracle):
    p = oracle.default_params(E=1e5, nu=0.3, V0=1.0)
    r = oracle.von_mises(p, np.eye(3), 1.0, 1.0, 0.0, 0.0)
    assert np.abs(r["sigma"]).max() < 1e-9 and np.allclose(r["F"], np.eye(3))
    assert r["alpha"] == 0.0 and r["psi"] == 0.0 and r["FP"] == 1.0
    th = (1 - p.theta) + p.theta
    assert r["g"] == pytest.approx(th)
    ang = 0.7
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    r = oracle.von_mises(p, R, 1.0, 1.0, 0.0, 0.0)
    assert np.abs(r["sigma"]).max() < 1e-8 and np.allclose(r["F"], R)


def test_von_mises_isotropic_state_is_a_fixed_point(oracle):
    """F = s R past the yield surface: the yield test on the full trial norm
    fires, but the trial stress has no deviatoric part, so the return mapping
    leaves F alone (the reference divides by the zero deviatoric norm and
    returns NaN).  The stress is the closed form V0 s tau_vol I, with
    tau_vol = kappa/2 (J^2 - 1) and J = s^3, times g under tension."""
    p = oracle.default_params(E=1e5, nu=0.3, V0=1.0, yield_stress=1000.0)
    kappa = 1e5 / (3 * (1 - 2 * 0.3))
    ang = 0.4
    R = np.array([[np.cos(ang), 0, -np.sin(ang)], [0, 1, 0], [np.sin(ang), 0, np.cos(ang)]])
    for s, rot in ((0.99, R), (0.99, np.eye(3)), (1.01, np.eye(3))):
        F = s * rot
        c = 0.9
        r = oracle.von_mises(p, F, c, 1.0, 0.0, 0.0)
        assert np.isfinite(r["F"]).all() and np.isfinite(r["sigma"]).all()
        assert np.abs(r["F"] - F).max() < 1e-12
        J = s ** 3
        g = c * c * (1 - p.theta) + p.theta
        tv = kappa / 2 * (J * J - 1) * (g if J >= 1 else 1.0)
        assert np.abs(r["sigma"] - s * tv * np.eye(3)).max() < 1e-9 * abs(s * tv)


def test_von_mises_small_strain_is_linear_elasticity(oracle):
    """For F = I + eps (symmetric, small) and an intact particle, the Kirchhoff
    stress is V0 (lambda tr(eps) I + 2 mu eps) to first order (kappa = lambda + 2mu/3)."""
    E, nu = 1e5, 0.3
    lam, mu, _ = lame(E, nu)
    p = oracle.default_params(E=E, nu=nu, V0=2.0, yield_stress=1e9)
    rng = np.random.default_rng(6)
    for _ in range(10):
        a = rng.standard_normal((3, 3))
        eps = 1e-6 * (a + a.T)
        r = oracle.von_mises(p, np.eye(3) + eps, 1.0, 1.0, 0.0, 0.0)
        lin = p.V0 * (lam * np.trace(eps) * np.eye(3) + 2 * mu * eps)
        assert np.abs(r["sigma"] - lin).max() < 1e-4 * np.abs(lin).max()


def test_von_mises_objectivity_and_symmetry(oracle):
    p = oracle.default_params(E=1e5, nu=0.3, V0=1.0, yield_stress=1e3)
    rng = np.random.default_rng(8)
    for _ in range(20):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.3:
            continue
        a = rng.standard_normal(3)
        a /= np.linalg.norm(a)
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        Q = np.eye(3) + np.sin(0.9) * K + (1 - np.cos(0.9)) * K @ K
        r1 = oracle.von_mises(p, F, 0.8, 0.7, 0.1, 0.0)
        r2 = oracle.von_mises(p, Q @ F, 0.8, 0.7, 0.1, 0.0)
        s = np.abs(r1["sigma"]).max()
        assert np.abs(r1["sigma"] - r1["sigma"].T).max() < 1e-10 * s
        assert np.abs(Q @ r1["sigma"] @ Q.T - r2["sigma"]).max() < 1e-10 * s
        assert np.abs(Q @ r1["F"] - r2["F"]).max() < 1e-10
        # the return mapping keeps the singular vectors: F_new F^-1 is symmetric
        M = r1["F"] @ np.linalg.inv(F)
        assert np.abs(M - M.T).max() < 1e-10 * np.abs(M).max()
        assert r1["alpha"] <= p.alpha_c


def test_phase_field_coefficients_closed_form(oracle):
    """FP = 4 L0 (1-theta) psi / G_c + 1 with L0 = dx/2, G_c = 3 while alpha/2 < 0.01;
    g = c^2 (1-theta) + theta."""
    p = oracle.default_params(E=1e5, nu=0.3, V0=1.0, N=128)
    psi = 123.0
    r = oracle.von_mises(p, np.eye(3), 0.6, 1.0, 0.0, psi)
    L0 = 0.5 / 128
    assert r["psi"] == psi
    assert r["FP"] == pytest.approx(4 * L0 * (1 - p.theta) * psi / 3.0 + 1, rel=1e-14)
    assert r["g"] == pytest.approx(0.36 * (1 - p.theta) + p.theta, rel=1e-14)


# ---------------------------------------------------------------- transfers
def bspline(d0):
    """quadratic B-spline weights and derivatives of nodes base+0..2 (textbook form)."""
    w = np.array([0.5 * (1.5 - d0) ** 2, 0.75 - (d0 - 1) ** 2, 0.5 * (d0 - 0.5) ** 2])
    g = np.array([-(1.5 - d0), -2 * (d0 - 1), d0 - 0.5])
    return w, g


def test_g2p_reproduces_affine_fields(oracle):
    """Quadratic B-splines reproduce linear fields: v_p = A x_p + b exactly,
    sum w v_i (x_i - x_p)^T = A dx^2/4 and sum v_i grad w^T = A."""
    rng = np.random.default_rng(9)
    N = 64
    dx = 1.0 / N
    for _ in range(20):
        x = rng.uniform(0.3, 0.7, 3)
        A = rng.standard_normal((3, 3))
        b = rng.standard_normal(3)
        vp, Bm, G = oracle.g2p_affine(N, x, A, b)
        assert np.abs(vp - (A @ x + b)).max() < 1e-12
        assert np.abs(Bm - A * dx * dx / 4).max() < 1e-14
        assert np.abs(G - A).max() < 1e-11


def test_p2g_conserves_mass_and_momentum(oracle):
    rng = np.random.default_rng(10)
    N = 64
    n = 500
    x = rng.uniform(0.4, 0.6, (n, 3))
    keys = oracle.cell_keys(x, N)
    perm, sk = oracle.sort_keys(keys)
    x = x[perm]
    t = oracle.page_tables(sk)
    v = rng.standard_normal((n, 3))
    B = 1e-5 * rng.standard_normal((n, 3, 3))
    sig = rng.standard_normal((n, 3, 3))
    sig = sig + sig.transpose(0, 2, 1)
    m = rng.uniform(0.5, 1.5, n)
    c = rng.uniform(0, 1, n)
    vol = rng.uniform(0.5, 1, n)
    FP = rng.uniform(1, 2, n)
    p = oracle.default_params(N=N)
    g = oracle.p2g_mls(p, sk, x, v, B, sig, m, c, vol, FP, 1e-4, t["page_key"])
    assert g[0].sum() == pytest.approx(m.sum(), rel=1e-13)
    P = (m[:, None] * v).sum(0)
    assert np.abs(g[1:4].sum(1) - P).max() < 1e-12 * np.abs(P).max()
    assert g[5].sum() == pytest.approx(n, rel=1e-13)
    assert g[4].sum() == pytest.approx(c.sum(), rel=1e-13)
    assert g[6].sum() == pytest.approx(vol.sum(), rel=1e-13)
    assert g[7].sum() == pytest.approx((vol * FP).sum(), rel=1e-13)


def lattice(n_cells, lo, N, rng, jitter=0.125):
    pts = []
    for i in range(lo, lo + n_cells):
        for j in range(lo, lo + n_cells):
            for k in range(lo, lo + n_cells):
                for q in range(8):
                    s = [0.25 + 0.5 * ((q >> (2 - a)) & 1) for a in range(3)]
                    pts.append([(c + s[a] + rng.uniform(-jitter, jitter)) / N for a, c in enumerate((i, j, k))])
    return np.array(pts)


def test_substep_free_flight_conservation(oracle):
    rng = np.random.default_rng(11)
    N = 64
    x = lattice(4, 30, N, rng)
    n = len(x)
    v = np.tile([0.3, -0.1, 0.2], (n, 1))
    m = np.full(n, 1e-3)
    p = oracle.default_params(N=N, V0=(1 / N) ** 3 / 8)
    sim = oracle.Sim(p, x, v, m)
    dt = 1e-4
    for _ in range(5):
        sim.substep(dt)
    cons = sim.conservation()
    assert cons[0] == pytest.approx(m.sum(), rel=1e-14)
    expect = (m[:, None] * v).sum(0) + m.sum() * np.array([0, -10.0, 0]) * 5 * dt
    assert np.abs(cons[1:] - expect).max() < 1e-12 * np.abs(expect).max()
    # rigid translation: no strain, no damage, one PCG iteration (x = c exactly)
    assert np.abs(sim.get("F") - np.eye(3)).max() < 1e-12
    assert sim.get("c").min() == 1.0
    assert sim.counts()["cg_iters"] == 1


def test_zero_weight_node_stays_inactive(oracle):
    """A particle with a coordinate exactly on a cell centre (x N = k + 1/2)
    gives its third node along that axis the weight 0.  Alone, that node has
    no mass and must stay out of the grid update and the damage solve: the
    particle then moves exactly as in free flight, v = v0 + g dt."""
    N = 64
    x = np.array([[20.5 / N, 30.5 / N, 40.25 / N]])
    v0 = np.array([[0.2, -0.1, 0.05]])
    p = oracle.default_params(N=N, V0=(1 / N) ** 3 / 8)
    sim = oracle.Sim(p, x, v0, np.array([1e-3]))
    dt = 1e-4
    sim.substep(dt)
    assert np.isfinite(sim.get("v")).all() and np.isfinite(sim.get("c")).all()
    assert np.abs(sim.get("v") - (v0 + np.array([0, -10.0, 0]) * dt)).max() < 1e-13
    w = sim.grid(8, stage=1)
    m = sim.grid(0, stage=1)
    touched = sim.grid(9, stage=1) > 0
    # 27 contributions; weight 0 on the third node of both cell-centre axes: 2 x 2 x 3 remain
    assert touched.sum() == 27 and (w > 0).sum() == 12
    assert np.all(m[touched & (w == 0)] == 0)
    assert sim.get("c")[0] == pytest.approx(1.0, abs=1e-12)


def test_block_max_dot_closed_forms(oracle):
    """The reference's PCG "dot" (getDotProduct, ExplicitTimeIntegrator.cu:731-747)
    sums 128-entry blocks and reduces the block sums with a maximum.  Closed
    forms that a plain dot, another block size, a sum of block maxima or a
    zero start of the maximum would each fail."""
    ones = np.ones(192)
    assert oracle.vdot(ones, ones, 0) == 128.0     # blocks of 128 and 64
    assert oracle.vdot(ones, ones, 1) == 192.0
    a = np.concatenate([-np.ones(128), 0.5 * np.ones(64)])
    assert oracle.vdot(a, np.ones(192), 0) == 32.0          # max(-128, 32)
    assert oracle.vdot(a, np.ones(192), 1) == -96.0
    b = np.concatenate([-np.ones(128), -4.0 * np.ones(64)])
    assert oracle.vdot(b, np.ones(192), 0) == -128.0        # max(-128, -256): no 0 start
    c = np.concatenate([np.ones(64), 3.0 * np.ones(64), 2.0 * np.ones(128), np.ones(64)])
    assert oracle.vdot(c, np.ones(320), 0) == 256.0         # blocks 256, 256, 64
    assert oracle.vdot(np.arange(64.0), np.ones(64), 0) == 2016.0   # one page: its sum


def test_pcg_matches_dense_solve(oracle):
    """Phase-field system (D + M K) x = b assembled densely from the particles
    with textbook B-spline weights; the oracle's PCG (true dot products, tight
    tolerance) must reach numpy.linalg.solve's answer."""
    rng = np.random.default_rng(12)
    N = 32
    dx = 1.0 / N
    x = lattice(2, 14, N, rng)
    n = len(x)
    v = 0.05 * rng.standard_normal((n, 3))
    m = np.full(n, 1e-3)
    p = oracle.default_params(N=N, V0=dx ** 3 / 8, dot_mode=1, cg_tol=1e-16, cg_max_iter=2000,
                              gravity=(0.0, 0.0, 0.0))
    sim = oracle.Sim(p, x, v, m)
    sim.set("psi", rng.uniform(0, 2e4, n))
    sim.set("c", rng.uniform(0.2, 1.0, n))
    x0 = sim.get("x").copy()
    dt = 1e-4
    sim.substep(dt)
    vol, FP = sim.get("vol"), sim.get("FP")
    c_grid = sim.grid(7, stage=1)
    xsol = sim.grid(7, stage=2) + c_grid
    rhs_o = sim.grid(10, stage=2)
    tag = sim.grid(8, stage=1) > 0  # active nodes: sum w > 0
    pages = sim.pages()

    def coords(pk):
        out = [0, 0, 0]
        for b in range(10):
            out[2] |= ((int(pk) >> (3 * b)) & 1) << b
            out[1] |= ((int(pk) >> (3 * b + 1)) & 1) << b
            out[0] |= ((int(pk) >> (3 * b + 2)) & 1) << b
        return out

    index = {}
    for pid, pk in enumerate(pages):
        px, py, pz = coords(pk)
        for cell in range(64):
            index[(4 * px + (cell >> 4), 4 * py + ((cell >> 2) & 3), 4 * pz + (cell & 3))] = pid * 64 + cell
    nn = len(pages) * 64
    Dg = np.zeros(nn)
    Kg = np.zeros((nn, nn))
    bg = np.zeros(nn)
    Mp = p.M
    for q in range(n):
        base = np.floor(x0[q] * N + 0.5).astype(int) - 1
        d0 = x0[q] * N - base
        ws, gs = zip(*(bspline(d) for d in d0))
        idx, wv, gv = [], [], []
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    idx.append(index[(base[0] + i, base[1] + j, base[2] + k)])
                    wv.append(ws[0][i] * ws[1][j] * ws[2][k])
                    gv.append(np.array([gs[0][i] * ws[1][j] * ws[2][k], ws[0][i] * gs[1][j] * ws[2][k],
                                        ws[0][i] * ws[1][j] * gs[2][k]]) / dx)
        for a, ia in enumerate(idx):
            Dg[ia] += vol[q] * wv[a] * (FP[q] * Mp + 1 / dt)
            bg[ia] += vol[q] * wv[a] * (c_grid[ia] / dt + Mp)
            for b, ib in enumerate(idx):
                Kg[ia, ib] += dx * dx * vol[q] * gv[a] @ gv[b]
    act = np.where(tag)[0]
    A = np.diag(Dg[act]) + Mp * Kg[np.ix_(act, act)]
    ref = np.linalg.solve(A, bg[act])
    assert np.abs(rhs_o[act] - bg[act]).max() < 1e-12 * np.abs(bg).max()
    assert np.abs(xsol[act] - ref).max() < 1e-9 * np.abs(ref).max()


# ---------------------------------------------------------------- other energies
def _fd_gradient(energy, F, h=1e-6):
    P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            E = np.zeros((3, 3))
            E[i, j] = h
            P[i, j] = (energy(F + E) - energy(F - E)) / (2 * h)
    return P


def test_fixed_corotated_is_energy_gradient(oracle):
    """sigma = V0 P F^T with P the gradient of the fixed corotated energy
    mu sum (s_i - 1)^2 + lambda/2 (J - 1)^2 (central differences)."""
    E, nu = 1e5, 0.3
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    p = oracle.default_params(E=E, nu=nu, V0=2.0, model=0)
    rng = np.random.default_rng(21)

    def psi(F):
        s = np.linalg.svd(F, compute_uv=False)
        return mu * ((s - 1) ** 2).sum() + lam / 2 * (np.linalg.det(F) - 1) ** 2

    for _ in range(20):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.3:
            continue
        sig = oracle.fixed_corotated(p, F)
        P = sig @ np.linalg.inv(F).T / 2.0
        Pfd = _fd_gradient(psi, F)
        assert np.abs(P - Pfd).max() < 1e-5 * np.abs(Pfd).max()
    ang = 0.3
    R = np.array([[1, 0, 0], [0, np.cos(ang), -np.sin(ang)], [0, np.sin(ang), np.cos(ang)]])
    assert np.abs(oracle.fixed_corotated(p, R)).max() < 1e-9


def test_neo_hookean_stress_is_energy_gradient(oracle):
    """The reference fills P_hat with the principal Kirchhoff stresses of the
    Neo-Hookean split g mu/2 (tr b_bar - 3) + w kappa/2 ((J^2 - 1)/2 - ln J)
    (w = g for J >= 1, 1 for J < 1) and assembles V0 U diag(P_hat) V^T F^T,
    which is V0 tau sqrt(F F^T): tau from central differences of the energy,
    the left stretch from the eigen decomposition of F F^T.  Rotations carry
    no stress."""
    E, nu = 1e5, 0.3
    mu, kappa = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
    p = oracle.default_params(E=E, nu=nu, V0=1.0, model=1)
    rng = np.random.default_rng(22)
    c = 0.7
    g = c * c * (1 - p.theta) + p.theta
    for target in (0.9, 1.1):
        for _ in range(10):
            A = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
            F = A * (target / np.linalg.det(A)) ** (1 / 3)   # J = target

            def psi(Fx):
                J = np.linalg.det(Fx)
                b = Fx @ Fx.T * J ** (-2 / 3)
                w = g if J >= 1 else 1.0
                return g * mu / 2 * (np.trace(b) - 3) + w * kappa / 2 * ((J * J - 1) / 2 - np.log(J))

            r = oracle.neo_hookean(p, F, c, 0.0)
            tau_fd = _fd_gradient(psi, F) @ F.T
            lam_b, Q = np.linalg.eigh(F @ F.T)
            stretch = Q @ np.diag(np.sqrt(lam_b)) @ Q.T
            expect = tau_fd @ stretch
            assert np.abs(r["sigma"] - expect).max() < 1e-5 * np.abs(expect).max()
            assert r["vol"] == pytest.approx(np.linalg.det(F), rel=1e-12)
    r = oracle.neo_hookean(p, np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]]), c, 0.0)
    assert np.abs(r["sigma"]).max() < 1e-9 and r["psi"] == 0.0


def test_neo_hookean_history_at_unit_volume(oracle):
    """J = 1: psi+ is the deviatoric energy mu/2 (sum s_i^2 - 3) alone, taken
    only when it exceeds the history, and FP = 4 L0 (1 - 0.001) psi / 3 + 1."""
    E, nu, N = 1e5, 0.3, 64
    mu = E / (2 * (1 + nu))
    p = oracle.default_params(E=E, nu=nu, V0=1.0, N=N, model=1)
    s = np.array([1.1, 1 / 1.1 / 0.95, 0.95])
    F = np.diag(s)
    dev = mu / 2 * ((s ** 2).sum() - 3)
    r = oracle.neo_hookean(p, F, 1.0, 0.0)
    assert r["psi"] == pytest.approx(dev, rel=1e-10)
    L0 = 0.5 / N
    assert r["FP"] == pytest.approx(4 * L0 * (1 - float(np.float32(0.001))) * dev / 3 + 1, rel=1e-12)
    r2 = oracle.neo_hookean(p, F, 1.0, 2 * dev)
    assert r2["psi"] == 2 * dev
