"""Closed-form pins of the oracle's branch quirks, boundaries and clamps (CPU).

Each case is built so that the quirk decides the answer: the expected value
comes from the mathematics of the special state (principal values of a
diagonal F, where U = V = I and C^-1 acts on the deviatoric and volumetric
parts separately), and each test also checks that the textbook alternative
(radial return, no cap, clamped alpha, component clamp, ...) would give a
different answer, so a transcription that drifted towards it fails.

References: applyVonMises (ExplicitTimeIntegrator.cu:132-370), the sticky
walls of collideWithGround (TimeIntegrator.cu:117-165), the grips
(PullGrid_kernal, TimeIntegrator.cu:169-261), the MLS speed clamp
(G2PKernal.cu:134-142), the damage update (G2PKernal.cu:155) and
computeDt (Simulator.cpp:200-225).
"""
import math

import numpy as np
import pytest

S23 = math.sqrt(2.0 / 3.0)


def lame(p):
    lam = p.E * p.nu / ((1 + p.nu) * (1 - 2 * p.nu))
    mu = p.E / (2 * (1 + p.nu))
    return lam, mu, lam + 2.0 * mu / 3.0


def yield_radius(p, alpha):
    """R(alpha) sqrt(2/3): the exponential hardening of the reference's R()."""
    Y = p.yield_stress
    return (S23 * Y - (S23 * Y - Y) * math.exp(-p.hardening * alpha)) * S23


def trial(p, eps):
    """Trial Kirchhoff stress of principal log strains eps (g = 1)."""
    lam, mu, kappa = lame(p)
    eps = np.asarray(eps, float)
    tr = eps.sum()
    tau = 2 * mu * eps + lam * tr
    return tau, tau.sum(), tau - tau.sum() / 3.0


def flowed_strain(p, eps, dlam):
    """Principal log strain after the reference's update with flow dlam: the
    deviatoric trial stress shrinks by taudiv = (2/3) tr(tau) dlam / |tau_dev|
    and C^-1 maps deviatoric stress by 1/(2 mu), the mean by 1/(3 kappa)."""
    lam, mu, kappa = lame(p)
    eps = np.asarray(eps, float)
    tau, tr_tau, dev = trial(p, eps)
    taudiv = (2.0 / 3.0) * tr_tau * dlam / np.linalg.norm(dev)
    e_dev = eps - eps.sum() / 3.0
    return e_dev * (1.0 - taudiv) + eps.sum() / 3.0


def vm(oracle, p, eps, alpha=0.0, psi=0.0, c=1.0):
    F = np.diag(np.exp(np.asarray(eps, float)))
    return oracle.von_mises(p, F, c, 1.0, alpha, psi)


def implied_Gc(p, r):
    """G_c from FP = 4 L0 (1 - theta) psi / G_c + 1 (ExplicitTimeIntegrator.cu:334)."""
    L0 = 0.5 / p.N
    return 4 * L0 * (1 - p.theta) * r["psi"] / (r["FP"] - 1.0)


def Gc_of(p, alpha):
    pp = alpha / 2
    return p.Gc0 if pp < p.ps else p.Gc0 * ((1 - p.residual_c) * math.exp(p.ps - pp) + p.residual_c)


# ------------------------------------------------------------ yield branch
def test_compressive_trial_past_yield_does_not_flow(oracle):
    """tr(tau) <= 0 makes bss <= 0, so dlambda = 0: a shear with a slight
    compression far past yield keeps F and alpha, where a radial return would
    flow.  (An exactly trace-free state is decided by the sign of the rounding
    of tr(tau): a positive rounding residue gives the capped flow below.)"""
    p = oracle.default_params(E=1e5, nu=0.3)
    a = 0.05
    eps = [a, 0.0, -a - 1e-6]
    tau, tr_tau, dev = trial(p, eps)
    assert tr_tau < 0 and abs(tr_tau) < 1e-3 * np.abs(tau).max()
    assert np.linalg.norm(tau) - yield_radius(p, 0.0) > 0.5 * np.linalg.norm(tau)   # far past yield
    r = vm(oracle, p, eps)
    assert np.allclose(r["F"], np.diag(np.exp(eps)), rtol=0, atol=1e-12)
    assert r["alpha"] == 0.0
    # a textbook radial return to the yield surface would shrink the deviator
    ret = np.exp(np.asarray(eps) * yield_radius(p, 0.0) / np.linalg.norm(dev) * 1.0)
    assert np.abs(np.diag(r["F"]) - ret).max() > 1e-3


def test_flow_increment_is_capped_at_100(oracle):
    """A tiny positive trace under a large deviator makes y / bss huge; the
    reference caps dlambda at 100 (:217-220) and stores alpha = alpha_c."""
    p = oracle.default_params(E=1e5, nu=0.3)
    eps = [0.02, 1e-6, -0.02 + 1.5e-6]
    tau, tr_tau, dev = trial(p, eps)
    lam, mu, kappa = lame(p)
    y = np.linalg.norm(tau) - yield_radius(p, 0.0)
    dl_uncapped = y / ((2.0 / 3.0) * tr_tau)
    assert dl_uncapped > 1e3
    r = vm(oracle, p, eps)
    expect = np.exp(flowed_strain(p, eps, 100.0))
    assert np.abs(np.diag(r["F"]) - expect).max() < 1e-9
    assert np.abs(r["F"] - np.diag(np.diag(r["F"]))).max() < 1e-12
    assert np.abs(expect - np.exp(flowed_strain(p, eps, dl_uncapped))).max() > 1e-3   # the cap matters
    assert r["alpha"] == p.alpha_c


def test_alpha_storage_clamps_but_fracture_energy_uses_local_alpha(oracle):
    """alpha = 0.99 and a flow of dlambda = 1: the stored alpha is clamped to
    alpha_c = 1 (:226-231), while G_c comes from the unclamped local alpha."""
    from scipy.optimize import brentq

    p = oracle.default_params(E=1e5, nu=0.3)
    alpha0, t = 0.99, 0.004
    lam, mu, kappa = lame(p)

    def excess(a):   # y / bss - 1 for eps = a (1, 0, -1) + t/3 (1, 1, 1)
        eps = np.array([a, 0.0, -a]) + t / 3.0
        tau, tr_tau, dev = trial(p, eps)
        return (np.linalg.norm(tau) - yield_radius(p, alpha0)) / ((2.0 / 3.0) * tr_tau) - 1.0

    a = brentq(excess, 1e-4, 0.5, xtol=1e-15)
    eps = np.array([a, 0.0, -a]) + t / 3.0
    psi0 = 1000.0   # history above the new energy: psi stays, FP = 4 L0 (1-theta) psi0 / G_c + 1
    r = vm(oracle, p, eps, alpha=alpha0, psi=psi0)
    assert r["alpha"] == p.alpha_c
    alpha_local = alpha0 + S23 * 1.0
    assert r["psi"] == psi0
    # (the reference's float literals, sqrt(2/3) and theta among them, move it by ~2e-8)
    assert implied_Gc(p, r) == pytest.approx(Gc_of(p, alpha_local), rel=1e-6)
    assert abs(Gc_of(p, alpha_local) - Gc_of(p, 1.0)) > 0.1   # clamped alpha would differ
    assert np.abs(np.diag(r["F"]) - np.exp(flowed_strain(p, eps, 1.0))).max() < 1e-9


def test_large_log_strain_components_are_rescaled_by_the_full_norm(oracle):
    """Components of the new log strain above 3 become eps 3 / |eps| (:265-270);
    the others stay, so this is neither a component clamp nor a norm clamp."""
    p = oracle.default_params(E=1e3, nu=0.3, yield_stress=1e6)   # no flow: eps_nn = eps
    eps = np.array([3.6, -1.6, -1.9])
    tau, tr_tau, dev = trial(p, eps)
    assert np.linalg.norm(tau) < yield_radius(p, 0.0)   # elastic: the branch is not taken
    r = vm(oracle, p, eps)
    assert np.allclose(np.diag(r["F"]), np.exp(eps), rtol=1e-12)   # rescale lives in the yield branch

    # a modulus that puts the trial stress 10 % past yield: a moderate flow
    tau1, _, _ = trial(oracle.default_params(E=1.0, nu=0.3), eps)
    p2 = oracle.default_params(E=1.1 * yield_radius(oracle.default_params(), 0.0) / np.linalg.norm(tau1), nu=0.3)
    r2 = vm(oracle, p2, eps)
    tau, tr_tau, dev = trial(p2, eps)
    y = np.linalg.norm(tau) - yield_radius(p2, 0.0)
    dl = min(y / ((2.0 / 3.0) * tr_tau), 100.0)
    e_nn = flowed_strain(p2, eps, dl)
    assert e_nn[0] > 3.0 and (e_nn[1:] <= 3.0).all()
    expect = e_nn.copy()
    expect[0] = e_nn[0] * 3.0 / np.linalg.norm(e_nn)
    assert np.allclose(np.diag(r2["F"]), np.exp(expect), rtol=1e-9)
    assert abs(expect[0] - 3.0) > 0.05                                   # not a clamp to 3
    assert abs(e_nn[1] * 3.0 / np.linalg.norm(e_nn) - e_nn[1]) > 0.05   # not a norm rescale of all


# ------------------------------------------------------------ walls and grips
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sticky_walls_sign_and_axis(oracle, axis):
    p = oracle.default_params(N=128)
    W, N = p.wall_cells, p.N

    def node(coord):
        c = [60, 60, 60]
        c[axis] = coord
        return c

    def vel(sign):
        v = np.array([0.3, -0.2, 0.1])
        v[axis] = sign * 1.5
        return v

    for coord, sign, stops in [(W, -1, True), (W, +1, False), (W + 1, -1, False), (0, -1, True),
                               (N - W + 1, +1, True), (N - W, +1, False), (N - W + 1, -1, False)]:
        out = oracle.node_boundary(p, *node(coord), vel(sign))
        if stops:
            assert (out == 0).all(), (axis, coord, sign, out)   # the whole velocity, not one component
        else:
            assert (out == vel(sign)).all(), (axis, coord, sign, out)


def test_grips_then_walls(oracle):
    p = oracle.default_params(N=128)
    p.grip_lo, p.grip_hi, p.grip_v = 20, 100, 0.5
    v = np.array([0.1, 0.2, -0.3])
    assert (oracle.node_boundary(p, 20, 60, 60, v) == [-0.5, 0, 0]).all()
    assert (oracle.node_boundary(p, 100, 60, 60, v) == [0.5, 0, 0]).all()
    assert (oracle.node_boundary(p, 21, 60, 60, v) == v).all()
    # a grip node inside the low wall moves out through it: the wall stops it
    assert (oracle.node_boundary(p, 3, 60, 60, v) == 0).all()
    # a grip node beside a y wall keeps v_y = 0, which does not point out
    assert (oracle.node_boundary(p, 10, 2, 60, v) == [-0.5, 0, 0]).all()
    p.grip_lo = -1   # off
    assert (oracle.node_boundary(p, 20, 60, 60, v) == v).all()


# ------------------------------------------------------------ G2P clamps
def test_mls_speed_clamp_keeps_direction(oracle):
    vmax = 11.137
    v = np.array([12.0, -16.0, 0.0])   # |v| = 20
    out = oracle.speed_clamp(vmax, v)
    assert np.linalg.norm(out) == pytest.approx(vmax, rel=1e-14)
    assert np.allclose(out / np.linalg.norm(out), v / 20.0, rtol=1e-14)
    slow = np.array([3.0, 4.0, 0.0])
    assert (oracle.speed_clamp(vmax, slow) == slow).all()


def test_damage_is_monotone_and_non_negative(oracle):
    assert oracle.damage_update(0.8, +0.3) == 0.8    # healing is ignored
    assert oracle.damage_update(0.8, -0.3) == pytest.approx(0.5, abs=1e-15)
    assert oracle.damage_update(0.8, -1.3) == 0.0    # never below zero
    assert oracle.damage_update(0.0, -0.1) == 0.0


# ------------------------------------------------------------ CFL
def test_cfl_dt_closed_forms(oracle):
    p = oracle.default_params(N=64, E=2e4, nu=0.25)
    n = 50
    rng = np.random.default_rng(3)
    x = 0.4 + 0.2 * rng.random((n, 3))
    m = np.full(n, 2.5e-6)
    v = np.zeros((n, 3))
    sim = oracle.Sim(p, x, v, m)
    dt, sp = sim.cfl_dt(1e-4, 0.3)
    assert dt == 1e-4 and sp == 0.0                  # at rest: the default
    v[7] = [3.0, 4.0, 0.0]                           # |v| = 5 is the largest
    v[9] = [1.0, 1.0, 1.0]
    sim = oracle.Sim(p, x, v, m)
    dt, sp = sim.cfl_dt(1e-2, 0.3)
    assert sp == pytest.approx(5.0, rel=1e-15)
    assert dt == pytest.approx(0.3 / 64 / 5.0, rel=1e-15)
    assert sim.cfl_dt(1e-4, 0.3)[0] == 1e-4          # the default when it is smaller
    lam, mu, kappa = lame(p)
    c = math.sqrt((lam + 2 * mu) * p.V0 / 2.5e-6)
    dtw, spw = sim.cfl_dt(1e-2, 0.3, wave=True)
    assert spw == pytest.approx(5.0 + c, rel=1e-14)
    assert dtw == pytest.approx(0.3 / 64 / (5.0 + c), rel=1e-14)
