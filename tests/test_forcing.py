"""Pins for the oracle's Ornstein-Uhlenbeck forcing (f_ForceCalc,
src/pf_force.f90:106-268, applied by f_RK3FormFIsotropic, src/pf_rk3.f90:2421-2467).

The draws come from the build's counter-based generator (the reference's
gfortran stream cannot be reproduced), so the random stream itself is parity
unpinned; what is pinned here: the deviates are N(0, 1) and independent, the OU
update's decay factor and amplitude (stationary variance 2 sigma^2 a / (1 -
(1 - a)^2) with a = dt / FORCE_TF), the reference's statement order on the i = 0
plane (the k = 0 line is updated twice per call), the FORCE_STOP cut-off, the
forced-mode lookup, the solenoidal projection of the force and its entry into
every RK stage (a constant force F drives a single mode by exactly dt * P F).
"""
import numpy as np
import pytest


def _orc(oracle_mod, **kw):
    base = dict(f_type=1, rk3_timestep=2, rk3_dt=0.1, nu=0.0, force_stop=1000, force_kf=2.5,
                force_sigma=0.0, force_tf=1.0, force_seed=7)
    base.update(kw)
    return oracle_mod.Oracle(16, **base)


def _hermitian_plane(rng, nk):
    """Random forcing arrays, conjugate symmetric on the i = 0 plane."""
    w = 2 * nk + 1
    f = [rng.standard_normal((nk + 1, w, w)) + 1j * rng.standard_normal((nk + 1, w, w)) for _ in range(3)]
    for a in f:
        for j in range(-nk, nk + 1):
            for k in range(-nk, nk + 1):
                if (j, k) < (-j, -k):
                    a[0, j + nk, k + nk] = np.conj(a[0, -j + nk, -k + nk])
        a[0, nk, nk] = a[0, nk, nk].real
    return f


def test_normal_pair_moments(oracle_mod):
    z = np.array([oracle_mod.normal_pair(11, k) for k in range(40000)])
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.03
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.02
    # fourth moment of N(0,1) is 3
    assert abs((z**4).mean() - 3.0) < 0.15


def test_ou_decay_and_i0_statement_order(oracle_mod):
    o = _orc(oracle_mod)
    nk = o.force_nk
    assert nk == 2
    f0 = _hermitian_plane(np.random.default_rng(1), nk)
    o.set_forcing(*f0)
    dt, tf = 0.1, 1.0
    a = dt / tf
    o.force_calc(5, dt)
    f1 = o.get_forcing()
    for c in range(3):
        # i >= 1: one update per call
        assert np.allclose(f1[c][1:], f0[c][1:] * (1 - a), rtol=0, atol=1e-15)
        for j in range(-nk, nk + 1):
            for k in range(-nk, nk + 1):
                got = f1[c][0, j + nk, k + nk]
                ref = f0[c][0, j + nk, k + nk]
                if j == 0 and k == 0:
                    want = np.conj(ref * (1 - a))          # copy onto itself flips Im
                elif k == 0:
                    want = ref * (1 - a) ** 2              # updated twice (j > 0 branch)
                else:
                    want = ref * (1 - a)
                assert abs(got - want) < 1e-15, (c, j, k, got, want)


def test_ou_stationary_variance(oracle_mod):
    sigma, tf, dt = 1.3, 1.0, 0.1
    o = _orc(oracle_mod, force_sigma=sigma, force_tf=tf)
    a = dt / tf
    samples = []
    for n in range(1500):
        o.force_calc(1, dt)
        if n >= 150:
            samples.append(np.concatenate([f[1:].ravel() for f in o.get_forcing()]))
    s = np.array(samples)
    var_re = s.real.var()
    var_im = s.imag.var()
    want = 2 * sigma**2 * a / (1 - (1 - a) ** 2)
    assert abs(var_re / want - 1) < 0.06, (var_re, want)
    assert abs(var_im / want - 1) < 0.06, (var_im, want)
    assert abs(s.real.mean()) < 0.1 * np.sqrt(want)


def test_force_stop_zeroes(oracle_mod):
    o = _orc(oracle_mod, force_sigma=1.0, force_stop=3)
    o.force_calc(2, 0.1)
    assert max(np.abs(f).max() for f in o.get_forcing()) > 0
    o.force_calc(3, 0.1)
    assert max(np.abs(f).max() for f in o.get_forcing()) == 0.0


def test_constant_force_drives_single_mode(oracle_mod):
    """nu = 0, u = 0, a constant force on k = (1, 0, 0) only: one RK3 step gives
    u(k) = dt * (F - k (k.F)/|k|^2) exactly (the single mode's nonlinear term is a
    gradient, removed by the projection; the RK3 weights sum to one)."""
    dt = 0.05
    o = _orc(oracle_mod, rk3_dt=dt, force_tf=1e300, force_kf=1.0)
    nk = o.force_nk
    assert nk == 1
    F = np.array([0.3 + 0.1j, 0.7 - 0.2j, -0.4 + 0.5j])
    f = [np.zeros((nk + 1, 2 * nk + 1, 2 * nk + 1), complex) for _ in range(3)]
    for c in range(3):
        f[c][1, nk, nk] = F[c]
    o.set_forcing(*f)
    z = np.zeros(o.shape, complex)
    o.set_state(z, z, z)
    o.step(1, 0.0)
    u = o.get_state()
    want = dt * np.array([0.0, F[1], F[2]])
    got = np.array([u[c][1, 0, 0] for c in range(3)])
    assert np.abs(got - want).max() < 1e-15
    rest = sum(np.abs(u[c]).sum() for c in range(3)) - np.abs(got).sum()
    assert rest < 1e-14
