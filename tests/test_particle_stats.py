"""Pins for the oracle's tracer statistics (p_StatsPrimary / p_StatsDispersion /
p_StatsRelativeVelocity, src/pp_stats.f90:119-390) and its particle_xp0
bookkeeping (src/pp_control.f90:1540-1588, src/pp_mpi.f90:1095-1129).

The expected values come from closed forms of flows whose tracer paths are
known: a uniform flow (the k = 0 mode, never updated), and a zero velocity
fluctuation in the sheared Rogallo frame, where tracers keep their moving-frame
coordinates while the frame (the mean shear) changes, across wraps and
remeshes.
"""
import numpy as np
import pytest

TWOPI = 2.0 * float(np.float32(np.pi))


def _uniform(o, U):
    u = [np.zeros(o.shape, complex) for _ in range(3)]
    for c in range(3):
        u[c][0, 0, 0] = U[c]
    o.set_state(*u)


def test_uniform_flow_record(oracle_mod):
    n, dt = 16, 0.05
    U = np.array([0.7, -0.4, 0.25])
    o = oracle_mod.Oracle(n, rk3_timestep=2, rk3_dt=dt, nu=0.01)
    _uniform(o, U)
    rng = np.random.default_rng(5)
    dx = TWOPI / n
    x0 = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(300, 3))
    o.set_particles(x0)
    o.particle_stats_config(True, keep_initial=True, ts_release=0)
    t = 0.0
    for s in range(1, 21):
        t = o.step(s, t)
        ts, r = o.particle_stats()
        assert ts == s
        # relative velocity of tracers = stage-1 velocity (uf = 0)
        for c in range(3):
            assert r[f"vrel{c + 1}"] == pytest.approx(U[c] ** 2, rel=1e-13)
        assert r["vrel_abs"] == pytest.approx(np.sqrt((U**2).sum()), rel=1e-13)
        for k in ("pken", "rep", "stk", "ste", "vs1", "vs2", "vs3"):
            assert r[k] == 0.0
        if s == 1:  # particle_up is zero before the first stage
            assert r["tke"] == 0.0 and r["vt1"] == 0.0 and r["xrel1"] == 0.0
            continue
        # primary statistics use the velocity of the previous step's last stage
        pairs = {"rs_u1u1": (0, 0), "rs_u2u2": (1, 1), "rs_u3u3": (2, 2),
                 "rs_u1u2": (0, 1), "rs_u1u3": (0, 2), "rs_u2u3": (1, 2)}
        for k, (a, b) in pairs.items():
            assert r[k] == pytest.approx(U[a] * U[b], rel=1e-12)
        assert r["tke"] == pytest.approx(0.5 * (U**2).sum(), rel=1e-12)
        for c in range(3):
            assert r[f"vt{c + 1}"] == pytest.approx(U[c], rel=1e-12)
        # displacement since release: (s - 1) dt U for every tracer, across wraps
        d = (s - 1) * dt * U
        for c in range(3):
            assert r[f"xrel{c + 1}"] == pytest.approx(d[c] ** 2, rel=1e-10)
            assert r[f"dfp{c + 1}"] == pytest.approx(0.5 * (2 * s - 3) * dt * U[c] ** 2, rel=1e-8)
        assert r["x1x2"] == pytest.approx(d[0] * d[1], rel=1e-10)
        assert r["x1x3"] == pytest.approx(d[0] * d[2], rel=1e-10)
        assert r["x2x3"] == pytest.approx(d[1] * d[2], rel=1e-10)
    # some tracers crossed the box boundary: xp0 followed their wraps
    xp, _ = o.get_particles()
    _, xp0 = o.get_particles_aux()
    moved = xp - x0
    assert np.any(np.abs(moved - 20 * dt * U) > 1.0)
    np.testing.assert_allclose(xp - xp0, np.broadcast_to(20 * dt * U, xp.shape), atol=1e-12)


def test_rrot_uniform_flow_record(oracle_mod):
    """R_ROT run (laboratory velocity on the sheared grid): tracers take the
    isotropic advance (p_RK3SolveParticleR, src/pp_rk3.f90:307-403), so each moves
    by U dt per step, while vt is binv particle_up (p_RK3AbsoluteVelocityR,
    720-773) and the dispersion binv (x - x0), with binv(1,3) = S t at the start
    of the step (FIXED dt: no remesh)."""
    n, dt, S = 16, 0.05, 0.8
    U = np.array([0.7, -0.4, 0.25])
    o = oracle_mod.Oracle(n, f_type=2, ns_deformed=3, rk3_timestep=2, rk3_dt=dt, nu=0.01, deform_s=S)
    _uniform(o, U)
    rng = np.random.default_rng(6)
    dx = TWOPI / n
    o.set_particles(rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(300, 3)))
    o.particle_stats_config(True, keep_initial=True, ts_release=0)
    t = 0.0
    for s in range(1, 13):
        t = o.step(s, t)
        ts, r = o.particle_stats()
        assert ts == s
        if s == 1:
            continue
        b13 = S * (s - 1) * dt
        assert r["vt1"] == pytest.approx(U[0] + b13 * U[2], rel=1e-12)
        assert r["vt2"] == pytest.approx(U[1], rel=1e-12)
        assert r["vt3"] == pytest.approx(U[2], rel=1e-12)
        assert r["rs_u1u3"] == pytest.approx(U[0] * U[2], rel=1e-12)
        d = (s - 1) * dt * U
        xl = np.array([d[0] + b13 * d[2], d[1], d[2]])
        for c in range(3):
            assert r[f"xrel{c + 1}"] == pytest.approx(xl[c] ** 2, rel=1e-10)
        assert r["x1x3"] == pytest.approx(xl[0] * xl[2], rel=1e-10)
    u = o.get_state()
    assert u[0][0, 0, 0] == U[0] and np.abs(u[0]).sum() == abs(U[0])


def test_release_step(oracle_mod):
    """Particles stay put before P_TS_RELEASE, xp0 is taken at the release, and
    statistics start one step later (tstep > P_TS_RELEASE)."""
    n, dt = 16, 0.05
    U = np.array([0.3, 0.2, -0.5])
    o = oracle_mod.Oracle(n, rk3_timestep=2, rk3_dt=dt, nu=0.01)
    _uniform(o, U)
    rng = np.random.default_rng(8)
    x0 = rng.uniform(1.0, 5.0, size=(40, 3))
    o.set_particles(x0)
    o.particle_stats_config(True, keep_initial=True, ts_release=3)
    t = 0.0
    for s in (1, 2):
        t = o.step(s, t)
        assert o.particle_stats()[0] == -1
    xp, _ = o.get_particles()
    np.testing.assert_array_equal(xp, x0)
    t = o.step(3, t)
    assert o.particle_stats()[0] == -1
    t = o.step(4, t)
    ts, r = o.particle_stats()
    assert ts == 4
    for c in range(3):
        assert r[f"xrel{c + 1}"] == pytest.approx((dt * U[c]) ** 2, rel=1e-10)
        assert r[f"vt{c + 1}"] == pytest.approx(U[c], rel=1e-12)


@pytest.mark.parametrize("decide", [False, True])
def test_sheared_frame_dispersion(oracle_mod, decide):
    """Velocity fluctuation zero in the sheared Rogallo frame: tracers keep their
    moving-frame coordinates (the frame carries the mean shear), so with an
    initial moving-frame offset D = x - x0 the laboratory displacement the
    reference reports, binv (x - x0), is D + S t D3 e1 for all t (t = simulation
    time; binv(1,3) = S t).  In DECIDE mode the frame is remeshed at tstep 4 and
    10 (x1 += x3, binv(1,3) -= 1); the displacement stays on the closed form only
    if xp0 is remeshed and wrapped with xp."""
    n, S = 16, 1.0
    if decide:  # dt = (0.5/S)/F_REMESH_T = 1/6
        kw = dict(f_type=2, rk3_timestep=1, deform_s=S, remesh_t=3, nu=0.005)
    else:
        kw = dict(f_type=2, rk3_timestep=2, rk3_dt=0.1, deform_s=S, nu=0.005)
    o = oracle_mod.Oracle(n, **kw)
    o.set_state(*[np.zeros(o.shape, complex) for _ in range(3)])
    rng = np.random.default_rng(3)
    dx = TWOPI / n
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(200, 3))
    D = rng.uniform(-1.0, 1.0, size=(200, 3))
    o.set_particles(xp)
    o.particle_stats_config(True, keep_initial=True, ts_release=0)
    o.set_particles_aux(xp0=xp - D)
    t = 0.0
    for s in range(1, 15):
        t_start = t  # the record is taken in the frame at the start of the step
        t = o.step(s, t)
        ts, r = o.particle_stats()
        assert ts == s
        L1 = D[:, 0] + S * t_start * D[:, 2]
        assert r["xrel1"] == pytest.approx(np.mean(L1**2), rel=1e-11)
        assert r["xrel2"] == pytest.approx(np.mean(D[:, 1] ** 2), rel=1e-11)
        assert r["xrel3"] == pytest.approx(np.mean(D[:, 2] ** 2), rel=1e-11)
        assert r["x1x2"] == pytest.approx(np.mean(L1 * D[:, 1]), rel=1e-9, abs=1e-13)
        assert r["x1x3"] == pytest.approx(np.mean(L1 * D[:, 2]), rel=1e-9, abs=1e-13)
        assert r["tke"] == 0.0 and r["vrel_abs"] == 0.0
    if decide:
        b, _ = o.get_frame()
        assert abs(b[6] + S * (t - 2.0 / S)) < 1e-12  # two remeshes: bmat(1,3) = -S (t - 2/S)
        xp_end, _ = o.get_particles()
        assert np.any(np.abs(xp_end[:, 0] - xp[:, 0]) > 1e-3)  # the remeshes moved x1


def test_stats_off_by_default(oracle_mod):
    o = oracle_mod.Oracle(16, rk3_timestep=2, rk3_dt=0.01, nu=0.01)
    _uniform(o, [0.1, 0.2, 0.3])
    o.set_particles(np.full((5, 3), 1.0))
    o.step(1, 0.0)
    assert o.particle_stats()[0] == -1
