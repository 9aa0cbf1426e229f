"""GPU tracer statistics (p_StatsPrimary / p_StatsDispersion /
p_StatsRelativeVelocity, src/pp_stats.f90:119-390) and the particle_up /
particle_xp0 bookkeeping, through the C-ABI, against the FP64 oracle on the same
inputs, plus a closed-form case that does not use the oracle."""
import sys

import numpy as np
import pytest

import pandora_amd as pa

pytestmark = pytest.mark.gpu

TWOPI = 2.0 * float(np.float32(np.pi))


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _wrapdiff(a, b):
    d = np.abs(a - b)
    return np.minimum(d, TWOPI - d)


def _compare_record(rg, ro, tol=1e-9):
    for k, v in ro.items():
        assert rg[k] == pytest.approx(v, rel=tol, abs=1e-13), k


def _run_pair(oracle_mod, n, steps, okw, gkw, seed, npart, ts_release=0):
    o = oracle_mod.Oracle(n, **okw)
    g = pa.Pandora(n, **gkw)
    u = pa.init_field_host(n, seed, 2.0 / 3.0, 5.0)
    o.set_state(*u)
    g.set_state(*u)
    rng = np.random.default_rng(seed)
    dx = TWOPI / n
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(npart, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    o.particle_stats_config(True, keep_initial=True, ts_release=ts_release)
    g.particle_stats_config(True, keep_initial=True, ts_release=ts_release)
    to = tg = 0.0
    for s in range(1, steps + 1):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        (t_o, ro), (t_g, rg) = o.particle_stats(), g.particle_stats()
        assert t_g == t_o
        if t_o >= 0:
            _compare_record(rg, ro)
    upo, x0o = o.get_particles_aux()
    upg, x0g = g.get_particles_aux()
    assert np.abs(upg - upo).max() < 1e-11
    assert _wrapdiff(x0g, x0o).max() < 1e-11
    g.close()
    return o


def test_particle_stats_isotropic_fp64(oracle_mod):
    note("tracer statistics, isotropic FP64 N=32, 3000 tracers, release at tstep 2, 6 steps")
    kw = dict(rk3_timestep=2, rk3_dt=0.01, nu=0.01)
    o = _run_pair(oracle_mod, 32, 6, kw, kw, 21, 3000, ts_release=2)
    assert o.particle_stats()[0] == 6


def test_particle_stats_shear_remesh_fp64(oracle_mod):
    note("tracer statistics, shear FP64 N=32, 2000 tracers, 12 steps (remesh at 4 and 10)")
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)
    o = _run_pair(oracle_mod, 32, 12, kw, kw, 7, 2000)
    t, r = o.particle_stats()
    assert t == 12 and r["xrel1"] > 0.0


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_particle_stats_uniform_flow_closed_form(prec):
    """k = 0 mode only: every tracer moves by U dt per step, so the record is
    U_i U_j, U, |U|, (s-1)^2 dt^2 U_i U_j and (2s-3) dt U_i^2 / 2, exactly up to
    rounding, also across the periodic wraps (xp0 follows them)."""
    note(f"tracer statistics closed form, uniform flow N=16, precision {prec}")
    n, dt = 16, 0.05
    U = np.array([0.7, -0.4, 0.25])
    g = pa.Pandora(n, rk3_timestep=2, rk3_dt=dt, nu=0.01, precision=prec)
    u = [np.zeros(g.shape, complex) for _ in range(3)]
    for c in range(3):
        u[c][0, 0, 0] = U[c]
    g.set_state(*u)
    rng = np.random.default_rng(5)
    dx = TWOPI / n
    x0 = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(5000, 3))
    g.set_particles(x0)
    g.particle_stats_config(True, keep_initial=True)
    t = 0.0
    tol = 1e-10 if prec == pa.PDR_FP64 else 1e-6  # FP32 field: U stored in single precision
    for s in range(1, 11):
        t = g.rk3_calc(s, t)
        ts, r = g.particle_stats()
        assert ts == s
        assert r["vrel_abs"] == pytest.approx(np.sqrt((U**2).sum()), rel=tol)
        if s == 1:
            continue
        assert r["rs_u1u3"] == pytest.approx(U[0] * U[2], rel=tol)
        assert r["vt2"] == pytest.approx(U[1], rel=tol)
        d = (s - 1) * dt * U
        assert r["xrel1"] == pytest.approx(d[0] ** 2, rel=tol)
        assert r["x2x3"] == pytest.approx(d[1] * d[2], rel=tol)
        assert r["dfp3"] == pytest.approx(0.5 * (2 * s - 3) * dt * U[2] ** 2, rel=100 * tol)
    xp, _ = g.get_particles()
    _, xp0 = g.get_particles_aux()
    assert np.any(np.abs(xp - x0 - 10 * dt * U) > 1.0)  # some tracers wrapped
    np.testing.assert_allclose(xp - xp0, np.broadcast_to(10 * dt * U, xp.shape), rtol=0, atol=1e-6)
    g.close()


def test_particles_wait_for_release():
    note("tracers before the release step stay put (N=16)")
    n = 16
    g = pa.Pandora(n, rk3_timestep=2, rk3_dt=0.01, nu=0.01)
    g.set_state(*pa.init_field_host(n, 4, 2.0 / 3.0, 4.0))
    x0 = np.random.default_rng(1).uniform(0.5, 5.5, size=(100, 3))
    g.set_particles(x0)
    g.particle_stats_config(True, keep_initial=False, ts_release=3)
    t = 0.0
    for s in (1, 2):
        t = g.rk3_calc(s, t)
    xp, _ = g.get_particles()
    np.testing.assert_array_equal(xp, x0)
    assert g.particle_stats()[0] == -1
    t = g.rk3_calc(3, t)
    xp, _ = g.get_particles()
    assert np.abs(xp - x0).max() > 0.0
    assert g.particle_stats()[0] == -1
    g.close()


def test_tracer_sorting_does_not_change_results():
    """The device re-sorts tracer slots by grid row; positions, particle_up_rk3,
    particle_up and particle_xp0 come back in the caller's numbering and are
    bit-identical to a run that never re-sorts (each tracer's arithmetic does
    not depend on its slot)."""
    note("tracer sort: every step vs never, shear FP32 N=64, 20000 tracers, 12 steps (remesh)")
    n = 64
    res = []
    for every in (0, 1, 5):
        g = pa.Pandora(n, f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005,
                       precision=pa.PDR_FP32_MIXED)
        g.set_state(*pa.init_field_host(n, 3, 2.0 / 3.0, 6.0))
        g.set_particle_sort(every)
        g.particle_stats_config(True, keep_initial=True)
        xp = np.random.default_rng(44).uniform(-0.05, TWOPI - 0.05, size=(20000, 3))
        g.set_particles(xp)
        t = 0.0
        for s in range(1, 13):
            t = g.rk3_calc(s, t)
        x, r = g.get_particles()
        up, x0 = g.get_particles_aux()
        res.append((x, r, up, x0, g.particle_stats()[1]))
        g.close()
    for other in res[1:]:
        for a, b in zip(res[0][:4], other[:4]):
            np.testing.assert_array_equal(a, b)
        for k, v in res[0][4].items():  # sums in another order: rounding only
            assert other[4][k] == pytest.approx(v, rel=1e-12, abs=1e-15), k


def test_empty_tracer_set_steps_and_reports_no_record():
    """An empty upload (np = 0) after tracers: the step runs the field path only,
    reading positions back is the 'no particles set' error, and no tracer record
    is formed."""
    note("empty tracer set after 100 tracers, shear N=32, 3 steps")
    n = 32
    g = pa.Pandora(n, f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)
    g.set_state(*pa.init_field_host(n, 6, 2.0 / 3.0, 6.0))
    g.set_particles(np.full((100, 3), 1.0))
    g.particle_stats_config(True, keep_initial=True)
    g.set_particles(np.zeros((0, 3)))
    t = 0.0
    for s in range(1, 4):
        t = g.rk3_calc(s, t)
    assert g.check_finite()
    with pytest.raises(pa.PdrError):
        g.get_particles()
    assert g.particle_stats()[0] == -1
    g.close()
