"""The tracer-stage boundary: a caller that drives the tracers itself, as the
reference's f_RK3RealSpace does (p_ControlRK3 then p_ControlFinishTimeStep,
src/pf_rk3.f90:1145-1176), or through p_RK3SolveParticle[BF|R]
(src/pp_rk3.f90:57, 169, 307), gets the in-stage advance's result bit for bit
(positions, particle_up_rk3, particle_up, particle_xp0 and the tracer
statistics record), and misuse is refused with the documented codes."""
import sys

import numpy as np
import pytest

import pandora_amd as pa

pytestmark = pytest.mark.gpu

TWOPI = 2.0 * float(np.float32(np.pi))
CSTAGE = (1.0 / 3.0, 3.0 / 4.0, 1.0)


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


CASES = {
    "iso": dict(f_type=1, rk3_timestep=2, rk3_dt=0.01, nu=0.01),
    "bfrot": dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005),
    "rrot": dict(f_type=2, ns_deformed=3, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005),
}
FORM = {"iso": 0, "bfrot": 1, "rrot": 2}


def _make(case, prec, n=32, release=2):
    g = pa.Pandora(n, precision=prec, **CASES[case])
    g.set_state(*pa.init_field_host(n, 9, 2.0 / 3.0, 5.0))
    dx = TWOPI / n
    xp = np.random.default_rng(4).uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(3000, 3))
    g.particle_stats_config(enable=True, keep_initial=True, ts_release=release)
    g.set_particles(xp)
    return g


def _result(g):
    xp, uprk = g.get_particles()
    up, x0 = g.get_particles_aux()
    return [g.get_state()[0], xp, uprk, up, x0], g.particle_stats()


def _same(a, b):
    (ra, sa), (rb, sb) = a, b
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)
    assert sa[0] == sb[0]
    for k in sa[1]:
        assert sa[1][k] == sb[1][k] or (np.isnan(sa[1][k]) and np.isnan(sb[1][k])), k


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
@pytest.mark.parametrize("case", ["iso", "bfrot", "rrot"])
def test_caller_driven_tracers_equal_in_stage(case, prec):
    """6 steps (release at step 2, two remeshes in the shear cases), tracer
    statistics on, initial positions kept: odd steps p_ControlRK3, even steps
    the frame's p_RK3SolveParticle*, each followed by p_ControlFinishTimeStep."""
    note(f"caller-driven tracers vs in-stage: {case} precision={prec}")
    nsteps = 6
    g = _make(case, prec)
    t = g.run(1, nsteps, 0.0)
    want = _result(g)
    g.close()

    g = _make(case, prec)
    g.set_tracers_in_stage(False)
    tt, dt = 0.0, 99999.0
    for s in range(1, nsteps + 1):
        t0 = tt
        for st in (1, 2, 3):
            dt = g.stage(st, s, tt, dt)
            if s >= 2:  # P_TS_RELEASE = 2: f_RK3RealSpace moves them from then on
                if s % 2:
                    g.control_rk3(st, s, dt)
                else:
                    g.solve_particle(FORM[case], 3000, st, dt)
                g.finish_time_step(s)
            tt = t0 + CSTAGE[st - 1] * dt
            g.update_b(s, tt)
    assert tt == t
    got = _result(g)
    g.close()
    _same(got, want)


def test_tracer_boundary_refusals():
    note("tracer boundary: refused calls")
    g = _make("bfrot", pa.PDR_FP64, release=0)
    with pytest.raises(pa.PdrError) as e:       # the stage advances them
        g.control_rk3(1, 1, 0.01)
    assert e.value.code == -6
    g.set_tracers_in_stage(False)
    with pytest.raises(pa.PdrError) as e:       # no tracer field yet
        g.control_rk3(1, 1, 0.01)
    assert e.value.code == -6
    dt = g.stage(1, 1, 0.0, 99999.0)
    with pytest.raises(pa.PdrError) as e:       # field of stage 1, not 2
        g.control_rk3(2, 1, dt)
    assert e.value.code == -6
    with pytest.raises(pa.PdrError) as e:       # field of tstep 1
        g.control_rk3(1, 2, dt)
    assert e.value.code == -6
    with pytest.raises(pa.PdrError) as e:       # BF_ROT run: p_RK3SolveParticleBF
        g.solve_particle(0, 3000, 1, dt)
    assert e.value.code == -1
    with pytest.raises(pa.PdrError) as e:       # every tracer has one owner
        g.solve_particle(1, 2999, 1, dt)
    assert e.value.code == -1
    g.solve_particle(1, 3000, 1, dt)
    g.finish_time_step(1)
    xp, _ = g.get_particles()
    dx = TWOPI / 32
    assert xp.min() >= -0.5 * dx and xp.max() <= TWOPI - 0.5 * dx
    g.close()
