"""Corners of the reference's parameter space against the FP64 oracle, through
the C-ABI: the DECIDE time step of isotropic runs (the CFL step capped at
FORCE_TF, 0.02 when the velocity is zero; src/pf_rk3.f90:3893-3907), the
rapid-distortion forms of the BF_ROT shear (F_TYPE 3 viscous, 4 inviscid), and
every tracer interpolation order P_INTERP_ORDER = 1..8 (odd orders centre the
molecule on the nearest node, src/pp_interp.f90:426-437)."""
import sys

import numpy as np
import pytest

import pandora_amd as pa
from conftest import rel

pytestmark = pytest.mark.gpu

TWOPI = 2.0 * float(np.float32(np.pi))


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _wrapdiff(a, b):
    d = np.abs(a - b)
    return np.minimum(d, TWOPI - d)


@pytest.mark.parametrize("force_tf,capped", [(0.004, True), (10.0, False)])
def test_decide_isotropic_timestep(oracle_mod, force_tf, capped):
    """RK3_TIMESTEP = DECIDE, F_TYPE = 1: dt = RK3_CFL dx / ucfl, capped at
    FORCE_TF.  With FORCE_TF = 0.004 the cap is active on every step (the CFL
    step is about 0.02); with 10 it never is.  10 steps: dt and simtime equal
    the oracle's, the state to 1e-12."""
    n = 32
    note(f"DECIDE isotropic N={n} FORCE_TF={force_tf}, 10 steps")
    kw = dict(f_type=1, rk3_timestep=1, rk3_cfl=0.5, nu=0.01, force_tf=force_tf)
    o = oracle_mod.Oracle(n, **kw)
    g = pa.Pandora(n, **kw)
    u = pa.init_field_host(n, 17, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 11):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        assert abs(g.dt - o.dt) <= 1e-14 * o.dt, s
        if capped:
            assert g.dt == force_tf
        else:
            assert g.dt < force_tf
            assert abs(g.dt - 0.5 * (TWOPI / n) / o.ucfl) <= 1e-14 * g.dt
    assert abs(tg - to) <= 1e-14 * to
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()


def test_decide_isotropic_zero_velocity_fallback(oracle_mod):
    """ucfl = 0 (a zero state): DECIDE and CFL both take dt = 0.02 (the
    reference's default-real literal, 0.0199999995529651...)."""
    note("DECIDE / CFL with a zero state")
    for ts in (1, 3):
        g = pa.Pandora(16, f_type=1, rk3_timestep=ts, rk3_cfl=0.5, force_tf=10.0)
        t = g.rk3_calc(1, 0.0)
        assert g.dt == float(np.float32(0.02))
        assert t == float(np.float32(0.02))
        g.close()


@pytest.mark.parametrize("f_type", [3, 4])
def test_bfrot_rdt_against_oracle(oracle_mod, f_type):
    """Rapid-distortion forms of the BF_ROT shear (F_TYPE 3: viscous, 4: inviscid;
    no nonlinear term, src/pf_rk3.f90:2494-2673): S = 20, F_REMESH_T = 3, 10 steps
    through two remeshes, with tracers, against the oracle."""
    n = 32
    note(f"BF_ROT RDT F_TYPE={f_type} N={n}, 10 steps with remesh + 1000 tracers")
    kw = dict(f_type=f_type, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)
    o = oracle_mod.Oracle(n, **kw)
    g = pa.Pandora(n, **kw)
    u = pa.init_field_host(n, 40 + f_type, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    dx = TWOPI / n
    xp = np.random.default_rng(f_type).uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(1000, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    to = tg = 0.0
    for s in range(1, 11):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        bo, bio = o.get_frame()
        bg, big = g.get_frame()
        assert np.array_equal(bo, bg) and np.array_equal(bio, big), s
    assert to == tg
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    for a, b in zip(g.get_rk(), o.get_rk()):
        assert rel(a, b) < 1e-12
    assert _wrapdiff(g.get_particles()[0], o.get_particles()[0]).max() < 1e-11
    g.close()


@pytest.mark.parametrize("f_type", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3, 5, 6, 7, 8])
def test_interpolation_orders_against_oracle(oracle_mod, order, f_type):
    """Tracers with P_INTERP_ORDER = order (a P x P x P molecule, Neville tableau
    along y, z, x; p_InterpFindMolecule's odd branch for odd P), isotropic and
    BF_ROT shear with a remesh, 6 steps: positions to 1e-11, particle_up_rk3 to
    1e-12 against the oracle; also tracers on the lower box faces and one
    crowded cell."""
    n = 32
    note(f"tracers P_INTERP_ORDER={order} f_type={f_type} N={n}, 6 steps")
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005) if f_type == 2 else \
        dict(nu=0.01, rk3_dt=0.01)
    o = oracle_mod.Oracle(n, interp_order=order, **kw)
    g = pa.Pandora(n, interp_order=order, **kw)
    u = pa.init_field_host(n, 70 + order, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    dx = TWOPI / n
    rng = np.random.default_rng(order)
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(2000, 3))
    for d in range(3):
        xp[50 * d:50 * d + 25, d] = -0.5 * dx
    xp[200:400] = xp[200]
    o.set_particles(xp)
    g.set_particles(xp)
    to = tg = 0.0
    for s in range(1, 7):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    xo, ro = o.get_particles()
    xg, rg = g.get_particles()
    assert _wrapdiff(xg, xo).max() < 1e-11
    assert rel(rg, ro) < 1e-12
    g.close()
