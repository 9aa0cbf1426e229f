"""Pins for the oracle's deformed isotropic form (F_TYPE = 5 over BF_ROT, R_ROT
and R_CONV, box stretches B110/B220/B330; src/pf_rk3.f90:2567-2648, 2751-2828,
3256-3349 and its R_ROT twin, 3905-3919,
src/pg_domain.f90:177-216, src/pf_fluidstats.f90:348-374).

The deformed form is checked against things it shares no branch with:

* with B = (1, 1, 1) it is the isotropic form (forcing included);
* a box stretched by B110 = 2 with NX = 2M points holding a field with even i
  only is the isotropic M^3 box seen twice: its state is bmat times the
  isotropic one (u1 / 2) at i = 2i', zero at odd i; its tracers move as the
  isotropic ones at x1 / 2; its shell sums (E(k) times the mode density) are
  the isotropic ones;
* one Fourier mode in the stretched box decays with the laboratory |k|:
  u(t_n) = u(0) R(-nu |bmat^T k|^2 dt)^n, R the stability polynomial of the
  three-stage third-order RK scheme, 1 + z + z^2/2 + z^3/6;
* the forcing of a stretched box after one stage from rest is the isotropic
  box's at (i/B110, j, k), made solenoidal by the continuity fix along k;
* the DECIDE time step of the stretched box uses min(DX, DY, DZ): half the
  isotropic one for the same field;
* over the R forms (laboratory velocities) the stretched box holds the
  unit-stretch run's state, forcing included, and unit-stretch R_ROT is the
  isotropic form.

CPU only (the oracle is FP64 C)."""
import numpy as np
import pytest

import pandora_amd as pa
from conftest import rel

TWOPI = 2.0 * float(np.float32(np.pi))
M = 24


def _stretched(u, b1=2):
    """The isotropic M^3 state as the state of the (b1 M, M, M) stretched box."""
    out = []
    for c, s in enumerate((1.0 / b1, 1.0, 1.0)):
        a = np.zeros((b1 * M // 3 + 1,) + u[c].shape[1:], np.complex128)
        a[0::b1] = s * u[c]
        out.append(a)
    return out


def test_unit_stretch_is_the_isotropic_form(oracle_mod):
    """B = (1, 1, 1), forcing on: 4 steps equal the isotropic form's."""
    kw = dict(rk3_timestep=2, rk3_dt=0.01, nu=0.01, force_stop=100, force_kf=2.5, force_tf=0.2,
              force_sigma=0.7, force_seed=9)
    u = pa.init_field_host(M, 13, 2.0 / 3.0, 4.0)
    oi = oracle_mod.Oracle(M, **kw)
    od = oracle_mod.Oracle(M, f_type=5, ns_deformed=1, **kw)
    for o in (oi, od):
        o.set_state(*u)
    ti = td = 0.0
    for s in range(1, 5):
        ti, td = oi.step(s, ti), od.step(s, td)
    assert ti == td
    for a, b in zip(od.get_state(), oi.get_state()):
        assert rel(a, b) < 1e-13


def test_stretched_box_is_the_isotropic_box_twice(oracle_mod):
    """B110 = 2, (48, 24, 24) against 24^3: state, statistics and tracers after
    4 steps."""
    kw = dict(rk3_timestep=2, rk3_dt=0.01, nu=0.01)
    u = pa.init_field_host(M, 11, 2.0 / 3.0, 4.0)
    oi = oracle_mod.Oracle(M, **kw)
    od = oracle_mod.Oracle(nx=2 * M, ny=M, nz=M, f_type=5, ns_deformed=1, b110=2, **kw)
    oi.set_state(*u)
    od.set_state(*_stretched(u))
    xp = -0.5 * TWOPI / M + np.random.default_rng(3).uniform(0.0, TWOPI, size=(400, 3))
    oi.set_particles(xp)
    xd = xp.copy()
    xd[:, 0] *= 0.5
    od.set_particles(xd)
    ti = td = 0.0
    for s in range(1, 5):
        ti, td = oi.step(s, ti), od.step(s, td)
    si, sd = oi.get_state(), od.get_state()
    for c, sc in enumerate((0.5, 1.0, 1.0)):
        assert rel(sd[c][0::2], sc * si[c]) < 1e-14
        assert np.abs(sd[c][1::2]).max() < 1e-16 * np.abs(si[c]).max() * 1e3
    xi, _ = oi.get_particles()
    xg, _ = od.get_particles()
    xg[:, 0] *= 2.0
    d = np.abs(xg - xi)
    assert np.minimum(d, TWOPI - d).max() < 1e-13
    (spi, qi), (spd, qd) = oi.stats(), od.stats()
    assert oi.kkmax == od.kkmax
    for k in ("ken", "ken1", "ken2", "ken3", "dis", "disskew"):
        assert rel(np.asarray(spd[k]) * np.asarray(spd["dens_mod"]),
                   np.asarray(spi[k]) * np.asarray(spi["dens_mod"])) < 1e-13, k
    for k in ("tke", "u1u1", "u2u2", "u3u3", "f_ken", "f_dis", "f_disskew"):
        assert qd[k] == pytest.approx(qi[k], rel=1e-13), k


def test_single_mode_decays_with_the_laboratory_wavenumber(oracle_mod):
    """u2 of mode (1, 0, 0) in a box with B110 = 2: |k_lab|^2 = 1/4."""
    nu, dt = 0.05, 0.02
    o = oracle_mod.Oracle(16, f_type=5, ns_deformed=1, b110=2, rk3_timestep=2, rk3_dt=dt, nu=nu)
    u = [np.zeros(o.shape, np.complex128) for _ in range(3)]
    u[1][1, 0, 0] = 0.3 + 0.1j
    o.set_state(*u)
    t = 0.0
    for s in range(1, 11):
        t = o.step(s, t)
    z = -nu * 0.25 * dt
    r = 1.0 + z + z * z / 2.0 + z**3 / 6.0
    got = o.get_state()[1][1, 0, 0]
    assert abs(got - u[1][1, 0, 0] * r**10) < 1e-14 * abs(got)
    # the unstretched rate would leave it a factor R(-nu dt)^10 / R(-nu dt / 4)^10 away
    assert abs(got - u[1][1, 0, 0] * (1 - nu * dt) ** 10) > 1e-4


def test_stretched_forcing_reads_the_isotropic_arrays_at_the_quotient(oracle_mod):
    """One stage from rest with forcing: the stretched box's modes (2i', j, k) hold
    the isotropic box's (i', j, k) after the continuity fix along the stored k,
    odd i stay zero."""
    kw = dict(rk3_timestep=2, rk3_dt=0.01, nu=0.01, force_stop=100, force_kf=2.5, force_tf=0.2,
              force_sigma=0.7, force_seed=5)
    oi = oracle_mod.Oracle(M, **kw)
    od = oracle_mod.Oracle(nx=2 * M, ny=M, nz=M, f_type=5, ns_deformed=1, b110=2, **kw)
    oi.stage(1, 1, 0.0, 0.01)
    od.stage(1, 1, 0.0, 0.01)
    assert all(np.array_equal(a, b) for a, b in zip(oi.get_forcing(), od.get_forcing()))
    si, sd = oi.get_state(), od.get_state()
    assert max(np.abs(a).max() for a in si) > 1e-4  # the forcing acted
    ki = np.arange(od.nki, dtype=np.float64)[0::2][:, None, None]
    kj = np.arange(od.nkj)
    kj = np.where(kj <= od.nkj // 2, kj, kj - od.nkj).astype(np.float64)[None, :, None]
    kk = np.arange(od.nkk)
    kk = np.where(kk <= od.nkk // 2, kk, kk - od.nkk).astype(np.float64)[None, None, :]
    k2 = ki**2 + kj**2 + kk**2
    kd = (ki * si[0] + kj * si[1] + kk * si[2]) / np.where(k2 > 0, k2, 1.0)
    want = [si[0] - ki * kd, si[1] - kj * kd, si[2] - kk * kd]
    for c in range(3):
        assert rel(sd[c][0::2], want[c]) < 1e-14
        assert not np.abs(sd[c][1::2]).any()


def test_decide_step_takes_the_smallest_spacing(oracle_mod):
    """RK3_TIMESTEP = DECIDE: the stretched (48, 24, 24) box's CFL step is half the
    24^3 box's for the same field (DY = 2 DX there); FORCE_TF does not cap it."""
    kw = dict(rk3_timestep=1, rk3_cfl=0.5, nu=0.01, force_tf=100.0)
    u = pa.init_field_host(M, 11, 2.0 / 3.0, 4.0)
    oi = oracle_mod.Oracle(M, **kw)
    od = oracle_mod.Oracle(nx=2 * M, ny=M, nz=M, f_type=5, ns_deformed=1, b110=2, deform_s=1.0,
                           remesh_t=1000, **kw)
    oi.set_state(*u)
    od.set_state(*_stretched(u))
    ti, td = oi.step(1, 0.0), od.step(1, 0.0)
    assert 0.0 < td < 0.1 and td * 2.0 == ti


@pytest.mark.parametrize("ns", [3, 2], ids=["R_ROT", "R_CONV"])
def test_rogallo_forms_stretched_twin_with_forcing(oracle_mod, ns):
    """The deformed branch of f_RK3FormFRRot with f_RK3RRotForcing (the R forms
    hold laboratory velocities): the (2M, M, M) box with B110 = 2 holds the
    unit-stretch M^3 run's state at i = 2i' through 4 forced steps; with unit
    stretch the R_ROT form is the isotropic one."""
    kw = dict(rk3_timestep=2, rk3_dt=0.01, nu=0.01, force_stop=100, force_kf=2.5, force_tf=0.2,
              force_sigma=0.7, force_seed=5)
    u = pa.init_field_host(M, 11, 2.0 / 3.0, 4.0)
    ou = oracle_mod.Oracle(M, f_type=5, ns_deformed=ns, **kw)
    od = oracle_mod.Oracle(nx=2 * M, ny=M, nz=M, f_type=5, ns_deformed=ns, b110=2, **kw)
    ou.set_state(*u)
    ud = [np.zeros(od.shape, np.complex128) for _ in range(3)]
    for c in range(3):
        ud[c][0::2] = u[c]
    od.set_state(*ud)
    tu = td = 0.0
    for s in range(1, 5):
        tu, td = ou.step(s, tu), od.step(s, td)
    su, sd = ou.get_state(), od.get_state()
    assert max(np.abs(a).max() for a in su) > 0.01
    for c in range(3):
        assert rel(sd[c][0::2], su[c]) < 1e-14
        assert np.abs(sd[c][1::2]).max() < 1e-13 * np.abs(su[c]).max()
    if ns == 3:
        oi = oracle_mod.Oracle(M, f_type=1, **kw)
        oi.set_state(*u)
        t = 0.0
        for s in range(1, 5):
            t = oi.step(s, t)
        for a, b in zip(su, oi.get_state()):
            assert rel(a, b) < 1e-13


# ---- stretched shear boxes (B110 > 1 with the shear forms) -------------------

def _shear_pair(oracle_mod, ns, T):
    kw = dict(f_type=2, ns_deformed=ns, rk3_timestep=1, deform_s=1.0, remesh_t=T, nu=0.01)
    u = pa.init_field_host(M, 11, 2.0 / 3.0, 4.0)
    ou = oracle_mod.Oracle(M, **kw)
    os_ = oracle_mod.Oracle(nx=2 * M, ny=M, nz=M, b110=2, **kw)
    ou.set_state(*u)
    scale = (0.5, 1.0, 1.0) if ns == 1 else (1.0, 1.0, 1.0)  # R forms: laboratory velocities
    us = [np.zeros(os_.shape, np.complex128) for _ in range(3)]
    for c in range(3):
        us[c][0::2] = scale[c] * u[c]
    os_.set_state(*us)
    return ou, os_, scale


def test_stretched_shear_box_is_its_twin_until_the_twin_remeshes(oracle_mod):
    """BF_ROT shear DNS with DECIDE (S = 1, F_REMESH_T = 5, so the unstretched box
    remeshes at step 6): the (48, 24, 24) box with B110 = 2 holds the 24^3 run's
    state at i = 2i' for the first 5 steps; the same lab shear, bmat(1,3) =
    -S t bmat(1,1).  (The R_ROT form is no such twin under shear: its state
    leaves the solenoidal set within a stage and its continuity fix projects
    along the stored k, which the stretch changes.)"""
    T = 5
    ou, os_, scale = _shear_pair(oracle_mod, 1, T)
    tu = ts = 0.0
    for s in range(1, T + 1):
        tu, ts = ou.step(s, tu), os_.step(s, ts)
    assert tu == ts
    su, ss = ou.get_state(), os_.get_state()
    for c in range(3):
        assert rel(ss[c][0::2], scale[c] * su[c]) < 1e-14
    bu, bs = ou.get_frame()[0], os_.get_frame()[0]
    assert bs[6] == pytest.approx(0.5 * bu[6], rel=1e-15)  # bmat(1,3), column-major


def test_stretched_shear_remesh_keeps_the_frame_within_half_a_unit(oracle_mod):
    """B110 = 2, F_REMESH_T = 5, 22 steps: the frame's shear entry bmat(1,3) =
    -S t bmat(1,1) stays within [-1/2, 1/2]; each remesh raises it by exactly 1
    (so the stored k3 move by whole multiples of i and keep their laboratory
    wavenumbers), and a remesh happens exactly when the entry would leave the
    interval (strain B110 / 2 since the last one)."""
    T = 5
    _, o, _ = _shear_pair(oracle_mod, 1, T)
    t = 0.0
    prev = o.get_frame()[0][6]
    jumps = []
    for s in range(1, 4 * T + 3):
        t = o.step(s, t)
        b = o.get_frame()[0][6]
        assert -0.5 - 1e-12 <= b <= 0.5 + 1e-12
        d = b - prev
        # 0.05 = S dt bmat(1,1) with dt = (1/2S)/F_REMESH_T
        if abs(d + 0.05) > 1e-12:
            assert d == pytest.approx(1.0 - 0.05, abs=1e-12)
            assert prev == pytest.approx(-0.5, abs=1e-12)
            jumps.append(s)
        prev = b
    assert jumps == [2 * T + 1]


def test_stretched_shear_remesh_keeps_the_laboratory_energy(oracle_mod):
    """Across the stretched box's remesh the laboratory-frame energy changes only
    by that of the modes the re-indexing drops (found by enumerating the map)."""
    T = 3
    _, o, _ = _shear_pair(oracle_mod, 1, T)
    t = 0.0
    for s in range(1, 2 * T + 1):
        t = o.step(s, t)
    c = oracle_mod.Oracle(nx=2 * M, ny=M, nz=M, b110=2, f_type=2, ns_deformed=1, rk3_timestep=1,
                          deform_s=1.0, remesh_t=T, nu=0.01)
    c.set_state(*o.get_state())
    c.set_frame(*o.get_frame())
    u = c.get_state()
    kmax = c.nkk // 2
    lost = np.zeros(c.shape, bool)
    for i in range(c.nki):
        for k in range(c.nkk):
            if oracle_mod.remesh_target(kmax, i, k) < 0:
                lost[i, :, k] = True
    b = np.asarray(c.get_frame()[1]).reshape(3, 3, order="F")  # binv
    ul = [sum(b[m, n] * u[n] for n in range(3)) for m in range(3)]
    w = np.where(np.arange(c.nki) > 0, 2.0, 1.0)[:, None, None]
    e_lost = 0.5 * sum((w * np.abs(a) ** 2)[lost].sum() for a in ul)
    e0 = c.stats()[1]["tke"]
    assert c.remesh(2 * T + 1)
    e1 = c.stats()[1]["tke"]
    assert e_lost > 1e-8 * e0  # the map drops something
    assert e1 == pytest.approx(e0 - e_lost, rel=1e-12)
