"""GPU parity of the deformed isotropic form (F_TYPE = 5 over BF_ROT, R_ROT and
R_CONV, boxes stretched by B110 / B220: src/pf_rk3.f90:2567-2648, 2751-2828 with
f_RK3BFRotForcing 3256-3349 and its R_ROT twin, the DECIDE step 3905-3919, the frame of src/pg_domain.f90:177-216,
the spectrum bins of src/pf_fluidstats.f90:348-374), and the shear forms on a
stretched box (the remesh clock of B110 F_REMESH_T, src/pf_rk3.f90:3958-4178,
src/pg_rk3.f90:280-385), against the oracle, whose deformed and stretched forms
are pinned by tests/test_deformed.py.  FP64 <= 1e-12 norm-wise
(tracers <= 1e-11), FP32-mixed <= 1e-5 (tracers <= 1e-6); statistics <= 1e-11."""
import sys

import numpy as np
import pytest

import pandora_amd as pa
from conftest import rel

pytestmark = pytest.mark.gpu

TWOPI = 2.0 * float(np.float32(np.pi))
FORCE = dict(force_stop=100, force_kf=2.5, force_sigma=0.5, force_tf=0.2, force_seed=7)


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _field(nx, ny, nz, b, seed):
    """An isotropic field of the unstretched box, laid on the stretched one
    (state = bmat u: each component divided by its stretch, indices multiplied)."""
    m = (nx // b[0], ny // b[1], nz // b[2])
    assert m[0] == m[1] == m[2]
    u = pa.init_field_host(m[0], seed, 2.0 / 3.0, 4.0)
    shape = pa.wave_shape(nx, ny, nz)
    nkj0, nkk0 = u[0].shape[1], u[0].shape[2]
    kj0 = np.arange(nkj0)
    kj0 = np.where(kj0 <= nkj0 // 2, kj0, kj0 - nkj0)
    kk0 = np.arange(nkk0)
    kk0 = np.where(kk0 <= nkk0 // 2, kk0, kk0 - nkk0)
    jj = (b[1] * kj0) % shape[1]
    kk = (b[2] * kk0) % shape[2]
    out = []
    for c in range(3):
        a = np.zeros(shape, np.complex128)
        a[np.ix_(b[0] * np.arange(u[c].shape[0]), jj, kk)] = u[c] / b[c]
        out.append(a)
    return out


def _tracers(count, seed, nx):
    return -0.5 * TWOPI / nx + np.random.default_rng(seed).uniform(0.0, TWOPI, size=(count, 3))


CASES = [
    # (nx, ny, nz), (B110, B220, B330), NS_DEFORMED: the runtime-radix kernels
    # (short non-cubic lines) and the length-specialised ones, an x and a y
    # stretch; BF_ROT, and the R forms (laboratory velocities)
    ((48, 24, 24), (2, 1, 1), 1),
    ((24, 48, 24), (1, 2, 1), 1),
    ((256, 128, 128), (2, 1, 1), 1),
    ((48, 24, 24), (2, 1, 1), 3),
    ((48, 24, 24), (2, 1, 1), 2),
]


@pytest.mark.parametrize("shape,b,ns", CASES, ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else f"ns{v}")
def test_deformed_forced_with_tracers_against_oracle(oracle_mod, shape, b, ns):
    """Forcing, 2000 tracers, FIXED dt, 5 steps (3 at 256x128x128); then the
    statistics of the final state."""
    nx, ny, nz = shape
    steps = 3 if nx >= 256 else 5
    note(f"deformed isotropic {shape} B={b} NS_DEFORMED={ns}: forcing + tracers, {steps} steps, "
         "oracle vs FP64 and FP32-mixed")
    kw = dict(f_type=5, ns_deformed=ns, rk3_timestep=2, rk3_dt=0.004, nu=0.01, b110=b[0], b220=b[1], b330=b[2])
    kw.update(FORCE)
    u = _field(nx, ny, nz, b, 40 + nx)
    if ns != 1:  # the R forms store the laboratory velocity
        u = [a * bc for a, bc in zip(u, b)]
    xp = _tracers(2000, nx + ny, nx)
    o = oracle_mod.Oracle(nx=nx, ny=ny, nz=nz, **kw)
    o.set_state(*u)
    o.set_particles(xp)
    to = 0.0
    for s in range(1, steps + 1):
        to = o.step(s, to)
    so = o.get_state()
    xo, _ = o.get_particles()
    spec_o, q_o = o.stats()
    rs_o = o.rs_stats()
    r1_o, li_o = o.rs_correlation()
    o.close()
    for prec in (pa.PDR_FP64, pa.PDR_FP32_MIXED):
        g = pa.Pandora(nx=nx, ny=ny, nz=nz, precision=prec, **kw)
        note(f"  precision={prec} forms={g.kernel_forms()} kkmax={g.kkmax}")
        g.set_state(*u)
        g.set_particles(xp)
        tg = 0.0
        for s in range(1, steps + 1):
            tg = g.rk3_calc(s, tg)
        assert tg == to
        sg = g.get_state()
        xg, _ = g.get_particles()
        d = np.abs(xg - xo)
        d = np.minimum(d, TWOPI - d)
        if prec == pa.PDR_FP64:
            for a, c in zip(sg, so):
                assert rel(a, c) < 1e-12
            assert d.max() < 1e-11
            spec_g, q_g = g.stats()
            for k in spec_o:
                assert rel(spec_g[k], spec_o[k]) < 1e-11, k
            for k in ("tke", "f_ken", "f_dis", "f_disskew", "fn_ken", "rn_tay"):
                assert q_g[k] == pytest.approx(q_o[k], rel=1e-11), k
            rs_g = g.rs_stats()
            for k in rs_o:
                assert rs_g[k] == pytest.approx(rs_o[k], rel=1e-11, abs=1e-14), k
            # lint11 carries dx_laboratoryframe = TWOPI B110 / NX
            r1_g, li_g = g.rs_correlation()
            assert np.abs(r1_g - r1_o).max() < 1e-12
            assert li_g == pytest.approx(li_o, rel=1e-11, abs=1e-13)
        else:
            for a, c in zip(sg, so):
                assert rel(a, c) < 1e-5
            assert d.max() < 1e-6
        g.close()


def test_deformed_decide_step_against_oracle(oracle_mod):
    """RK3_TIMESTEP = DECIDE: the CFL step on min(DX, DY, DZ) capped at FORCE_TF
    (the cap active in one run, inactive in the other), 3 steps."""
    nx, ny, nz = 48, 24, 24
    for tf in (100.0, 0.004):
        note(f"deformed isotropic DECIDE {(nx, ny, nz)} FORCE_TF={tf}")
        kw = dict(f_type=5, ns_deformed=1, rk3_timestep=1, rk3_cfl=0.5, nu=0.01, b110=2, deform_s=1.0,
                  remesh_t=1000, force_tf=tf)
        u = _field(nx, ny, nz, (2, 1, 1), 77)
        o = oracle_mod.Oracle(nx=nx, ny=ny, nz=nz, **kw)
        g = pa.Pandora(nx=nx, ny=ny, nz=nz, **kw)
        o.set_state(*u)
        g.set_state(*u)
        to = tg = 0.0
        for s in range(1, 4):
            to = o.step(s, to)
            tg = g.rk3_calc(s, tg)
            assert abs(tg - to) <= 1e-14 * to
        if tf < 1.0:
            assert to == pytest.approx(3 * tf, rel=1e-15)
        for a, c in zip(g.get_state(), o.get_state()):
            assert rel(a, c) < 1e-12
        g.close()
        o.close()


@pytest.mark.parametrize("shape,ns", [((48, 24, 24), 1), ((48, 24, 24), 3), ((256, 128, 128), 1)],
                         ids=["48x24x24-BF_ROT", "48x24x24-R_ROT", "256x128x128-BF_ROT"])
def test_stretched_shear_with_remesh_against_oracle(oracle_mod, shape, ns):
    """Shear DNS on a box stretched by B110 = 2 (S = 10, DECIDE with F_REMESH_T = 3,
    so dt = 1/60 and the remesh comes at step B110 F_REMESH_T + 1 = 7), 2000
    tracers, 8 steps: FP64 <= 1e-12 and the frame bit for bit, FP32-mixed <= 1e-5."""
    nx, ny, nz = shape
    note(f"stretched shear {shape} B110=2 NS_DEFORMED={ns}: remesh at step 7, 8 steps")
    kw = dict(f_type=2, ns_deformed=ns, rk3_timestep=1, deform_s=10.0, remesh_t=3, nu=0.005, b110=2)
    u = _field(nx, ny, nz, (2, 1, 1), 90 + nx)
    if ns != 1:
        u = [u[0] * 2.0, u[1], u[2]]
    xp = _tracers(2000, 5 + nx, nx)
    o = oracle_mod.Oracle(nx=nx, ny=ny, nz=nz, **kw)
    o.set_state(*u)
    o.set_particles(xp)
    to = 0.0
    for s in range(1, 9):
        to = o.step(s, to)
    assert all(np.isfinite(a).all() for a in o.get_state())
    so, fo = o.get_state(), o.get_frame()
    xo, _ = o.get_particles()
    o.close()
    for prec in (pa.PDR_FP64, pa.PDR_FP32_MIXED):
        g = pa.Pandora(nx=nx, ny=ny, nz=nz, precision=prec, **kw)
        g.set_state(*u)
        g.set_particles(xp)
        tg = 0.0
        for s in range(1, 9):
            tg = g.rk3_calc(s, tg)
        assert tg == to
        fg = g.get_frame()
        assert np.array_equal(fg[0], fo[0]) and np.array_equal(fg[1], fo[1])
        xg, _ = g.get_particles()
        d = np.abs(xg - xo)
        d = np.minimum(d, TWOPI - d)
        tol, ttol = (1e-12, 1e-11) if prec == pa.PDR_FP64 else (1e-5, 1e-6)
        for a, c in zip(g.get_state(), so):
            assert rel(a, c) < tol
        assert d.max() < ttol
        g.close()
