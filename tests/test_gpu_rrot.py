"""GPU parity of the Rogallo shear forms.  NS_R_ROT (rotational form: sheared-direction
vorticity, the isotropic omega x u of the unmapped velocity, f_RK3FormFRRot and
f_RK3ContinuityFixRRot, src/pf_rk3.f90:416-421, 1088-1089, 2680-2856, 2945-3249,
3620-3693), its tracers (p_RK3SolveParticleR / p_RK3AbsoluteVelocityR,
src/pp_rk3.f90:307-403, 720-773) and statistics (f_FluidstatsFsPrimaryQuantsRRot,
src/pf_fluidstats.f90:1801-1970); NS_R_CONV (convective form: nine raw-wavenumber
gradients, NL_i = sum bmat(k,j) du_i/dx_k u_j, src/pf_rk3.f90:422-427, 1717-1774,
then the same FormF).  Through the C-ABI against the FP64 oracle on the same inputs.  Criteria as for BF_ROT: FP64 norm-wise <= 1e-12 after 10 steps, the
frame and the remesh bit-exact, FP32-mixed E(k) within 1e-4 of FP64."""
import sys

import numpy as np
import pytest

import pandora_amd as pa
from conftest import rel

pytestmark = pytest.mark.gpu

TWOPI = 2.0 * float(np.float32(np.pi))
SHEAR = dict(f_type=2, ns_deformed=3, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _pair(oracle_mod, n, precision=pa.PDR_FP64, **kw):
    return oracle_mod.Oracle(n, **kw), pa.Pandora(n, precision=precision, **kw)


def _run(o, g, steps):
    to = tg = 0.0
    for s in range(1, steps + 1):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        bo, bio = o.get_frame()
        bg, big = g.get_frame()
        assert np.array_equal(bo, bg) and np.array_equal(bio, big), s
    assert to == tg


def test_rrot_pass1_vorticity_matches_oracle(oracle_mod):
    note("R_ROT pass 1 (sheared-direction vorticity + inverse z FFT) N=32, sheared frame")
    n = 32
    o, g = _pair(oracle_mod, n, **SHEAR)
    u = pa.init_field_host(n, 21, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    b, bi = o.get_frame()
    b[6], bi[6] = -0.3, 0.3
    o.set_frame(b, bi)
    g.set_frame(b, bi)
    o.prepare_realspace()
    for f in range(6):
        assert rel(g.debug_prepare(f), o.get_1r2w(f)) < 1e-13, f
    g.close()


@pytest.mark.parametrize("n", [32, 48])
def test_rrot_step_parity_fp64_with_remesh(oracle_mod, n):
    note(f"R_ROT shear DNS FP64 N={n}, 10 steps with remesh at 4 and 10")
    o, g = _pair(oracle_mod, n, **SHEAR)
    u = pa.init_field_host(n, 500 + n, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    _run(o, g, 10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()


@pytest.mark.parametrize("f_type", [3, 4])
def test_rrot_rdt_parity_fp64(oracle_mod, f_type):
    note(f"R_ROT RDT (F_TYPE={f_type}) FP64 N=32, 8 steps")
    kw = dict(SHEAR, f_type=f_type)
    o, g = _pair(oracle_mod, 32, **kw)
    u = pa.init_field_host(32, 40 + f_type, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    _run(o, g, 6)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()


def test_rrot_cfl_time_step_parity_fp64(oracle_mod):
    """RK3_TIMESTEP = CFL in the R_ROT form: dt from the laboratory-frame ucfl of
    stage 1 (binv u, as f_FluidstatsRs2D maps every sheared form); the frame runs
    on simtime (no remesh clock)."""
    note("R_ROT CFL time step FP64 N=32, 6 steps")
    kw = dict(f_type=2, ns_deformed=3, rk3_timestep=3, rk3_cfl=0.5, deform_s=2.0, nu=0.005)
    o, g = _pair(oracle_mod, 32, **kw)
    u = pa.init_field_host(32, 44, 2.0 / 3.0, 5.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 7):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        assert abs(g.dt - o.dt) < 1e-14 * o.dt, s
    assert abs(tg - to) < 1e-13
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()


@pytest.mark.parametrize("shape", [(64, 32, 48), (32, 32, 64)])
def test_rrot_non_cubic_box_parity(oracle_mod, shape):
    note(f"R_ROT non-cubic box {shape} FP64, 6 steps with a remesh")
    nx, ny, nz = shape
    o = oracle_mod.Oracle(nx=nx, ny=ny, nz=nz, **SHEAR)
    g = pa.Pandora(nx=nx, ny=ny, nz=nz, **SHEAR)
    rng = np.random.default_rng(31)
    u = [0.05 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)) for _ in range(3)]
    for c in range(3):  # real field: Hermitian i = 0 plane, zero mean
        p = u[c][0]
        u[c][0] = 0.5 * (p + np.conj(np.roll(p[::-1, ::-1], 1, axis=(0, 1))))
        u[c][0, 0, 0] = 0.0
    o.set_state(*u)
    g.set_state(*u)
    _run(o, g, 6)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()


def test_rrot_generic_kernel_path_parity(oracle_mod, monkeypatch):
    note("R_ROT generic kernels FP64 N=48, 6 steps")
    monkeypatch.setenv("PDR_GENERIC_KERNELS", "1")
    o, g = _pair(oracle_mod, 48, **dict(SHEAR, remesh_t=2))
    u = pa.init_field_host(48, 303, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    _run(o, g, 6)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()


def test_rrot_remesh_is_bit_exact(oracle_mod):
    """R forms move the modes without u1 += u3 (src/pf_rk3.f90:4155-4162)."""
    note("R_ROT remesh bit-exactness N=48")
    n, R = 48, 2
    o, g = _pair(oracle_mod, n, **dict(SHEAR, deform_s=1.0, remesh_t=R))
    rng = np.random.default_rng(5)
    u = [rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape) for _ in range(3)]
    o.set_state(*u)
    g.set_state(*u)
    assert o.remesh(R + 1) and g.remesh(R + 1)
    for a, b in zip(g.get_state(), o.get_state()):
        assert np.array_equal(a, b)
    g.close()


def test_rrot_fp32_mixed_spectrum_against_fp64(oracle_mod):
    note("R_ROT FP32-mixed N=64 shear, 10 steps (remesh at 4 and 10)")
    n = 64
    o, g = _pair(oracle_mod, n, precision=pa.PDR_FP32_MIXED, **SHEAR)
    u = pa.init_field_host(n, 78, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    _run(o, g, 10)
    eo, _, _, tot_o = o.spectrum()
    eg, _, _, tot_g = g.spectrum()
    live = eo > 1e-6 * eo.max()
    assert np.abs(eg[live] - eo[live]).max() / eo[live].max() < 1e-4
    assert abs(tot_g[0] - tot_o[0]) / tot_o[0] < 1e-5
    g.close()


def test_rrot_tracers_and_statistics_fp64(oracle_mod):
    """Tracer positions, particle_up, xp0 and the tracer record (vt = binv
    particle_up) over two remeshes, then the wave-space, real-space and 1-D
    statistics of the resulting state."""
    note("R_ROT tracers + tracer statistics FP64 N=32, 2000 tracers, 12 steps (remesh at 4 and 10)")
    n = 32
    o, g = _pair(oracle_mod, n, **SHEAR)
    u = pa.init_field_host(n, 7, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    dx = TWOPI / n
    xp = np.random.default_rng(8).uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(2000, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    o.particle_stats_config(True, keep_initial=True, ts_release=1)
    g.particle_stats_config(True, keep_initial=True, ts_release=1)
    to = tg = 0.0
    for s in range(1, 13):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        (t_o, ro), (t_g, rg) = o.particle_stats(), g.particle_stats()
        assert t_g == t_o
        if t_o >= 0:
            for k, v in ro.items():
                assert rg[k] == pytest.approx(v, rel=1e-9, abs=1e-13), (s, k)
    xo, _ = o.get_particles()
    xg, _ = g.get_particles()
    d = np.abs(xg - xo)
    assert np.minimum(d, TWOPI - d).max() < 1e-11
    upo, x0o = o.get_particles_aux()
    upg, x0g = g.get_particles_aux()
    assert np.abs(upg - upo).max() < 1e-11
    d0 = np.abs(x0g - x0o)
    assert np.minimum(d0, TWOPI - d0).max() < 1e-11
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    # statistics of the same state and frame
    o.set_state(*g.get_state())
    o.set_frame(*g.get_frame())
    sg, qg = g.stats()
    so, qo = o.stats()
    for k, v in qo.items():
        assert qg[k] == pytest.approx(v, rel=1e-12, abs=1e-300), k
    assert np.array_equal(sg["dens_mod"], so["dens_mod"])
    for k in ("ken", "ken1", "ken2", "ken3", "dis", "disskew"):
        assert rel(sg[k], so[k]) < 1e-12, k
    s1, s2, h = g.spectra_1d()
    o1, o2 = o.spectra_1d()
    np.testing.assert_allclose(s1, o1, rtol=1e-12)
    np.testing.assert_allclose(s2, o2, rtol=1e-12)
    for k, v in o.homogeneous().items():
        assert h[k] == pytest.approx(v, rel=1e-11), k
    rg_, ro_ = g.rs_stats(), o.rs_stats()
    for k, v in ro_.items():
        assert rg_[k] == pytest.approx(v, rel=1e-11, abs=1e-14), k
    g.close()


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_rrot_chunked_plane_pass_with_tracers(oracle_mod, monkeypatch, prec):
    """R_ROT through the chunked y/x passes (PDR_ZCHUNK_MB = 0.75: 5 planes per
    chunk at 48^3 FP64, the last partial) with tracers and a remesh."""
    note(f"R_ROT chunked plane pass N=48 prec={prec}, tracers, 6 steps")
    monkeypatch.setenv("PDR_ZCHUNK_MB", "0.75")
    n = 48
    o, g = _pair(oracle_mod, n, precision=prec, **dict(SHEAR, remesh_t=2))
    u = pa.init_field_host(n, 78, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    dx = TWOPI / n
    xp = np.random.default_rng(9).uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(500, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    _run(o, g, 6)
    tol = 1e-12 if prec == pa.PDR_FP64 else 1e-5
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < tol
    d = np.abs(g.get_particles()[0] - o.get_particles()[0])
    assert np.minimum(d, TWOPI - d).max() < (1e-11 if prec == pa.PDR_FP64 else 1e-5)
    g.close()


# ---- NS_R_CONV: nine gradients through three inverse sweeps, generic kernels ----

CONV = dict(SHEAR, ns_deformed=2)


def test_rconv_pass1_gradients_match_oracle(oracle_mod):
    note("R_CONV pass 1 (u and du/dx + inverse z FFT) N=32")
    n = 32
    o, g = _pair(oracle_mod, n, **CONV)
    u = pa.init_field_host(n, 22, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    o.prepare_realspace()
    for f in range(6):
        assert rel(g.debug_prepare(f), o.get_1r2w(f)) < 1e-13, f
    g.close()


@pytest.mark.parametrize("n", [32, 48])
def test_rconv_step_parity_fp64_with_remesh(oracle_mod, n):
    """N = 32 is alias-free (the oracle's R_CONV equals its R_ROT there); N = 48 =
    3K aliases, where the convective form takes its own path."""
    note(f"R_CONV shear DNS FP64 N={n}, 10 steps with remesh at 4 and 10")
    o, g = _pair(oracle_mod, n, **CONV)
    u = pa.init_field_host(n, 600 + n, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    _run(o, g, 10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()


def test_rconv_fp32_mixed_spectrum_against_fp64(oracle_mod):
    note("R_CONV FP32-mixed N=64 shear, 10 steps")
    n = 64
    o, g = _pair(oracle_mod, n, precision=pa.PDR_FP32_MIXED, **CONV)
    u = pa.init_field_host(n, 79, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    _run(o, g, 10)
    eo, _, _, tot_o = o.spectrum()
    eg, _, _, tot_g = g.spectrum()
    live = eo > 1e-6 * eo.max()
    assert np.abs(eg[live] - eo[live]).max() / eo[live].max() < 1e-4
    assert abs(tot_g[0] - tot_o[0]) / tot_o[0] < 1e-5
    g.close()


def test_rconv_tracers_and_statistics_fp64(oracle_mod):
    note("R_CONV tracers + tracer statistics FP64 N=32, 1000 tracers, 6 steps, then statistics")
    n = 32
    o, g = _pair(oracle_mod, n, **CONV)
    u = pa.init_field_host(n, 9, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    dx = TWOPI / n
    xp = np.random.default_rng(10).uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(1000, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    o.particle_stats_config(True, keep_initial=True)
    g.particle_stats_config(True, keep_initial=True)
    to = tg = 0.0
    for s in range(1, 7):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        (t_o, ro), (t_g, rg) = o.particle_stats(), g.particle_stats()
        assert t_g == t_o
        if t_o >= 0:
            for k, v in ro.items():
                assert rg[k] == pytest.approx(v, rel=1e-9, abs=1e-13), (s, k)
    xo, _ = o.get_particles()
    xg, _ = g.get_particles()
    d = np.abs(xg - xo)
    assert np.minimum(d, TWOPI - d).max() < 1e-11
    o.set_state(*g.get_state())
    o.set_frame(*g.get_frame())
    sg, qg = g.stats()
    so, qo = o.stats()
    for k, v in qo.items():
        assert qg[k] == pytest.approx(v, rel=1e-12, abs=1e-300), k
    rg_, ro_ = g.rs_stats(), o.rs_stats()
    for k, v in ro_.items():
        assert rg_[k] == pytest.approx(v, rel=1e-11, abs=1e-14), k
    g.close()


def test_rconv_rejects_an_x_line_beyond_the_lds():
    """12 (NX + 1) complex elements of LDS per x line: an FP64 1536-point line
    (295 KB) does not fit, and create says so before allocating anything."""
    note("R_CONV FP64 NX=1536 rejected for LDS")
    with pytest.raises(pa.PdrError):
        pa.Pandora(nx=1536, ny=16, nz=16, **CONV)


def test_unknown_form_is_rejected():
    note("NS_DEFORMED = 4 rejected")
    with pytest.raises(pa.PdrError):
        pa.Pandora(32, f_type=2, ns_deformed=4, rk3_timestep=1, deform_s=1.0, remesh_t=5)
