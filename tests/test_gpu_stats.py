"""GPU parity of the device statistics (f_FluidstatsFs, src/pf_fluidstats.f90:
1293-1361) against the oracle on the same state: isotropic and sheared
frames, FP64 and FP32-mixed storage.  Scalars and spectra FP64 within 1e-12
relative (different summation order); mode densities exact."""
import sys

import numpy as np
import pytest

import pandora_amd as pa
from conftest import rel

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _compare(g, o):
    sg, qg = g.stats()
    so, qo = o.stats()
    for name, v in qo.items():
        assert qg[name] == pytest.approx(v, rel=1e-12, abs=1e-300), name
    assert np.array_equal(sg["dens_mod"], so["dens_mod"])
    for name in ("ken", "ken1", "ken2", "ken3", "dis", "disskew"):
        assert rel(sg[name], so[name]) < 1e-12, name
    ek, dk, dsk, tot = g.spectrum()
    eo, do, dso, to = o.spectrum()
    assert rel(ek, eo) < 1e-12 and rel(dk, do) < 1e-12 and rel(dsk, dso) < 1e-12
    assert rel(tot, to) < 1e-12


@pytest.mark.parametrize("n", [32, 64])
def test_stats_isotropic_fp64(oracle_mod, n):
    note(f"device statistics, isotropic FP64 N={n}")
    g = pa.Pandora(n, nu=0.01)
    o = oracle_mod.Oracle(n, nu=0.01)
    u = pa.init_field_host(n, 40 + n, 2.0 / 3.0, 4.0)
    g.set_state(*u)
    o.set_state(*u)
    _compare(g, o)


def test_stats_sheared_frame_after_steps(oracle_mod):
    note("device statistics, shear N=48 after 4 steps (sheared frame, remesh_t=3)")
    n = 48
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)
    g = pa.Pandora(n, **kw)
    o = oracle_mod.Oracle(n, **kw)
    u = pa.init_field_host(n, 12, 2.0 / 3.0, 6.0)
    g.set_state(*u)
    o.set_state(*u)
    tg = to = 0.0
    for s in range(1, 5):
        tg = g.rk3_calc(s, tg)
        to = o.step(s, to)
    b, _ = g.get_frame()
    assert b[6] != 0.0
    o.set_state(*g.get_state())  # same state: the comparison isolates the statistics
    o.set_frame(*g.get_frame())
    _compare(g, o)


def test_stats_fp32_state(oracle_mod):
    note("device statistics from an FP32-mixed state N=64")
    n = 64
    g = pa.Pandora(n, nu=0.01, precision=pa.PDR_FP32_MIXED)
    o = oracle_mod.Oracle(n, nu=0.01)
    g.set_state(*pa.init_field_host(n, 5, 2.0 / 3.0, 4.0))
    o.set_state(*g.get_state())  # the FP32 values, widened
    _compare(g, o)


def test_stats_large_grid_runs():
    note("device statistics at 256^3 FP32-mixed (kkmax 147), finite and Parseval-consistent")
    n = 256
    g = pa.Pandora(n, nu=0.005, precision=pa.PDR_FP32_MIXED)
    g.init_field(3, 2.0 / 3.0, 6.0)
    sp, q = g.stats()
    assert np.isfinite(list(q.values())).all()
    assert q["f_ken"] == pytest.approx(q["tke"], rel=1e-12)
    assert 0.5 < q["tke"] < 1.5


def _compare_rs(g, o):
    qg, qo = g.rs_stats(), o.rs_stats()
    for name, v in qo.items():
        assert qg[name] == pytest.approx(v, rel=1e-11, abs=1e-14), name


@pytest.mark.parametrize("n", [32, 48])
def test_rs_stats_isotropic_fp64(oracle_mod, n):
    note(f"real-space statistics, isotropic FP64 N={n}")
    g = pa.Pandora(n, nu=0.01)
    o = oracle_mod.Oracle(n, nu=0.01)
    u = pa.init_field_host(n, 50 + n, 2.0 / 3.0, 4.0)
    g.set_state(*u)
    o.set_state(*u)
    _compare_rs(g, o)


def test_rs_stats_sheared_chunked(oracle_mod, monkeypatch):
    note("real-space statistics, shear N=48 sheared frame, chunked plane pass, and the generic path")
    n = 48
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)
    u = pa.init_field_host(n, 14, 2.0 / 3.0, 6.0)
    o = oracle_mod.Oracle(n, **kw)
    o.set_state(*u)
    b, bi = o.get_frame()
    b[6], bi[6] = -0.4, 0.4
    o.set_frame(b, bi)
    for env in ({"PDR_ZCHUNK_MB": "0.75"}, {"PDR_GENERIC_KERNELS": "1"}):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        g = pa.Pandora(n, **kw)
        g.set_state(*u)
        g.set_frame(b, bi)
        _compare_rs(g, o)
        g.close()
        for k in env:
            monkeypatch.delenv(k)


def test_rs_stats_fp32_state(oracle_mod):
    note("real-space statistics from an FP32-mixed state N=64")
    n = 64
    g = pa.Pandora(n, nu=0.01, precision=pa.PDR_FP32_MIXED)
    o = oracle_mod.Oracle(n, nu=0.01)
    g.set_state(*pa.init_field_host(n, 6, 2.0 / 3.0, 4.0))
    o.set_state(*g.get_state())
    qg, qo = g.rs_stats(), o.rs_stats()
    for name, v in qo.items():  # FP32 transforms: single-precision agreement
        assert qg[name] == pytest.approx(v, rel=2e-5, abs=1e-9), name


def test_cfl_dt_sheared_frame(oracle_mod):
    note("CFL time step with shear uses the laboratory velocity (N=48)")
    n = 48
    kw = dict(f_type=2, rk3_timestep=3, rk3_cfl=0.5, deform_s=5.0, remesh_t=100, nu=0.005)
    g = pa.Pandora(n, **kw)
    o = oracle_mod.Oracle(n, **kw)
    u = pa.init_field_host(n, 19, 2.0 / 3.0, 6.0)
    g.set_state(*u)
    o.set_state(*u)
    tg = to = 0.0
    for s in range(1, 4):
        tg = g.rk3_calc(s, tg)
        to = o.step(s, to)
        assert g.dt == pytest.approx(o.dt, rel=1e-13)
    assert tg == pytest.approx(to, rel=1e-13)
    assert g.rs_stats()["ucfl"] == pytest.approx(o.rs_stats()["ucfl"], rel=1e-12)


@pytest.mark.parametrize("shear", [False, True])
def test_spectra_1d_and_homogeneous(oracle_mod, shear):
    note(f"streamwise 1-D spectra and homogeneous record, N=48x32x40, shear={shear}")
    kw = dict(f_type=2, rk3_timestep=2, deform_s=2.0, nu=0.005) if shear else dict(nu=0.01)
    g = pa.Pandora(nx=48, ny=32, nz=40, **kw)
    o = oracle_mod.Oracle(nx=48, ny=32, nz=40, **kw)
    rng = np.random.default_rng(17)
    u = [rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape) for _ in range(3)]
    g.set_state(*u)
    o.set_state(*u)
    if shear:
        b, bi = o.get_frame()
        b[6], bi[6] = -0.3, 0.3
        o.set_frame(b, bi)
        g.set_frame(b, bi)
    s1, s2, h = g.spectra_1d()
    o1, o2 = o.spectra_1d()
    np.testing.assert_allclose(s1, o1, rtol=1e-12)
    np.testing.assert_allclose(s2, o2, rtol=1e-12)
    ho = o.homogeneous()
    for k, v in ho.items():
        assert h[k] == pytest.approx(v, rel=1e-11), k
    g.close()
