"""GPU parity of the forced isotropic step: OU forcing state advanced on the
device (f_ForceCalc, src/pf_force.f90:106-268) and applied in FormF
(src/pf_rk3.f90:2421-2467), against the FP64 oracle that draws the same
counter-based deviates.  Criterion as for the unforced step: FP64 <= 1e-12
norm-wise after 10 steps; forcing arrays <= 1e-13."""
import sys

import numpy as np
import pytest

import pandora_amd as pa
from conftest import rel

pytestmark = pytest.mark.gpu

FORCE = dict(force_stop=100, force_kf=2.5, force_sigma=0.5, force_tf=0.2, force_seed=3)


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _pair(oracle_mod, n, precision=pa.PDR_FP64, **kw):
    p = dict(f_type=1, rk3_timestep=2, rk3_dt=0.005, nu=0.01)
    p.update(FORCE)
    p.update(kw)
    o = oracle_mod.Oracle(n, **p)
    g = pa.Pandora(n, precision=precision, **p)
    return o, g


def test_forcing_draws_match_oracle(oracle_mod):
    note("forcing OU update on the device vs oracle, 3 stages")
    o, g = _pair(oracle_mod, 32)
    z = np.zeros(o.shape, complex)
    o.set_state(z, z, z)
    g.set_state(z, z, z)
    to = tg = 0.0
    to = o.step(1, to)
    tg = g.rk3_calc(1, tg)
    fo, fg = o.get_forcing(), g.get_forcing()
    assert max(np.abs(f).max() for f in fo) > 0.01
    for a, b in zip(fg, fo):
        assert rel(a, b) < 1e-13
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("n", [32, 48])
def test_forced_isotropic_parity_fp64(oracle_mod, n):
    note(f"forced isotropic FP64 N={n}, 10 steps")
    o, g = _pair(oracle_mod, n)
    u = pa.init_field_host(n, 300 + n, 2.0 / 3.0, 3.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 11):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    for a, b in zip(g.get_forcing(), o.get_forcing()):
        assert rel(a, b) < 1e-13


def test_forcing_set_get_and_stop(oracle_mod):
    note("forcing set/get round trip and FORCE_STOP")
    o, g = _pair(oracle_mod, 32, force_stop=3)
    rng = np.random.default_rng(5)
    nk = g.force_nk
    f = [rng.standard_normal((nk + 1, 2 * nk + 1, 2 * nk + 1)) + 0j for _ in range(3)]
    g.set_forcing(*f)
    for a, b in zip(g.get_forcing(), f):
        assert np.array_equal(a, b)
    u = pa.init_field_host(32, 9, 2.0 / 3.0, 3.0)
    g.set_state(*u)
    t = 0.0
    for s in range(1, 5):
        t = g.rk3_calc(s, t)
    assert max(np.abs(a).max() for a in g.get_forcing()) == 0.0


def test_forced_fp32_mixed_spectrum(oracle_mod):
    note("forced isotropic FP32-mixed N=64 vs FP64, 20 steps")
    n = 64
    g64 = pa.Pandora(n, precision=pa.PDR_FP64, f_type=1, rk3_timestep=2, rk3_dt=0.005, nu=0.01, **FORCE)
    g32 = pa.Pandora(n, precision=pa.PDR_FP32_MIXED, f_type=1, rk3_timestep=2, rk3_dt=0.005, nu=0.01,
                     **FORCE)
    u = pa.init_field_host(n, 77, 2.0 / 3.0, 3.0)
    g64.set_state(*u)
    g32.set_state(*u)
    t64 = t32 = 0.0
    for s in range(1, 21):
        t64 = g64.rk3_calc(s, t64)
        t32 = g32.rk3_calc(s, t32)
    e64 = g64.spectrum()[0]
    e32 = g32.spectrum()[0]
    m = e64 > 1e-12 * e64.max()
    assert np.max(np.abs(e32[m] - e64[m]) / e64[m]) < 1e-4
