"""The TS_CORR two-point correlation (r1i, lint11 of f_FluidstatsRs2D /
RsCollect, src/pf_fluidstats.f90:933-945, 1092-1106) on the device against the
oracle's real-space sums: isotropic, BF_ROT and R_ROT in a sheared frame after
steps with a remesh, a box length of radix 5, an FP32 state, and the stage-1 record."""
import sys

import numpy as np
import pytest

import pandora_amd as pa

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


CASES = {
    "iso32": (32, dict(nu=0.01, rk3_dt=0.005), 2),
    "iso40": (40, dict(nu=0.01, rk3_dt=0.005), 1),
    "bfrot48": (48, dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005), 3),
    "rrot32": (32, dict(f_type=2, ns_deformed=3, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005), 3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_correlation_matches_oracle(oracle_mod, case):
    n, kw, steps = CASES[case]
    note(f"two-point correlation {case}: N={n}, {steps} steps")
    u = pa.init_field_host(n, 60 + n, 2.0 / 3.0, 5.0)
    g = pa.Pandora(n, **kw)
    o = oracle_mod.Oracle(n, **kw)
    g.set_state(*u)
    o.set_state(*u)
    tg = to = 0.0
    for s in range(1, steps + 1):
        tg = g.rk3_calc(s, tg)
        to = o.step(s, to)
    r_g, l_g = g.rs_correlation()
    r_o, l_o = o.rs_correlation()
    assert len(r_g) == n // 2 + 1
    assert np.abs(r_g - r_o).max() < 1e-12
    assert l_g == pytest.approx(l_o, rel=1e-12)
    g.close()
    o.close()


def test_correlation_fp32_state(oracle_mod):
    n = 48
    note(f"two-point correlation of an FP32-mixed state, N={n}")
    u = pa.init_field_host(n, 5, 2.0 / 3.0, 5.0)
    g = pa.Pandora(n, nu=0.01, rk3_dt=0.005, precision=pa.PDR_FP32_MIXED)
    g.set_state(*u)
    o = oracle_mod.Oracle(n, nu=0.01, rk3_dt=0.005)
    o.set_state(*g.get_state())  # the FP32 state, widened
    r_g, l_g = g.rs_correlation()
    r_o, l_o = o.rs_correlation()
    assert np.abs(r_g - r_o).max() < 1e-12
    assert l_g == pytest.approx(l_o, rel=1e-12)
    g.close()


def test_stage_record_correlation(oracle_mod):
    """With the stage-1 record on, the record's correlation is that of the state
    stage 1 of the step started from (after its remesh check)."""
    n = 32
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005)
    note(f"stage-1 record correlation, BF_ROT N={n}, 4 steps")
    u = pa.init_field_host(n, 8, 2.0 / 3.0, 5.0)
    g = pa.Pandora(n, **kw)
    g.set_state(*u)
    g.set_stage_stats(1)
    assert g.stage_correlation()[0] == -1
    o = oracle_mod.Oracle(n, **kw)
    o.set_state(*u)
    tg = to = 0.0
    for s in range(1, 5):
        c = oracle_mod.Oracle(n, **kw)
        c.set_state(*o.get_state())
        c.set_frame(*o.get_frame())
        c.remesh(s)
        r_o, l_o = c.rs_correlation()
        c.close()
        tg = g.rk3_calc(s, tg)
        t_rec, r_g, l_g = g.stage_correlation()
        assert t_rec == s - 1
        assert np.abs(r_g - r_o).max() < 1e-12
        assert l_g == pytest.approx(l_o, rel=1e-12)
        to = o.step(s, to)
    g.close()
    o.close()
