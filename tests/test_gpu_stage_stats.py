"""Statistics of stage 1 inside the step (pdr_set_stage_stats / pdr_stage_stats):
the record of step N is f_FluidstatsFs((N-1)) of the state after step N-1 and
N's remesh check, and the real-space moments f_FluidstatsRs2D / RsCollect of
stage 1's own planes (src/pf_rk3.f90:58-60, 1073-1078, 1124-1143).  Checked
against the oracle's statistics of that state; the step's result does not
change when the record is taken."""
import sys

import numpy as np
import pytest

import pandora_amd as pa

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


CASES = {
    "iso": dict(f_type=1, rk3_timestep=2, rk3_dt=0.005, nu=0.01),
    "bfrot": dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005),
    "rrot": dict(f_type=2, ns_deformed=3, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005),
    "rconv": dict(f_type=2, ns_deformed=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005),
}


def _close(a, b, tol, what):
    for k in b:
        x, y = np.asarray(a[k], float), np.asarray(b[k], float)
        fin = np.isfinite(y)
        assert np.array_equal(np.isfinite(x), fin), (what, k)
        if not fin.any():  # the same entries are non-finite in both
            continue
        scale = max(np.abs(y[fin]).max(), 1e-300)
        assert np.abs(x[fin] - y[fin]).max() <= tol * scale, (what, k, x, y)


@pytest.mark.parametrize("case,env", [("iso", {}), ("bfrot", {}), ("bfrot", {"PDR_ZCHUNK_MB": "0.75"}),
                                      ("bfrot", {"PDR_ZSTREAMS": "2"}), ("rrot", {}), ("rconv", {})],
                         ids=["iso", "bfrot", "bfrot-chunked", "bfrot-two-streams", "rrot", "rconv"])
def test_stage_stats_match_oracle(oracle_mod, monkeypatch, case, env):
    """5 steps with the record every step (a remesh at step 3 in the shear
    cases): after each step the record equals the oracle's statistics of the
    state that stage 1 of that step started from."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    n = 48
    note(f"stage-1 statistics in the step: {case} {env} N={n}, 5 steps")
    kw = CASES[case]
    u = pa.init_field_host(n, 33, 2.0 / 3.0, 5.0)
    g = pa.Pandora(n, **kw)
    g.set_state(*u)
    g.set_stage_stats(1)
    assert g.stage_stats()[0] == -1
    o = oracle_mod.Oracle(n, **kw)
    o.set_state(*u)
    tg = to = 0.0
    for s in range(1, 6):
        # f_RK3Stage1's order: the remesh check, then f_FluidstatsFs; taken on a
        # copy of the oracle's state so that its own step remeshes once
        c = oracle_mod.Oracle(n, **kw)
        c.set_state(*o.get_state())
        c.set_frame(*o.get_frame())
        c.remesh(s)
        spec_o, q_o = c.stats()
        rs_o = c.rs_stats()
        s1_o, s2_o = c.spectra_1d()
        h_o = c.homogeneous() if kw["f_type"] != 1 else None
        c.close()
        tg = g.rk3_calc(s, tg)
        t_rec, q_g, spec_g, rs_g, s1_g, s2_g, h_g = g.stage_stats()
        assert t_rec == s - 1
        _close(q_g, q_o, 1e-11, f"wave-space step {s}")
        _close(spec_g, spec_o, 1e-11, f"spectra step {s}")
        _close(rs_g, rs_o, 1e-11, f"real-space step {s}")
        _close({"s1": s1_g, "s2": s2_g}, {"s1": s1_o, "s2": s2_o}, 1e-11, f"1-D spectra step {s}")
        if h_o is not None:
            _close(h_g, h_o, 1e-11, f"homogeneous record step {s}")
        to = o.step(s, to)
    assert tg == to
    g.close()


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_stage_stats_leave_the_step_unchanged(prec):
    """Shear with tracers: 6 steps with the record every 2 steps give the state,
    RK registers and tracers of the run without it, bit for bit; the record is
    that of step 5 (statistics step 4)."""
    n = 64
    note(f"stage-1 statistics do not change the step, precision={prec}")
    kw = dict(CASES["bfrot"], precision=prec)
    u = pa.init_field_host(n, 3, 2.0 / 3.0, 5.0)
    xp = np.random.default_rng(1).uniform(0.0, 6.0, size=(2000, 3))
    out = []
    for every in (0, 2):
        g = pa.Pandora(n, **kw)
        g.set_state(*u)
        g.set_particles(xp)
        g.set_stage_stats(every)
        g.run(1, 6, 0.0)
        out.append((g.get_state(), g.get_rk(), g.get_particles()[0], g.stage_stats()[0]))
        g.close()
    for a, b in zip(out[0][0] + out[0][1], out[1][0] + out[1][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == -1 and out[1][3] == 4
