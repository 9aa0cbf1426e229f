"""Batched ensembles (pdr_ensemble_*): members of uneven sizes advanced by one
graph launch per step must equal the same members stepped one by one through
pdr_step (the same device bodies, compiled into other kernels, so multiply-add
contraction may round differently: agreement to rounding), and the FP32-mixed
ensemble statistics must agree with the FP64 oracle (the decaying-ensemble case:
isotropic, k0 = 4, nu = 0.01, dt = 0.005)."""
import sys

import numpy as np
import pytest

import pandora_amd as pa
from conftest import rel

pytestmark = pytest.mark.gpu

SIZES = (32, 48, 64, 96)
KW = dict(rk3_timestep=2, rk3_dt=0.005, nu=0.01)


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _members(prec, sizes, seed0=0):
    out = []
    for k, n in enumerate(sizes):
        g = pa.Pandora(n, precision=prec, **KW)
        g.set_state(*pa.init_field_host(n, seed0 + k, 2.0 / 3.0, 4.0))
        out.append(g)
    return out


@pytest.mark.parametrize("prec", [pa.PDR_FP32_MIXED, pa.PDR_FP64])
def test_ensemble_equals_single_steps(prec):
    note(f"ensemble of 32/48/64/96 vs one-by-one pdr_step, 6 steps, precision {prec}")
    ens_m = _members(prec, SIZES)
    ref_m = _members(prec, SIZES)
    e = pa.Ensemble(ens_m)
    t = e.run(1, 6, np.zeros(len(SIZES)))
    for g in ref_m:
        tg = 0.0
        for s in range(1, 7):
            tg = g.rk3_calc(s, tg)
        assert tg == t[0]
    tol = 1e-6 if prec == pa.PDR_FP32_MIXED else 1e-13
    for a, b in zip(ens_m, ref_m):
        for x, y in zip(a.get_state(), b.get_state()):
            assert rel(x, y) < tol
    e.close()
    for g in ens_m + ref_m:
        g.close()


def test_ensemble_autotune_keeps_state_and_results():
    note("ensemble autotune: state restored, results unchanged (FP32, 8 members)")
    sizes = SIZES + SIZES
    ens_m = _members(pa.PDR_FP32_MIXED, sizes)
    ref_m = _members(pa.PDR_FP32_MIXED, sizes)
    e = pa.Ensemble(ens_m)
    before = [g.get_state() for g in ens_m]
    fy, fyf = e.autotune(2)
    assert fy in (1, 2, 3, 6) and fyf in (1, 3)
    for g, b in zip(ens_m, before):
        for x, y in zip(g.get_state(), b):
            np.testing.assert_array_equal(x, y)
    e.run(1, 4, np.zeros(len(sizes)))
    e.set_split(0, 0)
    e.run(5, 2, np.zeros(len(sizes)))
    for g in ref_m:
        g.run(1, 6, 0.0)
    for a, b in zip(ens_m, ref_m):
        for x, y in zip(a.get_state(), b.get_state()):
            assert rel(x, y) < 1e-6
    e.close()
    for g in ens_m + ref_m:
        g.close()


def test_ensemble_rejects_unsupported_members():
    note("ensemble argument checks")
    a = pa.Pandora(32, precision=pa.PDR_FP32_MIXED, **KW)
    b = pa.Pandora(32, precision=pa.PDR_FP64, **KW)
    c = pa.Pandora(32, f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=10, nu=0.01)
    d = pa.Pandora(32, **KW)
    d.set_particles(np.full((4, 3), 1.0))
    for bad in ([a, b], [c], [d], [a, a]):
        with pytest.raises(pa.PdrError):
            pa.Ensemble(bad)
    for g in (a, b, c, d):
        g.close()


def test_ensemble_statistics_against_fp64_oracle(oracle_mod):
    """Ensemble-mean E(k) and velocity-derivative (dissipation) skewness of an
    FP32-mixed ensemble (2 members of each size) after 10 steps against the
    FP64 oracle members."""
    note("ensemble FP32 vs FP64 oracle, 8 members, 10 steps")
    sizes = SIZES + SIZES
    ens_m = _members(pa.PDR_FP32_MIXED, sizes, seed0=100)
    e = pa.Ensemble(ens_m)
    e.run(1, 10, np.zeros(len(sizes)))
    for n in SIZES:
        idx = [k for k, m in enumerate(sizes) if m == n]
        eg, eo, sg, so = [], [], [], []
        for k in idx:
            o = oracle_mod.Oracle(n, **KW)
            o.set_state(*pa.init_field_host(n, 100 + k, 2.0 / 3.0, 4.0))
            to = 0.0
            for s in range(1, 11):
                to = o.step(s, to)
            spo, qo = o.stats()
            spg, qg = ens_m[k].stats()
            eg.append(spg["ken"])
            eo.append(spo["ken"])
            sg.append(qg["disskew"])
            so.append(qo["disskew"])
        mg, mo = np.mean(eg, axis=0), np.mean(eo, axis=0)
        big = mo > 1e-6 * mo.max()
        assert np.max(np.abs(mg[big] - mo[big]) / mo[big]) < 1e-4
        assert np.mean(sg) == pytest.approx(np.mean(so), rel=1e-4)
    e.close()
    for g in ens_m:
        g.close()
