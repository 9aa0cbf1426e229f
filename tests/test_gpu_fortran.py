"""The Fortran drop-in: pdr_fortran_driver runs TIMELOOP through the ISO_C_BINDING
shim (g_RK3Calc on odd steps, f_RK3Stage1/2/3 + g_RK3UpdateB on even steps) and
writes a PETSc-binary restart; the result must match the FP64 oracle.  With
tracers it either lets the stages advance them or drives them itself through
the shim's p_ControlRK3 / p_RK3SolveParticle[BF] / p_ControlFinishTimeStep
after each stage, as f_RK3RealSpace does (src/pf_rk3.f90:1145-1176)."""
import os
import subprocess

import numpy as np
import pytest

import pandora_amd as pa
from conftest import ROOT, rel

DRIVER = os.path.join(ROOT, "pandora-2.0_amd", "lib", "pdr_fortran_driver")


def test_fortran_driver_is_built():
    assert os.path.exists(DRIVER), "Fortran shim/driver not built (amdflang)"
    out = subprocess.run(["nm", DRIVER], capture_output=True, text=True, check=True).stdout
    assert "pdr_rk3_amd" in out


@pytest.mark.gpu
@pytest.mark.parametrize("ftype,nsteps", [(1, 4), (2, 5)])
def test_fortran_driver_matches_oracle(oracle_mod, tmp_path, ftype, nsteps):
    n, seed = 32, 11
    print(f"[gpu case] Fortran driver N={n} ftype={ftype} steps={nsteps}", flush=True)
    out = str(tmp_path / "restart.bin")
    r = subprocess.run([DRIVER, str(n), str(nsteps), str(ftype), str(seed), out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = pa.restart_read_host(out, pa.wave_shape(n))
    if ftype == 1:
        o = oracle_mod.Oracle(n, f_type=1, rk3_timestep=2, rk3_dt=0.005, nu=0.01)
        u = pa.init_field_host(n, seed, 2.0 / 3.0, 4.0)
    else:
        o = oracle_mod.Oracle(n, f_type=2, rk3_timestep=1, remesh_t=3, deform_s=20.0, nu=0.005)
        u = pa.init_field_host(n, seed, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    t = 0.0
    for s in range(1, nsteps + 1):
        t = o.step(s, t)
    simtime = float(r.stdout.split("simtime")[1].split()[0])
    assert abs(simtime - t) < 1e-15
    for a, b in zip(got, o.get_state()):
        assert rel(a, b) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("ftype", [1, 2])
def test_fortran_driver_tracers_match_oracle(oracle_mod, tmp_path, ftype):
    """500 tracers, 5 steps (remesh at step 4 in the shear run): mode 1 (in the
    stages) and mode 2 (the shim's particle entry points) give the same positions
    bit for bit, and both match the oracle to 1e-11."""
    n, seed, nsteps, np_ = 32, 13, 5, 500
    print(f"[gpu case] Fortran driver tracers N={n} ftype={ftype}", flush=True)
    res = {}
    for mode in (1, 2):
        out = str(tmp_path / f"restart{mode}.bin")
        r = subprocess.run([DRIVER, str(n), str(nsteps), str(ftype), str(seed), out, str(np_), str(mode)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        x0 = np.fromfile(out + ".x0", dtype=np.float64).reshape(np_, 3)
        x1 = np.fromfile(out + ".x1", dtype=np.float64).reshape(np_, 3)
        res[mode] = (x0, x1, pa.restart_read_host(out, pa.wave_shape(n)))
    assert np.array_equal(res[1][0], res[2][0])
    assert np.array_equal(res[1][1], res[2][1])
    for a, b in zip(res[1][2], res[2][2]):
        assert np.array_equal(a, b)
    if ftype == 1:
        o = oracle_mod.Oracle(n, f_type=1, rk3_timestep=2, rk3_dt=0.005, nu=0.01)
        u = pa.init_field_host(n, seed, 2.0 / 3.0, 4.0)
    else:
        o = oracle_mod.Oracle(n, f_type=2, rk3_timestep=1, remesh_t=3, deform_s=20.0, nu=0.005)
        u = pa.init_field_host(n, seed, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    o.set_particles(res[1][0])
    t = 0.0
    for s in range(1, nsteps + 1):
        t = o.step(s, t)
    xo, _ = o.get_particles()
    d = np.abs(res[1][1] - xo)
    twopi = 2.0 * float(np.float32(np.pi))
    assert np.minimum(d, twopi - d).max() < 1e-11
    for a, b in zip(res[1][2], o.get_state()):
        assert rel(a, b) < 1e-12
