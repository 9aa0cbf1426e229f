"""The length-specialised kernels change form with the line length: x tiles of
2 rows above 256 points, the z inverse that re-reads the velocity tile per
field and separate z-forward / FormF kernels above 512, 8-line (FP32) and
4/2-line (FP64) tiles for 768..1536-point lines, and x-tile twiddles read from
the global table above 1024.  Boxes with one long axis and two 128-point axes
run every such form at a size the FP64 oracle finishes in seconds; each is
checked against the oracle (FP64 GPU <= 1e-12 norm-wise, FP32-mixed <= 1e-5),
with shear, a remesh and tracers.  The full-size workloads follow: the FP64
drop-in case (256^3 shear, 262,144 tracers, 10 steps) against the oracle, and
the two largest FP32-mixed cases (1024^3 forced with 16,777,216 tracers,
1536^3 shear with 67,108,864 tracers) through size-independent properties."""
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


def _tracers(nx, ny, nz, count, seed):
    rng = np.random.default_rng(seed)
    lo = np.array([-0.5 * TWOPI / nx, -0.5 * TWOPI / ny, -0.5 * TWOPI / nz])
    return lo + rng.uniform(0.0, TWOPI, size=(count, 3))


# (nx, ny, nz): the long axis selects the form under test
LONG_AXIS = [
    (512, 128, 128),    # 2-row x tiles, 512-point x lines
    (128, 512, 128),    # 512-point y lines (inverse and forward y tiles)
    (128, 128, 512),    # register z inverse and fused z-forward + FormF at 512, two chunks in flight
    (768, 128, 128),    # mid-length x tile
    (128, 768, 128),    # mid-length y tiles
    (128, 128, 768),    # z inverse with the velocity re-read, separate z forward and FormF
    (1536, 128, 128),   # x twiddles from the global table
    (128, 1536, 128),   # long y tiles (FP32: 8 inverse / 4 forward lines; FP64: 2)
    (128, 128, 1536),   # long z tiles
]


@pytest.mark.parametrize("shape", LONG_AXIS, ids=lambda s: "x".join(map(str, s)))
def test_one_long_axis_shear_against_oracle(oracle_mod, shape):
    """BF_ROT shear (S = 100, F_REMESH_T = 4, so dt = 1.25e-3 and a remesh at step
    5), 4096 tracers, 6 steps: FP64 against the oracle to 1e-12 (state, frame
    bit for bit, tracers 1e-11), FP32-mixed against it to 1e-5 (tracers 1e-6)."""
    nx, ny, nz = shape
    note(f"one long axis {shape}: BF_ROT shear + 4096 tracers, 6 steps, oracle vs FP64 and FP32-mixed")
    kw = dict(f_type=2, rk3_timestep=1, deform_s=100.0, remesh_t=4, nu=0.005)
    u = pa.init_field_host(nx, 500 + nx + 3 * ny + 7 * nz, 0.05, 6.0, ny=ny, nz=nz)
    xp0 = _tracers(nx, ny, nz, 4096, nx + ny + nz)
    o = oracle_mod.Oracle(nx=nx, ny=ny, nz=nz, **kw)
    o.set_state(*u)
    o.set_particles(xp0)
    to = 0.0
    for s in range(1, 7):
        to = o.step(s, to)
    so = o.get_state()
    xo, ro = o.get_particles()
    fo = o.get_frame()
    o.close()
    for prec in (pa.PDR_FP64, pa.PDR_FP32_MIXED):
        g = pa.Pandora(nx=nx, ny=ny, nz=nz, precision=prec, **kw)
        forms = g.kernel_forms()
        note(f"  precision={prec} forms={forms}")
        assert forms["specialised"] == 1
        if nx >= 512:
            assert forms["xrows"] == 2
        assert forms["x_tw_global"] == (1 if nx > 1024 else 0)
        assert forms["zinv_register"] == (1 if nz <= 512 else 0)
        assert forms["fused_zformf"] == (1 if nz <= 512 else 0)
        g.set_state(*u)
        g.set_particles(xp0)
        tg = 0.0
        for s in range(1, 7):
            tg = g.rk3_calc(s, tg)
        assert tg == to
        fg = g.get_frame()
        assert np.array_equal(fg[0], fo[0]) and np.array_equal(fg[1], fo[1])
        xg, rg = g.get_particles()
        sg = g.get_state()
        g.close()
        if prec == pa.PDR_FP64:
            for a, b in zip(sg, so):
                assert rel(a, b) < 1e-12
            assert _wrapdiff(xg, xo).max() < 1e-11
            assert rel(rg, ro) < 1e-12
        else:
            for a, b in zip(sg, so):
                assert rel(a, b) < 1e-5
            assert _wrapdiff(xg, xo).max() < 1e-6


def test_cfg1_256_shear_fp64_with_tracers_against_oracle(oracle_mod):
    """The FP64 drop-in workload at full size: 256^3 homogeneous shear in the
    Rogallo frame (S = 1, nu = 0.005, F_REMESH_T = 250 so dt = 0.002, k0 = 6),
    262,144 tracers, 10 steps, against the oracle: Fourier coefficients
    <= 1e-12 norm-wise, tracer positions <= 1e-11, frame and simtime exact."""
    n, np_ = 256, 262144
    note(f"256^3 BF_ROT shear FP64 + {np_} tracers, 10 steps vs oracle")
    kw = dict(f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=250, nu=0.005)
    u = pa.init_field_host(n, 1, 2.0 / 3.0, 6.0)
    xp0 = _tracers(n, n, n, np_, 101)
    o = oracle_mod.Oracle(n, **kw)
    o.set_state(*u)
    o.set_particles(xp0)
    to = 0.0
    for s in range(1, 11):
        to = o.step(s, to)
    g = pa.Pandora(n, **kw)
    g.set_state(*u)
    g.set_particles(xp0)
    tg = g.run(1, 10, 0.0)
    assert tg == to
    fo, fg = o.get_frame(), g.get_frame()
    assert np.array_equal(fo[0], fg[0]) and np.array_equal(fo[1], fg[1])
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    for a, b in zip(g.get_rk(), o.get_rk()):
        assert rel(a, b) < 1e-12
    xg, rg = g.get_particles()
    xo, ro = o.get_particles()
    assert _wrapdiff(xg, xo).max() < 1e-11
    assert rel(rg, ro) < 1e-12
    g.close()
    o.close()


def _wave_properties(g, tol_div):
    """Finite, divergence-free (|k.u| against |k||u|, stored integer wavenumbers:
    the isotropic and BF_ROT continuity fix) and Hermitian on the i = 0 plane;
    slab by slab so that the host holds one copy of the state."""
    assert g.check_finite()
    u = g.get_state()
    nkj, nkk = g.nkj, g.nkk
    kj = np.arange(nkj, dtype=np.float64)
    kj = np.where(kj <= nkj // 2, kj, kj - nkj)
    kk = np.arange(nkk, dtype=np.float64)
    kk = np.where(kk <= nkk // 2, kk, kk - nkk)
    KJ, KK = np.meshgrid(kj, kk, indexing="ij")
    dmax = 0.0
    mmax = 0.0
    for i in range(g.nki):
        d = np.abs(float(i) * u[0][i] + KJ * u[1][i] + KK * u[2][i]).max()
        k = np.sqrt(float(i) ** 2 + KJ**2 + KK**2)
        m = (k * np.sqrt(np.abs(u[0][i]) ** 2 + np.abs(u[1][i]) ** 2 + np.abs(u[2][i]) ** 2)).max()
        dmax, mmax = max(dmax, d), max(mmax, m)
    assert dmax < tol_div * mmax, (dmax, mmax)
    jn = (-np.arange(nkj)) % nkj
    kn = (-np.arange(nkk)) % nkk
    for c in range(3):
        p = u[c][0]
        assert np.abs(p - np.conj(p[np.ix_(jn, kn)])).max() <= 1e-6 * np.abs(p).max()


def _tracers_inside(g, nx):
    xp, _ = g.get_particles()
    dx = TWOPI / nx
    assert np.isfinite(xp).all()
    assert xp.min() >= -0.5 * dx and xp.max() <= TWOPI - 0.5 * dx


def test_cfg2_forced_1024_fp32_full_size():
    """1024^3 forced isotropic turbulence in the FP32-mixed form (OU forcing of
    |k| <= 2.5, FORCE_TF = 0.2, sigma^2 T_F = 0.1; nu = 0.0008, dt = 5e-4, k0 = 3)
    with 16,777,216 tracers: 2 steps, then finite, divergence-free, Hermitian,
    energy within 10 % of the initial field, tracers inside the box."""
    n = 1024
    note("1024^3 forced FP32-mixed + 16777216 tracers, 2 steps (properties)")
    g = pa.Pandora(n, rk3_dt=0.0005, nu=0.0008, force_stop=1000, force_kf=2.5, force_tf=0.2,
                   force_sigma=float(np.sqrt(0.5)), force_seed=3, precision=pa.PDR_FP32_MIXED)
    g.init_field(3, 2.0 / 3.0, 3.0)
    k0 = 0.5 * sum(g.spectrum()[3][3:6])
    g.init_particles(103, 1 << 24)
    t = g.run(1, 2, 0.0)
    assert abs(t - 0.001) < 1e-15
    _wave_properties(g, 1e-5)
    k1 = 0.5 * sum(g.spectrum()[3][3:6])
    assert 0.9 * k0 < k1 < 1.1 * k0
    _tracers_inside(g, n)
    g.close()


def test_cfg4_shear_1536_fp32_full_size():
    """1536^3 homogeneous shear in the FP32-mixed form (S = 1, nu = 0.0003,
    F_REMESH_T = 2500 so dt = 2e-4) with 67,108,864 tracers: 2 steps, then the
    properties above."""
    n = 1536
    note("1536^3 shear FP32-mixed + 67108864 tracers, 2 steps (properties)")
    g = pa.Pandora(n, f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=2500, nu=0.0003,
                   precision=pa.PDR_FP32_MIXED)
    note(f"  forms={g.kernel_forms()}")
    g.init_field(4, 2.0 / 3.0, 6.0)
    k0 = 0.5 * sum(g.spectrum()[3][3:6])
    g.init_particles(104, 1 << 26)
    t = g.run(1, 2, 0.0)
    assert abs(t - 0.0004) < 1e-15
    _wave_properties(g, 1e-5)
    k1 = 0.5 * sum(g.spectrum()[3][3:6])
    assert 0.9 * k0 < k1 < 1.1 * k0
    _tracers_inside(g, n)
    g.close()
