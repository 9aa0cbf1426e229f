"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle on the same
seeded inputs.  Criterion: norm-wise relative difference ||a-b||_2/||b||_2 per
component (distinct FFT algorithms cannot be bit-identical); FP64 <= 1e-12 after
10 steps, the remesh index map and the shear clock bit-exact, FP32-mixed E(k)
within 1e-4 of FP64."""
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


def _pair(oracle_mod, n, **kw):
    o = oracle_mod.Oracle(n, f_type=kw.get("f_type", 1), rk3_timestep=kw.get("rk3_timestep", 2),
                          rk3_dt=kw.get("rk3_dt", 0.005), rk3_cfl=kw.get("rk3_cfl", 0.75),
                          nu=kw.get("nu", 0.01), deform_s=kw.get("deform_s", 0.0),
                          remesh_t=kw.get("remesh_t", 0), force_tf=kw.get("force_tf", 0.0))
    g = pa.Pandora(n, f_type=kw.get("f_type", 1), rk3_timestep=kw.get("rk3_timestep", 2),
                   rk3_dt=kw.get("rk3_dt", 0.005), rk3_cfl=kw.get("rk3_cfl", 0.75), nu=kw.get("nu", 0.01),
                   deform_s=kw.get("deform_s", 0.0), remesh_t=kw.get("remesh_t", 0),
                   force_tf=kw.get("force_tf", 0.0), precision=kw.get("precision", pa.PDR_FP64))
    return o, g


def test_device_properties():
    note("device properties")
    g = pa.Pandora(32)
    s = g.sizes()
    assert s[0] == 11 and s[1] == 21 and s[2] == 21
    g.close()


@pytest.mark.parametrize("f_type", [1, 2])
def test_pass1_vorticity_and_z_transform_match_oracle(oracle_mod, f_type):
    note(f"pass 1 (vorticity + inverse z FFT) N=32 f_type={f_type}")
    n = 32
    o, g = _pair(oracle_mod, n, f_type=f_type, rk3_timestep=2, deform_s=1.0)
    u = pa.init_field_host(n, 21, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    if f_type == 2:  # a sheared frame: bmat(1,3) = -0.3, binv(1,3) = 0.3
        b, bi = o.get_frame()
        b[6], bi[6] = -0.3, 0.3
        o.set_frame(b, bi)
        g.set_frame(b, bi)
    o.prepare_realspace()
    for f in range(6):
        assert rel(g.debug_prepare(f), o.get_1r2w(f)) < 1e-13, f


@pytest.mark.parametrize("n", [32, 48, 64])
def test_isotropic_step_parity_fp64(oracle_mod, n):
    note(f"isotropic FP64 N={n}, 10 steps")
    o, g = _pair(oracle_mod, n, nu=0.01, rk3_dt=0.005)
    u = pa.init_field_host(n, 100 + n, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 11):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    assert to == tg
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    for a, b in zip(g.get_rk(), o.get_rk()):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("n", [32, 48])
def test_shear_step_parity_fp64_with_remesh(oracle_mod, n):
    """Rogallo frame, RK3_TIMESTEP = DECIDE, F_REMESH_T = 3, S = 20 (dt = 0.5/(S F_REMESH_T)
    = 1/120): remeshes at steps 4 and 10."""
    note(f"shear BF_ROT FP64 N={n}, 10 steps with remesh")
    o, g = _pair(oracle_mod, n, f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)
    u = pa.init_field_host(n, 200 + n, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 11):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
        bo, bio = o.get_frame()
        bg, big = g.get_frame()
        assert np.array_equal(bo, bg) and np.array_equal(bio, big), s
    assert to == tg
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_remesh_is_bit_exact(oracle_mod, prec):
    """The remesh (k-index gather, u1 += u3, x1 += x3 with wrap, frame reset) is
    bit-exact against the oracle; in FP32 it is exact on FP32-representable data."""
    note(f"remesh bit-exactness N=48 precision={prec}")
    n, R = 48, 2
    o, g = _pair(oracle_mod, n, f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=R, precision=prec)
    rng = np.random.default_rng(5)
    u = [(rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)).astype(np.complex64)
         .astype(np.complex128) for _ in range(3)]
    o.set_state(*u)
    g.set_state(*u)
    dx = TWOPI / n
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(3000, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    for tstep in (R, R + 2):
        assert not g.remesh(tstep)
    assert o.remesh(R + 1)
    assert g.remesh(R + 1)
    bo, bio = o.get_frame()
    bg, big = g.get_frame()
    assert np.array_equal(bo, bg) and np.array_equal(bio, big)
    vo = o.get_state()
    vg = g.get_state()
    if prec == pa.PDR_FP64:
        for a, b in zip(vg, vo):
            assert np.array_equal(a, b)
    else:
        for a, b in zip(vg, vo):
            assert np.array_equal(a, b.astype(np.complex64).astype(np.complex128))
    xo, _ = o.get_particles()
    xg, _ = g.get_particles()
    assert np.array_equal(xg, xo)


def test_tracers_parity_shear_fp64(oracle_mod):
    note("tracers + shear FP64 N=32, 2000 tracers, 8 steps (remesh at 4)")
    n = 32
    o, g = _pair(oracle_mod, n, f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=3, nu=0.005)
    u = pa.init_field_host(n, 7, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    rng = np.random.default_rng(8)
    dx = TWOPI / n
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(2000, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    to = tg = 0.0
    for s in range(1, 9):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    xo, ro = o.get_particles()
    xg, rg = g.get_particles()
    d = np.abs(xg - xo)
    d = np.minimum(d, TWOPI - d)  # a wrap decided differently would show as ~L
    assert d.max() < 1e-11
    assert rel(rg, ro) < 1e-12
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12


def test_tracers_parity_isotropic_fp64(oracle_mod):
    note("tracers isotropic FP64 N=24, 500 tracers, 5 steps")
    n = 24
    o, g = _pair(oracle_mod, n, nu=0.01, rk3_dt=0.01)
    u = pa.init_field_host(n, 9, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    rng = np.random.default_rng(3)
    dx = TWOPI / n
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(500, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    to = tg = 0.0
    for s in range(1, 6):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    xo, _ = o.get_particles()
    xg, _ = g.get_particles()
    d = np.abs(xg - xo)
    d = np.minimum(d, TWOPI - d)
    assert d.max() < 1e-11


@pytest.mark.parametrize("f_type", [1, 2])
def test_tracers_on_box_faces_and_crowded_cell_fp64(oracle_mod, f_type):
    """Tracers on the lower faces (x = -dx/2), just below the upper faces, exactly on
    x3max (p_MPICorrectPosition keeps it: the x3 upper test is '>',
    src/pp_mpi.f90:1011-1057) and 1000 tracers in one cell: the molecule search
    wraps through the periodic box in every direction (p_InterpFindMolecule,
    src/pp_interp.f90:406-467)."""
    n = 32
    note(f"tracers on box faces + crowded cell FP64 N={n} f_type={f_type}, 6 steps")
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005) if f_type == 2 else \
        dict(nu=0.01, rk3_dt=0.01)
    o, g = _pair(oracle_mod, n, **kw)
    u = pa.init_field_host(n, 60 + f_type, 2.0 / 3.0, 6.0)
    o.set_state(*u)
    g.set_state(*u)
    dx = TWOPI / n
    rng = np.random.default_rng(20 + f_type)
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(3000, 3))
    for d in range(3):
        xp[100 * d:100 * d + 50, d] = -0.5 * dx
        xp[100 * d + 50:100 * d + 100, d] = np.nextafter(TWOPI - 0.5 * dx, 0.0)
    xp[300:340, 2] = TWOPI - 0.5 * dx  # exactly x3max
    xp[400:1400] = xp[400]
    o.set_particles(xp)
    g.set_particles(xp)
    to = tg = 0.0
    for s in range(1, 7):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    xo, ro = o.get_particles()
    xg, rg = g.get_particles()
    d = np.abs(xg - xo)
    d = np.minimum(d, TWOPI - d)
    assert d.max() < 1e-11
    assert rel(rg, ro) < 1e-12
    assert xg.min() >= -0.5 * dx and xg.max() <= TWOPI - 0.5 * dx


def test_single_mode_closed_form_on_gpu():
    """Independent of the oracle: one Fourier mode decays by R(z)^steps exactly."""
    note("single-mode closed form FP64 N=24")
    n, nu, dt = 24, 0.05, 0.01
    g = pa.Pandora(n, nu=nu, rk3_dt=dt)
    u = [np.zeros(g.shape, complex) for _ in range(3)]
    k = np.array([1.0, 2.0, 3.0])
    a = np.array([3.0 + 1j, -1.0 + 0.5j, 0.0])
    a[2] = -(k[0] * a[0] + k[1] * a[1]) / k[2]
    for c in range(3):
        u[c][1, 2, 3] = a[c]
    g.set_state(*u)
    t = 0.0
    for s in range(1, 4):
        t = g.rk3_calc(s, t)
    z = -nu * (k @ k) * dt
    R = 1 + z + z**2 / 2 + z**3 / 6
    got = g.get_state()
    for c in range(3):
        assert abs(got[c][1, 2, 3] - a[c] * R**3) < 1e-14 * abs(a).max()


def test_cfl_timestep_matches_oracle(oracle_mod):
    note("CFL time step N=32")
    n = 32
    o, g = _pair(oracle_mod, n, rk3_timestep=3, rk3_cfl=0.5)
    u = pa.init_field_host(n, 31, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    to = o.step(1, 0.0)
    tg = g.rk3_calc(1, 0.0)
    assert abs(g.dt - o.dt) < 1e-14 * o.dt
    assert abs(tg - to) < 1e-14


def test_fp32_mixed_spectrum_against_fp64(oracle_mod):
    """FP32-mixed vs the FP64 oracle: E(k) per shell within 1e-4 relative after 20 steps."""
    note("FP32-mixed N=64 isotropic, 20 steps")
    n = 64
    o, g = _pair(oracle_mod, n, nu=0.01, rk3_dt=0.005, precision=pa.PDR_FP32_MIXED)
    u = pa.init_field_host(n, 77, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 21):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    eo, do, so, tot_o = o.spectrum()
    eg, dg, sg, tot_g = g.spectrum()
    live = eo > 1e-6 * eo.max()
    shell = (np.abs(eg[live] - eo[live]) / eo[live]).max()
    note(f"  E(k) max relative difference per shell (shells >= 1e-6 of the peak): {shell:.3e}")
    assert shell < 1e-4
    assert abs(tot_g[0] - tot_o[0]) / tot_o[0] < 1e-5
    assert abs(tot_g[1] - tot_o[1]) / tot_o[1] < 1e-5


def test_restart_roundtrip_on_device(tmp_path):
    note("restart roundtrip N=32")
    g = pa.Pandora(32)
    g.init_field(5, 2.0 / 3.0, 4.0)
    a = g.get_state()
    path = str(tmp_path / "r.bin")
    g.restart_write(path)
    g.init_field(6, 2.0 / 3.0, 4.0)
    g.restart_read(path)
    b = g.get_state()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_scaled_restart_on_device(tmp_path, prec):
    """pdr_restart_read_scaled: a 16^3 restart onto the 32^3 context equals the
    host copy (FP64 exactly; FP32 storage rounds each value once)."""
    note(f"scaled restart 16^3 -> 32^3 precision={prec}")
    uc = pa.init_field_host(16, 5, 2.0 / 3.0, 3.0)
    path = str(tmp_path / "coarse.bin")
    pa.restart_write_host(path, *uc)
    g = pa.Pandora(32, precision=prec)
    g.restart_read_scaled(path, 2)
    want = pa.restart_read_scaled_host(path, 32, 32, 32, 2)
    for a, b in zip(g.get_state(), want):
        if prec == pa.PDR_FP64:
            assert np.array_equal(a, b)
        else:
            assert np.array_equal(a, b.astype(np.complex64).astype(np.complex128))
    g.close()


def test_error_paths_on_device():
    note("error paths")
    g = pa.Pandora(16)
    with pytest.raises(pa.PdrError) as e:
        g.get_particles()
    assert e.value.code == -6
    u = [np.zeros(g.shape, complex) for _ in range(3)]
    u[1][2, 3, 4] = np.nan
    g.set_state(*u)
    assert g.check_finite() is False
    u[1][2, 3, 4] = 0.0
    g.set_state(*u)
    assert g.check_finite() is True
    with pytest.raises(pa.PdrError) as e:
        pa.Pandora(2560, precision=pa.PDR_FP32_MIXED)  # fits the LDS, exceeds 288 GB
    assert e.value.code == -3


def _div_rel(u, shape):
    nki, nkj, nkk = shape
    kj = np.arange(nkj)
    kk = np.arange(nkk)
    kj = np.where(kj <= nkj // 2, kj, kj - nkj).astype(float)
    kk = np.where(kk <= nkk // 2, kk, kk - nkk).astype(float)
    KI, KJ, KK = np.meshgrid(np.arange(nki, dtype=float), kj, kk, indexing="ij")
    div = np.abs(KI * u[0] + KJ * u[1] + KK * u[2])
    mag = np.sqrt(KI**2 + KJ**2 + KK**2) * np.sqrt(sum(np.abs(a) ** 2 for a in u))
    return float(div.max() / mag.max())


def test_full_size_256_shear_fp64_with_tracers():
    """The FP64 drop-in workload at full size (256^3 shear, 262,144 tracers): size-independent
    properties after 3 steps (finite, divergence-free, Hermitian i=0 plane, tracers in the box)."""
    note("256^3 shear FP64 + 262144 tracers, 3 steps")
    n = 256
    g = pa.Pandora(n, f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=250, nu=0.005)
    g.init_field(1, 2.0 / 3.0, 6.0)
    g.init_particles(101, 262144)
    t = g.run(1, 3, 0.0)
    assert abs(t - 3 * 0.002) < 1e-15
    assert g.check_finite()
    u = g.get_state()
    assert _div_rel(u, g.shape) < 1e-13
    jn = (-np.arange(g.nkj)) % g.nkj
    kn = (-np.arange(g.nkk)) % g.nkk
    for c in range(3):
        p = u[c][0]
        assert np.abs(p - np.conj(p[np.ix_(jn, kn)])).max() < 1e-12 * np.abs(p).max()
    xp, _ = g.get_particles()
    dx = TWOPI / n
    assert xp.min() >= -0.5 * dx and xp.max() <= TWOPI - 0.5 * dx


def test_full_size_512_shear_fp32_with_tracers():
    """The benchmark workload (512^3 shear, 1M tracers, FP32-mixed) for 2 steps."""
    note("512^3 shear FP32-mixed + 1M tracers, 2 steps")
    n = 512
    g = pa.Pandora(n, f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=250, nu=0.005,
                   precision=pa.PDR_FP32_MIXED)
    g.init_field(2, 2.0 / 3.0, 6.0)
    g.init_particles(102, 1 << 20)
    e0 = g.spectrum()[3]
    g.run(1, 2, 0.0)
    assert g.check_finite()
    u = g.get_state()
    assert _div_rel(u, g.shape) < 1e-6
    e1 = g.spectrum()[3]
    k0 = 0.5 * (e0[3] + e0[4] + e0[5])
    k1 = 0.5 * (e1[3] + e1[4] + e1[5])
    assert 0.9 * k0 < k1 < 1.1 * k0


@pytest.mark.parametrize("f_type", [1, 2])
def test_generic_kernel_path_parity(oracle_mod, monkeypatch, f_type):
    """The runtime-radix kernels (used for lengths without a specialised kernel)
    stay at parity: forced with PDR_GENERIC_KERNELS=1."""
    note(f"generic kernels FP64 N=48 f_type={f_type}, 6 steps")
    monkeypatch.setenv("PDR_GENERIC_KERNELS", "1")
    n = 48
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005) if f_type == 2 else dict(nu=0.01)
    o, g = _pair(oracle_mod, n, **kw)
    u = pa.init_field_host(n, 300 + f_type, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 7):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12


def test_radix5_length_uses_generic_path(oracle_mod):
    """N = 40 = 2^3 * 5 has no specialised kernel: the generic path runs it."""
    note("N=40 (radix 5) isotropic FP64, 4 steps")
    n = 40
    o, g = _pair(oracle_mod, n, nu=0.01)
    u = pa.init_field_host(n, 41, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    to = tg = 0.0
    for s in range(1, 5):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_chunked_plane_pass_parity(oracle_mod, monkeypatch, prec):
    """The y/x passes run chunk by chunk over z with a small x-pencil buffer
    (PDR_ZCHUNK_MB = 0.75 MB: 5 planes per chunk at 48^3 FP64, the last chunk
    partial); shear with remesh and tracers stays at parity."""
    note(f"chunked plane pass N=48 prec={prec}, shear + tracers, 6 steps")
    monkeypatch.setenv("PDR_ZCHUNK_MB", "0.75")
    n = 48
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005)
    o, g = _pair(oracle_mod, n, precision=prec, **kw)
    u = pa.init_field_host(n, 77, 2.0 / 3.0, 4.0)
    o.set_state(*u)
    g.set_state(*u)
    rng = np.random.default_rng(8)
    dx = TWOPI / n
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(500, 3))
    o.set_particles(xp)
    g.set_particles(xp)
    to = tg = 0.0
    for s in range(1, 7):
        to = o.step(s, to)
        tg = g.rk3_calc(s, tg)
    tol = 1e-12 if prec == pa.PDR_FP64 else 1e-5
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < tol
    xo = o.get_particles()[0]
    xg = g.get_particles()[0]
    assert np.abs(xg - xo).max() < (1e-11 if prec == pa.PDR_FP64 else 1e-5)


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
@pytest.mark.parametrize("chunk_mb", ["", "0.75"])
def test_two_chunks_in_flight_match_serial(monkeypatch, prec, chunk_mb):
    """Odd z chunks on a second stream with their own x-pencil buffer
    (PDR_ZSTREAMS=2) give the serial passes' result bit for bit: the chunks
    touch disjoint planes and the CFL maximum is order-free (shear with remesh,
    CFL time step, tracers, N=48; two half-box chunks, and 5-plane chunks)."""
    note(f"two chunks in flight vs serial N=48 prec={prec} chunk_mb={chunk_mb or 'half box'}")
    n = 48
    kw = dict(f_type=2, rk3_timestep=3, rk3_cfl=0.5, deform_s=20.0, remesh_t=2, nu=0.005, precision=prec)
    u = pa.init_field_host(n, 91, 2.0 / 3.0, 4.0)
    rng = np.random.default_rng(12)
    dx = TWOPI / n
    xp = rng.uniform(-0.5 * dx, TWOPI - 0.5 * dx, size=(700, 3))
    if chunk_mb:
        monkeypatch.setenv("PDR_ZCHUNK_MB", chunk_mb)
    out = []
    for zs in ("1", "2"):
        monkeypatch.setenv("PDR_ZSTREAMS", zs)
        g = pa.Pandora(n, **kw)
        g.set_state(*u)
        g.set_particles(xp)
        t = g.run(1, 6, 0.0)
        out.append((t, g.get_state(), g.get_particles()))
        g.close()
    (ta, sa, pa_), (tb, sb, pb) = out
    assert ta == tb
    for a, b in zip(sa, sb):
        assert np.array_equal(a, b)
    assert np.array_equal(pa_[0], pb[0]) and np.array_equal(pa_[1], pb[1])


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_fused_zforward_formf_matches_separate(monkeypatch, prec):
    """The fused z-forward + FormF kernel gives the separate zfwd / formf
    kernels' result (shear with tracers, N=64, 4 steps)."""
    note(f"fused vs separate z-forward + FormF N=64 prec={prec}")
    n = 64
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005, precision=prec)
    u = pa.init_field_host(n, 31, 2.0 / 3.0, 4.0)
    out = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("PDR_FUSE_ZFORMF", fuse)
        g = pa.Pandora(n, **kw)
        g.set_state(*u)
        t = 0.0
        for s in range(1, 5):
            t = g.rk3_calc(s, t)
        out.append((g.get_state(), g.get_rk()))
        g.close()
    tol = 1e-14 if prec == pa.PDR_FP64 else 1e-6
    for a, b in zip(out[0][0] + out[0][1], out[1][0] + out[1][1]):
        assert rel(a, b) < tol


@pytest.mark.parametrize("shape", [(64, 32, 48), (32, 32, 64)])
def test_non_cubic_box_parity(oracle_mod, shape):
    """Non-cubic boxes: (64, 32, 48) cannot use the whole-plane y tiles (NX/3 != NY/3)
    and runs the runtime-radix kernels; (32, 32, 64) runs the specialised ones."""
    nx, ny, nz = shape
    note(f"non-cubic box {shape} FP64, 4 steps")
    kw = dict(rk3_timestep=2, rk3_dt=0.005, nu=0.01)
    g = pa.Pandora(nx=nx, ny=ny, nz=nz, **kw)
    o = oracle_mod.Oracle(nx=nx, ny=ny, nz=nz, **kw)
    rng = np.random.default_rng(5)
    u = [0.05 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)) for _ in range(3)]
    for c in range(3):  # real field: Hermitian i = 0 plane, zero mean
        p = u[c][0]
        u[c][0] = 0.5 * (p + np.conj(np.roll(p[::-1, ::-1], 1, axis=(0, 1))))
        u[c][0, 0, 0] = 0.0
    g.set_state(*u)
    o.set_state(*u)
    tg = to = 0.0
    for s in range(1, 5):
        tg = g.rk3_calc(s, tg)
        to = o.step(s, to)
    for a, b in zip(g.get_state(), o.get_state()):
        assert rel(a, b) < 1e-12
    g.close()
