"""Pins for the oracle's wave-space statistics (f_FluidstatsFs,
src/pf_fluidstats.f90:1293-1361, 1423-1792, 1977-2049, 2491-2558).

Checks that share no code with the oracle: Parseval against a real-space field
made with numpy's FFT (Reynolds stresses, TKE), the dissipation against
nu <|curl u|^2> in real space, the mode density against a brute-force count of
the full (Hermitian-completed) lattice, lab-frame binning against a count over
the transformed lattice, and algebraic identities of the secondary quantities
(Re_lambda^2 = 20 k^2 / (3 nu eps), eta^4 = nu^3 / eps, tau_eta^2 = nu / eps).
Real-space statistics (f_FluidstatsRs2D / RsCollect, src/pf_fluidstats.f90:
827-1146): closed forms for single Fourier modes (<cos^2> = 1/2, flatness 3/2,
maxima on the grid, the reference's unnormalised rs_rmsvort1) and Parseval
against the wave-space Reynolds stresses, in the isotropic and a sheared frame.
"""
import numpy as np
import pytest

ONEPI = float(np.float32(np.pi))


def _wvn(idx, kmax):
    idx = np.asarray(idx)
    return np.where(idx <= kmax, idx, idx - (2 * kmax + 1)).astype(np.float64)


def _random_state(o, seed):
    """Random solenoidal field with the i = 0 plane Hermitian and u(0) = 0."""
    rng = np.random.default_rng(seed)
    ki = np.arange(o.nki, dtype=float)
    kj = _wvn(np.arange(o.nkj), o.nkj // 2)
    kk = _wvn(np.arange(o.nkk), o.nkk // 2)
    KI, KJ, KK = np.meshgrid(ki, kj, kk, indexing="ij")
    k2 = KI**2 + KJ**2 + KK**2
    amp = np.exp(-k2 / 30.0)
    u = [amp * (rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)) for _ in range(3)]
    kd = (KI * u[0] + KJ * u[1] + KK * u[2]) / np.where(k2 > 0, k2, 1.0)
    u = [u[0] - KI * kd, u[1] - KJ * kd, u[2] - KK * kd]
    nj, nk = o.nkj, o.nkk
    for c in range(3):
        a = u[c]
        a[0, 0, 0] = 0.0
        for j in range(nj):
            for k in range(nk):
                jm, km = (-j) % nj, (-k) % nk
                if (j, k) < (jm, km):
                    a[0, jm, km] = np.conj(a[0, j, k])
                elif (j, k) == (jm, km):
                    a[0, j, k] = a[0, j, k].real
    return u


def _real_field(o, a):
    """u(x) = sum_k a(k) exp(i k.x) on the N^3 grid, with numpy (independent FFT)."""
    n = o.nx
    K = o.nki - 1
    full = np.zeros((n, n, n), complex)  # [x-freq][y-freq][z-freq]
    for i in range(o.nki):
        for j in range(o.nkj):
            jy = int(_wvn(j, o.nkj // 2))
            for k in range(o.nkk):
                kz = int(_wvn(k, o.nkk // 2))
                full[i % n, jy % n, kz % n] = a[i, j, k]
                if i > 0:
                    full[(-i) % n, (-jy) % n, (-kz) % n] = np.conj(a[i, j, k])
    r = np.fft.ifftn(full) * n**3
    assert np.abs(r.imag).max() < 1e-10 * max(np.abs(r.real).max(), 1e-300)
    return r.real, full


@pytest.fixture(scope="module")
def iso(oracle_mod):
    o = oracle_mod.Oracle(16, nu=0.013)
    u = _random_state(o, 3)
    o.set_state(*u)
    return o, u


def test_parseval_reynolds_stresses(iso):
    o, u = iso
    sp, q = o.stats()
    r = [_real_field(o, a)[0] for a in u]
    want = [np.mean(r[0] * r[0]), np.mean(r[1] * r[1]), np.mean(r[2] * r[2]),
            np.mean(r[0] * r[1]), np.mean(r[0] * r[2]), np.mean(r[1] * r[2])]
    got = [q[n] for n in ("u1u1", "u2u2", "u3u3", "u1u2", "u1u3", "u2u3")]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14)
    assert abs(q["tke"] - 0.5 * (want[0] + want[1] + want[2])) < 1e-13 * q["tke"]
    assert abs(q["f_ken"] - q["tke"]) < 1e-13 * q["tke"]  # every k != 0 mode is binned


def test_dissipation_real_space(iso):
    o, u = iso
    nu = 0.013
    n = o.nx
    _, fulls = zip(*[_real_field(o, a) for a in u])
    f = np.fft.fftfreq(n, 1.0 / n)
    KX, KY, KZ = np.meshgrid(f, f, f, indexing="ij")
    F1, F2, F3 = fulls
    w = [1j * (KY * F3 - KZ * F2), 1j * (KZ * F1 - KX * F3), 1j * (KX * F2 - KY * F1)]
    wr = [np.fft.ifftn(a).real * n**3 for a in w]
    eps = nu * np.mean(wr[0] ** 2 + wr[1] ** 2 + wr[2] ** 2)
    sp, q = o.stats()
    assert abs(q["f_dis"] - eps) < 1e-12 * eps


def test_mode_density_brute_force(iso):
    o, _ = iso
    sp, q = o.stats()
    K = o.nki - 1
    r = np.arange(-K, K + 1)
    X, Y, Z = np.meshgrid(r, r, r, indexing="ij")
    km = np.sqrt(X**2 + Y**2 + Z**2).ravel()
    kkmax = o.kkmax
    for kk in range(1, kkmax + 1):
        cnt = np.count_nonzero((km >= kk - 0.5) & (km < kk + 0.5))
        vol = 4.0 / 3.0 * ONEPI * ((kk + 0.5) ** 3 - (kk - 0.5) ** 3)
        assert sp["dens_mod"][kk - 1] == pytest.approx(cnt / vol, rel=1e-15, abs=0)


def test_secondary_identities(iso):
    o, _ = iso
    nu = 0.013
    sp, q = o.stats()
    k, eps = q["fn_ken"], q["fn_dis"]
    assert q["rn_tay"] ** 2 == pytest.approx(20.0 * k * k / (3.0 * nu * eps), rel=1e-13)
    assert q["ln_kol"] ** 4 == pytest.approx(nu**3 / eps, rel=1e-13)
    assert q["tn_kol"] ** 2 == pytest.approx(nu / eps, rel=1e-13)
    assert q["fn_ken"] == pytest.approx(sp["ken"].sum(), rel=1e-14)
    # Pope 6.225: L = pi / (2 u'^2) int E(k) / k dk, u'^2 = 2k/3
    L = ONEPI / (2.0 * (2.0 * k / 3.0)) * np.sum(sp["ken"] / np.arange(1, o.kkmax + 1))
    assert q["ln_int"] == pytest.approx(L, rel=1e-13)
    # disskew = S_eps * (15/eps)^{3/2} sqrt(nu) / 35 with S_eps = 4 nu^2 sum k^4 E
    ds = q["fn_disskew"] * (15.0 / eps) ** 1.5 * np.sqrt(nu) / 35.0
    assert q["disskew"] == pytest.approx(ds, rel=1e-13)


def test_sheared_frame_bins_by_lab_wavenumber(oracle_mod):
    o = oracle_mod.Oracle(16, f_type=2, rk3_timestep=2, deform_s=1.0, nu=0.01)
    u = _random_state(o, 9)
    o.set_state(*u)
    b, bi = o.get_frame()
    b[6], bi[6] = -0.7, 0.7   # bmat(1,3), binv(1,3): a sheared Rogallo frame
    o.set_frame(b, bi)
    sp, q = o.stats()
    K = o.nki - 1
    # counts over the compact storage with weight 2 for i > 0, |k_lab| with
    # k_lab = bmat^T k (k_lab3 = k3 + bmat(1,3) k1)
    ki = np.arange(0, K + 1)
    kj = np.arange(-K, K + 1)
    I, J, Kz = np.meshgrid(ki, kj, kj, indexing="ij")
    kl = np.sqrt(I**2.0 + J**2.0 + (Kz + (-0.7) * I) ** 2.0).ravel()
    wgt = np.where(I.ravel() > 0, 2, 1)
    for kk in range(1, o.kkmax + 1):
        m = (kl >= kk - 0.5) & (kl < kk + 0.5)
        vol = 4.0 / 3.0 * ONEPI * ((kk + 0.5) ** 3 - (kk - 0.5) ** 3)
        assert sp["dens_mod"][kk - 1] == pytest.approx(wgt[m].sum() / vol, rel=1e-15, abs=0)
    # lab velocity u_lab1 = u1 + binv(1,3) u3 enters u1u1
    w = np.where(np.arange(o.nki)[:, None, None] > 0, 2.0, 1.0)
    ul1 = u[0] + 0.7 * u[2]
    assert q["u1u1"] == pytest.approx(np.sum(w * np.abs(ul1) ** 2), rel=1e-13)


def test_rrot_frame_takes_the_state_as_laboratory_velocity(oracle_mod):
    """f_FluidstatsFsPrimaryQuantsRRot (src/pf_fluidstats.f90:1801-1970): R_ROT
    holds laboratory velocities, so its statistics of u equal the BF_ROT ones of
    bmat u (whose laboratory velocity binv bmat u is u again); same bins, same
    mode density."""
    r = oracle_mod.Oracle(16, f_type=2, ns_deformed=3, rk3_timestep=2, deform_s=1.0, nu=0.01)
    f = oracle_mod.Oracle(16, f_type=2, ns_deformed=1, rk3_timestep=2, deform_s=1.0, nu=0.01)
    u = _random_state(r, 9)
    b, bi = r.get_frame()
    b[6], bi[6] = -0.7, 0.7
    for o in (r, f):
        o.set_frame(b, bi)
    r.set_state(*u)
    f.set_state(u[0] - 0.7 * u[2], u[1], u[2])
    (sr, qr), (sf, qf) = r.stats(), f.stats()
    for k in sr:
        np.testing.assert_allclose(sr[k], sf[k], rtol=1e-12, atol=1e-14 * np.abs(sf[k]).max())
    for k in qr:
        assert qr[k] == pytest.approx(qf[k], rel=1e-12, abs=1e-15), k
    w = np.where(np.arange(r.nki)[:, None, None] > 0, 2.0, 1.0)
    assert qr["u1u1"] == pytest.approx(np.sum(w * np.abs(u[0]) ** 2), rel=1e-13)


# ---- real-space statistics (f_FluidstatsRs2D / RsCollect) --------------------

def _single_mode(o, comp, axis, amp):
    """u_comp = amp cos(x_axis) as compact wave arrays (i = 0 plane explicit)."""
    u = [np.zeros(o.shape, complex) for _ in range(3)]
    if axis == 0:
        u[comp][1, 0, 0] = amp / 2
    elif axis == 1:
        u[comp][0, 1, 0] = amp / 2
        u[comp][0, o.nkj - 1, 0] = amp / 2
    else:
        u[comp][0, 0, 1] = amp / 2
        u[comp][0, 0, o.nkk - 1] = amp / 2
    return u


def test_rs_stats_single_mode(oracle_mod):
    n, A = 16, 0.8
    o = oracle_mod.Oracle(n)
    o.set_state(*_single_mode(o, 1, 2, A))       # u2 = A cos z: w1 = A sin z
    q = o.rs_stats()
    assert q["rs_u2u2"] == pytest.approx(A * A / 2, rel=1e-14)
    assert q["rs_u1u1"] == 0.0 and q["rs_u3u3"] == 0.0
    assert q["rs_flat2"] == pytest.approx(1.5, rel=1e-13)   # <cos^4> / <cos^2>^2
    assert abs(q["rs_skew2"]) < 1e-13
    assert q["ucfl"] == pytest.approx(A, rel=1e-14)
    assert q["omegamax1"] == pytest.approx(A, rel=1e-14)     # sin = 1 on the grid (n % 4 == 0)
    # the reference leaves rs_rmsvort1 unnormalised: sqrt(sum w1^2) = sqrt(N^3 A^2 / 2)
    assert q["rs_rmsvort1"] == pytest.approx(np.sqrt(n**3 * A * A / 2), rel=1e-13)
    assert q["rs_flatvort1"] == pytest.approx(1.5 / n**6, rel=1e-12)
    o.set_state(*_single_mode(o, 2, 0, A))       # u3 = A cos x: w2 = A sin x, normalised
    q = o.rs_stats()
    assert q["rs_rmsvort2"] == pytest.approx(A / np.sqrt(2), rel=1e-14)
    assert q["rs_flatvort2"] == pytest.approx(1.5, rel=1e-13)


def test_rs_stats_parseval_iso_and_sheared(oracle_mod):
    o = oracle_mod.Oracle(16, nu=0.01)
    u = _random_state(o, 21)
    o.set_state(*u)
    q = o.rs_stats()
    _, fs = o.stats()
    for name in ("u1u1", "u2u2", "u3u3", "u1u2", "u1u3", "u2u3"):
        assert q["rs_" + name] == pytest.approx(fs[name], rel=1e-12, abs=1e-15)
    # Gaussian-like random field: flatness near 3
    for c in (1, 2, 3):
        assert 2.3 < q[f"rs_flat{c}"] < 3.7
    # sheared frame: laboratory velocity (binv(1,3) = 0.7) in both statistics
    s = oracle_mod.Oracle(16, f_type=2, rk3_timestep=2, deform_s=1.0, nu=0.01)
    s.set_state(*u)
    b, bi = s.get_frame()
    b[6], bi[6] = -0.7, 0.7
    s.set_frame(b, bi)
    q = s.rs_stats()
    _, fs = s.stats()
    for name in ("u1u1", "u3u3", "u1u3"):
        assert q["rs_" + name] == pytest.approx(fs[name], rel=1e-12)
    assert q["rs_u1u1"] != pytest.approx(o.rs_stats()["rs_u1u1"], rel=1e-3)


# ---- TS_CORR two-point correlation (f_FluidstatsRs2D / RsCollect) --------------

def _corr_numpy(a, b, n):
    """r1i(s) = sum a(x) b(x + s) / sum a^2 over the grid (x first axis, periodic),
    s = 0 .. (n-1)//2 + 1, with numpy from real-space fields."""
    den = float(np.sum(a * a))
    return np.array([float(np.sum(a * np.roll(b, -s, axis=0))) / den for s in range((n - 1) // 2 + 2)])


def test_rs_correlation_single_mode(oracle_mod):
    n, A = 16, 0.8
    o = oracle_mod.Oracle(n)
    dx = 2.0 * ONEPI / n
    o.set_state(*_single_mode(o, 0, 0, A))        # u1 = A cos x
    r, li = o.rs_correlation()
    s = np.arange(n // 2 + 1)
    assert np.abs(r - np.cos(2 * np.pi * s / n)).max() < 1e-14
    assert li == pytest.approx(dx * np.cos(2 * np.pi * s / n).sum(), rel=1e-13, abs=1e-15)
    o.set_state(*_single_mode(o, 0, 1, A))        # u1 = A cos y: no x dependence
    r, li = o.rs_correlation()
    assert np.abs(r - 1.0).max() < 1e-14
    assert li == pytest.approx(dx * (n // 2 + 1), rel=1e-14)


@pytest.mark.parametrize("n", [16, 15])
def test_rs_correlation_against_numpy(oracle_mod, n):
    """Random field, isotropic and in a sheared frame (binv(1,3) = 0.7, the
    laboratory u1 correlated with the stored u1); odd n: the reference loop's
    last separation (n-1)//2 + 1 enters lint11 but not the NX/2 + 1 array."""
    o = oracle_mod.Oracle(n, nu=0.01)
    u = _random_state(o, 40 + n)
    o.set_state(*u)
    dx = 2.0 * ONEPI / n
    u1, _ = _real_field(o, u[0])
    u3, _ = _real_field(o, u[2])
    for frame in ("iso", "sheared"):
        if frame == "sheared":
            o = oracle_mod.Oracle(n, f_type=2, rk3_timestep=2, deform_s=1.0, nu=0.01)
            o.set_state(*u)
            b, bi = o.get_frame()
            b[6], bi[6] = -0.7, 0.7
            o.set_frame(b, bi)
            a = u1 + 0.7 * u3
        else:
            a = u1
        want = _corr_numpy(a, u1, n)
        r, li = o.rs_correlation()
        assert len(r) == n // 2 + 1
        assert np.abs(r - want[:n // 2 + 1]).max() < 1e-12, frame
        assert li == pytest.approx(dx * want.sum(), rel=1e-12), frame
    # the stored u1 enters at x + s: the laboratory u1 there would give another r1i
    assert np.abs(r - _corr_numpy(a, a, n)[:n // 2 + 1]).max() > 1e-6


# ---- streamwise 1-D spectra (f_FluidstatsFs2D / FsCollect) --------------------

@pytest.mark.parametrize("shear", [False, True])
def test_spectra_1d_parseval(oracle_mod, shear):
    """The oracle transforms the velocity in z and y (zero padding, FFT, /NX) and
    sums 2 |u_lab|^2 over the (y, z) plane for each i.  By Parseval for the
    unnormalised y and z transforms this is 2 NY NZ / NX^2 sum_{j,k} |u_lab|^2,
    computed here from the wave state directly (a non-cubic box pins where each
    length enters; the sheared frame pins the laboratory-frame velocity)."""
    nx, ny, nz = 18, 12, 20
    kw = dict(f_type=2, rk3_timestep=2, deform_s=1.0) if shear else {}
    o = oracle_mod.Oracle(nx=nx, ny=ny, nz=nz, nu=0.01, **kw)
    rng = np.random.default_rng(31)
    u = [rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape) for _ in range(3)]
    o.set_state(*u)
    b13 = 0.0
    if shear:
        b, bi = o.get_frame()
        b[6], bi[6] = -0.45, 0.45
        o.set_frame(b, bi)
        b13 = 0.45
    s1, s2 = o.spectra_1d()
    ul1 = u[0] + b13 * u[2]
    e1 = 2.0 * ny * nz / nx**2 * np.sum(np.abs(ul1) ** 2, axis=(1, 2))
    e2 = 2.0 * ny * nz / nx**2 * np.sum(np.abs(u[1]) ** 2, axis=(1, 2))
    np.testing.assert_allclose(s1, e1, rtol=1e-13)
    np.testing.assert_allclose(s2, e2, rtol=1e-13)


def test_homogeneous_record(oracle_mod):
    """f_FluidstatsHomogeneous: S* = S tau_edy, K* = 2 u1u1 / (u2u2 + u3u3),
    l11 = pi E11(0) / (2 u1u1); ln_int_21, ln_int_12, ln_int_13 are never set
    in the reference, so L* = l11 / 0 and l22i = l33i = 0."""
    o = oracle_mod.Oracle(16, f_type=2, rk3_timestep=2, deform_s=2.5, nu=0.01)
    o.set_state(*_random_state(o, 12))
    _, q = o.stats()
    s1, _ = o.spectra_1d()
    h = o.homogeneous()
    assert h["sstar"] == pytest.approx(2.5 * q["tn_edy"], rel=1e-15)
    assert h["kstar"] == pytest.approx(2 * q["u1u1"] / (q["u2u2"] + q["u3u3"]), rel=1e-15)
    assert h["ln_int_11"] == pytest.approx(ONEPI * s1[0] / (2 * q["u1u1"]), rel=1e-15)
    assert h["lstar"] == float("inf") and h["ln_int_12"] == 0.0 and h["ln_int_13"] == 0.0
    for c in (1, 2, 3):
        assert h[f"fn_ken{c}"] == q[f"fn_ken{c}"] and h[f"fn_int{c}"] == q[f"fn_int{c}"]
        assert h[f"ln_int{c}"] == q[f"ln_int{c}"]
