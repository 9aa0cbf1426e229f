"""CPU tests of the C-ABI library: it loads, exports what include/pandora_amd.h
declares, validates parameters like the reference, and its host-side routines
(seeded initial field, PETSc-binary restart) are correct.  No GPU is used."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import pandora_amd as pa

TWOPI = 2.0 * float(np.float32(np.pi))
FOURPI = 4.0 * float(np.float32(np.pi))


def test_library_exports_every_header_symbol():
    L = pa.load_library()
    names = pa.exported_symbols_from_header()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", pa.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(L, n)
    assert b"gfx950" in L.pdr_version()


def _params(**kw):
    base = dict(nodes_x=32, nodes_y=32, nodes_z=32, f_type=1, ns_deformed=1, dealiasing=1,
                rk3_timestep=2, f_remesh_t=0, b110=1, p_interp_order=4, precision=0, force_stop=0,
                rk3_dt=0.005, rk3_cfl=0.75, f_nu=0.01, f_deform_s=0.0, force_tf=0.0)
    base.update(kw)
    return pa.Params(**base)


@pytest.mark.parametrize("kw,code", [
    (dict(nodes_x=28, nodes_y=28, nodes_z=28), -4),        # 7 is not a 2/3/5 radix
    (dict(nodes_x=15), -4),                               # odd x length (real transforms)
    (dict(f_type=6), -4),                                 # no such flow type
    (dict(f_type=5, ns_deformed=4), -4),                  # no such NS_DEFORMED form
    (dict(f_type=2, ns_deformed=4), -4),                  # no such NS_DEFORMED form
    (dict(dealiasing=4), -4),                             # no such dealiasing form
    (dict(dealiasing=3, les_cutoff=0.0), -1),             # LES cutoff must be > 0
    (dict(rk3_cfl=1.0), -1),                              # g_RK3Check abort condition
    (dict(force_stop=10, force_tf=0.0, force_kf=2.5), -1),  # OU time scale must be > 0
    (dict(force_stop=10, force_tf=0.2, force_kf=0.0), -1),  # no forced band
    (dict(f_type=2, rk3_timestep=1, f_remesh_t=0, f_deform_s=1.0), -1),
    (dict(p_interp_order=0), -4),                         # P_INTERP_ORDER outside 1..8
    (dict(p_interp_order=9), -4),
    (dict(precision=2), -1),
    (dict(b110=2), -4),                                   # stretches: deformed isotropic only
    (dict(b220=2), -4),
])
def test_create_rejects_invalid_parameters(kw, code):
    L = pa.load_library()
    h = C.c_void_p()
    p = _params(**kw)
    rc = L.pdr_create(C.byref(h), C.byref(p))
    assert rc == code
    assert not h.value
    assert len(L.pdr_last_error()) > 0


def test_null_arguments_are_errors():
    L = pa.load_library()
    assert L.pdr_create(None, None) == -1
    assert L.pdr_destroy(None) == 0


def _kgrid(shape):
    nki, nkj, nkk = shape
    kj = np.arange(nkj)
    kk = np.arange(nkk)
    kj = np.where(kj <= nkj // 2, kj, kj - nkj).astype(float)
    kk = np.where(kk <= nkk // 2, kk, kk - nkk).astype(float)
    return np.meshgrid(np.arange(nki, dtype=float), kj, kk, indexing="ij")


def test_initial_field_is_solenoidal_hermitian_with_exact_mode_energy():
    """Rogallo's construction is an isometry: |u(k)|^2 = E(|k|)/(4 pi |k|^2) for every
    mode off the k1 = k2 = 0 axis (src/pf_initialise.f90:161-219, 1081-1083)."""
    n, u0, kp = 24, 2.0 / 3.0, 4.0
    u = pa.init_field_host(n, 12345, u0, kp)
    KI, KJ, KK = _kgrid(u[0].shape)
    k2 = KI**2 + KJ**2 + KK**2
    kabs = np.sqrt(k2)
    div = KI * u[0] + KJ * u[1] + KK * u[2]
    assert np.abs(div).max() < 1e-15
    mag2 = sum(np.abs(a) ** 2 for a in u)
    ek = 16 * np.sqrt(2 / float(np.float32(np.pi))) * 1.5 * u0 * kabs**4 / kp**5 * np.exp(-2 * k2 / kp**2)
    on_axis = (KI == 0) & (KJ == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        expect = np.where(on_axis, 0.0, ek / (FOURPI * k2))
    assert np.abs(mag2 - expect).max() < 1e-14 * expect.max()
    nkj, nkk = u[0].shape[1:]
    jn = (-np.arange(nkj)) % nkj
    kn = (-np.arange(nkk)) % nkk
    for c in range(3):
        p = u[c][0]
        assert np.array_equal(p, np.conj(p[np.ix_(jn, kn)]))
    # reproducible and seed dependent
    v = pa.init_field_host(n, 12345, u0, kp)
    w = pa.init_field_host(n, 54321, u0, kp)
    assert all(np.array_equal(a, b) for a, b in zip(u, v))
    assert not np.array_equal(u[0], w[0])


@pytest.mark.parametrize("int64_header", [False, True])
def test_restart_format_roundtrip(tmp_path, int64_header):
    """Three PETSc Vec records: int32 1211214, PetscInt length, big-endian float64
    in natural order (src/pg_restart.f90:39-194)."""
    u = pa.init_field_host(16, 3, 2.0 / 3.0, 4.0)
    path = str(tmp_path / "restart.bin")
    pa.restart_write_host(path, *u, int64_header=int64_header)
    raw = open(path, "rb").read()
    n = u[0].size * 2
    hl = 8 if int64_header else 4
    assert len(raw) == 3 * (4 + hl + 8 * n)
    off = 0
    for c in range(3):
        assert int.from_bytes(raw[off:off + 4], "big") == 1211214
        assert int.from_bytes(raw[off + 4:off + 4 + hl], "big") == n
        data = np.frombuffer(raw[off + 4 + hl:off + 4 + hl + 8 * n], dtype=">f8")
        assert np.array_equal(data, u[c].view(np.float64).ravel())
        off += 4 + hl + 8 * n
    back = pa.restart_read_host(path, u[0].shape)
    assert all(np.array_equal(a, b) for a, b in zip(u, back))


def test_restart_rejects_wrong_grid(tmp_path):
    u = pa.init_field_host(16, 3, 2.0 / 3.0, 4.0)
    path = str(tmp_path / "r.bin")
    pa.restart_write_host(path, *u)
    with pytest.raises(pa.PdrError) as e:
        pa.restart_read_host(path, pa.wave_shape(24))
    assert e.value.code == -5
    with pytest.raises(pa.PdrError):
        pa.restart_read_host(str(tmp_path / "missing.bin"), u[0].shape)
