"""Scaled restart (F_RESTART_SCALE > 1): a coarse PETSc restart copied onto a finer
grid (f_InitialiseCopyRestart, src/pf_initialise.f90:1184-1299).  The oracle's
copy (oracle.copy_restart) is pinned by the physics it must keep: the fine field
is the same trigonometric polynomial, so on the coarse grid points it equals the
coarse field; the library's host form must then equal the oracle's copy bit for
bit, from a file written in the reference's PETSc binary format."""
import numpy as np
import pytest

import pandora_amd as pa


def _synth(u, n):
    """Real-space field of compact wave arrays (NKI, NKJ, NKK) on an n^3 grid
    (c2r convention: i > 0 counted twice, real part of the i = 0 plane)."""
    nki, nkj, nkk = u.shape
    ki = np.arange(nki)
    kj = np.where(np.arange(nkj) <= nkj // 2, np.arange(nkj), np.arange(nkj) - nkj)
    kk = np.where(np.arange(nkk) <= nkk // 2, np.arange(nkk), np.arange(nkk) - nkk)
    x = 2 * np.pi * np.arange(n) / n
    ex = np.exp(1j * np.outer(x, ki))
    ey = np.exp(1j * np.outer(x, kj))
    ez = np.exp(1j * np.outer(x, kk))
    f = np.einsum("ijk,xi,yj,zk->xyz", u, ex, ey, ez)
    f0 = np.einsum("jk,yj,zk->yz", u[0], ey, ez)
    return 2.0 * f.real - f0.real[None, :, :]


def _coarse_state(n, seed):
    return pa.init_field_host(n, seed, 2.0 / 3.0, 3.0)


@pytest.mark.parametrize("nc,scale", [(12, 2), (10, 3)])
def test_oracle_copy_keeps_the_field_on_the_coarse_points(oracle_mod, nc, scale):
    uc = _coarse_state(nc, 3)
    nf = nc * scale
    uf = oracle_mod.copy_restart(uc, nf, nf, nf, scale)
    for c in range(3):
        a = _synth(uc[c], nc)
        b = _synth(uf[c], nf)[::scale, ::scale, ::scale]
        assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()
        # nothing beyond the coarse band
        assert np.count_nonzero(uf[c]) == np.count_nonzero(uc[c])


def test_host_scaled_restart_equals_the_oracle_copy(oracle_mod, tmp_path):
    uc = _coarse_state(16, 5)
    path = str(tmp_path / "coarse.bin")
    pa.restart_write_host(path, *uc)
    for scale in (2, 3):
        nx, ny, nz = 16 * scale, 16 * scale, 16 * scale
        got = pa.restart_read_scaled_host(path, nx, ny, nz, scale)
        want = oracle_mod.copy_restart(uc, nx, ny, nz, scale)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        # the copied field stays solenoidal on the fine grid
        kj = np.where(np.arange(got[0].shape[1]) <= ny // 3, np.arange(got[0].shape[1]),
                      np.arange(got[0].shape[1]) - got[0].shape[1])
        kk = np.where(np.arange(got[0].shape[2]) <= nz // 3, np.arange(got[0].shape[2]),
                      np.arange(got[0].shape[2]) - got[0].shape[2])
        KI, KJ, KK = np.meshgrid(np.arange(got[0].shape[0]), kj, kk, indexing="ij")
        div = KI * got[0] + KJ * got[1] + KK * got[2]
        assert np.abs(div).max() < 1e-12


def test_scaled_restart_rejects_a_mismatched_file(tmp_path):
    uc = _coarse_state(16, 5)
    path = str(tmp_path / "coarse.bin")
    pa.restart_write_host(path, *uc)
    with pytest.raises(pa.PdrError):
        pa.restart_read_scaled_host(path, 64, 64, 64, 2)  # the file holds 16^3, not 32^3
