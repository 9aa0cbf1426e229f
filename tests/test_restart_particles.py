"""Particle restart file (g_RestartPetscParticle, src/pg_restart.f90:264-318):
the host writer's bytes parsed independently as PETSc Vec records, both header
widths, with and without particle_xp0, and the reader's placement by particle
number.  CPU only (host entry points of the library)."""
import numpy as np
import pytest

import pandora_amd as pa


def _records(path, hl):
    """PETSc Vec records of a file: (classid, n, values) each, parsed with numpy."""
    raw = open(path, "rb").read()
    out, o = [], 0
    while o < len(raw):
        cid = int(np.frombuffer(raw, ">i4", 1, o)[0])
        n = int(np.frombuffer(raw, ">i4" if hl == 4 else ">i8", 1, o + 4)[0])
        vals = np.frombuffer(raw, ">f8", n, o + 4 + hl).astype(np.float64)
        out.append((cid, n, vals))
        o += 4 + hl + 8 * n
    return out


@pytest.mark.parametrize("int64_header", [False, True])
@pytest.mark.parametrize("keep_initial", [False, True])
def test_particle_restart_layout(tmp_path, int64_header, keep_initial):
    rng = np.random.default_rng(5)
    n = 37
    xp, up, x0 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    path = str(tmp_path / "p.bin")
    pa.particles_restart_write_host(path, xp, up, x0 if keep_initial else None, int64_header)
    recs = _records(path, 8 if int64_header else 4)
    assert len(recs) == (11 if keep_initial else 8)
    assert all(cid == 1211214 and m == n for cid, m, _ in recs)
    for c in range(3):
        assert np.array_equal(recs[c][2], xp[:, c])
        assert np.array_equal(recs[3 + c][2], up[:, c])
        if keep_initial:
            assert np.array_equal(recs[8 + c][2], x0[:, c])
    assert np.array_equal(recs[6][2], np.zeros(n))                 # diameter: tracers
    assert np.array_equal(recs[7][2], np.arange(1, n + 1, dtype=float))  # particle numbers
    rx, ru, r0 = pa.particles_restart_read_host(path, n)
    assert np.array_equal(rx, xp) and np.array_equal(ru, up)
    assert (r0 is not None) == keep_initial
    if keep_initial:
        assert np.array_equal(r0, x0)


def test_particle_restart_reader_places_by_number(tmp_path):
    """A file whose records list the particles in another order (as a run on
    several ranks writes them) is read back in particle-number order."""
    rng = np.random.default_rng(6)
    n = 20
    xp, up = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    perm = rng.permutation(n)
    path = str(tmp_path / "p.bin")
    with open(path, "wb") as f:
        cols = [xp[perm, 0], xp[perm, 1], xp[perm, 2], up[perm, 0], up[perm, 1], up[perm, 2],
                np.zeros(n), (perm + 1).astype(float)]
        for col in cols:
            f.write(np.array([1211214, n], ">i4").tobytes())
            f.write(np.asarray(col, ">f8").tobytes())
    rx, ru, r0 = pa.particles_restart_read_host(path, n)
    assert np.array_equal(rx, xp) and np.array_equal(ru, up) and r0 is None


def test_particle_restart_reader_refuses_bad_files(tmp_path):
    n = 8
    path = str(tmp_path / "p.bin")
    xp = np.zeros((n, 3))
    pa.particles_restart_write_host(path, xp, xp)
    with pytest.raises(pa.PdrError):
        pa.particles_restart_read_host(path, n + 1)  # size does not match
    raw = bytearray(open(path, "rb").read())
    o = 7 * (8 + 8 * n) + 8  # first value of the status record
    raw[o:o + 8] = np.array([2.0], ">f8").tobytes()  # particle number 2 twice
    open(path, "wb").write(bytes(raw))
    with pytest.raises(pa.PdrError):
        pa.particles_restart_read_host(path, n)
