"""Particle restart through a context (pdr_particles_restart_write / _read):
the file holds the device's positions, particle_up and particle_xp0 exactly,
and a run continued from the fluid and particle restart files equals the
uninterrupted run bit for bit (particle_up_rk3 is not stored, and the first
stage of a step does not read it)."""
import sys

import numpy as np
import pytest

import pandora_amd as pa

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_particle_restart_continues_the_run(tmp_path, prec):
    n = 32
    note(f"particle restart: shear N={n}, 2000 tracers, keep initial positions, precision={prec}")
    kw = dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005, precision=prec)
    xp = np.random.default_rng(3).uniform(0.0, 6.0, size=(2000, 3))
    g = pa.Pandora(n, **kw)
    g.init_field(9, 2.0 / 3.0, 5.0)
    g.set_particles(xp)
    g.particle_stats_config(True, keep_initial=True)
    t = g.run(1, 3, 0.0)
    fluid, parts = str(tmp_path / "f.bin"), str(tmp_path / "p.bin")
    g.restart_write(fluid)
    g.particles_restart_write(parts)
    x_w, _ = g.get_particles()
    up_w, x0_w = g.get_particles_aux()
    rx, ru, r0 = pa.particles_restart_read_host(parts, len(xp))
    assert np.array_equal(rx, x_w) and np.array_equal(ru, up_w) and np.array_equal(r0, x0_w)
    frame = g.get_frame()
    t_end = g.run(4, 2, t)
    h = pa.Pandora(n, **kw)
    h.restart_read(fluid)
    h.set_frame(*frame)
    h.particle_stats_config(True, keep_initial=True)
    h.particles_restart_read(parts, len(xp))
    up_r, x0_r = h.get_particles_aux()
    assert np.array_equal(up_r, up_w) and np.array_equal(x0_r, x0_w)
    assert h.run(4, 2, t) == t_end
    for a, b in zip(g.get_state(), h.get_state()):
        assert np.array_equal(a, b)
    assert np.array_equal(g.get_particles()[0], h.get_particles()[0])
    assert np.array_equal(g.get_particles_aux()[1], h.get_particles_aux()[1])
    g.close()
    h.close()


def test_particle_restart_needs_particles(tmp_path):
    g = pa.Pandora(32)
    with pytest.raises(pa.PdrError):
        g.particles_restart_write(str(tmp_path / "p.bin"))
    g.close()
