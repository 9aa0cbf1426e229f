"""Lengths above 512 points, where the specialised path switches to its
long-line forms (the z inverse with the velocity re-read per field, the
separate z-forward and FormF kernels, 4-line tiles): 768^3 against the
runtime-radix kernels on the same input, an independent implementation of the
same transforms (FP64, 2 steps of shear with tracers)."""
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


def test_768_specialised_against_runtime_radix(monkeypatch):
    note("768^3 shear FP64 + 100k tracers: specialised vs runtime-radix kernels, 2 steps")
    n = 768
    kw = dict(f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=250, nu=0.005, precision=pa.PDR_FP64)
    res = []
    for generic in ("0", "1"):
        monkeypatch.setenv("PDR_GENERIC_KERNELS", generic)
        g = pa.Pandora(n, **kw)
        g.init_field(4, 2.0 / 3.0, 6.0)
        g.init_particles(104, 100000)
        g.run(1, 2, 0.0)
        assert g.check_finite()
        res.append((g.get_state(), g.get_particles()[0]))
        g.close()
    for a, b in zip(res[0][0], res[1][0]):
        assert rel(a, b) < 1e-12
    d = np.abs(res[0][1] - res[1][1])
    assert np.minimum(d, TWOPI - d).max() < 1e-11
