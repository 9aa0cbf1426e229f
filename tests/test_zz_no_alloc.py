"""pdr_stats_noalloc, the statistics entry point that allocates no device
memory (its partials live in the library's fixed device scratch array),
against pdr_stats on the same state: two calls in a row on one GPU, the
second on a smaller box, so the second call reuses the array the first one
left filled."""
import ctypes as C
import sys

import numpy as np
import pytest

import pandora_amd as pa

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _compare(g, tag):
    spec_a, q_a = g.stats()
    spec_b, q_b = g.stats_noalloc()
    for k, v in q_a.items():
        assert abs(q_b[k] - v) <= 1e-12 * max(abs(v), 1e-300), (tag, k, q_b[k], v)
    for k, v in spec_a.items():
        scale = max(np.abs(v).max(), 1e-300)
        assert np.abs(spec_b[k] - v).max() <= 1e-12 * scale, (tag, k)
    # mode counts are integers: exact whatever the summation order
    assert np.array_equal(spec_a["dens_mod"], spec_b["dens_mod"])


@pytest.mark.parametrize("prec", [pa.PDR_FP64, pa.PDR_FP32_MIXED])
def test_stats_noalloc_matches_stats(prec):
    cases = [
        (96, dict(f_type=2, rk3_timestep=1, deform_s=20.0, remesh_t=2, nu=0.005), 2),  # sheared frame
        (48, dict(f_type=1, rk3_dt=0.005, nu=0.01), 1),                                  # smaller, isotropic
    ]
    for n, kw, steps in cases:
        note(f"pdr_stats_noalloc against pdr_stats: N={n} {kw} precision={prec}")
        g = pa.Pandora(n, precision=prec, **kw)
        g.init_field(7 + n, 2.0 / 3.0, 5.0)
        g.run(1, steps, 0.0)
        _compare(g, f"N={n}")
        # a second call on the same state gives the same record
        _compare(g, f"N={n} again")
        g.close()


def test_stats_noalloc_refuses_a_missing_spectrum_buffer():
    g = pa.Pandora(32)
    g.init_field(1, 2.0 / 3.0, 4.0)
    q = np.zeros(len(pa.STATS_NAMES))
    fn = pa.load_library()["pdr_stats_noalloc"]  # a fresh function object: own argtypes
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    assert fn(g.h, q.ctypes.data, None) == -1  # PDR_ERR_ARG
    assert fn(None, q.ctypes.data, None) == -1
    g.close()
