"""FP32-mixed against FP64 at the benchmark workload: 512^3 homogeneous shear
(Rogallo BF_ROT, S = 1, nu = 0.005, dt = 0.002) with 1,048,576 tracers, 100 steps
from the same seeded field and tracer set.  The FP64 GPU form stands in for the
FP64 oracle at this size (it matches the oracle to 1e-12 at 32-64^3, and the
oracle would need about 40 minutes of CPU for these 100 steps).  Criterion: E(k)
within 1e-4 relative in every shell holding at least 1e-6 of the peak shell
energy; the error of every statistic and the tracer displacement error are
printed."""
import sys

import numpy as np
import pytest

import pandora_amd as pa

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def test_fp32_mixed_vs_fp64_512_shear_100_steps():
    note("FP32-mixed vs FP64, 512^3 shear + 1M tracers, 100 steps")
    n, ntr, steps = 512, 1 << 20, 100
    kw = dict(f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=250, nu=0.005)
    out = {}
    for prec in (pa.PDR_FP64, pa.PDR_FP32_MIXED):
        g = pa.Pandora(n, precision=prec, **kw)
        g.init_field(2, 2.0 / 3.0, 6.0)
        g.init_particles(102, ntr)
        x0 = g.get_particles()[0].copy()
        t = 0.0
        for s in range(1, steps + 1):
            t = g.rk3_calc(s, t)
        sp, q = g.stats()
        out[prec] = (sp, q, g.get_particles()[0], x0, t)
        g.close()
    sp64, q64, x64, x0, t64 = out[pa.PDR_FP64]
    sp32, q32, x32, _, t32 = out[pa.PDR_FP32_MIXED]
    assert t32 == t64
    e64, e32 = sp64["ken"], sp32["ken"]
    m = e64 >= 1e-6 * e64.max()
    err = np.abs(e32[m] - e64[m]) / e64[m]
    dx = 2 * np.pi / n
    d = np.abs(x32 - x64)
    d = np.minimum(d, 2 * np.pi - d)   # periodic
    disp = np.sqrt(np.mean(np.sum(d * d, axis=1))) / dx
    note(f"E(k) max rel err {err.max():.3e} over {m.sum()} shells; tracer rms |x32-x64|/dx {disp:.3e}")
    for name in ("tke", "fn_dis", "rn_tay", "disskew", "u1u3"):
        note(f"  {name}: fp64 {q64[name]:.10e} fp32 {q32[name]:.10e} rel {abs(q32[name]/q64[name]-1):.2e}")
    assert err.max() < 1e-4
    assert disp < 1e-3
