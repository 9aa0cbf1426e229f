"""Python binding of the MI355X-native PANDORA 2.0 RK3 step (C-ABI in include/pandora_amd.h).

The Fortran drop-in is ``pandora-2.0_amd/fortran/pdr_rk3_amd.f90``; this module binds
the same C entry points with ctypes for tests, benchmarks and scripting.  Method
names follow the reference interfaces they replace:

    ctx.rk3_calc(tstep, simtime)      g_RK3Calc        (src/pg_rk3.f90:82-129)
    ctx.stage(stage, tstep, t, dt)    f_RK3Stage1/2/3  (src/pf_rk3.f90:40,121,161)
    ctx.update_b(tstep, simtime)      g_RK3UpdateB     (src/pg_rk3.f90:280-385)

There is no CPU fallback: if the HIP library is missing or no GPU is visible the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PDR_LIBRARY selects another build of the library (e.g. one compiled with
# different tuning macros); the default is the in-tree build.
LIB_PATH = os.environ.get("PDR_LIBRARY") or os.path.join(_PKG_DIR, "lib", "libpandora_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG_DIR), "include", "pandora_amd.h")

PDR_FP64 = 0
PDR_FP32_MIXED = 1
F_ISOTROPIC, F_SHEAR_DNS, F_SHEAR_RDT_VISCOUS, F_SHEAR_RDT_INVISCID, F_ISOTROPIC_DEFORMED = 1, 2, 3, 4, 5
DEALIASING_TWOTHIRD, DEALIASING_SPHERICAL, DEALIASING_LES = 1, 2, 3
NS_BF_ROT, NS_R_CONV, NS_R_ROT = 1, 2, 3
RK3_TIMESTEP_DECIDE, RK3_TIMESTEP_FIXED, RK3_TIMESTEP_CFL = 1, 2, 3
ERRORS = {0: "PDR_OK", -1: "PDR_ERR_ARG", -2: "PDR_ERR_HIP", -3: "PDR_ERR_NOMEM",
          -4: "PDR_ERR_UNSUPPORTED", -5: "PDR_ERR_IO", -6: "PDR_ERR_STATE", -7: "PDR_ERR_NONFINITE"}

# float32 pi widened to double (src/pg_constants.f90:126)
ONEPI_REF = float(np.float32(np.pi))
TWOPI_REF = 2.0 * ONEPI_REF


class PdrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


# fields of pdr_stats_t and the rows of its spectrum array, in order
STATS_NAMES = ("tke", "u1u1", "u2u2", "u3u3", "u1u2", "u1u3", "u2u3", "f_ken", "f_dis", "f_disskew",
               "fn_ken", "fn_ken1", "fn_ken2", "fn_ken3", "fn_dis", "fn_disskew", "fn_int", "fn_int1",
               "fn_int2", "fn_int3", "tn_tur", "ln_tur", "ln_tay", "rn_tay", "ln_int", "ln_int1",
               "ln_int2", "ln_int3", "tn_edy", "ln_kol", "tn_kol", "disskew", "c_epsilon")
SPEC_NAMES = ("ken", "ken1", "ken2", "ken3", "dis", "disskew", "dens_mod")
# fields of pdr_rs_stats_t, in order
RS_NAMES = ("rs_u1u1", "rs_u2u2", "rs_u3u3", "rs_u1u2", "rs_u1u3", "rs_u2u3", "rs_skew1", "rs_skew2",
            "rs_skew3", "rs_flat1", "rs_flat2", "rs_flat3", "rs_rmsvort1", "rs_rmsvort2", "rs_rmsvort3",
            "rs_skewvort1", "rs_skewvort2", "rs_skewvort3", "rs_flatvort1", "rs_flatvort2",
            "rs_flatvort3", "omegamax1", "omegamax2", "omegamax3", "ucfl")
# fields of pdr_homog_t (the f_FluidstatsHomogeneous record), in order
HOMOG_NAMES = ("sstar", "kstar", "lstar", "ln_int1", "ln_int2", "ln_int3", "ln_int_11", "ln_int_12",
               "ln_int_13", "fn_ken1", "fn_ken2", "fn_ken3", "fn_int1", "fn_int2", "fn_int3")
# fields of pdr_pstats_t (the p_StatsWrite record), in order
PSTATS_NAMES = ("rs_u1u1", "rs_u2u2", "rs_u3u3", "rs_u1u2", "rs_u1u3", "rs_u2u3", "tke", "pken",
                "vrel1", "vrel2", "vrel3", "vrel_abs", "xrel1", "xrel2", "xrel3", "dfp1", "dfp2", "dfp3",
                "x1x2", "x1x3", "x2x3", "vt1", "vt2", "vt3", "vs1", "vs2", "vs3", "rep", "stk", "ste")


class Params(C.Structure):
    _fields_ = [
        ("nodes_x", C.c_int32), ("nodes_y", C.c_int32), ("nodes_z", C.c_int32),
        ("f_type", C.c_int32), ("ns_deformed", C.c_int32), ("dealiasing", C.c_int32),
        ("rk3_timestep", C.c_int32), ("f_remesh_t", C.c_int32), ("b110", C.c_int32),
        ("p_interp_order", C.c_int32), ("precision", C.c_int32), ("force_stop", C.c_int32),
        ("rk3_dt", C.c_double), ("rk3_cfl", C.c_double), ("f_nu", C.c_double),
        ("f_deform_s", C.c_double), ("force_tf", C.c_double), ("force_kf", C.c_double),
        ("force_sigma", C.c_double), ("force_seed", C.c_int64), ("les_cutoff", C.c_double),
        ("b220", C.c_int32), ("b330", C.c_int32),
    ]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run pandora-2.0_amd/build_lib.py)")
    L = C.CDLL(path)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    i32, i64, d = C.c_int32, C.c_int64, C.c_double
    sig = {
        "pdr_version": (C.c_char_p, []),
        "pdr_last_error": (C.c_char_p, []),
        "pdr_create": (C.c_int, [C.POINTER(vp), C.POINTER(Params)]),
        "pdr_destroy": (C.c_int, [vp]),
        "pdr_sizes": (C.c_int, [vp, np.ctypeslib.ndpointer(dtype=np.int64)]),
        "pdr_upload_state": (C.c_int, [vp, dp, dp, dp]),
        "pdr_download_state": (C.c_int, [vp, dp, dp, dp]),
        "pdr_upload_rk3": (C.c_int, [vp, dp, dp, dp]),
        "pdr_download_rk3": (C.c_int, [vp, dp, dp, dp]),
        "pdr_stats": (C.c_int, [vp, dp, dp]),
        "pdr_stats_noalloc": (C.c_int, [vp, dp, dp]),
        "pdr_rs_stats": (C.c_int, [vp, dp]),
        "pdr_rs_correlation": (C.c_int, [vp, dp, C.POINTER(d)]),
        "pdr_stage_correlation": (C.c_int, [vp, C.POINTER(i32), dp, C.POINTER(d)]),
        "pdr_spectra_1d": (C.c_int, [vp, dp, dp, dp]),
        "pdr_set_stage_stats": (C.c_int, [vp, i32]),
        "pdr_stage_stats": (C.c_int, [vp, C.POINTER(i32), dp, dp, dp, dp, dp, dp]),
        "pdr_ensemble_create": (C.c_int, [C.POINTER(vp), C.POINTER(vp), i32]),
        "pdr_ensemble_run": (C.c_int, [vp, i32, i32, dp]),
        "pdr_ensemble_set_split": (C.c_int, [vp, i32, i32]),
        "pdr_ensemble_autotune": (C.c_int, [vp, i32, C.POINTER(i32), C.POINTER(i32)]),
        "pdr_ensemble_destroy": (C.c_int, [vp]),
        "pdr_set_forcing": (C.c_int, [vp, i32, dp, dp, dp]),
        "pdr_get_forcing": (C.c_int, [vp, i32, dp, dp, dp]),
        "pdr_set_frame": (C.c_int, [vp, dp, dp]),
        "pdr_get_frame": (C.c_int, [vp, dp, dp]),
        "pdr_set_dt": (C.c_int, [vp, d]),
        "pdr_get_dt": (C.c_int, [vp, C.POINTER(d)]),
        "pdr_stage": (C.c_int, [vp, i32, i32, d, C.POINTER(d)]),
        "pdr_update_b": (C.c_int, [vp, i32, d]),
        "pdr_remesh": (C.c_int, [vp, i32, C.POINTER(i32)]),
        "pdr_dealias": (C.c_int, [vp]),
        "pdr_step": (C.c_int, [vp, i32, C.POINTER(d)]),
        "pdr_run": (C.c_int, [vp, i32, i32, C.POINTER(d)]),
        "pdr_synchronize": (C.c_int, [vp]),
        "pdr_check_finite": (C.c_int, [vp]),
        "pdr_particles_upload": (C.c_int, [vp, dp, i64]),
        "pdr_particles_download": (C.c_int, [vp, vp, vp]),
        "pdr_particles_download_aux": (C.c_int, [vp, vp, vp]),
        "pdr_particles_upload_aux": (C.c_int, [vp, vp, vp]),
        "pdr_particle_stats_config": (C.c_int, [vp, i32, i32, i32]),
        "pdr_set_particle_sort": (C.c_int, [vp, i32]),
        "pdr_set_tracers_in_stage": (C.c_int, [vp, i32]),
        "pdr_particles_control_rk3": (C.c_int, [vp, i32, i32, d]),
        "pdr_particles_solve": (C.c_int, [vp, i32, i64, i32, d]),
        "pdr_particles_finish_step": (C.c_int, [vp, i32]),
        "pdr_particle_stats": (C.c_int, [vp, C.POINTER(i32), dp]),
        "pdr_init_field": (C.c_int, [vp, C.c_uint64, d, d]),
        "pdr_init_particles": (C.c_int, [vp, C.c_uint64, i64]),
        "pdr_init_field_host": (C.c_int, [i32, i32, i32, C.c_uint64, d, d, dp, dp, dp]),
        "pdr_restart_write_host": (C.c_int, [C.c_char_p, i64, dp, dp, dp, i32]),
        "pdr_restart_read_host": (C.c_int, [C.c_char_p, i64, dp, dp, dp]),
        "pdr_particles_restart_write": (C.c_int, [vp, C.c_char_p, i32]),
        "pdr_particles_restart_read": (C.c_int, [vp, C.c_char_p, i64]),
        "pdr_particles_restart_write_host": (C.c_int, [C.c_char_p, i64, vp, vp, vp, i32]),
        "pdr_particles_restart_read_host": (C.c_int, [C.c_char_p, i64, vp, vp, vp, C.POINTER(i32)]),
        "pdr_restart_write": (C.c_int, [vp, C.c_char_p, i32]),
        "pdr_restart_read": (C.c_int, [vp, C.c_char_p]),
        "pdr_restart_read_scaled_host": (C.c_int, [C.c_char_p, i32, i32, i32, i32, dp, dp, dp]),
        "pdr_restart_read_scaled": (C.c_int, [vp, C.c_char_p, i32]),
        "pdr_spectrum": (C.c_int, [vp, dp, dp, dp, dp]),
        "pdr_debug_prepare": (C.c_int, [vp, i32, dp]),
        "pdr_kernel_forms": (C.c_int, [vp, np.ctypeslib.ndpointer(dtype=np.int32)]),
        "pdr_profile_enable": (C.c_int, [vp, i32]),
        "pdr_copy_bandwidth": (C.c_int, [i64, i32, C.POINTER(d)]),
        "pdr_hbm_probe": (C.c_int, [i32, i64, i32, C.POINTER(d)]),
        "pdr_profile_get": (C.c_int, [vp, i32, C.c_char_p, i32,
                                      np.ctypeslib.ndpointer(dtype=np.int64), dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols_from_header(path: str = HEADER_PATH):
    """Names of every function declared in include/pandora_amd.h."""
    import re
    txt = open(path).read()
    return sorted(set(re.findall(r"\b(pdr_[a-z0-9_]+)\s*\(", txt)))


def _check(rc):
    if rc != 0:
        raise PdrError(rc, load_library().pdr_last_error().decode())


def wave_shape(nx, ny=None, nz=None):
    """(NKI, NKJ, NKK) of the stored 2/3-dealiased modes (src/pg_domain.f90:220-228)."""
    ny = ny or nx
    nz = nz or nx
    return (nx // 3 + 1, 2 * (ny // 3) + 1, 2 * (nz // 3) + 1)


def init_field_host(n, seed, u0, kp, ny=None, nz=None):
    """Seeded Rogallo initial field computed on the host (no GPU): three complex128
    arrays shaped (NKI, NKJ, NKK)."""
    L = load_library()
    ny = ny or n
    nz = nz or n
    shape = wave_shape(n, ny, nz)
    out = [np.zeros(shape, np.complex128) for _ in range(3)]
    _check(L.pdr_init_field_host(n, ny, nz, seed, u0, kp, *[o.view(np.float64).reshape(-1) for o in out]))
    return out


def restart_write_host(path, u1, u2, u3, int64_header=False):
    L = load_library()
    arrs = [np.ascontiguousarray(a, np.complex128).view(np.float64).reshape(-1) for a in (u1, u2, u3)]
    _check(L.pdr_restart_write_host(path.encode(), arrs[0].size, *arrs, 1 if int64_header else 0))


def restart_read_host(path, shape):
    L = load_library()
    out = [np.zeros(shape, np.complex128) for _ in range(3)]
    flat = [o.view(np.float64).reshape(-1) for o in out]
    _check(L.pdr_restart_read_host(path.encode(), flat[0].size, *flat))
    return out


def particles_restart_write_host(path, xp, up, xp0=None, int64_header=False):
    """Particle restart file (g_RestartPetscParticle) from host arrays of shape
    (np, 3): positions, particle_up and, optionally, particle_xp0."""
    L = load_library()
    xp = np.ascontiguousarray(xp, np.float64)
    up = np.ascontiguousarray(up, np.float64)
    x0 = None if xp0 is None else np.ascontiguousarray(xp0, np.float64)
    _check(L.pdr_particles_restart_write_host(path.encode(), len(xp), xp.ctypes.data, up.ctypes.data,
                                              None if x0 is None else x0.ctypes.data, 1 if int64_header else 0))


def particles_restart_read_host(path, count):
    """(xp, up, xp0 or None) of a particle restart file of `count` particles,
    each (count, 3), in particle-number order."""
    L = load_library()
    xp, up, x0 = (np.zeros((count, 3)) for _ in range(3))
    has = C.c_int32(0)
    _check(L.pdr_particles_restart_read_host(path.encode(), count, xp.ctypes.data, up.ctypes.data,
                                             x0.ctypes.data, C.byref(has)))
    return xp, up, (x0 if has.value else None)


def restart_read_scaled_host(path, nx, ny, nz, scale):
    """Scaled restart onto the (nx, ny, nz) grid (F_RESTART_SCALE = scale): u1..u3
    as (NKI, NKJ, NKK) complex arrays of the fine grid."""
    L = load_library()
    shape = (nx // 3 + 1, 2 * (ny // 3) + 1, 2 * (nz // 3) + 1)
    out = [np.zeros(shape, np.complex128) for _ in range(3)]
    flat = [o.view(np.float64).reshape(-1) for o in out]
    _check(L.pdr_restart_read_scaled_host(path.encode(), nx, ny, nz, scale, *flat))
    return out


def copy_bandwidth(nbytes=1 << 30, iters=20):
    """Measured HBM copy bandwidth in GB/s (bytes read + written per second)."""
    g = C.c_double()
    _check(load_library().pdr_copy_bandwidth(nbytes, iters, C.byref(g)))
    return g.value


HBM_PROBES = ("copy", "read", "write", "triad")


def hbm_probe(kind, nbytes=1 << 32, iters=10):
    """Measured HBM rate in GB/s (bytes read + written per second) of one of the
    probe kernels: "copy", "read", "write" or "triad", on nbytes-byte buffers."""
    g = C.c_double()
    _check(load_library().pdr_hbm_probe(HBM_PROBES.index(kind), nbytes, iters, C.byref(g)))
    return g.value


class Pandora:
    """One resident box on one GPU.

    Wave-space arrays are complex128 numpy arrays shaped (NKI, NKJ, NKK), the C-order
    view of the reference's Fortran arr(0:1, k, j, i)."""

    def __init__(self, n=None, *, nx=None, ny=None, nz=None, f_type=F_ISOTROPIC, ns_deformed=1,
                 rk3_timestep=RK3_TIMESTEP_FIXED, rk3_dt=0.005, rk3_cfl=0.75, nu=0.01, deform_s=0.0,
                 remesh_t=0, b110=1, interp_order=4, precision=PDR_FP64, force_tf=0.0, force_stop=0,
                 force_kf=0.0, force_sigma=0.0, force_seed=0, dealiasing=1, les_cutoff=1.0, b220=1, b330=1):
        L = load_library()
        nx = nx or n
        ny = ny or n
        nz = nz or n
        self.params = Params(nx, ny, nz, f_type, ns_deformed, dealiasing, rk3_timestep, remesh_t, b110,
                             interp_order, precision, force_stop, rk3_dt, rk3_cfl, nu, deform_s,
                             force_tf, force_kf, force_sigma, force_seed, les_cutoff, b220, b330)
        self.force_nk = int(force_kf) if (f_type in (F_ISOTROPIC, F_ISOTROPIC_DEFORMED) and force_stop > 0) else -1
        h = C.c_void_p()
        _check(L.pdr_create(C.byref(h), C.byref(self.params)))
        self.h = h
        s = np.zeros(8, dtype=np.int64)
        _check(L.pdr_sizes(self.h, s))
        self.nki, self.nkj, self.nkk, self.mw, self.m1, self.kkmax = (int(v) for v in s[:6])
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (self.nki, self.nkj, self.nkk)
        self.precision = precision
        self._keep_x0 = False

    def close(self):
        if getattr(self, "h", None):
            load_library().pdr_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def L(self):
        return load_library()

    def sizes(self):
        s = np.zeros(8, dtype=np.int64)
        _check(self.L.pdr_sizes(self.h, s))
        return s

    # ---- state -------------------------------------------------------------
    def _wave(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.shape)
        return a.view(np.float64).reshape(-1)

    def set_state(self, u1, u2, u3):
        _check(self.L.pdr_upload_state(self.h, self._wave(u1), self._wave(u2), self._wave(u3)))

    def get_state(self):
        out = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        _check(self.L.pdr_download_state(self.h, *[o.view(np.float64).reshape(-1) for o in out]))
        return out

    def set_rk(self, r1, r2, r3):
        _check(self.L.pdr_upload_rk3(self.h, self._wave(r1), self._wave(r2), self._wave(r3)))

    def get_rk(self):
        out = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        _check(self.L.pdr_download_rk3(self.h, *[o.view(np.float64).reshape(-1) for o in out]))
        return out


    # ---- forcing (fluid_f1..3, src/pf_force.f90:27-33) -----------------------
    def _fshape(self):
        w = 2 * self.force_nk + 1
        return (self.force_nk + 1, w, w)   # C order (i, j, k) = Fortran f(0:1, k, j, i)

    def set_forcing(self, f1, f2, f3):
        a = [np.ascontiguousarray(f, dtype=np.complex128).reshape(self._fshape()).view(np.float64).reshape(-1)
             for f in (f1, f2, f3)]
        _check(self.L.pdr_set_forcing(self.h, self.force_nk, *a))

    def get_forcing(self):
        out = [np.zeros(self._fshape(), dtype=np.complex128) for _ in range(3)]
        v = [o.view(np.float64).reshape(-1) for o in out]
        _check(self.L.pdr_get_forcing(self.h, self.force_nk, *v))
        return out
    def set_frame(self, bmat, binv):
        _check(self.L.pdr_set_frame(self.h, np.ascontiguousarray(bmat, np.float64).reshape(9),
                                    np.ascontiguousarray(binv, np.float64).reshape(9)))

    def get_frame(self):
        b = np.zeros(9)
        bi = np.zeros(9)
        _check(self.L.pdr_get_frame(self.h, b, bi))
        return b, bi

    @property
    def dt(self):
        d = C.c_double()
        _check(self.L.pdr_get_dt(self.h, C.byref(d)))
        return d.value

    @dt.setter
    def dt(self, v):
        _check(self.L.pdr_set_dt(self.h, float(v)))

    # ---- stepping ----------------------------------------------------------
    def rk3_calc(self, tstep: int, simtime: float) -> float:
        """g_RK3Calc: one time step; returns the new simtime."""
        t = C.c_double(simtime)
        _check(self.L.pdr_step(self.h, tstep, C.byref(t)))
        return t.value

    step = rk3_calc

    def run(self, tstep0: int, nsteps: int, simtime: float) -> float:
        t = C.c_double(simtime)
        _check(self.L.pdr_run(self.h, tstep0, nsteps, C.byref(t)))
        return t.value

    def stage(self, stage: int, tstep: int, time: float, dt: float) -> float:
        d = C.c_double(dt)
        _check(self.L.pdr_stage(self.h, stage, tstep, time, C.byref(d)))
        return d.value

    def dealias(self):
        """Spherical / LES truncation of the state (no-op for DEALIASING = 1)."""
        _check(self.L.pdr_dealias(self.h))

    def remesh(self, tstep: int) -> bool:
        """f_RK3Remesh: remesh if this step needs it; returns whether it did."""
        done = C.c_int32(0)
        _check(self.L.pdr_remesh(self.h, tstep, C.byref(done)))
        return bool(done.value)

    def update_b(self, tstep: int, simtime: float):
        _check(self.L.pdr_update_b(self.h, tstep, simtime))

    def synchronize(self):
        _check(self.L.pdr_synchronize(self.h))

    def check_finite(self) -> bool:
        rc = self.L.pdr_check_finite(self.h)
        if rc == -7:
            return False
        _check(rc)
        return True

    # ---- tracers -----------------------------------------------------------
    def set_particles(self, xp):
        xp = np.ascontiguousarray(xp, np.float64).reshape(-1, 3)
        _check(self.L.pdr_particles_upload(self.h, xp.reshape(-1), xp.shape[0]))
        self.np = xp.shape[0]

    def init_particles(self, seed: int, np_: int):
        _check(self.L.pdr_init_particles(self.h, seed, np_))
        self.np = np_

    def get_particles(self):
        n = int(self.sizes()[7])
        xp = np.zeros((n, 3))
        ur = np.zeros((n, 3))
        _check(self.L.pdr_particles_download(self.h, xp.ctypes.data, ur.ctypes.data))
        return xp, ur

    def get_particles_aux(self):
        """(particle_up, particle_xp0 or None)"""
        n = int(self.sizes()[7])
        up = np.zeros((n, 3))
        _check(self.L.pdr_particles_download_aux(self.h, up.ctypes.data, None))
        if not self._keep_x0:
            return up, None
        x0 = np.zeros((n, 3))
        _check(self.L.pdr_particles_download_aux(self.h, None, x0.ctypes.data))
        return up, x0

    def set_particles_aux(self, up=None, xp0=None):
        a = None if up is None else np.ascontiguousarray(up, np.float64).reshape(-1, 3)
        b = None if xp0 is None else np.ascontiguousarray(xp0, np.float64).reshape(-1, 3)
        _check(self.L.pdr_particles_upload_aux(self.h, None if a is None else a.ctypes.data,
                                               None if b is None else b.ctypes.data))

    def set_particle_sort(self, every: int):
        """Re-sort the tracer slots by grid row every `every` steps (0: never)."""
        _check(self.L.pdr_set_particle_sort(self.h, int(every)))

    def set_tracers_in_stage(self, on: bool):
        """True (default): pdr_stage advances the tracers; False: the caller does,
        with control_rk3 / solve_particle and finish_time_step."""
        _check(self.L.pdr_set_tracers_in_stage(self.h, 1 if on else 0))

    def control_rk3(self, stage: int, tstep: int, deltat: float):
        """p_ControlRK3(stage, tstep, deltat) (src/pp_control.f90:518-580)."""
        _check(self.L.pdr_particles_control_rk3(self.h, stage, tstep, float(deltat)))

    def solve_particle(self, form: int, npin: int, stage: int, deltat: float):
        """p_RK3SolveParticle (form 0), ...BF (1), ...R (2) (src/pp_rk3.f90:57, 169, 307)."""
        _check(self.L.pdr_particles_solve(self.h, form, npin, stage, float(deltat)))

    def finish_time_step(self, tstep: int):
        """p_ControlFinishTimeStep(tstep) (src/pp_control.f90:1440-1450)."""
        _check(self.L.pdr_particles_finish_step(self.h, tstep))

    def particle_stats_config(self, enable=True, keep_initial=False, ts_release=0):
        """Tracer statistics in stage 1 (p_StatsPrimary/Dispersion/RelativeVelocity)."""
        _check(self.L.pdr_particle_stats_config(self.h, int(enable), int(keep_initial), int(ts_release)))
        self._keep_x0 = bool(keep_initial)

    def particle_stats(self):
        """(tstep, {name: value}) of the last tracer-statistics record (tstep -1: none)."""
        t = C.c_int32(-1)
        r = np.zeros(len(PSTATS_NAMES))
        _check(self.L.pdr_particle_stats(self.h, C.byref(t), r))
        return int(t.value), dict(zip(PSTATS_NAMES, r.tolist()))

    # ---- initial field / restart / statistics -------------------------------
    def init_field(self, seed: int, u0: float, kp: float):
        _check(self.L.pdr_init_field(self.h, seed, u0, kp))

    def restart_write(self, path: str, int64_header: bool = False):
        _check(self.L.pdr_restart_write(self.h, path.encode(), 1 if int64_header else 0))

    def restart_read(self, path: str):
        _check(self.L.pdr_restart_read(self.h, path.encode()))

    def particles_restart_write(self, path: str, int64_header: bool = False):
        _check(self.L.pdr_particles_restart_write(self.h, path.encode(), 1 if int64_header else 0))

    def particles_restart_read(self, path: str, count: int):
        _check(self.L.pdr_particles_restart_read(self.h, path.encode(), count))

    def restart_read_scaled(self, path: str, scale: int):
        _check(self.L.pdr_restart_read_scaled(self.h, path.encode(), scale))

    def stats(self):
        """f_FluidstatsFs on the device: (spectra dict of kkmax arrays, dict of
        the pdr_stats_t quantities)."""
        q = np.zeros(len(STATS_NAMES))
        spec = np.zeros(7 * self.kkmax)
        _check(self.L.pdr_stats(self.h, q, spec))
        return dict(zip(SPEC_NAMES, spec.reshape(7, self.kkmax))), dict(zip(STATS_NAMES, q))

    def stats_noalloc(self):
        """As stats(), through pdr_stats_noalloc (no device allocation: the
        partials live in the library's fixed device scratch array)."""
        q = np.zeros(len(STATS_NAMES))
        spec = np.zeros(7 * self.kkmax)
        _check(self.L.pdr_stats_noalloc(self.h, q, spec))
        return dict(zip(SPEC_NAMES, spec.reshape(7, self.kkmax))), dict(zip(STATS_NAMES, q))

    def rs_correlation(self):
        """(r1i over separations 0..NX/2, lint11): the TS_CORR two-point
        correlation of the current state on the device."""
        r = np.zeros(self.params.nodes_x // 2 + 1)
        li = C.c_double()
        _check(self.L.pdr_rs_correlation(self.h, r, C.byref(li)))
        return r, li.value

    def stage_correlation(self):
        """(tstep of the record or -1, r1i, lint11) of the stage-1 record."""
        r = np.zeros(self.params.nodes_x // 2 + 1)
        li = C.c_double()
        t = C.c_int32()
        _check(self.L.pdr_stage_correlation(self.h, C.byref(t), r, C.byref(li)))
        return t.value, r, li.value

    def spectra_1d(self):
        """(spec_1i, spec_2i, homogeneous-record dict) of the current state on the device."""
        a = np.zeros(self.nki)
        b = np.zeros(self.nki)
        h = np.zeros(len(HOMOG_NAMES))
        _check(self.L.pdr_spectra_1d(self.h, a, b, h))
        return a, b, dict(zip(HOMOG_NAMES, h.tolist()))

    def set_stage_stats(self, every: int):
        """Statistics of stage 1 inside the step every `every` steps (0: off)."""
        _check(self.L.pdr_set_stage_stats(self.h, int(every)))

    def stage_stats(self):
        """The last stage-1 record: (tstep, stats dict, spectra dict, rs dict,
        spec_1i, spec_2i, homogeneous dict); tstep -1 if none yet."""
        t = C.c_int32(-1)
        q = np.zeros(len(STATS_NAMES))
        spec = np.zeros(7 * self.kkmax)
        rs = np.zeros(len(RS_NAMES))
        a = np.zeros(self.nki)
        b = np.zeros(self.nki)
        h = np.zeros(len(HOMOG_NAMES))
        _check(self.L.pdr_stage_stats(self.h, C.byref(t), q, spec, rs, a, b, h))
        return (int(t.value), dict(zip(STATS_NAMES, q)), dict(zip(SPEC_NAMES, spec.reshape(7, self.kkmax))),
                dict(zip(RS_NAMES, rs)), a, b, dict(zip(HOMOG_NAMES, h.tolist())))

    def rs_stats(self):
        """f_FluidstatsRs2D/RsCollect of the current state on the device (dict of RS_NAMES)."""
        out = np.zeros(len(RS_NAMES))
        _check(self.L.pdr_rs_stats(self.h, out))
        return dict(zip(RS_NAMES, out))

    def spectrum(self):
        ek = np.zeros(self.kkmax)
        dk = np.zeros(self.kkmax)
        ds = np.zeros(self.kkmax)
        tot = np.zeros(6)
        _check(self.L.pdr_spectrum(self.h, ek, dk, ds, tot))
        return ek, dk, ds, tot

    def kernel_forms(self):
        """Kernel forms this box runs (pdr_kernel_forms) as a dict."""
        out = np.zeros(8, dtype=np.int32)
        _check(self.L.pdr_kernel_forms(self.h, out))
        keys = ("specialised", "zchunk", "two_chunks", "fused_zformf", "xrows", "x_tw_global",
                "zinv_register", "zinv_lines")
        return dict(zip(keys, (int(v) for v in out)))

    def debug_prepare(self, field: int):
        out = np.zeros((self.nki, self.nz, self.nkj), np.complex128)
        _check(self.L.pdr_debug_prepare(self.h, field, out.view(np.float64).reshape(-1)))
        return out

    # ---- profiling ---------------------------------------------------------
    def profile(self, enable: bool = True):
        _check(self.L.pdr_profile_enable(self.h, 1 if enable else 0))

    def profile_get(self):
        n = 16
        names = C.create_string_buffer(512)
        cnt = np.zeros(n, dtype=np.int64)
        ms = np.zeros(n)
        _check(self.L.pdr_profile_get(self.h, n, names, 512, cnt, ms))
        keys = names.value.decode().split(",")
        return {k: (int(cnt[i]), float(ms[i])) for i, k in enumerate(keys)}


class Ensemble:
    """Batched ensemble of member contexts (pdr_ensemble_*): one graph launch per
    RK3 step advances every member."""

    def __init__(self, members):
        self.L = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        _check(self.L.pdr_ensemble_create(C.byref(h), arr, len(self.members)))
        self.h = h

    def run(self, tstep0: int, nsteps: int, simtime):
        t = np.ascontiguousarray(simtime, dtype=np.float64).copy()
        _check(self.L.pdr_ensemble_run(self.h, tstep0, nsteps, t))
        return t

    def set_split(self, fpb_y: int, fpb_yf: int):
        _check(self.L.pdr_ensemble_set_split(self.h, fpb_y, fpb_yf))

    def autotune(self, steps: int = 5):
        a, b = C.c_int32(0), C.c_int32(0)
        _check(self.L.pdr_ensemble_autotune(self.h, steps, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            self.L.pdr_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
