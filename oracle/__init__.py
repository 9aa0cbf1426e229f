"""FP64 CPU oracle for PANDORA 2.0's pseudo-spectral RK3 step.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by ``bench.py``'s cpu_baseline leg, where it is the parity checker (and the
"port" CPU baseline).  The product in ``pandora-2.0_amd/`` never imports it.

The C restatement lives in ``oracle/pdr_oracle.c`` (each routine cites the
reference file:line it follows); this module is a thin ctypes wrapper.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liborc.so")
_lib = None


class OrcParams(C.Structure):
    _fields_ = [
        ("nodes_x", C.c_int32), ("nodes_y", C.c_int32), ("nodes_z", C.c_int32),
        ("f_type", C.c_int32), ("ns_deformed", C.c_int32), ("rk3_timestep", C.c_int32),
        ("rk3_dt", C.c_double), ("rk3_cfl", C.c_double), ("f_nu", C.c_double),
        ("f_deform_s", C.c_double), ("f_remesh_t", C.c_int32), ("b110", C.c_int32),
        ("force_tf", C.c_double), ("p_interp_order", C.c_int32), ("nthreads", C.c_int32),
        ("force_stop", C.c_int32), ("rrot_p3_two", C.c_int32), ("force_kf", C.c_double),
        ("force_sigma", C.c_double), ("force_seed", C.c_int64), ("dealiasing", C.c_int32),
        ("pad1", C.c_int32), ("les_cutoff", C.c_double), ("b220", C.c_int32), ("b330", C.c_int32),
    ]


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [vp]
        L.orc_sizes.argtypes = [vp, np.ctypeslib.ndpointer(dtype=np.int64)]
        for name in ("orc_set_state", "orc_get_state", "orc_set_rk", "orc_get_rk", "orc_nl_forward"):
            getattr(L, name).argtypes = [vp, dp, dp, dp]
        L.orc_set_frame.argtypes = [vp, dp, dp]
        L.orc_get_frame.argtypes = [vp, dp, dp]
        L.orc_set_dt.argtypes = [vp, C.c_double]
        L.orc_get_dt.restype = C.c_double
        L.orc_get_dt.argtypes = [vp]
        L.orc_get_ucfl.restype = C.c_double
        L.orc_get_ucfl.argtypes = [vp]
        L.orc_step.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.orc_stage.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.orc_update_b.argtypes = [vp, C.c_int, C.c_double]
        L.orc_prepare_realspace.argtypes = [vp]
        L.orc_realspace.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.orc_get_1r2w.argtypes = [vp, C.c_int, dp]
        L.orc_remesh.restype = C.c_int
        L.orc_remesh.argtypes = [vp, C.c_int]
        L.orc_dealias.restype = None
        L.orc_dealias.argtypes = [vp]
        L.orc_remesh_target.restype = C.c_int
        L.orc_remesh_target.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_set_particles.argtypes = [vp, dp, C.c_int64]
        L.orc_get_particles.argtypes = [vp, dp, dp]
        L.orc_get_tracer_field.argtypes = [vp, dp, dp, dp]
        L.orc_interpolate.argtypes = [vp, dp, dp, dp, dp, C.c_int64, dp]
        L.orc_correct_position.argtypes = [vp, dp, C.c_int64]
        L.orc_get_particles_aux.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.orc_set_particles_aux.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.orc_particle_stats_config.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.orc_particle_stats.argtypes = [vp, dp]
        L.orc_fft.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_spectrum.argtypes = [vp, dp, dp, dp, dp]
        L.orc_kkmax.restype = C.c_int
        L.orc_kkmax.argtypes = [vp]
        L.orc_stats_raw.argtypes = [vp, dp, dp]
        L.orc_stats_finish.argtypes = [C.c_int, C.c_double, dp, dp, dp, dp]
        L.orc_stats.argtypes = [vp, dp, dp]
        L.orc_rs_raw.argtypes = [vp, dp]
        L.orc_rs_finish.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_rs_stats.argtypes = [vp, dp]
        L.orc_rs_correlation.argtypes = [vp, dp, C.POINTER(C.c_double)]
        L.orc_spectra_1d.argtypes = [vp, dp, dp]
        L.orc_homogeneous.argtypes = [vp, dp]
        L.orc_force_nk.restype = C.c_int
        L.orc_force_nk.argtypes = [vp]
        L.orc_set_forcing.argtypes = [vp, dp, dp, dp]
        L.orc_get_forcing.argtypes = [vp, dp, dp, dp]
        L.orc_force_calc.argtypes = [vp, C.c_int, C.c_double]
        L.orc_normal_pair.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def fft(x: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised complex DFT with exponent sign*2*pi*i*n*k/N (FFTW convention)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    inp = x.view(np.float64).copy()
    out = np.zeros_like(inp)
    lib().orc_fft(len(x), sign, inp, out)
    return out.view(np.complex128)


# quantities of f_FluidstatsFs, in the order the oracle and the library return them
STATS_NAMES = ("tke", "u1u1", "u2u2", "u3u3", "u1u2", "u1u3", "u2u3", "f_ken", "f_dis", "f_disskew",
               "fn_ken", "fn_ken1", "fn_ken2", "fn_ken3", "fn_dis", "fn_disskew", "fn_int", "fn_int1",
               "fn_int2", "fn_int3", "tn_tur", "ln_tur", "ln_tay", "rn_tay", "ln_int", "ln_int1",
               "ln_int2", "ln_int3", "tn_edy", "ln_kol", "tn_kol", "disskew", "c_epsilon")
SPEC_NAMES = ("ken", "ken1", "ken2", "ken3", "dis", "disskew", "dens_mod")
# f_FluidstatsRsCollect quantities, in the order the oracle and the library return them
RS_NAMES = ("rs_u1u1", "rs_u2u2", "rs_u3u3", "rs_u1u2", "rs_u1u3", "rs_u2u3", "rs_skew1", "rs_skew2",
            "rs_skew3", "rs_flat1", "rs_flat2", "rs_flat3", "rs_rmsvort1", "rs_rmsvort2", "rs_rmsvort3",
            "rs_skewvort1", "rs_skewvort2", "rs_skewvort3", "rs_flatvort1", "rs_flatvort2",
            "rs_flatvort3", "omegamax1", "omegamax2", "omegamax3", "ucfl")
# f_FluidstatsHomogeneous record (src/pf_fluidstats.f90:2910-2932)
HOMOG_NAMES = ("sstar", "kstar", "lstar", "ln_int1", "ln_int2", "ln_int3", "ln_int_11", "ln_int_12",
               "ln_int_13", "fn_ken1", "fn_ken2", "fn_ken3", "fn_int1", "fn_int2", "fn_int3")
# p_StatsWrite record (src/pp_stats.f90:379-390)
PSTATS_NAMES = ("rs_u1u1", "rs_u2u2", "rs_u3u3", "rs_u1u2", "rs_u1u3", "rs_u2u3", "tke", "pken",
                "vrel1", "vrel2", "vrel3", "vrel_abs", "xrel1", "xrel2", "xrel3", "dfp1", "dfp2", "dfp3",
                "x1x2", "x1x3", "x2x3", "vt1", "vt2", "vt3", "vs1", "vs2", "vs3", "rep", "stk", "ste")


def stats_finish(kkmax: int, nu: float, raw: np.ndarray, uu: np.ndarray):
    """Normalisation and secondary quantities from raw shell sums."""
    spec = np.zeros(7 * kkmax)
    q = np.zeros(33)
    lib().orc_stats_finish(kkmax, nu, np.ascontiguousarray(raw, np.float64), np.ascontiguousarray(uu, np.float64),
                           spec, q)
    return dict(zip(SPEC_NAMES, spec.reshape(7, kkmax))), dict(zip(STATS_NAMES, q))


def normal_pair(seed: int, key: int):
    """The counter-based N(0,1) pair shared by the oracle and the library."""
    a, b = C.c_double(), C.c_double()
    lib().orc_normal_pair(seed, key, C.byref(a), C.byref(b))
    return a.value, b.value


def remesh_target(kmax: int, i: int, k: int) -> int:
    return lib().orc_remesh_target(kmax, i, k)


def copy_restart(coarse, nx, ny, nz, scale):
    """f_InitialiseCopyRestart (src/pf_initialise.f90:1184-1299) for one process
    owning every mode, with both j halves copied: the fine state is zero except
    i <= KIc, j and k with |j| <= KJc, |k| <= KKc, taken from the coarse arrays at
    the same signed wavenumbers (positive k at k, negative k shifted by
    NODES_KK - nkk_c; likewise j).  coarse: three (NKIc, NKJc, NKKc) arrays."""
    nkjf, nkkf = 2 * (ny // 3) + 1, 2 * (nz // 3) + 1
    kic, kjc, kkc = (nx // scale) // 3, (ny // scale) // 3, (nz // scale) // 3
    nkjc, nkkc = 2 * kjc + 1, 2 * kkc + 1
    kshift, jshift = nkkf - nkkc, nkjf - nkjc
    out = [np.zeros((nx // 3 + 1, nkjf, nkkf), np.complex128) for _ in range(3)]
    for i in range(kic + 1):
        jrows = [(j, j) for j in range(kjc + 1)] + [(j, j - jshift) for j in range(nkjf - kjc, nkjf)]
        for j, jc in jrows:
            for c in range(3):
                out[c][i, j, : kkc + 1] = coarse[c][i, jc, : kkc + 1]
                out[c][i, j, nkkf - kkc:] = coarse[c][i, jc, nkkf - kkc - kshift:]
    return out


class Oracle:
    """FP64 reference-step restatement.  Wave arrays are complex128 numpy
    arrays shaped (NKI, NKJ, NKK) = Fortran arr(0:1, k, j, i) in C order."""

    def __init__(self, n=None, *, nx=None, ny=None, nz=None, f_type=1, ns_deformed=1,
                 rk3_timestep=2, rk3_dt=0.005, rk3_cfl=0.75, nu=0.01, deform_s=0.0,
                 remesh_t=0, b110=1, force_tf=0.0, interp_order=4, nthreads=0, force_stop=0,
                 force_kf=0.0, force_sigma=0.0, force_seed=0, rrot_p3_two=0, dealiasing=1, les_cutoff=1.0,
                 b220=1, b330=1):
        nx = nx or n
        ny = ny or n
        nz = nz or n
        self.p = OrcParams(nx, ny, nz, f_type, ns_deformed, rk3_timestep, rk3_dt, rk3_cfl,
                           nu, deform_s, remesh_t, b110, force_tf, interp_order, nthreads,
                           force_stop, rrot_p3_two, force_kf, force_sigma, force_seed, dealiasing, 0,
                           les_cutoff, b220, b330)
        self.h = lib().orc_create(C.byref(self.p))
        if not self.h:
            raise ValueError(f"oracle: F_TYPE={f_type} with NS_DEFORMED={ns_deformed} is not restated")
        s = np.zeros(6, dtype=np.int64)
        lib().orc_sizes(self.h, s)
        self.nki, self.nkj, self.nkk, self.mw, self.m1, self.kkmax = (int(v) for v in s)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (self.nki, self.nkj, self.nkk)
        self.np = 0
        self._keep_x0 = False

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state -------------------------------------------------------------
    def _wave(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.shape)
        return a.view(np.float64).reshape(-1)

    # ---- statistics (f_FluidstatsFs) -----------------------------------------
    def stats_raw(self):
        raw = np.zeros(7 * self.kkmax)
        uu = np.zeros(6)
        lib().orc_stats_raw(self.h, raw, uu)
        return raw, uu

    def stats(self):
        """(spectra dict of kkmax arrays, dict of the 33 scalar quantities)"""
        spec = np.zeros(7 * self.kkmax)
        q = np.zeros(33)
        lib().orc_stats(self.h, spec, q)
        return dict(zip(SPEC_NAMES, spec.reshape(7, self.kkmax))), dict(zip(STATS_NAMES, q))

    def rs_raw(self):
        raw = np.zeros(25)
        lib().orc_rs_raw(self.h, raw)
        return raw

    def rs_stats(self):
        """f_FluidstatsRs2D/RsCollect of the current state (dict of RS_NAMES)."""
        out = np.zeros(25)
        lib().orc_rs_stats(self.h, out)
        return dict(zip(RS_NAMES, out))

    def rs_correlation(self):
        """(r1i over separations 0..NX/2, lint11): the TS_CORR two-point
        correlation of f_FluidstatsRs2D/RsCollect of the current state."""
        r = np.zeros(self.nx // 2 + 1)
        li = C.c_double()
        lib().orc_rs_correlation(self.h, r, C.byref(li))
        return r, li.value

    def spectra_1d(self):
        """(spec_1i, spec_2i) over i = 0..K (f_FluidstatsFs2D/FsCollect)."""
        a = np.zeros(self.nki)
        b = np.zeros(self.nki)
        lib().orc_spectra_1d(self.h, a, b)
        return a, b

    def homogeneous(self):
        out = np.zeros(len(HOMOG_NAMES))
        lib().orc_homogeneous(self.h, out)
        return dict(zip(HOMOG_NAMES, out.tolist()))

    # ---- forcing -----------------------------------------------------------
    @property
    def force_nk(self) -> int:
        """NK = aint(FORCE_KF), or -1 without forcing."""
        return lib().orc_force_nk(self.h)

    def _fshape(self):
        w = 2 * self.force_nk + 1
        return (self.force_nk + 1, w, w)   # C order (i, j, k) = Fortran f(0:1, k, j, i)

    def set_forcing(self, f1, f2, f3):
        a = [np.ascontiguousarray(f, dtype=np.complex128).reshape(self._fshape()).view(np.float64).reshape(-1)
             for f in (f1, f2, f3)]
        lib().orc_set_forcing(self.h, *a)

    def get_forcing(self):
        out = [np.zeros(self._fshape(), dtype=np.complex128) for _ in range(3)]
        v = [o.view(np.float64).reshape(-1) for o in out]
        lib().orc_get_forcing(self.h, *v)
        return out

    def force_calc(self, tstep: int, dt: float):
        lib().orc_force_calc(self.h, tstep, dt)

    def set_state(self, u1, u2, u3):
        lib().orc_set_state(self.h, self._wave(u1), self._wave(u2), self._wave(u3))

    def get_state(self):
        out = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        lib().orc_get_state(self.h, *[o.view(np.float64).reshape(-1) for o in out])
        return out

    def set_rk(self, r1, r2, r3):
        lib().orc_set_rk(self.h, self._wave(r1), self._wave(r2), self._wave(r3))

    def get_rk(self):
        out = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        lib().orc_get_rk(self.h, *[o.view(np.float64).reshape(-1) for o in out])
        return out

    def set_frame(self, bmat, binv):
        lib().orc_set_frame(self.h, np.ascontiguousarray(bmat, np.float64).reshape(9),
                            np.ascontiguousarray(binv, np.float64).reshape(9))

    def get_frame(self):
        b = np.zeros(9)
        bi = np.zeros(9)
        lib().orc_get_frame(self.h, b, bi)
        return b, bi

    @property
    def dt(self):
        return lib().orc_get_dt(self.h)

    @dt.setter
    def dt(self, v):
        lib().orc_set_dt(self.h, float(v))

    # ---- stepping ----------------------------------------------------------
    def step(self, tstep: int, simtime: float) -> float:
        t = C.c_double(simtime)
        lib().orc_step(self.h, tstep, C.byref(t))
        return t.value

    def stage(self, stage: int, tstep: int, time: float, dt: float) -> float:
        d = C.c_double(dt)
        lib().orc_stage(self.h, stage, tstep, time, C.byref(d))
        return d.value

    def prepare_realspace(self):
        lib().orc_prepare_realspace(self.h)

    def realspace(self, stage: int, tstep: int = 0, dt: float = 0.0) -> float:
        d = C.c_double(dt)
        lib().orc_realspace(self.h, stage, tstep, C.byref(d))
        return d.value

    def nl_forward(self):
        out = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        lib().orc_nl_forward(self.h, *[o.view(np.float64).reshape(-1) for o in out])
        return out

    def get_1r2w(self, field: int) -> np.ndarray:
        out = np.zeros((self.nki, self.nz, self.nkj), np.complex128)
        lib().orc_get_1r2w(self.h, field, out.view(np.float64).reshape(-1))
        return out

    @property
    def ucfl(self):
        return lib().orc_get_ucfl(self.h)

    def remesh(self, tstep: int) -> bool:
        return bool(lib().orc_remesh(self.h, tstep))

    def dealias(self):
        lib().orc_dealias(self.h)

    # ---- tracers -----------------------------------------------------------
    def set_particles(self, xp):
        xp = np.ascontiguousarray(xp, np.float64).reshape(-1, 3)
        self.np = xp.shape[0]
        lib().orc_set_particles(self.h, xp.reshape(-1), self.np)

    def get_particles(self):
        xp = np.zeros((self.np, 3))
        ur = np.zeros((self.np, 3))
        lib().orc_get_particles(self.h, xp.reshape(-1), ur.reshape(-1))
        return xp, ur

    def get_particles_aux(self):
        """(particle_up, particle_xp0 or None)"""
        up = np.zeros((self.np, 3))
        x0 = np.zeros((self.np, 3))
        lib().orc_get_particles_aux(self.h, up.ctypes.data, x0.ctypes.data)
        return up, (x0 if self._keep_x0 else None)

    def set_particles_aux(self, up=None, xp0=None):
        a = None if up is None else np.ascontiguousarray(up, np.float64).reshape(-1, 3)
        b = None if xp0 is None else np.ascontiguousarray(xp0, np.float64).reshape(-1, 3)
        lib().orc_set_particles_aux(self.h, None if a is None else a.ctypes.data,
                                    None if b is None else b.ctypes.data)

    def particle_stats_config(self, enable=True, keep_initial=False, ts_release=0):
        self._keep_x0 = bool(keep_initial)
        lib().orc_particle_stats_config(self.h, int(enable), int(keep_initial), int(ts_release))

    def particle_stats(self):
        """(tstep, {name: value}) of the last tracer-statistics record (tstep -1: none)."""
        r = np.zeros(30)
        t = lib().orc_particle_stats(self.h, r)
        return int(t), dict(zip(PSTATS_NAMES, r.tolist()))

    def tracer_field(self):
        out = [np.zeros((self.nz, self.ny, self.nx)) for _ in range(3)]
        lib().orc_get_tracer_field(self.h, *[o.reshape(-1) for o in out])
        return out

    def interpolate(self, f1, f2, f3, xp):
        xp = np.ascontiguousarray(xp, np.float64).reshape(-1, 3)
        uf = np.zeros_like(xp)
        fs = [np.ascontiguousarray(f, np.float64).reshape(-1) for f in (f1, f2, f3)]
        lib().orc_interpolate(self.h, *fs, xp.reshape(-1), xp.shape[0], uf.reshape(-1))
        return uf

    def correct_position(self, xp):
        xp = np.ascontiguousarray(xp, np.float64).reshape(-1, 3).copy()
        lib().orc_correct_position(self.h, xp.reshape(-1), xp.shape[0])
        return xp

    # ---- statistics --------------------------------------------------------
    def spectrum(self):
        ek = np.zeros(self.kkmax)
        dk = np.zeros(self.kkmax)
        ds = np.zeros(self.kkmax)
        tot = np.zeros(6)
        lib().orc_spectrum(self.h, ek, dk, ds, tot)
        return ek, dk, ds, tot
