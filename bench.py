"""Benchmark of the MI355X RK3 step on its headline workload.

Workload (N = 1 GPU): 512^3 homogeneous shear in the Rogallo frame (F_TYPE = 2,
NS_BF_ROT, S = 1, nu = 0.005, RK3_TIMESTEP = DECIDE with F_REMESH_T = 250, i.e.
dt = 0.002) with 1,048,576 passive tracers (P_INTERP_ORDER = 4), seeded
type-10 initial field (k0 = 6).  The timed value is the FP32-mixed form; the
FP64 form of the same case is timed beside it.  A "step" is one g_RK3Calc
(three RK stages) with all data resident in HBM.

Prints one JSON line (see the metric/config keys).  --gpus N > 1 is refused:
the path runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pandora-2.0_amd"))

METRIC = "grid-pt RK3 steps/s at 512³ shear, FP32-mixed vs FP64 GPU; % HBM peak"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def model_bytes_per_step(g, csz, rsz, ntr):
    """Algorithmic HBM bytes of one RK3 step in the reference's slab dataflow
    (compact wave state + 1r2w intermediates, x-y planes on chip):
    3 * ((15 Mw + 18 M1) * s_c + 96 N_tr + 3 N^3 s_r)."""
    n3 = g.nx * g.ny * g.nz
    return 3 * ((15 * g.mw + 18 * g.m1) * csz + (96 * ntr + 3 * n3 * rsz if ntr else 0))


def kernel_bytes(g, name, csz, rsz, ntr):
    """Compulsory bytes one launch of each kernel must move (its inputs + outputs)."""
    m2 = g.nz * g.ny * g.nki
    n3 = g.nx * g.ny * g.nz
    return {
        "zinv": (3 * g.mw + 6 * g.m1) * csz,
        "yinv": (6 * g.m1 + 6 * m2) * csz,
        "xpass": 9 * m2 * csz + (3 * n3 * rsz if ntr else 0),
        "yfwd": (3 * m2 + 3 * g.m1) * csz,
        "zfwd": (3 * g.m1 + 3 * g.mw) * csz,
        "formf": 15 * g.mw * csz,
        "zfwdformf": (3 * g.m1 + 12 * g.mw) * csz,
        "tracers": 96 * ntr + 3 * n3 * rsz,
    }.get(name)


# The y inverse, x pass and y forward run chunk by chunk over z when the box is
# split (from 512 planes up two half-box chunks are in flight on two streams):
# a launch then moves its chunk's share of the bytes.
CHUNKED = ("yinv", "xpass", "yfwd")


def chunks_per_stage(name, launches, steps):
    return max(1.0, launches / (3.0 * steps)) if name in CHUNKED else 1.0


def run_gpu(pa, n, precision, steps, warmup, ntr, profile):
    g = pa.Pandora(n, f_type=pa.F_SHEAR_DNS, rk3_timestep=pa.RK3_TIMESTEP_DECIDE, deform_s=1.0,
                   remesh_t=250, nu=0.005, precision=precision)
    g.init_field(2, 2.0 / 3.0, 6.0)
    if ntr:
        g.init_particles(102, ntr)
    t = 0.0
    if warmup:
        t = g.run(1, warmup, t)
    g.synchronize()
    if profile:
        g.profile(True)
    t0 = time.perf_counter()
    t = g.run(1 + warmup, steps, t)
    g.synchronize()
    t1 = time.perf_counter()
    prof = g.profile_get() if profile else {}
    ok = g.check_finite()
    g.close()
    return (t1 - t0), prof, ok


# The large single-box cases, run on request (--case): forced isotropic
# turbulence at 1024^3 and homogeneous shear at 1536^3, each FP32-mixed with its
# tracers, then the FP64 GPU form of the same case for the error of E(k),
# epsilon and the tracer positions (the two forms do not fit on the GPU at once,
# so they run one after the other from the same seeded field).
CASES = {
    "forced1024": dict(n=1024, tracers=16_777_216, seed=3, tseed=103, kp=3.0,
                       kw=dict(f_type=1, rk3_timestep=2, rk3_dt=0.0005, nu=0.0008, force_tf=0.2,
                               force_kf=2.5, force_sigma=0.5**0.5, force_stop=100000, force_seed=3),
                       label="1024^3 forced isotropic (OU |k|<=2.5, T_F=0.2, sigma^2 T_F=0.1), nu=0.0008, "
                             "dt=0.0005, 16,777,216 tracers"),
    "shear1536": dict(n=1536, tracers=67_108_864, seed=4, tseed=104, kp=6.0,
                      kw=dict(f_type=2, rk3_timestep=1, deform_s=1.0, remesh_t=2500, nu=0.0003),
                      label="1536^3 homogeneous shear, S=1, nu=0.0003, dt=0.0002 (F_REMESH_T=2500), "
                            "67,108,864 tracers"),
}


def run_case_prec(pa, name, tag, steps, warmup, out_path):
    """One precision of a large case, in its own process (the two forms of the
    1536^3 case do not fit on the GPU together, and a fresh process starts with
    all of the device memory free): timings, E(k), epsilon and the tracer
    positions go to out_path (.npz)."""
    import numpy as np
    c = CASES[name]
    n = c["n"]
    prec = pa.PDR_FP32_MIXED if tag == "fp32" else pa.PDR_FP64
    print(f"[bench] {name} {tag}: init", file=sys.stderr, flush=True)
    g = pa.Pandora(n, precision=prec, **c["kw"])
    g.init_field(c["seed"], 2.0 / 3.0, c["kp"])
    ntr = c["tracers"] if tag == "fp32" else c.get("fp64_tracers", c["tracers"])
    if ntr:
        g.init_particles(c["tseed"], ntr)
    t = g.run(1, warmup, 0.0) if warmup else 0.0
    g.synchronize()
    g.profile(True)
    print(f"[bench] {name} {tag}: {steps} timed steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    t = g.run(1 + warmup, steps, t)
    g.synchronize()
    dt = time.perf_counter() - t0
    prof = g.profile_get()
    spec, q = g.stats()
    xp = g.get_particles()[0] if ntr else np.zeros((0, 3))
    summary = {"steps_per_s": steps / dt, "grid_point_steps_per_s": steps / dt * n**3, "tracers": ntr,
               "kernels": {k: {"launches": v[0], "avg_ms": v[1] / v[0]} for k, v in prof.items() if v[0]},
               "finite": bool(g.check_finite()), "device_gib": int(g.sizes()[6]) / 2**30}
    g.close()
    np.savez(out_path, ek=spec["ken"], eps=np.array([q["fn_dis"]]), xp=xp, summary=json.dumps(summary))


def run_case(name, steps, warmup):
    import subprocess
    import tempfile

    import numpy as np
    c = CASES[name]
    n = c["n"]
    out = {"case": name, "workload": c["label"], "steps": steps, "warmup": warmup}
    data = {}
    with tempfile.TemporaryDirectory() as tmp:
        for tag in ("fp32", "fp64"):
            path = os.path.join(tmp, f"{tag}.npz")
            subprocess.run([sys.executable, os.path.abspath(__file__), "--case", name, "--case-prec", tag,
                            "--case-out", path, "--steps", str(steps), "--warmup", str(warmup)], check=True)
            with np.load(path) as z:
                data[tag] = {"ek": z["ek"], "eps": float(z["eps"][0]), "xp": z["xp"],
                             "summary": json.loads(str(z["summary"]))}
            out[tag] = data[tag]["summary"]
    e32, e64 = data["fp32"]["ek"], data["fp64"]["ek"]
    big = e64 > 1e-6 * e64.max()
    out["fp32_over_fp64"] = out["fp32"]["steps_per_s"] / out["fp64"]["steps_per_s"]
    out["ek_max_rel_err"] = float(np.max(np.abs(e32[big] - e64[big]) / e64[big]))
    out["eps_rel_err"] = abs(data["fp32"]["eps"] - data["fp64"]["eps"]) / abs(data["fp64"]["eps"])
    out["tracer_rms_err_dx"] = None
    if len(data["fp32"]["xp"]) and len(data["fp32"]["xp"]) == len(data["fp64"]["xp"]):
        d = np.abs(data["fp32"]["xp"] - data["fp64"]["xp"])
        L = 2.0 * float(np.float32(np.pi))
        d = np.minimum(d, L - d) / (L / n)
        out["tracer_rms_err_dx"] = float(np.sqrt(np.mean(np.sum(d * d, axis=1))))
    return out


ENS_SIZES = (32, 48, 64, 96)


def ensemble_bench(pa, per_size=16, steps=50, warmup=5):
    """The batching form: 16 decaying isotropic members each of 32^3, 48^3, 64^3
    and 96^3 (k0 = 4, nu = 0.01, dt = 0.005, FP32-mixed, seeds 0..63), advanced by
    one graph launch per RK3 step; timed beside stepping the same members one by
    one with pdr_run."""
    sizes = [n for n in ENS_SIZES for _ in range(per_size)]
    members = []
    for k, n in enumerate(sizes):
        g = pa.Pandora(n, rk3_timestep=pa.RK3_TIMESTEP_FIXED, rk3_dt=0.005, nu=0.01, precision=pa.PDR_FP32_MIXED)
        g.init_field(k, 2.0 / 3.0, 4.0)
        members.append(g)
    e = pa.Ensemble(members)
    split = e.autotune(3)
    t = e.run(1, warmup, [0.0] * len(sizes))
    t0 = time.perf_counter()
    t = e.run(1 + warmup, steps, t)
    t1 = time.perf_counter()
    ens_sps = steps / (t1 - t0)
    seq_steps = 5
    for g in members:
        g.synchronize()
    s0 = time.perf_counter()
    for g in members:
        g.run(1, seq_steps, 0.0)
    for g in members:
        g.synchronize()
    s1 = time.perf_counter()
    seq_sps = seq_steps / (s1 - s0)
    pts = sum(n**3 for n in sizes)

    class G:
        pass

    b = 0
    for n in sizes:
        g = G()
        g.nx = g.ny = g.nz = n
        g.nki, g.nkj, g.nkk = n // 3 + 1, 2 * (n // 3) + 1, 2 * (n // 3) + 1
        g.mw = g.nki * g.nkj * g.nkk
        g.m1 = g.nki * g.nkj * n
        b += model_bytes_per_step(g, 8, 4, 0)
    e.close()
    for g in members:
        g.close()
    return {"members": f"{per_size} each of " + ", ".join(f"{n}^3" for n in ENS_SIZES) + " (FP32-mixed)",
            "ensemble_steps_per_s": ens_sps, "grid_point_steps_per_s": ens_sps * pts,
            "one_by_one_steps_per_s": seq_sps, "batching_speedup": ens_sps / seq_sps,
            "autotuned_split": list(split), "model_bytes_per_step": b,
            "model_frac_of_peak": b * ens_sps / HBM_PEAK, "steps": steps}


def host_info():
    """The host CPU as /proc describes it: model, logical CPUs, the CPUs this
    process may run on, physical cores and memory."""
    info = {"logical_cpus": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        cores = set()
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and "cpu_model" not in info:
                    info["cpu_model"] = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
        if cores:
            info["physical_cores"] = len(cores)
    except OSError:
        pass
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    info["mem_gib"] = round(int(line.split()[1]) / 2**20, 1)
    except OSError:
        pass
    return info


def cpu_threads():
    """OpenMP threads of the CPU leg: OMP_NUM_THREADS when the environment sets
    it (the share of host cores given to this process), else every CPU this
    process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def _median(v):
    v = sorted(v)
    m = len(v) // 2
    return v[m] if len(v) % 2 else 0.5 * (v[m - 1] + v[m])


def cpu_baseline(n, ntr, small=256, small_tracers=262144, small_steps=5):
    """The reference's CPU path restated as the FP64 C oracle (OpenMP over the
    host cores), on the host of the GPU run:
    * the headline workload (n^3 shear, ntr tracers): one warm-up stage, then
      the three stages of one step, each timed; value = n^3 / their sum;
    * the FP64 drop-in workload (small^3 shear, small_tracers tracers): one
      warm-up step, then the median of small_steps timed steps with the
      statistics off, and of small_steps steps with the statistics of stage 1
      on (f_FluidstatsFs and the real-space moments before each step)."""
    sys.path.insert(0, ROOT)
    import numpy as np

    import oracle
    import pandora_amd as pa

    oracle.build()
    threads = cpu_threads()
    kw = dict(f_type=2, rk3_timestep=1, remesh_t=250, deform_s=1.0, nu=0.005, nthreads=threads)

    def make(nn, count, seed):
        u = pa.init_field_host(nn, 2, 2.0 / 3.0, 6.0)
        o = oracle.Oracle(nn, **kw)
        o.set_state(*u)
        del u
        rng = np.random.default_rng(seed)
        dx = pa.TWOPI_REF / nn
        o.set_particles(rng.uniform(-0.5 * dx, pa.TWOPI_REF - 0.5 * dx, size=(count, 3)))
        return o

    print(f"[bench] CPU leg: {n}^3 stages on {threads} threads", file=sys.stderr, flush=True)
    o = make(n, ntr, 102)
    dt = o.stage(1, 1, 0.0, 0.0)  # warm-up
    stage_s = []
    for stage, tstep in ((2, 1), (3, 1), (1, 2)):
        t0 = time.perf_counter()
        o.stage(stage, tstep, 0.0, dt)
        stage_s.append(time.perf_counter() - t0)
    o.close()
    step_s = sum(stage_s)
    out = {"value": n**3 / step_s, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
           "sample": f"{n}^3 shear + {ntr} tracers on the FP64 C oracle, {threads} OpenMP threads: one "
                     f"warm-up stage, then the three stages of one step ({', '.join(f'{t:.2f}' for t in stage_s)} s)",
           "host": host_info()}
    if small:
        print(f"[bench] CPU leg: {small}^3 steps, statistics off and on", file=sys.stderr, flush=True)
        o = make(small, small_tracers, 101)
        t = o.step(1, 0.0)  # warm-up
        runs = {"stats_off": [], "stats_on": []}
        s = 2
        for key in ("stats_off", "stats_on"):
            for _ in range(small_steps):
                t0 = time.perf_counter()
                if key == "stats_on":
                    o.stats()
                    o.rs_stats()
                t = o.step(s, t)
                runs[key].append(time.perf_counter() - t0)
                s += 1
        o.close()
        out["drop_in_256"] = {
            "workload": f"{small}^3 shear, {small_tracers} tracers, FP64, {threads} OpenMP threads",
            **{k: {"median_s_per_step": _median(v), "steps_per_s": 1.0 / _median(v),
                   "grid_point_steps_per_s": small**3 / _median(v), "samples_s": v} for k, v in runs.items()}}
    return out


def hbm_ceiling(pa, sizes_gib=(1, 4, 8), iters=10):
    """Measured HBM rates of the probe kernels (copy, read, write, triad) over
    several buffer sizes; the highest is the measured ceiling."""
    probes = {}
    best, where = 0.0, None
    for gib in sizes_gib:
        for kind in pa.HBM_PROBES:
            r = pa.hbm_probe(kind, gib << 30, iters)
            probes[f"{kind}_{gib}GiB"] = r
            if r > best:
                best, where = r, f"{kind}_{gib}GiB"
    return {"GBps": best, "best": where, "probes_GBps": probes}


def stats_on_off(pa, n, ntr, steps=5, repeats=3):
    """Steps/s of the FP32-mixed headline workload with the statistics of stage 1
    taken inside every step (f_FluidstatsFs of the starting state and the
    real-space moments of stage 1's planes, pdr_set_stage_stats(1)) and without,
    each the median of `repeats` timed runs of `steps` steps in one context."""
    out = {}
    for every in (0, 1):
        g = pa.Pandora(n, f_type=pa.F_SHEAR_DNS, rk3_timestep=pa.RK3_TIMESTEP_DECIDE, deform_s=1.0,
                       remesh_t=250, nu=0.005, precision=pa.PDR_FP32_MIXED)
        g.init_field(2, 2.0 / 3.0, 6.0)
        if ntr:
            g.init_particles(102, ntr)
        g.set_stage_stats(every)
        t = g.run(1, 2, 0.0)
        s = 3
        rates = []
        for _ in range(repeats):
            g.synchronize()
            t0 = time.perf_counter()
            t = g.run(s, steps, t)
            g.synchronize()
            rates.append(steps / (time.perf_counter() - t0))
            s += steps
        g.close()
        out["stats_on" if every else "stats_off"] = {"median_steps_per_s": _median(rates), "samples": rates}
    out["cost"] = 1.0 - out["stats_on"]["median_steps_per_s"] / out["stats_off"]["median_steps_per_s"]
    return out


def zstreams_ab(pa, n, ntr, steps=5, repeats=3):
    """Steps/s of the FP32-mixed headline workload with the z chunks of a stage
    one after another (PDR_ZSTREAMS=1) and two in flight (2), each the median
    of `repeats` timed runs of `steps` steps in one context."""
    env = os.environ.get("PDR_ZSTREAMS")
    out = {}
    try:
        for zs in ("1", "2"):
            os.environ["PDR_ZSTREAMS"] = zs
            g = pa.Pandora(n, f_type=pa.F_SHEAR_DNS, rk3_timestep=pa.RK3_TIMESTEP_DECIDE, deform_s=1.0,
                           remesh_t=250, nu=0.005, precision=pa.PDR_FP32_MIXED)
            g.init_field(2, 2.0 / 3.0, 6.0)
            if ntr:
                g.init_particles(102, ntr)
            t = g.run(1, 2, 0.0)
            s = 3
            rates = []
            for _ in range(repeats):
                g.synchronize()
                t0 = time.perf_counter()
                t = g.run(s, steps, t)
                g.synchronize()
                rates.append(steps / (time.perf_counter() - t0))
                s += steps
            g.close()
            out[f"zstreams_{zs}"] = {"median_steps_per_s": _median(rates), "min": min(rates), "max": max(rates),
                                     "samples": rates}
    finally:
        if env is None:
            os.environ.pop("PDR_ZSTREAMS", None)
        else:
            os.environ["PDR_ZSTREAMS"] = env
    a, b = out["zstreams_1"], out["zstreams_2"]
    out["gain"] = b["median_steps_per_s"] / a["median_steps_per_s"] - 1.0
    out["spread"] = max((a["max"] - a["min"]) / a["median_steps_per_s"], (b["max"] - b["min"]) / b["median_steps_per_s"])
    out["steps"], out["repeats"] = steps, repeats
    return out


# ---------------------------------------------------------------------------
# rocprofv3 summaries (kernel trace + FETCH_SIZE / WRITE_SIZE counter runs of
# this script), written to profiles/; the traffic JSON feeds roofline.traffic.
# ---------------------------------------------------------------------------
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "rk3_step_512_shear_fp32_traffic.json")
KERNEL_CLASS = {"zinv": "zinv", "zinv_r": "zinv", "yinv": "yinv", "xpass": "xpass", "yfwd": "yfwd",
                "zfwd": "zfwd", "formf": "formf", "zfwd_formf": "zfwdformf", "tracers": "tracers",
                "tracers_col": "tracers", "copy": "copy"}


def _kernel_key(name):
    import re
    m = re.search(r"(k2?_[a-z0-9_]+)", name)
    if not m:
        return None, name[:40]
    base = m.group(1)
    cls = KERNEL_CLASS.get(base.split("_", 1)[1])
    return cls, base


def rocpd_summary(trace_db, fetch_db, write_db, out_json, sq_dbs=(), copy_bytes=1 << 30):
    """Markdown tables of a kernel trace, of per-dispatch FETCH_SIZE /
    WRITE_SIZE (KiB), calibrated by the copy kernel k_copy (copy_bytes read and
    written per launch, bench.py's bandwidth probe), and of the SQ_ counters of
    the sq_dbs runs (per-dispatch sums, averaged over dispatches)."""
    import collections
    import sqlite3

    import re

    rows = sqlite3.connect(trace_db).execute("select name, duration from kernels").fetchall()
    agg = collections.OrderedDict()
    for name, dur in rows:
        cls, base = _kernel_key(name)
        m = re.search(r"k2?_[a-z0-9_]+<(float|double)", name)
        if m:
            base += "<f32>" if m.group(1) == "float" else "<f64>"
        if re.search(r"k2_xpass<[a-z]+, \d+, true>", name):
            base += " with moments"  # the stage-1 x pass of a step with statistics
        a = agg.setdefault(base, [0, 0.0, 1e30, 0.0])
        a[0] += 1
        a[1] += dur
        a[2] = min(a[2], dur)
        a[3] = max(a[3], dur)
    tot = sum(a[1] for a in agg.values())
    print("| kernel | calls | total ms | avg ms | min ms | max ms | % |")
    print("|---|---|---|---|---|---|---|")
    for base, a in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| {base} | {a[0]} | {a[1] / 1e6:.3f} | {a[1] / a[0] / 1e6:.4f} | {a[2] / 1e6:.4f} | "
              f"{a[3] / 1e6:.4f} | {100 * a[1] / tot:.1f} |")

    def per_dispatch(db, counter):
        con = sqlite3.connect(db)
        vals = collections.defaultdict(lambda: collections.defaultdict(float))
        for kn, cn, did, v in con.execute(
                "select kernel_name, counter_name, dispatch_id, value from counters_collection"):
            if cn == counter:
                vals[_kernel_key(kn)[1]][did] += float(v)
        return {k: sum(d.values()) / len(d) for k, d in vals.items()}

    def counters(db):
        con = sqlite3.connect(db)
        return sorted({r[0] for r in con.execute("select distinct counter_name from counters_collection")})

    if not (fetch_db and write_db):  # kernel trace only
        return
    fetch = per_dispatch(fetch_db, "FETCH_SIZE")
    write = per_dispatch(write_db, "WRITE_SIZE")
    ff = copy_bytes / (fetch["k_copy"] * 1024.0)
    wf = copy_bytes / (write["k_copy"] * 1024.0)
    print()
    print(f"calibration (k_copy, {copy_bytes} B in and out): FETCH_SIZE x{ff:.3f}, WRITE_SIZE x{wf:.3f}")
    print()
    print("| kernel | FETCH_SIZE KiB | read GB | WRITE_SIZE KiB | write GB | traffic GB |")
    print("|---|---|---|---|---|---|")
    kern = {}
    for base in sorted(set(fetch) | set(write)):
        cls, _ = _kernel_key(base)
        fk, wk = fetch.get(base, 0.0), write.get(base, 0.0)
        rb, wb = fk * 1024 * ff, wk * 1024 * wf
        print(f"| {base} | {fk:,.0f} | {rb / 1e9:.3f} | {wk:,.0f} | {wb / 1e9:.3f} | {(rb + wb) / 1e9:.3f} |")
        if cls and cls != "copy":
            kern[cls] = {"kernel": base, "fetch_kib": fk, "write_kib": wk, "read_bytes": rb,
                         "write_bytes": wb, "traffic_bytes": rb + wb}
    sq = collections.defaultdict(dict)
    for db in sq_dbs:
        for cn in counters(db):
            for base, v in per_dispatch(db, cn).items():
                sq[base][cn] = v
    if sq:
        names = sorted({c for d in sq.values() for c in d})
        print()
        print("| kernel | " + " | ".join(names) + " |")
        print("|---|" + "---|" * len(names))
        for base in sorted(sq):
            print(f"| {base} | " + " | ".join(f"{sq[base][c]:.3g}" if c in sq[base] else "-" for c in names)
                  + " |")
    if out_json:
        with open(out_json, "w") as f:
            json.dump({"workload": {"n": 512, "tracers": 1 << 20, "precision": "fp32-mixed"},
                       "calibration": {"fetch_factor": ff, "write_factor": wf, "copy_bytes": copy_bytes},
                       "kernels": kern}, f, indent=1)


def measured_traffic(kernel, n, ntr):
    """HBM bytes per launch of `kernel` from the committed counter summary, if
    it was measured on this workload."""
    try:
        with open(TRAFFIC_JSON) as f:
            d = json.load(f)
    except OSError:
        return None
    w = d.get("workload", {})
    if w.get("n") != n or w.get("tracers") != ntr or kernel not in d.get("kernels", {}):
        return None
    return d["kernels"][kernel]["traffic_bytes"]


def main():
    if "--rocpd-summary" in sys.argv:
        sp = argparse.ArgumentParser()
        sp.add_argument("--rocpd-summary", action="store_true")
        sp.add_argument("--trace", required=True)
        sp.add_argument("--fetch", default=None)
        sp.add_argument("--write", default=None)
        sp.add_argument("--json", default=None)
        sp.add_argument("--sq", nargs="*", default=[])
        a = sp.parse_args()
        rocpd_summary(a.trace, a.fetch, a.write, a.json, a.sq)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--tracers", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble run")
    ap.add_argument("--no-ab", action="store_true",
                    help="skip the one/two chunk-stream and the statistics on/off comparisons")
    ap.add_argument("--no-cpu-small", action="store_true",
                    help="CPU leg: skip the 256^3 statistics off/on medians")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the batched-ensemble measurement and print its JSON object")
    ap.add_argument("--case", choices=sorted(CASES), default=None,
                    help="run one large single-box case (FP32-mixed and FP64) and print its JSON object")
    ap.add_argument("--case-prec", choices=("fp32", "fp64"), default=None, help=argparse.SUPPRESS)
    ap.add_argument("--case-out", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this path runs on one GPU (the whole box stays resident on one "
              "MI355X); multi-GPU runs are not supported", file=sys.stderr)
        sys.exit(2)
    if args.case and not args.case_prec:  # the parent only starts the two children
        print(json.dumps(run_case(args.case, args.steps, args.warmup)), flush=True)
        return
    import pandora_amd as pa

    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(pa)}), flush=True)
        return
    if args.case and args.case_prec:
        run_case_prec(pa, args.case, args.case_prec, args.steps, args.warmup, args.case_out)
        return
    n, ntr = args.n, args.tracers
    print(f"[bench] FP32-mixed {n}^3 shear + {ntr} tracers: {args.warmup} warmup, {args.steps} timed",
          file=sys.stderr, flush=True)
    # the headline: no per-kernel events on the streams
    dt32, _, ok32 = run_gpu(pa, n, pa.PDR_FP32_MIXED, args.steps, args.warmup, ntr, False)
    sps32 = args.steps / dt32
    # Per-kernel times from a short profiled run with the chunks of a stage one
    # after another (the default at this size), whose launch durations are each
    # kernel's own: the rates and the roofline come from it.
    rsteps = max(3, args.steps // 4)
    print(f"[bench] FP32-mixed {n}^3, profiled: {rsteps} steps, chunks one after another",
          file=sys.stderr, flush=True)
    env_zs = os.environ.get("PDR_ZSTREAMS")
    os.environ["PDR_ZSTREAMS"] = "1"
    try:
        _, prof, _ = run_gpu(pa, n, pa.PDR_FP32_MIXED, rsteps, 1, ntr, True)
    finally:
        if env_zs is None:
            del os.environ["PDR_ZSTREAMS"]
        else:
            os.environ["PDR_ZSTREAMS"] = env_zs
    ab = None
    if not args.no_ab:
        print("[bench] one against two z chunks in flight (medians)", file=sys.stderr, flush=True)
        ab = zstreams_ab(pa, n, ntr)
    stats = None
    if not args.no_ab:
        print("[bench] statistics of stage 1 in every step, on and off (medians)", file=sys.stderr, flush=True)
        stats = stats_on_off(pa, n, ntr)
    fp64 = None
    if not args.no_fp64:
        k64 = max(3, args.steps // 3)
        print(f"[bench] FP64 {n}^3: 1 warmup, {k64} timed", file=sys.stderr, flush=True)
        dt64, _, ok64 = run_gpu(pa, n, pa.PDR_FP64, k64, 1, ntr, False)
        sps64 = k64 / dt64
        print(f"[bench] FP64 {n}^3, profiled: {rsteps} steps, chunks one after another", file=sys.stderr, flush=True)
        os.environ["PDR_ZSTREAMS"] = "1"
        try:
            _, prof64, _ = run_gpu(pa, n, pa.PDR_FP64, rsteps, 1, ntr, True)
        finally:
            if env_zs is None:
                del os.environ["PDR_ZSTREAMS"]
            else:
                os.environ["PDR_ZSTREAMS"] = env_zs
        fp64 = {"steps_per_s": sps64, "grid_point_steps_per_s": sps64 * n**3, "ms_per_step": 1e3 / sps64,
                "finite": bool(ok64), "prof": prof64}

    class G:
        pass

    g = G()
    g.nx = g.ny = g.nz = n
    g.nki, g.nkj, g.nkk = n // 3 + 1, 2 * (n // 3) + 1, 2 * (n // 3) + 1
    g.mw = g.nki * g.nkj * g.nkk
    g.m1 = g.nki * g.nkj * n
    dom, best = None, -1.0
    for k, (cnt, ms) in prof.items():
        if cnt and ms > best and kernel_bytes(g, k, 8, 4, ntr):
            dom, best = k, ms
    roof = None

    def per_kernel(pr, csz, rsz, nsteps):
        out = {}
        for k, (cnt, ms) in pr.items():
            if cnt:
                b = kernel_bytes(g, k, csz, rsz, ntr)
                ch = chunks_per_stage(k, cnt, nsteps)
                if b:
                    b /= ch
                out[k] = {"launches": cnt, "chunks_per_stage": ch, "avg_ms": ms / cnt,
                          "GBps": (b / (ms / cnt * 1e-3) / 1e9) if b else None}
        return out

    kernels = per_kernel(prof, 8, 4, rsteps)
    if fp64:
        fp64["kernels"] = per_kernel(fp64.pop("prof"), 16, 8, rsteps)
    if dom:
        cnt, ms = prof[dom]
        avg = ms / cnt * 1e-3
        ch = chunks_per_stage(dom, cnt, rsteps)
        b = kernel_bytes(g, dom, 8, 4, ntr) / ch
        traffic = measured_traffic(dom, n, ntr)
        if traffic:
            traffic /= ch
        roof = {"bound": "hbm", "kernel": dom, "achieved": b / avg / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": b / avg / HBM_PEAK, "traffic": traffic,
                "traffic_source": (os.path.relpath(TRAFFIC_JSON, ROOT) + " (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                   "copy-calibrated, whole-box launch"
                                   + (f", divided by the {ch:g} chunks of a stage" if ch > 1 else "") + ")")
                if traffic else None,
                "bytes_per_launch": b, "chunks_per_stage": ch, "avg_launch_ms": avg * 1e3,
                "measured_in": f"{rsteps} timed steps with the chunks of a stage one after another "
                               "(PDR_ZSTREAMS=1), after the headline run"}
    ens = None
    if not args.no_ensemble:
        print("[bench] batched ensemble 16 x {32,48,64,96}^3", file=sys.stderr, flush=True)
        ens = ensemble_bench(pa)
    bstep = model_bytes_per_step(g, 8, 4, ntr)
    copy_gbps = pa.copy_bandwidth(1 << 30, 20)
    print("[bench] HBM ceiling probes", file=sys.stderr, flush=True)
    ceiling = hbm_ceiling(pa)
    if roof:
        roof["frac_of_measured_ceiling"] = roof["achieved"] * 1e9 / (ceiling["GBps"] * 1e9)
        roof["measured_ceiling_GBps"] = ceiling["GBps"]
    for k in kernels.values():
        if k.get("GBps"):
            k["frac_of_peak"] = k["GBps"] * 1e9 / HBM_PEAK
            k["frac_of_measured_ceiling"] = k["GBps"] / ceiling["GBps"]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, ntr, small=0 if args.no_cpu_small else 256)
    out = {
        "metric": METRIC,
        "value": sps32 * n**3,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt32 / args.steps,
        "steps_per_s": sps32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed (f64 frame, dt, viscous factors, tracer positions)",
        "data": "synthetic: seeded Rogallo type-10 field (k0=6) + uniform random tracers",
        "config": {"workload": f"{n}^3 homogeneous shear, Rogallo frame BF_ROT, S=1, nu=0.005, "
                               f"dt=0.002 (F_REMESH_T=250), {ntr} tracers, P_INTERP_ORDER=4",
                   "n": n, "tracers": ntr, "precision": "fp32-mixed"},
        "fp64": fp64,
        "fp32_over_fp64": (sps32 / fp64["steps_per_s"]) if fp64 else None,
        "step_model": {"bytes_per_step": bstep, "achieved_GBps": bstep * sps32 / 1e9,
                       "frac_of_peak": bstep * sps32 / HBM_PEAK,
                       "frac_of_measured_ceiling": bstep * sps32 / (ceiling["GBps"] * 1e9)},
        "kernels": kernels,
        "measured_copy_GBps": copy_gbps,
        "hbm_ceiling": ceiling,
        "zstreams_ab": ab,
        "stage_statistics": stats,
        "roofline": roof,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "finite": bool(ok32),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
