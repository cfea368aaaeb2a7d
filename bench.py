"""Benchmark of the nonlinear-term path on one MI355X.

A step is one pass of the full nonlinear-term path over all radial levels of
the workload (inverse transforms 5 x QST + scalar + gradient, fused FFTs and
grid products, forward transforms 3 x QST + scalar), with the coefficient
inputs resident in HBM.  The metric is radial-level transform pairs per
second: levels x steps / time, for the split-precision form (scaled FP16 x3
Legendre MFMA, FP32 grids); the FP64 plain form is timed beside it and the
relative L2 error of every output field against it is reported.  Prints one
JSON line (last line of stdout).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--nlev 97]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SHT pairs/s at l_max=255 (split-prec vs FP64 GPU) at ≤1e-6 rel err vs FP64"
PEAK = {"fp64": 78.6, "fp16": 2500.0}   # dense MFMA TFLOP/s
PEAK_HBM_GBS = 8000.0

# launch order of mhx_nl_step: (role, kind, Legendre weight, sources / comps)
# Legendre weights are flops per (N_h * lm_max) per level (FMA = 2), charged
# the FP64-equivalent count in both forms.
LAUNCHES = [
    ("syn_sg", "syn", 12, 1),
    ("syn_qst", "syn", 20, 3),   # u: w dw z
    ("syn_qst", "syn", 20, 4),   # curl u: z dz w ddw
    ("syn_qst", "syn", 20, 3),   # du/dr: dw ddw dz
    ("syn_qst", "syn", 20, 3),   # B: b db aj
    ("syn_qst", "syn", 20, 4),   # curl B: aj dj b ddb
    ("grid", "grid", 0, 0),
    ("ana_qst", "ana", 20, 3),
    ("ana_qst", "ana", 20, 3),
    ("ana_qst", "ana", 20, 3),
    ("ana_sh", "ana", 4, 1),
]
# kernels behind each role (prefix of the rocprofv3 kernel name)
KERNELS = {
    "split16": {"syn_sg": "k_nsyn16p<0", "syn_qst": "k_nsyn16p<", "grid": "k_grid_tile",
                "ana_qst": "k_nana16p<0", "ana_sh": "k_nana16p<1"},
    # l_max < 255 (n_phi != 768): the one-row slot grid stage
    "split16_slots": {"syn_sg": "k_nsyn16p<0", "syn_qst": "k_nsyn16p<", "grid": "k_nl_slots<float",
                      "ana_qst": "k_nana16p<0", "ana_sh": "k_nana16p<1"},
    # l_max > 255: the chunked split Legendre kernels
    "split16_chunked": {"syn_sg": "k_nsyn16<0", "syn_qst": "k_nsyn16<", "grid": "k_nl_slots<float",
                        "ana_qst": "k_nana16<0", "ana_sh": "k_nana16<1"},
    "fp64": {"syn_sg": "k_syn<2>", "syn_qst": "k_syn<3>", "grid": "k_nl_slots<double",
             "ana_qst": "k_ana<2>", "ana_sh": "k_ana<0>"},
}
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "nl_step_l255_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic(kind, w, nsrc, nh, lm_max, n_theta, n_phi, n_m, nlev, word, tword, p_only=False,
                rec=False):
    """(flops, HBM bytes) of one launch.  Legendre: w N_h lm_max flops per
    level; bytes = coefficient arrays (16 B per lm) + Fourier comps (n_theta x
    n_m complex of 2 word bytes) + the tables read once (P and dP, or P alone
    for the split kernels that fold dP onto P; tword bytes per entry: 8 FP64,
    4 for the FP16 hi/lo pair; none for a synthesis that generates its tables
    by recurrence).  Grid stage: 18 Fourier comps in, 10 out, 3
    du/dr grids (n_theta x n_phi) out."""
    four = n_theta * n_m * 2 * word
    coef = 16 * lm_max
    if kind == "grid":
        return 0.0, float(((18 + 10) * four + 3 * n_theta * n_phi * word) * nlev)
    flops = float(w * nh * lm_max * nlev)
    ntab = 1 if (w == 4 or p_only) else 2
    if kind == "syn" and rec:
        ntab = 0
    if kind == "syn":
        nout = 3
        byts = (nsrc * coef + nout * four) * nlev + ntab * tword * nh * lm_max
    else:
        nout = 3 if w == 20 else 1
        byts = (nsrc * four + nout * coef) * nlev + ntab * tword * nh * lm_max
    return flops, float(byts)


def host_cpu_model():
    """The host CPU's model name (lscpu), or the /proc/cpuinfo entry."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        m = re.search(r"Model name:\s*(.+)", out)
        if m:
            return m.group(1).strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            m = re.search(r"model name\s*:\s*(.+)", f.read())
            return m.group(1).strip() if m else None
    except OSError:
        return None


def cpu_baseline(l_max, nlev_sample, nrep=5):
    """The reference's own OpenMP native path (oracle/_ref, the reference's
    Fortran modules) on this host's cores over a bounded sample of levels:
    one warm-up pass, then the median of nrep timed passes."""
    import numpy as np
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if oracle.have_ref():
        with tempfile.TemporaryDirectory() as td:
            out = os.path.join(td, "bench.txt")
            t0 = time.time()
            oracle.run_ref("bench", l_max, 1, 20, "/dev/null", out, nlev=nlev_sample, nrep=nrep,
                           threads=threads)
            wall = time.time() - t0
            lines = open(out).read().split("\n")
        kv = dict(re.findall(r"(\w+)=\s*([^\s]+)", lines[0]))
        times = [float(x) for x in lines[2:] if x.strip()]
        med = float(np.median(times))
        return {"value": nlev_sample / med, "unit": "radial-level transform pairs/s",
                "cores": int(kv["threads"]), "kind": "reference", "host_cpu": host_cpu_model(),
                "pass_times_s": times,
                "sample": "%d radial levels of the l_max=%d nonlinear step (the reference's "
                          "nl_level chain, its own inputs), 1 warm-up + median of %d timed "
                          "passes (%.1f s wall)" % (nlev_sample, l_max, len(times), wall)}
    # without the reference build: the FP64 oracle, single thread, one level
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from magic_amd import workload
    from nl_reference import oracle_step, radial
    nt, nph = workload.grid_size(l_max)
    o = oracle.Plan(l_max, nt, nph)
    inputs = workload.nl_inputs(l_max, 1)
    t0 = time.time()
    oracle_step(o, inputs, *radial(1), [l_max])
    dt = time.time() - t0
    return {"value": 1.0 / dt, "unit": "radial-level transform pairs/s", "cores": 1,
            "kind": "port", "sample": "1 radial level of the l_max=%d nonlinear step" % l_max}


def run_td(L, nlev, tout, rad, lcut, reps=20):
    """The lm-space assembly after the step (mhx_get_td) on the step's FP64
    outputs and inputs, timed on a side stream; algorithmic bytes per level
    and lm: 10 complex in, 7 out."""
    import torch

    from magic_amd import sht, workload
    nt, nph = workload.grid_size(L)
    p = sht.Plan(L, nt, nph, max_levels=nlev)
    dev = tout[0].device
    ins = [t.clone() for t in tout[:3]] + [tout[4].clone()] + [t.clone() for t in tout[6:9]] + \
          [torch.ones_like(tout[0]) for _ in range(3)]
    outs = [torch.zeros_like(tout[0]) for _ in range(7)]
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    epsc = [0.5] * nlev
    for _ in range(3):
        sht.get_td(p, ins, outs, rad[0], rad[1], epsc, 1.0e3, lcut, stream=st)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        sht.get_td(p, ins, outs, rad[0], rad[1], epsc, 1.0e3, lcut, stream=st)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    p.close()
    byts = 17.0 * 16.0 * p.lm_max * nlev
    return {"kernel": "k_get_td", "avg_ms_per_call": ms, "bound": "hbm",
            "algorithmic_bytes": byts, "GB/s": byts / (ms * 1e-3) / 1e9,
            "frac_of_hbm_peak": byts / (ms * 1e-3) / 8.0e12}


# the Boussinesq MHD switch set with curl advection (the boussBenchSat
# sample's: src/rIter.f90:483-707 with l_adv_curl, no double curl)
RLOOP_BOUSS_MHD = {"l_conv": 1, "l_conv_nl": 1, "l_heat": 1, "l_heat_nl": 1, "l_mag": 1,
                   "l_mag_LF": 1, "l_mag_nl": 1, "l_adv_curl": 1, "l_corr": 1,
                   "l_cour_alf_damp": 1, "LFfac": 200.0, "CorFac": 1000.0, "opm": 0.2,
                   "courfac": 0.8, "alffac": 0.35, "oek": 1000.0, "ra": 1.1e5, "opr": 1.0}


def run_rloop(L, nlev, tin, precision, switches=None, reps=10):
    """The reference-side radial loop (mhx_rloop, the body of the
    rIter_batched_t plugin: syntheses, grid products, Courant steps,
    analyses and the lm-space assembly of every interior level) on the same
    fields in the given precision, host-synchronous per call as the
    reference's loop."""
    import numpy as np
    import torch

    from magic_amd import sht, workload
    nt, nph = workload.grid_size(L)
    p = sht.Plan(L, nt, nph, precision=precision, max_levels=nlev)
    dev = tin[0].device
    fields = dict(zip(sht.NL_INPUTS, tin))
    names = ("dwdt", "dzdt", "dpdt", "dsdt", "dVSrLM", "dbdt", "djdt", "dVxBhLM")
    outs = {n: torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for n in names}
    r = np.linspace(0.55, 1.45, nlev)
    one, zero = np.ones(nlev), np.zeros(nlev)
    rad = {"r": r, "or1": 1 / r, "or2": 1 / r ** 2, "or4": 1 / r ** 4, "orho1": one, "beta": zero,
           "otemp1": one, "visc": one, "lambda": one, "temp0": one, "epscProf": zero,
           "delxr2": np.full(nlev, 1e-3), "delxh2": np.full(nlev, 2e-2),
           "nR": np.arange(2, nlev + 2, dtype=np.int32), "l_R": np.full(nlev, L, dtype=np.int32)}
    dtr = torch.zeros(nlev, dtype=torch.float64, device=dev)
    dth = torch.zeros(nlev, dtype=torch.float64, device=dev)

    sw = RLOOP_BOUSS_MHD if switches is None else switches

    def call():
        sht.rloop(p, sw, nlev, rad, fields, outs, dtr, dth)
    for _ in range(3):
        call()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        call()
        ts.append(time.perf_counter() - t0)
    p.close()
    return {"entry": "mhx_rloop",
            "switches": "Boussinesq MHD, " + ("curl" if sw["l_adv_curl"] else "u.grad u") + " advection",
            "levels": nlev, "median_ms_per_call": 1e3 * float(np.median(ts)),
            "timing": "host clock, synchronous call"}


def run_diag(L, nlev, tin, reps=20):
    """Two FP64 calls next to the step, host-synchronous as the reference
    makes them, timed per call with the host clock (launches, the copy of
    the results and the wait included): the kinetic-energy sums over all
    levels (mhx_get_e_kin; algorithmic bytes 3 complex per level and lm) and
    the boundary products of one boundary (mhx_get_br_v_bcs on device grids:
    3 grids read, 2 written, then the sphertor analysis)."""
    import ctypes
    import time

    import numpy as np
    import torch

    from magic_amd import sht, workload
    nt, nph = workload.grid_size(L)
    p = sht.Plan(L, nt, nph, max_levels=1)
    r = np.linspace(0.54, 1.54, nlev)
    or2, orho = 1.0 / r ** 2, np.ones(nlev)
    w, dw, z = tin[0], tin[1], tin[3]

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps * 1e3

    ms_ek = timed(lambda: sht.get_e_kin_levels(p, w, dw, z, or2, orho))
    by_ek = 48.0 * p.lm_max * nlev
    g = torch.rand((3, nph, nt), dtype=torch.float64, device=tin[0].device) - 0.5
    oa = torch.zeros(p.lm_max, dtype=torch.complex128, device=g.device)
    ob = torch.zeros_like(oa)
    vp = ctypes.c_void_p

    def brv():
        sht._check(p._L.mhx_get_br_v_bcs(p.h, vp(g[0].data_ptr()), vp(g[1].data_ptr()),
                                         vp(g[2].data_ptr()), ctypes.c_double(0.7),
                                         ctypes.c_double(0.42), ctypes.c_double(1.0),
                                         vp(oa.data_ptr()), vp(ob.data_ptr())))
    ms_bv = timed(brv)
    p.close()
    return ({"entry": "mhx_get_e_kin", "levels": nlev, "avg_ms_per_call": ms_ek,
             "algorithmic_bytes": by_ek, "GB/s": by_ek / (ms_ek * 1e-3) / 1e9,
             "timing": "host clock, synchronous call"},
            {"entry": "mhx_get_br_v_bcs", "avg_ms_per_call": ms_bv,
             "grid": [nt, nph], "timing": "host clock, synchronous call"})


def run_form(precision, L, nlev, steps, warmup, tin, rad, lcut, courant=False):
    import numpy as np
    import torch

    from magic_amd import sht, workload
    nt, nph = workload.grid_size(L)
    p = sht.Plan(L, nt, nph, precision=precision, max_levels=nlev)
    dev = tin[0].device
    tout = [torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for _ in range(10)]
    # with courant: the step with the fused Courant maxima of every level
    # (courfac 2.5, alffac 1, Alfven damping of opm = 0.2 over delxr2 = 1e-3,
    # delxh2 = 2e-2), which returns when the maxima are on the host
    cvr = np.full(nlev, 0.36 / 1e-3)
    cvh = np.full(nlev, 0.36 / 2e-2)

    def step(stream=None):
        if courant:
            sht.nl_step_courant(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut, 2.5, 1.0,
                                cvr, cvh)
        else:
            sht.nl_step(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut, stream=stream)
    # steps issued on a side stream: each call returns without waiting (the
    # device outputs are ordered on the stream), so the host prepares the
    # next step while the current one runs
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())  # inputs and outputs were made there
    for _ in range(max(warmup, 2)):  # the second call captures the step's launch graph
        step(stream=st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(stream=st)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # per-launch durations (HIP events between the launches) in a separate pass
    p.set_timing(True)
    nt_steps = max(3, min(steps, 10))
    lt = np.zeros(len(LAUNCHES))
    for _ in range(nt_steps):
        step()
        lt += np.array(p.launch_times())
    p.set_timing(False)
    lt /= nt_steps
    split = precision in (sht.PREC_SPLIT16, sht.PREC_SPLIT16_REC)
    word, tword = (4, 4) if split else (8, 8)
    nh, lm_max, n_m = nt // 2, p.lm_max, p.n_m_max
    per_kernel = {}
    total_flop = flop_ms = 0.0
    for (role, kind, w, nsrc), t in zip(LAUNCHES, lt):
        fl, by = algorithmic(kind, w, nsrc, nh, lm_max, nt, nph, n_m, nlev, word, tword,
                             p_only=split and L <= 255,
                             rec=precision == sht.PREC_SPLIT16_REC)
        d = per_kernel.setdefault(role, {"ms": 0.0, "n": 0, "flop": 0.0, "bytes": 0.0})
        d["ms"] += t
        d["n"] += 1
        d["flop"] += fl
        d["bytes"] += by
        if kind != "grid":
            total_flop += fl
            flop_ms += t
    kernels = {}
    for role, v in per_kernel.items():
        ms = v["ms"] / v["n"]
        kernels[role] = {"avg_ms": ms, "launches": v["n"],
                         "TFLOP/s": v["flop"] / v["n"] / (ms * 1e-3) / 1e12,
                         "GB/s": v["bytes"] / v["n"] / (ms * 1e-3) / 1e9}
    outs = [t.cpu().numpy() for t in tout]
    td = run_td(L, nlev, tout, rad, lcut) if precision == sht.PREC_FP64 else None
    diag = run_diag(L, nlev, tin) if precision == sht.PREC_FP64 else None
    rl = None
    if L <= 255:
        rl = {"curl": run_rloop(L, nlev, tin, precision)}
        if precision == sht.PREC_FP64:
            # u.grad u advection: outside the fused step's switch sets, the
            # loop's own kernels
            rl["ugradu"] = run_rloop(L, nlev, tin, precision, dict(RLOOP_BOUSS_MHD, l_adv_curl=0))
    p.close()
    return {"td": td, "diag": diag, "rloop": rl, "pairs": nlev * steps / dt, "ms_per_step": 1e3 * dt / steps,
            "per_kernel": per_kernel, "kernels": kernels,
            "legendre_tflops": total_flop / (flop_ms * 1e-3) / 1e12, "outs": outs}


# workloads of the bench and the seed stream each one's synthetic inputs are
# drawn from (the stream number keeps the inputs identical across names)
WORKLOADS = {"shell255": 2, "batch64": 3, "shell511": 4, "shell1023": 5}


PROFILED = (255, 97)  # (l_max, n_r) of the committed PMC summary


def measured_traffic(form_key, role, workload):
    """HBM bytes per launch of the role's kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE scaled by a plain copy) when it
    covers this workload, else None."""
    if workload != PROFILED:
        return None
    try:
        data = json.load(open(TRAFFIC_JSON))
    except (OSError, ValueError):
        return None
    pref = KERNELS[form_key][role]
    hits = [v for k, v in data.get("kernels", {}).items() if k.startswith(pref)]
    if not hits:
        return None
    return sum(h["bytes"] for h in hits) / len(hits)


def roofline(form, form_key, workload):
    """Roofline of the dominant kernel (largest total time): MFMA-bound when
    its arithmetic intensity exceeds the ridge of its datapath, else HBM."""
    peak_key = "fp16" if form_key.startswith("split16") else "fp64"
    pk = form["per_kernel"]
    dom = max(pk, key=lambda k: pk[k]["ms"])
    d = pk[dom]
    avg_ms = d["ms"] / d["n"]
    flop, byts = d["flop"] / d["n"], d["bytes"] / d["n"]
    ridge = PEAK[peak_key] * 1e12 / (PEAK_HBM_GBS * 1e9)
    if flop > 0 and flop / byts >= ridge:
        achieved = flop / (avg_ms * 1e-3) / 1e12
        r = {"bound": "mfma", "achieved": achieved, "peak": PEAK[peak_key], "unit": "TFLOP/s",
             "frac": achieved / PEAK[peak_key]}
    else:
        achieved = byts / (avg_ms * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
             "frac": achieved / PEAK_HBM_GBS}
    r.update({"kernel": KERNELS[form_key][dom], "role": dom, "avg_launch_ms": avg_ms,
              "algorithmic_bytes": byts, "algorithmic_flop": flop,
              "traffic": measured_traffic(form_key, dom, workload)})
    return r


def bench_batch(args, dev):
    """The uneven batch (64 round-trip jobs, l_max 63..255, split form) from
    one captured graph; pairs are counted per (job, level)."""
    import numpy as np
    import torch

    from magic_amd import sht, workload
    specs = workload.batch_specs()
    rng = np.random.default_rng(5)
    jobs, host = [], []
    for L, kind, nf, nlev in specs:
        nc = 3 if kind == sht.JOB_VECTOR else 1
        x = workload.batch_coeffs(rng, L, nc, nf, nlev)
        host.append(x)
        jobs.append((L, kind, torch.from_numpy(x).to(dev),
                     torch.zeros(x.shape, dtype=torch.complex128, device=dev)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = sht.Batch(jobs)
    t_make = time.perf_counter() - t0
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    for _ in range(max(args.warmup, 1)):
        b.execute(st)
    torch.cuda.synchronize()
    # round-trip error against the identity (l >= 1 for S, T and Q of vectors)
    err = 0.0
    for (L, kind, nf, nlev), x, j in zip(specs, host, jobs):
        y = j[3].cpu().numpy()
        ls = np.concatenate([np.arange(m, L + 1) for m in range(L + 1)])
        want = x.copy()
        if kind == sht.JOB_VECTOR:
            want[0] = ls * (ls + 1.0) * x[0]
            y, want = y[..., ls >= 1], want[..., ls >= 1]
        for c in range(x.shape[0]):
            err = max(err, float(np.linalg.norm(y[c] - want[c]) / np.linalg.norm(want[c])))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.execute(st)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    pairs = sum(nlev for _, _, _, nlev in specs)
    b.status()
    form, shapes = b.shapes()
    # algorithmic bytes and FP64-equivalent Legendre flops of one execute:
    # coefficients in and out, the Fourier rows written and read on both
    # sides, and (per-job launches only) the grids written and read
    byts = flops = 0.0
    for L, kind, nf, nlev in specs:
        nc = 3 if kind == sht.JOB_VECTOR else 1
        nt = 3 * (L + 1) // 2
        nph, nm, lm, R = 2 * nt, L + 1, (L + 1) * (L + 2) // 2, nf * nlev
        four, grid = nc * R * nm * nt * 8, nc * R * nph * nt * 4
        byts += 2 * nc * R * lm * 16 + 2 * four + (0 if form == 1 else 2 * grid) + 2 * four
        flops += (40 if nc == 3 else 8) * (nt // 2) * lm * R
    tune = b.tuning()
    b.close()
    achieved = byts / dt / 1e9
    log("batch: %d jobs, %d (job, level) pairs, %.3f ms per execute, max round-trip err %.2e, "
        "made in %.1f s; form %d, shapes %s, tuned %s" % (len(specs), pairs, dt * 1e3, err, t_make, form,
                                                          shapes, tune))
    one = form == 1
    return {
        "metric": "SHT round-trip pairs/s, uneven batch of 64 jobs (split form, one %s)"
                  % ("persistent launch" if one else "graph of per-job launches"),
        "value": pairs / dt, "unit": "(job, radial level) round-trip pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16x3-f32acc", "data": "synthetic",
        "config": {"workload": "64 independent round-trip jobs (scal_to_spat + scal_to_SH or "
                               "torpol_to_spat + spat_to_qst), l_max in {63,127,191,255}, "
                               "n_theta = 3(l_max+1)/2, n_phi = 2 n_theta, n_r 17..65, "
                               "1..4 fields, (l+1)^-2 spectra",
                   "jobs": len(specs), "pairs_per_execute": pairs},
        "max_round_trip_rel_l2_err": err,
        "legendre_tflops_fp64_equiv": flops / dt / 1e12,
        "launch_form": "persistent" if one else "graph",
        "tile_shapes": [{"l_max": t[0], "orders_per_item": t[1], "level_groups_per_item": t[2],
                         "level_groups_per_item_from_split": t[3], "split_order": t[4],
                         "theta_rows_per_row_item": t[5]} for t in shapes] if one else None,
        "tuning": None if one else [{"l_max": t[0], "synthesis_form": t[1], "analysis_form": t[2],
                                     "fft_rows_per_block": t[3]} for t in tune],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "kernel": ("k_pbatch (algorithmic bytes / time of the one launch)" if one else
                                "whole batch graph (algorithmic bytes of all launches / time)")},
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 50 timed steps after 10 untimed ones: the clocks have settled before the
    # timed region (the first few launches of a run are up to 15 % slower)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--l_max", type=int, default=255)
    ap.add_argument("--nlev", type=int, default=97)
    ap.add_argument("--cpu-levels", type=int, default=None,
                    help="levels of the CPU baseline sample (default: all 97 at l_max=255, 8 at "
                         "l_max=511, 3 at l_max=1023, else 24)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["both", "split16", "fp64"], default="both")
    ap.add_argument("--courant", action="store_true",
                    help="time the step with the Courant maxima fused into its grid stage "
                         "(mhx_nl_step_courant; synchronous per step)")
    ap.add_argument("--tables", choices=["recurrence", "stored"], default="recurrence",
                    help="shell1023: split-form Legendre tables generated in-kernel by the FP64 "
                         "recurrence (default) or read from stored FP16 tables")
    ap.add_argument("--spectrum", default=None,
                    help="amplitude spectrum of the synthesised fields: l^-5/3 (default), "
                         "slopes-0-3 (default of shell1023: field f with (l+1)^-(f mod 4)), "
                         "white-to-l^-3 ((1+(l/64)^2)^-3/2), l+1^-<s>")
    ap.add_argument("--workload", choices=list(WORKLOADS), default="shell255",
                    help="shell255: the nonlinear step at l_max=255 x 97 (headline); batch64: "
                         "the uneven batch of 64 transform jobs in one launch; shell511: the "
                         "nonlinear step at l_max=511 x 161 with the l=1 poloidal field x100; "
                         "shell1023: the nonlinear step at l_max=1023 x 64, spectral slopes "
                         "0-3 per field, tables generated in-kernel")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the bench, run one device-to-device copy of this many MiB "
                         "(a kernel of known traffic, for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        log("bench.py measures one GPU only: the transform path is independent per radial "
            "level, so several GPUs would each run their own share of the levels; only "
            "--gpus 1 is measured")
        sys.exit(2)

    import numpy as np
    import torch

    from magic_amd import sht, workload

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    dev = torch.device("cuda:0")
    if args.workload == "batch64":
        print(json.dumps(bench_batch(args, dev)))
        return
    L, nlev = args.l_max, args.nlev
    dipole, kind = 1.0, "l^-5/3"
    if args.workload == "shell511":
        L, nlev, dipole = 511, 161, 100.0
    if args.workload == "shell1023":
        L, nlev, kind = 1023, 64, "slopes-0-3"
    kind = args.spectrum or kind
    nt, nph = workload.grid_size(L)
    log("l_max=%d %dx%d nlev=%d spectrum %s" % (L, nt, nph, nlev, kind))
    inputs = workload.nl_inputs(L, nlev, stream=WORKLOADS[args.workload], kind=kind, dipole=dipole)
    r = np.linspace(0.55, 1.45, nlev)
    rad = (1.0 / r ** 2, 1.0 / r ** 4, 1.0 / (1.0 + 0.1 * r), 0.8)
    lcut = [L] * nlev
    tin = [torch.from_numpy(x).to(dev) for x in inputs]
    torch.cuda.synchronize()

    forms = {}
    if args.form in ("both", "fp64"):
        forms["fp64"] = run_form(sht.PREC_FP64, L, nlev, args.steps, args.warmup, tin, rad, lcut,
                                 courant=args.courant)
        log("fp64: %.3f ms/step, %.1f pairs/s" % (forms["fp64"]["ms_per_step"],
                                                  forms["fp64"]["pairs"]))
    if args.form in ("both", "split16"):
        # the l_max=1023 case generates its tables in-kernel (--tables stored
        # runs the stored-table kernels instead)
        rec = args.workload == "shell1023" and args.tables == "recurrence"
        forms["split16"] = run_form(sht.PREC_SPLIT16_REC if rec else sht.PREC_SPLIT16, L, nlev,
                                    args.steps, args.warmup, tin,
                                    rad, lcut, courant=args.courant)
        log("split16: %.3f ms/step, %.1f pairs/s" % (forms["split16"]["ms_per_step"],
                                                     forms["split16"]["pairs"]))
    main_key = "split16" if "split16" in forms else "fp64"
    kern_key = main_key
    if main_key == "split16":
        # the theta-row tile grid stage runs on the n_phi = 768 grid unless
        # MHX_GRID_FORM=slots selects the one-row kernel
        tile = nph == 768 and os.environ.get("MHX_GRID_FORM") != "slots"
        kern_key = "split16_chunked" if L > 255 else ("split16" if tile else "split16_slots")
    main = forms[main_key]
    errs = None
    if "fp64" in forms and "split16" in forms:
        errs = {}
        for k, name in enumerate(sht.NL_OUTPUTS):
            a, b = forms["split16"]["outs"][k], forms["fp64"]["outs"][k]
            errs[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        log("max rel L2 error split vs fp64: %.3e" % max(errs.values()))
        # per l-band of 64 (the large-shell cases state their error by band)
        ls = np.concatenate([np.arange(m, L + 1) for m in range(L + 1)])
        band_errs = {}
        for k, name in enumerate(sht.NL_OUTPUTS):
            a, b = forms["split16"]["outs"][k], forms["fp64"]["outs"][k]
            row = []
            for b0 in range(0, L + 1, 64):
                sel = (ls >= b0) & (ls < b0 + 64)
                den = np.linalg.norm(b[:, sel])
                row.append(float(np.linalg.norm(a[:, sel] - b[:, sel]) / den) if den > 0 else 0.0)
            band_errs[name] = row

    cpu = None
    if not args.no_cpu:
        ncpu = args.cpu_levels or {255: 97, 511: 8, 1023: 3}.get(L, 24)
        log("cpu baseline (%d levels)" % ncpu)
        try:
            cpu = cpu_baseline(L, ncpu)
        except Exception as e:  # reported, not fatal for the GPU figure
            log("cpu baseline failed: %r" % (e,))
    res = {
        "metric": METRIC,
        "value": main["pairs"],
        "unit": "radial-level transform pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x3-f32acc" if main_key == "split16" else "f64",
        "data": "synthetic",
        "config": {"workload": "full nonlinear step (5 QST + scalar + gradient inverse, "
                               "fused FFTs + grid products, 3 QST + scalar forward)",
                   "l_max": L, "n_theta": nt, "n_phi": nph, "n_r": nlev,
                   "dipole_l1_factor": dipole, "spectrum": kind, "precision": main_key,
                   "tables": ("recurrence" if args.workload == "shell1023" and args.tables == "recurrence"
                              else "stored"),
                   "courant_maxima": bool(args.courant)},
        "legendre_tflops_fp64_equiv": main["legendre_tflops"],
        "kernels": main["kernels"],
        "roofline": roofline(main, kern_key, (L, nlev)),
        "cpu_baseline": cpu,
    }
    if "fp64" in forms and forms["fp64"].get("td"):
        # the next stage of the time step (src/get_td.f90), FP64, not in `value`
        res["td_assembly"] = forms["fp64"]["td"]
        if forms["fp64"].get("diag"):
            res["e_kin"], res["boundary_products"] = forms["fp64"]["diag"]
        # the radial loop the reference-side plugin runs, per plan precision
        # (not in `value`)
        rls = {k: f["rloop"] for k, f in forms.items() if f.get("rloop")}
        if rls:
            res["radial_loop"] = rls
            for k, v in rls.items():
                for kk, vv in v.items():
                    log("mhx_rloop %s %s: %.3f ms per call" % (k, kk, vv["median_ms_per_call"]))
        log("get_td: %.3f ms per call (%.0f GB/s)" % (res["td_assembly"]["avg_ms_per_call"],
                                                       res["td_assembly"]["GB/s"]))
    if "fp64" in forms and main_key != "fp64":
        f = forms["fp64"]
        res["fp64_form"] = {"value": f["pairs"], "ms_per_step": f["ms_per_step"],
                            "legendre_tflops": f["legendre_tflops"], "kernels": f["kernels"],
                            "roofline": roofline(f, "fp64", (L, nlev))}
        res["speedup_vs_fp64"] = main["pairs"] / f["pairs"]
    if errs is not None:
        res["rel_l2_err_vs_fp64"] = errs
        if L > 255:
            res["rel_l2_err_vs_fp64_per_l_band_of_64"] = band_errs
        res["max_rel_l2_err_vs_fp64"] = max(errs.values())
    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib << 18  # float32 elements
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB read + %d MiB written" % (args.calib_copy_mib,
                                                                args.calib_copy_mib))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
