"""Benchmark of the nonlinear-term path on one MI355X.

A step is one pass of the full nonlinear-term path over all radial levels of
the workload (inverse transforms 5 x QST + scalar + gradient, grid products,
forward transforms 3 x QST + scalar), with the coefficient inputs resident in
HBM.  The metric is radial-level transform pairs per second: levels x steps
/ time.  Prints one JSON line (last line of stdout).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--nlev 97]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SHT pairs/s at l_max=255 (split-prec vs FP64 GPU) at ≤1e-6 rel err vs FP64"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# launch order of mhx_nl_step and the algorithmic work of each launch:
# Legendre weights are flops per (N_h * lm_max) per level (FMA = 2), FFT and
# product launches count bytes per level.
LAUNCHES = [
    ("syn_scal_grad", "k_syn<SYN_SG>", "flop", 12),
    ("syn_qst_u", "k_syn<SYN_QST>", "flop", 20),
    ("syn_qst_curl_u", "k_syn<SYN_QST>", "flop", 20),
    ("syn_qst_du", "k_syn<SYN_QST>", "flop", 20),
    ("syn_qst_b", "k_syn<SYN_QST>", "flop", 20),
    ("syn_qst_curl_b", "k_syn<SYN_QST>", "flop", 20),
    ("grid_stage", "k_nl_fused", "fused", 0),
    ("ana_qst_adv", "k_ana<ANA_QST>", "flop", 20),
    ("ana_qst_vs", "k_ana<ANA_QST>", "flop", 20),
    ("ana_qst_vxb", "k_ana<ANA_QST>", "flop", 20),
    ("ana_sh_ugs", "k_ana<ANA_SH>", "flop", 4),
]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic(kind, w, nh, lm_max, n_theta, n_phi, n_m, nlev):
    if kind == "flop":
        return w * nh * lm_max * nlev
    # fused grid stage, bytes: 18 Fourier comps in, 10 out (n_theta x n_m
    # complex128 each) and the 3 du/dr grids written (n_theta x n_phi float64)
    return ((18 + 10) * n_theta * n_m * 16 + 3 * n_theta * n_phi * 8) * nlev


def cpu_baseline(l_max, nlev_sample):
    """The reference's own OpenMP native path (oracle/_ref, the reference's
    Fortran modules) on this host's cores over a bounded sample of levels."""
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if oracle.have_ref():
        with tempfile.TemporaryDirectory() as td:
            out = os.path.join(td, "bench.txt")
            t0 = time.time()
            oracle.run_ref("bench", l_max, 1, 20, "/dev/null", out, nlev=nlev_sample, nrep=1,
                           threads=threads)
            wall = time.time() - t0
            line = open(out).readline()
        import re
        kv = dict(re.findall(r"(\w+)=\s*([^\s]+)", line))
        mean_s = float(kv["mean_s"])
        return {"value": nlev_sample / mean_s, "unit": "radial-level transform pairs/s",
                "cores": int(kv["threads"]), "kind": "reference",
                "sample": "%d radial levels of the l_max=%d nonlinear step, 1 warm-up + 1 timed "
                          "pass (%.1f s wall)" % (nlev_sample, l_max, wall)}
    # without the reference build: the FP64 oracle, single thread, one level
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from magic_amd import workload
    from nl_reference import make_inputs, oracle_step, radial
    nt, nph = workload.grid_size(l_max)
    o = oracle.Plan(l_max, nt, nph)
    inputs = make_inputs(l_max, 1)
    t0 = time.time()
    oracle_step(o, inputs, *radial(1), [l_max])
    dt = time.time() - t0
    return {"value": 1.0 / dt, "unit": "radial-level transform pairs/s", "cores": 1,
            "kind": "port", "sample": "1 radial level of the l_max=%d nonlinear step" % l_max}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--l_max", type=int, default=255)
    ap.add_argument("--nlev", type=int, default=97)
    ap.add_argument("--cpu-levels", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        log("bench.py measures one GPU only: the transform path is independent per radial "
            "level and this build runs on a single MI355X (see DESIGN.md for sharding)")
        sys.exit(2)

    import numpy as np
    import torch

    from magic_amd import sht, workload

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    dev = torch.device("cuda:0")
    L, nlev = args.l_max, args.nlev
    nt, nph = workload.grid_size(L)
    log("plan l_max=%d %dx%d nlev=%d" % (L, nt, nph, nlev))
    p = sht.Plan(L, nt, nph, max_levels=nlev)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from nl_reference import make_inputs, radial
    inputs = make_inputs(L, nlev)
    rad = radial(nlev)
    lcut = [L] * nlev
    tin = [torch.from_numpy(x).to(dev) for x in inputs]
    tout = [torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for _ in range(10)]
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        sht.nl_step(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut)
    torch.cuda.synchronize()
    log("warmup done")

    p.set_timing(True)
    lt = np.zeros(len(LAUNCHES))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sht.nl_step(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut)
        lt += np.array(p.launch_times())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    p.set_timing(False)
    lt /= args.steps
    ms_per_step = 1e3 * dt / args.steps
    pairs = nlev * args.steps / dt
    log("timed %d steps: %.3f ms/step, %.1f pairs/s" % (args.steps, ms_per_step, pairs))

    # per-kernel averages and the dominant kernel
    nh, lm_max, n_m = nt // 2, p.lm_max, p.n_m_max
    per_kernel = {}
    for (name, kern, kind, w), t in zip(LAUNCHES, lt):
        a = algorithmic(kind, w, nh, lm_max, nt, nph, n_m, nlev)
        d = per_kernel.setdefault(kern, {"ms": 0.0, "n": 0, "work": 0.0, "kind": kind})
        d["ms"] += t
        d["n"] += 1
        d["work"] += a
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms"])
    dk = per_kernel[dom]
    avg_ms = dk["ms"] / dk["n"]
    work = dk["work"] / dk["n"]
    if dk["kind"] == "flop":
        achieved = work / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS}
    else:
        achieved = work / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS}
    roof.update({"kernel": dom, "avg_launch_ms": avg_ms, "traffic": None})
    total_flop = sum(algorithmic(k, w, nh, lm_max, nt, nph, n_m, nlev)
                     for _, _, k, w in LAUNCHES if k == "flop")
    kernels = {k: {"avg_ms": v["ms"] / v["n"], "launches": v["n"],
                   ("TFLOP/s" if v["kind"] == "flop" else "GB/s"):
                       (v["work"] / (v["ms"] * 1e-3) / (1e12 if v["kind"] == "flop" else 1e9))}
               for k, v in per_kernel.items()}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (%d levels)" % args.cpu_levels)
        try:
            cpu = cpu_baseline(L, args.cpu_levels)
        except Exception as e:  # reported, not fatal for the GPU figure
            log("cpu baseline failed: %r" % (e,))
    res = {
        "metric": METRIC,
        "value": pairs,
        "unit": "radial-level transform pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "full nonlinear step (5 QST + scalar + gradient inverse, "
                               "fused FFTs + grid products, 3 QST + scalar forward), "
                               "FP64 plain form",
                   "l_max": L, "n_theta": nt, "n_phi": nph, "n_r": nlev,
                   "precision": "fp64"},
        "legendre_tflops": total_flop / (sum(t for (_, _, k, _), t in zip(LAUNCHES, lt)
                                             if k == "flop") * 1e-3) / 1e12,
        "kernels": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
