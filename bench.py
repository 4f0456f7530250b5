"""Benchmark of the SIMPLE lid-driven-cavity hot path on one MI355X.

Workload (N=1): the 1024 x 1024 cavity, Re 1000, QUICK convection, 4th-order diffusion,
multicolour Gauss-Seidel momentum, geometric-multigrid V(2,2) p', zero initial fields, run to the
absolute SIMPLE residual 1e-5 (the stop rule of the reference's run_outer_loop). A step is one
SIMPLE outer iteration over the whole grid, replayed as a captured hipGraph; the timed region is
bracketed by device synchronisations and the fields are resident in HBM throughout.

Forms: plain (FP64 multigrid, bit-identical to the reference's CPU path) and mixed (FP32 V-cycle
inside FP64 defect correction).

Also reported on the same JSON line: the roofline of the dominant kernel (one colour pass of the
u-momentum multicolour sweep) from hipEvent timings on the solver's stream, and a CPU baseline:
the oracle (a restatement of the reference's OpenMP CPU path) timed on this host's cores on a
bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fixed|tol] [--forms plain,mixed]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simple-hpc_amd", "py"))

HBM_PEAK_GBS = 8000.0


def config_text(n, form, re=1000, maxit=10000000):
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = {re}\nconvection_scheme = QUICK\n"
            f"diffusion_order = 4\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_pre = 2\nmg_post = 2\nerror_tolerance = 1e-5\nmaxit = {maxit}\nbackend = gpu\n"
            f"write_fields = False\ncheck_every = 500\n")
    if form in ("mixed", "learned"):
        text += "mg_precision = mixed\n"
    return text


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def run_form(sg, n, form, mode, steps, warmup, knobs=None):
    p, _ = sg.params_from_config(config_text(n, form))
    if knobs:  # given launch shapes: "tile,threads,mid,lds" (skips the autotuner)
        t, th, mid, lds = (int(x) for x in knobs.split(","))
        tune = (f"best momentum tile: given ({t})\n"
                f"best: tail_threads {th} mid_cells {mid} lds_cells {lds} : given\n")
    else:
        s = sg.Solver(p)
        s.advance(3)
        tune = s.autotune(reps=10)
        log(f"{n}^2 {form}: autotune -> {tune.strip().splitlines()[-1]}")
        s.close()
    s = sg.Solver(p)  # fresh fields; the tuned launch shapes
    for line in tune.strip().splitlines():
        f = line.split()
        if line.startswith("best momentum tile"):
            s.set_tuning(sg.TUNE_MOM_TILE, int(f[-1].strip("()")))
        elif line.startswith("best:"):
            s.set_tuning(sg.TUNE_TAIL_THREADS, int(f[2]))
            s.set_tuning(sg.TUNE_MID_CELLS, int(f[4]))
            s.set_tuning(sg.TUNE_LDS_CELLS, int(f[6]))
    if mode == "fixed":
        s.advance(warmup)
        log(f"{n}^2 {form}: timing {steps} outer iterations")
        t0 = time.perf_counter()
        done, lg, conv = s.advance(steps)
        t = time.perf_counter() - t0
        return s, {"form": form, "iterations": done, "seconds": t, "ms_per_iter": 1e3 * t / done,
                   "last_residuals": [float(x) for x in lg[-1]],
                   "coarse_sweeps_last": s.coarse_sweeps(),
                   "autotune": [l for l in tune.strip().splitlines() if l.startswith("best")]}
    log(f"{n}^2 {form}: to tolerance")
    t0 = time.perf_counter()
    total, conv, last = 0, False, None
    next_report = time.perf_counter() + 20
    while not conv and total < p.maxit:
        k, lg, conv = s.advance(2000)
        total += k
        last = lg[-1] if k else last
        if time.perf_counter() > next_report:
            log(f"  {form}: {total} iterations, residuals {last}")
            next_report += 20
        if k == 0:
            break
    t = time.perf_counter() - t0
    return s, {"form": form, "iterations": total, "seconds": t, "ms_per_iter": 1e3 * t / total,
               "converged": bool(conv), "last_residuals": [float(x) for x in last],
               "autotune": [l for l in tune.strip().splitlines() if l.startswith("best")]}


def cpu_baseline(sg, n, seconds=12.0):
    """Oracle outer iterations at the full size on all host cores for about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = os.cpu_count() or 1
    pyoracle.set_threads(cores)
    p, _ = sg.params_from_config(config_text(n, "plain"))
    o = pyoracle.Oracle(p)
    k, t = 0, 0.0
    while t < seconds and k < 500:
        dt, _ = o.run_fixed(1)
        t += dt
        k += 1
    return {"value": k / t, "unit": "outer iterations/s", "cores": cores, "kind": "port",
            "sample": f"{k} FP64 outer iterations of the {n}^2 workload from zero fields "
                      f"(oracle restating the reference's OpenMP path, {cores} threads), {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--mode", default="fixed", choices=["fixed", "tol"])
    ap.add_argument("--forms", default="plain")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--knobs", default=None,
                    help="launch shapes 'tile,tail_threads,mid_cells,lds_cells' instead of "
                         "autotuning (profiling runs)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU: the multi-GPU decomposition of the cavity is not built "
              "(DESIGN.md describes how it would shard).", file=sys.stderr)
        sys.exit(2)
    import simple_gpu as sg
    results = []
    probe = None
    for form in a.forms.split(","):
        s, r = run_form(sg, a.n, form, a.mode, a.steps, a.warmup, a.knobs)
        results.append(r)
        if probe is None:
            sec, bytes_, per = s.probe(sg.K_SWEEP_U, reps=200)
            probe = (sec, bytes_, per)
        s.close()
    head = results[0]
    sec, bytes_, per = probe
    gbs = bytes_ / sec / 1e9
    if a.mode == "tol":
        metric = f"time to SIMPLE residual 1e-5, {a.n}^2 Re1000 ({head['form']})"
        value, unit, hib = head["seconds"], "s", False
    else:
        metric = f"outer iterations/s, {a.n}^2 Re1000 ({head['form']})"
        value, unit, hib = head["iterations"] / head["seconds"], "outer iterations/s", True
    out = {
        "metric": metric, "value": value, "unit": unit,
        "n_gpus": 1, "steps": head["iterations"], "warmup": a.warmup if a.mode == "fixed" else 0,
        "ms_per_step": head["ms_per_iter"], "higher_is_better": hib, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (lid-driven cavity from zero fields, lid velocity 1)",
        "config": {"workload": f"lid-driven cavity {a.n}x{a.n}, Re 1000, QUICK, 4th-order "
                               f"diffusion, gs_multicolour momentum, mg V(2,2) p', tol 1e-5",
                   "forms": results},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_mom_sweep<5> (one u multicolour sweep, all colours)",
                     "kernel_us": sec * 1e6, "bytes_per_launch": bytes_,
                     "launches_per_iteration": per},
    }
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(sg, a.n)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
