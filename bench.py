"""Benchmark of the SIMPLE lid-driven-cavity hot path on one MI355X.

Workload (N=1): the 1024 x 1024 cavity, Re 1000, QUICK convection, 4th-order diffusion,
multicolour Gauss-Seidel momentum, geometric-multigrid V(2,2) p' (FP64 plain form), zero initial
fields. A step is one SIMPLE outer iteration over the whole grid, replayed as a captured hipGraph.
The timed region is bracketed by device synchronisations; inputs are resident in HBM.

Also reported (same JSON line): the roofline of the dominant kernel (one colour pass of the
u-momentum multicolour sweep) from hipEvent timings on the solver's stream, and a CPU baseline:
the oracle (a restatement of the reference's OpenMP CPU path) timed on this host's cores on a
bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form plain|mixed]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simple-hpc_amd", "py"))

HBM_PEAK_GBS = 8000.0


def config_text(n=1024, form="plain", maxit=1000000):
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = 1000\nconvection_scheme = QUICK\n"
            f"diffusion_order = 4\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_pre = 2\nmg_post = 2\nerror_tolerance = 1e-5\nmaxit = {maxit}\nbackend = gpu\n"
            f"write_fields = False\n")
    if form in ("mixed", "learned"):
        text += "mg_precision = mixed\n"
    return text


def cpu_baseline(p, seconds=12.0):
    """Oracle outer iterations at the full size on all host cores for about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = os.cpu_count() or 1
    pyoracle.set_threads(cores)
    o = pyoracle.Oracle(p)
    n, t = 0, 0.0
    while t < seconds and n < 200:
        dt, _ = o.run_fixed(1)
        t += dt
        n += 1
    return {"value": n / t, "unit": "outer iterations/s", "cores": cores, "kind": "port",
            "sample": f"{n} FP64 outer iterations of the same 1024^2 workload from zero fields "
                      f"(oracle, OpenMP, {cores} threads), {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--form", default="plain", choices=["plain", "mixed"])
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the multi-GPU decomposition is not built "
              "(see DESIGN.md, 'How the path would shard').", file=sys.stderr)
        sys.exit(2)
    import simple_gpu as sg
    p, _ = sg.params_from_config(config_text(a.n, a.form))
    s = sg.Solver(p)
    print(f"bench: {a.n}^2 {a.form}: warmup {a.warmup}", file=sys.stderr, flush=True)
    s.advance(a.warmup)
    print(f"bench: timing {a.steps} outer iterations", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    done, log, conv = s.advance(a.steps)
    t = time.perf_counter() - t0
    ms = 1e3 * t / max(done, 1)
    sec, bytes_, per = s.probe(sg.K_SWEEP_U, reps=100)
    gbs = bytes_ / sec / 1e9
    out = {
        "metric": "outer iterations/s, 1024^2 Re1000 plain FP64 (time-to-1e-5 = iterations x ms)",
        "value": done / t,
        "unit": "outer iterations/s",
        "n_gpus": 1, "steps": done, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (zero initial fields, lid velocity 1)",
        "config": {"workload": f"lid-driven cavity {a.n}x{a.n}, Re 1000, QUICK, 4th-order "
                               f"diffusion, gs_multicolour momentum, mg V(2,2) p', {a.form}",
                   "last_residuals": [float(x) for x in log[-1]] if done else None},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_mom_pass<5> (u colour pass)",
                     "kernel_us": sec * 1e6, "bytes_per_launch": bytes_},
    }
    if not a.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(p)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
